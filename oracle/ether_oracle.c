/*
 * ether_oracle.c - CPU FP64 oracle (TEST INFRASTRUCTURE ONLY; see ether_oracle.h).
 *
 * Plain C restatement of the reference's hot-path routines.  Every function
 * names the reference file:line it follows.  Nothing here is linked into the
 * product library libether_gpu.so.
 */
#include "ether_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,       */
/* "Parallel random numbers: as easy as 1, 2, 3", SC'11, section 4.2).      */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void o_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c[0];
        uint64_t p1 = (uint64_t)PHILOX_M1 * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

void o_rng(uint64_t seed, uint32_t id, uint32_t replica, uint32_t step, uint32_t purpose,
           uint32_t attempt, uint32_t out[4]) {
    uint32_t ctr[4] = {id, replica, step, (purpose << 16) | (attempt & 0xffffu)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    o_philox4x32_10(ctr, key, out);
}

/* uniform in [-1,1) from the top 24 bits; get_random_num(-1,1) (src/get_random_num.f90:29-30) */
static double sym24(uint32_t x24) { return (double)((int32_t)x24 - (1 << 23)) * 0x1p-23; }
/* acceptance uniform in (0,1] */
static double eta24(uint32_t w) { return (double)((w >> 8) + 1u) * 0x1p-24; }

/* ------------------------------------------------------------------------ */
/* Neighbour search: src/getting_nbd.f90:46-82 with src/search_nbd.f90:34-58 */
/* ------------------------------------------------------------------------ */
static int search_nbd(const double* p0, const double* p1, double d, double tol,
                      const double scabc[3][3], const int32_t* open_bc) {
    static const int shift[3] = {0, 1, -1};
    int nx = open_bc[0] ? 1 : 3, ny = open_bc[1] ? 1 : 3, nz = open_bc[2] ? 1 : 3;
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b)
            for (int c = 0; c < nz; ++c) {
                double r2 = 0.0;
                for (int x = 0; x < 3; ++x) {
                    double img = scabc[0][x] * shift[a] + scabc[1][x] * shift[b] +
                                 scabc[2][x] * shift[c] + p1[x];
                    double dr = p0[x] - img;
                    r2 += dr * dr;
                }
                if (fabs(d - sqrt(r2)) <= tol) return 1;
            }
    return 0;
}

int o_neighbours_brute(const double* abc, int32_t n_basis, const double* basis_xyz,
                       const int32_t* sc, const int32_t* open_bc, int32_t n_shells,
                       const double* shell_d, double tol, int32_t capacity,
                       int32_t* nn_count, int32_t* nn_idx) {
    int32_t n = sc[0] * sc[1] * sc[2] * n_basis;
    double* pos = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    if (!pos) return -1;
    /* site order and positions: src/generate_supercell.f90:115-132 */
    int32_t s = 0;
    for (int k = 0; k < sc[2]; ++k)
        for (int j = 0; j < sc[1]; ++j)
            for (int i = 0; i < sc[0]; ++i)
                for (int l = 0; l < n_basis; ++l, ++s)
                    for (int x = 0; x < 3; ++x)
                        pos[3 * s + x] = abc[0 * 3 + x] * i + abc[1 * 3 + x] * j +
                                         abc[2 * 3 + x] * k + basis_xyz[3 * l + x];
    double scabc[3][3];
    for (int a = 0; a < 3; ++a)
        for (int x = 0; x < 3; ++x) scabc[a][x] = sc[a] * abc[a * 3 + x];
    int rc = 0;
    for (int32_t i = 0; i < n && rc == 0; ++i)
        for (int32_t sh = 0; sh < n_shells && rc == 0; ++sh) {
            int32_t cnt = 0;
            for (int32_t j = 0; j < n; ++j) {
                if (j == i) continue;
                if (search_nbd(pos + 3 * i, pos + 3 * j, shell_d[sh], tol, scabc, open_bc)) {
                    if (cnt >= capacity) { rc = -2; break; }
                    nn_idx[((size_t)sh * n + i) * capacity + cnt] = j;
                    ++cnt;
                }
            }
            nn_count[(size_t)sh * n + i] = cnt;
        }
    free(pos);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Hamiltonian: src/Hamiltonian.f90:20-54 (dispatch), JSiSj :57-105,        */
/* gmbSH :108-124, siaS2 :127-143.                                           */
/* ------------------------------------------------------------------------ */
static const double* jvec(const o_model* m, int sh, int a, int b) {
    return m->j_exc + (((size_t)sh * m->n_species + (a - 1)) * m->n_species + (b - 1)) * 3;
}

/* JSiSj: sum over shells and neighbours of (Si o Sj) . J */
static double jsisj(const o_model* m, const double* S, int32_t i, const double* si) {
    double e = 0.0;
    int a = m->species[i];
    for (int sh = 0; sh < m->n_shells; ++sh) {
        int cnt = m->nn_count[(size_t)sh * m->n_sites + i];
        const int32_t* nb = m->nn_idx + ((size_t)sh * m->n_sites + i) * m->capacity;
        for (int k = 0; k < cnt; ++k) {
            int j = nb[k];
            const double* J = jvec(m, sh, a, m->species[j]);
            const double* sj = S + 3 * (size_t)j;
            e += si[0] * sj[0] * J[0] + si[1] * sj[1] * J[1] + si[2] * sj[2] * J[2];
        }
    }
    return e;
}

static double gmbsh(const o_model* m, int32_t i, const double* si) {
    double s = m->spin[m->species[i] - 1];
    return -m->g_factor * m->mu_b * s * (si[0] * m->h[0] + si[1] * m->h[1] + si[2] * m->h[2]);
}

static double sias2(const o_model* m, int32_t i, const double* si2) {
    const double* d = m->sia_vec + 3 * (size_t)(m->species[i] - 1);
    return d[0] * si2[0] + d[1] * si2[1] + d[2] * si2[2];
}

double o_site_energy(const o_model* m, const double* S, int32_t i) {
    const double* si = S + 3 * (size_t)i;
    double si2[3] = {si[0] * si[0], si[1] * si[1], si[2] * si[2]};
    double e = jsisj(m, S, i, si);
    if (m->zeeman) e += gmbsh(m, i, si);
    if (m->sia && m->model == 1) e += sias2(m, i, si2);
    return e;
}

double o_delta_energy(const o_model* m, const double* S, int32_t i, const double* s_new) {
    const double* so = S + 3 * (size_t)i;
    double d[3] = {s_new[0] - so[0], s_new[1] - so[1], s_new[2] - so[2]};
    double d2[3] = {s_new[0] * s_new[0] - so[0] * so[0], s_new[1] * s_new[1] - so[1] * so[1],
                    s_new[2] * s_new[2] - so[2] * so[2]};
    double e = jsisj(m, S, i, d);
    if (m->zeeman) e += gmbsh(m, i, d);
    if (m->sia && m->model == 1) e += sias2(m, i, d2);
    return e;
}

/* JSiSj_ovrr: A = sum_j J o Sj (src/get_ovrr_vec.f90:41-85) */
void o_exchange_field(const o_model* m, const double* S, int32_t i, double* h3) {
    h3[0] = h3[1] = h3[2] = 0.0;
    int a = m->species[i];
    for (int sh = 0; sh < m->n_shells; ++sh) {
        int cnt = m->nn_count[(size_t)sh * m->n_sites + i];
        const int32_t* nb = m->nn_idx + ((size_t)sh * m->n_sites + i) * m->capacity;
        for (int k = 0; k < cnt; ++k) {
            int j = nb[k];
            const double* J = jvec(m, sh, a, m->species[j]);
            const double* sj = S + 3 * (size_t)j;
            h3[0] += J[0] * sj[0];
            h3[1] += J[1] * sj[1];
            h3[2] += J[2] * sj[2];
        }
    }
}

/* get_tot_energy (src/get_tot_energy.f90:35-45), split into its three terms */
void o_total_energy(const o_model* m, const double* S, double* parts3) {
    double ex = 0.0, ze = 0.0, si = 0.0;
    for (int32_t i = 0; i < m->n_sites; ++i) {
        const double* s = S + 3 * (size_t)i;
        double s2[3] = {s[0] * s[0], s[1] * s[1], s[2] * s[2]};
        ex += jsisj(m, S, i, s);
        if (m->zeeman) ze += gmbsh(m, i, s);
        if (m->sia && m->model == 1) si += sias2(m, i, s2);
    }
    parts3[0] = ex; parts3[1] = ze; parts3[2] = si;
}

void o_ising_bond_sums(const o_model* m, const double* S, const int32_t* class_of_edge,
                       int32_t n_classes, int64_t* bonds) {
    for (int k = 0; k < n_classes; ++k) bonds[k] = 0;
    for (int32_t i = 0; i < m->n_sites; ++i) {
        int si = S[3 * (size_t)i + 2] > 0 ? 1 : -1;
        for (int sh = 0; sh < m->n_shells; ++sh) {
            int cnt = m->nn_count[(size_t)sh * m->n_sites + i];
            size_t base = ((size_t)sh * m->n_sites + i) * m->capacity;
            for (int k = 0; k < cnt; ++k) {
                int c = class_of_edge[base + k];
                if (c < 0) continue;
                int sj = S[3 * (size_t)m->nn_idx[base + k] + 2] > 0 ? 1 : -1;
                bonds[c] += si * sj;
            }
        }
    }
}

/* get_tot_magnetisation (src/get_tot_magnetisation.f90:35-38) */
void o_magnetisation(const o_model* m, const double* S, double* M3) {
    M3[0] = M3[1] = M3[2] = 0.0;
    for (int32_t i = 0; i < m->n_sites; ++i) {
        double w = m->stagger[i] * m->spin[m->species[i] - 1];
        for (int x = 0; x < 3; ++x) M3[x] += w * S[3 * (size_t)i + x];
    }
}

/* ------------------------------------------------------------------------ */
/* Proposals: src/update_spin_details.f90:28-51, George_Marsaglia :29-39     */
/* ------------------------------------------------------------------------ */
static void normalise(double* v) {
    /* ref2(1:3) = ref2(1:3)/sqrt(dot_product(...)) (src/update_spin_details.f90:51) */
    double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= n; v[1] /= n; v[2] /= n;
}

static double wrap_pi(double p) {
    const double two_pi = 6.283185307179586476925286766559;
    return p - two_pi * floor((p + 3.14159265358979323846264338327950) / two_pi);
}

void o_propose(const o_model* m, uint64_t seed, uint32_t id, uint32_t replica, uint32_t step,
               uint32_t purpose, const double* s_old, double phi_old, double* s_new,
               double* phi_new, double* eta) {
    uint32_t w[4];
    o_rng(seed, id, replica, step, purpose, 0, w);
    *eta = eta24(w[2]);
    if (m->model == 0) {           /* Ising: ref2(3) = -ref1(3) */
        s_new[0] = s_old[0]; s_new[1] = s_old[1]; s_new[2] = -s_old[2];
        *phi_new = phi_old;
        return;
    }
    if (!m->angle) {
        /* Marsaglia 1972: v1,v2 ~ U(-1,1), retry while s >= 1.  Two candidate
         * pairs per Philox block; the acceptance test is exact in integers. */
        for (uint32_t att = 0;; ++att) {
            if (att > 0) o_rng(seed, id, replica, step, purpose, att, w);
            uint32_t cx[2] = {w[0] >> 8, ((w[3] >> 16) << 8) | (w[0] & 0xffu)};
            uint32_t cy[2] = {w[1] >> 8, ((w[3] & 0xffffu) << 8) | (w[1] & 0xffu)};
            for (int c = 0; c < 2; ++c) {
                int64_t a1 = (int64_t)cx[c] - (1 << 23), a2 = (int64_t)cy[c] - (1 << 23);
                int64_t q = a1 * a1 + a2 * a2;
                if (q < ((int64_t)1 << 46)) {
                    double v1 = a1 * 0x1p-23, v2 = a2 * 0x1p-23, s = q * 0x1p-46;
                    double r = sqrt(1.0 - s);
                    s_new[0] = 2.0 * v1 * r;
                    s_new[1] = 2.0 * v2 * r;
                    s_new[2] = 1.0 - 2.0 * s;
                    normalise(s_new);
                    *phi_new = 0.0;   /* ref2(5) = 0 */
                    return;
                }
            }
        }
    }
    /* ANGLE: S_z ~ U(-1,1); phi' = phi + dphi*U(-1,1) (src/update_spin_details.f90:39-46) */
    double sz = sym24(w[0] >> 8);
    double phi = wrap_pi(phi_old + m->dphi * sym24(w[1] >> 8));
    double r = sqrt(1.0 - sz * sz);
    s_new[0] = r * cos(phi);
    s_new[1] = r * sin(phi);
    s_new[2] = sz;
    normalise(s_new);
    *phi_new = phi;
}

/* generate_supercell start state (src/generate_supercell.f90:34-63): Ising
 * atom(i,3) = (-1)**i by structure-file index; XYZ George_Marsaglia, or for
 * ANGLE S_z = get_random_num(-1,1) and phi = get_random_num(0, 2 pi). */
void o_start_state(uint64_t seed, int32_t model, int32_t angle, int32_t nb,
                   const int32_t* file_index, double* S, double* phi) {
    o_model m;
    memset(&m, 0, sizeof m);
    m.model = 1;
    for (int32_t l = 0; l < nb; ++l) {
        double* s = S + 3 * (size_t)l;
        phi[l] = 0.0;
        if (model == 0) {
            s[0] = 0.0; s[1] = 0.0;
            s[2] = (file_index[l] % 2 == 0) ? 1.0 : -1.0;
        } else if (!angle) {
            double old[3] = {0.0, 0.0, 1.0}, eta;
            o_propose(&m, seed, (uint32_t)l, 0, 0, O_RNG_START, old, 0.0, s, &phi[l], &eta);
        } else {
            uint32_t w[4];
            o_rng(seed, (uint32_t)l, 0, 0, O_RNG_START, 0, w);
            double sz = sym24(w[0] >> 8);
            double p = 6.283185307179586476925286766559 * (double)(w[1] >> 8) * 0x1p-24;
            double r = sqrt(1.0 - sz * sz);
            s[0] = r * cos(p); s[1] = r * sin(p); s[2] = sz;
            normalise(s);
            phi[l] = p;
        }
    }
}

uint32_t o_ising_threshold(double beta, double dE) {
    double p = exp(-beta * dE);
    if (!(p < 1.0)) return 0xffffffffu;
    double t = floor(p * 4294967296.0);
    if (t > 4294967294.0) t = 4294967294.0;
    return (uint32_t)t;
}

/* Marsaglia point from an accepted 24-bit pair (src/George_Marsaglia.f90:36-38) */
static void marsaglia_from(uint32_t cx, uint32_t cy, double* s_new) {
    int64_t a1 = (int64_t)cx - (1 << 23), a2 = (int64_t)cy - (1 << 23);
    double v1 = a1 * 0x1p-23, v2 = a2 * 0x1p-23, s = (a1 * a1 + a2 * a2) * 0x1p-46;
    double r = sqrt(1.0 - s);
    s_new[0] = 2.0 * v1 * r;
    s_new[1] = 2.0 * v2 * r;
    s_new[2] = 1.0 - 2.0 * s;
    normalise(s_new);
}

static int in_disc(uint32_t cx, uint32_t cy) {
    int64_t a1 = (int64_t)cx - (1 << 23), a2 = (int64_t)cy - (1 << 23);
    return a1 * a1 + a2 * a2 < ((int64_t)1 << 46);
}

void o_propose_site(const o_model* m, uint64_t seed, uint32_t id, int32_t R, uint32_t step,
                    const double* s_old, const double* phi_old, double* s_new, double* phi_new,
                    double* eta) {
    if (m->model == 0 || m->angle) {
        for (int32_t r = 0; r < R; ++r)
            o_propose(m, seed, id, (uint32_t)r, step, O_RNG_METROPOLIS, s_old + 3 * (size_t)r,
                      phi_old ? phi_old[r] : 0.0, s_new + 3 * (size_t)r, phi_new + r, eta + r);
        return;
    }
    /* Marsaglia candidates pooled over each group of 64 consecutive replicas:
     * a replica whose first pair is outside the disc takes, in replica order,
     * the second pair of the replicas whose second pair is inside; if those run
     * out it draws attempts 1, 2, ... itself (same rule as the device wave). */
    for (int32_t g0 = 0; g0 < R; g0 += 64) {
        int32_t g1 = g0 + 64 < R ? g0 + 64 : R;
        uint32_t dx[64], dy[64];
        int nd = 0;
        for (int32_t r = g0; r < g1; ++r) {
            uint32_t w[4];
            o_rng(seed, id, (uint32_t)r, step, O_RNG_METROPOLIS, 0, w);
            uint32_t x1 = ((w[3] >> 16) << 8) | (w[0] & 0xffu);
            uint32_t y1 = ((w[3] & 0xffffu) << 8) | (w[1] & 0xffu);
            if (in_disc(x1, y1)) { dx[nd] = x1; dy[nd] = y1; ++nd; }
        }
        int nj = 0;
        for (int32_t r = g0; r < g1; ++r) {
            uint32_t w[4];
            o_rng(seed, id, (uint32_t)r, step, O_RNG_METROPOLIS, 0, w);
            eta[r] = eta24(w[2]);
            phi_new[r] = 0.0;
            double* sn = s_new + 3 * (size_t)r;
            if (in_disc(w[0] >> 8, w[1] >> 8)) {
                marsaglia_from(w[0] >> 8, w[1] >> 8, sn);
                continue;
            }
            if (nj < nd) {
                marsaglia_from(dx[nj], dy[nj], sn);
                ++nj;
                continue;
            }
            ++nj;
            for (uint32_t att = 1;; ++att) {
                o_rng(seed, id, (uint32_t)r, step, O_RNG_METROPOLIS, att, w);
                uint32_t x0 = w[0] >> 8, y0 = w[1] >> 8;
                uint32_t x1 = ((w[3] >> 16) << 8) | (w[0] & 0xffu);
                uint32_t y1 = ((w[3] & 0xffffu) << 8) | (w[1] & 0xffu);
                if (in_disc(x0, y0)) { marsaglia_from(x0, y0, sn); break; }
                if (in_disc(x1, y1)) { marsaglia_from(x1, y1, sn); break; }
            }
        }
    }
}

int64_t o_sweep_batch(const o_model* m, double* S, double* phi, int32_t R, const int32_t* order,
                      int32_t n, const int32_t* rng_id, const double* beta, uint64_t seed,
                      uint32_t step, int64_t* acc) {
    const size_t N = (size_t)m->n_sites;
    double* so = (double*)malloc(sizeof(double) * 3 * (size_t)R);
    double* sn = (double*)malloc(sizeof(double) * 3 * (size_t)R);
    double* po = (double*)malloc(sizeof(double) * (size_t)R);
    double* pn = (double*)malloc(sizeof(double) * (size_t)R);
    double* eta = (double*)malloc(sizeof(double) * (size_t)R);
    int64_t total = 0;
    for (int32_t t = 0; t < n; ++t) {
        const int32_t i = order[t];
        for (int32_t r = 0; r < R; ++r) {
            for (int x = 0; x < 3; ++x) so[3 * r + x] = S[(r * N + i) * 3 + x];
            po[r] = phi ? phi[r * N + i] : 0.0;
        }
        if (m->model != 0) o_propose_site(m, seed, (uint32_t)rng_id[i], R, step, so, po, sn, pn, eta);
        for (int32_t r = 0; r < R; ++r) {
            double* Sr = S + r * N * 3;
            double* s = Sr + 3 * (size_t)i;
            int ok;
            if (m->model == 0) {
                /* Ising: word r&3 of the block with id (slot, r>>2) */
                uint32_t w[4];
                o_rng(seed, (uint32_t)rng_id[i], (uint32_t)r >> 2, step, O_RNG_METROPOLIS, 0, w);
                double snew[3] = {s[0], s[1], -s[2]};
                uint32_t thr = o_ising_threshold(beta[r], o_delta_energy(m, Sr, i, snew));
                ok = w[r & 3] < thr || thr == 0xffffffffu;
                if (ok) s[2] = -s[2];
            } else {
                double dE = o_delta_energy(m, Sr, i, sn + 3 * (size_t)r);
                ok = -beta[r] * dE > log(eta[r]);
                if (ok) {
                    for (int x = 0; x < 3; ++x) s[x] = sn[3 * r + x];
                    if (phi) phi[r * N + i] = pn[r];
                }
            }
            if (ok) {
                ++total;
                if (acc) acc[r] += 1;
            }
        }
    }
    free(so); free(sn); free(po); free(pn); free(eta);
    return total;
}

/* Metropolis (src/Metropolis.f90:35-62): propose, dE, accept if -beta*dE > log(eta).
 * One replica on its own (a group of one: no candidate pooling). */
int64_t o_sweep(const o_model* m, double* S, double* phi, const int32_t* order, int32_t n,
                const int32_t* rng_id, double beta, uint64_t seed, uint32_t replica,
                uint32_t step) {
    int64_t acc = 0;
    for (int32_t t = 0; t < n; ++t) {
        int32_t i = order[t];
        double* so = S + 3 * (size_t)i;
        double sn[3], phin, eta;
        if (m->model == 0) {
            uint32_t w[4];
            o_rng(seed, (uint32_t)rng_id[i], replica >> 2, step, O_RNG_METROPOLIS, 0, w);
            sn[0] = so[0]; sn[1] = so[1]; sn[2] = -so[2];
            double dE = o_delta_energy(m, S, i, sn);
            uint32_t thr = o_ising_threshold(beta, dE);
            uint32_t u = w[replica & 3u];
            if (u < thr || thr == 0xffffffffu) {
                so[2] = sn[2];
                ++acc;
            }
            continue;
        }
        double ph = phi ? phi[i] : 0.0;
        o_propose(m, seed, (uint32_t)rng_id[i], replica, step, O_RNG_METROPOLIS, so, ph, sn,
                  &phin, &eta);
        double dE = o_delta_energy(m, S, i, sn);
        if (-beta * dE > log(eta)) {
            so[0] = sn[0]; so[1] = sn[1]; so[2] = sn[2];
            if (phi) phi[i] = phin;
            ++acc;
        }
    }
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Overrelaxation: src/overrelaxation.f90:30-49                              */
/* ------------------------------------------------------------------------ */
void o_reflect(const o_model* m, double* S, int32_t i) {
    double A[3];
    o_exchange_field(m, S, i, A);
    if (m->zeeman) {
        double s = m->spin[m->species[i] - 1];
        for (int x = 0; x < 3; ++x) A[x] += -m->g_factor * m->mu_b * s * m->h[x];
    }
    double aa = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
    if (!(aa > 0.0)) return;
    double* si = S + 3 * (size_t)i;
    double f = 2.0 * (A[0] * si[0] + A[1] * si[1] + A[2] * si[2]) / aa;
    double r[3] = {f * A[0] - si[0], f * A[1] - si[1], f * A[2] - si[2]};
    normalise(r);
    si[0] = r[0]; si[1] = r[1]; si[2] = r[2];
}

void o_overrelax_pass(const o_model* m, double* S, double* phi, const int32_t* order,
                      int32_t n, const int32_t* rng_id, double frac, uint64_t seed,
                      uint32_t replica, uint32_t step, uint32_t pass) {
    for (int32_t t = 0; t < n; ++t) {
        int32_t i = order[t];
        if (frac < 1.0) {
            uint32_t w[4];
            o_rng(seed, (uint32_t)rng_id[i], replica, step, O_RNG_OVRR, pass, w);
            if (!((double)(w[0] >> 8) * 0x1p-24 < frac)) continue;
        }
        o_reflect(m, S, i);
        if (phi) phi[i] = atan2(S[3 * (size_t)i + 1], S[3 * (size_t)i]);
    }
}

/* ------------------------------------------------------------------------ */
/* One chain through the MCS loop of src/ether.f90:159-203 (no tempering):   */
/* OVRR when stepi%ovrr_iv==0 and stepi%coa!=0 (:163-169), sweep (:172),     */
/* sample when stepi>eq and stepi%coa==0 (src/observable.f90:30-43).         */
/* mom[8] += {1, e, e^2, e^4, m, m^2, m^4, accepted/N per sweep}.            */
/* ------------------------------------------------------------------------ */
void o_run_chain(const o_model* m, double* S, double* phi, const int32_t* order, int32_t n,
                 const int32_t* rng_id, double beta, uint64_t seed, uint32_t replica,
                 int32_t mcs, int32_t eq, int32_t coa, int32_t ovrr_iv, double ovrr_p,
                 double* mom) {
    const double N = (double)m->n_sites;
    for (int32_t stepi = 1; stepi <= mcs; ++stepi) {
        const uint32_t sweep = (uint32_t)(stepi - 1);
        if (ovrr_iv > 0 && stepi % ovrr_iv == 0 && stepi % coa != 0) {
            int full = (int)floor(ovrr_p);
            double frac = ovrr_p - full;
            int passes = full + (frac > 0.0 ? 1 : 0);
            for (int p = 0; p < passes; ++p)
                o_overrelax_pass(m, S, phi, order, n, rng_id, p < full ? 1.0 : frac, seed, replica,
                                 sweep, (uint32_t)p);
        }
        int64_t acc = o_sweep(m, S, phi, order, n, rng_id, beta, seed, replica, sweep);
        mom[7] += (double)acc / N;
        if (stepi > eq && stepi % coa == 0) {
            double parts[3], M[3];
            o_total_energy(m, S, parts);
            o_magnetisation(m, S, M);
            double e = (parts[0] + parts[1] + parts[2]) / (2.0 * N);
            double mm = sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]) / N;
            mom[0] += 1.0;
            mom[1] += e;
            mom[2] += e * e;
            mom[3] += e * e * e * e;
            mom[4] += mm;
            mom[5] += mm * mm;
            mom[6] += mm * mm * mm * mm;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Estimators                                                                */
/* ------------------------------------------------------------------------ */
void o_repeat_observables(const double* mom, double beta, int32_t n_sites, double* out6) {
    double n = mom[0];
    /* get_eng 'avg_eng' (src/get_eng.f90:43-56) */
    double e = mom[1] / n, e2 = mom[2] / n, e4 = mom[3] / n;
    out6[0] = e;
    out6[1] = 1.0 - (1.0 / 3.0) * (e4 / (e2 * e2));
    out6[2] = beta * beta * (e2 - e * e) * n_sites;
    /* get_mag 'avg_mag' (src/get_mag.f90:44-55) */
    double mm = mom[4] / n, m2 = mom[5] / n, m4 = mom[6] / n;
    out6[3] = mm;
    out6[4] = 1.0 - (1.0 / 3.0) * (m4 / (m2 * m2));
    out6[5] = beta * (m2 - mm * mm) * n_sites;
}

void o_over_repeats(const double* per, int32_t n_repeat, double* mean6, double* sem6) {
    /* src/observables_per_repeat.f90:29-69 */
    double r1 = 1.0 / (n_repeat * (n_repeat - 1.0));
    for (int c = 0; c < 6; ++c) {
        double s = 0.0;
        for (int r = 0; r < n_repeat; ++r) s += per[6 * r + c];
        double mean = s / n_repeat, e = 0.0;
        for (int r = 0; r < n_repeat; ++r) {
            double d = per[6 * r + c] - mean;
            e += d * d;
        }
        mean6[c] = mean;
        sem6[c] = sqrt(r1 * e);
    }
}

/* src/get_spiral_state.f90:47-82 */
int64_t o_spiral_partners(int32_t n, const double* pos, const double* dir, double dis,
                          double latency_deg, int32_t cap, int32_t* partners) {
    const double to_angle = 180.0 / (4.0 * atan(1.0));   /* src/init.f90:44,55 */
    int64_t count = 0;
    for (int32_t i = 0; i < n; ++i) {
        int32_t found = 0;
        for (int32_t k = 0; k < cap; ++k) partners[(int64_t)i * cap + k] = -1;
        for (int32_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double r[3], d2 = 0.0;
            for (int a = 0; a < 3; ++a) {
                r[a] = pos[3 * j + a] - pos[3 * i + a];
                d2 += r[a] * r[a];
            }
            double d = sqrt(d2);
            if (!(d <= dis)) continue;
            double c = 0.0;
            for (int a = 0; a < 3; ++a) c += (r[a] / d) * dir[a];
            if (!(acos(c) * to_angle <= latency_deg)) continue;
            if (found == cap) return -1;
            partners[(int64_t)i * cap + found++] = j;
            ++count;
        }
    }
    return count;
}

/* src/get_spiral_state.f90:84-120 */
double o_spiral_psi(int32_t n, const double* S, const int32_t* partners, int32_t cap,
                    int64_t spiral_count) {
    double ss[3] = {0.0, 0.0, 0.0};
    for (int32_t i = 0; i < n; ++i) {
        const double* a = S + 3 * i;
        for (int32_t k = 0; k < cap; ++k) {
            int32_t j = partners[(int64_t)i * cap + k];
            if (j < 0) break;
            const double* b = S + 3 * j;
            ss[0] += a[1] * b[2] - a[2] * b[1];
            ss[1] += a[2] * b[0] - a[0] * b[2];
            ss[2] += a[0] * b[1] - a[1] * b[0];
        }
    }
    return sqrt(ss[0] * ss[0] + ss[1] * ss[1] + ss[2] * ss[2]) / (double)spiral_count;
}

/* src/get_spiral_state.f90:122-140 */
void o_spiral_observables(const double* sums, double n_samples, double beta, int64_t spiral_count,
                          double* out3) {
    double p = sums[0] / n_samples, p2 = sums[1] / n_samples, p4 = sums[2] / n_samples;
    out3[0] = p;
    out3[1] = 1.5 - 0.5 * (p4 / (p2 * p2));
    out3[2] = beta * (p2 - p * p) * (double)spiral_count;
}

/* src/parallel_tempering.f90:77-89 */
int o_pt_accept(double beta, double beta_p, double E, double E_p, double u) {
    double delta = -(beta_p - beta) * (E - E_p);
    if (delta > 0.0) return 1;
    return u < exp(delta);
}

/* ------------------------------------------------------------------------ */
/* MLP 1-64-64-2 (no reference counterpart)                                  */
/* ------------------------------------------------------------------------ */
#define H 64
static void mlp_layers(const double* p, const double** W1, const double** b1,
                       const double** W2, const double** b2, const double** W3,
                       const double** b3) {
    *W1 = p; *b1 = p + H; *W2 = p + 2 * H; *b2 = *W2 + H * H; *W3 = *b2 + H; *b3 = *W3 + 2 * H;
}

void o_mlp_forward(const double* p, int32_t n, const double* x, double* y) {
    const double *W1, *b1, *W2, *b2, *W3, *b3;
    mlp_layers(p, &W1, &b1, &W2, &b2, &W3, &b3);
    for (int32_t s = 0; s < n; ++s) {
        double h1[H], h2[H];
        for (int i = 0; i < H; ++i) h1[i] = tanh(W1[i] * x[s] + b1[i]);
        for (int o = 0; o < H; ++o) {
            double a = b2[o];
            for (int i = 0; i < H; ++i) a += W2[o * H + i] * h1[i];
            h2[o] = tanh(a);
        }
        for (int o = 0; o < 2; ++o) {
            double a = b3[o];
            for (int i = 0; i < H; ++i) a += W3[o * H + i] * h2[i];
            y[2 * s + o] = a;
        }
    }
}

double o_mlp_loss_grad(const double* p, int32_t n, const double* x, const double* y,
                       const double* w, double* grad) {
    const double *W1, *b1, *W2, *b2, *W3, *b3;
    mlp_layers(p, &W1, &b1, &W2, &b2, &W3, &b3);
    double *gW1 = grad, *gb1 = grad + H, *gW2 = grad + 2 * H, *gb2 = gW2 + H * H,
           *gW3 = gb2 + H, *gb3 = gW3 + 2 * H;
    memset(grad, 0, sizeof(double) * O_MLP_NPARAM);
    double loss = 0.0;
    for (int32_t s = 0; s < n; ++s) {
        double h1[H], h2[H], out[2], d2[H], d1[H];
        for (int i = 0; i < H; ++i) h1[i] = tanh(W1[i] * x[s] + b1[i]);
        for (int o = 0; o < H; ++o) {
            double a = b2[o];
            for (int i = 0; i < H; ++i) a += W2[o * H + i] * h1[i];
            h2[o] = tanh(a);
        }
        double dout[2];
        for (int o = 0; o < 2; ++o) {
            double a = b3[o];
            for (int i = 0; i < H; ++i) a += W3[o * H + i] * h2[i];
            out[o] = a;
            double r = out[o] - y[2 * s + o];
            loss += w[2 * s + o] * r * r;
            dout[o] = 2.0 * w[2 * s + o] * r / n;
        }
        for (int o = 0; o < 2; ++o) {
            gb3[o] += dout[o];
            for (int i = 0; i < H; ++i) gW3[o * H + i] += dout[o] * h2[i];
        }
        for (int i = 0; i < H; ++i) {
            double g = dout[0] * W3[i] + dout[1] * W3[H + i];
            d2[i] = g * (1.0 - h2[i] * h2[i]);
        }
        for (int o = 0; o < H; ++o) {
            gb2[o] += d2[o];
            for (int i = 0; i < H; ++i) gW2[o * H + i] += d2[o] * h1[i];
        }
        for (int i = 0; i < H; ++i) {
            double g = 0.0;
            for (int o = 0; o < H; ++o) g += d2[o] * W2[o * H + i];
            d1[i] = g * (1.0 - h1[i] * h1[i]);
            gb1[i] += d1[i];
            gW1[i] += d1[i] * x[s];
        }
    }
    return loss / n;
}
#undef H

/* ------------------------------------------------------------------------ */
/* S(q): utility/Structure_Factor/structure_factor.f90:150-177              */
/* ------------------------------------------------------------------------ */
/* S(q) with the per-component moduli of the reference's output line
 * (utility/Structure_Factor/structure_factor.f90:152-177): out4 = {sum_a
 * |S_a(q)|^2 / N^2, |S_x(q)|/N, |S_y(q)|/N, |S_z(q)|/N}. */
void o_structure_factor_parts(int32_t n, const double* pos, const double* S, const double* q3, double* out4) {
    double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
    for (int32_t i = 0; i < n; ++i) {
        double ph = q3[0] * pos[3 * i] + q3[1] * pos[3 * i + 1] + q3[2] * pos[3 * i + 2];
        double c = cos(ph), s = sin(ph);
        for (int a = 0; a < 3; ++a) {
            re[a] += S[3 * i + a] * c;
            im[a] += S[3 * i + a] * s;
        }
    }
    out4[0] = 0.0;
    for (int a = 0; a < 3; ++a) {
        double m2 = re[a] * re[a] + im[a] * im[a];
        out4[0] += m2;
        out4[1 + a] = sqrt(m2) / (double)n;
    }
    out4[0] /= (double)n * (double)n;
}

double o_structure_factor(int32_t n, const double* pos, const double* S, const double* q3) {
    double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
    for (int32_t i = 0; i < n; ++i) {
        double ph = q3[0] * pos[3 * i] + q3[1] * pos[3 * i + 1] + q3[2] * pos[3 * i + 2];
        double c = cos(ph), s = sin(ph);
        for (int a = 0; a < 3; ++a) {
            re[a] += S[3 * i + a] * c;
            im[a] += S[3 * i + a] * s;
        }
    }
    double t = 0.0;
    for (int a = 0; a < 3; ++a) t += re[a] * re[a] + im[a] * im[a];
    return t / ((double)n * (double)n);
}
