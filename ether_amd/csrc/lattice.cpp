// lattice.cpp - host-side neighbour construction for the sweep engine.
//
// Same capture rule as the reference's getting_nbd/search_nbd
// (src/getting_nbd.f90:46-82, src/search_nbd.f90:34-58): site j != i is a
// shell-s neighbour of i when, for some image shift in {0,+1,-1} along each
// closed axis (only 0 along open axes), | d_s - |r_i - (r_j + shift.L.A)| |
// <= tol; j is stored once per shell and the list is in ascending j.
//
// Instead of the reference's O(N^2 * 27) scan this uses translation symmetry:
// the displacement r_j + shift.L.A - r_i depends only on the cell offset
// D = (cell_j + shift.L) - cell_i and the two basis sites, so a per-basis
// template {(D, l')} is built once per shell and applied to every cell; the
// shift is recovered as floor((cell_i + D)/L) and the entry is kept only when
// that shift is one the reference tries.  Cost O(N * template).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ether_gpu.h"
#include "lattice_internal.h"
#include "philox.h"

namespace ether {
thread_local std::string g_last_error;
void set_global_error(const std::string& s) { g_last_error = s; }
const std::string& global_error() { return g_last_error; }
}  // namespace ether

namespace {

using ether::TemplateEntry;

bool invert3(const double* a, double* inv) {
    // a is row-major 3x3; inv = a^{-1}
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                 a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (!(std::fabs(det) > 1e-300)) return false;
    double id = 1.0 / det;
    inv[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    inv[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    inv[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    inv[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    inv[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    inv[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    inv[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    inv[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    inv[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return true;
}

int floordiv(int a, int b) {
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

}  // namespace

namespace ether {

int build_templates(const double* abc, int32_t n_basis, const double* basis_xyz, const int32_t* sc,
                    const int32_t* open_bc, int32_t n_shells, const double* shell_d, double tol,
                    LatticeTemplates& out, std::string& err) {
    if (!abc || !basis_xyz || !sc || !open_bc || !shell_d) {
        err = "null pointer argument";
        return ETHER_ERR_ARG;
    }
    if (n_basis <= 0 || n_shells <= 0 || sc[0] <= 0 || sc[1] <= 0 || sc[2] <= 0 || !(tol >= 0.0)) {
        err = "invalid sizes";
        return ETHER_ERR_ARG;
    }
    if ((long long)sc[0] * sc[1] * sc[2] * n_basis > (1LL << 24)) {
        err = "more than 2^24 sites";
        return ETHER_ERR_LIMIT;
    }
    double inv[9];
    if (!invert3(abc, inv)) {
        err = "singular lattice vectors";
        return ETHER_ERR_ARG;
    }
    // |column k of A^{-1}| bounds |D_k| for a displacement of length r: D = p.A^{-1}
    double colnorm[3];
    for (int k = 0; k < 3; ++k)
        colnorm[k] = std::sqrt(inv[0 * 3 + k] * inv[0 * 3 + k] + inv[1 * 3 + k] * inv[1 * 3 + k] +
                               inv[2 * 3 + k] * inv[2 * 3 + k]);
    double maxdiff = 0.0;
    for (int a = 0; a < n_basis; ++a)
        for (int b = 0; b < n_basis; ++b) {
            double r2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                double d = basis_xyz[3 * a + x] - basis_xyz[3 * b + x];
                r2 += d * d;
            }
            maxdiff = std::max(maxdiff, std::sqrt(r2));
        }
    out.n_shells = n_shells;
    out.n_basis = n_basis;
    out.t.assign((size_t)n_shells * n_basis, {});
    for (int sh = 0; sh < n_shells; ++sh) {
        const double d = shell_d[sh];
        int lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            int b = (int)std::ceil((d + tol + maxdiff) * colnorm[k]) + 1;
            int L = sc[k];
            // shifts the reference tries: {0,+-1} closed, {0} open
            int lim_lo = open_bc[k] ? -(L - 1) : -(2 * L - 1);
            int lim_hi = open_bc[k] ? (L - 1) : (2 * L - 1);
            lo[k] = std::max(-b, lim_lo);
            hi[k] = std::min(b, lim_hi);
        }
        for (int la = 0; la < n_basis; ++la)
            for (int lb = 0; lb < n_basis; ++lb)
                for (int d0 = lo[0]; d0 <= hi[0]; ++d0)
                    for (int d1 = lo[1]; d1 <= hi[1]; ++d1)
                        for (int d2 = lo[2]; d2 <= hi[2]; ++d2) {
                            double r2 = 0.0;
                            for (int x = 0; x < 3; ++x) {
                                double img = abc[0 * 3 + x] * d0 + abc[1 * 3 + x] * d1 +
                                             abc[2 * 3 + x] * d2 + basis_xyz[3 * lb + x];
                                double dr = basis_xyz[3 * la + x] - img;
                                r2 += dr * dr;
                            }
                            if (std::fabs(d - std::sqrt(r2)) <= tol)
                                out.t[(size_t)sh * n_basis + la].push_back({{d0, d1, d2}, lb});
                        }
    }
    return ETHER_OK;
}

bool periodic_colouring(const LatticeTemplates& tp, const int32_t* sc, const int32_t* open_bc,
                        PeriodicColouring& out) {
    static const int kPeriods[] = {1, 2, 3, 4, 6};
    const int nb = tp.n_basis;
    bool found = false;
    out = PeriodicColouring();
    for (int i0 : kPeriods)
        for (int i1 : kPeriods)
            for (int i2 : kPeriods) {
                const int m[3] = {i0, i1, i2};
                bool ok = true;
                for (int k = 0; k < 3 && ok; ++k)
                    if (!open_bc[k] && sc[k] % m[k] != 0) ok = false;
                if (!ok) continue;
                const int cells = m[0] * m[1] * m[2];
                const int Q = cells * nb;
                if (Q > 1024) continue;
                if (found && (Q >= (int)out.qcol.size() && out.n_colours <= 2)) continue;
                // the quotient graph: vertex (la, residue r) joins (lb, r + d mod m)
                std::vector<std::vector<int>> adj(Q);
                for (int sh = 0; sh < tp.n_shells && ok; ++sh)
                    for (int la = 0; la < nb && ok; ++la)
                        for (const TemplateEntry& e : tp.t[(size_t)sh * nb + la]) {
                            // an entry that lands on the site itself is no bond (j != i)
                            bool self = e.lb == la;
                            for (int k = 0; k < 3 && self; ++k)
                                if (open_bc[k] ? e.d[k] != 0 : e.d[k] % sc[k] != 0) self = false;
                            if (self) continue;
                            for (int r = 0; r < cells && ok; ++r) {
                                const int r0 = r % m[0], r1 = (r / m[0]) % m[1], r2 = r / (m[0] * m[1]);
                                auto md = [](int x, int p) { return ((x % p) + p) % p; };
                                const int w0 = md(r0 + e.d[0], m[0]), w1 = md(r1 + e.d[1], m[1]),
                                          w2 = md(r2 + e.d[2], m[2]);
                                const int v = la + nb * r, w = e.lb + nb * (w0 + m[0] * (w1 + m[1] * w2));
                                if (v == w) {
                                    ok = false;   // a bond within one quotient vertex
                                    break;
                                }
                                adj[v].push_back(w);
                                adj[w].push_back(v);
                            }
                        }
                if (!ok) continue;
                // greedy in two orders (basis innermost, basis outermost) and
                // DSATUR (Brelaz 1979: next the vertex seeing the most colours,
                // ties to the most neighbours, then the lowest index), the best kept
                std::vector<int> best;
                int best_n = 1 << 30;
                {
                    std::vector<int> col(Q, -1), deg(Q);
                    std::vector<std::vector<char>> seen(Q);
                    std::vector<int> sat(Q, 0);
                    for (int v = 0; v < Q; ++v) {
                        std::sort(adj[v].begin(), adj[v].end());
                        adj[v].erase(std::unique(adj[v].begin(), adj[v].end()), adj[v].end());
                        deg[v] = (int)adj[v].size();
                    }
                    int n_col = 0;
                    for (int step = 0; step < Q; ++step) {
                        int v = -1;
                        for (int u = 0; u < Q; ++u) {
                            if (col[u] >= 0) continue;
                            if (v < 0 || sat[u] > sat[v] || (sat[u] == sat[v] && deg[u] > deg[v])) v = u;
                        }
                        std::vector<char> used(n_col + 1, 0);
                        for (int w : adj[v])
                            if (col[w] >= 0) used[col[w]] = 1;
                        int c = 0;
                        while (used[c]) ++c;
                        col[v] = c;
                        n_col = std::max(n_col, c + 1);
                        for (int w : adj[v]) {
                            if ((int)seen[w].size() <= c) seen[w].resize(c + 1, 0);
                            if (!seen[w][c]) {
                                seen[w][c] = 1;
                                ++sat[w];
                            }
                        }
                    }
                    best_n = n_col;
                    best.swap(col);
                }
                for (int ord = 0; ord < 2; ++ord) {
                    std::vector<int> col(Q, -1);
                    int n_col = 0;
                    for (int q = 0; q < Q; ++q) {
                        const int v = ord == 0 ? q : (q % cells) * nb + q / cells;
                        std::vector<char> used(n_col + 1, 0);
                        for (int w : adj[v])
                            if (col[w] >= 0 && col[w] <= n_col) used[col[w]] = 1;
                        int c = 0;
                        while (used[c]) ++c;
                        col[v] = c;
                        n_col = std::max(n_col, c + 1);
                    }
                    if (n_col < best_n) {
                        best_n = n_col;
                        best.swap(col);
                    }
                }
                if (!found || best_n < out.n_colours || (best_n == out.n_colours && Q < (int)out.qcol.size())) {
                    found = true;
                    out.n_colours = best_n;
                    out.qcol.swap(best);
                    for (int k = 0; k < 3; ++k) out.m[k] = m[k];
                }
            }
    return found;
}

}  // namespace ether

extern "C" int ether_build_neighbours(const double* abc, int32_t n_basis, const double* basis_xyz,
                                      const int32_t* sc, const int32_t* open_bc,
                                      int32_t n_shells, const double* shell_d, double tol,
                                      int32_t capacity, int32_t* nn_count, int32_t* nn_idx) {
    using ether::set_global_error;
    if (!nn_count || !nn_idx || capacity <= 0) {
        set_global_error("ether_build_neighbours: null output or capacity <= 0");
        return ETHER_ERR_ARG;
    }
    ether::LatticeTemplates tp;
    std::string err;
    const int rc = ether::build_templates(abc, n_basis, basis_xyz, sc, open_bc, n_shells, shell_d, tol, tp, err);
    if (rc != ETHER_OK) {
        set_global_error("ether_build_neighbours: " + err);
        return rc;
    }
    const int n = sc[0] * sc[1] * sc[2] * n_basis;
    std::vector<int> nbr;
    nbr.reserve(64);
    for (int sh = 0; sh < n_shells; ++sh) {
        for (int ck = 0; ck < sc[2]; ++ck)
            for (int cj = 0; cj < sc[1]; ++cj)
                for (int ci = 0; ci < sc[0]; ++ci)
                    for (int la = 0; la < n_basis; ++la) {
                        const int c[3] = {ci, cj, ck};
                        const int i = ((ck * sc[1] + cj) * sc[0] + ci) * n_basis + la;
                        nbr.clear();
                        for (const TemplateEntry& t : tp.t[(size_t)sh * n_basis + la]) {
                            int cell[3];
                            bool ok = true;
                            for (int k = 0; k < 3 && ok; ++k) {
                                int u = c[k] + t.d[k];
                                int s = floordiv(u, sc[k]);
                                if (open_bc[k] ? (s != 0) : (s < -1 || s > 1)) ok = false;
                                cell[k] = u - s * sc[k];
                            }
                            if (!ok) continue;
                            int j = ((cell[2] * sc[1] + cell[1]) * sc[0] + cell[0]) * n_basis + t.lb;
                            if (j == i) continue;
                            nbr.push_back(j);
                        }
                        std::sort(nbr.begin(), nbr.end());
                        nbr.erase(std::unique(nbr.begin(), nbr.end()), nbr.end());
                        if ((int)nbr.size() > capacity) {
                            char buf[160];
                            std::snprintf(buf, sizeof buf,
                                          "ether_build_neighbours: site %d shell %d has %d "
                                          "neighbours > nbd_capacity %d",
                                          i + 1, sh + 1, (int)nbr.size(), capacity);
                            set_global_error(buf);
                            return ETHER_ERR_LIMIT;
                        }
                        nn_count[(size_t)sh * n + i] = (int)nbr.size();
                        int32_t* out = nn_idx + ((size_t)sh * n + i) * capacity;
                        for (size_t q = 0; q < nbr.size(); ++q) out[q] = nbr[q];
                        for (int q = (int)nbr.size(); q < capacity; ++q) out[q] = -1;
                    }
    }
    return ETHER_OK;
}

extern "C" int ether_colour_lattice(const double* abc, int32_t n_basis, const double* basis_xyz,
                                    const int32_t* sc, const int32_t* open_bc, int32_t n_shells,
                                    const double* shell_d, double tol, int32_t* colour, int32_t* n_colours) {
    using ether::set_global_error;
    if (!colour || !n_colours) {
        set_global_error("ether_colour_lattice: null output");
        return ETHER_ERR_ARG;
    }
    ether::LatticeTemplates tp;
    std::string err;
    const int rc = ether::build_templates(abc, n_basis, basis_xyz, sc, open_bc, n_shells, shell_d, tol, tp, err);
    if (rc != ETHER_OK) {
        set_global_error("ether_colour_lattice: " + err);
        return rc;
    }
    ether::PeriodicColouring pc;
    *n_colours = 0;
    if (!ether::periodic_colouring(tp, sc, open_bc, pc)) return ETHER_OK;
    const int nb = n_basis;
    for (int ck = 0; ck < sc[2]; ++ck)
        for (int cj = 0; cj < sc[1]; ++cj)
            for (int ci = 0; ci < sc[0]; ++ci)
                for (int la = 0; la < nb; ++la) {
                    const int i = ((ck * sc[1] + cj) * sc[0] + ci) * nb + la;
                    colour[i] = pc.qcol[la + nb * (ci % pc.m[0] + pc.m[0] * (cj % pc.m[1] + pc.m[1] * (ck % pc.m[2])))];
                }
    *n_colours = pc.n_colours;
    return ETHER_OK;
}

// ---------------------------------------------------------------------------
// Spiral-state partners (src/get_spiral_state.f90:47-82, case
// 'ions_for_spiral'): site j != i is a partner of i when |r_j - r_i| <= dis
// and the angle between (r_j - r_i)/|r_j - r_i| and dir, acos(u.dir) in
// degrees, is <= latency.  Positions are the supercell's Cartesian positions
// without periodic images, dir is used as given, and partners are listed in
// ascending j.  A uniform grid of cells no smaller than `dis` replaces the
// O(N^2) scan.
// ---------------------------------------------------------------------------
extern "C" int ether_build_spiral_partners(int32_t n_sites, const double* pos, const double* dir,
                                           double dis, double latency_deg, int32_t capacity,
                                           int32_t* partners, int64_t* count) {
    using ether::set_global_error;
    if (n_sites <= 0 || !pos || !dir || !partners || capacity <= 0 || !(dis > 0.0)) {
        set_global_error("ether_build_spiral_partners: bad arguments");
        return ETHER_ERR_ARG;
    }
    const int N = n_sites;
    double lo[3], hi[3];
    for (int x = 0; x < 3; ++x) lo[x] = hi[x] = pos[x];
    for (int i = 0; i < N; ++i)
        for (int x = 0; x < 3; ++x) {
            lo[x] = std::min(lo[x], pos[3 * i + x]);
            hi[x] = std::max(hi[x], pos[3 * i + x]);
        }
    // cell edge h >= dis, so every partner lies in the 27 cells around a site;
    // h grows until the grid holds at most about 4 cells per site
    double h = dis;
    int nb[3];
    for (;;) {
        double cells = 1.0;
        for (int x = 0; x < 3; ++x) {
            nb[x] = (int)std::min(1048576.0, std::floor((hi[x] - lo[x]) / h) + 1.0);
            cells *= nb[x];
        }
        if (cells <= 4.0 * N + 64.0) break;
        h *= 1.25;
    }
    auto bin_of = [&](const double* p, int* b) {
        for (int x = 0; x < 3; ++x) b[x] = std::min(nb[x] - 1, std::max(0, (int)std::floor((p[x] - lo[x]) / h)));
    };
    const size_t nbins = (size_t)nb[0] * nb[1] * nb[2];
    std::vector<int> head(nbins + 1, 0), items(N);
    std::vector<int> bin(N);
    for (int i = 0; i < N; ++i) {
        int b[3];
        bin_of(pos + 3 * i, b);
        bin[i] = (b[2] * nb[1] + b[1]) * nb[0] + b[0];
        head[bin[i] + 1]++;
    }
    for (size_t q = 0; q < nbins; ++q) head[q + 1] += head[q];
    {
        std::vector<int> fill(head.begin(), head.end() - 1);
        for (int i = 0; i < N; ++i) items[fill[bin[i]]++] = i;   // ascending i within a bin
    }
    const double to_deg = 180.0 / M_PI;
    int64_t total = 0;
    std::vector<int> cand;
    for (int i = 0; i < N; ++i) {
        int b[3];
        bin_of(pos + 3 * i, b);
        cand.clear();
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int x = b[0] + dx, y = b[1] + dy, z = b[2] + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= nb[0] || y >= nb[1] || z >= nb[2]) continue;
                    const size_t q = ((size_t)z * nb[1] + y) * nb[0] + x;
                    for (int t = head[q]; t < head[q + 1]; ++t) cand.push_back(items[t]);
                }
        std::sort(cand.begin(), cand.end());
        int found = 0;
        int32_t* out = partners + (size_t)i * capacity;
        for (int k = 0; k < capacity; ++k) out[k] = -1;
        for (int j : cand) {
            if (j == i) continue;
            const double r[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                                 pos[3 * j + 2] - pos[3 * i + 2]};
            const double d = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (!(d <= dis)) continue;
            const double u[3] = {r[0] / d, r[1] / d, r[2] / d};   // as the reference: uni_vec = r21/dis
            const double ang = std::acos(u[0] * dir[0] + u[1] * dir[1] + u[2] * dir[2]) * to_deg;
            if (!(ang <= latency_deg)) continue;
            if (found == capacity) {
                set_global_error("ether_build_spiral_partners: more partners than spiral_capacity for site " +
                                 std::to_string(i + 1));
                return ETHER_ERR_LIMIT;
            }
            out[found++] = j;
            ++total;
        }
    }
    if (count) *count = total;
    return ETHER_OK;
}

// ---------------------------------------------------------------------------
// start state of the basis sites (src/generate_supercell.f90:34-63)
// ---------------------------------------------------------------------------
extern "C" int ether_start_state(uint64_t seed, int32_t model, int32_t angle, int32_t n_basis,
                                 const int32_t* file_index, double* S, double* phi) {
    if (n_basis <= 0 || !file_index || !S || !phi) {
        ether::set_global_error("ether_start_state: bad argument");
        return ETHER_ERR_ARG;
    }
    constexpr uint32_t kRngStart = 5;
    for (int32_t l = 0; l < n_basis; ++l) {
        double* s = S + 3 * (size_t)l;
        double p = 0.0;
        if (model == ETHER_MODEL_ISING) {
            // atom(i,3) = (-1)**i
            s[0] = 0.0;
            s[1] = 0.0;
            s[2] = (file_index[l] % 2 == 0) ? 1.0 : -1.0;
        } else if (!angle) {
            // George_Marsaglia: (v1, v2) uniform in the unit disc, by rejection
            // on two 24-bit candidate pairs per Philox block
            bool done = false;
            for (uint32_t att = 0; !done; ++att) {
                const ether::u32x4 w = ether::rng_draw(seed, (uint32_t)l, 0, 0, kRngStart, att);
                const uint32_t cx[2] = {w.x >> 8, ((w.w >> 16) << 8) | (w.x & 0xffu)};
                const uint32_t cy[2] = {w.y >> 8, ((w.w & 0xffffu) << 8) | (w.y & 0xffu)};
                for (int c = 0; c < 2 && !done; ++c) {
                    const int64_t a1 = (int64_t)cx[c] - (1 << 23), a2 = (int64_t)cy[c] - (1 << 23);
                    const int64_t q = a1 * a1 + a2 * a2;
                    if (q >= ((int64_t)1 << 46)) continue;
                    const double v1 = a1 * 0x1p-23, v2 = a2 * 0x1p-23, sq = q * 0x1p-46;
                    const double r = std::sqrt(1.0 - sq);
                    s[0] = 2.0 * v1 * r;
                    s[1] = 2.0 * v2 * r;
                    s[2] = 1.0 - 2.0 * sq;
                    done = true;
                }
            }
        } else {
            const ether::u32x4 w = ether::rng_draw(seed, (uint32_t)l, 0, 0, kRngStart, 0);
            const double sz = (double)((int32_t)(w.x >> 8) - (1 << 23)) * 0x1p-23;
            p = 6.283185307179586476925286766559 * (double)(w.y >> 8) * 0x1p-24;
            const double r = std::sqrt(1.0 - sz * sz);
            s[0] = r * std::cos(p);
            s[1] = r * std::sin(p);
            s[2] = sz;
        }
        if (model != ETHER_MODEL_ISING) {
            const double n = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
            s[0] /= n;
            s[1] /= n;
            s[2] /= n;
        }
        phi[l] = p;
    }
    return ETHER_OK;
}
