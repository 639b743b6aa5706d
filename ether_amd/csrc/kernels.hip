// kernels.hip - gfx950 kernels of the Monte Carlo sweep engine.
//
//   xyz_sweep     one colour phase of graph-coloured single-spin Metropolis for
//                 XYZ spins (src/Metropolis.f90:35-62 with the proposals of
//                 src/update_spin_details.f90:28-51 and the energy change of
//                 src/Hamiltonian.f90:20-143)
//   ising_sweep   the same for Ising spins, 4 sites per thread, with exact
//                 integer acceptance thresholds
//   xyz_ovrr      one colour phase of overrelaxation (src/overrelaxation.f90:30-49)
//   measure       per-replica energy and magnetisation partial sums
//                 (src/get_tot_energy.f90:35-45, src/get_tot_magnetisation.f90:35-38)
//   finalize      fixed-order reduction of the partials plus the observable()
//                 moment update (src/observable.f90:30-43, src/get_eng.f90:36-39,
//                 src/get_mag.f90:36-40)
//   pt_swap       replica exchange as a slot permutation (src/parallel_tempering.f90:50-108)
//   init          fresh_spins / random start (src/fresh_spins.f90:31-38)
//   local_fields  exchange field and site energy of one replica (for parity checks)
//
// All of these are HBM/L2-bound gathers: none is GEMM-shaped.  Each thread owns
// one site (four for Ising) of one replica; a colour phase never writes a
// site that another lane of the same phase reads.
#include <hip/hip_runtime.h>

#include "engine_common.h"
#include "philox.h"

namespace ether {

// ---------------------------------------------------------------------------
// random-number helpers (24-bit uniforms, identical to the oracle's)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sym24f(uint32_t x24) {
    return (float)((int32_t)x24 - (1 << 23)) * 0x1p-23f;  // exact in FP32
}
__device__ __forceinline__ float eta24f(uint32_t w) {
    return (float)((w >> 8) + 1u) * 0x1p-24f;  // (0,1], exact
}
__device__ __forceinline__ float wrap_pif(float p) {
    const float two_pi = 6.28318530717958647692f;
    return p - two_pi * floorf((p + 3.14159265358979323846f) / two_pi);
}

__device__ __forceinline__ void normalise3(float& x, float& y, float& z) {
    float inv = 1.0f / sqrtf(x * x + y * y + z * z);
    x *= inv;
    y *= inv;
    z *= inv;
}

// George_Marsaglia (src/George_Marsaglia.f90:29-39): two candidate pairs per
// Philox block, exact integer test a1^2 + a2^2 < 2^46 (i.e. v1^2+v2^2 < 1).
__device__ __forceinline__ void marsaglia(u32x4 w, uint64_t seed, uint32_t id, uint32_t r,
                                          uint32_t step, uint32_t purpose, float& x, float& y,
                                          float& z) {
    for (uint32_t att = 0;; ++att) {
        if (att) w = rng_draw(seed, id, r, step, purpose, att);
        uint32_t cx[2] = {w.x >> 8, ((w.w >> 16) << 8) | (w.x & 0xffu)};
        uint32_t cy[2] = {w.y >> 8, ((w.w & 0xffffu) << 8) | (w.y & 0xffu)};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            int32_t a1 = (int32_t)cx[c] - (1 << 23), a2 = (int32_t)cy[c] - (1 << 23);
            int64_t q = (int64_t)a1 * a1 + (int64_t)a2 * a2;
            if (q < (int64_t(1) << 46)) {
                float v1 = (float)a1 * 0x1p-23f, v2 = (float)a2 * 0x1p-23f;
                float s = (float)q * 0x1p-46f;
                float rt = sqrtf(1.0f - s);
                x = 2.0f * v1 * rt;
                y = 2.0f * v2 * rt;
                z = 1.0f - 2.0f * s;
                normalise3(x, y, z);
                return;
            }
        }
    }
}

// ANGLE move (src/update_spin_details.f90:39-46)
__device__ __forceinline__ void angle_move(u32x4 w, float dphi, float phi_old, float& x, float& y,
                                           float& z, float& phi) {
    float sz = sym24f(w.x >> 8);
    phi = wrap_pif(phi_old + dphi * sym24f(w.y >> 8));
    float rt = sqrtf(fmaxf(0.0f, 1.0f - sz * sz));
    float sn, cs;
    sincosf(phi, &sn, &cs);
    x = rt * cs;
    y = rt * sn;
    z = sz;
    normalise3(x, y, z);
}

// ---------------------------------------------------------------------------
// XYZ Metropolis, one colour phase
// ---------------------------------------------------------------------------
template <bool ANGLE, bool ZEE, bool SIA>
__global__ __launch_bounds__(kBlock) void k_xyz_sweep(XyzArgs a) {
    __shared__ float sJ[3 * (kMaxClasses + 1)];
    __shared__ float sZ[3 * kMaxSpecies];
    __shared__ float sD[3 * kMaxSpecies];
    for (int t = threadIdx.x; t < 3 * (a.n_classes + 1); t += kBlock) sJ[t] = a.jclass[t];
    if (ZEE)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sZ[t] = a.zvec[t];
    if (SIA)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sD[t] = a.dsia[t];
    __syncthreads();

    int r, tile;
    if (!map_block(a.R, a.tiles, a.xcd_map, r, tile)) return;
    const int slot = a.c_begin + tile * kBlock + threadIdx.x;
    const bool active = slot < a.c_end;
    int accepted = 0;
    if (active) {
        const size_t base = (size_t)r * a.N;
        const float* __restrict__ px = a.sx + base;
        const float* __restrict__ py = a.sy + base;
        const float* __restrict__ pz = a.sz + base;
        float x = px[slot], y = py[slot], z = pz[slot];
        float hx = 0.f, hy = 0.f, hz = 0.f;
        for (int k = 0; k < a.D; ++k) {
            uint32_t e = a.ell[(size_t)k * a.ell_stride + slot];
            uint32_t j = e & 0xffffffu;
            uint32_t c = e >> 24;
            hx = fmaf(sJ[3 * c + 0], px[j], hx);
            hy = fmaf(sJ[3 * c + 1], py[j], hy);
            hz = fmaf(sJ[3 * c + 2], pz[j], hz);
        }
        int sp = (ZEE || SIA) ? a.sp[slot] : 0;
        if (ZEE) {
            hx += sZ[3 * sp + 0];
            hy += sZ[3 * sp + 1];
            hz += sZ[3 * sp + 2];
        }
        const float beta = a.beta_slot[a.slot_of[r]];
        u32x4 w = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, a.step, RNG_METROPOLIS, 0);
        const float eta = eta24f(w.z);
        float nx, ny, nz, nphi = 0.f;
        if (ANGLE) {
            angle_move(w, a.dphi, a.phi[base + slot], nx, ny, nz, nphi);
        } else {
            marsaglia(w, a.seed, (uint32_t)slot, (uint32_t)r, a.step, RNG_METROPOLIS, nx, ny, nz);
        }
        float dE = (nx - x) * hx + (ny - y) * hy + (nz - z) * hz;
        if (SIA) {
            dE += sD[3 * sp + 0] * (nx * nx - x * x) + sD[3 * sp + 1] * (ny * ny - y * y) +
                  sD[3 * sp + 2] * (nz * nz - z * z);
        }
        // accept if -beta*dE > log(eta)   (src/Metropolis.f90:52)
        if (-beta * dE > __logf(eta)) {
            a.sx[base + slot] = nx;
            a.sy[base + slot] = ny;
            a.sz[base + slot] = nz;
            if (ANGLE) a.phi[base + slot] = nphi;
            accepted = 1;
        }
    }
    // acceptance is booked to the temperature slot, as the reference books it per rank
    int cnt = __syncthreads_count(accepted);
    if (threadIdx.x == 0 && cnt) atomicAdd(a.acc + a.slot_of[r], (unsigned long long)cnt);
}

// ---------------------------------------------------------------------------
// Overrelaxation, one colour phase: S' = 2(A.S)/(A.A) A - S, renormalised.
// ---------------------------------------------------------------------------
template <bool ANGLE, bool ZEE>
__global__ __launch_bounds__(kBlock) void k_xyz_ovrr(XyzArgs a) {
    __shared__ float sJ[3 * (kMaxClasses + 1)];
    __shared__ float sZ[3 * kMaxSpecies];
    for (int t = threadIdx.x; t < 3 * (a.n_classes + 1); t += kBlock) sJ[t] = a.jclass[t];
    if (ZEE)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sZ[t] = a.zvec[t];
    __syncthreads();
    int r, tile;
    if (!map_block(a.R, a.tiles, a.xcd_map, r, tile)) return;
    const int slot = a.c_begin + tile * kBlock + threadIdx.x;
    if (slot >= a.c_end) return;
    if (a.frac < 1.0f) {
        u32x4 w = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, a.step, RNG_OVRR, a.pass);
        if (!((float)(w.x >> 8) * 0x1p-24f < a.frac)) return;
    }
    const size_t base = (size_t)r * a.N;
    const float* __restrict__ px = a.sx + base;
    const float* __restrict__ py = a.sy + base;
    const float* __restrict__ pz = a.sz + base;
    float hx = 0.f, hy = 0.f, hz = 0.f;
    for (int k = 0; k < a.D; ++k) {
        uint32_t e = a.ell[(size_t)k * a.ell_stride + slot];
        uint32_t j = e & 0xffffffu;
        uint32_t c = e >> 24;
        hx = fmaf(sJ[3 * c + 0], px[j], hx);
        hy = fmaf(sJ[3 * c + 1], py[j], hy);
        hz = fmaf(sJ[3 * c + 2], pz[j], hz);
    }
    if (ZEE) {
        int sp = a.sp[slot];
        hx += sZ[3 * sp + 0];
        hy += sZ[3 * sp + 1];
        hz += sZ[3 * sp + 2];
    }
    float aa = hx * hx + hy * hy + hz * hz;
    if (!(aa > 0.f)) return;
    float x = px[slot], y = py[slot], z = pz[slot];
    float f = 2.0f * (hx * x + hy * y + hz * z) / aa;
    float nx = f * hx - x, ny = f * hy - y, nz = f * hz - z;
    normalise3(nx, ny, nz);
    a.sx[base + slot] = nx;
    a.sy[base + slot] = ny;
    a.sz[base + slot] = nz;
    if (ANGLE) a.phi[base + slot] = atan2f(ny, nx);
}

// ---------------------------------------------------------------------------
// Ising Metropolis, one colour phase.  Thread t owns the aligned slot group
// [4g, 4g+4) and takes word (slot & 3) of the Philox block for id = g.
// dE depends only on (species, sigma_i, sum_j stride[class_j] * sigma_j), so
// the acceptance test is an exact integer compare with a host-built threshold.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_ising_sweep(IsingArgs a) {
    __shared__ int sStride[kMaxIsingClasses + 1];
    __shared__ uint32_t sThr[kIsingLdsTable];
    int r, tile;
    const bool live = map_block(a.R, a.tiles, a.xcd_map, r, tile);
    const bool lds_table = a.table_per_slot <= kIsingLdsTable;
    const uint32_t* thr = nullptr;
    if (live) thr = a.thr + (size_t)a.slot_of[r] * a.table_per_slot;
    for (int t = threadIdx.x; t <= a.n_classes; t += kBlock) sStride[t] = a.stride[t];
    if (live && lds_table)
        for (int t = threadIdx.x; t < a.table_per_slot; t += kBlock) sThr[t] = thr[t];
    __syncthreads();
    if (!live) return;

    const int g = (a.c_begin >> 2) + tile * kBlock + threadIdx.x;
    const int s0 = 4 * g;
    int accepted = 0;
    if (s0 < a.c_end) {
        const size_t base = (size_t)r * a.N;
        int8_t* __restrict__ sp8 = a.s + base;
        u32x4 w = rng_draw(a.seed, (uint32_t)g, (uint32_t)r, a.step, RNG_METROPOLIS, 0);
        const uint32_t words[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int slot = s0 + q;
            if (slot < a.c_begin || slot >= a.c_end) continue;
            int idx = a.offset;
            for (int k = 0; k < a.D; ++k) {
                uint32_t e = a.ell[(size_t)k * a.ell_stride + slot];
                idx += sStride[e >> 24] * (int)sp8[e & 0xffffffu];
            }
            const int sigma = sp8[slot];
            const int spc = a.sp[slot];
            const int t = (spc * 2 + (sigma > 0)) * a.P + idx;
            const uint32_t th = lds_table ? sThr[t] : thr[t];
            if (words[q] < th || th == 0xffffffffu) {
                sp8[slot] = (int8_t)(-sigma);
                ++accepted;
            }
        }
    }
    // block sum of accepted moves
    __shared__ int sAcc;
    if (threadIdx.x == 0) sAcc = 0;
    __syncthreads();
    if (accepted) atomicAdd(&sAcc, accepted);
    __syncthreads();
    if (threadIdx.x == 0 && sAcc) atomicAdd(a.acc + a.slot_of[r], (unsigned long long)sAcc);
}

// ---------------------------------------------------------------------------
// measurement partials
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide fixed-order sum of n values per thread into out[0..n).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out) {
    __shared__ double red[kBlock / 64][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) red[wv][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) s += red[q][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

template <bool ZEE, bool SIA>
__global__ __launch_bounds__(kBlock) void k_xyz_measure(MeasArgs a) {
    __shared__ float sJ[3 * (kMaxClasses + 1)];
    __shared__ float sZ[3 * kMaxSpecies];
    __shared__ float sD[3 * kMaxSpecies];
    for (int t = threadIdx.x; t < 3 * (a.n_classes + 1); t += kBlock) sJ[t] = a.jclass[t];
    if (ZEE)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sZ[t] = a.zvec[t];
    if (SIA)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sD[t] = a.dsia[t];
    __syncthreads();
    const int r = blockIdx.x / a.tiles, tile = blockIdx.x % a.tiles;
    const size_t base = (size_t)r * a.N;
    const float* __restrict__ px = a.sx + base;
    const float* __restrict__ py = a.sy + base;
    const float* __restrict__ pz = a.sz + base;
    double v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < kMeasPerThread; ++q) {
        const int slot = tile * kMeasTile + q * kBlock + threadIdx.x;
        if (slot >= a.N) continue;
        float hx = 0.f, hy = 0.f, hz = 0.f;
        for (int k = 0; k < a.D; ++k) {
            uint32_t e = a.ell[(size_t)k * a.ell_stride + slot];
            uint32_t j = e & 0xffffffu;
            uint32_t c = e >> 24;
            hx = fmaf(sJ[3 * c + 0], px[j], hx);
            hy = fmaf(sJ[3 * c + 1], py[j], hy);
            hz = fmaf(sJ[3 * c + 2], pz[j], hz);
        }
        float x = px[slot], y = py[slot], z = pz[slot];
        v[0] += (double)(x * hx + y * hy + z * hz);
        int sp = a.sp[slot];
        if (ZEE) v[1] += (double)(x * sZ[3 * sp] + y * sZ[3 * sp + 1] + z * sZ[3 * sp + 2]);
        if (SIA)
            v[2] += (double)(sD[3 * sp] * x * x + sD[3 * sp + 1] * y * y + sD[3 * sp + 2] * z * z);
        float wgt = a.w[slot];
        v[3] += (double)(wgt * x);
        v[4] += (double)(wgt * y);
        v[5] += (double)(wgt * z);
    }
    block_sum<6>(v, a.partial + ((size_t)r * a.tiles + tile) * kRecord);
}

__global__ __launch_bounds__(kBlock) void k_ising_measure(MeasArgs a) {
    __shared__ long long red[kBlock / 64][2 * kMaxIsingClasses + 2];
    const int r = blockIdx.x / a.tiles, tile = blockIdx.x % a.tiles;
    const size_t base = (size_t)r * a.N;
    const int8_t* __restrict__ s = a.s + base;
    const int K = a.n_classes;
    long long bonds[kMaxIsingClasses + 1];
    long long zs[kMaxIsingClasses];
#pragma unroll
    for (int k = 0; k <= kMaxIsingClasses; ++k) bonds[k] = 0;
#pragma unroll
    for (int k = 0; k < kMaxIsingClasses; ++k) zs[k] = 0;
    double mz = 0.0;
#pragma unroll
    for (int q = 0; q < kMeasPerThread; ++q) {
        const int slot = tile * kMeasTile + q * kBlock + threadIdx.x;
        if (slot >= a.N) continue;
        const int si = s[slot];
        for (int k = 0; k < a.D; ++k) {
            uint32_t e = a.ell[(size_t)k * a.ell_stride + slot];
            int c = (int)(e >> 24);
            int sj = s[e & 0xffffffu];
#pragma unroll
            for (int kk = 1; kk <= kMaxIsingClasses; ++kk)
                if (kk == c) bonds[kk] += si * sj;
        }
        const int sp = a.sp[slot];
#pragma unroll
        for (int kk = 0; kk < kMaxIsingClasses; ++kk)
            if (kk == sp) zs[kk] += si;
        mz += (double)(a.w[slot] * (float)si);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = 1; k <= K; ++k) {
        long long t = wave_sum_ll(bonds[k]);
        if (lane == 0) red[wv][k] = t;
    }
    for (int k = 0; k < a.nsp; ++k) {
        long long t = wave_sum_ll(zs[k]);
        if (lane == 0) red[wv][kMaxIsingClasses + 1 + k] = t;
    }
    double m = wave_sum(mz);
    __shared__ double redm[kBlock / 64];
    if (lane == 0) redm[wv] = m;
    __syncthreads();
    double* out = a.partial + ((size_t)r * a.tiles + tile) * kRecord;
    long long* outl = reinterpret_cast<long long*>(out);
    if (threadIdx.x >= 1 && threadIdx.x <= K) {
        long long t = 0;
        for (int q = 0; q < kBlock / 64; ++q) t += red[q][threadIdx.x];
        outl[threadIdx.x] = t;
    }
    if (threadIdx.x >= 16 && threadIdx.x < 16 + a.nsp) {
        int k = threadIdx.x - 16;
        long long t = 0;
        for (int q = 0; q < kBlock / 64; ++q) t += red[q][kMaxIsingClasses + 1 + k];
        outl[16 + k] = t;
    }
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) t += redm[q];
        out[31] = t;
    }
}

// One wave per replica: fixed-order reduction of the tile partials, energy
// assembly and the observable() moments.
__global__ __launch_bounds__(64) void k_finalize(FinArgs a) {
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    const double* p = a.partial + (size_t)r * a.tiles * kRecord;
    double ex, ez, es, mx, my, mz;
    if (a.model == 1) {
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (int t = lane; t < a.tiles; t += 64)
            for (int k = 0; k < 6; ++k) v[k] += p[(size_t)t * kRecord + k];
        for (int k = 0; k < 6; ++k) v[k] = wave_sum(v[k]);
        ex = v[0]; ez = v[1]; es = v[2]; mx = v[3]; my = v[4]; mz = v[5];
    } else {
        const long long* pl = reinterpret_cast<const long long*>(p);
        long long b[kMaxIsingClasses + 1], z[kMaxIsingClasses];
        for (int k = 0; k <= kMaxIsingClasses; ++k) b[k] = 0;
        for (int k = 0; k < kMaxIsingClasses; ++k) z[k] = 0;
        double m = 0.0;
        for (int t = lane; t < a.tiles; t += 64) {
            for (int k = 1; k <= a.n_classes; ++k) b[k] += pl[(size_t)t * kRecord + k];
            for (int k = 0; k < a.nsp; ++k) z[k] += pl[(size_t)t * kRecord + 16 + k];
            m += p[(size_t)t * kRecord + 31];
        }
        ex = 0.0;
        for (int k = 1; k <= a.n_classes; ++k) {
            b[k] = wave_sum_ll(b[k]);
            ex += a.ising_j[k] * (double)b[k];
            if (lane == 0) a.bonds[(size_t)r * (a.n_classes + 1) + k] = b[k];
        }
        ez = 0.0;
        for (int k = 0; k < a.nsp; ++k) {
            z[k] = wave_sum_ll(z[k]);
            ez += a.ising_zc[k] * (double)z[k];
        }
        es = 0.0;
        mx = 0.0; my = 0.0;
        mz = wave_sum(m);
    }
    if (lane != 0) return;
    const double etot = ex + ez + es;
    const double ephys = 0.5 * ex + ez + es;
    a.parts[4 * r + 0] = ex;
    a.parts[4 * r + 1] = ez;
    a.parts[4 * r + 2] = es;
    a.parts[4 * r + 3] = ephys;
    a.M[3 * r + 0] = mx;
    a.M[3 * r + 1] = my;
    a.M[3 * r + 2] = mz;
    a.E_pt[r] = a.pt_etot ? etot : ephys;
    if (a.sample) {
        // get_eng 'eng' / get_mag 'mag', booked to the replica's temperature slot
        const double e = etot / (2.0 * a.N);
        const double m = sqrt(mx * mx + my * my + mz * mz) / a.N;
        double* mo = a.moments + 8 * (size_t)a.slot_of[r];
        const double e2 = e * e, m2 = m * m;
        mo[0] += 1.0;
        mo[1] += e;
        mo[2] += e2;
        mo[3] += e2 * e2;
        mo[4] += m;
        mo[5] += m2;
        mo[6] += m2 * m2;
    }
}

// ---------------------------------------------------------------------------
// replica exchange: one thread per ladder, slots permuted, spins untouched
// ---------------------------------------------------------------------------
__global__ void k_pt_swap(PtArgs a) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.n_groups) return;
    const int s0 = a.gstart[g], s1 = a.gstart[g + 1];
    int nacc = 0;
    for (int k = s0 + a.parity; k + 1 < s1; k += 2) {
        const int ra = a.rep_at[k], rb = a.rep_at[k + 1];
        const double b = a.beta_slot[k], bp = a.beta_slot[k + 1];
        const double E = a.E_pt[ra], Ep = a.E_pt[rb];
        const double delta = -(bp - b) * (E - Ep);
        bool accept = delta > 0.0;
        if (!accept) {
            u32x4 w = rng_draw(a.seed, (uint32_t)k, (uint32_t)g, a.swap_index, RNG_PT, 0);
            const double u = (double)(w.x >> 8) * 0x1p-24;
            accept = u < exp(delta);
        }
        a.att[k] += 1;
        if (accept) {
            a.acc[k] += 1;
            a.rep_at[k] = rb;
            a.rep_at[k + 1] = ra;
            a.slot_of[ra] = k + 1;
            a.slot_of[rb] = k;
            ++nacc;
        }
    }
    // cold -> hot -> cold round trips of replica labels
    const int cold = (a.beta_slot[s0] > a.beta_slot[s1 - 1]) ? s0 : s1 - 1;
    const int hot = (cold == s0) ? s1 - 1 : s0;
    if (cold != hot) {
        const int rc = a.rep_at[cold], rh = a.rep_at[hot];
        if (a.trip_state[rc] == 2) a.trips[rc] += 1;
        a.trip_state[rc] = 1;
        if (a.trip_state[rh] == 1) a.trip_state[rh] = 2;
    }
    if (nacc) atomicAdd(a.n_accepted, nacc);
}

// ---------------------------------------------------------------------------
// initial states
// ---------------------------------------------------------------------------


__global__ __launch_bounds__(kBlock) void k_init(InitArgs a) {
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (size_t)a.N * a.R) return;
    const int r = (int)(t / a.N), slot = (int)(t % a.N);
    u32x4 w = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, a.step, RNG_INIT, 0);
    if (a.model == 0) {
        if (a.mode == 1) a.s[t] = (int8_t)(-a.s[t]);
        else a.s[t] = (w.x >> 31) ? (int8_t)1 : (int8_t)-1;
        return;
    }
    float x, y, z;
    if (!a.angle) {
        marsaglia(w, a.seed, (uint32_t)slot, (uint32_t)r, a.step, RNG_INIT, x, y, z);
        if (a.phi) a.phi[t] = 0.f;
    } else {
        float ph;
        if (a.mode == 1) {
            angle_move(w, a.dphi, a.phi[t], x, y, z, ph);
        } else {
            angle_move(w, 3.14159265358979323846f, 0.0f, x, y, z, ph);
        }
        a.phi[t] = ph;
    }
    a.sx[t] = x;
    a.sy[t] = y;
    a.sz[t] = z;
}

// ---------------------------------------------------------------------------
// local fields and site energies of one replica
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_local_fields(MeasArgs a, int r, float* h, float* e,
                                                         int zee, int sia, int model) {
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= a.N) return;
    const size_t base = (size_t)r * a.N;
    float hx = 0.f, hy = 0.f, hz = 0.f, x, y, z;
    if (model == 1) {
        for (int k = 0; k < a.D; ++k) {
            uint32_t en = a.ell[(size_t)k * a.ell_stride + slot];
            uint32_t j = en & 0xffffffu, c = en >> 24;
            hx = fmaf(a.jclass[3 * c + 0], a.sx[base + j], hx);
            hy = fmaf(a.jclass[3 * c + 1], a.sy[base + j], hy);
            hz = fmaf(a.jclass[3 * c + 2], a.sz[base + j], hz);
        }
        x = a.sx[base + slot];
        y = a.sy[base + slot];
        z = a.sz[base + slot];
    } else {
        for (int k = 0; k < a.D; ++k) {
            uint32_t en = a.ell[(size_t)k * a.ell_stride + slot];
            uint32_t j = en & 0xffffffu, c = en >> 24;
            hz = fmaf(a.jclass[3 * c + 2], (float)a.s[base + j], hz);
        }
        x = 0.f;
        y = 0.f;
        z = (float)a.s[base + slot];
    }
    h[3 * slot + 0] = hx;
    h[3 * slot + 1] = hy;
    h[3 * slot + 2] = hz;
    int sp = a.sp[slot];
    float en = x * hx + y * hy + z * hz;
    if (zee) en += x * a.zvec[3 * sp] + y * a.zvec[3 * sp + 1] + z * a.zvec[3 * sp + 2];
    if (sia && model == 1)
        en += a.dsia[3 * sp] * x * x + a.dsia[3 * sp + 1] * y * y + a.dsia[3 * sp + 2] * z * z;
    e[slot] = en;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_xyz_sweep(const XyzArgs& a, bool angle, bool zee, bool sia, hipStream_t st) {
    dim3 grid(blocks_for(a.R, a.tiles, a.xcd_map)), block(kBlock);
#define L(A, Z, S) hipLaunchKernelGGL((k_xyz_sweep<A, Z, S>), grid, block, 0, st, a)
    if (angle) {
        if (zee) { if (sia) L(true, true, true); else L(true, true, false); }
        else { if (sia) L(true, false, true); else L(true, false, false); }
    } else {
        if (zee) { if (sia) L(false, true, true); else L(false, true, false); }
        else { if (sia) L(false, false, true); else L(false, false, false); }
    }
#undef L
}

void launch_xyz_ovrr(const XyzArgs& a, bool angle, bool zee, hipStream_t st) {
    dim3 grid(blocks_for(a.R, a.tiles, a.xcd_map)), block(kBlock);
    if (angle) {
        if (zee) hipLaunchKernelGGL((k_xyz_ovrr<true, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_xyz_ovrr<true, false>), grid, block, 0, st, a);
    } else {
        if (zee) hipLaunchKernelGGL((k_xyz_ovrr<false, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_xyz_ovrr<false, false>), grid, block, 0, st, a);
    }
}

void launch_ising_sweep(const IsingArgs& a, hipStream_t st) {
    dim3 grid(blocks_for(a.R, a.tiles, a.xcd_map)), block(kBlock);
    hipLaunchKernelGGL(k_ising_sweep, grid, block, 0, st, a);
}

void launch_measure(const MeasArgs& a, int model, bool zee, bool sia, hipStream_t st) {
    dim3 grid(a.R * a.tiles), block(kBlock);
    if (model == 0) {
        hipLaunchKernelGGL(k_ising_measure, grid, block, 0, st, a);
        return;
    }
    if (zee) {
        if (sia) hipLaunchKernelGGL((k_xyz_measure<true, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_xyz_measure<true, false>), grid, block, 0, st, a);
    } else {
        if (sia) hipLaunchKernelGGL((k_xyz_measure<false, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_xyz_measure<false, false>), grid, block, 0, st, a);
    }
}

void launch_finalize(const FinArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize, dim3(a.R), dim3(64), 0, st, a);
}

void launch_pt_swap(const PtArgs& a, hipStream_t st) {
    int blocks = (a.n_groups + 63) / 64;
    hipLaunchKernelGGL(k_pt_swap, dim3(blocks), dim3(64), 0, st, a);
}

void launch_init(const InitArgs& a, hipStream_t st) {
    size_t n = (size_t)a.N * a.R;
    hipLaunchKernelGGL(k_init, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, a);
}

void launch_local_fields(const MeasArgs& a, int r, float* h, float* e, int zee, int sia,
                         int model, hipStream_t st) {
    hipLaunchKernelGGL(k_local_fields, dim3((a.N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, a,
                       r, h, e, zee, sia, model);
}

}  // namespace ether
