// kernels.hip - gfx950 kernels of the Monte Carlo sweep engine.
//
//   k_xyz_sweep     one colour phase of graph-coloured single-spin Metropolis for
//                   XYZ spins (src/Metropolis.f90:35-62 with the proposals of
//                   src/update_spin_details.f90:28-51 and the energy change of
//                   src/Hamiltonian.f90:20-143)
//   k_ising_sweep   the same for Ising spins, four replicas per lane (packed
//                   int8), exact integer acceptance thresholds
//   k_xyz_ovrr      one colour phase of overrelaxation (src/overrelaxation.f90:30-49)
//   k_*_measure     per-replica energy and magnetisation partial sums
//                   (src/get_tot_energy.f90:35-45, src/get_tot_magnetisation.f90:35-38);
//                   for XYZ the same record is usually booked by the sweep that
//                   precedes the measurement (k_xyz_sweep<..., MEAS = true>)
//   k_finalize      fixed-order reduction of the partials plus the observable()
//                   moment update (src/observable.f90:30-43, src/get_eng.f90:36-39,
//                   src/get_mag.f90:36-40)
//   k_pt_swap       replica exchange as a slot permutation (src/parallel_tempering.f90:50-108)
//   k_init          fresh_spins / random start (src/fresh_spins.f90:31-38)
//
// Work mapping: a wavefront owns one site (slot) of 64 consecutive replicas
// (256 replicas for Ising, four per lane) and walks a few consecutive sites of
// its colour class.  The site, its neighbour list and its coupling classes are
// wave-uniform (scalar loads); every spin access of the wave is one contiguous
// segment of the replica-major planes.  None of this is GEMM-shaped; the
// kernels are bound by HBM/L2 traffic and by the VALU work of the generator.
#include <atomic>
#include <hip/hip_runtime.h>

#include "engine_common.h"
#include "philox.h"

// Every fused multiply-add in this file is written as fmaf and the compiler
// contracts nothing else: a kernel's rounding then does not depend on how the
// code around an expression is arranged, so the variants of one update (launch
// shapes, lanes per site, prefetching, batches) give bit-identical spins.
#pragma clang fp contract(off)

namespace ether {

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

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
    return fmaf(-two_pi, floorf((p + 3.14159265358979323846f) / two_pi), p);
}
// a.b, accumulated x, y, z in order
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return fmaf(az, bz, fmaf(ay, by, ax * bx));
}
// Unit-length renormalisation with the hardware reciprocal square root
// (1 ulp): every vector it is applied to has length 1 to within a few ulp
// already, and the spins are FP32 to begin with.
__device__ __forceinline__ void normalise3(float& x, float& y, float& z) {
    const float inv = __builtin_amdgcn_rsqf(dot3(x, y, z, x, y, z));
    x *= inv;
    y *= inv;
    z *= inv;
}
// round(v * scale) as an integer-valued double.  scale is a power of two
// chosen on the host so that |v * scale| < 2^50 for every site term and the
// sum of the terms of all sites stays below 2^51: adding 1.5 * 2^52 puts the
// value in the binade whose spacing is 1, so the addition rounds it to an
// integer and the subtraction is exact.  Sums of such terms are exact in any
// order.
__device__ __forceinline__ double fixed_point(float v, double scale) {
    const double magic = 6755399441055744.0;   // 1.5 * 2^52
    return fma((double)v, scale, magic) - magic;
}

// Marsaglia's disc test a1^2 + a2^2 < 2^46 on 24-bit coordinates, exact.
// The FP32 sum is within 2^24 of the integer sum (|a| <= 2^23 converts
// exactly; two roundings of at most 2^-24 relative), so outside a band of
// 2^26 around the boundary its decision is the integer one; a wave with a
// lane inside the band (about one pair in 10^5) takes the integer test.
__device__ __forceinline__ bool in_disc(uint32_t cx, uint32_t cy) {
    const int32_t a1 = (int32_t)cx - (1 << 23), a2 = (int32_t)cy - (1 << 23);
    const float f1 = (float)a1, f2 = (float)a2;
    const float q = fmaf(f1, f1, f2 * f2);
    bool inside = q < 0x1p46f;
    const bool near = fabsf(q - 0x1p46f) <= 0x1p26f;
    if (__any(near)) {
        if (near) inside = (int64_t)a1 * a1 + (int64_t)a2 * a2 < (int64_t(1) << 46);
    }
    return inside;
}
// Marsaglia's point from an accepted pair (src/George_Marsaglia.f90:36-38), renormalised
__device__ __forceinline__ void marsaglia_point(uint32_t cx, uint32_t cy, float& x, float& y, float& z) {
    const int32_t a1 = (int32_t)cx - (1 << 23), a2 = (int32_t)cy - (1 << 23);
    const float v1 = (float)a1 * 0x1p-23f, v2 = (float)a2 * 0x1p-23f;   // exact
    const float s = fmaf(v1, v1, v2 * v2);
    const float rt = __builtin_amdgcn_sqrtf(1.0f - s);
    x = 2.0f * v1 * rt;
    y = 2.0f * v2 * rt;
    z = fmaf(-2.0f, s, 1.0f);
    normalise3(x, y, z);
}
// two candidate pairs per Philox block
__device__ __forceinline__ void candidates(u32x4 w, uint32_t& x0, uint32_t& y0, uint32_t& x1,
                                           uint32_t& y1) {
    x0 = w.x >> 8;
    y0 = w.y >> 8;
    x1 = ((w.w >> 16) << 8) | (w.x & 0xffu);
    y1 = ((w.w & 0xffffu) << 8) | (w.y & 0xffu);
}

// Per-lane Marsaglia with retries (attempt >= first): used by the start-state
// kernel and as the fallback of the pooled proposal.
__device__ void marsaglia_retry(uint64_t seed, uint32_t id, uint32_t r, uint32_t step,
                                uint32_t purpose, uint32_t first, float& x, float& y, float& z) {
    for (uint32_t att = first;; ++att) {
        u32x4 w = rng_draw(seed, id, r, step, purpose, att);
        uint32_t x0, y0, x1, y1;
        candidates(w, x0, y0, x1, y1);
        if (in_disc(x0, y0)) { marsaglia_point(x0, y0, x, y, z); return; }
        if (in_disc(x1, y1)) { marsaglia_point(x1, y1, x, y, z); return; }
    }
}

// Pooled Marsaglia proposal of one wavefront (64 replicas of one site).
// Every lane draws two candidate pairs from its own Philox block.  A lane
// whose first pair falls outside the disc takes, in lane order, the second
// pair of the lanes whose second pair is inside ("donors", in lane order).
// The assignment depends only on the inside/outside bits, so each lane still
// gets an independent uniform point of the disc; only if donors run out does
// a lane draw more blocks itself (attempts 1, 2, ...).  Saves the second
// Philox block that almost every wave would otherwise pay for one lane.
__device__ __forceinline__ void marsaglia_pooled(u32x4 w, bool active, uint64_t seed, uint32_t id,
                                                 uint32_t r, uint32_t step, uint64_t* pool,
                                                 float& x, float& y, float& z) {
    uint32_t x0, y0, x1, y1;
    candidates(w, x0, y0, x1, y1);
    const bool ok0 = in_disc(x0, y0), ok1 = in_disc(x1, y1);
    const uint64_t needy = __ballot(active && !ok0);
    const uint64_t donors = __ballot(active && ok1);
    const int lane = lane_id();
    const uint64_t below = lanes_below(lane);
    // branch-free exchange: non-donors write to their own spare entry, every
    // lane reads, and a select keeps what a needy lane is owed
    const bool give = active && ok1;
    pool[give ? __popcll(donors & below) : kWave + lane] = ((uint64_t)x1 << 32) | y1;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const int j = __popcll(needy & below);
    const bool take = active && !ok0 && j < __popcll(donors);
    const uint64_t v = pool[j];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const uint32_t ux = take ? (uint32_t)(v >> 32) : x0, uy = take ? (uint32_t)v : y0;
    const bool have = ok0 || take;
    float px, py, pz;
    marsaglia_point(ux, uy, px, py, pz);
    if (__any(active && !have)) {
        if (active && !have) marsaglia_retry(seed, id, r, step, RNG_METROPOLIS, 1, px, py, pz);
    }
    if (active) {
        x = px;
        y = py;
        z = pz;
    }
}

// ANGLE move (src/update_spin_details.f90:39-46)
__device__ __forceinline__ void angle_move(u32x4 w, float dphi, float phi_old, float& x, float& y,
                                           float& z, float& phi) {
    float sz = sym24f(w.x >> 8);
    phi = wrap_pif(phi_old + dphi * sym24f(w.y >> 8));
    float rt = sqrtf(fmaxf(0.0f, fmaf(-sz, sz, 1.0f)));
    float sn, cs;
    sincosf(phi, &sn, &cs);
    x = rt * cs;
    y = rt * sn;
    z = sz;
    normalise3(x, y, z);
}

// per-lane counters of the four waves of a block, summed into the slot counters
// Lane l of every wave of the block holds the counts of replicas
// r_lane + q (q < nrep_lane); r_lane is the same in every wave of the block.
// The sum goes to copy blockIdx % kAccCopies of the counters: device-scope
// atomics on one address serialise, and every workgroup books every slot of
// its replica group (the host adds the copies).
__device__ __forceinline__ void book_acceptance(unsigned* sAcc, int nrep_lane, const unsigned* mine,
                                                int r_lane, int R, const int* slot_of,
                                                unsigned long long* acc) {
    const int w = wave_id(), lane = lane_id();
    for (int q = 0; q < nrep_lane; ++q) sAcc[(w * kWave + lane) * nrep_lane + q] = mine[q];
    __syncthreads();
    if (w == 0) {
        for (int q = 0; q < nrep_lane; ++q) {
            unsigned t = 0;
            for (int v = 0; v < kWavesPerBlock; ++v) t += sAcc[(v * kWave + lane) * nrep_lane + q];
            const int r = r_lane + q;
            if (t && r < R)
                atomicAdd(acc + (size_t)(blockIdx.x % kAccCopies) * R + slot_of[r], (unsigned long long)t);
        }
    }
}

// Spin planes are replica-major: the value of (slot, replica r) sits at
// plane[slot * Rp + r].  A wave's site is uniform, so the row's byte offset
// is a scalar (soffset) and the lane adds its own byte offset r*4 (voffset):
// buffer loads and stores through a per-plane resource need no per-lane
// address arithmetic, and an offset past the plane reads 0 instead of
// faulting.  The host keeps every plane under 2 GiB.
typedef __amdgpu_buffer_rsrc_t Plane;
__device__ __forceinline__ Plane plane_of(const float* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float plane_ld(Plane p, uint32_t roff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(p, (int)roff, (int)soff, 0));
}
__device__ __forceinline__ void plane_st(Plane p, uint32_t roff, uint32_t soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), p, (int)roff, (int)soff, 0);
}
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
// A wave-uniform row of a table that no kernel writes (the ELL neighbour
// table), read through the constant address space: the compiler then fetches
// it with scalar loads into SGPRs instead of a vector load and readfirstlane.
typedef const __attribute__((address_space(4))) uint32_t* ConstRow;
__device__ __forceinline__ ConstRow const_row(const uint32_t* p) { return (ConstRow)p; }
typedef const __attribute__((address_space(4))) float* ConstF32;
typedef const __attribute__((address_space(4))) uint8_t* ConstU8;

// Coupling of class c as (Jxx, Jyy, Jzz, 0), staged in LDS (one 16-byte read)
__device__ __forceinline__ void stage_couplings(float4* sJ4, const float* jclass, int n_classes) {
    for (int t = threadIdx.x; t <= n_classes; t += kBlock)
        sJ4[t] = make_float4(jclass[3 * t + 0], jclass[3 * t + 1], jclass[3 * t + 2], 0.f);
}

// A short row's gathers (DM <= 8) and their sum, apart: work that does not
// depend on the neighbours (the proposal's Philox block) goes between
// issuing the loads and waiting for them.  Same order of accumulation as
// local_field, so the field is bit-identical.
template <int DM>
struct RowVals {
    float x[DM], y[DM], z[DM];
    uint32_t cl[DM];
};
// A site's ELL row into registers: wave-uniform (scalar loads) in the
// one-site-per-wave kernels, per lane in the lanes kernels.  The lanes sweep
// loads the next step's rows while it updates the current sites, so the
// neighbour loads of a step need not wait for their rows (batches of 58-66
// simple-cubic, triangular and pyrochlore lattices of 8..32 replicas: 2.5-4%
// faster).
template <int DM>
__device__ __forceinline__ void load_row(const uint32_t* nb, uint32_t (&e)[DM]) {
    const ConstRow row = const_row(nb);
#pragma unroll
    for (int k = 0; k < DM; ++k) e[k] = row[k];
}
template <int DM>
__device__ __forceinline__ void load_row_lanes(const uint32_t* nb, uint32_t (&e)[DM]) {
#pragma unroll
    for (int k = 0; k < DM; ++k) e[k] = nb[k];
}

template <int DM>
__device__ __forceinline__ void gather_row(Plane px, Plane py, Plane pz, const uint32_t (&row)[DM],
                                           uint32_t row_bytes, uint32_t roff, RowVals<DM>& v) {
#pragma unroll
    for (int k = 0; k < DM; ++k) {
        const uint32_t e = row[k];
        const uint32_t so = (e & 0xffffffu) * row_bytes;
        v.cl[k] = e >> 24;
        v.x[k] = plane_ld(px, roff, so);
        v.y[k] = plane_ld(py, roff, so);
        v.z[k] = plane_ld(pz, roff, so);
    }
}
template <int DM>
__device__ __forceinline__ void sum_row(const RowVals<DM>& v, const float4* sJ4, float& hx, float& hy,
                                        float& hz) {
    hx = 0.f;
    hy = 0.f;
    hz = 0.f;
#pragma unroll
    for (int k = 0; k < DM; ++k) {
        const float4 j = sJ4[v.cl[k]];
        hx = fmaf(j.x, v.x[k], hx);
        hy = fmaf(j.y, v.y[k], hy);
        hz = fmaf(j.z, v.z[k], hz);
    }
}

// The exchange field from the neighbours of earlier colours only (slots
// below c_begin): the fused measurement's bond term with three or more
// colours.  Same operations in the same order as sum_row on the entries it
// keeps, so where every neighbour is earlier it equals the full field.
template <int DM>
__device__ __forceinline__ void sum_row_lower(const RowVals<DM>& v, const uint32_t (&row)[DM], int c_begin,
                                              const float4* sJ4, float& lx, float& ly, float& lz) {
    lx = 0.f;
    ly = 0.f;
    lz = 0.f;
#pragma unroll
    for (int k = 0; k < DM; ++k) {
        const bool lo = (int)(row[k] & 0xffffffu) < c_begin;
        const float4 j = sJ4[v.cl[k]];
        lx = fmaf(lo ? j.x : 0.f, v.x[k], lx);
        ly = fmaf(lo ? j.y : 0.f, v.y[k], ly);
        lz = fmaf(lo ? j.z : 0.f, v.z[k], lz);
    }
}

// Neighbours gathered per chunk of a long row: all DM loads of a row in
// flight at once cost registers (130 VGPRs at DM = 18, 3 waves per SIMD);
// chunks of 6 with a scheduling barrier between them keep 4 waves per SIMD
// and ran the 16^3-cell pyrochlore phase 5% faster.  The accumulation order
// is the same, so the field is bit-identical.
constexpr int field_chunk(int DM) {
    return DM <= 8 ? DM : (DM % 6 == 0 ? 6 : (DM % 8 == 0 ? 8 : DM));
}

// Exchange field h = sum_k J_{c_k} o S_{j_k} of the lane's replica at the site
// whose ELL row is nb (src/Hamiltonian.f90:72-103 / src/get_ovrr_vec.f90:64-82).
// The row entries are wave-uniform (read to scalar registers), so neighbour
// row addresses and coupling classes cost no vector instructions.  With
// DM > 0 the row has exactly DM entries (padded with class 0, J = 0) and all
// neighbour loads are issued before the first use, so their latencies overlap
// each other and the generator work that follows.  Accumulation order is
// k = 0..D-1 in both forms.
template <int DM, bool LOWER = false>
__device__ __forceinline__ void local_field(Plane px, Plane py, Plane pz, const uint32_t* nb, int D,
                                            uint32_t row_bytes, uint32_t roff, const float4* sJ4,
                                            float& hx, float& hy, float& hz, int c_begin = 0,
                                            float* lower = nullptr) {
    hx = 0.f;
    hy = 0.f;
    hz = 0.f;
    float lx = 0.f, ly = 0.f, lz = 0.f;   // LOWER: the part from slots below c_begin (sum_row_lower)
    const ConstRow row = const_row(nb);
    if constexpr (DM > 0) {
        constexpr int CH = field_chunk(DM);
#pragma unroll
        for (int c0 = 0; c0 < DM; c0 += CH) {
            float vx[CH], vy[CH], vz[CH];
            uint32_t cl[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) {
                const uint32_t e = row[c0 + k];
                const uint32_t so = (e & 0xffffffu) * row_bytes;
                cl[k] = e >> 24;
                vx[k] = plane_ld(px, roff, so);
                vy[k] = plane_ld(py, roff, so);
                vz[k] = plane_ld(pz, roff, so);
            }
#pragma unroll
            for (int k = 0; k < CH; ++k) {
                const float4 j = sJ4[cl[k]];
                hx = fmaf(j.x, vx[k], hx);
                hy = fmaf(j.y, vy[k], hy);
                hz = fmaf(j.z, vz[k], hz);
                if constexpr (LOWER) {
                    const bool lo = (int)(row[c0 + k] & 0xffffffu) < c_begin;
                    lx = fmaf(lo ? j.x : 0.f, vx[k], lx);
                    ly = fmaf(lo ? j.y : 0.f, vy[k], ly);
                    lz = fmaf(lo ? j.z : 0.f, vz[k], lz);
                }
            }
            if (CH < DM) __builtin_amdgcn_sched_barrier(0);   // keep the chunks apart
        }
    } else {
        for (int k = 0; k < D; ++k) {
            const uint32_t e = row[k];
            const uint32_t so = (e & 0xffffffu) * row_bytes;
            const float4 j = sJ4[e >> 24];
            const float vx = plane_ld(px, roff, so), vy = plane_ld(py, roff, so), vz = plane_ld(pz, roff, so);
            hx = fmaf(j.x, vx, hx);
            hy = fmaf(j.y, vy, hy);
            hz = fmaf(j.z, vz, hz);
            if constexpr (LOWER) {
                const bool lo = (int)(e & 0xffffffu) < c_begin;
                lx = fmaf(lo ? j.x : 0.f, vx, lx);
                ly = fmaf(lo ? j.y : 0.f, vy, ly);
                lz = fmaf(lo ? j.z : 0.f, vz, lz);
            }
        }
    }
    if constexpr (LOWER) {
        lower[0] = lx;
        lower[1] = ly;
        lower[2] = lz;
    }
}

// Reflection of S about the local field A: S' = 2(A.S)/(A.A) A - S, then
// renormalised (src/overrelaxation.f90:44-49).  Returns false (S unchanged)
// where A vanishes.
__device__ __forceinline__ bool reflect_about(float hx, float hy, float hz, float& x, float& y,
                                              float& z) {
    const float aa = dot3(hx, hy, hz, hx, hy, hz);
    if (!(aa > 0.f)) return false;
    const float f = 2.0f * dot3(hx, hy, hz, x, y, z) * __builtin_amdgcn_rcpf(aa);
    x = fmaf(f, hx, -x);
    y = fmaf(f, hy, -y);
    z = fmaf(f, hz, -z);
    normalise3(x, y, z);
    return true;
}

// Energy change of S -> S' in the local field h (src/Hamiltonian.f90:72-124,
// field already holding the Zeeman term) and the single-ion term
// D.(S'^2 - S^2) (src/Hamiltonian.f90:127-143).
__device__ __forceinline__ float delta_field(float nx, float ny, float nz, float x, float y, float z,
                                             float hx, float hy, float hz) {
    return dot3(nx - x, ny - y, nz - z, hx, hy, hz);
}
__device__ __forceinline__ float delta_sia(const float* D, float nx, float ny, float nz, float x,
                                           float y, float z) {
    return dot3(D[0], D[1], D[2], fmaf(nx, nx, -(x * x)), fmaf(ny, ny, -(y * y)), fmaf(nz, nz, -(z * z)));
}

// Fused measurement: the record terms of one site's final spin f, added to
// the lane's fixed-point sums m (see XyzArgs::meas).  In a phase whose
// neighbours are all of earlier colours (final already) the exchange term is
// f.hl, hl the site's exchange field, counted twice (every bond is booked
// once, from its later-coloured end); in a phase whose neighbours all come
// later it is left to them.  The others are the standalone measurement's
// per-site terms.
__device__ __forceinline__ void book_site(double (&m)[kMeasWords], const XyzArgs& a, bool zee,
                                          bool sia, const float* sZ, const float* sD, int sp,
                                          int slot, float fx, float fy, float fz, const float* hl) {
    if (a.meas_lower) m[0] += fixed_point(2.0f * dot3(fx, fy, fz, hl[0], hl[1], hl[2]), a.mscale[0]);
    if (zee) m[1] += fixed_point(dot3(fx, fy, fz, sZ[3 * sp], sZ[3 * sp + 1], sZ[3 * sp + 2]), a.mscale[1]);
    if (sia)
        m[2] += fixed_point(dot3(sD[3 * sp], sD[3 * sp + 1], sD[3 * sp + 2], fx * fx, fy * fy, fz * fz),
                            a.mscale[2]);
    const float wgt = ((ConstF32)a.w)[slot];
    m[3] += fixed_point(wgt * fx, a.mscale[3]);
    m[4] += fixed_point(wgt * fy, a.mscale[3]);
    m[5] += fixed_point(wgt * fz, a.mscale[3]);
}

// Adds a workgroup's fixed-point sums to macc.  Lanes of one replica (one
// per wave, or one in each of the sps lane groups of `per` lanes of a wave in
// the lanes kernels: lane s * per + q holds replica r_base + q) are summed
// through LDS, then one
// int64 atomic per (replica, word) into copy blockIdx % kMeasCopies of the
// sums: device-scope atomics on one address serialise, and hundreds of
// workgroups book every replica.  The terms are integers whose sums stay
// below 2^51, so FP64 adds them exactly and the result is the same in any
// order.  Called by every thread of the workgroup.
__device__ __forceinline__ void book_measurement(double* buf /* [kMeasWords][kBlock] */,
                                                 const double (&m)[kMeasWords], int per, int sps, int r_base,
                                                 const XyzArgs& a) {
#pragma unroll
    for (int k = 0; k < kMeasWords; ++k) buf[k * kBlock + threadIdx.x] = m[k];
    __syncthreads();
    const int w = wave_id(), lane = lane_id();
    for (int k = w; k < kMeasWords; k += kWavesPerBlock) {
        // words that are zero for the whole launch take no atomics
        if ((k == 0 && !a.meas_lower) || (k == 1 && !a.zee) || (k == 2 && !a.sia)) continue;
        const int r = r_base + lane;
        if (lane < per && r < a.R) {
            double t = 0.0;
            for (int v = 0; v < kWavesPerBlock; ++v)
                for (int q = 0; q < sps; ++q) t += buf[k * kBlock + v * kWave + q * per + lane];
            const size_t copy = blockIdx.x % kMeasCopies;
            atomicAdd(a.macc + (copy * a.Rp + r) * kMeasWords + k, (unsigned long long)(long long)t);
        }
    }
}

// The same sums, written to the workgroup's own record instead (three or more
// colours): unit bb of the phase, its lane l holding replica l of group
// bb % G, kMeasWords int64 words per lane, stored replica-major (the records
// of one replica in one phase are contiguous, for the finalize).  Every
// workgroup of the phase writes its record, so the finalize reads records of
// the last measuring sweep only and nothing needs clearing.  Called by every
// thread of the workgroup.
__device__ __forceinline__ void book_records(double* buf /* [kMeasWords][kBlock] */,
                                             const double (&m)[kMeasWords], int per, int sps, int bb,
                                             const XyzArgs& a) {
#pragma unroll
    for (int k = 0; k < kMeasWords; ++k) buf[k * kBlock + threadIdx.x] = m[k];
    __syncthreads();
    const int g = bb % a.G, u = bb / a.G;
    long long* rec = a.mrec + (((size_t)a.mrec_phase * a.G + g) * kWave * a.mrec_pp + u) * kMeasWords;
    for (int o = threadIdx.x; o < per * kMeasWords; o += kBlock) {
        const int l = o / kMeasWords, k = o - l * kMeasWords;
        double t = 0.0;
        for (int v = 0; v < kWavesPerBlock; ++v)
            for (int q = 0; q < sps; ++q) t += buf[k * kBlock + v * kWave + q * per + l];
        rec[(size_t)l * a.mrec_pp * kMeasWords + k] = (long long)t;
    }
}

// ---------------------------------------------------------------------------
// XYZ Metropolis, one colour phase.  With a.fuse set the phase first reflects
// each site about its local field (selected with probability a.frac, the OVRR
// draw of the split pass) and proposes from the reflected spin: one pass over
// the spins instead of an OVRR phase and a sweep phase.  Within a phase the
// neighbours are fixed, so both moves use the same field.  With MEAS the
// phase also books the fused measurement (book_site / book_measurement).
// ---------------------------------------------------------------------------
template <bool ANGLE, int DM, bool MEAS>
__device__ __forceinline__ void xyz_sweep_block(const XyzArgs& a, int bb) {
    __shared__ float4 sJ4[kMaxClasses + 1];
    __shared__ float sZ[3 * kMaxSpecies];
    __shared__ float sD[3 * kMaxSpecies];
    __shared__ uint64_t sPool[kWavesPerBlock][2 * kWave];
    __shared__ unsigned sAcc[kBlock];
    const bool ZEE = a.zee != 0, SIA = a.sia != 0;   // wave-uniform
    stage_couplings(sJ4, a.jclass, a.n_classes);
    if (ZEE)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sZ[t] = a.zvec[t];
    if (SIA)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sD[t] = a.dsia[t];
    __syncthreads();

    const int g = bb % a.G, chunk = bb / a.G;
    const int w = wave_id(), lane = lane_id();
    const int r = g * kWave + lane;
    const uint32_t roff = 4u * (uint32_t)r, rowb = 4u * (uint32_t)a.Rp;
    const Plane px = plane_of(a.sx, a.plane_bytes), py = plane_of(a.sy, a.plane_bytes),
                pz = plane_of(a.sz, a.plane_bytes);
    const Plane pp = plane_of(ANGLE ? a.phi : a.sx, a.plane_bytes);
    const bool rv = r < a.R;
    const float beta = rv ? a.beta_slot[a.slot_of[r]] : 0.f;
    const uint32_t step = a.ctr[0] + a.step_off;
    uint64_t* pool = sPool[w];
    unsigned accepted = 0;
    double m[kMeasWords] = {0, 0, 0, 0, 0, 0};

    // the update of one site from its spin (x, y, z, phi) and local exchange
    // field (hl: the part from earlier colours, for the fused measurement)
    auto update = [&](int slot, float x, float y, float z, float cphi, float hx, float hy, float hz,
                      const float* hl, u32x4 wr) {
        const uint32_t so = (uint32_t)slot * rowb;
        const int sp = (ZEE || SIA) ? a.sp[slot] : 0;
        if (ZEE) {
            hx += sZ[3 * sp + 0];
            hy += sZ[3 * sp + 1];
            hz += sZ[3 * sp + 2];
        }
        bool reflected = false;
        if (a.fuse) {                                  // wave-uniform
            bool sel = true;
            if (a.frac < 1.0f) {
                const u32x4 wo = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, step, RNG_OVRR, a.pass);
                sel = (float)(wo.x >> 8) * 0x1p-24f < a.frac;
            }
            if (sel) {
                reflected = reflect_about(hx, hy, hz, x, y, z);
                if (ANGLE && reflected) cphi = atan2f(y, x);
            }
        }
        const float eta = eta24f(wr.z);
        float nx = 0.f, ny = 0.f, nz = 1.f, nphi = 0.f;
        if (ANGLE) {
            angle_move(wr, a.dphi, cphi, nx, ny, nz, nphi);
        } else {
            marsaglia_pooled(wr, rv, a.seed, (uint32_t)slot, (uint32_t)r, step, pool, nx, ny, nz);
        }
        float dE = delta_field(nx, ny, nz, x, y, z, hx, hy, hz);
        if (SIA) dE += delta_sia(sD + 3 * sp, nx, ny, nz, x, y, z);
        // accept if -beta*dE > log(eta)   (src/Metropolis.f90:52); one store
        // path for an accepted move and for a rejected move after a reflection
        const bool acc = rv && -beta * dE > __logf(eta);
        accepted += acc ? 1u : 0u;
        if (acc || (rv && reflected)) {
            plane_st(px, roff, so, acc ? nx : x);
            plane_st(py, roff, so, acc ? ny : y);
            plane_st(pz, roff, so, acc ? nz : z);
            if (ANGLE) plane_st(pp, roff, so, acc ? nphi : cphi);
        }
        if constexpr (MEAS) {
            if (acc) {
                x = nx;
                y = ny;
                z = nz;
            }
            book_site(m, a, ZEE, SIA, sZ, sD, sp, slot, x, y, z, hl);
        }
    };

    const int s0 = a.c_begin + (chunk * kWavesPerBlock + w) * a.spw;
    for (int q = 0; q < a.spw; ++q) {
        const int slot = s0 + q;                       // wave-uniform
        if (slot >= a.c_end) break;
        const uint32_t so = (uint32_t)slot * rowb;
        const float x = plane_ld(px, roff, so), y = plane_ld(py, roff, so), z = plane_ld(pz, roff, so);
        const float cphi = ANGLE ? plane_ld(pp, roff, so) : 0.f;
        const uint32_t* nb = a.ell + (size_t)slot * a.D;
        float hx, hy, hz;
        float hl[3] = {0.f, 0.f, 0.f};   // MEAS: the exchange field from earlier colours (book_site)
        u32x4 wr;
        if constexpr (DM > 0 && DM <= 8) {
            // the row at the step: loading the next site's row one step ahead
            // (scalar registers) cost a wave per SIMD on the 32^3 Heisenberg
            // lattice with 190 replicas (102 SGPRs) and
            // gained nothing
            uint32_t row[DM];
            load_row<DM>(nb, row);
            RowVals<DM> v;
            gather_row<DM>(px, py, pz, row, rowb, roff, v);
            wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, step, RNG_METROPOLIS, 0);
            // an empty asm that takes the block: the compiler computes it
            // here, while the gathers are in flight, instead of sinking it
            // to its first use after the field sum
            asm volatile("" : "+v"(wr.x), "+v"(wr.y), "+v"(wr.z), "+v"(wr.w));
            sum_row<DM>(v, sJ4, hx, hy, hz);
            if constexpr (MEAS) sum_row_lower<DM>(v, row, a.c_begin, sJ4, hl[0], hl[1], hl[2]);
        } else {
            if constexpr (MEAS)
                local_field<DM, true>(px, py, pz, nb, a.D, rowb, roff, sJ4, hx, hy, hz, a.c_begin, hl);
            else
                local_field<DM>(px, py, pz, nb, a.D, rowb, roff, sJ4, hx, hy, hz);
            wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, step, RNG_METROPOLIS, 0);
        }
        update(slot, x, y, z, cphi, hx, hy, hz, hl, wr);
    }
    // acceptance is booked to the temperature slot, as the reference books it per rank
    book_acceptance(sAcc, 1, &accepted, r, a.R, a.slot_of, a.acc);
    if constexpr (MEAS) {
        __shared__ double sMeas[kMeasWords * kBlock];
        if (!a.mrec)
            book_measurement(sMeas, m, kWave, 1, g * kWave, a);
        else if (a.c_end > a.c_begin)   // an empty phase (of a batch) owns no records
            book_records(sMeas, m, kWave, 1, bb, a);
    }
}

template <bool ANGLE, int DM, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_xyz_sweep(XyzArgs a) {
    xyz_sweep_block<ANGLE, DM, MEAS>(a, block_index(a.nblocks, a.xcd_map, a.G, a.nchunk));
}

// Batched form: workgroup b runs local workgroup items[b].y of the launch
// described by args[items[b].x] (one lattice's colour phase); x < 0 pads.
template <bool ANGLE, int DM, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_xyz_sweep_batch(const XyzArgs* __restrict__ args,
                                                           const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const XyzArgs a = args[it.x];
    xyz_sweep_block<ANGLE, DM, MEAS>(a, it.y);
}

// ---------------------------------------------------------------------------
// Overrelaxation, one colour phase: S' = 2(A.S)/(A.A) A - S, renormalised.
// ---------------------------------------------------------------------------
template <bool ANGLE, int DM>
__device__ __forceinline__ void xyz_ovrr_block(const XyzArgs& a, int bb) {
    __shared__ float4 sJ4[kMaxClasses + 1];
    __shared__ float sZ[3 * kMaxSpecies];
    const bool ZEE = a.zee != 0;
    stage_couplings(sJ4, a.jclass, a.n_classes);
    if (ZEE)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sZ[t] = a.zvec[t];
    __syncthreads();
    const int g = bb % a.G, chunk = bb / a.G;
    const int w = wave_id(), lane = lane_id();
    const int r = g * kWave + lane;
    if (r >= a.R) return;
    const uint32_t roff = 4u * (uint32_t)r, rowb = 4u * (uint32_t)a.Rp;
    const Plane px = plane_of(a.sx, a.plane_bytes), py = plane_of(a.sy, a.plane_bytes),
                pz = plane_of(a.sz, a.plane_bytes);
    const Plane pp = plane_of(ANGLE ? a.phi : a.sx, a.plane_bytes);
    const uint32_t step = a.ctr[0] + a.step_off;
    const int s0 = a.c_begin + (chunk * kWavesPerBlock + w) * a.spw;
    for (int q = 0; q < a.spw; ++q) {
        const int slot = s0 + q;
        if (slot >= a.c_end) break;
        if (a.frac < 1.0f) {
            u32x4 wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, step, RNG_OVRR, a.pass);
            if (!((float)(wr.x >> 8) * 0x1p-24f < a.frac)) continue;
        }
        const uint32_t so = (uint32_t)slot * rowb;
        float hx, hy, hz;
        local_field<DM>(px, py, pz, a.ell + (size_t)slot * a.D, a.D, rowb, roff, sJ4, hx, hy, hz);
        if (ZEE) {
            const int sp = a.sp[slot];
            hx += sZ[3 * sp + 0];
            hy += sZ[3 * sp + 1];
            hz += sZ[3 * sp + 2];
        }
        float nx = plane_ld(px, roff, so), ny = plane_ld(py, roff, so), nz = plane_ld(pz, roff, so);
        if (!reflect_about(hx, hy, hz, nx, ny, nz)) continue;
        plane_st(px, roff, so, nx);
        plane_st(py, roff, so, ny);
        plane_st(pz, roff, so, nz);
        if (ANGLE) plane_st(pp, roff, so, atan2f(ny, nx));
    }
}

template <bool ANGLE, int DM>
__global__ __launch_bounds__(kBlock) void k_xyz_ovrr(XyzArgs a) {
    xyz_ovrr_block<ANGLE, DM>(a, block_index(a.nblocks, a.xcd_map, a.G, a.nchunk));
}

template <bool ANGLE, int DM>
__global__ __launch_bounds__(kBlock) void k_xyz_ovrr_batch(const XyzArgs* __restrict__ args,
                                                          const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const XyzArgs a = args[it.x];
    xyz_ovrr_block<ANGLE, DM>(a, it.y);
}

// ---------------------------------------------------------------------------
// Ising Metropolis, one colour phase.  Lane l of group g owns replicas
// r0..r0+3 (r0 = 4(64g + l)) of the wave's site as one 32-bit word of int8
// spins and takes word r&3 of the Philox block with id (slot, r>>2).
// Down-spin counts of each coupling class are summed per byte (SWAR); dE then
// depends only on (species, sigma_i, sum_k stride_k n_k), and the acceptance
// test is an exact integer compare with a host-built threshold.
// ---------------------------------------------------------------------------
// v added to class c's per-byte counter and 1 to its bond count.  BOUNDED
// (rows of at most 8): the loop stops at the context's class count nc
// (uniform), which bounds c, so a one-class lattice pays one compare per
// neighbour instead of eight; longer unrolled rows keep the eight-way test
// (the bounded form of 18 neighbours crashes this compiler's register
// coalescer).
template <bool BOUNDED>
__device__ __forceinline__ void class_add(uint32_t (&d)[kMaxIsingClasses + 1], int (&cnt)[kMaxIsingClasses + 1],
                                          int nc, int c, uint32_t v) {
    if constexpr (BOUNDED) {
#pragma unroll
        for (int k = 1; k <= kMaxIsingClasses; ++k) {
            if (k > nc) break;
            const bool on = k == c;
            d[k] += on ? v : 0u;
            cnt[k] += on ? 1 : 0;
        }
    } else {
#pragma unroll
        for (int k = 1; k <= kMaxIsingClasses; ++k)
            if (k == c) d[k] += v;
#pragma unroll
        for (int k = 1; k <= kMaxIsingClasses; ++k)
            if (k == c) cnt[k] += 1;
    }
}

// The flip mask of one site's word (byte b is 0xFE where replica b accepts)
// from its per-byte down counts d and bond counts cnt per class, its word v
// and its Philox block wr.  The four thresholds are looked up together: from
// LDS when the tables are staged there (a uniform branch), else by buffer
// loads from the table in HBM (a distinct load, so the two are not merged
// into one flat load behind a pointer select).  A byte past the last replica
// (tbase < 0, uniform) reads table 0 at an index inside it (its down counts
// lie in [0, cnt]) and never flips.
__device__ __forceinline__ uint32_t ising_flip(const IsingArgs& a, const int* sStride, const uint32_t* sThr,
                                               const int (&tbase)[4], const uint32_t (&d)[kMaxIsingClasses + 1],
                                               const int (&cnt)[kMaxIsingClasses + 1], int spc, uint32_t v,
                                               u32x4 wr) {
    int ti[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        int idx = a.offset;
#pragma unroll
        for (int k = 1; k <= kMaxIsingClasses; ++k) {
            if (k > a.n_classes) break;
            const int down = (int)((d[k] >> (8 * b)) & 0xffu);
            idx += sStride[k] * (cnt[k] - 2 * down);
        }
        const int sigma = (int)(int8_t)((v >> (8 * b)) & 0xffu);
        ti[b] = max(tbase[b], 0) + (spc * 2 + (sigma > 0)) * a.P + idx;
    }
    uint32_t th[4];
    if (a.table_in_lds) {
#pragma unroll
        for (int b = 0; b < 4; ++b) th[b] = sThr[ti[b]];
    } else {
        const __amdgpu_buffer_rsrc_t tb = __builtin_amdgcn_make_buffer_rsrc(
            (void*)a.thr, (short)0, (int)((size_t)a.R * a.table_per_slot * 4u), 0x00020000);
#pragma unroll
        for (int b = 0; b < 4; ++b) th[b] = __builtin_amdgcn_raw_buffer_load_b32(tb, 4 * ti[b], 0, 0);
    }
    const uint32_t words[4] = {wr.x, wr.y, wr.z, wr.w};
    uint32_t flip = 0u;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if (tbase[b] >= 0 && (words[b] < th[b] || th[b] == 0xffffffffu)) flip |= 0xFEu << (8 * b);   // +1 <-> -1 as int8: x ^ 0xFE
    return flip;
}

template <int DM>
__device__ __forceinline__ void ising_sweep_block(const IsingArgs& a, int bb) {
    __shared__ int sStride[kMaxIsingClasses + 1];
    __shared__ uint32_t sThr[kIsingLdsTable];
    __shared__ unsigned sAcc[kBlock * 4];
    for (int t = threadIdx.x; t <= a.n_classes; t += kBlock) sStride[t] = a.stride[t];
    if (a.table_in_lds)
        for (int t = threadIdx.x; t < a.R * a.table_per_slot; t += kBlock) sThr[t] = a.thr[t];
    __syncthreads();
    const int g4 = bb % a.G4, chunk = bb / a.G4;
    const int w = wave_id(), lane = lane_id();
    const int rid = g4 * kWave + lane;          // Philox id of this lane's 4 replicas
    const int r0 = 4 * rid;
    const bool lv = r0 < a.Rp;
    int tbase[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = r0 + q;
        tbase[q] = (r < a.R) ? a.slot_of[r] * a.table_per_slot : -1;
    }
    unsigned accepted[4] = {0, 0, 0, 0};
    const uint32_t step = a.ctr[0] + a.step_off;
    const int s0 = a.c_begin + (chunk * kWavesPerBlock + w) * a.spw;
    for (int q = 0; q < a.spw; ++q) {
        const int slot = s0 + q;
        if (slot >= a.c_end) break;
        if (!lv) continue;
        uint32_t d[kMaxIsingClasses + 1];
        int cnt[kMaxIsingClasses + 1];
#pragma unroll
        for (int k = 0; k <= kMaxIsingClasses; ++k) {
            d[k] = 0;
            cnt[k] = 0;
        }
        const ConstRow nb = const_row(a.ell + (size_t)slot * a.D);   // wave-uniform row
        uint32_t* own = reinterpret_cast<uint32_t*>(a.s + (size_t)slot * a.Rp + r0);
        uint32_t v;
        u32x4 wr;
        if constexpr (DM > 0) {
            // issue every neighbour load and the own load first (the row is
            // padded to DM with class 0), then draw the Philox block while
            // they are in flight (the empty asm keeps the draw here instead
            // of at its first use)
            uint32_t nv[DM];
#pragma unroll
            for (int k = 0; k < DM; ++k)
                nv[k] = *reinterpret_cast<const uint32_t*>(a.s + (size_t)(nb[k] & 0xffffffu) * a.Rp + r0);
            v = *own;
            wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)rid, step, RNG_METROPOLIS, 0);
            asm volatile("" : "+v"(wr.x), "+v"(wr.y), "+v"(wr.z), "+v"(wr.w));
#pragma unroll
            for (int k = 0; k < DM; ++k) {
                const int c = (int)(nb[k] >> 24);
                class_add<(DM > 0 && DM <= 8)>(d, cnt, a.n_classes, c, (nv[k] >> 1) & 0x01010101u);   // 1 per byte where sigma = -1
            }
        } else {
            for (int k = 0; k < a.D; ++k) {
                const uint32_t e = nb[k];
                const int c = (int)(e >> 24);
                if (c == 0) continue;
                const uint32_t nbv = *reinterpret_cast<const uint32_t*>(a.s + (size_t)(e & 0xffffffu) * a.Rp + r0);
                class_add<(DM > 0 && DM <= 8)>(d, cnt, a.n_classes, c, (nbv >> 1) & 0x01010101u);
            }
            v = *own;
            wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)rid, step, RNG_METROPOLIS, 0);
        }
        const uint32_t flip = ising_flip(a, sStride, sThr, tbase, d, cnt, a.sp[slot], v, wr);
#pragma unroll
        for (int b = 0; b < 4; ++b) accepted[b] += (flip >> (8 * b + 1)) & 1u;
        if (flip) *own = v ^ flip;
    }
    book_acceptance(sAcc, 4, accepted, r0, a.R, a.slot_of, a.acc);
}

template <int DM>
__global__ __launch_bounds__(kBlock) void k_ising_sweep(IsingArgs a) {
    ising_sweep_block<DM>(a, block_index(a.nblocks, a.xcd_map, a.G4, a.nchunk));
}

template <int DM>
__global__ __launch_bounds__(kBlock) void k_ising_sweep_batch(const IsingArgs* __restrict__ args,
                                                             const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const IsingArgs a = args[it.x];
    ising_sweep_block<DM>(a, it.y);
}

// ---------------------------------------------------------------------------
// Ising sweeps with the lattice resident in LDS.  One workgroup holds one
// word of four replicas (r0 = 4 * blockIdx.x .. r0 + 3) for every site: the
// N words are loaded once, n_sweeps sweeps run over the colour classes with a
// workgroup barrier after each phase (a phase reads only other colours, so
// the in-place update is race-free), and the words are written back once.
// Every update is the one k_ising_sweep makes: the same Philox block (slot,
// word, sweep step), the same per-byte down counts per coupling class and the
// same integer threshold, so the spins are bit-identical to launching the
// colour phases one by one.  With meas set, the final state's record (bond
// sums per class over full rows, spin sums per species, M_z) is written as
// tile 0 of the measurement partials for k_finalize (integers, exact in FP64).
// ---------------------------------------------------------------------------
// class_add with a class that differs between lanes: every class takes a
// select, so the counters stay in registers (an indexed update would put them
// in scratch)
// NC > 0: the context has exactly NC classes (one class needs no select);
// NC = 0: up to kMaxIsingClasses, bounded by nc
template <int NC>
__device__ __forceinline__ void class_add_lane(uint32_t (&d)[kMaxIsingClasses + 1],
                                               int (&cnt)[kMaxIsingClasses + 1], int nc, int c, uint32_t v) {
    if constexpr (NC == 1) {
        const uint32_t on = (uint32_t)(c == 1);     // padding entries are class 0
        d[1] += v * on;
        cnt[1] += (int)on;
    } else {
#pragma unroll
        for (int k = 1; k <= (NC > 0 ? NC : kMaxIsingClasses); ++k) {
            if (NC > 0 || k <= nc) {                // uniform: the context's class count
                const uint32_t on = (uint32_t)(k == c);
                d[k] += v * on;
                cnt[k] += (int)on;
            }
        }
    }
}

// TL: the four replicas' threshold tables are staged in LDS (a.table_in_lds);
// a template argument, so that the lookups are LDS reads and not flat loads
// through a pointer chosen at run time
template <int DM, int NC, bool TL>
__global__ void __launch_bounds__(1024) k_ising_resident(IsingResArgs a) {
    constexpr int KC = NC > 0 ? NC : kMaxIsingClasses;   // classes the loops visit
    extern __shared__ __attribute__((aligned(16))) uint32_t sRes[];
    uint32_t* sS = sRes;                               // [N] words
    uint32_t* sThr = sRes + a.N;                       // [4][table_per_slot] when staged
    __shared__ unsigned sAccB[4];
    __shared__ int sCnt[(kMaxIsingClasses + kMaxIsingClasses) * 4];
    __shared__ double sMz[16][4];
    __shared__ int sCb[kMaxResColours + 1];           // colour boundaries (an indexed argument would go to scratch)
    const int rid = blockIdx.x, r0 = 4 * rid;
    const int t = threadIdx.x, nt = blockDim.x;
    int tbase[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int r = r0 + b;
        tbase[b] = (r < a.R) ? (TL ? b * a.table_per_slot : a.slot_of[r] * a.table_per_slot) : -1;
    }
    for (int q = t; q < a.N; q += nt) sS[q] = *reinterpret_cast<const uint32_t*>(a.s + (size_t)q * a.Rp + r0);
    if (TL)
        for (int b = 0; b < 4; ++b) {
            const int r = r0 + b;
            if (r >= a.R) continue;
            const uint32_t* src = a.thr + (size_t)a.slot_of[r] * a.table_per_slot;
            for (int q = t; q < a.table_per_slot; q += nt) sThr[b * a.table_per_slot + q] = src[q];
        }
    if (t < 4) sAccB[t] = 0u;
    if (t <= kMaxResColours) sCb[t] = a.cbeg[t];
    for (int q = t; q < 2 * kMaxIsingClasses * 4; q += nt) sCnt[q] = 0;
    __syncthreads();
    unsigned accepted[4] = {0u, 0u, 0u, 0u};
    const uint32_t step0 = a.ctr[0] + a.step_off;
    // per-byte down counts (rows padded with class 0 up to DM) per coupling class of one site's row
    auto gather = [&](int slot, uint32_t (&d)[kMaxIsingClasses + 1], int (&cnt)[kMaxIsingClasses + 1]) {
#pragma unroll
        for (int kk = 0; kk <= kMaxIsingClasses; ++kk) {
            d[kk] = 0u;
            cnt[kk] = 0;
        }
        const uint32_t* nb = a.ell + (size_t)slot * a.D;
        auto add = [&](int j) {
            const uint32_t e = nb[j];
            const int cl = (int)(e >> 24);
            const uint32_t nv = sS[e & 0xffffffu];
            class_add_lane<NC>(d, cnt, a.n_classes, cl, (nv >> 1) & 0x01010101u);   // 1 per byte where sigma = -1
        };
        if constexpr (DM > 0) {
#pragma unroll
            for (int j = 0; j < DM; ++j) add(j);
        } else {
            for (int j = 0; j < a.D; ++j) add(j);
        }
    };
    // the flip mask of one site's word: byte b is 0xFE where replica b
    // accepts.  The four thresholds are looked up together; a byte past the
    // last replica (tbase < 0, uniform) reads table 0 at an index inside it
    // (its down counts still lie in [0, cnt]) and never flips.
    auto decide = [&](int slot, const uint32_t (&d)[kMaxIsingClasses + 1], const int (&cnt)[kMaxIsingClasses + 1],
                      uint32_t v, u32x4 wr) {
        const uint32_t words[4] = {wr.x, wr.y, wr.z, wr.w};
        const int spc = a.sp[slot];
        uint32_t th[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int idx = a.offset;
#pragma unroll
            for (int kk = 1; kk <= KC; ++kk) {
                if (NC == 0 && kk > a.n_classes) break;
                const int down = (int)((d[kk] >> (8 * b)) & 0xffu);
                idx += a.stride[kk] * (cnt[kk] - 2 * down);
            }
            const int sigma = (int)(int8_t)((v >> (8 * b)) & 0xffu);
            const int ti = max(tbase[b], 0) + (spc * 2 + (sigma > 0)) * a.P + idx;
            th[b] = TL ? sThr[ti] : a.thr[ti];
        }
        uint32_t flip = 0u;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (tbase[b] >= 0 && (words[b] < th[b] || th[b] == 0xffffffffu))
                flip |= 0xFEu << (8 * b);   // +1 <-> -1 as int8: x ^ 0xFE
        return flip;
    };
    for (int k = 0; k < a.n_sweeps; ++k) {
        const uint32_t step = step0 + (uint32_t)k;
        for (int c = 0; c < a.ncol; ++c) {
            const int c_end = sCb[c + 1];
            // two sites of the class per thread in flight: sites of one colour
            // share no bond, so both rows are read before either word is
            // written and the updates are those of one site at a time.  A
            // thread without a second site repeats the first and drops it.
            for (int sa = sCb[c] + t; sa < c_end; sa += 2 * nt) {
                const bool hb = sa + nt < c_end;
                const int sb = hb ? sa + nt : sa;
                uint32_t da[kMaxIsingClasses + 1], db[kMaxIsingClasses + 1];
                int ca[kMaxIsingClasses + 1], cb[kMaxIsingClasses + 1];
                gather(sa, da, ca);
                gather(sb, db, cb);
                const uint32_t va = sS[sa], vb = sS[sb];
                const u32x4 wa = rng_draw(a.seed, (uint32_t)sa, (uint32_t)rid, step, RNG_METROPOLIS, 0);
                const u32x4 wb = rng_draw(a.seed, (uint32_t)sb, (uint32_t)rid, step, RNG_METROPOLIS, 0);
                const uint32_t fa = decide(sa, da, ca, va, wa);
                const uint32_t fb = hb ? decide(sb, db, cb, vb, wb) : 0u;
#pragma unroll
                for (int b = 0; b < 4; ++b) accepted[b] += ((fa >> (8 * b + 1)) & 1u) + ((fb >> (8 * b + 1)) & 1u);
                if (fa) sS[sa] = va ^ fa;
                if (fb) sS[sb] = vb ^ fb;
            }
            __syncthreads();
        }
    }
    for (int q = t; q < a.N; q += nt) *reinterpret_cast<uint32_t*>(a.s + (size_t)q * a.Rp + r0) = sS[q];
    // acceptances: integers, summed in LDS, one device atomic per replica
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if (accepted[b]) atomicAdd(&sAccB[b], accepted[b]);
    if (a.meas) {
        // bond sums per class over full rows (each bond from both ends, the
        // reference's double sum), spin sums per species, M_z per replica
        int bond[kMaxIsingClasses + 1][4], zs[kMaxIsingClasses][4];
        double mz[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk <= kMaxIsingClasses; ++kk)
#pragma unroll
            for (int b = 0; b < 4; ++b) bond[kk][b] = 0;
#pragma unroll
        for (int kk = 0; kk < kMaxIsingClasses; ++kk)
#pragma unroll
            for (int b = 0; b < 4; ++b) zs[kk][b] = 0;
        for (int slot = t; slot < a.N; slot += nt) {
            const uint32_t* nb = a.ell + (size_t)slot * a.D;
            const uint32_t v = sS[slot];
            uint32_t d[kMaxIsingClasses + 1];
            int cnt[kMaxIsingClasses + 1];
#pragma unroll
            for (int kk = 0; kk <= kMaxIsingClasses; ++kk) {
                d[kk] = 0u;
                cnt[kk] = 0;
            }
            for (int j = 0; j < a.D; ++j) {
                const uint32_t e = nb[j];
                const int cl = (int)(e >> 24);
                class_add_lane<NC>(d, cnt, a.n_classes, cl, ((sS[e & 0xffffffu] ^ v) >> 1) & 0x01010101u);   // 1 per byte where they differ
            }
            const int spc = a.sp[slot];
            const float w = a.w[slot];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
#pragma unroll
                for (int kk = 1; kk <= KC; ++kk)
                    bond[kk][b] += cnt[kk] - 2 * (int)((d[kk] >> (8 * b)) & 0xffu);
                const int sigma = (int)(int8_t)((v >> (8 * b)) & 0xffu);
#pragma unroll
                for (int kk = 0; kk < kMaxIsingClasses; ++kk) zs[kk][b] += sigma * (int)(kk == spc);
                mz[b] += (double)(sigma > 0 ? w : -w);
            }
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
#pragma unroll
            for (int kk = 1; kk <= KC; ++kk)
                if (kk <= a.n_classes && bond[kk][b]) atomicAdd(&sCnt[(kk - 1) * 4 + b], bond[kk][b]);
#pragma unroll
            for (int kk = 0; kk < kMaxIsingClasses; ++kk)
                if (kk < a.nsp && zs[kk][b]) atomicAdd(&sCnt[(kMaxIsingClasses + kk) * 4 + b], zs[kk][b]);
            // M_z: a fixed-order wave reduction, then the waves in order (the
            // sum of +-w in FP64 is exact in any order anyway)
            double m = mz[b];
            for (int off = 32; off > 0; off >>= 1) m += __shfl_xor(m, off);
            if ((t & 63) == 0) sMz[t >> 6][b] = m;
        }
    }
    __syncthreads();
    if (t < 4) {
        const int b = t, r = r0 + b;
        if (r < a.R && sAccB[b])
            atomicAdd(a.acc + (size_t)(blockIdx.x % kAccCopies) * a.R + a.slot_of[r], (unsigned long long)sAccB[b]);
        if (a.meas && r < a.Rp) {
            double* out = a.partial + (size_t)r * a.rec;    // tile 0
            int o = 0;
            for (int kk = 0; kk < a.n_classes; ++kk) out[o++] = (double)sCnt[kk * 4 + b];
            for (int kk = 0; kk < a.nsp; ++kk) out[o++] = (double)sCnt[(kMaxIsingClasses + kk) * 4 + b];
            double m = 0.0;
            for (int wv = 0; wv < (nt + 63) / 64; ++wv) m += sMz[wv][b];
            out[o] = m;
        }
    }
}

// ---------------------------------------------------------------------------
// The resident Ising sweeps with one workgroup per replica instead of per
// word: for fewer replicas than compute units a word's workgroup keeps one CU
// busy with four replicas while most CUs idle (16^3 sites, 190 replicas:
// 48 of 256), so each
// replica gets a workgroup of its own, its int8 spin of every site and its
// threshold table in LDS.  Every update is the word kernel's for byte
// b = r & 3 of word r >> 2: the same Philox block (slot, word, sweep), its
// word b, the same down counts and integer threshold; the four workgroups of
// a word each draw the block (the generator is a pure function of the
// counter), so the spins are bit-identical.  The record of a measurement
// (bond sums per class over full rows, spin sums per species, M_z) is the
// word kernel's for this replica, summed in the same thread order.
// ---------------------------------------------------------------------------
template <int DM, int NC, bool TL>
__device__ __forceinline__ void ising_resident1_block(const IsingResArgs& a, int r) {
    constexpr int KC = NC > 0 ? NC : kMaxIsingClasses;
    extern __shared__ __attribute__((aligned(16))) uint32_t sRes1[];
    int8_t* sS = reinterpret_cast<int8_t*>(sRes1);                 // [N] spins
    uint32_t* sThr = sRes1 + (a.N + 15) / 16 * 4;                   // [table_per_slot] when staged
    __shared__ unsigned sAcc1;
    __shared__ int sCnt1[2 * kMaxIsingClasses];
    __shared__ double sMz1[16];
    __shared__ int sCb[kMaxResColours + 1];
    const int rid = r >> 2, b = r & 3;
    const int t = threadIdx.x, nt = blockDim.x;
    const int tbase = TL ? 0 : a.slot_of[r] * a.table_per_slot;
    for (int q = t; q < a.N; q += nt) sS[q] = a.s[(size_t)q * a.Rp + r];
    if (TL) {
        const uint32_t* src = a.thr + (size_t)a.slot_of[r] * a.table_per_slot;
        for (int q = t; q < a.table_per_slot; q += nt) sThr[q] = src[q];
    }
    if (t == 0) sAcc1 = 0u;
    if (t <= kMaxResColours) sCb[t] = a.cbeg[t];
    for (int q = t; q < 2 * kMaxIsingClasses; q += nt) sCnt1[q] = 0;
    __syncthreads();
    unsigned accepted = 0u;
    const uint32_t step0 = a.ctr[0] + a.step_off;
    // down counts and bonds per coupling class of one site's row (rows padded
    // with class 0 up to DM)
    auto gather = [&](int slot, int (&dn)[kMaxIsingClasses + 1], int (&cnt)[kMaxIsingClasses + 1]) {
#pragma unroll
        for (int kk = 0; kk <= kMaxIsingClasses; ++kk) {
            dn[kk] = 0;
            cnt[kk] = 0;
        }
        const uint32_t* nb = a.ell + (size_t)slot * a.D;
        // NC = 0 (up to 8 classes): one byte per class in two 64-bit words
        // (at most 255 neighbours per class, checked at commit)
        uint64_t cp = 0, dp = 0;
        auto add = [&](int j) {
            const uint32_t e = nb[j];
            const int cl = (int)(e >> 24);
            const int down = sS[e & 0xffffffu] < 0 ? 1 : 0;
            if constexpr (NC == 1) {
                const int on = cl == 1;
                dn[1] += down & on;
                cnt[1] += on;
            } else if constexpr (NC == 2) {
#pragma unroll
                for (int kk = 1; kk <= 2; ++kk) {
                    const int on = kk == cl;
                    dn[kk] += down & on;
                    cnt[kk] += on;
                }
            } else {
                const uint64_t on = cl != 0 ? 1u : 0u;     // padding entries are class 0
                const int sh = (8 * (cl - 1)) & 63;
                cp += on << sh;
                dp += (on & (uint64_t)down) << sh;
            }
        };
        if constexpr (DM > 0) {
#pragma unroll
            for (int j = 0; j < DM; ++j) add(j);
        } else {
            for (int j = 0; j < a.D; ++j) add(j);
        }
        if constexpr (NC == 0) {
#pragma unroll
            for (int kk = 1; kk <= kMaxIsingClasses; ++kk) {
                cnt[kk] = (int)((cp >> (8 * (kk - 1))) & 0xffu);
                dn[kk] = (int)((dp >> (8 * (kk - 1))) & 0xffu);
            }
        }
    };
    // threshold lookup and test of one site (r < R always: the grid is R)
    auto decide = [&](int slot, const int (&dn)[kMaxIsingClasses + 1], const int (&cnt)[kMaxIsingClasses + 1],
                      int v, u32x4 wr) {
        int idx = a.offset;
#pragma unroll
        for (int kk = 1; kk <= KC; ++kk) {
            if (NC == 0 && kk > a.n_classes) break;
            idx += a.stride[kk] * (cnt[kk] - 2 * dn[kk]);
        }
        const int ti = tbase + (a.sp[slot] * 2 + (v > 0)) * a.P + idx;
        const uint32_t th = TL ? sThr[ti] : a.thr[ti];
        const uint32_t w = b == 0 ? wr.x : (b == 1 ? wr.y : (b == 2 ? wr.z : wr.w));
        return w < th || th == 0xffffffffu;
    };
    for (int k = 0; k < a.n_sweeps; ++k) {
        const uint32_t step = step0 + (uint32_t)k;
        for (int c = 0; c < a.ncol; ++c) {
            const int c_end = sCb[c + 1];
            // two sites of the class per thread in flight, as in the word kernel
            for (int sa = sCb[c] + t; sa < c_end; sa += 2 * nt) {
                const bool hb = sa + nt < c_end;
                const int sb = hb ? sa + nt : sa;
                int da[kMaxIsingClasses + 1], db[kMaxIsingClasses + 1];
                int ca[kMaxIsingClasses + 1], cb[kMaxIsingClasses + 1];
                gather(sa, da, ca);
                gather(sb, db, cb);
                const int va = sS[sa], vb = sS[sb];
                const u32x4 wa = rng_draw(a.seed, (uint32_t)sa, (uint32_t)rid, step, RNG_METROPOLIS, 0);
                const u32x4 wb = rng_draw(a.seed, (uint32_t)sb, (uint32_t)rid, step, RNG_METROPOLIS, 0);
                const bool fa = decide(sa, da, ca, va, wa);
                const bool fb = hb && decide(sb, db, cb, vb, wb);
                accepted += (unsigned)fa + (unsigned)fb;
                if (fa) sS[sa] = (int8_t)(-va);
                if (fb) sS[sb] = (int8_t)(-vb);
            }
            __syncthreads();
        }
    }
    for (int q = t; q < a.N; q += nt) a.s[(size_t)q * a.Rp + r] = sS[q];
    if (accepted) atomicAdd(&sAcc1, accepted);
    if (a.meas) {
        int bond[kMaxIsingClasses + 1], zs[kMaxIsingClasses];
        double mz = 0.0;
#pragma unroll
        for (int kk = 0; kk <= kMaxIsingClasses; ++kk) bond[kk] = 0;
#pragma unroll
        for (int kk = 0; kk < kMaxIsingClasses; ++kk) zs[kk] = 0;
        for (int slot = t; slot < a.N; slot += nt) {
            const uint32_t* nb = a.ell + (size_t)slot * a.D;
            const int v = sS[slot];
            int diff[kMaxIsingClasses + 1], cnt[kMaxIsingClasses + 1];
#pragma unroll
            for (int kk = 0; kk <= kMaxIsingClasses; ++kk) {
                diff[kk] = 0;
                cnt[kk] = 0;
            }
            uint64_t cp = 0, dp = 0;                   // NC = 0: a byte per class, as in gather
            for (int j = 0; j < a.D; ++j) {
                const uint32_t e = nb[j];
                const int cl = (int)(e >> 24);
                const int d = sS[e & 0xffffffu] != v ? 1 : 0;
                if constexpr (NC > 0) {
#pragma unroll
                    for (int kk = 1; kk <= KC; ++kk) {
                        const int on = kk == cl;
                        diff[kk] += d & on;
                        cnt[kk] += on;
                    }
                } else {
                    const uint64_t on = cl != 0 ? 1u : 0u;
                    const int sh = (8 * (cl - 1)) & 63;
                    cp += on << sh;
                    dp += (on & (uint64_t)d) << sh;
                }
            }
            if constexpr (NC == 0) {
#pragma unroll
                for (int kk = 1; kk <= kMaxIsingClasses; ++kk) {
                    cnt[kk] = (int)((cp >> (8 * (kk - 1))) & 0xffu);
                    diff[kk] = (int)((dp >> (8 * (kk - 1))) & 0xffu);
                }
            }
            const int spc = a.sp[slot];
#pragma unroll
            for (int kk = 1; kk <= KC; ++kk) bond[kk] += cnt[kk] - 2 * diff[kk];
#pragma unroll
            for (int kk = 0; kk < kMaxIsingClasses; ++kk) zs[kk] += v * (int)(kk == spc);
            mz += (double)(v > 0 ? a.w[slot] : -a.w[slot]);
        }
#pragma unroll
        for (int kk = 1; kk <= KC; ++kk)
            if (kk <= a.n_classes && bond[kk]) atomicAdd(&sCnt1[kk - 1], bond[kk]);
#pragma unroll
        for (int kk = 0; kk < kMaxIsingClasses; ++kk)
            if (kk < a.nsp && zs[kk]) atomicAdd(&sCnt1[kMaxIsingClasses + kk], zs[kk]);
        // M_z: the word kernel's fixed-order wave reduction, then the waves in order
        double m = mz;
        for (int off = 32; off > 0; off >>= 1) m += __shfl_xor(m, off);
        if ((t & 63) == 0) sMz1[t >> 6] = m;
    }
    __syncthreads();
    if (t == 0) {
        if (sAcc1)
            atomicAdd(a.acc + (size_t)(r % kAccCopies) * a.R + a.slot_of[r], (unsigned long long)sAcc1);
        if (a.meas) {
            double* out = a.partial + (size_t)r * a.rec;    // tile 0
            int o = 0;
            for (int kk = 0; kk < a.n_classes; ++kk) out[o++] = (double)sCnt1[kk];
            for (int kk = 0; kk < a.nsp; ++kk) out[o++] = (double)sCnt1[kMaxIsingClasses + kk];
            double mm = 0.0;
            for (int wv = 0; wv < (nt + 63) / 64; ++wv) mm += sMz1[wv];
            out[o] = mm;
        }
    }
}

template <int DM, int NC, bool TL>
__global__ void __launch_bounds__(1024) k_ising_resident1(IsingResArgs a) {
    ising_resident1_block<DM, NC, TL>(a, blockIdx.x);
}

// The same for the lattices of a batch: workgroup j runs replica items[j].y
// of lattice items[j].x, whose spins (and table) take the first part of the
// launch's dynamic LDS (sized for the batch's largest lattice).  The block
// size is the batch's largest; a lattice's sweeps do not depend on it.
template <int DM, int NC, bool TL>
__global__ void __launch_bounds__(1024) k_ising_resident1_batch(const IsingResArgs* __restrict__ args,
                                                                const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    // read in place: a private copy would put its indexed arrays (colour
    // bounds, class strides) in scratch
    ising_resident1_block<DM, NC, TL>(args[it.x], it.y);
}

// ---------------------------------------------------------------------------
// Several sites per wavefront, for contexts with few replicas: LPS lanes per
// site (XYZ: a.lps, any count from the replicas up, 16 or 32 by default;
// Ising: 1 << a.lps_shift = 8, 16 or 32 lanes of 4 replicas), so a wave
// updates 64 / LPS whole sites per step instead of
// one site with mostly idle lanes.  Each lane's replica, Philox counter,
// neighbour order and arithmetic are those of the one-site-per-wave kernels,
// and the pooled Marsaglia proposal pools within the lanes of one site (the
// same replicas as a 64-lane pool when R <= LPS), so results are bit-identical.
// The site differs between lane groups: neighbour rows are addressed per lane.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float plane_ldv(Plane p, uint32_t voff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(p, (int)voff, 0, 0));
}
__device__ __forceinline__ void plane_stv(Plane p, uint32_t voff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), p, (int)voff, 0, 0);
}

// gather_row with a per-lane row (the lanes kernels)
template <int DM>
__device__ __forceinline__ void gather_row_lanes(Plane px, Plane py, Plane pz, const uint32_t (&nb)[DM],
                                                 uint32_t row_bytes, uint32_t roff, RowVals<DM>& v) {
#pragma unroll
    for (int k = 0; k < DM; ++k) {
        const uint32_t e = nb[k];
        const uint32_t vo = (e & 0xffffffu) * row_bytes + roff;
        v.cl[k] = e >> 24;
        v.x[k] = plane_ldv(px, vo);
        v.y[k] = plane_ldv(py, vo);
        v.z[k] = plane_ldv(pz, vo);
    }
}

template <int DM, bool LOWER = false>
__device__ __forceinline__ void local_field_lanes(Plane px, Plane py, Plane pz, const uint32_t* nb,
                                                  int D, uint32_t row_bytes, uint32_t roff,
                                                  const float4* sJ4, float& hx, float& hy, float& hz,
                                                  int c_begin = 0, float* lower = nullptr) {
    hx = 0.f;
    hy = 0.f;
    hz = 0.f;
    float lx = 0.f, ly = 0.f, lz = 0.f;   // LOWER: as in local_field
    if constexpr (DM > 0) {
        constexpr int CH = field_chunk(DM);
#pragma unroll
        for (int c0 = 0; c0 < DM; c0 += CH) {
            float vx[CH], vy[CH], vz[CH];
            uint32_t cl[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) {
                const uint32_t e = nb[c0 + k];
                const uint32_t vo = (e & 0xffffffu) * row_bytes + roff;
                cl[k] = e >> 24;
                vx[k] = plane_ldv(px, vo);
                vy[k] = plane_ldv(py, vo);
                vz[k] = plane_ldv(pz, vo);
            }
#pragma unroll
            for (int k = 0; k < CH; ++k) {
                const float4 j = sJ4[cl[k]];
                hx = fmaf(j.x, vx[k], hx);
                hy = fmaf(j.y, vy[k], hy);
                hz = fmaf(j.z, vz[k], hz);
                if constexpr (LOWER) {
                    const bool lo = (int)(nb[c0 + k] & 0xffffffu) < c_begin;
                    lx = fmaf(lo ? j.x : 0.f, vx[k], lx);
                    ly = fmaf(lo ? j.y : 0.f, vy[k], ly);
                    lz = fmaf(lo ? j.z : 0.f, vz[k], lz);
                }
            }
            if (CH < DM) __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        for (int k = 0; k < D; ++k) {
            const uint32_t e = nb[k];
            const uint32_t vo = (e & 0xffffffu) * row_bytes + roff;
            const float4 j = sJ4[e >> 24];
            const float vx = plane_ldv(px, vo), vy = plane_ldv(py, vo), vz = plane_ldv(pz, vo);
            hx = fmaf(j.x, vx, hx);
            hy = fmaf(j.y, vy, hy);
            hz = fmaf(j.z, vz, hz);
            if constexpr (LOWER) {
                const bool lo = (int)(e & 0xffffffu) < c_begin;
                lx = fmaf(lo ? j.x : 0.f, vx, lx);
                ly = fmaf(lo ? j.y : 0.f, vy, ly);
                lz = fmaf(lo ? j.z : 0.f, vz, lz);
            }
        }
    }
    if constexpr (LOWER) {
        lower[0] = lx;
        lower[1] = ly;
        lower[2] = lz;
    }
}

// marsaglia_pooled restricted to the lanes of `gmask` (one site's lanes); the
// pool slots of the group start at gbase
__device__ __forceinline__ void marsaglia_pooled_group(u32x4 w, bool active, uint64_t gmask, int gbase,
                                                       uint64_t seed, uint32_t id, uint32_t r,
                                                       uint32_t step, uint64_t* pool, float& x,
                                                       float& y, float& z) {
    uint32_t x0, y0, x1, y1;
    candidates(w, x0, y0, x1, y1);
    const bool ok0 = in_disc(x0, y0), ok1 = in_disc(x1, y1);
    const uint64_t needy = __ballot(active && !ok0) & gmask;
    const uint64_t donors = __ballot(active && ok1) & gmask;
    const int lane = lane_id();
    const uint64_t below = lanes_below(lane) & gmask;
    if (active && ok1) pool[gbase + __popcll(donors & below)] = ((uint64_t)x1 << 32) | y1;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    uint32_t ux = x0, uy = y0;
    bool have = ok0;
    if (active && !ok0) {
        const int j = __popcll(needy & below);
        if (j < __popcll(donors)) {
            const uint64_t v = pool[gbase + j];
            ux = (uint32_t)(v >> 32);
            uy = (uint32_t)v;
            have = true;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (!active) return;
    if (have) marsaglia_point(ux, uy, x, y, z);
    else marsaglia_retry(seed, id, r, step, RNG_METROPOLIS, 1, x, y, z);
}

template <bool ANGLE, int DM, bool MEAS>
__device__ __forceinline__ void xyz_sweep_lanes_block(const XyzArgs& a, int bb) {
    __shared__ float4 sJ4[kMaxClasses + 1];
    __shared__ float sZ[3 * kMaxSpecies];
    __shared__ float sD[3 * kMaxSpecies];
    __shared__ uint64_t sPool[kWavesPerBlock][kWave];
    __shared__ unsigned sAcc[kBlock];
    const bool ZEE = a.zee != 0, SIA = a.sia != 0;
    stage_couplings(sJ4, a.jclass, a.n_classes);
    if (ZEE)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sZ[t] = a.zvec[t];
    if (SIA)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sD[t] = a.dsia[t];
    __syncthreads();

    const int w = wave_id(), lane = lane_id();
    const int L = a.lps, S = a.sps;                            // lanes per site, sites per step
    const int sub = lane / L, r = lane - sub * L;              // lanes from S * L idle (sub == S)
    const bool lok = sub < S;
    const uint64_t gmask = L >= kWave ? ~0ull : (((1ull << L) - 1ull) << (sub * L));
    const int gbase = sub * L;
    const uint32_t roff = 4u * (uint32_t)r, rowb = 4u * (uint32_t)a.Rp;
    const Plane px = plane_of(a.sx, a.plane_bytes), py = plane_of(a.sy, a.plane_bytes),
                pz = plane_of(a.sz, a.plane_bytes);
    const Plane pp = plane_of(ANGLE ? a.phi : a.sx, a.plane_bytes);
    const bool rok = lok && r < a.R;
    const float beta = rok ? a.beta_slot[a.slot_of[r]] : 0.f;
    const uint32_t step = a.ctr[0] + a.step_off;
    uint64_t* pool = sPool[w];
    unsigned accepted = 0;
    double m[kMeasWords] = {0, 0, 0, 0, 0, 0};
    const int s0 = a.c_begin + (bb * kWavesPerBlock + w) * a.spw * S;
    // the lane's site at step base (an idle lane group reads a valid row)
    auto lane_slot = [&](int base) { return lok && base + sub < a.c_end ? base + sub : base; };
    constexpr bool ROWS = DM > 0 && DM <= 8;                   // rows held in registers, one step ahead
    uint32_t row_n[ROWS ? DM : 1];
    if constexpr (ROWS)
        if (s0 < a.c_end) load_row_lanes<DM>(a.ell + (size_t)lane_slot(s0) * a.D, row_n);
    for (int q = 0; q < a.spw; ++q) {
        const int base = s0 + q * S;                           // wave-uniform
        if (base >= a.c_end) break;
        const bool sv = lok && base + sub < a.c_end;
        const int slot = lane_slot(base);
        const bool rv = rok && sv;
        const uint32_t vo = (uint32_t)slot * rowb + roff;
        float x = plane_ldv(px, vo), y = plane_ldv(py, vo), z = plane_ldv(pz, vo);
        float cphi = ANGLE ? plane_ldv(pp, vo) : 0.f;
        const uint32_t* nb = a.ell + (size_t)slot * a.D;
        float hx, hy, hz;
        float hl[3] = {0.f, 0.f, 0.f};   // MEAS: the exchange field from earlier colours
        u32x4 wr;
        if constexpr (ROWS) {
            uint32_t row[DM];
#pragma unroll
            for (int k = 0; k < DM; ++k) row[k] = row_n[k];
            RowVals<DM> v;
            gather_row_lanes<DM>(px, py, pz, row, rowb, roff, v);
            // the next step's rows (the last step reads its own again)
            const int ns = q + 1 < a.spw && base + S < a.c_end ? lane_slot(base + S) : slot;
            load_row_lanes<DM>(a.ell + (size_t)ns * a.D, row_n);
            wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, step, RNG_METROPOLIS, 0);
            asm volatile("" : "+v"(wr.x), "+v"(wr.y), "+v"(wr.z), "+v"(wr.w));   // as in xyz_sweep_block
            sum_row<DM>(v, sJ4, hx, hy, hz);
            if constexpr (MEAS) sum_row_lower<DM>(v, row, a.c_begin, sJ4, hl[0], hl[1], hl[2]);
        } else {
            if constexpr (MEAS)
                local_field_lanes<DM, true>(px, py, pz, nb, a.D, rowb, roff, sJ4, hx, hy, hz, a.c_begin, hl);
            else
                local_field_lanes<DM>(px, py, pz, nb, a.D, rowb, roff, sJ4, hx, hy, hz);
            wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, step, RNG_METROPOLIS, 0);
        }
        const int sp = (ZEE || SIA) ? a.sp[slot] : 0;
        if (ZEE) {
            hx += sZ[3 * sp + 0];
            hy += sZ[3 * sp + 1];
            hz += sZ[3 * sp + 2];
        }
        bool reflected = false;
        if (a.fuse) {
            bool sel = true;
            if (a.frac < 1.0f) {
                const u32x4 wo = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, step, RNG_OVRR, a.pass);
                sel = (float)(wo.x >> 8) * 0x1p-24f < a.frac;
            }
            if (sel) {
                reflected = reflect_about(hx, hy, hz, x, y, z);
                if (ANGLE && reflected) cphi = atan2f(y, x);
            }
        }
        const float eta = eta24f(wr.z);
        float nx = 0.f, ny = 0.f, nz = 1.f, nphi = 0.f;
        if (ANGLE) {
            angle_move(wr, a.dphi, cphi, nx, ny, nz, nphi);
        } else {
            marsaglia_pooled_group(wr, rv, gmask, gbase, a.seed, (uint32_t)slot, (uint32_t)r, step, pool,
                                   nx, ny, nz);
        }
        float dE = delta_field(nx, ny, nz, x, y, z, hx, hy, hz);
        if (SIA) dE += delta_sia(sD + 3 * sp, nx, ny, nz, x, y, z);
        const bool acc = rv && -beta * dE > __logf(eta);
        if (acc) {
            plane_stv(px, vo, nx);
            plane_stv(py, vo, ny);
            plane_stv(pz, vo, nz);
            if (ANGLE) plane_stv(pp, vo, nphi);
            ++accepted;
        } else if (rv && reflected) {
            plane_stv(px, vo, x);
            plane_stv(py, vo, y);
            plane_stv(pz, vo, z);
            if (ANGLE) plane_stv(pp, vo, cphi);
        }
        if constexpr (MEAS) {
            if (sv) {                                  // an idle lane group books nothing
                if (acc) {
                    x = nx;
                    y = ny;
                    z = nz;
                }
                book_site(m, a, ZEE, SIA, sZ, sD, sp, slot, x, y, z, hl);
            }
        }
    }
    book_acceptance(sAcc, 1, &accepted, r, a.R, a.slot_of, a.acc);
    if constexpr (MEAS) {
        __shared__ double sMeas[kMeasWords * kBlock];
        if (!a.mrec)
            book_measurement(sMeas, m, L, S, 0, a);
        else if (a.c_end > a.c_begin)
            book_records(sMeas, m, L, S, bb, a);
    }
}

template <bool ANGLE, int DM, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_xyz_sweep_lanes(XyzArgs a) {
    xyz_sweep_lanes_block<ANGLE, DM, MEAS>(a, block_index(a.nblocks, a.xcd_map, 1, a.nchunk));
}

template <bool ANGLE, int DM, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_xyz_sweep_lanes_batch(const XyzArgs* __restrict__ args,
                                                                 const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const XyzArgs a = args[it.x];
    xyz_sweep_lanes_block<ANGLE, DM, MEAS>(a, it.y);
}

template <int DM>
__device__ __forceinline__ void ising_sweep_lanes_block(const IsingArgs& a, int bb) {
    __shared__ int sStride[kMaxIsingClasses + 1];
    __shared__ uint32_t sThr[kIsingLdsTable];
    __shared__ unsigned sAcc[kBlock * 4];
    for (int t = threadIdx.x; t <= a.n_classes; t += kBlock) sStride[t] = a.stride[t];
    if (a.table_in_lds)
        for (int t = threadIdx.x; t < a.R * a.table_per_slot; t += kBlock) sThr[t] = a.thr[t];
    __syncthreads();
    const int w = wave_id(), lane = lane_id();
    const int ls = a.lps_shift, S = kWave >> ls;
    const int sub = lane >> ls;
    const int rid = lane & ((1 << ls) - 1);                     // Philox id of the lane's 4 replicas
    const int r0 = 4 * rid;
    const bool lv = r0 < a.Rp;
    int tbase[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = r0 + q;
        tbase[q] = (r < a.R) ? a.slot_of[r] * a.table_per_slot : -1;
    }
    unsigned accepted[4] = {0, 0, 0, 0};
    const uint32_t step = a.ctr[0] + a.step_off;
    const int s0 = a.c_begin + (bb * kWavesPerBlock + w) * a.spw * S;
    for (int q = 0; q < a.spw; ++q) {
        const int base = s0 + q * S;
        if (base >= a.c_end) break;
        const int slot = base + sub;
        if (!lv || slot >= a.c_end) continue;                   // no wave-level operation below
        uint32_t d[kMaxIsingClasses + 1];
        int cnt[kMaxIsingClasses + 1];
#pragma unroll
        for (int k = 0; k <= kMaxIsingClasses; ++k) {
            d[k] = 0;
            cnt[k] = 0;
        }
        const uint32_t* nb = a.ell + (size_t)slot * a.D;
        uint32_t* own = reinterpret_cast<uint32_t*>(a.s + (size_t)slot * a.Rp + r0);
        uint32_t v;
        u32x4 wr;
        if constexpr (DM > 0) {
            uint32_t nv[DM];
#pragma unroll
            for (int k = 0; k < DM; ++k)
                nv[k] = *reinterpret_cast<const uint32_t*>(a.s + (size_t)(nb[k] & 0xffffffu) * a.Rp + r0);
            v = *own;
            wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)rid, step, RNG_METROPOLIS, 0);
            asm volatile("" : "+v"(wr.x), "+v"(wr.y), "+v"(wr.z), "+v"(wr.w));   // as in ising_sweep_block
#pragma unroll
            for (int k = 0; k < DM; ++k) {
                const int c = (int)(nb[k] >> 24);
                class_add<(DM > 0 && DM <= 8)>(d, cnt, a.n_classes, c, (nv[k] >> 1) & 0x01010101u);
            }
        } else {
            for (int k = 0; k < a.D; ++k) {
                const uint32_t e = nb[k];
                const int c = (int)(e >> 24);
                if (c == 0) continue;
                const uint32_t nbv = *reinterpret_cast<const uint32_t*>(a.s + (size_t)(e & 0xffffffu) * a.Rp + r0);
                class_add<(DM > 0 && DM <= 8)>(d, cnt, a.n_classes, c, (nbv >> 1) & 0x01010101u);
            }
            v = *own;
            wr = rng_draw(a.seed, (uint32_t)slot, (uint32_t)rid, step, RNG_METROPOLIS, 0);
        }
        const uint32_t flip = ising_flip(a, sStride, sThr, tbase, d, cnt, a.sp[slot], v, wr);
#pragma unroll
        for (int b = 0; b < 4; ++b) accepted[b] += (flip >> (8 * b + 1)) & 1u;
        if (flip) *own = v ^ flip;
    }
    book_acceptance(sAcc, 4, accepted, r0, a.R, a.slot_of, a.acc);
}

template <int DM>
__global__ __launch_bounds__(kBlock) void k_ising_sweep_lanes(IsingArgs a) {
    ising_sweep_lanes_block<DM>(a, block_index(a.nblocks, a.xcd_map, 1, a.nchunk));
}

template <int DM>
__global__ __launch_bounds__(kBlock) void k_ising_sweep_lanes_batch(const IsingArgs* __restrict__ args,
                                                                   const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const IsingArgs a = args[it.x];
    ising_sweep_lanes_block<DM>(a, it.y);
}

// Sum over the workgroup's waves of each lane's value (waves in order), written
// by wave 0 to out[k]; called by every thread of the workgroup.
__device__ __forceinline__ void block_record(double* buf /* [kWavesPerBlock][kWave] */, double v,
                                             double* out_k) {
    const int w = wave_id(), lane = lane_id();
    buf[w * kWave + lane] = v;
    __syncthreads();
    if (w == 0) {
        double t = buf[lane];
        for (int q = 1; q < kWavesPerBlock; ++q) t += buf[q * kWave + lane];
        *out_k = t;
    }
    __syncthreads();
}

// block_record of the four replicas of a lane at once (Ising, four per
// lane): out_k[b * stride] gets replica b's sum; lanes that hold no replica
// (valid false) write nothing
__device__ __forceinline__ void block_record4(double* buf /* [kWavesPerBlock][kWave][4] */,
                                              const double (&v)[4], bool valid, double* out_k,
                                              int stride) {
    const int w = wave_id(), lane = lane_id();
#pragma unroll
    for (int b = 0; b < 4; ++b) buf[(w * kWave + lane) * 4 + b] = v[b];
    __syncthreads();
    if (w == 0 && valid) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            double t = buf[lane * 4 + b];
            for (int q = 1; q < kWavesPerBlock; ++q) t += buf[(q * kWave + lane) * 4 + b];
            out_k[b * stride] = t;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// measurement partials: workgroup = (tile of sites, replica group of 64), the
// tile split between its waves and their sums added in the workgroup
// record per (tile, replica): XYZ {E_exch, E_z, E_sia, Mx, My, Mz};
// Ising {bonds_1..K, zsum_1..nsp, Mz} (integers held exactly in FP64)
// ---------------------------------------------------------------------------
template <bool ZEE, bool SIA>
__device__ __forceinline__ void xyz_measure_block(const MeasArgs& a, int blk) {
    constexpr int NB = 4;
    __shared__ float sJ[3 * (kMaxClasses + 1)];
    __shared__ float sZ[3 * kMaxSpecies];
    __shared__ float sD[3 * kMaxSpecies];
    for (int t = threadIdx.x; t < 3 * (a.n_classes + 1); t += kBlock) sJ[t] = a.jclass[t];
    if (ZEE)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sZ[t] = a.zvec[t];
    if (SIA)
        for (int t = threadIdx.x; t < 3 * a.nsp; t += kBlock) sD[t] = a.dsia[t];
    __shared__ double sRed[kWavesPerBlock * kWave];
    __syncthreads();
    const int g = blk % a.G, tile = blk / a.G;             // workgroup-uniform
    const int r = g * kWave + lane_id();
    const uint32_t roff = 4u * (uint32_t)r, rowb = 4u * (uint32_t)a.Rp;
    const Plane px = plane_of(a.sx, a.plane_bytes), py = plane_of(a.sy, a.plane_bytes),
                pz = plane_of(a.sz, a.plane_bytes);
    const ConstF32 wtab = (ConstF32)a.w;
    const ConstU8 sptab = (ConstU8)a.sp;
    double v[6] = {0, 0, 0, 0, 0, 0};
    double ss[3] = {0, 0, 0};
    const int per_wave = (a.tile_sites + kWavesPerBlock - 1) / kWavesPerBlock;
    const int s_begin = tile * a.tile_sites + wave_id() * per_wave;
    const int s_end = min(min(a.N, (tile + 1) * a.tile_sites), s_begin + per_wave);
    // one site's own spin and the first NB entries of its measurement row
    // (gathering 8 at a time was slower at 16^3 pyrochlore cells: 223 vs 182 us)
    struct SiteM {
        float x, y, z;
        float vx[NB], vy[NB], vz[NB];
        uint32_t cl[NB];
    };
    const ConstRow rown = const_row(a.ell_m_n);
    auto chunk = [&](ConstRow nb, int n, int k0, float* vx, float* vy, float* vz, uint32_t* cl) {
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const uint32_t e = (k0 + k < n) ? nb[k0 + k] : 0u;   // 0: class 0, zero coupling
            const uint32_t o = (e & 0xffffffu) * rowb;
            cl[k] = e >> 24;
            vx[k] = plane_ld(px, roff, o);
            vy[k] = plane_ld(py, roff, o);
            vz[k] = plane_ld(pz, roff, o);
        }
    };
    auto load = [&](int slot, SiteM& d) {
        const uint32_t so = (uint32_t)slot * rowb;
        d.x = plane_ld(px, roff, so);
        d.y = plane_ld(py, roff, so);
        d.z = plane_ld(pz, roff, so);
        chunk(const_row(a.ell_m + (size_t)slot * a.D_m), (int)rown[slot], 0, d.vx, d.vy, d.vz, d.cl);
    };
    auto accumulate = [&](int slot, const SiteM& d) {
        const float x = d.x, y = d.y, z = d.z;
        float hx = 0.f, hy = 0.f, hz = 0.f;
        auto add = [&](const float* vx, const float* vy, const float* vz, const uint32_t* cl) {
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                hx = fmaf(sJ[3 * cl[k] + 0], vx[k], hx);
                hy = fmaf(sJ[3 * cl[k] + 1], vy[k], hy);
                hz = fmaf(sJ[3 * cl[k] + 2], vz[k], hz);
            }
        };
        add(d.vx, d.vy, d.vz, d.cl);
        const ConstRow nb = const_row(a.ell_m + (size_t)slot * a.D_m);   // wave-uniform row
        const int n = (int)rown[slot];
        for (int k0 = NB; k0 < n; k0 += NB) {
            float vx[NB], vy[NB], vz[NB];
            uint32_t cl[NB];
            chunk(nb, n, k0, vx, vy, vz, cl);
            add(vx, vy, vz, cl);
        }
        v[0] += (double)((float)a.bond_factor * dot3(x, y, z, hx, hy, hz));
        const int sp = (ZEE || SIA) ? (int)sptab[slot] : 0;
        if (ZEE) v[1] += (double)dot3(x, y, z, sZ[3 * sp], sZ[3 * sp + 1], sZ[3 * sp + 2]);
        if (SIA) v[2] += (double)dot3(sD[3 * sp], sD[3 * sp + 1], sD[3 * sp + 2], x * x, y * y, z * z);
        const float wgt = wtab[slot];
        v[3] += (double)(wgt * x);
        v[4] += (double)(wgt * y);
        v[5] += (double)(wgt * z);
        // spiral state: sum of S_i x S_j over the site's partners (src/get_spiral_state.f90:88-112)
        if (a.spiral_cap) {
            const int* pj = a.spiral + (size_t)slot * a.spiral_cap;
            for (int k = 0; k < a.spiral_cap; ++k) {
                const int j = pj[k];
                if (j < 0) break;
                const uint32_t o = (uint32_t)j * rowb;
                const double bx = plane_ld(px, roff, o), by = plane_ld(py, roff, o), bz = plane_ld(pz, roff, o);
                ss[0] += (double)y * bz - (double)z * by;
                ss[1] += (double)z * bx - (double)x * bz;
                ss[2] += (double)x * by - (double)y * bx;
            }
        }
    };
    // two sites at a time: the second site's loads are in flight while the
    // first is summed (sites are still added in order)
    for (int slot = s_begin; slot < s_end; slot += 2) {
        const bool two = slot + 1 < s_end;                  // wave-uniform
        SiteM A, B;
        load(slot, A);
        load(two ? slot + 1 : slot, B);
        accumulate(slot, A);
        if (two) accumulate(slot + 1, B);
    }
    double* out = a.partial + ((size_t)tile * a.Rp + r) * a.rec;
#pragma unroll
    for (int k = 0; k < 6; ++k) block_record(sRed, v[k], out + k);
    if (a.spiral_cap) {
#pragma unroll
        for (int k = 0; k < 3; ++k) block_record(sRed, ss[k], out + 6 + k);
    }
}

template <bool ZEE, bool SIA>
__global__ __launch_bounds__(kBlock) void k_xyz_measure(MeasArgs a) {
    xyz_measure_block<ZEE, SIA>(a, blockIdx.x);
}

template <bool ZEE, bool SIA>
__global__ __launch_bounds__(kBlock) void k_xyz_measure_batch(const MeasArgs* __restrict__ args,
                                                             const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const MeasArgs a = args[it.x];
    xyz_measure_block<ZEE, SIA>(a, it.y);
}

// Ising measurement with four replicas per lane (one 32-bit word of int8
// spins, as the sweep holds them): workgroup = (tile of sites, group of 256
// replicas).  Per site and coupling class the bonds whose two spins differ
// are counted per byte (SWAR: ((s_j ^ s_i) >> 1) & 0x01010101 is 1 per byte
// where they differ), then added into 16-bit fields (even and odd bytes
// apart); a class's bond sum is then (its bonds) - 2 (the differing ones).
// The spin sum per species comes from the down spins the same way.  All
// integers, exact in any order.
__device__ __forceinline__ void ising_measure_block(const MeasArgs& a, int blk) {
    __shared__ double sRed[kWavesPerBlock * kWave * 4];
    const int G4 = (a.Rp + 4 * kWave - 1) / (4 * kWave);
    const int g4 = blk % G4, tile = blk / G4;             // workgroup-uniform
    const int r0 = 4 * (g4 * kWave + lane_id());
    const bool lv = r0 < a.Rp;
    const ConstF32 wtab = (ConstF32)a.w;
    const ConstU8 sptab = (ConstU8)a.sp;
    const ConstRow rown = const_row(a.ell_m_n);
    uint32_t de[kMaxIsingClasses + 1], dodd[kMaxIsingClasses + 1];   // differing bonds
    int nbond[kMaxIsingClasses + 1];                                 // bonds (wave-uniform)
    uint32_t ze[kMaxIsingClasses], zodd[kMaxIsingClasses];           // down spins
    int nsite[kMaxIsingClasses];                                     // sites (wave-uniform)
#pragma unroll
    for (int k = 0; k <= kMaxIsingClasses; ++k) {
        de[k] = dodd[k] = 0u;
        nbond[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < kMaxIsingClasses; ++k) {
        ze[k] = zodd[k] = 0u;
        nsite[k] = 0;
    }
    double mz[4] = {0.0, 0.0, 0.0, 0.0};
    const int per_wave = (a.tile_sites + kWavesPerBlock - 1) / kWavesPerBlock;
    const int s_begin = tile * a.tile_sites + wave_id() * per_wave;
    const int s_end = min(min(a.N, (tile + 1) * a.tile_sites), s_begin + per_wave);
    for (int slot = s_begin; slot < s_end; ++slot) {
        const uint32_t v = lv ? *reinterpret_cast<const uint32_t*>(a.s + (size_t)slot * a.Rp + r0) : 0u;
        const ConstRow nb = const_row(a.ell_m + (size_t)slot * a.D_m);   // wave-uniform row
        const int n = (int)rown[slot];
        uint32_t d[kMaxIsingClasses + 1];                                // this site's, per byte
#pragma unroll
        for (int k = 0; k <= kMaxIsingClasses; ++k) d[k] = 0u;
        for (int k0 = 0; k0 < n; k0 += 4) {
            uint32_t sj[4];
            int cl[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t e = (k0 + k < n) ? nb[k0 + k] : 0u;   // 0: class 0, no bond
                cl[k] = (int)(e >> 24);
                sj[k] = lv ? *reinterpret_cast<const uint32_t*>(a.s + (size_t)(e & 0xffffffu) * a.Rp + r0) : 0u;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                class_add<true>(d, nbond, a.n_classes, cl[k], ((sj[k] ^ v) >> 1) & 0x01010101u);
            }
        }
#pragma unroll
        for (int kk = 1; kk <= kMaxIsingClasses; ++kk) {
            de[kk] += d[kk] & 0x00ff00ffu;
            dodd[kk] += (d[kk] >> 8) & 0x00ff00ffu;
        }
        const int sp = (int)sptab[slot];
        const uint32_t down = (v >> 1) & 0x01010101u;
#pragma unroll
        for (int kk = 0; kk < kMaxIsingClasses; ++kk)
            if (kk == sp) {
                ze[kk] += down & 0x00ff00ffu;
                zodd[kk] += (down >> 8) & 0x00ff00ffu;
                nsite[kk] += 1;
            }
        const float w = wtab[slot];
#pragma unroll
        for (int b = 0; b < 4; ++b) mz[b] += (double)(((down >> (8 * b)) & 1u) ? -w : w);
    }
    // per replica b of the lane: 16-bit field b/2 of the even (b even) or odd word
    auto field = [](uint32_t even, uint32_t odd, int b) -> int {
        const uint32_t w = (b & 1) ? odd : even;
        return (int)((b & 2) ? (w >> 16) : (w & 0xffffu));
    };
    // each bond was counted once, from its endpoint with the smaller site index
    // (bond_factor 2), or on full rows (1)
    double* out = a.partial + ((size_t)tile * a.Rp + r0) * a.rec;
    int o = 0;
    double t[4];
#pragma unroll
    for (int k = 1; k <= kMaxIsingClasses; ++k)
        if (k <= a.n_classes) {
#pragma unroll
            for (int b = 0; b < 4; ++b)
                t[b] = (double)((nbond[k] - 2 * field(de[k], dodd[k], b)) * a.bond_factor);
            block_record4(sRed, t, lv, out + o++, a.rec);
        }
#pragma unroll
    for (int k = 0; k < kMaxIsingClasses; ++k)
        if (k < a.nsp) {
#pragma unroll
            for (int b = 0; b < 4; ++b) t[b] = (double)(nsite[k] - 2 * field(ze[k], zodd[k], b));
            block_record4(sRed, t, lv, out + o++, a.rec);
        }
    block_record4(sRed, mz, lv, out + o, a.rec);
}

// Ising measurement with one replica per lane (int8 loads): for fewer than
// 128 replicas, where a four-replica lane layout would leave most lanes idle
// (a batch of small lattices ran 5% slower with it).
__device__ __forceinline__ void ising_measure_narrow_block(const MeasArgs& a, int blk) {
    __shared__ double sRed[kWavesPerBlock * kWave];
    const int g = blk % a.G, tile = blk / a.G;             // workgroup-uniform
    const int r = g * kWave + lane_id();
    const ConstF32 wtab = (ConstF32)a.w;
    const ConstU8 sptab = (ConstU8)a.sp;
    // bond and spin sums are integers: counted in int32 (a lane sees at most
    // a tile's sites times their rows) and written as doubles, exact; the
    // class and species loops stop at the context's counts (uniform), so a
    // one-class lattice pays one compare-and-add per bond
    int bonds[kMaxIsingClasses + 1];
    int zs[kMaxIsingClasses];
#pragma unroll
    for (int k = 0; k <= kMaxIsingClasses; ++k) bonds[k] = 0;
#pragma unroll
    for (int k = 0; k < kMaxIsingClasses; ++k) zs[k] = 0;
    double mz = 0.0;
    const int per_wave = (a.tile_sites + kWavesPerBlock - 1) / kWavesPerBlock;
    const int s_begin = tile * a.tile_sites + wave_id() * per_wave;
    const int s_end = min(min(a.N, (tile + 1) * a.tile_sites), s_begin + per_wave);
    // four sites of the walk at a time, so that their loads are in flight
    // together (a walk of a few sites otherwise waits for memory once per
    // site); the magnetisation is added in site order
    constexpr int U = 4;
    for (int slot0 = s_begin; slot0 < s_end; slot0 += U) {
        int si[U], n[U], nmax = 0;
        ConstRow nb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = slot0 + u < s_end;                // wave-uniform
            const int slot = in ? slot0 + u : slot0;
            si[u] = a.s[(size_t)slot * a.Rp + r];
            nb[u] = const_row(a.ell_m + (size_t)slot * a.D_m);   // wave-uniform rows
            n[u] = in ? (int)const_row(a.ell_m_n)[slot] : 0;
            nmax = max(nmax, n[u]);
        }
        for (int k0 = 0; k0 < nmax; k0 += 4) {
            int sj[U][4], cl[U][4];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t e = (k0 + k < n[u]) ? nb[u][k0 + k] : 0u;   // 0: class 0, no bond
                    cl[u][k] = (int)(e >> 24);
                    sj[u][k] = a.s[(size_t)(e & 0xffffffu) * a.Rp + r];
                }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int prod = si[u] * sj[u][k];
#pragma unroll
                    for (int kk = 1; kk <= kMaxIsingClasses; ++kk) {
                        if (kk > a.n_classes) break;
                        bonds[kk] += (kk == cl[u][k]) ? prod : 0;
                    }
                }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (slot0 + u >= s_end) break;
            const int slot = slot0 + u;
            const int sp = (int)sptab[slot];
#pragma unroll
            for (int kk = 0; kk < kMaxIsingClasses; ++kk) {
                if (kk >= a.nsp) break;
                zs[kk] += (kk == sp) ? si[u] : 0;
            }
            mz += (double)(wtab[slot] * (float)si[u]);
        }
    }
    // each bond was counted once, from its endpoint with the smaller slot
    double* out = a.partial + ((size_t)tile * a.Rp + r) * a.rec;
    int o = 0;
#pragma unroll
    for (int k = 1; k <= kMaxIsingClasses; ++k)
        if (k <= a.n_classes) block_record(sRed, (double)bonds[k] * (double)a.bond_factor, out + o++);
#pragma unroll
    for (int k = 0; k < kMaxIsingClasses; ++k)
        if (k < a.nsp) block_record(sRed, (double)zs[k], out + o++);
    block_record(sRed, mz, out + o);
}

__global__ __launch_bounds__(kBlock) void k_ising_measure(MeasArgs a) {
    ising_measure_block(a, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_ising_measure_narrow(MeasArgs a) {
    ising_measure_narrow_block(a, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_ising_measure_batch(const MeasArgs* __restrict__ args,
                                                               const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const MeasArgs a = args[it.x];
    ising_measure_block(a, it.y);
}

__global__ __launch_bounds__(kBlock) void k_ising_measure_narrow_batch(const MeasArgs* __restrict__ args,
                                                               const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const MeasArgs a = args[it.x];
    ising_measure_narrow_block(a, it.y);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of one value per thread of the workgroup in a fixed order (the four
// waves' lanes in order, then the wave); the result is valid in wave 0.
// Called by every thread of the workgroup.
__device__ __forceinline__ double block_sum(double* buf /* [kBlock] */, double v) {
    buf[threadIdx.x] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < kWave) {
        for (int q = 0; q < kWavesPerBlock; ++q) t += buf[q * kWave + threadIdx.x];
        t = wave_sum(t);
    }
    __syncthreads();
    return t;
}

// One workgroup per replica: fixed-order reduction of the tile partials
// (thread t adds tiles t, t + kBlock, ...), energy assembly and the
// observable() moments.
__device__ __forceinline__ void finalize_block(const FinArgs& a, int blk) {
    __shared__ double sFin[kBlock];
    const int r = blk;
    const int tid = threadIdx.x, lane = lane_id();
    double ex = 0.0, ez = 0.0, es = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    double psi = 0.0;
    if (a.model == 1 && a.mrec) {
        // fused measurement with per-workgroup records: replica r is lane
        // r % 64 of the units bb = g + u * groups (g = r / 64) of every phase
        // whose bb lies below the stride, contiguous per phase; the terms are
        // integers, so any order is exact
        const int g = r / kWave, lane = r % kWave;
        const int cnt = (a.mrec_stride - g + a.mrec_groups - 1) / a.mrec_groups;
        double v[kMeasWords] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < a.mrec_phases; ++q) {
            const long long* p =
                a.mrec + (((size_t)q * a.mrec_groups + g) * kWave + lane) * a.mrec_pp * kMeasWords;
            for (int u = tid; u < cnt; u += kBlock) {
#pragma unroll
                for (int k = 0; k < kMeasWords; ++k) v[k] += (double)p[(size_t)u * kMeasWords + k];
            }
        }
        for (int k = 0; k < kMeasWords; ++k) v[k] = block_sum(sFin, v[k]);
        if (tid >= kWave) return;
        ex = v[0] / a.mscale[0];
        ez = v[1] / a.mscale[1];
        es = v[2] / a.mscale[2];
        mx = v[3] / a.mscale[3];
        my = v[4] / a.mscale[3];
        mz = v[5] / a.mscale[3];
    } else if (a.model == 1 && a.macc) {
        // fused measurement: exact fixed-point sums (lane c < kMeasCopies of
        // wave 0 reads copy c), cleared for the next one
        if (tid >= kWave) return;
        double v[kMeasWords] = {0, 0, 0, 0, 0, 0};
        unsigned long long* q = a.macc + ((size_t)lane * a.Rp + r) * kMeasWords;
        if (lane < kMeasCopies) {
#pragma unroll
            for (int k = 0; k < kMeasWords; ++k) {
                v[k] = (double)(long long)q[k];
                q[k] = 0ull;
            }
        }
#pragma unroll
        for (int k = 0; k < kMeasWords; ++k) v[k] = wave_sum(v[k]);
        ex = v[0] / a.mscale[0];
        ez = v[1] / a.mscale[1];
        es = v[2] / a.mscale[2];
        mx = v[3] / a.mscale[3];
        my = v[4] / a.mscale[3];
        mz = v[5] / a.mscale[3];
    } else if (a.model == 1) {
        const int nv = a.spiral_count > 0.0 ? 9 : 6;
        double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        // four tiles per thread in flight: their loads are issued before the
        // first is added (tiles t, t + kBlock, t + 2 kBlock, t + 3 kBlock in order)
        for (int t0 = tid; t0 < a.tiles; t0 += 4 * kBlock) {
            double w[4][9];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = t0 + u * kBlock;
                const double* p = a.partial + ((size_t)t * a.Rp + r) * a.rec;
#pragma unroll
                for (int k = 0; k < 9; ++k) w[u][k] = (t < a.tiles && k < nv) ? p[k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < 9; ++k) v[k] += w[u][k];
        }
        for (int k = 0; k < nv; ++k) v[k] = block_sum(sFin, v[k]);
        if (tid >= kWave) return;
        ex = v[0]; ez = v[1]; es = v[2]; mx = v[3]; my = v[4]; mz = v[5];
        // |sum S_i x S_j| / spiral_count (src/get_spiral_state.f90:114-117)
        if (nv == 9) psi = sqrt(v[6] * v[6] + v[7] * v[7] + v[8] * v[8]) / a.spiral_count;
    } else {
        for (int k = 0; k < a.rec; ++k) {
            double s = 0.0;
            for (int t = tid; t < a.tiles; t += kBlock) s += a.partial[((size_t)t * a.Rp + r) * a.rec + k];
            s = block_sum(sFin, s);
            if (k < a.n_classes) {
                ex += a.ising_j[k + 1] * s;
                if (tid == 0) a.bonds[(size_t)r * (a.n_classes + 1) + k + 1] = (long long)s;
            } else if (k < a.n_classes + a.nsp) {
                ez += a.ising_zc[k - a.n_classes] * s;
            } else {
                mz = s;
            }
        }
    }
    if (tid != 0) return;
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
        if (a.spiral_count > 0.0) {
            double* ms = a.mom_sp + 3 * (size_t)a.slot_of[r];
            const double p2 = psi * psi;
            ms[0] += psi;
            ms[1] += p2;
            ms[2] += p2 * p2;
        }
    }
    if (a.spiral_count > 0.0) a.psi[r] = psi;
}

__global__ __launch_bounds__(kBlock) void k_finalize(FinArgs a) {
    finalize_block(a, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_finalize_batch(const FinArgs* __restrict__ args,
                                                      const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const FinArgs a = args[it.x];
    finalize_block(a, it.y);
}

// ---------------------------------------------------------------------------
// replica exchange: one thread per ladder, slots permuted, spins untouched
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pt_swap_block(const PtArgs& a, int blk) {
    const int g = blk * blockDim.x + threadIdx.x;
    if (g >= a.n_groups) return;
    const int s0 = a.gstart[g], s1 = a.gstart[g + 1];
    const uint32_t swap_index = a.ctr[1] + a.pt_off;
    const int parity = a.parity >= 0 ? a.parity : (int)((a.ctr[2] + a.pt_off + 1u) & 1u);
    int nacc = 0;
    for (int k = s0 + parity; k + 1 < s1; k += 2) {
        const int ra = a.rep_at[k], rb = a.rep_at[k + 1];
        const double b = a.beta_slot[k], bp = a.beta_slot[k + 1];
        const double E = a.E_pt[ra], Ep = a.E_pt[rb];
        const double delta = -(bp - b) * (E - Ep);
        bool accept = delta > 0.0;
        if (!accept) {
            u32x4 w = rng_draw(a.seed, (uint32_t)k, (uint32_t)g, swap_index, RNG_PT, 0);
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

__global__ void k_pt_swap(PtArgs a) {
    pt_swap_block(a, blockIdx.x);
}

__global__ void k_pt_swap_batch(const PtArgs* __restrict__ args, const int2* __restrict__ items) {
    const int2 it = items[blockIdx.x];
    if (it.x < 0) return;
    const PtArgs a = args[it.x];
    pt_swap_block(a, it.y);
}

// ---------------------------------------------------------------------------
// initial states
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_init(InitArgs a) {
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (size_t)a.N * a.Rp) return;
    const int slot = (int)(t / a.Rp), r = (int)(t % a.Rp);
    if (r >= a.R) return;
    if (a.model == 0) {
        if (a.mode == 1) {
            a.s[t] = (int8_t)(-a.s[t]);
        } else {
            u32x4 w = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, a.step, RNG_INIT, 0);
            a.s[t] = (w.x >> 31) ? (int8_t)1 : (int8_t)-1;
        }
        return;
    }
    float x, y, z;
    if (!a.angle) {
        marsaglia_retry(a.seed, (uint32_t)slot, (uint32_t)r, a.step, RNG_INIT, 0, x, y, z);
        if (a.phi) a.phi[t] = 0.f;
    } else {
        u32x4 w = rng_draw(a.seed, (uint32_t)slot, (uint32_t)r, a.step, RNG_INIT, 0);
        float ph;
        if (a.mode == 1) angle_move(w, a.dphi, a.phi[t], x, y, z, ph);
        else angle_move(w, 3.14159265358979323846f, 0.0f, x, y, z, ph);
        a.phi[t] = ph;
    }
    a.sx[t] = x;
    a.sy[t] = y;
    a.sz[t] = z;
}

// ---------------------------------------------------------------------------
// local fields and site energies of one replica; spin extract / insert
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_local_fields(MeasArgs a, int r, float* h, float* e,
                                                         int zee, int sia, int model) {
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= a.N) return;
    const size_t row = (size_t)slot * a.Rp + r;
    float hx = 0.f, hy = 0.f, hz = 0.f, x, y, z;
    const uint32_t* nb = a.ell + (size_t)slot * a.D;
    if (model == 1) {
        for (int k = 0; k < a.D; ++k) {
            const uint32_t en = nb[k];
            const size_t jr = (size_t)(en & 0xffffffu) * a.Rp + r;
            const uint32_t c = en >> 24;
            hx = fmaf(a.jclass[3 * c + 0], a.sx[jr], hx);
            hy = fmaf(a.jclass[3 * c + 1], a.sy[jr], hy);
            hz = fmaf(a.jclass[3 * c + 2], a.sz[jr], hz);
        }
        x = a.sx[row];
        y = a.sy[row];
        z = a.sz[row];
    } else {
        for (int k = 0; k < a.D; ++k) {
            const uint32_t en = nb[k];
            const uint32_t c = en >> 24;
            hz = fmaf(a.jclass[3 * c + 2], (float)a.s[(size_t)(en & 0xffffffu) * a.Rp + r], hz);
        }
        x = 0.f;
        y = 0.f;
        z = (float)a.s[row];
    }
    h[3 * slot + 0] = hx;
    h[3 * slot + 1] = hy;
    h[3 * slot + 2] = hz;
    const int sp = a.sp[slot];
    float en = dot3(x, y, z, hx, hy, hz);
    if (zee) en += dot3(x, y, z, a.zvec[3 * sp], a.zvec[3 * sp + 1], a.zvec[3 * sp + 2]);
    if (sia && model == 1)
        en += dot3(a.dsia[3 * sp], a.dsia[3 * sp + 1], a.dsia[3 * sp + 2], x * x, y * y, z * z);
    e[slot] = en;
}

__global__ void k_extract(const float* plane, const int8_t* s, int N, int Rp, int r, float* out) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= N) return;
    const size_t i = (size_t)slot * Rp + r;
    out[slot] = plane ? plane[i] : (float)s[i];
}

__global__ void k_insert(float* plane, int8_t* s, int N, int Rp, int r0, int r1, const float* in) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = r1 - r0;
    if (t >= (size_t)N * nr) return;
    const int slot = (int)(t / nr), r = r0 + (int)(t % nr);
    const size_t i = (size_t)slot * Rp + r;
    if (plane) plane[i] = in[slot];
    else s[i] = in[slot] > 0.f ? (int8_t)1 : (int8_t)-1;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// the ELL width buckets that have unrolled instantiations (0 = runtime loop)
int ell_bucket(int D) {
    static const int b[] = {4, 6, 8, 12, 18, 24, 32};
    for (int x : b)
        if (D <= x) return x;
    return 0;
}

#define ETHER_DISPATCH_DM(D, CALL)        \
    switch (D) {                          \
        case 4: CALL(4); break;           \
        case 6: CALL(6); break;           \
        case 8: CALL(8); break;           \
        case 12: CALL(12); break;         \
        case 18: CALL(18); break;         \
        case 24: CALL(24); break;         \
        case 32: CALL(32); break;         \
        default: CALL(0); break;          \
    }

thread_local LaunchRecorder* g_recorder = nullptr;

static int sweep_units(int groups, int n_sites, int spw) {
    const int chunk = kWavesPerBlock * spw;
    return n_sites > 0 ? groups * ((n_sites + chunk - 1) / chunk) : 0;
}

template <class A>
static void record(int kind, int variant, int units, const A& a, int block = 0, int lds = 0) {
    LaunchRecord r;
    r.kind = kind;
    r.variant = variant;
    r.units = units;
    r.bytes = (int)sizeof(A);
    r.block = block;
    r.lds = lds;
    *reinterpret_cast<A*>(&r.u) = a;
    g_recorder->recs.push_back(r);
}

// sites a wave updates per step of its loop
static int sites_per_step(int lps_shift) { return kWave >> lps_shift; }

template <bool MEAS>
static void launch_xyz_sweep_t(const XyzArgs& a, bool angle, bool lanes, hipStream_t st) {
    dim3 grid(a.nblocks), block(kBlock);
    if (lanes) {
#define LA(DM) hipLaunchKernelGGL((k_xyz_sweep_lanes<true, DM, MEAS>), grid, block, 0, st, a)
#define LM(DM) hipLaunchKernelGGL((k_xyz_sweep_lanes<false, DM, MEAS>), grid, block, 0, st, a)
        if (angle) { ETHER_DISPATCH_DM(a.D, LA) }
        else { ETHER_DISPATCH_DM(a.D, LM) }
#undef LA
#undef LM
        return;
    }
#define LA(DM) hipLaunchKernelGGL((k_xyz_sweep<true, DM, MEAS>), grid, block, 0, st, a)
#define LM(DM) hipLaunchKernelGGL((k_xyz_sweep<false, DM, MEAS>), grid, block, 0, st, a)
    if (angle) { ETHER_DISPATCH_DM(a.D, LA) }
    else { ETHER_DISPATCH_DM(a.D, LM) }
#undef LA
#undef LM
}

void launch_xyz_sweep(const XyzArgs& a, bool angle, bool, bool, hipStream_t st) {
    const bool lanes = a.lps < kWave;
    if (g_recorder) {
        // a sweep writing measurement records runs every unit of its stride,
        // so that each record the finalize reads is rewritten
        record(LK_XYZ_SWEEP, (a.meas ? 262144 : 0) | (angle ? 65536 : 0) | (lanes ? 131072 : 0) | a.D,
               a.mrec ? a.nblocks : sweep_units(a.G, a.c_end - a.c_begin, a.spw * a.sps), a);
        return;
    }
    if (a.meas) launch_xyz_sweep_t<true>(a, angle, lanes, st);
    else launch_xyz_sweep_t<false>(a, angle, lanes, st);
}

void launch_xyz_ovrr(const XyzArgs& a, bool angle, bool, hipStream_t st) {
    if (g_recorder) {
        record(LK_XYZ_OVRR, (angle ? 65536 : 0) | a.D, sweep_units(a.G, a.c_end - a.c_begin, a.spw), a);
        return;
    }
    dim3 grid(a.nblocks), block(kBlock);
#define LA(DM) hipLaunchKernelGGL((k_xyz_ovrr<true, DM>), grid, block, 0, st, a)
#define LM(DM) hipLaunchKernelGGL((k_xyz_ovrr<false, DM>), grid, block, 0, st, a)
    if (angle) { ETHER_DISPATCH_DM(a.D, LA) }
    else { ETHER_DISPATCH_DM(a.D, LM) }
#undef LA
#undef LM
}

void launch_ising_sweep(const IsingArgs& a, hipStream_t st) {
    const bool lanes = a.lps_shift < 6;
    if (g_recorder) {
        record(LK_ISING_SWEEP, (lanes ? 131072 : 0) | a.D,
               sweep_units(a.G4, a.c_end - a.c_begin, a.spw * sites_per_step(a.lps_shift)), a);
        return;
    }
    if (lanes) {
#define LI(DM) hipLaunchKernelGGL((k_ising_sweep_lanes<DM>), dim3(a.nblocks), dim3(kBlock), 0, st, a)
        ETHER_DISPATCH_DM(a.D, LI)
#undef LI
        return;
    }
#define LI(DM) hipLaunchKernelGGL((k_ising_sweep<DM>), dim3(a.nblocks), dim3(kBlock), 0, st, a)
    ETHER_DISPATCH_DM(a.D, LI)
#undef LI
}

template <bool TL>
static void launch_ising_resident_tl(const IsingResArgs& a, dim3 grid, dim3 blk, size_t lds_bytes, hipStream_t st) {
    if (a.n_classes == 1) {
#define LR(DM) hipLaunchKernelGGL((k_ising_resident<DM, 1, TL>), grid, blk, lds_bytes, st, a)
        ETHER_DISPATCH_DM(a.D, LR)
#undef LR
    } else if (a.n_classes == 2) {
#define LR(DM) hipLaunchKernelGGL((k_ising_resident<DM, 2, TL>), grid, blk, lds_bytes, st, a)
        ETHER_DISPATCH_DM(a.D, LR)
#undef LR
    } else {
#define LR(DM) hipLaunchKernelGGL((k_ising_resident<DM, 0, TL>), grid, blk, lds_bytes, st, a)
        ETHER_DISPATCH_DM(a.D, LR)
#undef LR
    }
}

template <bool TL>
static void launch_ising_resident1_tl(const IsingResArgs& a, dim3 grid, dim3 blk, size_t lds_bytes, hipStream_t st) {
    if (a.n_classes == 1) {
#define LR(DM) hipLaunchKernelGGL((k_ising_resident1<DM, 1, TL>), grid, blk, lds_bytes, st, a)
        ETHER_DISPATCH_DM(a.D, LR)
#undef LR
    } else if (a.n_classes == 2) {
#define LR(DM) hipLaunchKernelGGL((k_ising_resident1<DM, 2, TL>), grid, blk, lds_bytes, st, a)
        ETHER_DISPATCH_DM(a.D, LR)
#undef LR
    } else {
#define LR(DM) hipLaunchKernelGGL((k_ising_resident1<DM, 0, TL>), grid, blk, lds_bytes, st, a)
        ETHER_DISPATCH_DM(a.D, LR)
#undef LR
    }
}

// variant of a recorded resident launch: (coupling-class instance << 16) | D
static int resident_variant(const IsingResArgs& a) {
    return ((a.n_classes == 1 ? 1 : (a.n_classes == 2 ? 2 : 0)) << 16) | a.D;
}

void launch_ising_resident(const IsingResArgs& a, int block, size_t lds_bytes, bool per_replica, hipStream_t st) {
    if (g_recorder) {
        // batches run the per-replica kernel with the table in LDS
        record(LK_ISING_RESIDENT, resident_variant(a), a.R, a, block, (int)lds_bytes);
        return;
    }
    const dim3 blk((unsigned)block);
    if (per_replica) {
        const dim3 grid((unsigned)a.R);
        if (a.table_in_lds) launch_ising_resident1_tl<true>(a, grid, blk, lds_bytes, st);
        else launch_ising_resident1_tl<false>(a, grid, blk, lds_bytes, st);
        return;
    }
    const dim3 grid((unsigned)(a.Rp / 4));
    if (a.table_in_lds) launch_ising_resident_tl<true>(a, grid, blk, lds_bytes, st);
    else launch_ising_resident_tl<false>(a, grid, blk, lds_bytes, st);
}

void launch_measure(const MeasArgs& a, int model, bool zee, bool sia, hipStream_t st) {
    // one workgroup per (tile, replica group): 64 replicas (XYZ, and Ising
    // below 128 replicas), 256 (Ising, four per lane)
    const bool wide = model == 0 && a.Rp >= 2 * kWave;
    const long long blocks = (long long)a.tiles * (wide ? (a.Rp + 4 * kWave - 1) / (4 * kWave) : a.G);
    if (g_recorder) {
        record(LK_MEASURE, (wide ? 8 : 0) | (model << 2) | ((zee ? 1 : 0) << 1) | (sia ? 1 : 0), (int)blocks, a);
        return;
    }
    dim3 grid((unsigned)blocks), block(kBlock);
    if (model == 0) {
        if (wide) hipLaunchKernelGGL(k_ising_measure, grid, block, 0, st, a);
        else hipLaunchKernelGGL(k_ising_measure_narrow, grid, block, 0, st, a);
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
    if (g_recorder) {
        record(LK_FINALIZE, 0, a.R, a);
        return;
    }
    hipLaunchKernelGGL(k_finalize, dim3(a.R), dim3(kBlock), 0, st, a);
}

void launch_pt_swap(const PtArgs& a, hipStream_t st) {
    int blocks = (a.n_groups + 63) / 64;
    if (g_recorder) {
        record(LK_PT, 0, blocks, a);
        return;
    }
    hipLaunchKernelGGL(k_pt_swap, dim3(blocks), dim3(64), 0, st, a);
}

template <bool MEAS>
static void launch_xyz_sweep_batch_t(const XyzArgs* a, const int2* d_items, int variant, dim3 grid,
                                     hipStream_t st) {
    const dim3 block(kBlock);
    if (variant & 131072) {
#define LA(DM) hipLaunchKernelGGL((k_xyz_sweep_lanes_batch<true, DM, MEAS>), grid, block, 0, st, a, d_items)
#define LM(DM) hipLaunchKernelGGL((k_xyz_sweep_lanes_batch<false, DM, MEAS>), grid, block, 0, st, a, d_items)
        if (variant & 65536) { ETHER_DISPATCH_DM(variant & 65535, LA) }
        else { ETHER_DISPATCH_DM(variant & 65535, LM) }
#undef LA
#undef LM
        return;
    }
#define LA(DM) hipLaunchKernelGGL((k_xyz_sweep_batch<true, DM, MEAS>), grid, block, 0, st, a, d_items)
#define LM(DM) hipLaunchKernelGGL((k_xyz_sweep_batch<false, DM, MEAS>), grid, block, 0, st, a, d_items)
    if (variant & 65536) { ETHER_DISPATCH_DM(variant & 65535, LA) }
    else { ETHER_DISPATCH_DM(variant & 65535, LM) }
#undef LA
#undef LM
}

// the batched resident kernel of one instance; LDS beyond the default
// per-workgroup limit is opted into once per instance (grant_batch_lds does
// it ahead of a graph capture)
template <int DM, int NC>
static void grant_resident1_batch(int lds) {
    constexpr int kMaxDevices = 64;
    static std::atomic<int> granted[kMaxDevices];   // per device: the attribute is the current device's
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
    if (lds > granted[dev].load() &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ising_resident1_batch<DM, NC, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds) == hipSuccess)
        granted[dev].store(lds);
}
template <int DM, int NC>
static void launch_resident1_batch(const IsingResArgs* a, const int2* d_items, dim3 grid, dim3 blk, int lds,
                                   hipStream_t st) {
    grant_resident1_batch<DM, NC>(lds);
    // without the grant the launch fails, which the caller's hipGetLastError reports
    hipLaunchKernelGGL((k_ising_resident1_batch<DM, NC, true>), grid, blk, (size_t)lds, st, a, d_items);
}

void grant_batch_lds(int kind, int variant, int lds) {
    if (kind != LK_ISING_RESIDENT) return;
    const int nc = variant >> 16;
#define LR(DM) grant_resident1_batch<DM, 1>(lds)
#define LS(DM) grant_resident1_batch<DM, 2>(lds)
#define LT(DM) grant_resident1_batch<DM, 0>(lds)
    if (nc == 1) { ETHER_DISPATCH_DM(variant & 65535, LR) }
    else if (nc == 2) { ETHER_DISPATCH_DM(variant & 65535, LS) }
    else { ETHER_DISPATCH_DM(variant & 65535, LT) }
#undef LR
#undef LS
#undef LT
}

void launch_batch(int kind, int variant, const void* d_args, const int2* d_items, int n_items,
                  int threads, int lds, hipStream_t st) {
    if (n_items <= 0) return;
    const dim3 grid(n_items), block(kBlock);
    switch (kind) {
    case LK_ISING_RESIDENT: {
        const IsingResArgs* a = static_cast<const IsingResArgs*>(d_args);
        const dim3 blk((unsigned)threads);
        const int nc = variant >> 16;
        if (nc == 1) {
#define LR(DM) launch_resident1_batch<DM, 1>(a, d_items, grid, blk, lds, st)
            ETHER_DISPATCH_DM(variant & 65535, LR)
#undef LR
        } else if (nc == 2) {
#define LR(DM) launch_resident1_batch<DM, 2>(a, d_items, grid, blk, lds, st)
            ETHER_DISPATCH_DM(variant & 65535, LR)
#undef LR
        } else {
#define LR(DM) launch_resident1_batch<DM, 0>(a, d_items, grid, blk, lds, st)
            ETHER_DISPATCH_DM(variant & 65535, LR)
#undef LR
        }
        break;
    }
    case LK_XYZ_SWEEP: {
        const XyzArgs* a = static_cast<const XyzArgs*>(d_args);
        if (variant & 262144) launch_xyz_sweep_batch_t<true>(a, d_items, variant, grid, st);
        else launch_xyz_sweep_batch_t<false>(a, d_items, variant, grid, st);
        break;
    }
    case LK_XYZ_OVRR: {
        const XyzArgs* a = static_cast<const XyzArgs*>(d_args);
#define LA(DM) hipLaunchKernelGGL((k_xyz_ovrr_batch<true, DM>), grid, block, 0, st, a, d_items)
#define LM(DM) hipLaunchKernelGGL((k_xyz_ovrr_batch<false, DM>), grid, block, 0, st, a, d_items)
        if (variant & 65536) { ETHER_DISPATCH_DM(variant & 65535, LA) }
        else { ETHER_DISPATCH_DM(variant & 65535, LM) }
#undef LA
#undef LM
        break;
    }
    case LK_ISING_SWEEP: {
        const IsingArgs* a = static_cast<const IsingArgs*>(d_args);
        if (variant & 131072) {
#define LI(DM) hipLaunchKernelGGL((k_ising_sweep_lanes_batch<DM>), grid, block, 0, st, a, d_items)
            ETHER_DISPATCH_DM(variant & 65535, LI)
#undef LI
            break;
        }
#define LI(DM) hipLaunchKernelGGL((k_ising_sweep_batch<DM>), grid, block, 0, st, a, d_items)
        ETHER_DISPATCH_DM(variant & 65535, LI)
#undef LI
        break;
    }
    case LK_MEASURE: {
        const MeasArgs* a = static_cast<const MeasArgs*>(d_args);
        switch (variant) {
        case 0: case 1: case 2: case 3:
            hipLaunchKernelGGL(k_ising_measure_narrow_batch, grid, block, 0, st, a, d_items); break;
        case 8: case 9: case 10: case 11:
            hipLaunchKernelGGL(k_ising_measure_batch, grid, block, 0, st, a, d_items); break;
        case 4: hipLaunchKernelGGL((k_xyz_measure_batch<false, false>), grid, block, 0, st, a, d_items); break;
        case 5: hipLaunchKernelGGL((k_xyz_measure_batch<false, true>), grid, block, 0, st, a, d_items); break;
        case 6: hipLaunchKernelGGL((k_xyz_measure_batch<true, false>), grid, block, 0, st, a, d_items); break;
        default: hipLaunchKernelGGL((k_xyz_measure_batch<true, true>), grid, block, 0, st, a, d_items); break;
        }
        break;
    }
    case LK_FINALIZE:
        hipLaunchKernelGGL(k_finalize_batch, grid, dim3(kBlock), 0, st,
                           static_cast<const FinArgs*>(d_args), d_items);
        break;
    case LK_PT:
        hipLaunchKernelGGL(k_pt_swap_batch, grid, dim3(64), 0, st, static_cast<const PtArgs*>(d_args),
                           d_items);
        break;
    default:
        break;
    }
}

void launch_init(const InitArgs& a, hipStream_t st) {
    size_t n = (size_t)a.N * a.Rp;
    hipLaunchKernelGGL(k_init, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, a);
}

void launch_local_fields(const MeasArgs& a, int r, float* h, float* e, int zee, int sia,
                         int model, hipStream_t st) {
    hipLaunchKernelGGL(k_local_fields, dim3((a.N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, a,
                       r, h, e, zee, sia, model);
}

__global__ void k_set_ctr(uint32_t* ctr, uint32_t step, uint32_t swaps, uint32_t swap_count) {
    ctr[0] = step;
    ctr[1] = swaps;
    ctr[2] = swap_count;
}

__global__ void k_advance_ctr(uint32_t* ctr, uint32_t dstep, uint32_t dswaps) {
    ctr[0] += dstep;
    ctr[1] += dswaps;
    ctr[2] += dswaps;
}

__global__ void k_ctr_batch(uint32_t* const* ctrs, int n, const uint32_t* vals, uint32_t dstep,
                            uint32_t dswaps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* c = ctrs[i];
    if (vals) {
        c[0] = vals[3 * i + 0];
        c[1] = vals[3 * i + 1];
        c[2] = vals[3 * i + 2];
    } else {
        c[0] += dstep;
        c[1] += dswaps;
        c[2] += dswaps;
    }
}

void launch_ctr_batch(uint32_t* const* d_ctrs, int n, const uint32_t* d_vals, uint32_t dstep,
                      uint32_t dswaps, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ctr_batch, dim3((n + 63) / 64), dim3(64), 0, st, d_ctrs, n, d_vals, dstep,
                       dswaps);
}

void launch_set_ctr(uint32_t* ctr, uint32_t step, uint32_t swaps, uint32_t swap_count,
                    hipStream_t st) {
    hipLaunchKernelGGL(k_set_ctr, dim3(1), dim3(1), 0, st, ctr, step, swaps, swap_count);
}

void launch_advance_ctr(uint32_t* ctr, uint32_t dstep, uint32_t dswaps, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_ctr, dim3(1), dim3(1), 0, st, ctr, dstep, dswaps);
}

void launch_extract(const float* plane, const int8_t* s, int N, int Rp, int r, float* out,
                    hipStream_t st) {
    hipLaunchKernelGGL(k_extract, dim3((N + 255) / 256), dim3(256), 0, st, plane, s, N, Rp, r, out);
}

void launch_insert(float* plane, int8_t* s, int N, int Rp, int r0, int r1, const float* in,
                   hipStream_t st) {
    size_t n = (size_t)N * (r1 - r0);
    hipLaunchKernelGGL(k_insert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, plane, s, N,
                       Rp, r0, r1, in);
}

}  // namespace ether
