// sq.hip - static structure factor on the matrix cores.
//
// Reference: utility/Structure_Factor/structure_factor.f90:141-177 computes
//   S_a(q) = sum_i S_a(r_i) exp(i q k.r_i),  k = b1 + b2 + b3  (:116)
// on a q-line q = q_start + m dq (:127,148) by a direct O(Q*N) sum, and writes
// sum_a |S_a|^2 / N^2 and |S_a| / N per component (:175-177).
//
// Two forms, both split-BF16 GEMMs on the MFMA units with FP32 accumulation:
//
// Grid.  On the commensurate grid q = 2 pi sum_k m_k b_k / L_k (a_i.b_j =
// delta_ij) of an L1 x L2 x L3 supercell with n_b basis sites at fractional
// coordinates f_l, q.r = 2 pi sum_k m_k (n_k + f_lk) / L_k, so the phase
// factorises over the lattice axes and S_a is a 3-D DFT of every sublattice,
// taken one axis at a time with the basis offset folded into the axis
// matrix:  U(m) = sum_n W_l(m, n) Z(n),  W_l(m, n) = exp(2 pi i m (n + f_lk) / L_k);
// the sublattices are then summed per q.  With one site per cell the last
// axis stage sums |U_a|^2 over the components in registers.
//
// Line.  The reference's own form: W(m, i) = exp(i (q_start + m dq) p_i) with
// p_i = k.r_i (FP64 on the device, the phase reduced to [-pi, pi] in FP64
// before its FP32 sine and cosine), contracted with the spin planes over all
// sites; the components' |U_a|^2 and |U_a| are formed in registers.
//
// Layout.  The engine keeps spins replica-major (site row of Rp replicas), and
// every stage keeps that: a GEMM's columns are the Rp replicas of a row, so
// every operand row and every result row is one contiguous, coalesced stretch
// of memory.  The first grid stage reads the spin planes themselves through
// the site -> slot map (no gather pass); the line form reads them in slot
// order (the sum runs over all sites, so the order is free).
//
// Arithmetic.  A stage is one fused complex GEMM:
//   U_re = C Z_re - S Z_im,  U_im = S Z_re + C Z_im
// (the Z_im products skipped where the input is real).  Operands are split
// into BF16 hi/lo parts (x = hi + lo, hi = bf16(x), lo = bf16(x - hi)); a
// product is hi*hi + hi*lo + lo*hi accumulated in FP32 with gfx950's
// v_mfma_f32_16x16x32_bf16 (one instruction per 32-deep K step), which keeps
// about 16 mantissa bits.  C and S of the grid are split once on the host; the
// line phases, the spin and the U operands are split as they are staged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ether_gpu.h"

namespace ether {
hipStream_t ctx_stream(ether_ctx* c);
void*& ctx_sq(ether_ctx* c);
int ctx_fail(ether_ctx* c, int code, const std::string& msg);
int ctx_model(ether_ctx* c);
int ctx_ev_begin(ether_ctx* c, int kind);
void ctx_ev_end(ether_ctx* c, int k);
int ctx_N(ether_ctx* c);
int ctx_R(ether_ctx* c);
int ctx_Rp(ether_ctx* c);
bool ctx_committed(ether_ctx* c);
const float* ctx_spin_plane(ether_ctx* c, int comp);
const int8_t* ctx_ising(ether_ctx* c);
const int* ctx_slot_of_site(ether_ctx* c);
const int* ctx_site_of_slot_dev(ether_ctx* c);
int ctx_device(ether_ctx* c);
bool ctx_dft_resident(ether_ctx* c);
}  // namespace ether

using namespace ether;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int BKP = BK + 8;      // LDS row pitch in shorts: 80 bytes, 16-byte aligned fragments, no conflicts
constexpr int kMaxAxis = 2048;   // longest lattice axis (the first stage stages its row map in LDS)

// round-to-nearest-even FP32 -> BF16 (finite inputs), used on host and device
__host__ __device__ inline unsigned short bf16_rne(float x) {
    unsigned int u;
    __builtin_memcpy(&u, &x, 4);
    unsigned int r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(r >> 16);
}
__host__ __device__ inline float bf16_to_f(unsigned short h) {
    unsigned int u = ((unsigned int)h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

struct DftArgs {
    // grid stages: input row n of (component a, batch i, o) is site
    // s = i + inner*(n + L*o):
    //   first stage: spin plane a (float, or the int8 Ising plane) at row map[s]
    //   later stages: in_re/in_im + (a*N + s)*Rp
    // line form: input row n is slot n of the spin planes
    const float* plane[3];
    const int8_t* s8;
    const int* map;
    const float* in_re;
    const float* in_im;
    float* out_re;            // U planes, [a][site][Rp]; unused in the norm stage
    float* out_im;
    float* norm;              // norm stage: [site][Rp] sum over a of |U_a|^2
    const unsigned short* w;  // grid: [sets][4][Mp][Kp]: C hi, C lo, S hi, S lo, zero padded
    const int* wsel;          // grid: W set of sublattice l = (batch index i) % nb; null: set 0
    int nb;
    int nsets;                // grid: W sets of the axis (distinct sublattice offsets along it)
    int L, M, Mp, Kp, inner, outer, ncomp, N, R, Rp;
    int tiles_m, tiles_n;
    // line form
    const double* p;          // [N] k.r of the site in each slot
    double q0, dq;
    float inv_n;
    float* line_out;          // [R][M][4]: sum_a |S_a|^2/N^2, |S_x|/N, |S_y|/N, |S_z|/N
};

__device__ __forceinline__ float4_t mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, float4_t c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
}

// sign flip of eight BF16 values (exact)
__device__ __forceinline__ bf16x8 neg8(bf16x8 x) {
    uint4 u;
    __builtin_memcpy(&u, &x, 16);
    u.x ^= 0x80008000u;
    u.y ^= 0x80008000u;
    u.z ^= 0x80008000u;
    u.w ^= 0x80008000u;
    bf16x8 y;
    __builtin_memcpy(&y, &u, 16);
    return y;
}

__device__ __forceinline__ bf16x8 frag(const unsigned short* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}

// One output tile (BM rows m x BN replicas) of one DFT stage.
// FIRST: the input is the real spin planes (through the site -> slot map);
// LAST: the components are summed as |U|^2 in registers and written as one
// plane; LINE: the reference's q-line (rows m are q points, the K dimension
// is every slot, the phases are computed as they are staged).  Four waves,
// each 32 x 32 as 2 x 2 MFMA 16x16 tiles; the operands of the next K step are
// loaded into registers while the current one is multiplied.
template <bool FIRST, bool LAST, bool LINE>
__global__ __launch_bounds__(256) void k_dft_stage(DftArgs g) {
    constexpr bool REAL = FIRST || LINE;
    // K-step tiles in LDS, k contiguous so that an MFMA operand fragment (8
    // consecutive k) is one 16-byte read: W as C hi, C lo, S hi, S lo; the
    // input split into hi/lo once, when it is staged
    __shared__ __attribute__((aligned(16))) unsigned short Aw[4][BM][BKP];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[REAL ? 1 : 2][2][BN][BKP];   // [re, im][hi, lo][replica][k]
    __shared__ int sRow[(FIRST && !LINE) ? kMaxAxis : 1];      // slot of each input row
    int bid = blockIdx.x;
    const int tn = bid % g.tiles_n;
    bid /= g.tiles_n;
    const int tm = bid % g.tiles_m;
    bid /= g.tiles_m;
    const int i = bid % g.inner;
    bid /= g.inner;
    const int o = bid % g.outer;
    const int a_only = bid / g.outer;              // component (stages that store U)
    const int i0 = tm * BM, j0 = tn * BN;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wi = (wv >> 1) * 32, wj = (wv & 1) * 32;
    const int kf = 8 * (lane >> 4);
    const long long L = g.L;
    const long long site0 = i + (long long)g.inner * L * o;   // site of row n: site0 + inner*n
    if constexpr (FIRST && !LINE) {
        for (int n = t; n < g.L; n += 256) sRow[n] = g.map[site0 + (long long)g.inner * n];
        __syncthreads();
    }
    const unsigned short* wset = g.w;
    if constexpr (!LINE)
        if (g.wsel) wset += (size_t)g.wsel[i % g.nb] * 4 * g.Mp * g.Kp;
    // per-thread slices of the K-step operand tiles:
    //   A: rows ii = t/4, 8 consecutive k at ka = (t%4)*8, for the 4 matrices
    //   B: rows kk = t/16 and kk+16, 4 consecutive replicas at cb = (t%16)*4
    const int ia = t >> 2, ka = (t & 3) * 8;
    const int kr = t >> 4, cb = (t & 15) * 4;
    uint4 pa0, pa1, pa2, pa3;   // named, not an array: the array went to scratch
    float4 pr[2], pi[2];
    auto fetch = [&](int a, int k0) {
        if constexpr (LINE) {
            // W(m, n) = exp(i q_m p_n): FP64 phase reduced to [-pi, pi], FP32 sine and cosine
            const int m = i0 + ia;
            const double q = g.q0 + (double)m * g.dq;
            unsigned short v[4][8];
            #pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int n = k0 + ka + u;
                float cs = 0.f, sn = 0.f;
                if (n < g.L && m < g.M) {
                    const double th = q * g.p[n];
                    const double red = th - 6.283185307179586476925286766559 *
                                                rint(th * 0.15915494309189533576888376337251);
                    sincosf((float)red, &sn, &cs);
                }
                v[0][u] = bf16_rne(cs);
                v[1][u] = bf16_rne(cs - bf16_to_f(v[0][u]));
                v[2][u] = bf16_rne(sn);
                v[3][u] = bf16_rne(sn - bf16_to_f(v[2][u]));
            }
            __builtin_memcpy(&pa0, v[0], 16);
            __builtin_memcpy(&pa1, v[1], 16);
            __builtin_memcpy(&pa2, v[2], 16);
            __builtin_memcpy(&pa3, v[3], 16);
        } else {
            const size_t P = (size_t)g.Mp * g.Kp;
            const unsigned short* src = wset + (size_t)(i0 + ia) * g.Kp + k0 + ka;
            pa0 = *(const uint4*)(src);
            pa1 = *(const uint4*)(src + P);
            pa2 = *(const uint4*)(src + 2 * P);
            pa3 = *(const uint4*)(src + 3 * P);
        }
        #pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int n = k0 + kr + 16 * h;
            float4 vr = make_float4(0.f, 0.f, 0.f, 0.f), vi = vr;
            if (n < g.L) {
                if constexpr (FIRST || LINE) {
                    long long row;
                    if constexpr (LINE) row = (long long)n * g.Rp + j0 + cb;
                    else row = (long long)sRow[n] * g.Rp + j0 + cb;
                    if (g.s8) {
                        const char4 c4 = *(const char4*)(g.s8 + row);
                        vr = make_float4((float)c4.x, (float)c4.y, (float)c4.z, (float)c4.w);
                    } else {
                        // a select, not an index: an indexed argument array would go to scratch
                        const float* pl = a == 0 ? g.plane[0] : (a == 1 ? g.plane[1] : g.plane[2]);
                        vr = *(const float4*)(pl + row);
                    }
                } else {
                    const long long off = ((long long)a * g.N + site0 + (long long)g.inner * n) * g.Rp + j0 + cb;
                    vr = *(const float4*)(g.in_re + off);
                    vi = *(const float4*)(g.in_im + off);
                }
            }
            pr[h] = vr;
            pi[h] = vi;
        }
    };
    auto put = [&](int part, int k, int col, float x) {
        const unsigned short hb = bf16_rne(x);
        Bs[part][0][col][k] = hb;
        Bs[part][1][col][k] = bf16_rne(x - bf16_to_f(hb));
    };
    auto stage = [&]() {
        *(uint4*)&Aw[0][ia][ka] = pa0;
        *(uint4*)&Aw[1][ia][ka] = pa1;
        *(uint4*)&Aw[2][ia][ka] = pa2;
        *(uint4*)&Aw[3][ia][ka] = pa3;
        #pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = kr + 16 * h;
            put(0, k, cb + 0, pr[h].x);
            put(0, k, cb + 1, pr[h].y);
            put(0, k, cb + 2, pr[h].z);
            put(0, k, cb + 3, pr[h].w);
            if constexpr (!REAL) {
                put(1, k, cb + 0, pi[h].x);
                put(1, k, cb + 1, pi[h].y);
                put(1, k, cb + 2, pi[h].z);
                put(1, k, cb + 3, pi[h].w);
            }
        }
    };
    float4_t nrm[2][2];
    #pragma unroll
    for (int p = 0; p < 2; ++p)
        #pragma unroll
        for (int q = 0; q < 2; ++q) nrm[p][q] = (float4_t){0.f, 0.f, 0.f, 0.f};
    const int a_begin = (LAST || LINE) ? 0 : a_only, a_end = (LAST || LINE) ? g.ncomp : a_only + 1;
    const int nsteps = (g.L + BK - 1) / BK;
    fetch(a_begin, 0);
    for (int a = a_begin; a < a_end; ++a) {
        float4_t ur[2][2], ui[2][2];
        #pragma unroll
        for (int p = 0; p < 2; ++p)
            #pragma unroll
            for (int q = 0; q < 2; ++q) {
                ur[p][q] = (float4_t){0.f, 0.f, 0.f, 0.f};
                ui[p][q] = (float4_t){0.f, 0.f, 0.f, 0.f};
            }
        for (int ks = 0; ks < nsteps; ++ks) {
            __syncthreads();                       // the previous step's tiles are consumed
            stage();
            __syncthreads();
            // next step (or the next component's first step) into registers
            if (ks + 1 < nsteps) fetch(a, (ks + 1) * BK);
            else if (a + 1 < a_end) fetch(a + 1, 0);
            // operand fragments of v_mfma_f32_16x16x32_bf16: lane l holds
            // A[l%16][8*(l/16) + v] and B[8*(l/16) + v][l%16], v = 0..7
            bf16x8 ch[2], cl[2], sh[2], sl[2];
            #pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int row = wi + p * 16 + (lane & 15);
                ch[p] = frag(&Aw[0][row][kf]);
                cl[p] = frag(&Aw[1][row][kf]);
                sh[p] = frag(&Aw[2][row][kf]);
                sl[p] = frag(&Aw[3][row][kf]);
            }
            #pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int col = wj + q * 16 + (lane & 15);
                const bf16x8 rh = frag(&Bs[0][0][col][kf]);
                const bf16x8 rl = frag(&Bs[0][1][col][kf]);
                bf16x8 ih, il;
                if constexpr (!REAL) {
                    ih = frag(&Bs[REAL ? 0 : 1][0][col][kf]);
                    il = frag(&Bs[REAL ? 0 : 1][1][col][kf]);
                }
                #pragma unroll
                for (int p = 0; p < 2; ++p) {
                    ur[p][q] = mfma3(ch[p], cl[p], rh, rl, ur[p][q]);
                    ui[p][q] = mfma3(sh[p], sl[p], rh, rl, ui[p][q]);
                    if constexpr (!REAL) {
                        ur[p][q] = mfma3(neg8(sh[p]), neg8(sl[p]), ih, il, ur[p][q]);
                        ui[p][q] = mfma3(ch[p], cl[p], ih, il, ui[p][q]);
                    }
                }
            }
        }
        // result fragment: lane l holds D[4*(l/16) + v][l%16]
        #pragma unroll
        for (int p = 0; p < 2; ++p)
            #pragma unroll
            for (int q = 0; q < 2; ++q)
                #pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int m = i0 + wi + p * 16 + 4 * (lane >> 4) + v;
                    const int r = j0 + wj + q * 16 + (lane & 15);
                    const float n2 = fmaf(ui[p][q][v], ui[p][q][v], ur[p][q][v] * ur[p][q][v]);
                    if constexpr (LAST || LINE) {
                        nrm[p][q][v] += n2;
                        if constexpr (LINE) {
                            // |S_a|/N of component a (the Ising plane is S_z)
                            const int comp = g.s8 ? 2 : a;
                            if (m < g.M && r < g.R)
                                g.line_out[((size_t)r * g.M + m) * 4 + 1 + comp] = sqrtf(n2) * g.inv_n;
                        }
                    } else {
                        if (m < g.M) {
                            const long long off = ((long long)a * g.N + site0 + (long long)g.inner * m) * g.Rp + r;
                            g.out_re[off] = ur[p][q][v];
                            g.out_im[off] = ui[p][q][v];
                        }
                    }
                }
    }
    if constexpr (LAST || LINE)
        #pragma unroll
        for (int p = 0; p < 2; ++p)
            #pragma unroll
            for (int q = 0; q < 2; ++q)
                #pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int m = i0 + wi + p * 16 + 4 * (lane >> 4) + v;
                    const int r = j0 + wj + q * 16 + (lane & 15);
                    if (m >= g.M) continue;
                    if constexpr (LINE) {
                        if (r < g.R) g.line_out[((size_t)r * g.M + m) * 4] = nrm[p][q][v] * g.inv_n * g.inv_n;
                    } else {
                        g.norm[(site0 + (long long)g.inner * m) * g.Rp + r] = nrm[p][q][v];
                    }
                }
}

// One grid axis stage for axes of at most 96 cells, with every W matrix of
// the axis resident in LDS (the C hi/lo and S hi/lo parts of each offset set,
// staged once per workgroup).  Each wave computes all rows of one output slab
// (the MB*16 >= L rows of one line, 16 replica columns) on its own: its B
// operand fragments (8 consecutive k of one replica per lane) are read
// straight from memory into registers, one K step ahead of the products, so
// the input is read once per stage and the waves share nothing but W (and, in
// the first stage, each wave's own copy of its line's row -> slot map).
// Waves walk the items (line, component unless LAST, slab) with a stride of
// the whole grid.  Same products, in the same order per output element, as
// k_dft_stage.
template <int MB, bool FIRST, bool LAST>
__global__ __launch_bounds__(256, 2) void k_dft_axis(DftArgs g, int items, int nslab) {
    extern __shared__ __attribute__((aligned(16))) unsigned short sW[];   // [nsets][4][MB*16][KP]
    __shared__ int sRowW[4][FIRST ? MB * 16 : 1];                         // slot of each row of the wave's line
    constexpr int ROWS = MB * 16;
    const int KP = g.Kp + 8;                       // row pitch (shorts): 16-byte rows, no bank conflicts
    {
        const int rowv = g.Kp / 8;                 // uint4 per row
        const int total = g.nsets * 4 * ROWS * rowv;
        for (int e = threadIdx.x; e < total; e += 256) {
            const int k8 = e % rowv, rm = e / rowv;
            const int m = rm % ROWS, sq = rm / ROWS;   // sq = set*4 + matrix
            *reinterpret_cast<uint4*>(sW + ((size_t)sq * ROWS + m) * KP + 8 * k8) =
                *reinterpret_cast<const uint4*>(g.w + ((size_t)sq * g.Mp + m) * g.Kp + 8 * k8);
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nwaves = gridDim.x * 4;
    const int kq = 8 * (lane >> 4);                // this lane's 8 k within a 32-deep step
    const int nsteps = g.Kp / BK;
    for (long long it = blockIdx.x * 4 + wv; it < items; it += nwaves) {
        int rest = (int)it;
        const int slab = rest % nslab;
        rest /= nslab;
        int a_only = 0;
        if constexpr (!LAST) {
            a_only = rest % g.ncomp;
            rest /= g.ncomp;
        }
        const int i = rest % g.inner, o = rest / g.inner;
        const long long site0 = i + (long long)g.inner * g.L * o;
        const int r = slab * 16 + (lane & 15);
        const int set = g.wsel ? g.wsel[i % g.nb] : 0;
        const unsigned short* W = sW + (size_t)set * 4 * ROWS * KP;
        const int a_begin = LAST ? 0 : a_only, a_end = LAST ? g.ncomp : a_only + 1;
        if constexpr (FIRST) {
            // the line's row -> slot map, read by this wave only
            __builtin_amdgcn_wave_barrier();           // the previous item's reads are done
            for (int n = lane; n < g.L; n += 64) sRowW[wv][n] = g.map[site0 + (long long)g.inner * n];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
        }
        // raw input of one k step: rows k0 + kq .. k0 + kq + 7 of replica r,
        // by buffer loads (the host keeps every plane under 2 GiB): a lane
        // offset plus a wave-uniform row stride; rows past the axis (n >= L)
        // read 0 through an offset past the buffer
        auto fetch = [&](int a, int k0, float (&xr)[8], float (&xi)[8]) {
            if constexpr (FIRST) {
                const float* pl = g.s8 ? nullptr : (a == 0 ? g.plane[0] : (a == 1 ? g.plane[1] : g.plane[2]));
                const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                    g.s8 ? (void*)g.s8 : (void*)pl, (short)0, (int)((size_t)g.N * g.Rp * (g.s8 ? 1 : 4)), 0x00020000);
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const int n = k0 + kq + v;
                    const int slot = n < g.L ? sRowW[wv][n] : -1;
                    if (g.s8) {
                        const uint32_t off = slot < 0 ? 0x80000000u : (uint32_t)(slot * g.Rp + r);
                        xr[v] = (float)(int8_t)__builtin_amdgcn_raw_buffer_load_b8(rs, (int)off, 0, 0);
                    } else {
                        const uint32_t off = slot < 0 ? 0x80000000u : 4u * (uint32_t)(slot * g.Rp + r);
                        xr[v] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0));
                    }
                    xi[v] = 0.f;
                }
            } else {
                const int bytes = (int)((size_t)g.ncomp * g.N * g.Rp * 4);
                const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc((void*)g.in_re, (short)0, bytes, 0x00020000);
                const __amdgpu_buffer_rsrc_t ri = __builtin_amdgcn_make_buffer_rsrc((void*)g.in_im, (short)0, bytes, 0x00020000);
                const uint32_t stride = 4u * (uint32_t)g.inner * (uint32_t)g.Rp;   // bytes per row n
                const uint32_t base = 4u * (uint32_t)(((long long)a * g.N + site0) * g.Rp + r) + (uint32_t)kq * stride;
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const bool in = k0 + kq + v < g.L;
                    const uint32_t off = in ? base + (uint32_t)(k0 + v) * stride : 0x80000000u;
                    xr[v] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rr, (int)off, 0, 0));
                    xi[v] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ri, (int)off, 0, 0));
                }
            }
        };
        auto split = [&](const float (&x)[8], bf16x8& hi, bf16x8& lo) {
            unsigned short h[8], l[8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                h[v] = bf16_rne(x[v]);
                l[v] = bf16_rne(x[v] - bf16_to_f(h[v]));
            }
            __builtin_memcpy(&hi, h, 16);
            __builtin_memcpy(&lo, l, 16);
        };
        float xr[8], xi[8];
        float4_t nrm[MB];
#pragma unroll
        for (int p = 0; p < MB; ++p) nrm[p] = (float4_t){0.f, 0.f, 0.f, 0.f};
        fetch(a_begin, 0, xr, xi);
        for (int a = a_begin; a < a_end; ++a) {
            float4_t ur[MB], ui[MB];
#pragma unroll
            for (int p = 0; p < MB; ++p) {
                ur[p] = (float4_t){0.f, 0.f, 0.f, 0.f};
                ui[p] = (float4_t){0.f, 0.f, 0.f, 0.f};
            }
            for (int ks = 0; ks < nsteps; ++ks) {
                bf16x8 rh, rl, ih, il;
                split(xr, rh, rl);
                if constexpr (!FIRST) split(xi, ih, il);
                // the next step's input (or the next component's first) while
                // this one multiplies (two steps ahead took more registers than
                // it saved in waiting)
                if (ks + 1 < nsteps) fetch(a, (ks + 1) * BK, xr, xi);
                else if (a + 1 < a_end) fetch(a + 1, 0, xr, xi);
                const int kf = ks * BK + kq;
#pragma unroll
                for (int p = 0; p < MB; ++p) {
                    const unsigned short* row = W + (size_t)(p * 16 + (lane & 15)) * KP + kf;
                    const bf16x8 ch = frag(row), cl = frag(row + (size_t)ROWS * KP),
                                 sh = frag(row + (size_t)2 * ROWS * KP), sl = frag(row + (size_t)3 * ROWS * KP);
                    ur[p] = mfma3(ch, cl, rh, rl, ur[p]);
                    ui[p] = mfma3(sh, sl, rh, rl, ui[p]);
                    if constexpr (!FIRST) {
                        ur[p] = mfma3(neg8(sh), neg8(sl), ih, il, ur[p]);
                        ui[p] = mfma3(ch, cl, ih, il, ui[p]);
                    }
                }
            }
            // result fragment: lane l holds D[4*(l/16) + v][l%16]
#pragma unroll
            for (int p = 0; p < MB; ++p)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int m = p * 16 + 4 * (lane >> 4) + v;
                    if constexpr (LAST) {
                        nrm[p][v] += fmaf(ui[p][v], ui[p][v], ur[p][v] * ur[p][v]);
                    } else if (m < g.M) {
                        const long long off = ((long long)a * g.N + site0 + (long long)g.inner * m) * g.Rp + r;
                        g.out_re[off] = ur[p][v];
                        g.out_im[off] = ui[p][v];
                    }
                }
        }
        if constexpr (LAST)
#pragma unroll
            for (int p = 0; p < MB; ++p)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int m = p * 16 + 4 * (lane >> 4) + v;
                    if (m < g.M) g.norm[(site0 + (long long)g.inner * m) * g.Rp + r] = nrm[p][v];
                }
    }
}

// several sites per cell: norm[c][r] = sum_a |sum_l U_a(l + nb*c)|^2 over the
// cells c of the grid (the axis stages left U per sublattice)
__global__ __launch_bounds__(256) void k_sq_basis_norm(const float* re, const float* im, int N, int nb, int Rp,
                                                       int ncomp, float* norm) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long cells = N / nb;
    if (t >= cells * Rp) return;
    const long long c = t / Rp;
    const int r = (int)(t % Rp);
    float s = 0.f;
    for (int a = 0; a < ncomp; ++a) {
        float xr = 0.f, xi = 0.f;
        for (int l = 0; l < nb; ++l) {
            const long long off = ((long long)a * N + c * nb + l) * Rp + r;
            xr += re[off];
            xi += im[off];
        }
        s = fmaf(xr, xr, fmaf(xi, xi, s));
    }
    norm[c * Rp + r] = s;
}

// out[r][q] = norm[q][r] * scale for r < R: 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void k_sq_transpose(const float* norm, int Q, int R, int Rp, float scale,
                                                    float* out) {
    __shared__ float tile[32][33];
    const int q0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int q = q0 + k, r = r0 + tx;
        tile[k][tx] = (q < Q && r < R) ? norm[(long long)q * Rp + r] : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, q = q0 + tx;
        if (r < R && q < Q) out[(long long)r * Q + q] = tile[tx][k] * scale;
    }
}

// fill a float buffer (the line output's components that a spin model lacks)
__global__ void k_sq_fill(float* p, long long n, float v) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t < n) p[t] = v;
}

// p[slot] = k . r of the site in that slot, in FP64
__global__ void k_sq_line_p(const double* pos, const int* site_of_slot, int N, double kx, double ky, double kz,
                            double* p) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= N) return;
    const double* r = pos + 3 * (size_t)site_of_slot[q];
    p[q] = fma(kx, r[0], fma(ky, r[1], kz * r[2]));
}

struct SqState {
    // grid
    int N = 0, R = 0, L[3] = {0, 0, 0}, nb = 0;
    std::vector<double> frac;
    float* d_z[4] = {nullptr, nullptr, nullptr, nullptr};  // U re/im planes, ping/pong
    unsigned short* d_w[3] = {nullptr, nullptr, nullptr};   // per-axis split C/S matrices (one set per offset)
    int* d_wsel[3] = {nullptr, nullptr, nullptr};           // per-axis W set of each sublattice
    int nsets[3] = {0, 0, 0};                               // per-axis number of W sets
    int n_cu = 0;                                           // compute units of the device
    size_t lds_max = 0;                                     // LDS a workgroup may use
    float* d_norm = nullptr;
    int* d_map = nullptr;
    float* d_out = nullptr;
    // line
    double* d_pos = nullptr;
    double* d_p = nullptr;
    float* d_lout = nullptr;
    size_t lout_n = 0;
    int p_n = 0;
};

void sq_free_grid(SqState* s) {
    for (int k = 0; k < 4; ++k) (void)hipFree(s->d_z[k]);
    for (int k = 0; k < 3; ++k) {
        (void)hipFree(s->d_w[k]);
        (void)hipFree(s->d_wsel[k]);
        s->d_w[k] = nullptr;
        s->d_wsel[k] = nullptr;
    }
    for (int k = 0; k < 4; ++k) s->d_z[k] = nullptr;
    (void)hipFree(s->d_norm);
    (void)hipFree(s->d_map);
    (void)hipFree(s->d_out);
    s->d_norm = s->d_out = nullptr;
    s->d_map = nullptr;
    s->N = 0;
}

}  // namespace

namespace ether {
void sq_destroy(void* p) {
    SqState* s = static_cast<SqState*>(p);
    sq_free_grid(s);
    (void)hipFree(s->d_pos);
    (void)hipFree(s->d_p);
    (void)hipFree(s->d_lout);
    delete s;
}
}  // namespace ether

#define HC(c, e)                                                                          \
    do {                                                                                  \
        hipError_t _e = (e);                                                              \
        if (_e != hipSuccess) return ctx_fail(c, ETHER_ERR_HIP, std::string(#e) + ": " +  \
                                                                   hipGetErrorString(_e)); \
    } while (0)

static SqState* sq_state(ether_ctx* c) {
    SqState* s = static_cast<SqState*>(ctx_sq(c));
    if (!s) {
        s = new SqState();
        ctx_sq(c) = s;
    }
    return s;
}

// The W-resident axis kernel for an axis of L <= 96 cells: MB = rows / 16
// with the rows L rounded up to 16 (one block) or 32 (longer axes would need
// more registers than two waves per SIMD leave); LDS for every W set of the
// axis; 0 if it does not apply.
static int axis_mb(int L) {
    if (L <= 16) return 1;
    if (L > 96) return 0;
    return (L + 31) / 32 * 2;
}
static size_t axis_lds(int mb, int Kp, int nsets) { return (size_t)nsets * 4 * mb * 16 * (Kp + 8) * 2; }

// Launches one axis stage: as many workgroups as can be resident at once
// (LDS and registers, from the occupancy query), at most one per four items.
template <int MB, bool FIRST, bool LAST>
static hipError_t launch_axis_t(const DftArgs& g, long long items, int nslab, int n_cu, size_t lds, hipStream_t st) {
    const void* fn = (const void*)k_dft_axis<MB, FIRST, LAST>;
    // above the default dynamic LDS limit of a launch the kernel must ask for it
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds);
    if (e != hipSuccess) return e;
    const long long grid = std::max<long long>(1, std::min<long long>((items + 3) / 4, (long long)std::max(1, per_cu) * n_cu));
    hipLaunchKernelGGL((k_dft_axis<MB, FIRST, LAST>), dim3((unsigned)grid), dim3(256), lds, st, g, (int)items, nslab);
    return hipGetLastError();
}
template <int MB>
static hipError_t launch_axis_mb(const DftArgs& g, bool first, bool last, long long items, int nslab, int n_cu,
                                 size_t lds, hipStream_t st) {
    if (first && last) return launch_axis_t<MB, true, true>(g, items, nslab, n_cu, lds, st);
    if (first) return launch_axis_t<MB, true, false>(g, items, nslab, n_cu, lds, st);
    if (last) return launch_axis_t<MB, false, true>(g, items, nslab, n_cu, lds, st);
    return launch_axis_t<MB, false, false>(g, items, nslab, n_cu, lds, st);
}
static hipError_t launch_axis(const DftArgs& g, int mb, bool first, bool last, long long items, int nslab, int n_cu,
                              size_t lds, hipStream_t st) {
    switch (mb) {
        case 1: return launch_axis_mb<1>(g, first, last, items, nslab, n_cu, lds, st);
        case 2: return launch_axis_mb<2>(g, first, last, items, nslab, n_cu, lds, st);
        case 4: return launch_axis_mb<4>(g, first, last, items, nslab, n_cu, lds, st);
        default: return launch_axis_mb<6>(g, first, last, items, nslab, n_cu, lds, st);
    }
}

static void launch_stage(const DftArgs& g, bool first, bool last, long long blocks, hipStream_t st) {
    const dim3 grid((unsigned)blocks), block(256);
    if (first && last) hipLaunchKernelGGL((k_dft_stage<true, true, false>), grid, block, 0, st, g);
    else if (first) hipLaunchKernelGGL((k_dft_stage<true, false, false>), grid, block, 0, st, g);
    else if (last) hipLaunchKernelGGL((k_dft_stage<false, true, false>), grid, block, 0, st, g);
    else hipLaunchKernelGGL((k_dft_stage<false, false, false>), grid, block, 0, st, g);
}

// The grid form for nb sites per cell at fractional offsets frac[3*nb].
static int sq_grid(ether_ctx* c, const int32_t* sc, int nb, const double* frac, float* out) {
    if (!c) return ctx_fail(nullptr, ETHER_ERR_ARG, "null context");
    if (!ctx_committed(c)) return ctx_fail(c, ETHER_ERR_STATE, "context not committed");
    if (!sc || !out || nb <= 0) return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor: null pointer");
    const int N = ctx_N(c), R = ctx_R(c), Rp = ctx_Rp(c);
    if (sc[0] <= 0 || sc[1] <= 0 || sc[2] <= 0 || (long long)sc[0] * sc[1] * sc[2] * nb != N)
        return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor: L1*L2*L3*n_basis must equal the sites");
    if ((long long)N * Rp * 3 > (1LL << 31))
        return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor: too many spins");
    if (sc[0] > kMaxAxis || sc[1] > kMaxAxis || sc[2] > kMaxAxis)
        return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor: a lattice axis is longer than 2048 cells");
    const bool ising = ctx_model(c) == ETHER_MODEL_ISING;
    const int ncomp = ising ? 1 : 3;
    const int Q = N / nb;
    hipStream_t st = ctx_stream(c);
    SqState* s = sq_state(c);
    std::vector<double> fr((size_t)3 * nb, 0.0);
    if (frac)
        for (int q = 0; q < 3 * nb; ++q) fr[q] = frac[q];
    const bool fresh = s->N != N || s->R != R || s->nb != nb || s->L[0] != sc[0] || s->L[1] != sc[1] ||
                       s->L[2] != sc[2] || s->frac != fr;
    if (fresh) {
        sq_free_grid(s);
        s->N = N;
        s->R = R;
        s->nb = nb;
        s->frac = fr;
        for (int k = 0; k < 3; ++k) s->L[k] = sc[k];
        const size_t plane = (size_t)N * Rp * 3;
        for (int k = 0; k < 4; ++k) HC(c, hipMalloc(&s->d_z[k], plane * sizeof(float)));
        HC(c, hipMalloc(&s->d_norm, sizeof(float) * (size_t)Q * Rp));
        HC(c, hipMalloc(&s->d_map, sizeof(int) * N));
        HC(c, hipMalloc(&s->d_out, sizeof(float) * (size_t)Q * R));
        std::vector<std::vector<unsigned short>> host_w(3);
        std::vector<std::vector<int>> host_sel(3);
        for (int k = 0; k < 3; ++k) {
            const int L = sc[k];
            const size_t Mp = ((size_t)L + BM - 1) / BM * BM, Kp = ((size_t)L + BK - 1) / BK * BK, P = Mp * Kp;
            // one matrix set per distinct offset of the sublattices along axis k
            std::vector<double> offs;
            std::vector<int> sel(nb);
            for (int l = 0; l < nb; ++l) {
                const double f = fr[3 * l + k];
                size_t u = 0;
                while (u < offs.size() && offs[u] != f) ++u;
                if (u == offs.size()) offs.push_back(f);
                sel[l] = (int)u;
            }
            std::vector<unsigned short>& w = host_w[k];
            w.assign(offs.size() * 4 * P, 0);
            for (size_t u = 0; u < offs.size(); ++u)
                for (int m = 0; m < L; ++m)
                    for (int n = 0; n < L; ++n) {
                        // exact argument reduction of the integer part: (m*n mod L) / L,
                        // then the offset's phase m*f / L
                        const double th = 2.0 * M_PI * ((double)((long long)m * n % L) + m * offs[u]) / L;
                        const float cv = (float)std::cos(th), sv = (float)std::sin(th);
                        unsigned short* ws = w.data() + u * 4 * P;
                        const size_t e = (size_t)m * Kp + n;
                        ws[e] = bf16_rne(cv);
                        ws[P + e] = bf16_rne(cv - bf16_to_f(ws[e]));
                        ws[2 * P + e] = bf16_rne(sv);
                        ws[3 * P + e] = bf16_rne(sv - bf16_to_f(ws[2 * P + e]));
                    }
            HC(c, hipMalloc(&s->d_w[k], sizeof(unsigned short) * w.size()));
            HC(c, hipMemcpyAsync(s->d_w[k], w.data(), sizeof(unsigned short) * w.size(), hipMemcpyHostToDevice, st));
            host_sel[k] = sel;
            s->nsets[k] = (int)offs.size();
            HC(c, hipMalloc(&s->d_wsel[k], sizeof(int) * nb));
            HC(c, hipMemcpyAsync(s->d_wsel[k], host_sel[k].data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
        }
        HC(c, hipMemcpyAsync(s->d_map, ctx_slot_of_site(c), sizeof(int) * N, hipMemcpyHostToDevice, st));
        int dev = 0, ncu = 0, lds = 0;
        HC(c, hipGetDevice(&dev));
        HC(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HC(c, hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        s->n_cu = ncu;
        s->lds_max = (size_t)lds;
        HC(c, hipStreamSynchronize(st));   // the host staging vectors go out of scope
    }
    const int ev = ctx_ev_begin(c, ETHER_TIME_SQ);
    // the transformed axes (length > 1), in order; a lattice of one cell
    // still needs one stage to form the norm
    int axes[3], na = 0;
    for (int k = 0; k < 3; ++k)
        if (sc[k] > 1) axes[na++] = k;
    if (na == 0) axes[na++] = 0;
    int cur = -1;   // index of the input U pair (0 or 2); -1: the spin planes
    for (int st_i = 0; st_i < na; ++st_i) {
        const int k = axes[st_i];
        DftArgs g{};
        g.L = g.M = sc[k];
        g.Mp = (g.L + BM - 1) / BM * BM;
        g.Kp = (g.L + BK - 1) / BK * BK;
        g.inner = nb;
        for (int q = 0; q < k; ++q) g.inner *= sc[q];
        g.outer = N / (g.inner * g.L);
        g.ncomp = ncomp;
        g.N = N;
        g.R = R;
        g.Rp = Rp;
        g.nb = nb;
        g.nsets = s->nsets[k];
        const bool first = st_i == 0;
        // with several sites per cell the sublattices are summed after the
        // last axis, so every axis stage stores U
        const bool last = st_i == na - 1 && nb == 1;
        g.w = s->d_w[k];
        g.wsel = s->d_wsel[k];
        for (int a = 0; a < 3; ++a) g.plane[a] = ising ? nullptr : ctx_spin_plane(c, a);
        g.s8 = ising ? ctx_ising(c) : nullptr;
        g.map = s->d_map;
        if (!first) {
            g.in_re = s->d_z[cur];
            g.in_im = s->d_z[cur + 1];
        }
        const int nxt = cur == 0 ? 2 : 0;
        if (!last) {
            g.out_re = s->d_z[nxt];
            g.out_im = s->d_z[nxt + 1];
        } else {
            g.norm = s->d_norm;
        }
        const int mb = ctx_dft_resident(c) ? axis_mb(g.L) : 0;
        const size_t lds = mb ? axis_lds(mb, g.Kp, g.nsets) : 0;
        // its buffer loads address a plane in 32-bit byte offsets
        const bool fits = (size_t)ncomp * N * Rp * 4 < ((size_t)1 << 31);
        if (mb && lds <= s->lds_max && fits) {
            // W resident in LDS: one wave per (line, component unless last, 16 replicas)
            const int nslab = Rp / 16;
            const long long items = (long long)g.inner * g.outer * (last ? 1 : ncomp) * nslab;
            if (items > 0x7fffffffLL) return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor: grid too large");
            HC(c, launch_axis(g, mb, first, last, items, nslab, s->n_cu, lds, st));
        } else {
            g.tiles_m = (g.M + BM - 1) / BM;
            g.tiles_n = (Rp + BN - 1) / BN;
            const long long blocks =
                (long long)g.tiles_m * g.tiles_n * g.inner * g.outer * (last ? 1 : ncomp);
            if (blocks > 0x7fffffffLL) return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor: grid too large");
            launch_stage(g, first, last, blocks, st);
        }
        HC(c, hipGetLastError());
        cur = nxt;
    }
    if (nb > 1) {
        const long long n = (long long)Q * Rp;
        hipLaunchKernelGGL(k_sq_basis_norm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s->d_z[cur],
                           s->d_z[cur + 1], N, nb, Rp, ncomp, s->d_norm);
    }
    hipLaunchKernelGGL(k_sq_transpose, dim3((Q + 31) / 32, (R + 31) / 32), dim3(256), 0, st, s->d_norm, Q, R,
                       Rp, (float)(1.0 / ((double)N * N)), s->d_out);
    ctx_ev_end(c, ev);
    HC(c, hipGetLastError());
    HC(c, hipMemcpyAsync(out, s->d_out, sizeof(float) * (size_t)Q * R, hipMemcpyDeviceToHost, st));
    HC(c, hipStreamSynchronize(st));
    return ETHER_OK;
}

extern "C" int ether_structure_factor(ether_ctx* c, const int32_t* sc, float* out) {
    return sq_grid(c, sc, 1, nullptr, out);
}

extern "C" int ether_structure_factor_basis(ether_ctx* c, const int32_t* sc, int32_t n_basis,
                                            const double* frac, float* out) {
    if (!frac) return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor_basis: null offsets");
    return sq_grid(c, sc, n_basis, frac, out);
}

// The line form on device buffers that the caller provides (the context's
// cached ones, or the no-allocation entry point's static scratch): pos[3N]
// (filled here from the host), p[N] (k.r per slot) and lout[R*n_q*4].
// Enqueues the copy, the phase kernel and the contraction, then copies lout
// to the host and synchronises.
static int sq_line_run(ether_ctx* c, const double* pos, const double* k_vec, double q_start, double dq,
                       int32_t n_q, double* d_pos, double* d_p, float* d_lout, float* out) {
    const int N = ctx_N(c), R = ctx_R(c), Rp = ctx_Rp(c);
    const bool ising = ctx_model(c) == ETHER_MODEL_ISING;
    hipStream_t st = ctx_stream(c);
    HC(c, hipMemcpyAsync(d_pos, pos, sizeof(double) * 3 * (size_t)N, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_sq_line_p, dim3((N + 255) / 256), dim3(256), 0, st, d_pos, ctx_site_of_slot_dev(c), N,
                       k_vec[0], k_vec[1], k_vec[2], d_p);
    const long long nout = (long long)R * n_q * 4;
    // the components a spin model lacks (Ising: S_x, S_y) read zero
    if (ising)
        hipLaunchKernelGGL(k_sq_fill, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, d_lout, nout, 0.f);
    const int ev = ctx_ev_begin(c, ETHER_TIME_SQ);
    DftArgs g{};
    g.L = N;
    g.M = n_q;
    g.inner = g.outer = 1;
    g.ncomp = ising ? 1 : 3;
    g.N = N;
    g.R = R;
    g.Rp = Rp;
    g.nb = 1;
    for (int a = 0; a < 3; ++a) g.plane[a] = ising ? nullptr : ctx_spin_plane(c, a);
    g.s8 = ising ? ctx_ising(c) : nullptr;
    g.p = d_p;
    g.q0 = q_start;
    g.dq = dq;
    g.inv_n = (float)(1.0 / (double)N);
    g.line_out = d_lout;
    g.tiles_m = (n_q + BM - 1) / BM;
    g.tiles_n = (Rp + BN - 1) / BN;
    const long long blocks = (long long)g.tiles_m * g.tiles_n;
    hipLaunchKernelGGL((k_dft_stage<true, true, true>), dim3((unsigned)blocks), dim3(256), 0, st, g);
    ctx_ev_end(c, ev);
    HC(c, hipGetLastError());
    HC(c, hipMemcpyAsync(out, d_lout, sizeof(float) * (size_t)nout, hipMemcpyDeviceToHost, st));
    HC(c, hipStreamSynchronize(st));
    return ETHER_OK;
}

static int sq_line_check(ether_ctx* c, const double* pos, const double* k_vec, double q_start, double q_stop,
                         double dq, int32_t n_q, const float* out) {
    if (!c) return ctx_fail(nullptr, ETHER_ERR_ARG, "null context");
    if (!ctx_committed(c)) return ctx_fail(c, ETHER_ERR_STATE, "context not committed");
    if (!pos || !k_vec || !out) return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor_line: null pointer");
    if (!(dq != 0.0) || !std::isfinite(dq) || !std::isfinite(q_start) || !std::isfinite(q_stop))
        return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor_line: dq must be nonzero and finite");
    // total_q = abs(nint((q_stop - q_start)/dq)) + 1
    const double x = (q_stop - q_start) / dq;
    if (!(std::fabs(x) < 1e9)) return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor_line: too many q points");
    const long long want = std::llabs((long long)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5))) + 1;
    if (n_q != want)
        return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor_line: n_q must be |nint((q_stop-q_start)/dq)| + 1 = " +
                                              std::to_string(want));
    if ((long long)ctx_N(c) * ctx_Rp(c) > (1LL << 31) || (long long)ctx_R(c) * n_q * 4 > (1LL << 31))
        return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor_line: too many spins or q points");
    return ETHER_OK;
}

extern "C" int ether_structure_factor_line(ether_ctx* c, const double* pos, const double* k_vec, double q_start,
                                           double q_stop, double dq, int32_t n_q, float* out) {
    int rc = sq_line_check(c, pos, k_vec, q_start, q_stop, dq, n_q, out);
    if (rc) return rc;
    const int N = ctx_N(c), R = ctx_R(c);
    SqState* s = sq_state(c);
    if (s->p_n != N) {
        (void)hipFree(s->d_p);
        (void)hipFree(s->d_pos);
        s->d_p = s->d_pos = nullptr;
        s->p_n = 0;
        HC(c, hipMalloc(&s->d_p, sizeof(double) * N));
        HC(c, hipMalloc(&s->d_pos, sizeof(double) * 3 * (size_t)N));
        s->p_n = N;
    }
    const size_t nout = (size_t)R * n_q * 4;
    if (s->lout_n < nout) {
        (void)hipFree(s->d_lout);
        s->d_lout = nullptr;
        s->lout_n = 0;
        HC(c, hipMalloc(&s->d_lout, sizeof(float) * nout));
        s->lout_n = nout;
    }
    return sq_line_run(c, pos, k_vec, q_start, dq, n_q, s->d_pos, s->d_p, s->d_lout, out);
}

// ---------------------------------------------------------------------------
// The line form without allocation.  Its device buffers (positions, k.r,
// output staging) are carved from one static device array of each GPU, whose
// address is looked up once per device when a context is created there.
// Calls on one GPU share the array: each call fills what it reads and
// synchronises before returning, and calls that overlap on one GPU (from two
// host threads) race.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;   // device -> address of its copy of `scratch`

size_t align256(size_t x) { return (x + 255) / 256 * 256; }
}  // namespace

namespace ether {
// Looks up (once) the address of `scratch` on `device`, with that device current.
int sq_scratch_setup(int device, std::string& err) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return ETHER_OK;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) {
        err = "hipGetDevice failed";
        return ETHER_ERR_HIP;
    }
    if (cur != device && hipSetDevice(device) != hipSuccess) {
        err = "hipSetDevice failed";
        return ETHER_ERR_HIP;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
        err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        return ETHER_ERR_HIP;
    }
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    return ETHER_OK;
}
}  // namespace ether

extern "C" int ether_structure_factor_line_noalloc(ether_ctx* c, const double* pos, const double* k_vec,
                                                   double q_start, double q_stop, double dq, int32_t n_q,
                                                   float* out) {
    int rc = sq_line_check(c, pos, k_vec, q_start, q_stop, dq, n_q, out);
    if (rc) return rc;
    const size_t N = (size_t)ctx_N(c), R = (size_t)ctx_R(c);
    const size_t off_p = align256(sizeof(double) * 3 * N);
    const size_t off_l = off_p + align256(sizeof(double) * N);
    const size_t need = off_l + sizeof(float) * R * (size_t)n_q * 4;
    if (need > kScratchBytes)
        return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor_line_noalloc: needs " + std::to_string(need) +
                                                " bytes of device scratch (32 N + 16 R n_q), more than the " +
                                                std::to_string(kScratchBytes) + " it has");
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = g_scratch_ptr.find(ctx_device(c));
        if (it != g_scratch_ptr.end()) base = it->second;
    }
    if (!base) return ctx_fail(c, ETHER_ERR_STATE, "ether_structure_factor_line_noalloc: no scratch on this device");
    return sq_line_run(c, pos, k_vec, q_start, dq, n_q, reinterpret_cast<double*>(base),
                       reinterpret_cast<double*>(base + off_p), reinterpret_cast<float*>(base + off_l), out);
}
