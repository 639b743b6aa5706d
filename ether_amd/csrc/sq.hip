// sq.hip - static structure factor on the matrix cores.
//
// Reference: utility/Structure_Factor/structure_factor.f90:150-177 computes
//   S_a(q) = sum_r S_a(r) exp(i q.r),  out = sum_a |S_a(q)|^2 / N^2
// by a direct O(Q*N) sum with complex exponentials.  For a one-site-per-cell
// supercell and the commensurate grid q = 2*pi*sum_k m_k b_k / L_k
// (a_i.b_j = delta_ij) the phase factorises, q.r = 2*pi*sum_k m_k n_k / L_k,
// so S_a is a 3-D DFT of each spin plane, taken one lattice axis at a time:
//   U(m) = sum_n W(m, n) Z(n),  W = C + iS,  C/S(m, n) = cos/sin(2 pi m n / L).
//
// Layout.  The engine keeps spins replica-major (site row of Rp replicas), and
// every stage keeps that: a stage is, for each (component, sites before the
// axis, sites after it), one GEMM whose columns are the Rp replicas of a row,
// so every operand row and every result row is one contiguous, coalesced
// stretch of memory.  The first stage reads the spin planes themselves through
// the site -> slot map (no gather pass); the last stage does not store U but
// sums |U_a|^2 over the components in registers and writes one FP32 value per
// (q, replica); a tiled transpose then writes out[replica][q].
//
// Arithmetic.  Each stage is one fused complex GEMM on the MFMA units:
//   U_re = C Z_re - S Z_im,  U_im = S Z_re + C Z_im
// (the Z_im products skipped where the input is real).  Operands are split
// into BF16 hi/lo parts (x = hi + lo, hi = bf16(x), lo = bf16(x - hi)); a
// product is hi*hi + hi*lo + lo*hi accumulated in FP32
// (v_mfma_f32_16x16x16_bf16), which keeps about 16 mantissa bits.  C and S
// are split once on the host; the spin and U operands are split as they are
// staged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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
}  // namespace ether

using namespace ether;

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BM = 64, BN = 64, BK = 32;
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
    // input row n of (component a, batch i, o) is site s = i + inner*(n + L*o):
    //   first stage: spin plane a (float, or the int8 Ising plane) at row map[s]
    //   later stages: in_re/in_im + (a*N + s)*Rp
    const float* plane[3];
    const int8_t* s8;
    const int* map;
    const float* in_re;
    const float* in_im;
    float* out_re;            // U planes, [a][site][Rp]; unused in the norm stage
    float* out_im;
    float* norm;              // norm stage: [site][Rp] sum over a of |U_a|^2
    const unsigned short* w;  // [4][Mp][Kp]: C hi, C lo, S hi, S lo, zero padded
    int L, Mp, Kp, inner, outer, ncomp, N, Rp;
    int tiles_m, tiles_n;
};

__device__ __forceinline__ float4_t mfma3(short4_t ah, short4_t al, short4_t bh, short4_t bl, float4_t c) {
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
}

__device__ __forceinline__ short4_t neg4(short4_t x) {
    short4_t y;
    for (int v = 0; v < 4; ++v) y[v] = (short)((unsigned short)x[v] ^ 0x8000u);   // exact
    return y;
}

// One output tile (BM rows m of the axis x BN replicas) of one DFT stage.
// FIRST: the input is the real spin planes (through the site -> slot map);
// LAST: the components are summed as |U|^2 in registers and written as one
// plane.  Four waves, each 32 x 32 as 2 x 2 MFMA 16x16 tiles; the operands of
// the next K step are loaded into registers while the current one is
// multiplied.
template <bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_dft_stage(DftArgs g) {
    constexpr bool REAL = FIRST;
    // K-step tiles in LDS, k contiguous so that an MFMA operand fragment (4
    // consecutive k) is one 8-byte read: W as C hi, C lo, S hi, S lo; the
    // input split into hi/lo once, when it is staged
    __shared__ __attribute__((aligned(16))) short Aw[4][BM][BK + 4];
    __shared__ __attribute__((aligned(16))) short Bs[REAL ? 1 : 2][2][BN][BK + 4];   // [re, im][hi, lo][replica][k]
    __shared__ int sRow[FIRST ? kMaxAxis : 1];      // slot of each input row
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
    const int kb = 4 * (lane >> 4);
    const long long L = g.L;
    const long long site0 = i + (long long)g.inner * L * o;   // site of row n: site0 + inner*n
    if constexpr (FIRST) {
        for (int n = t; n < g.L; n += 256) sRow[n] = g.map[site0 + (long long)g.inner * n];
        __syncthreads();
    }
    // per-thread slices of the K-step operand tiles:
    //   A: rows ii = t/4, 8 consecutive k at ka = (t%4)*8, for the 4 matrices
    //   B: rows kk = t/16 and kk+16, 4 consecutive replicas at cb = (t%16)*4
    const int ia = t >> 2, ka = (t & 3) * 8;
    const int kr = t >> 4, cb = (t & 15) * 4;
    uint4 pa[4];
    float4 pr[2], pi[2];
    auto fetch = [&](int a, int k0) {
        for (int w = 0; w < 4; ++w)
            pa[w] = *(const uint4*)(g.w + ((size_t)w * g.Mp + i0 + ia) * g.Kp + k0 + ka);
        for (int h = 0; h < 2; ++h) {
            const int n = k0 + kr + 16 * h;
            float4 vr = make_float4(0.f, 0.f, 0.f, 0.f), vi = vr;
            if (n < g.L) {
                if constexpr (FIRST) {
                    const long long row = (long long)sRow[n] * g.Rp + j0 + cb;
                    if (g.s8) {
                        const char4 c4 = *(const char4*)(g.s8 + row);
                        vr = make_float4((float)c4.x, (float)c4.y, (float)c4.z, (float)c4.w);
                    } else {
                        vr = *(const float4*)(g.plane[a] + row);
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
        Bs[part][0][col][k] = (short)hb;
        Bs[part][1][col][k] = (short)bf16_rne(x - bf16_to_f(hb));
    };
    auto stage = [&]() {
        for (int w = 0; w < 4; ++w) {
            const uint2* src = (const uint2*)&pa[w];
            *(uint2*)&Aw[w][ia][ka] = src[0];
            *(uint2*)&Aw[w][ia][ka + 4] = src[1];
        }
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
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q) nrm[p][q] = (float4_t){0.f, 0.f, 0.f, 0.f};
    const int a_begin = LAST ? 0 : a_only, a_end = LAST ? g.ncomp : a_only + 1;
    const int nsteps = (g.L + BK - 1) / BK;
    fetch(a_begin, 0);
    for (int a = a_begin; a < a_end; ++a) {
        float4_t ur[2][2], ui[2][2];
        for (int p = 0; p < 2; ++p)
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
            for (int k16 = 0; k16 < BK; k16 += 16) {
                // operand fragments of v_mfma_f32_16x16x16_bf16: lane l holds
                // A[l%16][4*(l/16) + v] and B[4*(l/16) + v][l%16], v = 0..3
                const int kf = k16 + kb;
                short4_t ch[2], cl[2], sh[2], sl[2];
                for (int p = 0; p < 2; ++p) {
                    const int row = wi + p * 16 + (lane & 15);
                    ch[p] = *(const short4_t*)&Aw[0][row][kf];
                    cl[p] = *(const short4_t*)&Aw[1][row][kf];
                    sh[p] = *(const short4_t*)&Aw[2][row][kf];
                    sl[p] = *(const short4_t*)&Aw[3][row][kf];
                }
                for (int q = 0; q < 2; ++q) {
                    const int col = wj + q * 16 + (lane & 15);
                    const short4_t rh = *(const short4_t*)&Bs[0][0][col][kf];
                    const short4_t rl = *(const short4_t*)&Bs[0][1][col][kf];
                    short4_t ih, il;
                    if constexpr (!REAL) {
                        ih = *(const short4_t*)&Bs[1][0][col][kf];
                        il = *(const short4_t*)&Bs[1][1][col][kf];
                    }
                    for (int p = 0; p < 2; ++p) {
                        ur[p][q] = mfma3(ch[p], cl[p], rh, rl, ur[p][q]);
                        ui[p][q] = mfma3(sh[p], sl[p], rh, rl, ui[p][q]);
                        if constexpr (!REAL) {
                            ur[p][q] = mfma3(neg4(sh[p]), neg4(sl[p]), ih, il, ur[p][q]);
                            ui[p][q] = mfma3(ch[p], cl[p], ih, il, ui[p][q]);
                        }
                    }
                }
            }
        }
        // result fragment: lane l holds D[4*(l/16) + v][l%16]
        for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q)
                for (int v = 0; v < 4; ++v) {
                    if constexpr (LAST) {
                        nrm[p][q][v] = fmaf(ui[p][q][v], ui[p][q][v],
                                            fmaf(ur[p][q][v], ur[p][q][v], nrm[p][q][v]));
                    } else {
                        const int m = i0 + wi + p * 16 + 4 * (lane >> 4) + v;
                        const int r = j0 + wj + q * 16 + (lane & 15);
                        if (m < g.L) {
                            const long long off = ((long long)a * g.N + site0 + (long long)g.inner * m) * g.Rp + r;
                            g.out_re[off] = ur[p][q][v];
                            g.out_im[off] = ui[p][q][v];
                        }
                    }
                }
    }
    if constexpr (LAST)
        for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q)
                for (int v = 0; v < 4; ++v) {
                    const int m = i0 + wi + p * 16 + 4 * (lane >> 4) + v;
                    const int r = j0 + wj + q * 16 + (lane & 15);
                    if (m < g.L) g.norm[(site0 + (long long)g.inner * m) * g.Rp + r] = nrm[p][q][v];
                }
}

// out[r][q] = norm[q][r] * scale for r < R: 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void k_sq_transpose(const float* norm, int N, int R, int Rp, float scale,
                                                    float* out) {
    __shared__ float tile[32][33];
    const int q0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int q = q0 + k, r = r0 + tx;
        tile[k][tx] = (q < N && r < R) ? norm[(long long)q * Rp + r] : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, q = q0 + tx;
        if (r < R && q < N) out[(long long)r * N + q] = tile[tx][k] * scale;
    }
}

struct SqState {
    int N = 0, R = 0, L[3] = {0, 0, 0};
    float* d_z[4] = {nullptr, nullptr, nullptr, nullptr};  // U re/im planes, ping/pong
    unsigned short* d_w[3] = {nullptr, nullptr, nullptr};   // per-axis split C/S matrices
    float* d_norm = nullptr;
    int* d_map = nullptr;
    float* d_out = nullptr;
};

}  // namespace

namespace ether {
void sq_destroy(void* p) {
    SqState* s = static_cast<SqState*>(p);
    for (int k = 0; k < 4; ++k) (void)hipFree(s->d_z[k]);
    for (int k = 0; k < 3; ++k) (void)hipFree(s->d_w[k]);
    (void)hipFree(s->d_norm);
    (void)hipFree(s->d_map);
    (void)hipFree(s->d_out);
    delete s;
}
}  // namespace ether

#define HC(c, e)                                                                          \
    do {                                                                                  \
        hipError_t _e = (e);                                                              \
        if (_e != hipSuccess) return ctx_fail(c, ETHER_ERR_HIP, std::string(#e) + ": " +  \
                                                                   hipGetErrorString(_e)); \
    } while (0)

extern "C" int ether_structure_factor(ether_ctx* c, const int32_t* sc, float* out) {
    if (!c) return ctx_fail(nullptr, ETHER_ERR_ARG, "null context");
    if (!ctx_committed(c)) return ctx_fail(c, ETHER_ERR_STATE, "context not committed");
    if (!sc || !out) return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor: null pointer");
    const int N = ctx_N(c), R = ctx_R(c), Rp = ctx_Rp(c);
    if (sc[0] <= 0 || sc[1] <= 0 || sc[2] <= 0 || (long long)sc[0] * sc[1] * sc[2] != N)
        return ctx_fail(c, ETHER_ERR_ARG,
                        "ether_structure_factor: needs one site per cell (L1*L2*L3 == N)");
    if ((long long)N * Rp * 3 > (1LL << 31))
        return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor: too many spins");
    if (sc[0] > kMaxAxis || sc[1] > kMaxAxis || sc[2] > kMaxAxis)
        return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor: a lattice axis is longer than 2048 cells");
    const bool ising = ctx_model(c) == ETHER_MODEL_ISING;
    const int ncomp = ising ? 1 : 3;
    hipStream_t st = ctx_stream(c);
    SqState* s = static_cast<SqState*>(ctx_sq(c));
    bool fresh = !s || s->N != N || s->R != R || s->L[0] != sc[0] || s->L[1] != sc[1] ||
                 s->L[2] != sc[2];
    if (fresh) {
        if (s) sq_destroy(s);
        s = new SqState();
        ctx_sq(c) = s;
        s->N = N;
        s->R = R;
        for (int k = 0; k < 3; ++k) s->L[k] = sc[k];
        const size_t plane = (size_t)N * Rp * 3;
        for (int k = 0; k < 4; ++k) HC(c, hipMalloc(&s->d_z[k], plane * sizeof(float)));
        HC(c, hipMalloc(&s->d_norm, sizeof(float) * (size_t)N * Rp));
        HC(c, hipMalloc(&s->d_map, sizeof(int) * N));
        HC(c, hipMalloc(&s->d_out, sizeof(float) * (size_t)N * R));
        for (int k = 0; k < 3; ++k) {
            const int L = sc[k];
            const size_t Mp = ((size_t)L + BM - 1) / BM * BM, Kp = ((size_t)L + BK - 1) / BK * BK;
            std::vector<unsigned short> w(4 * Mp * Kp, 0);
            for (int m = 0; m < L; ++m)
                for (int n = 0; n < L; ++n) {
                    // exact argument reduction: (m*n mod L) / L
                    const double th = 2.0 * M_PI * (double)((long long)m * n % L) / L;
                    const float cv = (float)std::cos(th), sv = (float)std::sin(th);
                    const size_t e = (size_t)m * Kp + n, P = Mp * Kp;
                    w[e] = bf16_rne(cv);
                    w[P + e] = bf16_rne(cv - bf16_to_f(w[e]));
                    w[2 * P + e] = bf16_rne(sv);
                    w[3 * P + e] = bf16_rne(sv - bf16_to_f(w[2 * P + e]));
                }
            HC(c, hipMalloc(&s->d_w[k], sizeof(unsigned short) * w.size()));
            HC(c, hipMemcpyAsync(s->d_w[k], w.data(), sizeof(unsigned short) * w.size(),
                                 hipMemcpyHostToDevice, st));
        }
        HC(c, hipMemcpyAsync(s->d_map, ctx_slot_of_site(c), sizeof(int) * N, hipMemcpyHostToDevice, st));
        HC(c, hipStreamSynchronize(st));   // the host staging vectors go out of scope
    }
    const int ev = ctx_ev_begin(c, ETHER_TIME_SQ);
    // the transformed axes (length > 1), in order; a 1 x 1 x 1 lattice still
    // needs one stage to form the norm
    int axes[3], na = 0;
    for (int k = 0; k < 3; ++k)
        if (sc[k] > 1) axes[na++] = k;
    if (na == 0) axes[na++] = 0;
    int cur = -1;   // index of the input U pair (0 or 2); -1: the spin planes
    for (int st_i = 0; st_i < na; ++st_i) {
        const int k = axes[st_i];
        DftArgs g{};
        g.L = sc[k];
        g.Mp = (g.L + BM - 1) / BM * BM;
        g.Kp = (g.L + BK - 1) / BK * BK;
        g.inner = 1;
        for (int q = 0; q < k; ++q) g.inner *= sc[q];
        g.outer = N / (g.inner * g.L);
        g.ncomp = ncomp;
        g.N = N;
        g.Rp = Rp;
        const bool first = st_i == 0, last = st_i == na - 1;
        g.w = s->d_w[k];
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
        g.tiles_m = (g.L + BM - 1) / BM;
        g.tiles_n = (Rp + BN - 1) / BN;
        const long long blocks =
            (long long)g.tiles_m * g.tiles_n * g.inner * g.outer * (last ? 1 : ncomp);
        if (blocks > 0x7fffffffLL) return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor: grid too large");
        const dim3 grid((unsigned)blocks), block(256);
        if (first && last) hipLaunchKernelGGL((k_dft_stage<true, true>), grid, block, 0, st, g);
        else if (first) hipLaunchKernelGGL((k_dft_stage<true, false>), grid, block, 0, st, g);
        else if (last) hipLaunchKernelGGL((k_dft_stage<false, true>), grid, block, 0, st, g);
        else hipLaunchKernelGGL((k_dft_stage<false, false>), grid, block, 0, st, g);
        HC(c, hipGetLastError());
        cur = nxt;
    }
    hipLaunchKernelGGL(k_sq_transpose, dim3((N + 31) / 32, (R + 31) / 32), dim3(256), 0, st, s->d_norm, N, R,
                       Rp, (float)(1.0 / ((double)N * N)), s->d_out);
    ctx_ev_end(c, ev);
    HC(c, hipGetLastError());
    HC(c, hipMemcpyAsync(out, s->d_out, sizeof(float) * (size_t)N * R, hipMemcpyDeviceToHost, st));
    HC(c, hipStreamSynchronize(st));
    return ETHER_OK;
}
