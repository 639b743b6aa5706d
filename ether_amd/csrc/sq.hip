// sq.hip - static structure factor on the matrix cores.
//
// Reference: utility/Structure_Factor/structure_factor.f90:150-177 computes
//   S_a(q) = sum_r S_a(r) exp(i q.r),  out = sum_a |S_a(q)|^2 / N^2
// by a direct O(Q*N) sum with complex exponentials.  For a one-site-per-cell
// supercell and the commensurate grid q = 2*pi*sum_k m_k b_k / L_k
// (a_i.b_j = delta_ij) the phase factorises, q.r = 2*pi*sum_k m_k n_k / L_k,
// so S_a is a 3-D DFT of each spin plane.  It is evaluated as one dense
// contraction per lattice axis with the L_k x L_k cos/sin matrices:
//   U_re = Z_re x C - Z_im x S,   U_im = Z_re x S + Z_im x C   (along axis k)
// Each contraction is a strided, batched GEMM on the MFMA units.  Operands are
// split into BF16 hi/lo parts (x = hi + lo, hi = bf16(x), lo = bf16(x - hi))
// and the product is hi*hi + hi*lo + lo*hi accumulated in FP32
// (v_mfma_f32_16x16x16_bf16), which keeps about 16 mantissa bits.
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

constexpr int BM = 64, BN = 64, BK = 16;

struct GemmArgs {
    // C[b](i,j) = alpha * sum_k A[b](i,k) B[b](k,j) + beta * C[b](i,j)
    const float* A;
    const float* B;
    float* C;
    int M, N, K;
    long long a_rs, a_cs, a_bs;
    long long b_rs, b_cs, b_bs;
    long long c_rs, c_cs, c_bs;
    int batch;
    float alpha, beta;
};

__device__ __forceinline__ unsigned short bf16_rne(float x) {
    unsigned int u = __float_as_uint(x);
    unsigned int r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(r >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) {
    return __uint_as_float(((unsigned int)h) << 16);
}

__global__ __launch_bounds__(256) void k_gemm_split_bf16(GemmArgs g) {
    __shared__ float As[BM][BK + 1];
    __shared__ float Bs[BK][BN + 1];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * BM, j0 = blockIdx.x * BN;
    const float* A = g.A + (long long)b * g.a_bs;
    const float* B = g.B + (long long)b * g.b_bs;
    float* C = g.C + (long long)b * g.c_bs;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wi = (wv >> 1) * 32, wj = (wv & 1) * 32;
    float4_t acc[2][2];
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q) acc[p][q] = (float4_t){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        for (int e = t; e < BM * BK; e += 256) {
            int ii = e / BK, kk = e % BK;
            int gi = i0 + ii, gk = k0 + kk;
            As[ii][kk] = (gi < g.M && gk < g.K) ? A[gi * g.a_rs + gk * g.a_cs] : 0.f;
        }
        for (int e = t; e < BK * BN; e += 256) {
            int kk = e / BN, jj = e % BN;
            int gk = k0 + kk, gj = j0 + jj;
            Bs[kk][jj] = (gk < g.K && gj < g.N) ? B[gk * g.b_rs + gj * g.b_cs] : 0.f;
        }
        __syncthreads();
        // operand fragments of v_mfma_f32_16x16x16_bf16: lane l holds
        // A[l%16][4*(l/16) + v] and B[4*(l/16) + v][l%16], v = 0..3
        const int kb = 4 * (lane >> 4);
        short4_t ah[2], al[2], bh[2], bl[2];
        for (int p = 0; p < 2; ++p) {
            const int row = wi + p * 16 + (lane & 15);
            for (int v = 0; v < 4; ++v) {
                float x = As[row][kb + v];
                unsigned short h = bf16_rne(x);
                ah[p][v] = (short)h;
                al[p][v] = (short)bf16_rne(x - bf16_to_f(h));
            }
        }
        for (int q = 0; q < 2; ++q) {
            const int col = wj + q * 16 + (lane & 15);
            for (int v = 0; v < 4; ++v) {
                float x = Bs[kb + v][col];
                unsigned short h = bf16_rne(x);
                bh[q][v] = (short)h;
                bl[q][v] = (short)bf16_rne(x - bf16_to_f(h));
            }
        }
        for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q) {
                acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[p], bh[q], acc[p][q], 0, 0, 0);
                acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[p], bl[q], acc[p][q], 0, 0, 0);
                acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[p], bh[q], acc[p][q], 0, 0, 0);
            }
        __syncthreads();
    }
    // result fragment: lane l holds D[4*(l/16) + v][l%16]
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q)
            for (int v = 0; v < 4; ++v) {
                int gi = i0 + wi + p * 16 + 4 * (lane >> 4) + v;
                int gj = j0 + wj + q * 16 + (lane & 15);
                if (gi < g.M && gj < g.N) {
                    float* dst = C + gi * g.c_rs + gj * g.c_cs;
                    float val = g.alpha * acc[p][q][v];
                    *dst = (g.beta != 0.f) ? val + g.beta * (*dst) : val;
                }
            }
}

// The batch goes on grid.z, which holds at most 65535 blocks: larger batches
// are launched in chunks with the operand pointers advanced.
void gemm(const GemmArgs& g, hipStream_t st) {
    constexpr int kMaxZ = 65535;
    for (int b0 = 0; b0 < g.batch; b0 += kMaxZ) {
        GemmArgs h = g;
        h.batch = std::min(kMaxZ, g.batch - b0);
        h.A = g.A + (long long)b0 * g.a_bs;
        h.B = g.B + (long long)b0 * g.b_bs;
        h.C = g.C + (long long)b0 * g.c_bs;
        dim3 grid((h.N + BN - 1) / BN, (h.M + BM - 1) / BM, h.batch);
        hipLaunchKernelGGL(k_gemm_split_bf16, grid, dim3(256), 0, st, h);
    }
}

// gather spins of every replica (replica-major device planes, row stride Rp)
// into site-ordered planes Z[r*3 + a][site]
__global__ void k_gather(const float* sx, const float* sy, const float* sz, const int8_t* s,
                         const int* slot_of_site, int N, int R, int Rp, float* Z) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * R) return;
    int r = (int)(t / N), i = (int)(t % N);
    long long src = (long long)slot_of_site[i] * Rp + r;
    if (s) {
        Z[((long long)r * 3 + 0) * N + i] = 0.f;
        Z[((long long)r * 3 + 1) * N + i] = 0.f;
        Z[((long long)r * 3 + 2) * N + i] = (float)s[src];
    } else {
        Z[((long long)r * 3 + 0) * N + i] = sx[src];
        Z[((long long)r * 3 + 1) * N + i] = sy[src];
        Z[((long long)r * 3 + 2) * N + i] = sz[src];
    }
}

__global__ void k_sq_norm(const float* Ur, const float* Ui, int N, int R, float inv_n2, float* out) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * R) return;
    int r = (int)(t / N), q = (int)(t % N);
    float acc = 0.f;
    for (int a = 0; a < 3; ++a) {
        long long o = ((long long)r * 3 + a) * N + q;
        acc += Ur[o] * Ur[o] + Ui[o] * Ui[o];
    }
    out[t] = acc * inv_n2;
}

struct SqState {
    int N = 0, R = 0, L[3] = {0, 0, 0};
    float* d_z[4] = {nullptr, nullptr, nullptr, nullptr};  // re/im planes, ping/pong
    float* d_cs[3] = {nullptr, nullptr, nullptr};           // per-axis cos matrices
    float* d_sn[3] = {nullptr, nullptr, nullptr};           // per-axis sin matrices
    int* d_map = nullptr;
    float* d_out = nullptr;
};

}  // namespace

namespace ether {
void sq_destroy(void* p) {
    SqState* s = static_cast<SqState*>(p);
    for (int k = 0; k < 4; ++k) (void)hipFree(s->d_z[k]);
    for (int k = 0; k < 3; ++k) {
        (void)hipFree(s->d_cs[k]);
        (void)hipFree(s->d_sn[k]);
    }
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

// Apply U = Z x_k F along axis k (complex, planar) with GEMMs.
static void dft_axis(const SqState* s, int k, const float* zr, const float* zi, float* ur, float* ui,
                     bool zi_zero, hipStream_t st) {
    const int L = s->L[k];
    long long inner = 1;
    for (int q = 0; q < k; ++q) inner *= s->L[q];
    const long long total = (long long)s->N * s->R * 3;
    const long long outer = total / (inner * L);
    const float* Cm = s->d_cs[k];
    const float* Sm = s->d_sn[k];
    auto run = [&](const float* z, const float* F, float* u, float alpha, float beta) {
        GemmArgs g;
        if (k == 0) {
            // U[o][m] = sum_n Z[o][n] F[n][m]
            g.A = z; g.a_rs = L; g.a_cs = 1; g.a_bs = 0;
            g.B = F; g.b_rs = L; g.b_cs = 1; g.b_bs = 0;
            g.C = u; g.c_rs = L; g.c_cs = 1; g.c_bs = 0;
            g.M = (int)outer; g.N = L; g.K = L; g.batch = 1;
        } else {
            // for each outer o: U_o[m][i] = sum_n F[n][m] Z_o[n][i]
            g.A = F; g.a_rs = 1; g.a_cs = L; g.a_bs = 0;
            g.B = z; g.b_rs = inner; g.b_cs = 1; g.b_bs = L * inner;
            g.C = u; g.c_rs = inner; g.c_cs = 1; g.c_bs = L * inner;
            g.M = L; g.N = (int)inner; g.K = L; g.batch = (int)outer;
        }
        g.alpha = alpha;
        g.beta = beta;
        gemm(g, st);
    };
    run(zr, Cm, ur, 1.f, 0.f);
    if (!zi_zero) run(zi, Sm, ur, -1.f, 1.f);
    run(zr, Sm, ui, 1.f, 0.f);
    if (!zi_zero) run(zi, Cm, ui, 1.f, 1.f);
}

extern "C" int ether_structure_factor(ether_ctx* c, const int32_t* sc, float* out) {
    if (!c) return ctx_fail(nullptr, ETHER_ERR_ARG, "null context");
    if (!ctx_committed(c)) return ctx_fail(c, ETHER_ERR_STATE, "context not committed");
    if (!sc || !out) return ctx_fail(c, ETHER_ERR_ARG, "ether_structure_factor: null pointer");
    const int N = ctx_N(c), R = ctx_R(c);
    if (sc[0] <= 0 || sc[1] <= 0 || sc[2] <= 0 || (long long)sc[0] * sc[1] * sc[2] != N)
        return ctx_fail(c, ETHER_ERR_ARG,
                        "ether_structure_factor: needs one site per cell (L1*L2*L3 == N)");
    if ((long long)N * R * 3 > (1LL << 31))
        return ctx_fail(c, ETHER_ERR_LIMIT, "ether_structure_factor: too many spins");
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
        const size_t plane = (size_t)N * R * 3;
        for (int k = 0; k < 4; ++k) HC(c, hipMalloc(&s->d_z[k], plane * sizeof(float)));
        HC(c, hipMalloc(&s->d_map, sizeof(int) * N));
        HC(c, hipMalloc(&s->d_out, sizeof(float) * (size_t)N * R));
        for (int k = 0; k < 3; ++k) {
            const int L = sc[k];
            std::vector<float> cs((size_t)L * L), sn((size_t)L * L);
            for (int n = 0; n < L; ++n)
                for (int m = 0; m < L; ++m) {
                    // exact argument reduction: (m*n mod L) / L
                    double th = 2.0 * M_PI * (double)((long long)m * n % L) / L;
                    cs[(size_t)n * L + m] = (float)std::cos(th);
                    sn[(size_t)n * L + m] = (float)std::sin(th);
                }
            HC(c, hipMalloc(&s->d_cs[k], sizeof(float) * L * L));
            HC(c, hipMalloc(&s->d_sn[k], sizeof(float) * L * L));
            HC(c, hipMemcpyAsync(s->d_cs[k], cs.data(), sizeof(float) * L * L, hipMemcpyHostToDevice, st));
            HC(c, hipMemcpyAsync(s->d_sn[k], sn.data(), sizeof(float) * L * L, hipMemcpyHostToDevice, st));
        }
        HC(c, hipMemcpyAsync(s->d_map, ctx_slot_of_site(c), sizeof(int) * N, hipMemcpyHostToDevice, st));
    }
    const long long tot = (long long)N * R;
    const bool ising = ctx_model(c) == ETHER_MODEL_ISING;
    const int ev = ctx_ev_begin(c, ETHER_TIME_SQ);
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       ctx_spin_plane(c, 0), ctx_spin_plane(c, 1), ctx_spin_plane(c, 2),
                       ising ? ctx_ising(c) : (const int8_t*)nullptr, s->d_map, N, R, ctx_Rp(c),
                       s->d_z[0]);
    HC(c, hipGetLastError());
    // ping-pong: (z0 re, z1 im) -> (z2, z3) -> (z0, z1) ...
    int cur = 0;
    bool zi_zero = true;
    for (int k = 0; k < 3; ++k) {
        if (sc[k] == 1) continue;
        int nxt = cur ^ 2;
        dft_axis(s, k, s->d_z[cur], s->d_z[cur + 1], s->d_z[nxt], s->d_z[nxt + 1], zi_zero, st);
        zi_zero = false;
        cur = nxt;
    }
    HC(c, hipGetLastError());
    if (zi_zero) HC(c, hipMemsetAsync(s->d_z[cur + 1], 0, sizeof(float) * N * R * 3, st));
    hipLaunchKernelGGL(k_sq_norm, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, s->d_z[cur],
                       s->d_z[cur + 1], N, R, (float)(1.0 / ((double)N * N)), s->d_out);
    ctx_ev_end(c, ev);
    HC(c, hipGetLastError());
    HC(c, hipMemcpyAsync(out, s->d_out, sizeof(float) * N * R, hipMemcpyDeviceToHost, st));
    HC(c, hipStreamSynchronize(st));
    return ETHER_OK;
}
