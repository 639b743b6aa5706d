// mlp.hip - learned tempering ladder: a 1-64-64-2 tanh MLP surrogate of
// (<E>(T), log C_v(T)) trained on the engine's own measurements, and the
// placement of a ladder with uniform expected swap acceptance.
//
// No reference counterpart: the reference's ladder is linear
// (src/ether.f90:79-84) and its adaptive-ladder idea is a commented stub
// (src/startup.f90:115-122).  The ladder only chooses temperatures; every
// swap still passes the exact acceptance test, so the sampled distribution
// is unchanged.
//
// Kernels (one workgroup of 256 lanes; the network has 4418 parameters and
// the batch is one sample per ladder temperature, so this is latency work):
//   k_mlp_forward    inference, one lane per sample, weights staged in LDS
//   k_mlp_train      `iters` full-batch Adam steps in one launch; each lane
//                    owns a fixed set of parameters and accumulates their
//                    gradients in registers over the batch (deterministic,
//                    no atomics); per-sample activations are shared in LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/ether_gpu.h"
#include "philox.h"

namespace ether {
hipStream_t ctx_stream(ether_ctx* c);
void*& ctx_mlp(ether_ctx* c);
void*& ctx_sq(ether_ctx* c);
int ctx_fail(ether_ctx* c, int code, const std::string& msg);
void sq_destroy(void* p);
}  // namespace ether

using namespace ether;

namespace {

constexpr int H = 64;
constexpr int NP = ETHER_MLP_NPARAM;  // 4418
constexpr int kT = 256;
constexpr int kOwn = (NP + kT - 1) / kT;  // parameters owned per lane (18)
constexpr int OFF_W1 = 0, OFF_B1 = H, OFF_W2 = 2 * H, OFF_B2 = 2 * H + H * H,
              OFF_W3 = 3 * H + H * H, OFF_B3 = 5 * H + H * H;
static_assert(OFF_B3 + 2 == NP, "parameter layout");

struct MlpState {
    float* d_p = nullptr;      // parameters
    float* d_m = nullptr;      // Adam first moments
    float* d_v = nullptr;      // Adam second moments
    float* d_x = nullptr;      // batch buffers
    float* d_y = nullptr;
    float* d_w = nullptr;
    float* d_out = nullptr;    // forward outputs / grads / loss
    int cap = 0;
    int t = 0;                 // Adam step counter
};

// Forward + backward of one sample with 256 lanes; activations in LDS.
// Returns the sample's weighted squared error in lane 0.
__device__ float sample_pass(const float* P, float x, float y0, float y1, float w0, float w1,
                             float inv_n, float* h1, float* h2, float* d1, float* d2,
                             float* dout, bool backward) {
    const int t = threadIdx.x;
    if (t < H) h1[t] = tanhf(P[OFF_W1 + t] * x + P[OFF_B1 + t]);
    __syncthreads();
    if (t < H) {
        float a = P[OFF_B2 + t];
        const float* row = P + OFF_W2 + t * H;
#pragma unroll 16
        for (int i = 0; i < H; ++i) a = fmaf(row[i], h1[i], a);
        h2[t] = tanhf(a);
    }
    __syncthreads();
    float err = 0.f;
    if (t < 2) {
        float a = P[OFF_B3 + t];
        const float* row = P + OFF_W3 + t * H;
        for (int i = 0; i < H; ++i) a = fmaf(row[i], h2[i], a);
        float r = a - (t == 0 ? y0 : y1);
        float w = t == 0 ? w0 : w1;
        dout[t] = 2.f * w * r * inv_n;
        dout[2 + t] = a;
        err = w * r * r;
    }
    __syncthreads();
    if (!backward) return err;
    if (t < H) d2[t] = (dout[0] * P[OFF_W3 + t] + dout[1] * P[OFF_W3 + H + t]) * (1.f - h2[t] * h2[t]);
    __syncthreads();
    if (t < H) {
        float g = 0.f;
        for (int o = 0; o < H; ++o) g = fmaf(d2[o], P[OFF_W2 + o * H + t], g);
        d1[t] = g * (1.f - h1[t] * h1[t]);
    }
    __syncthreads();
    return err;
}

// gradient contribution of parameter p for the sample whose activations are in LDS
__device__ __forceinline__ float param_grad(int p, float x, const float* h1, const float* h2,
                                            const float* d1, const float* d2, const float* dout) {
    if (p < OFF_B1) return d1[p] * x;
    if (p < OFF_W2) return d1[p - OFF_B1];
    if (p < OFF_B2) {
        int q = p - OFF_W2;
        return d2[q / H] * h1[q % H];
    }
    if (p < OFF_W3) return d2[p - OFF_B2];
    if (p < OFF_B3) {
        int q = p - OFF_W3;
        return dout[q / H] * h2[q % H];
    }
    return dout[p - OFF_B3];
}

__global__ __launch_bounds__(kT) void k_mlp_train(float* __restrict__ params, float* __restrict__ m,
                                                  float* __restrict__ v, const float* __restrict__ X,
                                                  const float* __restrict__ Y,
                                                  const float* __restrict__ W, int n, int iters,
                                                  float lr, int t0, float* __restrict__ loss_out,
                                                  float* __restrict__ grad_out) {
    __shared__ float P[NP];
    __shared__ float h1[H], h2[H], d1[H], d2[H], dout[4];
    __shared__ float red[kT / 64];
    const int t = threadIdx.x;
    for (int p = t; p < NP; p += kT) P[p] = params[p];
    float mo[kOwn], vo[kOwn];
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
        int p = t + q * kT;
        mo[q] = (p < NP) ? m[p] : 0.f;
        vo[q] = (p < NP) ? v[p] : 0.f;
    }
    __syncthreads();
    const float inv_n = 1.0f / (float)n;
    float loss = 0.f;
    for (int it = 0; it < iters; ++it) {
        float g[kOwn];
#pragma unroll
        for (int q = 0; q < kOwn; ++q) g[q] = 0.f;
        float lsum = 0.f;
        for (int s = 0; s < n; ++s) {
            const float x = X[s];
            lsum += sample_pass(P, x, Y[2 * s], Y[2 * s + 1], W[2 * s], W[2 * s + 1], inv_n, h1, h2,
                                d1, d2, dout, true);
#pragma unroll
            for (int q = 0; q < kOwn; ++q) {
                int p = t + q * kT;
                if (p < NP) g[q] += param_grad(p, x, h1, h2, d1, d2, dout);
            }
            __syncthreads();
        }
        // loss of this iteration (before the update)
        float ls = lsum;
        for (int o = 32; o > 0; o >>= 1) ls += __shfl_xor(ls, o, 64);
        if ((t & 63) == 0) red[t >> 6] = ls;
        __syncthreads();
        loss = (red[0] + red[1] + red[2] + red[3]) * inv_n;
        if (grad_out) {
#pragma unroll
            for (int q = 0; q < kOwn; ++q) {
                int p = t + q * kT;
                if (p < NP && grad_out) grad_out[p] = g[q];
            }
        }
        if (lr > 0.f) {
            // Adam (Kingma & Ba 2015), beta1 0.9, beta2 0.999, eps 1e-8
            const int step = t0 + it + 1;
            const float c1 = 1.f - powf(0.9f, (float)step);
            const float c2 = 1.f - powf(0.999f, (float)step);
#pragma unroll
            for (int q = 0; q < kOwn; ++q) {
                int p = t + q * kT;
                if (p < NP) {
                    mo[q] = 0.9f * mo[q] + 0.1f * g[q];
                    vo[q] = 0.999f * vo[q] + 0.001f * g[q] * g[q];
                    float mh = mo[q] / c1, vh = vo[q] / c2;
                    P[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
                }
            }
        }
        __syncthreads();
    }
    for (int p = t; p < NP; p += kT) params[p] = P[p];
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
        int p = t + q * kT;
        if (p < NP) {
            m[p] = mo[q];
            v[p] = vo[q];
        }
    }
    if (t == 0 && loss_out) *loss_out = loss;
}

__global__ __launch_bounds__(kT) void k_mlp_forward(const float* __restrict__ params,
                                                    const float* __restrict__ X, int n,
                                                    float* __restrict__ Yout) {
    __shared__ float P[NP];
    for (int p = threadIdx.x; p < NP; p += kT) P[p] = params[p];
    __syncthreads();
    const int s = blockIdx.x * kT + threadIdx.x;
    if (s >= n) return;
    const float x = X[s];
    float h1[H];
#pragma unroll 8
    for (int i = 0; i < H; ++i) h1[i] = tanhf(P[OFF_W1 + i] * x + P[OFF_B1 + i]);
    float o0 = P[OFF_B3], o1 = P[OFF_B3 + 1];
    for (int o = 0; o < H; ++o) {
        float a = P[OFF_B2 + o];
        const float* row = P + OFF_W2 + o * H;
#pragma unroll 16
        for (int i = 0; i < H; ++i) a = fmaf(row[i], h1[i], a);
        float h2 = tanhf(a);
        o0 = fmaf(P[OFF_W3 + o], h2, o0);
        o1 = fmaf(P[OFF_W3 + H + o], h2, o1);
    }
    Yout[2 * s] = o0;
    Yout[2 * s + 1] = o1;
}

MlpState* state(ether_ctx* c) { return static_cast<MlpState*>(ctx_mlp(c)); }

int ensure_batch(ether_ctx* c, MlpState* s, int n) {
    if (n <= s->cap) return ETHER_OK;
    (void)hipFree(s->d_x);
    (void)hipFree(s->d_y);
    (void)hipFree(s->d_w);
    (void)hipFree(s->d_out);
    s->d_x = s->d_y = s->d_w = s->d_out = nullptr;
    int cap = std::max(n, 256);
    if (hipMalloc(&s->d_x, sizeof(float) * cap) != hipSuccess ||
        hipMalloc(&s->d_y, sizeof(float) * 2 * cap) != hipSuccess ||
        hipMalloc(&s->d_w, sizeof(float) * 2 * cap) != hipSuccess ||
        hipMalloc(&s->d_out, sizeof(float) * std::max(2 * cap, NP + 1)) != hipSuccess)
        return ctx_fail(c, ETHER_ERR_HIP, "mlp: hipMalloc failed");
    s->cap = cap;
    return ETHER_OK;
}

int need_mlp(ether_ctx* c) {
    if (!c) return ctx_fail(nullptr, ETHER_ERR_ARG, "null context");
    if (!state(c)) return ctx_fail(c, ETHER_ERR_STATE, "MLP not initialised (call ether_mlp_init)");
    return ETHER_OK;
}

#define HCHK(c, e)                                                                        \
    do {                                                                                  \
        hipError_t _e = (e);                                                              \
        if (_e != hipSuccess) return ctx_fail(c, ETHER_ERR_HIP, std::string(#e) + ": " +  \
                                                                   hipGetErrorString(_e)); \
    } while (0)

int upload_batch(ether_ctx* c, MlpState* s, int n, const float* x, const float* y, const float* w) {
    hipStream_t st = ctx_stream(c);
    int rc = ensure_batch(c, s, n);
    if (rc) return rc;
    HCHK(c, hipMemcpyAsync(s->d_x, x, sizeof(float) * n, hipMemcpyHostToDevice, st));
    if (y) HCHK(c, hipMemcpyAsync(s->d_y, y, sizeof(float) * 2 * n, hipMemcpyHostToDevice, st));
    if (w) HCHK(c, hipMemcpyAsync(s->d_w, w, sizeof(float) * 2 * n, hipMemcpyHostToDevice, st));
    return ETHER_OK;
}

}  // namespace

namespace ether {
void ether_ext_destroy_impl(ether_ctx* c) {
    MlpState* s = state(c);
    if (s) {
        (void)hipFree(s->d_p);
        (void)hipFree(s->d_m);
        (void)hipFree(s->d_v);
        (void)hipFree(s->d_x);
        (void)hipFree(s->d_y);
        (void)hipFree(s->d_w);
        (void)hipFree(s->d_out);
        delete s;
        ctx_mlp(c) = nullptr;
    }
    if (ctx_sq(c)) {
        sq_destroy(ctx_sq(c));
        ctx_sq(c) = nullptr;
    }
}
}  // namespace ether

void ether_ext_destroy(ether_ctx* c) { ether::ether_ext_destroy_impl(c); }

extern "C" int ether_mlp_init(ether_ctx* c, uint64_t seed) {
    if (!c) return ctx_fail(nullptr, ETHER_ERR_ARG, "null context");
    MlpState* s = state(c);
    if (!s) {
        s = new MlpState();
        ctx_mlp(c) = s;
        if (hipMalloc(&s->d_p, sizeof(float) * NP) != hipSuccess ||
            hipMalloc(&s->d_m, sizeof(float) * NP) != hipSuccess ||
            hipMalloc(&s->d_v, sizeof(float) * NP) != hipSuccess)
            return ctx_fail(c, ETHER_ERR_HIP, "mlp: hipMalloc failed");
    }
    // Glorot-uniform weights, zero biases, drawn from the counter-based generator
    std::vector<float> p(NP, 0.f);
    auto uni = [&](uint32_t k) {
        u32x4 w = rng_draw(seed, k, 0xffffffffu, 0, 7, 0);
        return (float)((double)(w.x >> 8) * 0x1p-23 - 1.0);
    };
    const float a1 = std::sqrt(6.f / (1 + H)), a2 = std::sqrt(6.f / (H + H)), a3 = std::sqrt(6.f / (H + 2));
    for (int i = 0; i < H; ++i) p[OFF_W1 + i] = a1 * uni(OFF_W1 + i);
    for (int i = 0; i < H * H; ++i) p[OFF_W2 + i] = a2 * uni(OFF_W2 + i);
    for (int i = 0; i < 2 * H; ++i) p[OFF_W3 + i] = a3 * uni(OFF_W3 + i);
    hipStream_t st = ctx_stream(c);
    HCHK(c, hipMemcpyAsync(s->d_p, p.data(), sizeof(float) * NP, hipMemcpyHostToDevice, st));
    HCHK(c, hipMemsetAsync(s->d_m, 0, sizeof(float) * NP, st));
    HCHK(c, hipMemsetAsync(s->d_v, 0, sizeof(float) * NP, st));
    HCHK(c, hipStreamSynchronize(st));
    s->t = 0;
    return ETHER_OK;
}

extern "C" int ether_mlp_set_params(ether_ctx* c, const float* params) {
    int rc = need_mlp(c);
    if (rc) return rc;
    if (!params) return ctx_fail(c, ETHER_ERR_ARG, "ether_mlp_set_params: null");
    MlpState* s = state(c);
    hipStream_t st = ctx_stream(c);
    HCHK(c, hipMemcpyAsync(s->d_p, params, sizeof(float) * NP, hipMemcpyHostToDevice, st));
    HCHK(c, hipMemsetAsync(s->d_m, 0, sizeof(float) * NP, st));
    HCHK(c, hipMemsetAsync(s->d_v, 0, sizeof(float) * NP, st));
    HCHK(c, hipStreamSynchronize(st));
    s->t = 0;
    return ETHER_OK;
}

extern "C" int ether_mlp_get_params(ether_ctx* c, float* params) {
    int rc = need_mlp(c);
    if (rc) return rc;
    if (!params) return ctx_fail(c, ETHER_ERR_ARG, "ether_mlp_get_params: null");
    hipStream_t st = ctx_stream(c);
    HCHK(c, hipMemcpyAsync(params, state(c)->d_p, sizeof(float) * NP, hipMemcpyDeviceToHost, st));
    HCHK(c, hipStreamSynchronize(st));
    return ETHER_OK;
}

extern "C" int ether_mlp_loss_grad(ether_ctx* c, int32_t n, const float* x, const float* y,
                                   const float* w, float* loss, float* grad) {
    int rc = need_mlp(c);
    if (rc) return rc;
    if (n <= 0 || !x || !y || !w) return ctx_fail(c, ETHER_ERR_ARG, "ether_mlp_loss_grad: bad args");
    MlpState* s = state(c);
    if ((rc = upload_batch(c, s, n, x, y, w))) return rc;
    hipStream_t st = ctx_stream(c);
    // one iteration with lr = 0: gradients and loss, parameters unchanged
    hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kT), 0, st, s->d_p, s->d_m, s->d_v, s->d_x,
                       s->d_y, s->d_w, n, 1, 0.0f, s->t, s->d_out + NP, s->d_out);
    HCHK(c, hipGetLastError());
    std::vector<float> buf(NP + 1);
    HCHK(c, hipMemcpyAsync(buf.data(), s->d_out, sizeof(float) * (NP + 1), hipMemcpyDeviceToHost, st));
    HCHK(c, hipStreamSynchronize(st));
    if (grad) std::copy(buf.begin(), buf.begin() + NP, grad);
    if (loss) *loss = buf[NP];
    return ETHER_OK;
}

extern "C" int ether_mlp_forward(ether_ctx* c, int32_t n, const float* x, float* y) {
    int rc = need_mlp(c);
    if (rc) return rc;
    if (n <= 0 || !x || !y) return ctx_fail(c, ETHER_ERR_ARG, "ether_mlp_forward: bad args");
    MlpState* s = state(c);
    if ((rc = upload_batch(c, s, n, x, nullptr, nullptr))) return rc;
    hipStream_t st = ctx_stream(c);
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + kT - 1) / kT), dim3(kT), 0, st, s->d_p, s->d_x, n,
                       s->d_out);
    HCHK(c, hipGetLastError());
    HCHK(c, hipMemcpyAsync(y, s->d_out, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, st));
    HCHK(c, hipStreamSynchronize(st));
    return ETHER_OK;
}

extern "C" int ether_mlp_train(ether_ctx* c, int32_t n, const float* x, const float* y,
                               const float* w, int32_t iters, float lr, float* final_loss) {
    int rc = need_mlp(c);
    if (rc) return rc;
    if (n <= 0 || iters <= 0 || !x || !y || !w || !(lr > 0.f))
        return ctx_fail(c, ETHER_ERR_ARG, "ether_mlp_train: bad args");
    MlpState* s = state(c);
    if ((rc = upload_batch(c, s, n, x, y, w))) return rc;
    hipStream_t st = ctx_stream(c);
    hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kT), 0, st, s->d_p, s->d_m, s->d_v, s->d_x,
                       s->d_y, s->d_w, n, iters, lr, s->t, s->d_out + NP, (float*)nullptr);
    HCHK(c, hipGetLastError());
    s->t += iters;
    float l = 0.f;
    HCHK(c, hipMemcpyAsync(&l, s->d_out + NP, sizeof(float), hipMemcpyDeviceToHost, st));
    HCHK(c, hipStreamSynchronize(st));
    if (final_loss) *final_loss = l;
    return ETHER_OK;
}

extern "C" int ether_ladder_place(ether_ctx* c, double x_mid, double x_half, double logcv_scale,
                                  double t_lo, double t_hi, int32_t n_temps, int32_t mode,
                                  double* T_out) {
    const bool log_t = (mode & ETHER_PLACE_LOG_T) != 0, length = (mode & ETHER_PLACE_LENGTH) != 0;
    int rc = need_mlp(c);
    if (rc) return rc;
    if (!T_out || n_temps < 2 || !(t_hi > t_lo) || !(t_lo > 0.0) || !(x_half > 0.0))
        return ctx_fail(c, ETHER_ERR_ARG, "ether_ladder_place: bad arguments");
    // surrogate on a fine grid, uniform in log T
    const int G = 4096;
    std::vector<float> xg(G), yg(2 * G);
    std::vector<double> tg(G);
    for (int g = 0; g < G; ++g) {
        tg[g] = t_lo * std::pow(t_hi / t_lo, (double)g / (G - 1));
        xg[g] = (float)(((log_t ? std::log(tg[g]) : tg[g]) - x_mid) / x_half);
    }
    rc = ether_mlp_forward(c, G, xg.data(), yg.data());
    if (rc) return rc;
    // thermodynamic length  s(T) = int_{t_lo}^{T} sqrt(C_v)/T' dT'  (trapezoid), or
    // the surrogate's own s(T), made non-decreasing
    std::vector<double> cum(G, 0.0);
    if (length) {
        double run = (double)yg[1];
        for (int g = 1; g < G; ++g) {
            run = std::max(run, (double)yg[2 * g + 1]);
            cum[g] = run - (double)yg[1];
        }
    } else {
        auto f = [&](int g) { return std::exp(0.5 * logcv_scale * (double)yg[2 * g + 1]) / tg[g]; };
        for (int g = 1; g < G; ++g) cum[g] = cum[g - 1] + 0.5 * (f(g) + f(g - 1)) * (tg[g] - tg[g - 1]);
    }
    const double total = cum[G - 1];
    if (!(total > 0.0) || !std::isfinite(total))
        return ctx_fail(c, ETHER_ERR_STATE, "ether_ladder_place: surrogate gives no usable C_v");
    // equal steps of s, hottest first
    int g = 0;
    for (int k = 0; k < n_temps; ++k) {
        const double target = total * (double)(n_temps - 1 - k) / (n_temps - 1);
        if (k == 0) { T_out[k] = t_hi; continue; }
        if (k == n_temps - 1) { T_out[k] = t_lo; continue; }
        g = 0;
        while (g + 1 < G && cum[g + 1] < target) ++g;
        const double w = (target - cum[g]) / std::max(cum[g + 1] - cum[g], 1e-300);
        T_out[k] = tg[g] + w * (tg[g + 1] - tg[g]);
    }
    return ETHER_OK;
}
