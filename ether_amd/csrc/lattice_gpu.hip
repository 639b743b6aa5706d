// lattice_gpu.hip - the neighbour lists and the periodic colouring built on
// the GPU: the same capture rule as ether_build_neighbours
// (src/getting_nbd.f90:46-82, src/search_nbd.f90:34-58), one thread per
// (site, shell) applying the shell's template for the site's basis atom.
//
// The templates (offsets (d, lb) whose displacement d.A + b_lb - b_la lies
// within tol of the shell distance) and the periodic colouring come from the
// host code shared with the host builder (lattice_internal.h); they are small
// (per basis atom and shell).  Per site the thread wraps each offset into the
// supercell, keeps it when the wrap is an image the reference tries ({0, +1,
// -1} closed, {0} open) and the site is not i itself, sorts the indices and
// drops repeats (a site is stored once per shell, ascending), as the host
// builder does.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ether_gpu.h"
#include "lattice_internal.h"

namespace ether {
void set_global_error(const std::string& s);
}

namespace {

constexpr int kMaxTemplate = 256;   // offsets per (shell, basis atom) one thread can hold

struct BuildArgs {
    const int4* ent;      // template entries (d0, d1, d2, lb)
    const int* ent_beg;   // [n_shells * nb + 1] start of each (shell, la) list
    int sc[3];
    int open_bc[3];
    int nb, n, n_shells, capacity;
    int* nn_count;        // [n_shells * n]
    int* nn_idx;          // [n_shells * n * capacity]
    const int* qcol;      // periodic colouring (nullable)
    int m[3];
    int* colour;          // [n] (nullable)
    int* overflow;        // smallest (shell * n + site) whose list exceeds capacity
};

__device__ __forceinline__ int floordiv_d(int a, int b) {
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

__global__ __launch_bounds__(256) void k_build_neighbours(BuildArgs a) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)a.n * a.n_shells) return;
    const int sh = (int)(t / a.n), i = (int)(t % a.n);
    const int la = i % a.nb, cell = i / a.nb;
    const int c[3] = {cell % a.sc[0], (cell / a.sc[0]) % a.sc[1], cell / (a.sc[0] * a.sc[1])};
    if (sh == 0 && a.colour && a.qcol)
        a.colour[i] = a.qcol[la + a.nb * (c[0] % a.m[0] + a.m[0] * (c[1] % a.m[1] + a.m[1] * (c[2] % a.m[2])))];
    int buf[kMaxTemplate];
    int cnt = 0;
    const int b = a.ent_beg[sh * a.nb + la], e = a.ent_beg[sh * a.nb + la + 1];
    for (int q = b; q < e; ++q) {
        const int4 en = a.ent[q];
        const int dd[3] = {en.x, en.y, en.z};
        int cc[3];
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int u = c[k] + dd[k];
            const int s = floordiv_d(u, a.sc[k]);
            if (a.open_bc[k] ? (s != 0) : (s < -1 || s > 1)) ok = false;
            cc[k] = u - s * a.sc[k];
        }
        if (!ok) continue;
        const int j = ((cc[2] * a.sc[1] + cc[1]) * a.sc[0] + cc[0]) * a.nb + en.w;
        if (j == i) continue;
        // insertion into the sorted list, repeats dropped
        int p = cnt;
        while (p > 0 && buf[p - 1] > j) --p;
        if (p > 0 && buf[p - 1] == j) continue;
        for (int r = cnt; r > p; --r) buf[r] = buf[r - 1];
        buf[p] = j;
        ++cnt;
    }
    if (cnt > a.capacity) {
        atomicMin(a.overflow, (int)t);
        return;
    }
    a.nn_count[t] = cnt;
    int* out = a.nn_idx + t * a.capacity;
    for (int r = 0; r < a.capacity; ++r) out[r] = r < cnt ? buf[r] : -1;
}

}  // namespace

#define HG(e)                                                                              \
    do {                                                                                   \
        hipError_t _e = (e);                                                               \
        if (_e != hipSuccess) {                                                            \
            ether::set_global_error(std::string("ether_build_neighbours_device: ") + #e + \
                                    ": " + hipGetErrorString(_e));                         \
            rc = ETHER_ERR_HIP;                                                            \
            goto done;                                                                     \
        }                                                                                  \
    } while (0)

extern "C" int ether_build_neighbours_device(const double* abc, int32_t n_basis, const double* basis_xyz,
                                             const int32_t* sc, const int32_t* open_bc, int32_t n_shells,
                                             const double* shell_d, double tol, int32_t capacity,
                                             int32_t* nn_count, int32_t* nn_idx, int32_t* colour,
                                             int32_t* n_colours) {
    using ether::set_global_error;
    if (!nn_count || !nn_idx || capacity <= 0) {
        set_global_error("ether_build_neighbours_device: null output or capacity <= 0");
        return ETHER_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_global_error("ether_build_neighbours_device: no HIP device visible");
        return ETHER_ERR_NODEV;
    }
    ether::LatticeTemplates tp;
    std::string err;
    int rc = ether::build_templates(abc, n_basis, basis_xyz, sc, open_bc, n_shells, shell_d, tol, tp, err);
    if (rc != ETHER_OK) {
        set_global_error("ether_build_neighbours_device: " + err);
        return rc;
    }
    const int n = sc[0] * sc[1] * sc[2] * n_basis;
    std::vector<int4> ent;
    std::vector<int> beg(1, 0);
    for (const auto& lst : tp.t) {
        if ((int)lst.size() > kMaxTemplate) {
            set_global_error("ether_build_neighbours_device: more than 256 candidate offsets for one shell and "
                             "basis atom (use ether_build_neighbours)");
            return ETHER_ERR_LIMIT;
        }
        for (const auto& e : lst) ent.push_back(make_int4(e.d[0], e.d[1], e.d[2], e.lb));
        beg.push_back((int)ent.size());
    }
    ether::PeriodicColouring pc;
    const bool coloured = colour && n_colours && ether::periodic_colouring(tp, sc, open_bc, pc);
    if (n_colours) *n_colours = coloured ? pc.n_colours : 0;
    const size_t nlist = (size_t)n * n_shells;
    int4* d_ent = nullptr;
    int *d_beg = nullptr, *d_cnt = nullptr, *d_idx = nullptr, *d_qcol = nullptr, *d_col = nullptr,
        *d_over = nullptr;
    int over = 0x7fffffff;
    BuildArgs a{};
    HG(hipMalloc(&d_ent, sizeof(int4) * (ent.size() + 1)));
    HG(hipMalloc(&d_beg, sizeof(int) * beg.size()));
    HG(hipMalloc(&d_cnt, sizeof(int) * nlist));
    HG(hipMalloc(&d_idx, sizeof(int) * nlist * capacity));
    HG(hipMalloc(&d_over, sizeof(int)));
    if (coloured) {
        HG(hipMalloc(&d_qcol, sizeof(int) * pc.qcol.size()));
        HG(hipMalloc(&d_col, sizeof(int) * n));
        HG(hipMemcpy(d_qcol, pc.qcol.data(), sizeof(int) * pc.qcol.size(), hipMemcpyHostToDevice));
    }
    if (!ent.empty()) HG(hipMemcpy(d_ent, ent.data(), sizeof(int4) * ent.size(), hipMemcpyHostToDevice));
    HG(hipMemcpy(d_beg, beg.data(), sizeof(int) * beg.size(), hipMemcpyHostToDevice));
    HG(hipMemcpy(d_over, &over, sizeof(int), hipMemcpyHostToDevice));
    a.ent = d_ent;
    a.ent_beg = d_beg;
    for (int k = 0; k < 3; ++k) {
        a.sc[k] = sc[k];
        a.open_bc[k] = open_bc[k] ? 1 : 0;
        a.m[k] = coloured ? pc.m[k] : 1;
    }
    a.nb = n_basis;
    a.n = n;
    a.n_shells = n_shells;
    a.capacity = capacity;
    a.nn_count = d_cnt;
    a.nn_idx = d_idx;
    a.qcol = d_qcol;
    a.colour = d_col;
    a.overflow = d_over;
    hipLaunchKernelGGL(k_build_neighbours, dim3((unsigned)((nlist + 255) / 256)), dim3(256), 0, 0, a);
    HG(hipGetLastError());
    HG(hipMemcpy(&over, d_over, sizeof(int), hipMemcpyDeviceToHost));
    if (over != 0x7fffffff) {
        // the first list over capacity in (shell, site) order; the host builder words the message
        rc = ether_build_neighbours(abc, n_basis, basis_xyz, sc, open_bc, n_shells, shell_d, tol, capacity,
                                    nn_count, nn_idx);
        if (rc == ETHER_OK) {
            set_global_error("ether_build_neighbours_device: a neighbour list exceeds nbd_capacity");
            rc = ETHER_ERR_LIMIT;
        }
        goto done;
    }
    HG(hipMemcpy(nn_count, d_cnt, sizeof(int) * nlist, hipMemcpyDeviceToHost));
    HG(hipMemcpy(nn_idx, d_idx, sizeof(int) * nlist * capacity, hipMemcpyDeviceToHost));
    if (coloured) HG(hipMemcpy(colour, d_col, sizeof(int) * n, hipMemcpyDeviceToHost));
done:
    (void)hipFree(d_ent);
    (void)hipFree(d_beg);
    (void)hipFree(d_cnt);
    (void)hipFree(d_idx);
    (void)hipFree(d_over);
    (void)hipFree(d_qcol);
    (void)hipFree(d_col);
    return rc;
}
