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

namespace ether {
thread_local std::string g_last_error;
void set_global_error(const std::string& s) { g_last_error = s; }
const std::string& global_error() { return g_last_error; }
}  // namespace ether

namespace {

struct TemplateEntry {
    int d[3];
    int lb;
};

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

extern "C" int ether_build_neighbours(const double* abc, int32_t n_basis, const double* basis_xyz,
                                      const int32_t* sc, const int32_t* open_bc,
                                      int32_t n_shells, const double* shell_d, double tol,
                                      int32_t capacity, int32_t* nn_count, int32_t* nn_idx) {
    using ether::set_global_error;
    if (!abc || !basis_xyz || !sc || !open_bc || !shell_d || !nn_count || !nn_idx) {
        set_global_error("ether_build_neighbours: null pointer argument");
        return ETHER_ERR_ARG;
    }
    if (n_basis <= 0 || n_shells <= 0 || capacity <= 0 || sc[0] <= 0 || sc[1] <= 0 ||
        sc[2] <= 0 || !(tol >= 0.0)) {
        set_global_error("ether_build_neighbours: invalid sizes");
        return ETHER_ERR_ARG;
    }
    const long long n_ll = (long long)sc[0] * sc[1] * sc[2] * n_basis;
    if (n_ll > (1LL << 24)) {
        set_global_error("ether_build_neighbours: more than 2^24 sites");
        return ETHER_ERR_LIMIT;
    }
    const int n = (int)n_ll;
    double inv[9];
    if (!invert3(abc, inv)) {
        set_global_error("ether_build_neighbours: singular lattice vectors");
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

    std::vector<int> nbr;
    nbr.reserve(64);
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
        // template per basis site la
        std::vector<std::vector<TemplateEntry>> tmpl(n_basis);
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
                                tmpl[la].push_back({{d0, d1, d2}, lb});
                        }
        for (int ck = 0; ck < sc[2]; ++ck)
            for (int cj = 0; cj < sc[1]; ++cj)
                for (int ci = 0; ci < sc[0]; ++ci)
                    for (int la = 0; la < n_basis; ++la) {
                        const int c[3] = {ci, cj, ck};
                        const int i = ((ck * sc[1] + cj) * sc[0] + ci) * n_basis + la;
                        nbr.clear();
                        for (const TemplateEntry& t : tmpl[la]) {
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
