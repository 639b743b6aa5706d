// lattice_internal.h - the neighbour templates and the periodic colouring
// shared by the host builder (lattice.cpp) and the device builder
// (lattice_gpu.hip).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace ether {

// One neighbour of basis site la in one shell: the site of basis lb in the
// cell at offset d (in cells, before the periodic wrap).
struct TemplateEntry {
    int d[3];
    int lb;
};

// Per (shell, basis site) the offsets (d, lb) whose displacement
// d.A + b_lb - b_la has a length within tol of the shell distance
// (src/search_nbd.f90:34-58), over every offset the reference's images
// {0, +1, -1} x L (closed axes) or {0} (open axes) can produce.
struct LatticeTemplates {
    int n_shells = 0, n_basis = 0;
    std::vector<std::vector<TemplateEntry>> t;   // [shell * n_basis + la]
};

int build_templates(const double* abc, int32_t n_basis, const double* basis_xyz, const int32_t* sc,
                    const int32_t* open_bc, int32_t n_shells, const double* shell_d, double tol,
                    LatticeTemplates& out, std::string& err);

// A proper colouring of the neighbour graph that repeats with the lattice:
// colour(site) = qcol[la + n_basis * (c0 % m0 + m0 * (c1 % m1 + m1 * (c2 % m2)))]
// with m_k dividing L_k on closed axes.  Tries periods {1, 2, 3, 4, 6} per
// axis and keeps the colouring with the fewest colours (then the shortest
// period).  Returns false when no period gives a proper colouring.
struct PeriodicColouring {
    int m[3] = {1, 1, 1};
    int n_colours = 0;
    std::vector<int> qcol;
};
bool periodic_colouring(const LatticeTemplates& tp, const int32_t* sc, const int32_t* open_bc,
                        PeriodicColouring& out);

}  // namespace ether
