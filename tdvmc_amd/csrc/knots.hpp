// Host-side pair-spline tables: the boundary-condition factors of the cubic
// families, and for non-uniform knots (USE_NURBS with NURBS_GRID) the tables
// the kernels look intervals and basis pieces up in.
#pragma once

#include <string>
#include <vector>

namespace tdvmc {

// Cells of the interval lookup: [0, rc] in ncell - 1 cells of width w below
// the smallest knot spacing (so a cell holds at most one knot), and one last
// cell for every r beyond rc.  Cell c = min((int)(r * cinv), ncell - 1):
//   (b, j, g_j, g_{j+1}):  r <= b -> interval j, offset s = r - g_j;
//                          r >  b -> interval j + 1, offset s = r - g_{j+1}.
// Interval P_s - 1 is the pseudo-interval beyond rc (constant pieces).
struct NuKnots {
    std::vector<double> grid;   // [P_s] the knots g_0 = 0 < ... < g_{P_s-1} = rc
    std::vector<double> beta;   // [P_s][4][4]: piece m (spline j + m) of interval j as
                                //   sum_p beta[j][m][p] s^p, s = r - g_j; row P_s - 1:
                                //   the values at rc of splines P_s-1 .. P_s+1 (p = 0 only)
    std::vector<double> cells;  // [ncell][4]: b, j, g_j, g_{j+1}
    int ncell = 0;
    double cinv = 0.0;          // cells per unit length
};

constexpr int kMaxNuCells = 512;

// Boundary-condition factors of cubic family 1..4 at the origin
// (SetBoundaryConditions3_1D_OR_k, src/SplineFactory.cpp:377-515) or at the
// cutoff (_CO_k, :517-656), the uniform branch that BosonsBulk passes for
// uniform and NURBS knots alike (BosonsBulk.cpp:76-77): row i holds parameter
// i's weights on splines 0..2 (origin) or S-3..S-1 (cutoff), rows past the
// count are zero.  1: no condition (3 rows), 2: u' = 0 (2), 3: u'' = 0 (2),
// 4: u' = u'' = 0 (1).  Returns the row count, 0 for an unknown family.
int bc_factors(int family, bool at_cutoff, double f[3][3]);

// Validates grid (P_s values, g_0 = 0, strictly increasing, g_{P_s-1} <=
// rmax) and fills the tables.  Returns false with a message on bad input.
bool nu_build(const std::vector<double>& grid, double rmax, NuKnots* out, std::string* err);

}  // namespace tdvmc
