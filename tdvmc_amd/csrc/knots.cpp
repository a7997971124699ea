// Non-uniform pair-spline knots: basis pieces and the interval lookup cells.
//
// The knot vector follows BosonsBulk::SetNodes (src/PhysicalSystems/BosonsBulk.cpp:35-44):
// the P_s grid values, extended by three mirrored knots at each end,
//   n = (-g_3, -g_2, -g_1, g_0, ..., g_{P-1}, 2g_{P-1}-g_{P-2}, 2g_{P-1}-g_{P-3}, 2g_{P-1}-g_{P-4}),
// and cubic B-spline i lives on n_i .. n_{i+4}.  A distance r lies in grid
// interval j when g_j < r <= g_{j+1} (the lower_bound - 1 of
// BosonsBulk.cpp:197-198), clamped to 0 .. P_s - 2, and splines j .. j+3 are
// the non-zero ones there.  Their pieces on interval j come from the
// Cox-de Boor recursion (C. de Boor, "A Practical Guide to Splines", ch. IX)
// carried out on cubic polynomials in s = r - g_j.

#include "knots.hpp"

#include <algorithm>
#include <cmath>

namespace tdvmc {

namespace {

double ext_knot(const std::vector<double>& g, int i)
{
    const int P = (int)g.size();
    if (i < 3) return -g[3 - i];
    if (i < P + 3) return g[i - 3];
    return 2.0 * g[P - 1] - g[P - 1 - (i - (P + 2))];
}

struct Poly {
    double c[4] = {0.0, 0.0, 0.0, 0.0};
};

// (a0 + a1 s) * p, degree kept below 4 by construction
Poly mul_lin(double a0, double a1, const Poly& p)
{
    Poly q;
    for (int k = 0; k < 4; k++) {
        q.c[k] += a0 * p.c[k];
        if (k + 1 < 4) q.c[k + 1] += a1 * p.c[k];
    }
    return q;
}

Poly add(const Poly& a, const Poly& b)
{
    Poly q;
    for (int k = 0; k < 4; k++) q.c[k] = a.c[k] + b.c[k];
    return q;
}

double eval(const Poly& p, double s) { return p.c[0] + s * (p.c[1] + s * (p.c[2] + s * p.c[3])); }

// pieces of splines j .. j+3 on grid interval j, as cubics in s = r - g_j
void pieces(const std::vector<double>& g, int j, Poly out[4])
{
    const int e = j + 3;   // extended interval n_e .. n_{e+1}
    const double ne = ext_knot(g, e);
    // B[k][m]: order-k spline e - k + m (m = 0 .. k)
    Poly B[4][4];
    B[0][0].c[0] = 1.0;
    for (int k = 1; k <= 3; k++) {
        for (int m = 0; m <= k; m++) {
            const int i = e - k + m;
            Poly v;
            if (m > 0) {
                // (r - n_i) / (n_{i+k} - n_i) * B_{i,k-1}
                const double ni = ext_knot(g, i), nik = ext_knot(g, i + k);
                const double inv = 1.0 / (nik - ni);
                v = add(v, mul_lin((ne - ni) * inv, inv, B[k - 1][m - 1]));
            }
            if (m < k) {
                // (n_{i+k+1} - r) / (n_{i+k+1} - n_{i+1}) * B_{i+1,k-1}
                const double ni1 = ext_knot(g, i + 1), nik1 = ext_knot(g, i + k + 1);
                const double inv = 1.0 / (nik1 - ni1);
                v = add(v, mul_lin((nik1 - ne) * inv, -inv, B[k - 1][m]));
            }
            B[k][m] = v;
        }
    }
    for (int m = 0; m < 4; m++) out[m] = B[3][m];
}

}  // namespace

bool nu_build(const std::vector<double>& grid, double rmax, NuKnots* out, std::string* err)
{
    const int P = (int)grid.size();
    if (P < 4) {
        *err = "NURBS_GRID needs at least 4 knots";
        return false;
    }
    if (grid[0] != 0.0) {
        *err = "NURBS_GRID must start at 0 (the mirrored knots assume it)";
        return false;
    }
    double dmin = INFINITY;
    for (int k = 0; k + 1 < P; k++) {
        const double d = grid[k + 1] - grid[k];
        if (!(d > 0.0)) {
            *err = "NURBS_GRID must be strictly increasing";
            return false;
        }
        dmin = std::min(dmin, d);
    }
    const double rc = grid[P - 1];
    if (!(rc <= rmax * (1.0 + 1e-12))) {
        *err = "NURBS_GRID ends beyond half the box";
        return false;
    }
    NuKnots& K = *out;
    K.grid = grid;
    // basis pieces per interval, then the values at rc for the pseudo-interval
    K.beta.assign((size_t)P * 16, 0.0);
    for (int j = 0; j + 1 < P; j++) {
        Poly pc[4];
        pieces(grid, j, pc);
        for (int m = 0; m < 4; m++)
            for (int p = 0; p < 4; p++) K.beta[((size_t)j * 4 + m) * 4 + p] = pc[m].c[p];
    }
    {
        Poly pc[4];
        pieces(grid, P - 2, pc);
        const double hl = grid[P - 1] - grid[P - 2];
        for (int m = 0; m < 3; m++) K.beta[((size_t)(P - 1) * 4 + m) * 4] = eval(pc[m + 1], hl);
    }
    // lookup cells: width below the smallest spacing, so at most one knot per cell
    const int nin = (int)std::ceil(rc / (0.999 * dmin));
    if (nin + 1 > kMaxNuCells) {
        *err = "NURBS_GRID spacing too fine: more than " + std::to_string(kMaxNuCells - 1) +
               " lookup cells of the smallest spacing fit into [0, rc]";
        return false;
    }
    K.ncell = nin + 1;
    K.cinv = nin / rc;
    const double w = rc / nin;
    K.cells.assign((size_t)K.ncell * 4, 0.0);
    for (int c = 0; c < nin; c++) {
        const double x0 = c * w, x1 = (c + 1) * w;
        int j = (int)(std::lower_bound(grid.begin(), grid.end(), x0) - grid.begin()) - 1;
        j = std::max(0, std::min(P - 2, j));
        double* e = &K.cells[(size_t)c * 4];
        e[0] = grid[j + 1] < x1 ? grid[j + 1] : 1e30;
        e[1] = j;
        e[2] = grid[j];
        e[3] = grid[j + 1];
    }
    double* e = &K.cells[(size_t)nin * 4];   // r >= rc: interval P-2 up to rc, beyond it P-1
    e[0] = rc;
    e[1] = P - 2;
    e[2] = grid[P - 2];
    e[3] = rc;
    return true;
}

int bc_factors(int family, bool at_cutoff, double f[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int m = 0; m < 3; m++) f[i][m] = 0.0;
    // each family's rows as (row, spline, weight) entries
    struct E { int i, m; double v; };
    static const E or_1[] = {{0, 0, 1}, {1, 1, 1}, {2, 2, 1}};
    static const E or_2[] = {{0, 1, 1}, {1, 0, 1}, {1, 2, 1}};
    static const E or_3[] = {{0, 0, -1}, {0, 2, 1}, {1, 0, 2}, {1, 1, 1}};
    static const E co_2[] = {{0, 0, 1}, {0, 2, 1}, {1, 1, 1}};
    static const E co_3[] = {{0, 0, 1}, {0, 2, -1}, {1, 1, 1}, {1, 2, 2}};
    static const E all_4[] = {{0, 0, 1}, {0, 1, 1}, {0, 2, 1}};
    const E* e = nullptr;
    int ne = 0, rows = 0;
    switch (family) {
    case 1: e = or_1; ne = 3; rows = 3; break;   // the same identity at both ends
    case 2: e = at_cutoff ? co_2 : or_2; ne = 3; rows = 2; break;
    case 3: e = at_cutoff ? co_3 : or_3; ne = 4; rows = 2; break;
    case 4: e = all_4; ne = 3; rows = 1; break;
    default: return 0;
    }
    for (int k = 0; k < ne; k++) f[e[k].i][e[k].m] = e[k].v;
    return rows;
}

}  // namespace tdvmc
