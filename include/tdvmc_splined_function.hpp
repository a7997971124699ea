// tdvmc::SplinedFunction -- the pair spline of a context in the shape of the
// reference's data holder WFParts::SplinedFunction (src/WFParts/SplinedFunction.h:10-40):
// the same member names and index order, so code written against that holder
// reads the same fields here.
//
//   nodes            the knot vector with three mirrored knots at each end
//                    (uniform: i rc / (P - 1), i = -3 .. P + 2, BosonsBulk.cpp:60-66;
//                    USE_NURBS: NURBS_GRID extended as BosonsBulk::SetNodes, :35-44)
//   splineWeights    [s][p][c]: piece p of cubic B-spline s (on knot interval
//                    s + p) as sum_c w r^c, the layout of SplineFactory::GetWeights3
//                    (SplineFactory.cpp:54-101); computed here from the Cox-de Boor
//                    recursion on the knots, not from that routine's closed forms
//   bcFactorsStart / bcFactorsEnd, np1 .. np3, numberOf*Parameters
//                    the boundary conditions at 0 and rc: BosonsBulk's zero-slope
//                    pair (SetBoundaryConditions3_1D_OR_2 / _CO_2 with uniform =
//                    true, BosonsBulk.cpp:75-82), or in of() the context's
//                    families (tdvmc_spline_bc)
//   splineSums       [s]: basis sums B_s = sum over pairs r_ij <= rc of B_s(r_ij)
//                    for one configuration (fill_sums: tdvmc_basis_sums on the GPU)
//
// The per-particle derivative sums splineSumsD / splineSumsD2 and the trial-move
// members splineSumsNew / sumOldPerBin / sumNewPerBin stay empty: the library
// keeps per-walker gradients and move deltas on the device (tdvmc_eval_configs
// returns the contracted gradient and Laplacian).
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "tdvmc.h"

namespace tdvmc {

class SplinedFunction {
public:
    int splineOrder = 3;
    int numOfSplineParts = 4;

    std::vector<double> nodes;
    double nodeSpacing = 0.0;   // uniform knots: rc / (P - 1); non-uniform: 0

    int numberOfSplines = 0;
    std::vector<double> splineSums;
    std::vector<std::vector<std::vector<double>>> splineSumsD;
    std::vector<std::vector<double>> splineSumsD2;

    std::vector<double> splineSumsNew;
    std::vector<double> sumOldPerBin;
    std::vector<double> sumNewPerBin;

    int numberOfSpecialParametersStart = 0;
    int numberOfSpecialParametersEnd = 0;
    int numberOfStandardParameters = 0;
    int np1 = 0, np2 = 0, np3 = 0;
    std::vector<std::vector<double>> bcFactorsStart;
    std::vector<std::vector<double>> bcFactorsEnd;

    std::vector<std::vector<std::vector<double>>> splineWeights;

    // The pair spline of P parameters on [0, rc]: uniform knots (grid empty) or
    // the P knots of NURBS_GRID (grid.back() = rc).
    SplinedFunction(int P, double rc, const std::vector<double>& grid = {})
    {
        if (P < 4) throw std::invalid_argument("SplinedFunction: need at least 4 parameters");
        if (!grid.empty() && (int)grid.size() != P)
            throw std::invalid_argument("SplinedFunction: the grid must hold P knots");
        if (grid.empty()) {
            for (int i = -3; i < P + 3; i++) nodes.push_back(i * rc / (P - 1.0));
            nodeSpacing = rc / (P - 1.0);
        } else {
            nodes = grid;
            for (int i = 0; i < 3; i++) {
                nodes.insert(nodes.begin(), -grid[i + 1]);
                nodes.push_back(2.0 * grid[P - 1] - grid[P - 2 - i]);
            }
        }
        numberOfSplines = (int)nodes.size() - 4;
        splineWeights.assign(numberOfSplines, std::vector<std::vector<double>>(4, std::vector<double>(4, 0.0)));
        for (int s = 0; s < numberOfSplines; s++)
            for (int p = 0; p < 4; p++) piece(s, p, splineWeights[s][p]);
        // BosonsBulk's zero-slope pair (of() replaces it by the context's rows)
        bcFactorsStart = {{0.0, 1.0, 0.0}, {1.0, 0.0, 1.0}};
        bcFactorsEnd = {{1.0, 0.0, 1.0}, {0.0, 1.0, 0.0}};
        set_counts(P);
        splineSums.assign(numberOfSplines, 0.0);
    }

    // numberOf*Parameters and np1 .. np3 from the factor rows (BosonsBulk.cpp:78-83)
    void set_counts(int P)
    {
        numberOfSpecialParametersStart = (int)bcFactorsStart.size();
        numberOfSpecialParametersEnd = (int)bcFactorsEnd.size();
        numberOfStandardParameters = P - numberOfSpecialParametersStart - numberOfSpecialParametersEnd;
        np1 = numberOfSpecialParametersStart;
        np2 = np1 + numberOfStandardParameters;
        np3 = np2 + numberOfSpecialParametersEnd;
    }

    // The context's pair spline (its config's N_PARAM, rc and NURBS_GRID, and
    // the boundary-condition rows of SPLINE_BC_ORIGIN / SPLINE_BC_CUTOFF).
    static SplinedFunction of(tdvmc_ctx* ctx, const std::vector<double>& grid = {})
    {
        tdvmc_info info{};
        if (tdvmc_get_info(ctx, &info) != TDVMC_OK) throw std::runtime_error(tdvmc_last_error(ctx));
        if (info.use_nurbs && grid.empty())
            throw std::invalid_argument("SplinedFunction::of: pass the NURBS_GRID of a USE_NURBS config");
        SplinedFunction f(info.P_pair, info.rc, info.use_nurbs ? grid : std::vector<double>{});
        double a[9], b[9];
        int na = 0, nb = 0;
        if (tdvmc_spline_bc(ctx, a, &na, b, &nb) != TDVMC_OK) throw std::runtime_error(tdvmc_last_error(ctx));
        f.bcFactorsStart.assign(na, std::vector<double>(3));
        f.bcFactorsEnd.assign(nb, std::vector<double>(3));
        for (int i = 0; i < na; i++)
            for (int m = 0; m < 3; m++) f.bcFactorsStart[i][m] = a[i * 3 + m];
        for (int i = 0; i < nb; i++)
            for (int m = 0; m < 3; m++) f.bcFactorsEnd[i][m] = b[i * 3 + m];
        f.set_counts(info.P_pair);
        return f;
    }

    // splineSums of configuration R [N][DIM] (FP64 evaluation on the device).
    void fill_sums(tdvmc_ctx* ctx, const double* R)
    {
        if (tdvmc_basis_sums(ctx, R, 1, TDVMC_FP64, splineSums.data()) != TDVMC_OK)
            throw std::runtime_error(tdvmc_last_error(ctx));
    }

    // sum_p ... of spline s at r (0 outside its support), from splineWeights
    double value(int s, double r) const
    {
        for (int p = 0; p < 4; p++) {
            const double a = nodes[s + p], b = nodes[s + p + 1];
            if (r >= a && r <= b && b > a) {
                const auto& w = splineWeights[s][p];
                return w[0] + r * (w[1] + r * (w[2] + r * w[3]));
            }
        }
        return 0.0;
    }

private:
    // Cox-de Boor on the knots of spline s alone: B_{s,3} restricted to the
    // interval nodes[s + p] .. nodes[s + p + 1], as a cubic in r.
    void piece(int s, int p, std::vector<double>& w) const
    {
        // P[k][m]: order-k spline s + m (m = 0 .. 3 - k) on that interval
        double P[4][4][4] = {};
        for (int m = 0; m < 4; m++)
            if (m == p) P[0][m][0] = 1.0;
        for (int k = 1; k <= 3; k++) {
            for (int m = 0; m + k <= 3; m++) {
                const int i = s + m;
                double* out = P[k][m];
                const double d1 = nodes[i + k] - nodes[i], d2 = nodes[i + k + 1] - nodes[i + 1];
                for (int c = 0; c < 4; c++) out[c] = 0.0;
                if (d1 > 0.0) {
                    // (r - n_i) / d1 * P[k-1][m]
                    for (int c = 0; c < 4; c++) {
                        const double v = P[k - 1][m][c] / d1;
                        out[c] -= nodes[i] * v;
                        if (c + 1 < 4) out[c + 1] += v;
                    }
                }
                if (d2 > 0.0) {
                    // (n_{i+k+1} - r) / d2 * P[k-1][m+1]
                    for (int c = 0; c < 4; c++) {
                        const double v = P[k - 1][m + 1][c] / d2;
                        out[c] += nodes[i + k + 1] * v;
                        if (c + 1 < 4) out[c + 1] -= v;
                    }
                }
            }
        }
        for (int c = 0; c < 4; c++) w[c] = P[3][0][c];
    }
};

}  // namespace tdvmc
