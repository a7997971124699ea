// tdvmc::SplinedFunction (include/tdvmc_splined_function.hpp) built from a
// pair-spline description, printed for comparison with the reference's own
// tables (tests/golden/spline_weights.json, by tests/test_splined_function_cpu.py).
// Input (stdin): P rc G, then G knots (G = 0: uniform knots).  Output lines:
// "nodes ...", "w s p c0 c1 c2 c3" per piece, "bcs ...", "bce ...",
// "np np1 np2 np3 nstd nsplines", "value s r B_s(r)" for a few sample points.
#include "tdvmc_splined_function.hpp"

#include <cstdio>
#include <iostream>

int main()
{
    int P = 0, G = 0;
    double rc = 0.0;
    std::cin >> P >> rc >> G;
    std::vector<double> grid(G);
    for (auto& v : grid) std::cin >> v;
    tdvmc::SplinedFunction f(P, rc, grid);
    std::printf("nodes");
    for (double v : f.nodes) std::printf(" %.17g", v);
    std::printf("\n");
    for (int s = 0; s < f.numberOfSplines; s++)
        for (int p = 0; p < 4; p++) {
            const auto& w = f.splineWeights[s][p];
            std::printf("w %d %d %.17g %.17g %.17g %.17g\n", s, p, w[0], w[1], w[2], w[3]);
        }
    std::printf("bcs");
    for (const auto& r : f.bcFactorsStart)
        for (double v : r) std::printf(" %.17g", v);
    std::printf("\nbce");
    for (const auto& r : f.bcFactorsEnd)
        for (double v : r) std::printf(" %.17g", v);
    std::printf("\nnp %d %d %d %d %d\n", f.np1, f.np2, f.np3, f.numberOfStandardParameters, f.numberOfSplines);
    return 0;
}
