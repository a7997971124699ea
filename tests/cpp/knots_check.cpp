// The non-uniform knot tables of the library (tdvmc_amd/csrc/knots.cpp)
// looked up the way the kernels do: cell c = min((int)(r * cinv), ncell - 1),
// interval j or j + 1 by the cell's knot b, then the four basis pieces of
// that interval at s = r - g_j.  Input (stdin): P, the P knots, rmax, n, then
// n distances.  Output per distance: "j b0 b1 b2 b3 b0' b1' b2' b3'" (values
// and d/dr of the four pieces; interval P_s - 1 is the pseudo-interval
// beyond rc).  Built by __graft_entry__.build(); compared
// with the oracle's Cox-de Boor basis by tests/test_knots_cpu.py.
// With the single argument "bc": the boundary-condition factors of
// tdvmc::bc_factors instead, one line per family and end:
// "origin|cutoff k rows f00 f01 f02 f10 ..." (tests/test_bc_families_cpu.py).
#include "../../tdvmc_amd/csrc/knots.hpp"

#include <cstdio>
#include <iostream>
#include <string>

int main(int argc, char** argv)
{
    if (argc == 2 && std::string(argv[1]) == "bc") {
        for (int side = 0; side < 2; side++)
            for (int k = 1; k <= 4; k++) {
                double f[3][3];
                const int rows = tdvmc::bc_factors(k, side == 1, f);
                std::printf("%s %d %d", side ? "cutoff" : "origin", k, rows);
                for (int i = 0; i < rows; i++)
                    for (int m = 0; m < 3; m++) std::printf(" %.17g", f[i][m]);
                std::printf("\n");
            }
        return 0;
    }
    int P = 0, n = 0;
    std::cin >> P;
    std::vector<double> g(P);
    for (auto& v : g) std::cin >> v;
    double rmax = 0.0;
    std::cin >> rmax >> n;
    tdvmc::NuKnots K;
    std::string err;
    if (!tdvmc::nu_build(g, rmax, &K, &err)) {
        std::printf("error %s\n", err.c_str());
        return 0;
    }
    std::printf("ncell %d\n", K.ncell);
    for (int q = 0; q < n; q++) {
        double r;
        std::cin >> r;
        int c = (int)(r * K.cinv);
        if (c > K.ncell - 1) c = K.ncell - 1;
        const double* e = &K.cells[(size_t)c * 4];
        const bool right = r > e[0];
        const int j = (int)e[1] + (right ? 1 : 0);
        const double s = r - (right ? e[3] : e[2]);
        std::printf("%d", j);
        for (int m = 0; m < 4; m++) {
            const double* b = &K.beta[((size_t)j * 4 + m) * 4];
            std::printf(" %.17g", b[0] + s * (b[1] + s * (b[2] + s * b[3])));
        }
        for (int m = 0; m < 4; m++) {
            const double* b = &K.beta[((size_t)j * 4 + m) * 4];
            std::printf(" %.17g", b[1] + s * (2.0 * b[2] + s * 3.0 * b[3]));
        }
        std::printf("\n");
    }
    return 0;
}
