// Drives the C++ layer (include/tdvmc.hpp) the way the reference's driver
// drives IPhysicalSystem: per-configuration evaluations, a trial-move
// quotient, and one batched measurement filling SimulationStepData.
//
//   adapter_check <config.json> <input.txt>
// input.txt: "W N D P", then uR (P values), uI (P values), phiR phiI, then
// W*N*D coordinates, then "p x y z" (particle index and its new position for
// the quotient on walker 0).  Prints one "key index value" line per result.
#include "tdvmc.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: adapter_check <config.json> <input.txt>\n");
        return 2;
    }
    try {
        std::ifstream in(argv[2]);
        int W, N, D, P;
        in >> W >> N >> D >> P;
        std::vector<double> uR(P), uI(P);
        double phiR, phiI;
        for (auto& v : uR) in >> v;
        for (auto& v : uI) in >> v;
        in >> phiR >> phiI;
        std::vector<double> flat((size_t)W * N * D);
        for (auto& v : flat) in >> v;
        int p;
        std::vector<double> newpos(D);
        in >> p;
        for (auto& v : newpos) in >> v;
        if (!in) {
            std::fprintf(stderr, "bad input file\n");
            return 2;
        }

        tdvmc::Context ctx(argv[1]);
        if (ctx.N() != N || ctx.DIM() != D || ctx.P() != P) {
            std::fprintf(stderr, "input shape does not match the config\n");
            return 2;
        }
        tdvmc::PhysicalSystem sys(ctx);
        std::printf("P 0 %d\n", P);
        for (int w = 0; w < W; w++) {
            std::vector<std::vector<double>> R(N, std::vector<double>(D));
            for (int n = 0; n < N; n++)
                for (int a = 0; a < D; a++) R[n][a] = flat[((size_t)w * N + n) * D + a];
            sys.CalculateExpectationValues(R, uR, uI, phiR, phiI);
            std::printf("expo %d %.17g\nER %d %.17g\nEI %d %.17g\n", w, sys.GetExponent(), w,
                        sys.GetLocalEnergyR(), w, sys.GetLocalEnergyI());
            const auto O = sys.GetLocalOperators();
            for (int k = 0; k < P; k++) std::printf("O %d %.17g\n", w * P + k, O[k]);
            const auto M = sys.GetLocalOperatorsMatrix();
            std::printf("OO %d %.17g\n", w, M[P - 1][0]);
        }
        // trial move of particle p of walker 0
        std::vector<std::vector<double>> R(N, std::vector<double>(D));
        for (int n = 0; n < N; n++)
            for (int a = 0; a < D; a++) R[n][a] = flat[(size_t)n * D + a];
        std::vector<double> old = R[p];
        R[p] = newpos;
        const double q = sys.CalculateWFQuotient(R, uR, uI, phiR, phiI, p, old);
        std::printf("quotient 0 %.17g\n", q);
        sys.AcceptMove();
        std::printf("exponent_after_accept 0 %.17g\n", sys.GetExponent());

        // batched measurement of all walkers -> SimulationStepData
        ctx.set_params(uR, uI, phiR, phiI);
        ctx.set_walkers(flat);
        tdvmc::SimulationStepData s = ctx.measure(TDVMC_FP64);
        std::printf("mean_ER 0 %.17g\nmean_EI 0 %.17g\nn_samples 0 %lld\n", s.localEnergyR, s.localEnergyI,
                    (long long)s.n_samples);
        for (int k = 0; k < P; k++) std::printf("mean_O %d %.17g\n", k, s.localOperators[k]);
        for (int k = 0; k < P; k++) std::printf("S_diag %d %.17g\n", k, s.S[k][k]);
        // errors surface as exceptions with the library's message
        try {
            std::vector<double> bad(P + 1);
            ctx.set_params(bad, bad, 0.0, 0.0);
            std::printf("error_caught 0 0\n");
        } catch (const tdvmc::Error& e) {
            std::printf("error_caught 0 %d\n", e.code());
        }
    } catch (const tdvmc::Error& e) {
        std::fprintf(stderr, "tdvmc error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
