// Drives two plug-ins through the reference's own interface pointer type,
// PhysicalSystems::IPhysicalSystem* (src/PhysicalSystems/IPhysicalSystem.h):
// the reference's BosonsBulk (compiled from its sources into
// oracle/_ref/libref.so) and this library's GpuBulkBosons
// (include/tdvmc_iphysicalsystem.hpp), with the calls the driver makes and in
// its order: InitSystem, CalculateWavefunction, CalculateWFQuotient /
// AcceptMove on a short Metropolis chain driven by the reference plug-in's
// decisions, CalculateExpectationValues, CalculateAdditionalSystemProperties,
// the correlated-sampling overloads, GetCenterOfMass.
//
//   iphys_check <config.json>
// The config describes a square-well BosonsBulk system (N, DIM, LBOX,
// N_PARAM, SYSTEM_PARAMS, GR_BIN_COUNT, PARAMS_REAL, PARAMS_IMAGINARY).
// Prints "key value" lines: the largest relative differences between the two
// plug-ins and counts; the caller checks them.
#include "tdvmc_iphysicalsystem.hpp"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>

extern "C" void* ref_new_bosonsbulk(int n, int dim, double L, int nparam, const double* sysparams, int nsys,
                                    int gr_bins);

namespace {

double num(const std::string& text, const std::string& key)
{
    const size_t k = text.find("\"" + key + "\"");
    const size_t c = text.find(':', k);
    return std::stod(text.substr(c + 1));
}
std::vector<double> arr(const std::string& text, const std::string& key)
{
    const size_t k = text.find("\"" + key + "\"");
    const size_t a = text.find('[', k), b = text.find(']', a);
    std::vector<double> v;
    std::stringstream ss(text.substr(a + 1, b - a - 1));
    std::string tok;
    while (std::getline(ss, tok, ',')) v.push_back(std::stod(tok));
    return v;
}
double rel(double a, double b, double scale) { return std::fabs(a - b) / std::max(scale, 1e-300); }
double vrel(const std::vector<double>& a, const std::vector<double>& b)
{
    double m = 0.0, s = 0.0;
    for (size_t i = 0; i < b.size(); i++) {
        m = std::max(m, std::fabs(a[i] - b[i]));
        s = std::max(s, std::fabs(b[i]));
    }
    return a.size() == b.size() ? m / std::max(s, 1e-300) : 1e300;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: iphys_check <config.json>\n");
        return 2;
    }
    std::ifstream f(argv[1]);
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string cfg = buf.str();
    const int N = (int)num(cfg, "N"), D = (int)num(cfg, "DIM"), P = (int)num(cfg, "N_PARAM");
    const int bins = (int)num(cfg, "GR_BIN_COUNT");
    const double L = num(cfg, "LBOX");
    std::vector<double> sp = arr(cfg, "SYSTEM_PARAMS"), uR = arr(cfg, "PARAMS_REAL"), uI = arr(cfg, "PARAMS_IMAGINARY");
    const double phiR = 0.2, phiI = -0.1;
    try {
        using PhysicalSystems::IPhysicalSystem;
        IPhysicalSystem* ref = static_cast<IPhysicalSystem*>(ref_new_bosonsbulk(N, D, L, P, sp.data(),
                                                                                 (int)sp.size(), bins));
        IPhysicalSystem* gpu = new PhysicalSystems::GpuBulkBosons(sp, "", argv[1]);
        gpu->InitSystem();

        std::mt19937_64 gen(7);
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        std::normal_distribution<double> nrm(0.0, 1.0);
        std::vector<std::vector<double> > R(N, std::vector<double>(D));
        for (auto& r : R)
            for (auto& x : r) x = (uni(gen) - 0.5) * L;

        double e_expo = 0, e_wf = 0, e_q = 0, e_ER = 0, e_EI = 0, e_O = 0, e_OO = 0, e_OE = 0, e_oth = 0;
        ref->CalculateWavefunction(R, uR, uI, phiR, phiI);
        gpu->CalculateWavefunction(R, uR, uI, phiR, phiI);
        e_expo = std::max(e_expo, rel(gpu->GetExponent(), ref->GetExponent(), std::fabs(ref->GetExponent()) + 1));
        e_wf = std::max(e_wf, rel(std::log(gpu->GetWf()), std::log(ref->GetWf()), std::fabs(ref->GetExponent()) + 1));
        e_O = std::max(e_O, vrel(gpu->GetLocalOperators(), ref->GetLocalOperators()));

        // a chain driven by the reference plug-in's accept decisions
        int accepted = 0, moves = 4 * N;
        for (int t = 0; t < moves; t++) {
            const int p = (int)(uni(gen) * N) % N;
            std::vector<double> old = R[p];
            for (int a = 0; a < D; a++) R[p][a] += 0.35 * nrm(gen);
            const double qr = ref->CalculateWFQuotient(R, uR, uI, phiR, phiI, p, old);
            const double qg = gpu->CalculateWFQuotient(R, uR, uI, phiR, phiI, p, old);
            e_q = std::max(e_q, rel(qg, qr, qr));
            e_expo = std::max(e_expo, rel(gpu->GetExponentNew(), ref->GetExponentNew(),
                                          std::fabs(ref->GetExponentNew()) + 1));
            if (qr >= uni(gen)) {
                ref->AcceptMove();
                gpu->AcceptMove();
                accepted++;
                e_expo = std::max(e_expo, rel(gpu->GetExponent(), ref->GetExponent(),
                                              std::fabs(ref->GetExponent()) + 1));
            } else {
                R[p] = old;
            }
        }
        ref->CalculateExpectationValues(R, uR, uI, phiR, phiI);
        gpu->CalculateExpectationValues(R, uR, uI, phiR, phiI);
        // E_L relative to the size of its cancelling parts (kinetic sums and V)
        std::vector<double> ro = ref->GetOtherExpectationValues();
        double esc = std::fabs(ro[1]);
        for (int k = 4; k < 9; k++) esc += std::fabs(ro[k]);
        e_ER = rel(gpu->GetLocalEnergyR(), ref->GetLocalEnergyR(), esc);
        e_EI = rel(gpu->GetLocalEnergyI(), ref->GetLocalEnergyI(), esc);
        e_O = std::max(e_O, vrel(gpu->GetLocalOperators(), ref->GetLocalOperators()));
        auto Mg = gpu->GetLocalOperatorsMatrix(), Mr = ref->GetLocalOperatorsMatrix();
        for (int i = 0; i < P; i++) e_OO = std::max(e_OO, vrel(Mg[i], Mr[i]));
        e_OE = std::max(vrel(gpu->GetLocalOperatorlocalEnergyR(), ref->GetLocalOperatorlocalEnergyR()),
                        vrel(gpu->GetLocalOperatorlocalEnergyI(), ref->GetLocalOperatorlocalEnergyI()));
        std::vector<double> go = gpu->GetOtherExpectationValues();
        for (int k : {0, 1, 3, 4, 5, 6, 7, 8}) e_oth = std::max(e_oth, rel(go[k], ro[k], esc + std::fabs(ro[3])));

        // g(r), S(k) in the plug-ins' own observables
        ref->CalculateAdditionalSystemProperties(R, uR, uI, phiR, phiI);
        gpu->CalculateAdditionalSystemProperties(R, uR, uI, phiR, phiI);
        auto oc_r = ref->GetAdditionalObservablesClone(), oc_g = gpu->GetAdditionalObservablesClone();
        double e_gr = 1e300, e_sk = 1e300;
        if (oc_r.observables.size() == 2 && oc_g.observables.size() == 2) {
            auto* gr_r = dynamic_cast<Observables::ObservableVsOnGrid*>(oc_r.observables[0]);
            auto* gr_g = dynamic_cast<Observables::ObservableVsOnGrid*>(oc_g.observables[0]);
            auto* sk_r = dynamic_cast<Observables::ObservableVsOnGrid*>(oc_r.observables[1]);
            auto* sk_g = dynamic_cast<Observables::ObservableVsOnGrid*>(oc_g.observables[1]);
            e_gr = vrel(gr_g->observablesV[0].values, gr_r->observablesV[0].values);
            e_sk = vrel(sk_g->observablesV[0].values, sk_r->observablesV[0].values);
            std::printf("gr_bins %zu %zu\n", gr_g->observablesV[0].values.size(), gr_r->observablesV[0].values.size());
            std::printf("gr_grid_err %.3e\n", vrel(gr_g->grid.gridCenterPoints, gr_r->grid.gridCenterPoints));
        }
        oc_r.Destroy();
        oc_g.Destroy();

        // correlated-sampling overloads on a kept sample equal the direct calls
        std::vector<ICorrelatedSamplingData*> cs(1);
        gpu->InitCorrelatedSamplingData(cs);
        cs[0]->R = R;
        gpu->FillCorrelatedSamplingData(cs[0]);
        std::vector<double> uR2 = uR;
        for (auto& v : uR2) v *= 1.01;
        gpu->CalculateWavefunction(cs[0], uR2, uI, phiR, phiI);
        const double ex_cs = gpu->GetExponent();
        gpu->CalculateExpectationValues(cs[0], uR2, uI, phiR, phiI);
        const double er_cs = gpu->GetLocalEnergyR();
        gpu->CalculateWavefunction(R, uR2, uI, phiR, phiI);
        gpu->CalculateExpectationValues(R, uR2, uI, phiR, phiI);
        const double e_cs = std::max(rel(ex_cs, gpu->GetExponent(), 1.0), rel(er_cs, gpu->GetLocalEnergyR(), 1.0));
        delete cs[0];
        auto com = gpu->GetCenterOfMass(R);

        std::printf("accepted %d\nmoves %d\n", accepted, moves);
        std::printf("err_expo %.3e\nerr_wf %.3e\nerr_quotient %.3e\n", e_expo, e_wf, e_q);
        std::printf("err_ER %.3e\nerr_EI %.3e\nerr_O %.3e\nerr_OO %.3e\nerr_OE %.3e\nerr_other %.3e\n", e_ER, e_EI,
                    e_O, e_OO, e_OE, e_oth);
        std::printf("err_gr %.3e\nerr_sk %.3e\nerr_cs %.3e\ncom %g\n", e_gr, e_sk, e_cs,
                    std::fabs(com[0]) + std::fabs(com[1]) + std::fabs(com[2]));
        std::printf("nic %d\nnormalization %d\nn_other %d\n", gpu->USE_NIC ? 1 : 0,
                    gpu->USE_NORMALIZATION_AND_PHASE ? 1 : 0, gpu->GetNumOfOtherExpectationValues());
        delete gpu;
    } catch (const tdvmc::Error& e) {
        std::fprintf(stderr, "tdvmc error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
