// ref_harness.cpp -- C entry points over the reference's own BosonsBulk
// physics plug-in, compiled from the sources where they lie in the reference
// checkout (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: used in this
// container to write golden vectors (tests/golden/make_golden.py), to pin the
// C restatement in tdvmc_oracle.c, and as the CPU baseline of bench.py.
//
// The reference defines its shape globals in src/TDVMC.cpp:84-89 and binds
// the *_DIM function pointers in Init (src/TDVMC.cpp:567-613); TDVMC.cpp needs
// MPI, so this harness defines the same globals and binds the pointers itself.
// BosonsBulk::InitSystem reads kVectors<D>D.json / kNorm<D>D.csv for its S(k)
// observable only (src/PhysicalSystems/BosonsBulk.cpp:124-137); the harness
// writes a small file set of that format into a scratch directory so the
// plug-in can be constructed without the reference's data directory.

#include "PhysicalSystems/BosonsBulk.h"
#include "SplineFactory.h"
#include "CSDataBulkSplines.h"
#include "ConfigItem.h"
#include "Utils.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

using namespace std;

string OUT_DIR;
int N;
int DIM;
int N_PARAM;
int IMAGINARY_TIME;
double LBOX;
double LBOX_R;
double LBOX_2;

namespace {

void bind_dim_functions(int dim)
{
    switch (dim) {
    case 1:
        VectorDotProduct_DIM = VectorDotProduct_1D;
        VectorNorm2_DIM = VectorNorm2_1D;
        VectorNorm_DIM = VectorNorm_1D;
        VectorDisplacement_DIM = VectorDisplacement_1D;
        GetVectorNIC_DIM = GetVectorNIC_1D;
        VectorDiffNIC_DIM = VectorDiffNIC_1D;
        VectorDisplacementNIC_DIM = VectorDisplacementNIC_1D;
        AssignVector_DIM = AssignVector_1D;
        break;
    case 2:
        VectorDotProduct_DIM = VectorDotProduct_2D;
        VectorNorm2_DIM = VectorNorm2_2D;
        VectorNorm_DIM = VectorNorm_2D;
        VectorDisplacement_DIM = VectorDisplacement_2D;
        GetVectorNIC_DIM = GetVectorNIC_2D;
        VectorDiffNIC_DIM = VectorDiffNIC_2D;
        VectorDisplacementNIC_DIM = VectorDisplacementNIC_2D;
        AssignVector_DIM = AssignVector_2D;
        break;
    default:
        VectorDotProduct_DIM = VectorDotProduct_3D;
        VectorNorm2_DIM = VectorNorm2_3D;
        VectorNorm_DIM = VectorNorm_3D;
        VectorDisplacement_DIM = VectorDisplacement_3D;
        GetVectorNIC_DIM = GetVectorNIC_3D;
        VectorDiffNIC_DIM = VectorDiffNIC_3D;
        VectorDisplacementNIC_DIM = VectorDisplacementNIC_3D;
        AssignVector_DIM = AssignVector_3D;
        break;
    }
}

// The first 50 shells of wave vectors in the format of kVectors<D>D.json:
// every integer vector n != 0 with non-negative components, grouped by |n|^2
// in increasing order, lexicographic within a shell (the order of the
// reference's config/kVectors<D>D.json), and kNorm<D>D.csv = |n|.
string write_kvector_dir(int dim)
{
    char tmpl[] = "/tmp/tdvmc_refk_XXXXXX";
    char* d = mkdtemp(tmpl);
    if (!d) return "";
    string dir = string(d) + "/";
    ofstream kv(dir + "kVectors" + to_string(dim) + "D.json");
    ofstream kn(dir + "kNorm" + to_string(dim) + "D.csv");
    kn.precision(17);
    kv << "{\"data\": [";
    int shells = 0;
    const int M = 64;
    for (int n2 = 1; shells < 50; n2++) {
        vector<vector<int>> sh;
        int v[3] = {0, 0, 0};
        const int lim[3] = {M, dim > 1 ? M : 1, dim > 2 ? M : 1};
        for (v[0] = 0; v[0] < lim[0]; v[0]++)
            for (v[1] = 0; v[1] < lim[1]; v[1]++)
                for (v[2] = 0; v[2] < lim[2]; v[2]++)
                    if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] == n2) sh.push_back(vector<int>(v, v + dim));
        if (sh.empty()) continue;
        kv << (shells ? "," : "") << "[";
        for (size_t q = 0; q < sh.size(); q++) {
            kv << (q ? "," : "") << "[";
            for (int a = 0; a < dim; a++) kv << (a ? "," : "") << sh[q][a];
            kv << "]";
        }
        kv << "]";
        kn << sqrt((double)n2) << "\n";
        shells++;
    }
    kv << "]}";
    kv.close();
    kn.close();
    return dir;
}

struct Setup {
    PhysicalSystems::BosonsBulk* sys = nullptr;
    string dir;
};

Setup make_system(int n, int dim, double L, int nparam, const double* sysparams, int nsys, double time,
                  int gr_bins = 50, const double* grid = nullptr)
{
    N = n;
    DIM = dim;
    N_PARAM = nparam;
    LBOX = L;
    LBOX_R = 1.0 / L;
    LBOX_2 = L / 2.0;
    IMAGINARY_TIME = 1;
    bind_dim_functions(dim);
    Setup s;
    s.dir = write_kvector_dir(dim);
    vector<double> p(sysparams, sysparams + nsys);
    s.sys = new PhysicalSystems::BosonsBulk(p, s.dir);
    // USE_NURBS: the factory hands NURBS_GRID to SetNodes before InitSystem (src/TDVMC.cpp:2845-2850)
    if (grid) s.sys->SetNodes(vector<double>(grid, grid + nparam));
    s.sys->SetPairDistributionBinCount(gr_bins);
    s.sys->InitSystem();
    s.sys->SetTime(time);
    return s;
}

// The plug-in is never deleted: its base destructor (IPhysicalSystem.h:71-74)
// deletes observables that BosonsBulk registers as member addresses
// (BosonsBulk.cpp:154-155), and the reference itself never destroys it.
void free_system(Setup& s)
{
    s.sys = nullptr;
    if (!s.dir.empty()) {
        unlink((s.dir + "kVectors" + to_string(DIM) + "D.json").c_str());
        unlink((s.dir + "kNorm" + to_string(DIM) + "D.csv").c_str());
        rmdir(s.dir.c_str());
    }
}

vector<vector<double>> to_R(const double* R)
{
    vector<vector<double>> out(N, vector<double>(DIM));
    for (int i = 0; i < N; i++)
        for (int a = 0; a < DIM; a++) out[i][a] = R[i * DIM + a];
    return out;
}

}  // namespace

extern "C" {

// Reference NIC displacement (src/Utils.cpp:376-382) with global box L.
double ref_displacement_nic(const double* ri, const double* rj, int dim, double L, double* out)
{
    LBOX = L;
    LBOX_R = 1.0 / L;
    LBOX_2 = L / 2.0;
    vector<double> a(ri, ri + dim), b(rj, rj + dim), r(dim);
    double d = VectorDisplacementNIC(a, b, r);
    for (int i = 0; i < dim; i++) out[i] = r[i];
    return d;
}

// Evaluate W fed configurations with BosonsBulk: full recompute of the basis
// sums (CalculateWavefunction) followed by CalculateExpectationValues.
// Outputs per walker: exponent, O[P], E_R, E_I, other[9],
// sD[S][N][D] and sD2[S][N] (S = P + 2 splines; may be NULL).
int ref_bosonsbulk_eval(int n, int dim, double L, int nparam, const double* sysparams, int nsys,
                        double time, int64_t W, const double* R, const double* uR, const double* uI,
                        double phiR, double phiI, double* expo, double* O, double* ER, double* EI,
                        double* other, double* sD, double* sD2)
{
    Setup s = make_system(n, dim, L, nparam, sysparams, nsys, time);
    vector<double> vR(uR, uR + nparam), vI(uI, uI + nparam);
    const int S = nparam + 2;
    for (int64_t w = 0; w < W; w++) {
        auto Rw = to_R(R + w * n * dim);
        s.sys->CalculateWavefunction(Rw, vR, vI, phiR, phiI);
        s.sys->CalculateExpectationValues(Rw, vR, vI, phiR, phiI);
        expo[w] = s.sys->GetExponent();
        auto lo = s.sys->GetLocalOperators();
        for (int k = 0; k < nparam; k++) O[w * nparam + k] = lo[k];
        ER[w] = s.sys->GetLocalEnergyR();
        EI[w] = s.sys->GetLocalEnergyI();
        auto ov = s.sys->GetOtherExpectationValues();
        for (int k = 0; k < 9; k++) other[w * 9 + k] = ov[k];
        if (sD || sD2) {
            CSDataBulkSplines cs;
            s.sys->FillCorrelatedSamplingData(&cs);
            for (int sp = 0; sp < S; sp++)
                for (int i = 0; i < n; i++) {
                    if (sD)
                        for (int a = 0; a < dim; a++)
                            sD[((w * S + sp) * n + i) * dim + a] = cs.splineSumsD[sp][i][a];
                    if (sD2) sD2[(w * S + sp) * n + i] = cs.splineSumsD2[sp][i];
                }
        }
    }
    free_system(s);
    return 0;
}

// A BosonsBulk plug-in with the shape globals and k-vector files set up, as
// the driver's factory and Init leave it (src/TDVMC.cpp:2850, :567-613), for
// a caller that drives it through PhysicalSystems::IPhysicalSystem*
// (tests/cpp/iphys_check.cpp).  Never deleted (see free_system).
void* ref_new_bosonsbulk(int n, int dim, double L, int nparam, const double* sysparams, int nsys, int gr_bins)
{
    Setup s = make_system(n, dim, L, nparam, sysparams, nsys, 0.0, gr_bins);
    return static_cast<PhysicalSystems::IPhysicalSystem*>(s.sys);
}

// ref_bosonsbulk_eval on the non-uniform knots of grid [nparam] (USE_NURBS with
// NURBS_GRID; BosonsBulk::SetNodes, BosonsBulk.cpp:35-44).
int ref_bosonsbulk_eval_nurbs(int n, int dim, double L, int nparam, const double* grid, const double* sysparams,
                              int nsys, double time, int64_t W, const double* R, const double* uR, const double* uI,
                              double phiR, double phiI, double* expo, double* O, double* ER, double* EI,
                              double* other, double* sD, double* sD2)
{
    Setup s = make_system(n, dim, L, nparam, sysparams, nsys, time, 50, grid);
    vector<double> vR(uR, uR + nparam), vI(uI, uI + nparam);
    const int S = nparam + 2;
    for (int64_t w = 0; w < W; w++) {
        auto Rw = to_R(R + w * n * dim);
        s.sys->CalculateWavefunction(Rw, vR, vI, phiR, phiI);
        s.sys->CalculateExpectationValues(Rw, vR, vI, phiR, phiI);
        expo[w] = s.sys->GetExponent();
        auto lo = s.sys->GetLocalOperators();
        for (int k = 0; k < nparam; k++) O[w * nparam + k] = lo[k];
        ER[w] = s.sys->GetLocalEnergyR();
        EI[w] = s.sys->GetLocalEnergyI();
        auto ov = s.sys->GetOtherExpectationValues();
        for (int k = 0; k < 9; k++) other[w * 9 + k] = ov[k];
        CSDataBulkSplines cs;
        s.sys->FillCorrelatedSamplingData(&cs);
        for (int sp = 0; sp < S; sp++)
            for (int i = 0; i < n; i++) {
                for (int a = 0; a < dim; a++) sD[((w * S + sp) * n + i) * dim + a] = cs.splineSumsD[sp][i][a];
                sD2[(w * S + sp) * n + i] = cs.splineSumsD2[sp][i];
            }
    }
    free_system(s);
    return 0;
}

// Incremental move of one particle (CalculateWFQuotient, BosonsBulk.cpp:553-657):
// R is the configuration before the move; particle p goes to newpos.
// Returns exponentNew; *quotient receives exp(2 (exponentNew - exponent)).
double ref_bosonsbulk_move(int n, int dim, double L, int nparam, const double* sysparams, int nsys,
                           const double* R, const double* uR, int p, const double* newpos, double* quotient)
{
    Setup s = make_system(n, dim, L, nparam, sysparams, nsys, 0.0);
    vector<double> vR(uR, uR + nparam), vI(nparam, 0.0);
    auto Rw = to_R(R);
    s.sys->CalculateWavefunction(Rw, vR, vI, 0.0, 0.0);
    vector<double> old = Rw[p];
    for (int a = 0; a < dim; a++) Rw[p][a] = newpos[a];
    *quotient = s.sys->CalculateWFQuotient(Rw, vR, vI, 0.0, 0.0, p, old);
    double e = s.sys->GetExponentNew();
    free_system(s);
    return e;
}

// CPU baseline: one walker doing n_samples x (n_therm single-particle
// Metropolis moves, then one measurement) through the reference plug-in's own
// incremental ratio and measurement; the move loop mirrors DoMetropolisStep
// (src/TDVMC.cpp:858-916) with mt19937_64 and std::normal_distribution.
// Returns the accepted-move count; *mean_E receives the sample mean of E_R.
int64_t ref_bosonsbulk_sample(int n, int dim, double L, int nparam, const double* sysparams, int nsys,
                              double* R, const double* uR, double step, uint64_t seed,
                              int64_t n_samples, int64_t n_therm, double* mean_E)
{
    Setup s = make_system(n, dim, L, nparam, sysparams, nsys, 0.0);
    vector<double> vR(uR, uR + nparam), vI(nparam, 0.0);
    auto Rw = to_R(R);
    mt19937_64 gen(seed);
    uniform_real_distribution<double> uni(0.0, 1.0);
    normal_distribution<double> nrm(0.0, 1.0);
    uniform_int_distribution<int> pick(0, n - 1);
    s.sys->CalculateWavefunction(Rw, vR, vI, 0.0, 0.0);
    int64_t acc = 0;
    double esum = 0.0;
    for (int64_t smp = 0; smp < n_samples; smp++) {
        for (int64_t t = 0; t < n_therm; t++) {
            int p = pick(gen);
            vector<double> old(Rw[p]);
            for (int a = 0; a < dim; a++) Rw[p][a] += nrm(gen) * step;
            double q = s.sys->CalculateWFQuotient(Rw, vR, vI, 0.0, 0.0, p, old);
            bool ok = std::isfinite(q) && std::isfinite(s.sys->GetExponentNew());
            if (!ok || q < uni(gen)) {
                Rw[p] = old;
            } else {
                s.sys->AcceptMove();
                acc++;
            }
        }
        s.sys->CalculateExpectationValues(Rw, vR, vI, 0.0, 0.0);
        esum += s.sys->GetLocalEnergyR();
    }
    for (int i = 0; i < n; i++)
        for (int a = 0; a < dim; a++) R[i * dim + a] = Rw[i][a];
    *mean_E = n_samples > 0 ? esum / (double)n_samples : 0.0;
    free_system(s);
    return acc;
}

// Additional observables of W fed configurations
// (BosonsBulk::CalculateAdditionalSystemProperties, BosonsBulk.cpp:474-520):
// the pair distribution with gr_bins bins requested (GR_BIN_COUNT) and S(k)
// on the 50 shells written above.  *n_gr receives the grid's bin count as the
// reference computes it (Grid.cpp:17-31); gr is [W][n_gr], sk [W][50],
// gr_grid the bin centres [n_gr], sk_grid the shell norms |k| [50].
int ref_bosonsbulk_observables(int n, int dim, double L, int nparam, const double* sysparams, int nsys,
                               int64_t W, const double* R, int gr_bins, int gr_cap, int* n_gr, double* gr,
                               double* sk, double* gr_grid, double* sk_grid)
{
    Setup s = make_system(n, dim, L, nparam, sysparams, nsys, 0.0, gr_bins);
    vector<double> vR(nparam, 0.0), vI(nparam, 0.0);
    int ng = 0;
    for (int64_t w = 0; w < W; w++) {
        auto Rw = to_R(R + w * n * dim);
        s.sys->CalculateAdditionalSystemProperties(Rw, vR, vI, 0.0, 0.0);
        auto oc = s.sys->GetAdditionalObservablesClone();
        auto* g = dynamic_cast<Observables::ObservableVsOnGrid*>(oc.observables[0]);
        auto* q = dynamic_cast<Observables::ObservableVsOnGrid*>(oc.observables[1]);
        if (!g || !q) return -1;
        ng = g->grid.count;
        if (ng > gr_cap || (int)q->observablesV[0].values.size() != 50) return -2;
        for (int b = 0; b < ng; b++) gr[w * ng + b] = g->observablesV[0].values[b];
        for (int k = 0; k < 50; k++) sk[w * 50 + k] = q->observablesV[0].values[k];
        if (w == 0) {
            for (int b = 0; b < ng; b++) gr_grid[b] = g->grid.gridCenterPoints[b];
            for (int k = 0; k < 50; k++) sk_grid[k] = q->grid.grid[k];
        }
        oc.Destroy();
    }
    *n_gr = ng;
    free_system(s);
    return 0;
}

// The reference's file writers (src/Utils.cpp:818-1016) with OUT_DIR = dir.
// SplineFactory::GetWeights (GetWeights3, src/SplineFactory.cpp:54-101) and the
// zero-slope boundary-condition factors with uniform = true
// (SetBoundaryConditions3_1D_OR_2 / _CO_2, :423-456, :563-596) on the knot
// vector nodes [n]: w [n - 4][4][4], bcs [2][3], bce [2][3].
int ref_spline_weights(const double* nodes, int n, double* w, double* bcs, double* bce)
{
    vector<double> nv(nodes, nodes + n);
    auto W = SplineFactory::GetWeights(nv);
    for (size_t s = 0; s < W.size(); s++)
        for (int p = 0; p < 4; p++)
            for (int c = 0; c < 4; c++) w[(s * 4 + p) * 4 + c] = W[s][p][c];
    vector<vector<double> > a, b;
    SplineFactory::SetBoundaryConditions3_1D_OR_2(nv, a, true);
    SplineFactory::SetBoundaryConditions3_1D_CO_2(nv, b, nv[n - 4], true);
    for (int i = 0; i < 2; i++)
        for (int k = 0; k < 3; k++) {
            bcs[i * 3 + k] = a[i][k];
            bce[i * 3 + k] = b[i][k];
        }
    return (int)W.size();
}

// The reference's cubic boundary-condition factors of one family
// (SetBoundaryConditions3_1D_OR_k / _CO_k, src/SplineFactory.cpp:377-656)
// on the knot vector nodes [n] with the given uniform flag: rows into
// f [rows][3]; returns the row count, or -1 for an unknown family.
int ref_bc_factors(int family, int at_cutoff, const double* nodes, int n, int uniform, double* f)
{
    vector<double> nv(nodes, nodes + n);
    vector<vector<double> > b;
    const double rc = nv[n - 4];
    const bool u = uniform != 0;
    if (!at_cutoff) {
        switch (family) {
        case 1: SplineFactory::SetBoundaryConditions3_1D_OR_1(nv, b, u); break;
        case 2: SplineFactory::SetBoundaryConditions3_1D_OR_2(nv, b, u); break;
        case 3: SplineFactory::SetBoundaryConditions3_1D_OR_3(nv, b, u); break;
        case 4: SplineFactory::SetBoundaryConditions3_1D_OR_4(nv, b, u); break;
        default: return -1;
        }
    } else {
        switch (family) {
        case 1: SplineFactory::SetBoundaryConditions3_1D_CO_1(nv, b, rc, u); break;
        case 2: SplineFactory::SetBoundaryConditions3_1D_CO_2(nv, b, rc, u); break;
        case 3: SplineFactory::SetBoundaryConditions3_1D_CO_3(nv, b, rc, u); break;
        case 4: SplineFactory::SetBoundaryConditions3_1D_CO_4(nv, b, rc, u); break;
        default: return -1;
        }
    }
    for (size_t i = 0; i < b.size(); i++)
        for (int k = 0; k < 3; k++) f[i * 3 + k] = b[i][k];
    return (int)b.size();
}

int ref_write_vector(const char* dir, const char* name, const double* v, int n, const char* header)
{
    OUT_DIR = dir;
    vector<double> d(v, v + n);
    WriteDataToFile(d, name, header, 1);
    return 0;
}

int ref_append_vector(const char* dir, const char* name, const double* v, int n)
{
    OUT_DIR = dir;
    vector<double> d(v, v + n);
    AppendDataToFile(d, name);
    return 0;
}

int ref_append_scalar(const char* dir, const char* name, double x)
{
    OUT_DIR = dir;
    AppendDataToFile(x, name);
    return 0;
}

// ConfigItem::getJsonString (src/ConfigItem.cpp:82-106) of one value of each
// type: kind 0 string (s), 1 int (i), 2 double (d[0]), 3 int array (ia[0..n)),
// 4 double array (d[0..n)); the text goes to out (cap bytes).
int ref_config_json(int kind, const char* s, int i, const double* d, const int* ia, int n, char* out, int cap)
{
    string sv = s ? s : "";
    int iv = i;
    double dv = d ? d[0] : 0.0;
    vector<int> va(ia, ia + (ia ? n : 0));
    vector<double> vd(d, d + (d ? n : 0));
    string r;
    if (kind == 0) r = ConfigItem("X", &sv, STRING).getJsonString();
    else if (kind == 1) r = ConfigItem("X", &iv, INT).getJsonString();
    else if (kind == 2) r = ConfigItem("X", &dv, DOUBLE).getJsonString();
    else if (kind == 3) r = ConfigItem("X", va, ARR_INT).getJsonString();
    else r = ConfigItem("X", vd, ARR_DOUBLE).getJsonString();
    if ((int)r.size() + 1 > cap) return -1;
    memcpy(out, r.c_str(), r.size() + 1);
    return (int)r.size();
}

}  // extern "C"
