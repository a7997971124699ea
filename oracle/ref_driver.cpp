// ref_driver.cpp -- C entry points over the reference driver's own TDVP
// linear-algebra functions (src/TDVMC.cpp), compiled with the whole reference
// source tree where it lies (see oracle/Makefile, target `drv`; the driver's
// main is renamed to reference_main at compile time, nothing else changes).
// TEST INFRASTRUCTURE ONLY: used in this container by
// tests/golden/make_golden.py to write golden vectors for the right-hand
// sides, both linear solvers, phi-dot, the gauge shift, the Lennard-Jones
// pair potential and the successive-run coordinate files.  Never shipped to,
// or loaded on, a GPU machine.
//
// The functions read the driver's globals (src/TDVMC.cpp:84-160): the moment
// sums localOperators, localEnergyR/I, localOperatorsMatrix,
// localOperatorlocalEnergyR/I, and the settings N_PARAM, USE_PARAM_START,
// USED_PARAM_COUNT, IMAGINARY_TIME, LINEAR_EQUATION_SOLVER_TYPE,
// USE_PRECONDITIONING.  BuildSystemOfEquationsForParameters reads
// sys->USE_NORMALIZATION_AND_PHASE (:1545), so a BosonsBulk plug-in is
// constructed once for `sys`.

#include "PhysicalSystems/BosonsBulk.h"
#include "Potentials/LJ.h"
#include "Potentials/LJ_He_He.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

using namespace std;

// globals defined in src/TDVMC.cpp
extern int N, DIM, N_PARAM, USE_PARAM_START, USE_PARAM_END, USED_PARAM_COUNT, IMAGINARY_TIME;
extern int LINEAR_EQUATION_SOLVER_TYPE, USE_PRECONDITIONING;
extern double LBOX, LBOX_R, LBOX_2;
extern vector<double> localOperators;
extern double localEnergyR, localEnergyI;
extern vector<vector<double> > localOperatorsMatrix;
extern vector<double> localOperatorlocalEnergyR, localOperatorlocalEnergyI;
extern bool doNotAcceptStep;
extern PhysicalSystems::IPhysicalSystem* sys;

extern string OUT_DIR, configDirectory;

// functions defined in src/TDVMC.cpp (no header declares them)
void WriteParticleInputFile(string fileName, vector<vector<double> >& R);
bool LoadLastPositionsFromFile(string filename, vector<vector<double> >& R);
void MoveCoordinatesToFirstCell(vector<vector<double> >& R);
void BuildSystemOfEquationsForParameters(vector<vector<double> >& matrix, vector<double>& energiesReal,
                                         vector<double>& energiesImag);
void SolveForParametersDot(vector<double>& uDotR, vector<double>& uDotI, double* phiDotR, double* phiDotI);
void CalculatePhiDot(vector<double>& uDotR, vector<double>& uDotI, double* phiDotR, double* phiDotI);
void AdjustParameters(vector<double>& uR, vector<double>& uI, double* phiR, double* phiI);

namespace {

// A BosonsBulk for `sys` (1-D box; its InitSystem reads kVectors1D.json and
// kNorm1D.csv, written here into a scratch directory in that file format).
void ensure_sys()
{
    if (sys) return;
    char tmpl[] = "/tmp/tdvmc_refd_XXXXXX";
    char* d = mkdtemp(tmpl);
    string dir = string(d ? d : "/tmp") + "/";
    {
        ofstream kv(dir + "kVectors1D.json");
        kv << "{\"data\": [";
        for (int s = 1; s <= 50; s++) kv << (s > 1 ? "," : "") << "[[" << s << "]]";
        kv << "]}";
        ofstream kn(dir + "kNorm1D.csv");
        for (int s = 1; s <= 50; s++) kn << s << "\n";
    }
    N = 4;
    DIM = 1;
    N_PARAM = 6;
    LBOX = 8.0;
    LBOX_R = 1.0 / LBOX;
    LBOX_2 = LBOX / 2.0;
    vector<double> p = {0.5, 1.0};
    auto* b = new PhysicalSystems::BosonsBulk(p, dir);
    b->SetPairDistributionBinCount(10);
    b->InitSystem();
    sys = b;
}

// The driver's moment sums from FP64 means: O [P], E = (E_R, E_I),
// OO [P][P] = <O O^T>, OER / OEI [P] = <O E_R>, <O E_I>.
void set_moments(int P, const double* O, const double* E, const double* OO, const double* OER, const double* OEI)
{
    localOperators.assign(O, O + P);
    localEnergyR = E[0];
    localEnergyI = E[1];
    localOperatorsMatrix.assign(P, vector<double>(P));
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++) localOperatorsMatrix[i][j] = OO[(size_t)i * P + j];
    localOperatorlocalEnergyR.assign(OER, OER + P);
    localOperatorlocalEnergyI.assign(OEI, OEI + P);
}

void set_shape(int P, int imag, int solver, int precondition)
{
    ensure_sys();
    N_PARAM = P;
    USE_PARAM_START = 0;
    USE_PARAM_END = P;
    USED_PARAM_COUNT = P;
    IMAGINARY_TIME = imag;
    LINEAR_EQUATION_SOLVER_TYPE = solver;
    USE_PRECONDITIONING = precondition;
    doNotAcceptStep = false;
}

}  // namespace

extern "C" {

// BuildSystemOfEquationsForParameters (src/TDVMC.cpp:1539-1554 ->
// :1506-1537): S [P][P], bR [P], bI [P].
int refd_build_rhs(int P, int imag, const double* O, const double* E, const double* OO, const double* OER,
                   const double* OEI, double* S, double* bR, double* bI)
{
    set_shape(P, imag, 0, 0);
    set_moments(P, O, E, OO, OER, OEI);
    vector<vector<double> > m;
    vector<double> er, ei;
    BuildSystemOfEquationsForParameters(m, er, ei);
    for (int i = 0; i < P; i++) {
        bR[i] = er[i];
        bI[i] = ei[i];
        for (int j = 0; j < P; j++) S[(size_t)i * P + j] = m[i][j];
    }
    return 0;
}

// SolveForParametersDot (src/TDVMC.cpp:1713-1828) with CalculatePhiDot
// (:1658-1682) from phi-dot = 0, as CalculateNextParametersEuler calls it
// (:1834-1853).  Returns the doNotAcceptStep flag of the Cholesky (:1580-1590).
int refd_solve(int P, int imag, int solver, int precondition, const double* O, const double* E, const double* OO,
               const double* OER, const double* OEI, double* xR, double* xI, double* phiDot)
{
    set_shape(P, imag, solver, precondition);
    set_moments(P, O, E, OO, OER, OEI);
    vector<double> uR, uI;
    double pR = 0.0, pI = 0.0;
    SolveForParametersDot(uR, uI, &pR, &pI);
    for (int i = 0; i < P; i++) {
        xR[i] = uR[i];
        xI[i] = uI[i];
    }
    phiDot[0] = pR;
    phiDot[1] = pI;
    return doNotAcceptStep ? 1 : 0;
}

// CalculatePhiDot alone (src/TDVMC.cpp:1658-1682) on given alpha-dot, from phi-dot = 0.
int refd_phi_dot(int P, int imag, const double* O, const double* E, const double* xR, const double* xI,
                 double* phiDot)
{
    set_shape(P, imag, 0, 0);
    localOperators.assign(O, O + P);
    localEnergyR = E[0];
    localEnergyI = E[1];
    vector<double> uR(xR, xR + P), uI(xI, xI + P);
    double pR = 0.0, pI = 0.0;
    CalculatePhiDot(uR, uI, &pR, &pI);
    phiDot[0] = pR;
    phiDot[1] = pI;
    return 0;
}

// AdjustParameters (src/TDVMC.cpp:2531-2567) over the first n_param entries
// (the driver's N_PARAM) of uR / uI, in place.
int refd_adjust(int n_param, double* uR, double* uI, int P)
{
    set_shape(n_param, 1, 0, 0);
    vector<double> vR(uR, uR + P), vI(uI, uI + P);
    double pR = 0.0, pI = 0.0;
    AdjustParameters(vR, vI, &pR, &pI);
    for (int i = 0; i < P; i++) {
        uR[i] = vR[i];
        uI[i] = vI[i];
    }
    return 0;
}

// The successive-run coordinate file as the driver writes it
// (src/TDVMC.cpp:4007-4018): MoveCoordinatesToFirstCell (AlignCoordinates for
// a periodic system, :2569-2582) when align != 0, then WriteParticleInputFile
// (:493-504) as out_dir + name + ".dat".
int refd_write_particle_file(const char* out_dir, const char* name, const double* R, int n, int dim, double L,
                             int align)
{
    N = n;
    DIM = dim;
    LBOX = L;
    LBOX_R = 1.0 / L;
    LBOX_2 = L / 2.0;
    OUT_DIR = out_dir;
    vector<vector<double> > Rv(n, vector<double>(dim));
    for (int i = 0; i < n; i++)
        for (int a = 0; a < dim; a++) Rv[i][a] = R[i * dim + a];
    if (align) MoveCoordinatesToFirstCell(Rv);
    WriteParticleInputFile(name, Rv);
    return 0;
}

// LoadLastPositionsFromFile (src/TDVMC.cpp:640-670): config_dir + name + ".csv".
int refd_load_positions(const char* config_dir, const char* name, double* R, int n, int dim)
{
    N = n;
    DIM = dim;
    configDirectory = config_dir;
    vector<vector<double> > Rv(n, vector<double>(dim, 0.0));
    const bool ok = LoadLastPositionsFromFile(name, Rv);
    for (int i = 0; i < n; i++)
        for (int a = 0; a < dim; a++) R[i * dim + a] = Rv[i][a];
    return ok ? 0 : 1;
}

// LJ::GetPotential (src/Potentials/LJ.cpp:16-22) with the He-He constants of
// LJ_He_He (src/Potentials/LJ_He_He.cpp:13-14), or with unit constants.
int refd_lj(int he, const double* r, int n, double* v)
{
    Potentials::LJ unit;
    Potentials::LJ_He_He he4;
    Potentials::IPairPotential* p = he ? static_cast<Potentials::IPairPotential*>(&he4) : &unit;
    for (int i = 0; i < n; i++) v[i] = p->GetPotential(r[i]);
    return 0;
}

}  // extern "C"
