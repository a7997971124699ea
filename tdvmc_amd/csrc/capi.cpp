// C ABI of the TDVMC hot path: context, configuration, device buffers and the
// host-side orchestration of the kernels (see include/tdvmc.h).

#include "../../include/tdvmc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.hpp"
#include "kernels.hpp"
#include "knots.hpp"

using namespace tdvmc;

// Fixed device scratch of tdvmc_adjoint_noalloc (one copy per GPU, held until
// the process ends): the per-walker evaluation outputs the adjoint contracts.
// In host code `scratch` names a host-side stand-in; its device address comes
// from hipGetSymbolAddress, once per device (init_scratch).
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

thread_local std::string g_create_error;

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

struct Prof {
    double ms = 0.0;
    int64_t launches = 0;
};

template <typename T>
struct DevTables {
    T* spl = nullptr;
    T* dspl = nullptr;
    T* kv = nullptr;
    T* alpha = nullptr;
    T* kag = nullptr;
    bool dirty = true;
};

}  // namespace

struct tdvmc_ctx {
    Config cfg;
    KSys sy{};
    NuKnots knots;                    // non-uniform knots (USE_NURBS)
    int bc_family[2] = {2, 2};        // SPLINE_BC_ORIGIN, SPLINE_BC_CUTOFF
    // moments over one column per k / -k pair (the two slots hold the same
    // |rho_k|^2): computed column -> O column, and -> its full indices
    std::vector<int> mom_cols, mom_outmap;
    int* dMomCols = nullptr;
    int* dMomOut = nullptr;
    double* dMeanCols = nullptr;
    double* dNuCell = nullptr;        // device copies of knots.cells / knots.beta
    double* dNuBeta = nullptr;
    std::vector<double> pot;          // SYSTEM_PARAMS
    int pot_type = 0;
    double time = 0.0;
    int step = 0;
    int precision = TDVMC_FP64;
    int ode_type = 0, imag_time = 0, precondition = 0, spw = 1;
    int solver_type = 0;              // LINEAR_EQUATION_SOLVER_TYPE (0 Cholesky, 1 + mean subtraction)
    int normalize_wf = 0, adjust_params = 0;
    int accept_check = 0, accept_type = 0;
    double reg_eps = 1e-3, dt = 0.0;
    double* dXprev = nullptr;         // parameters before the step (acceptance check)
    int64_t therm_moves = 0, init_moves = 0;
    uint64_t seed = 1;
    int lanes = 0;                    // lanes per walker set by the caller (0: tuned or default)
    int lanes_tuned[2] = {0, 0};      // chosen by tdvmc_autotune per precision (0: not tuned)
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev_join = nullptr;     // orders a caller's stream against the context stream (*_dev entries)
    int device = 0;

    std::vector<double> uR, uI;
    double phiR = 0.0, phiI = 0.0;
    bool host_params_stale = false;

    int64_t W = 0;
    double* dR = nullptr;
    double* dRho = nullptr;     // [W][Kh][2] rho_k of the k representatives left by the last sweep (float or double)
    unsigned long long* dAcc = nullptr;
    uint32_t* dEpoch = nullptr;
    double* dUR = nullptr;
    double* dUI = nullptr;
    double* dX = nullptr;       // packed parameters [2P+2]
    double* dXtmp = nullptr;
    double* dK[4] = {nullptr, nullptr, nullptr, nullptr};
    DevTables<float> tf;
    DevTables<double> td;
    std::vector<float> kv_f;          // [Kh][4] representatives
    std::vector<double> kv_d;
    std::vector<int> kmap;            // [Kh][2] parameter slots of k and -k
    int* dKmap = nullptr;

    int64_t cap_samples = 0;
    double* dSamp = nullptr;          // [spw][W][N][D] sample configurations kept for reuse (spw > 1)
    int64_t samp_elems = 0;
    double* dO64 = nullptr;
    float* dO32 = nullptr;
    double* dER = nullptr;
    double* dEI = nullptr;
    double* dExpo = nullptr;
    double* dOther = nullptr;

    double* dMean = nullptr;
    double* dS = nullptr;
    double* dF = nullptr;
    MomentsWork mw{};
    int64_t mw_samples = 0;

    double* dWork = nullptr;
    int* dStatus = nullptr;

    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;
    std::string graph_key;

    bool prof = false;
    bool capturing = false;     // building a profiled graph from captured segments
    bool cap_active = false;    // a stream capture is open
    bool in_capture = false;    // enqueueing a TDVP step into a graph (its replays count the trials)
    hipGraph_t seg_top = nullptr;
    hipGraphNode_t seg_last = nullptr;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
        bool owned;   // false: the events belong to the captured graph
    };
    std::vector<Pending> pending;
    std::vector<Pending> graph_ev;   // event pairs recorded by nodes of the captured graph
    std::map<std::string, Prof> prof_acc;
    int64_t trials_total = 0;
    // bumped whenever a value that a batch's device descriptors copy changes
    // (the walker count and the buffers sized by it, the pair potential in force)
    uint64_t gen = 0;

    // pair distribution and S(k) (BosonsBulk's additional observables), set up on first use
    struct Obs {
        bool ready = false;
        int count = 0, nshell = 0, V = 0, nmax = 0;
        double gmax = 0.0, spacing = 0.0;
        std::vector<double> scale, centres, kgrid;
        int* dVec = nullptr;                  // [V][4]
        int* dOff = nullptr;                  // [nshell + 1]
        double* dScale = nullptr;             // [count]
        unsigned long long* dTot = nullptr;   // [count]
        double* dSkAcc = nullptr;             // [nshell]
        double* dSkW = nullptr;               // [W][nshell]
        int64_t skw_rows = 0;
    } obs;
};

namespace {

#define HIPC(expr)                                                                  \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);           \
            return TDVMC_ERR_HIP;                                                   \
        }                                                                           \
    } while (0)

#define CHECK_ARG(cond, msg)            \
    do {                                \
        if (!(cond)) {                  \
            ctx->err = (msg);           \
            return TDVMC_ERR_BAD_ARG;   \
        }                               \
    } while (0)

int P_of(const tdvmc_ctx* ctx) { return ctx->sy.P; }

template <typename T>
hipError_t dalloc(T** p, size_t n)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (n == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// K shortest non-zero integer vectors ordered by (|n|^2, lexicographic n);
// both n and -n are kept.
std::vector<std::vector<int>> shortest_lattice_vectors(int D, int K)
{
    std::vector<std::vector<int>> out;
    if (K <= 0) return out;
    for (int m = 1;; m++) {
        std::vector<std::vector<int>> vs;
        std::vector<int> v(D, -m);
        while (true) {
            bool nz = false;
            for (int a = 0; a < D; a++) nz |= v[a] != 0;
            if (nz) vs.push_back(v);
            int a = D - 1;
            while (a >= 0 && v[a] == m) {
                v[a] = -m;
                a--;
            }
            if (a < 0) break;
            v[a]++;
        }
        auto n2 = [](const std::vector<int>& x) {
            int s = 0;
            for (int e : x) s += e * e;
            return s;
        };
        std::sort(vs.begin(), vs.end(), [&](const std::vector<int>& a, const std::vector<int>& b) {
            int na = n2(a), nb = n2(b);
            if (na != nb) return na < nb;
            return a < b;
        });
        if ((int)vs.size() >= K && n2(vs[K - 1]) < (m + 1) * (m + 1)) {
            vs.resize(K);
            return vs;
        }
    }
}

void resolve_potential(tdvmc_ctx* ctx)
{
    const auto& p = ctx->pot;
    double a = p.size() > 0 ? p[0] : 0.0, b = p.size() > 1 ? p[1] : 0.0;
    if ((ctx->pot_type == 1 || ctx->pot_type == 2) && p.size() > 4 && ctx->time >= p[2]) {
        a = p[3];
        b = p[4];
    }
    if (a != ctx->sy.pot_a || b != ctx->sy.pot_b) ctx->gen++;
    ctx->sy.pot_a = a;
    ctx->sy.pot_b = b;
}

// Fresh sampling stages of one TDVP step of ODE type `ode` (the stages that
// sweep): Euler 1, predictor-corrector 2 (k1 and the corrector), RK4 4; the
// variants reusing the samples of k1 (types 3 and 5) sample once.
int fresh_stages(int ode) { return ode == 2 ? 4 : (ode == 1 ? 2 : 1); }

int build_system(tdvmc_ctx* ctx)
{
    Config& c = ctx->cfg;
    KSys& s = ctx->sy;
    if (c.DIM < 1 || c.DIM > 3) {
        ctx->err = "DIM must be 1, 2 or 3";
        return TDVMC_ERR_BAD_CONFIG;
    }
    if (c.N < 2) {
        ctx->err = "N must be >= 2";
        return TDVMC_ERR_BAD_CONFIG;
    }
    if (c.N_PARAM < 4) {
        ctx->err = "N_PARAM (pair-spline parameters) must be >= 4";
        return TDVMC_ERR_BAD_CONFIG;
    }
    if (c.USE_PARAM_START != 0 || c.USE_PARAM_END != 0) {
        ctx->err = "USE_PARAM_START/END: only the full parameter set is implemented";
        return TDVMC_ERR_BAD_CONFIG;
    }
    if (c.LINEAR_EQUATION_SOLVER_TYPE != 0 && c.LINEAR_EQUATION_SOLVER_TYPE != 1) {
        ctx->err = "LINEAR_EQUATION_SOLVER_TYPE must be 0 or 1";
        return TDVMC_ERR_BAD_CONFIG;
    }
    if (c.USE_PARAMETER_ACCEPTANCE_CHECK == 1 &&
        (c.PARAMETER_ACCEPTANCE_CHECK_TYPE < 0 || c.PARAMETER_ACCEPTANCE_CHECK_TYPE > 2)) {
        ctx->err = "PARAMETER_ACCEPTANCE_CHECK_TYPE must be 0, 1 or 2";
        return TDVMC_ERR_BAD_CONFIG;
    }
    s.N = c.N;
    s.D = c.DIM;
    s.P_s = c.N_PARAM;
    s.K = c.K_VECTORS;
    s.one_body = c.ONE_BODY ? 1 : 0;
    s.off_pair = s.one_body;
    s.off_k = s.off_pair + s.P_s;
    s.P = s.off_k + s.K;
    double L = c.LBOX;
    if (L <= 0 && c.TRAP_OMEGA == 0.0 && c.RHO > 0) L = std::pow((double)c.N / c.RHO, 1.0 / c.DIM);
    s.periodic = L > 0 ? 1 : 0;
    s.L = L > 0 ? L : 0.0;
    s.Linv = L > 0 ? 1.0 / L : 0.0;
    s.rc = c.SPLINE_CUTOFF > 0 ? c.SPLINE_CUTOFF : L / 2.0;
    if (!(s.rc > 0)) {
        ctx->err = "open systems need SPLINE_CUTOFF > 0";
        return TDVMC_ERR_BAD_CONFIG;
    }
    if (!s.periodic && s.K > 0) {
        ctx->err = "K_VECTORS needs a periodic box (LBOX > 0)";
        return TDVMC_ERR_BAD_CONFIG;
    }
    if (s.K < 0 || s.K > 512) {
        ctx->err = "K_VECTORS must be in [0, 512]";
        return TDVMC_ERR_BAD_CONFIG;
    }
    s.nu = 0;
    s.ncell = 0;
    s.cinv = 0.0;
    if (c.USE_NURBS) {
        // USE_NURBS with NURBS_GRID (BosonsBulk::SetNodes, src/TDVMC.cpp:2845-2850):
        // the spline ends at the last knot (maxDistance, BosonsBulk.cpp:72)
        if (!s.periodic) {
            ctx->err = "USE_NURBS: non-uniform knots are implemented for the periodic bulk system";
            return TDVMC_ERR_BAD_CONFIG;
        }
        if ((int)c.NURBS_GRID.size() != c.N_PARAM) {
            ctx->err = "USE_NURBS: NURBS_GRID must hold N_PARAM knots";
            return TDVMC_ERR_BAD_CONFIG;
        }
        std::string e;
        if (!nu_build(c.NURBS_GRID, L / 2.0, &ctx->knots, &e)) {
            ctx->err = "USE_NURBS: " + e;
            return TDVMC_ERR_BAD_CONFIG;
        }
        s.nu = 1;
        s.ncell = ctx->knots.ncell;
        s.cinv = ctx->knots.cinv;
        s.rc = c.NURBS_GRID.back();
    }
    // boundary conditions of the pair spline (BosonsBulk::InitSystem,
    // BosonsBulk.cpp:75-91, with the family selectable): the knots follow from
    // N_PARAM, so the two families must leave P_s + 2 - 6 + n0 + n1 = P_s
    // parameters, as the reference's count check requires
    {
        const int f0 = c.SPLINE_BC_ORIGIN ? c.SPLINE_BC_ORIGIN : 2;
        const int f1 = c.SPLINE_BC_CUTOFF ? c.SPLINE_BC_CUTOFF : 2;
        s.bc_n0 = bc_factors(f0, false, s.bc_f0);
        s.bc_n1 = bc_factors(f1, true, s.bc_f1);
        if (s.bc_n0 == 0 || s.bc_n1 == 0) {
            ctx->err = "SPLINE_BC_ORIGIN and SPLINE_BC_CUTOFF must be 1, 2, 3 or 4";
            return TDVMC_ERR_BAD_CONFIG;
        }
        if (s.bc_n0 + s.bc_n1 != 4) {
            ctx->err = "SPLINE_BC_ORIGIN " + std::to_string(f0) + " with SPLINE_BC_CUTOFF " + std::to_string(f1) +
                       " fixes " + std::to_string(s.bc_n0 + s.bc_n1) + " boundary parameters; N_PARAM knots need 4 " +
                       "(BosonsBulk.cpp:84-91: need " + std::to_string(c.N_PARAM - 4 + s.bc_n0 + s.bc_n1) +
                       " parameters, got " + std::to_string(c.N_PARAM) + ")";
            return TDVMC_ERR_BAD_CONFIG;
        }
        ctx->bc_family[0] = f0;
        ctx->bc_family[1] = f1;
    }
    s.h = s.rc / (s.P_s - 1);
    s.hinv = 1.0 / s.h;
    s.hlast = s.nu ? c.NURBS_GRID[s.P_s - 1] - c.NURBS_GRID[s.P_s - 2] : s.h;
    s.hbar2_2m = c.HBAR2_2M > 0 ? c.HBAR2_2M : 1.0;
    s.trap_omega = c.TRAP_OMEGA;
    std::string pp = c.PAIR_POTENTIAL;
    if (pp.empty()) {
        if (c.SYSTEM_TYPE == "BosonsBulkDamped") pp = "GAUSSIAN";
        else if (c.SYSTEM_TYPE == "BosonsBulk" || c.SYSTEM_TYPE.empty()) pp = "SQUARE_WELL";
        else pp = "NONE";
    }
    if (pp == "NONE") ctx->pot_type = 0;
    else if (pp == "SQUARE_WELL") ctx->pot_type = 1;
    else if (pp == "GAUSSIAN") ctx->pot_type = 2;
    else if (pp == "LJ") ctx->pot_type = 3;
    else {
        ctx->err = "unknown PAIR_POTENTIAL '" + pp + "'";
        return TDVMC_ERR_BAD_CONFIG;
    }
    ctx->pot = c.SYSTEM_PARAMS;
    if (ctx->pot_type != 0 && ctx->pot.size() < 2) {
        ctx->err = "SYSTEM_PARAMS needs the two pair-potential parameters";
        return TDVMC_ERR_BAD_CONFIG;
    }
    s.pot_type = ctx->pot_type;
    s.pot_cutoff = s.periodic ? s.rc : 1e300;
    s.outer_idx = s.off_pair + s.P_s - 1;
    s.move_order = (c.MOVE_ORDER == "SEQUENTIAL") ? 1 : 0;
    s.step = c.MC_STEP;
    resolve_potential(ctx);

    ctx->precision = (c.PRECISION == "MIXED") ? TDVMC_MIXED : TDVMC_FP64;
    ctx->ode_type = c.ODE_SOLVER_TYPE;
    ctx->imag_time = c.IMAGINARY_TIME;
    ctx->precondition = c.USE_PRECONDITIONING;
    ctx->reg_eps = c.REG_EPS > 0 ? c.REG_EPS : 1e-3;
    ctx->solver_type = c.LINEAR_EQUATION_SOLVER_TYPE;
    ctx->normalize_wf = c.USE_NORMALIZE_WF == 1;
    ctx->adjust_params = c.USE_ADJUST_PARAMETERS == 1;
    ctx->accept_check = c.USE_PARAMETER_ACCEPTANCE_CHECK == 1;
    ctx->accept_type = c.PARAMETER_ACCEPTANCE_CHECK_TYPE;
    ctx->dt = c.TIMESTEP;
    ctx->therm_moves = c.MC_NTHERMSTEPS > 0 ? c.MC_NTHERMSTEPS : 0;
    ctx->init_moves = c.MC_NINITIALIZATIONSTEPS > 0 ? c.MC_NINITIALIZATIONSTEPS : 0;
    // samples per walker and stage: SAMPLES_PER_WALKER, else MC_NSTEPS (the
    // reference draws MC_NSTEPS samples per rank and stage, src/TDVMC.cpp:1073-1080,
    // and a walker here is what a rank is there), else 1
    ctx->spw = c.SAMPLES_PER_WALKER > 0 ? c.SAMPLES_PER_WALKER : (c.MC_NSTEPS > 0 ? c.MC_NSTEPS : 1);
    ctx->seed = c.SEED > 0 ? (uint64_t)c.SEED : 1u;

    // parameters: PARAMS_REAL / PARAMS_IMAGINARY of length P, or of the pair block only
    const int P = s.P;
    ctx->uR.assign(P, 0.0);
    ctx->uI.assign(P, 0.0);
    auto fill = [&](const std::vector<double>& src, std::vector<double>& dst) {
        if ((int)src.size() == s.P_s && s.P != s.P_s) {
            for (int k = 0; k < s.P_s; k++) dst[s.off_pair + k] = src[k];
        } else {
            for (int k = 0; k < P && k < (int)src.size(); k++) dst[k] = src[k];
        }
    };
    fill(c.PARAMS_REAL, ctx->uR);
    fill(c.PARAMS_IMAGINARY, ctx->uI);
    ctx->phiR = c.PARAM_PHIR;
    ctx->phiI = c.PARAM_PHII;

    // k-vectors in parameter order; each k whose -k is also in the set is
    // paired with it (the first of the two is the representative).  Stored
    // per representative: n/L (phase in revolutions) and |k|^2.
    auto ns = shortest_lattice_vectors(s.D, s.K);
    std::vector<int> partner(s.K, -1);
    for (int i = 0; i < s.K; i++)
        for (int j = i + 1; j < s.K && partner[i] < 0; j++) {
            bool neg = true;
            for (int a = 0; a < s.D; a++) neg &= ns[j][a] == -ns[i][a];
            if (neg && partner[j] < 0) {
                partner[i] = j;
                partner[j] = i;
            }
        }
    ctx->kmap.clear();
    for (int i = 0; i < s.K; i++) {
        if (partner[i] >= 0 && partner[i] < i) continue;   // represented by its partner
        ctx->kmap.push_back(i);
        ctx->kmap.push_back(partner[i]);
    }
    ctx->mom_cols.clear();
    ctx->mom_outmap.clear();
    for (int k = 0; k < s.off_k; k++) {   // one-body and pair parameters: their own columns
        ctx->mom_cols.push_back(k);
        ctx->mom_outmap.push_back(k);
        ctx->mom_outmap.push_back(-1);
    }
    for (size_t r = 0; r + 1 < ctx->kmap.size(); r += 2) {
        ctx->mom_cols.push_back(s.off_k + ctx->kmap[r]);
        ctx->mom_outmap.push_back(s.off_k + ctx->kmap[r]);
        ctx->mom_outmap.push_back(ctx->kmap[r + 1] >= 0 ? s.off_k + ctx->kmap[r + 1] : -1);
    }
    s.Kh = (int)ctx->kmap.size() / 2;
    ctx->kv_d.assign(4 * (size_t)std::max(s.Kh, 1), 0.0);
    ctx->kv_f.assign(4 * (size_t)std::max(s.Kh, 1), 0.0f);
    for (int k = 0; k < s.Kh; k++) {
        const auto& n = ns[ctx->kmap[2 * k]];
        double n2 = 0;
        for (int a = 0; a < s.D; a++) {
            ctx->kv_d[4 * k + a] = n[a] * s.Linv;
            n2 += (double)n[a] * n[a];
        }
        const double tp = 2.0 * M_PI * s.Linv;
        ctx->kv_d[4 * k + 3] = tp * tp * n2;
    }
    for (size_t i = 0; i < ctx->kv_d.size(); i++) ctx->kv_f[i] = (float)ctx->kv_d[i];
    if (ctx->kmap.empty()) ctx->kmap.assign(2, -1);

    // resource limits of the kernels
    const size_t lds_max = 160 * 1024;
    if (eval_lds_bytes<double>(s, true) > lds_max) {
        ctx->err = "system too large for the evaluation kernel's LDS budget";
        return TDVMC_ERR_BAD_CONFIG;
    }
    if (sweep_lds_bytes<double>(s, 16) > lds_max) {
        ctx->err = "system too large for the sweep kernel's LDS budget (spline table or particle count)";
        return TDVMC_ERR_BAD_CONFIG;
    }
    return TDVMC_OK;
}

int upload_params(tdvmc_ctx* ctx)
{
    const int P = P_of(ctx);
    std::vector<double> packed(2 * P + 2);
    std::copy(ctx->uR.begin(), ctx->uR.end(), packed.begin());
    std::copy(ctx->uI.begin(), ctx->uI.end(), packed.begin() + P);
    packed[2 * P] = ctx->phiR;
    packed[2 * P + 1] = ctx->phiI;
    HIPC(hipMemcpyAsync(ctx->dX, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    // device copies [P + 1]: the parameters, then phi R / phi I
    HIPC(launch_unpack_params(P, ctx->dX, ctx->dUR, ctx->dUI, ctx->stream));
    ctx->tf.dirty = ctx->td.dirty = true;
    ctx->host_params_stale = false;
    return TDVMC_OK;
}

int pull_params(tdvmc_ctx* ctx)
{
    if (!ctx->host_params_stale) return TDVMC_OK;
    const int P = P_of(ctx);
    std::vector<double> packed(2 * P + 2);
    HIPC(hipMemcpyAsync(packed.data(), ctx->dX, packed.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    std::copy(packed.begin(), packed.begin() + P, ctx->uR.begin());
    std::copy(packed.begin() + P, packed.begin() + 2 * P, ctx->uI.begin());
    ctx->phiR = packed[2 * P];
    ctx->phiI = packed[2 * P + 1];
    ctx->host_params_stale = false;
    return TDVMC_OK;
}

template <typename T>
KTables<T> tables_of(tdvmc_ctx* ctx, DevTables<T>& t)
{
    KTables<T> k;
    k.spl = t.spl;
    k.dspl = t.dspl;
    k.kv = t.kv;
    k.alpha = t.alpha;
    k.kag = t.kag;
    k.kmap = ctx->dKmap;
    k.uR = ctx->dUR;
    k.uI = ctx->dUI;
    k.nucell = ctx->dNuCell;
    k.nubeta = ctx->dNuBeta;
    return k;
}

template <typename T>
int ensure_tables(tdvmc_ctx* ctx, DevTables<T>& t, bool force)
{
    if (!t.dirty && !force) return TDVMC_OK;
    HIPC(launch_prep_tables<T>(ctx->sy, ctx->dUR, ctx->dUI, ctx->dKmap, t.kv, t.spl, t.dspl, t.alpha, t.kag,
                               ctx->dNuBeta, ctx->stream));
    t.dirty = false;
    return TDVMC_OK;
}

// Kernel timing with HIP events on the context stream.  A profiled TDVP step
// is captured in segments: each timing point closes the open capture, appends
// it to the step's graph as a child-graph node, appends an event-record node
// and reopens the capture, so every replay of the graph re-records the events.
int seg_begin(tdvmc_ctx* ctx)
{
    HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->cap_active = true;
    return TDVMC_OK;
}

int seg_cut(tdvmc_ctx* ctx)
{
    hipGraph_t g = nullptr;
    ctx->cap_active = false;
    HIPC(hipStreamEndCapture(ctx->stream, &g));
    size_t n = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &n);
    if (e == hipSuccess && n > 0) {
        hipGraphNode_t node;
        e = hipGraphAddChildGraphNode(&node, ctx->seg_top, ctx->seg_last ? &ctx->seg_last : nullptr,
                                      ctx->seg_last ? 1 : 0, g);
        if (e == hipSuccess) ctx->seg_last = node;
    }
    (void)hipGraphDestroy(g);
    HIPC(e);
    return TDVMC_OK;
}

int seg_event(tdvmc_ctx* ctx, hipEvent_t ev)
{
    int rc = seg_cut(ctx);
    if (rc) return rc;
    hipGraphNode_t node;
    HIPC(hipGraphAddEventRecordNode(&node, ctx->seg_top, ctx->seg_last ? &ctx->seg_last : nullptr,
                                    ctx->seg_last ? 1 : 0, ev));
    ctx->seg_last = node;
    return seg_begin(ctx);
}

int prof_begin(tdvmc_ctx* ctx, const char* name, hipEvent_t* a)
{
    if (!ctx->prof) return TDVMC_OK;
    HIPC(hipEventCreate(a));
    (void)name;
    if (ctx->capturing) {
        int rc = seg_event(ctx, *a);
        if (rc) {
            (void)hipEventDestroy(*a);
            return rc;
        }
        return TDVMC_OK;
    }
    HIPC(hipEventRecord(*a, ctx->stream));
    return TDVMC_OK;
}

int prof_end(tdvmc_ctx* ctx, const char* name, hipEvent_t a)
{
    if (!ctx->prof) return TDVMC_OK;
    hipEvent_t b;
    HIPC(hipEventCreate(&b));
    if (ctx->capturing) {
        // the pair is owned by the graph from here on
        ctx->graph_ev.push_back({name, a, b, false});
        return seg_event(ctx, b);
    }
    HIPC(hipEventRecord(b, ctx->stream));
    ctx->pending.push_back({name, a, b, true});
    return TDVMC_OK;
}

int prof_flush(tdvmc_ctx* ctx)
{
    if (ctx->pending.empty()) return TDVMC_OK;
    HIPC(hipStreamSynchronize(ctx->stream));
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, p.a, p.b));
        auto& acc = ctx->prof_acc[p.name];
        acc.ms += ms;
        acc.launches++;
        if (p.owned) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
    }
    ctx->pending.clear();
    return TDVMC_OK;
}

void drop_graph(tdvmc_ctx* ctx)
{
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    if (ctx->graph_src) (void)hipGraphDestroy(ctx->graph_src);
    ctx->graph = nullptr;
    ctx->graph_src = nullptr;
    for (auto& p : ctx->graph_ev) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    ctx->graph_ev.clear();
    ctx->graph_key.clear();
}

int ensure_samples(tdvmc_ctx* ctx, int64_t n)
{
    if (n <= ctx->cap_samples) return TDVMC_OK;
    const int P = P_of(ctx);
    HIPC(dalloc(&ctx->dO64, (size_t)n * P));
    HIPC(dalloc(&ctx->dO32, (size_t)n * P));
    HIPC(dalloc(&ctx->dER, (size_t)n));
    HIPC(dalloc(&ctx->dEI, (size_t)n));
    HIPC(dalloc(&ctx->dExpo, (size_t)n));
    HIPC(dalloc(&ctx->dOther, (size_t)n * 9));
    ctx->cap_samples = n;
    return TDVMC_OK;
}

int ensure_moments(tdvmc_ctx* ctx, int64_t n)
{
    if (n <= ctx->mw_samples) return TDVMC_OK;
    const int P = (int)ctx->mom_cols.size();   // the computed columns (k / -k pairs folded)
    size_t x0, p0, c0, x1, p1, c1;
    moments_work_bytes(n, P, 0, &x0, &p0, &c0);
    moments_work_bytes(n, P, 1, &x1, &p1, &c1);
    size_t xb = std::max(x0, x1), pb = std::max(p0, p1), cb = std::max(c0, c1);
    void* xt = ctx->mw.xt;
    if (xt) (void)hipFree(xt);
    ctx->mw.xt = nullptr;
    HIPC(hipMalloc(&ctx->mw.xt, xb));
    HIPC(dalloc(&ctx->mw.partial, (pb + 7) / 8));
    HIPC(dalloc(&ctx->mw.colsum, (cb + 7) / 8));
    ctx->mw.xt_bytes = xb;
    ctx->mw.partial_bytes = pb;
    ctx->mw.colsum_bytes = cb;
    ctx->mw_samples = n;
    return TDVMC_OK;
}

int resolve_precision(tdvmc_ctx* ctx, int precision, int* out)
{
    if (precision == TDVMC_PRECISION_DEFAULT) precision = ctx->precision;
    if (precision != TDVMC_FP64 && precision != TDVMC_MIXED) {
        ctx->err = "precision must be TDVMC_FP64 or TDVMC_MIXED";
        return TDVMC_ERR_BAD_ARG;
    }
    *out = precision;
    return TDVMC_OK;
}

int enqueue_sweep(tdvmc_ctx* ctx, int prec, int64_t n_moves, uint64_t seed, bool want_rho = false)
{
    if (n_moves <= 0) return TDVMC_OK;
    SweepArgs a;
    a.rho_out = (want_rho && ctx->sy.Kh > 0) ? static_cast<void*>(ctx->dRho) : nullptr;
    a.R = ctx->dR;
    a.W = ctx->W;
    a.n_moves = n_moves;
    a.seed = seed;
    a.epoch = ctx->dEpoch;
    a.acc = ctx->dAcc;
    a.lanes_per_walker = ctx->lanes ? ctx->lanes : ctx->lanes_tuned[prec == TDVMC_MIXED ? 1 : 0];
    hipEvent_t ev{};
    int rc = prof_begin(ctx, "sweep", &ev);
    if (rc) return rc;
    if (prec == TDVMC_MIXED) {
        HIPC(launch_sweep<float>(ctx->sy, tables_of(ctx, ctx->tf), a, ctx->stream));
    } else {
        HIPC(launch_sweep<double>(ctx->sy, tables_of(ctx, ctx->td), a, ctx->stream));
    }
    rc = prof_end(ctx, "sweep", ev);
    if (rc) return rc;
    HIPC(launch_add_u32(ctx->dEpoch, 1u, ctx->stream));
    if (!ctx->in_capture) ctx->trials_total += n_moves * ctx->W;
    return TDVMC_OK;
}

// eval of the context walkers into sample rows [row0, row0 + W); use_rho:
// take rho_k from the sweep that just ran instead of recomputing it
int enqueue_eval(tdvmc_ctx* ctx, int prec, int64_t row0, bool use_rho = false, const double* R = nullptr)
{
    const bool rho = use_rho && ctx->sy.Kh > 0;
    if (!R) R = ctx->dR;
    const int P = P_of(ctx);
    hipEvent_t ev{};
    int rc = prof_begin(ctx, "eval", &ev);
    if (rc) return rc;
    if (prec == TDVMC_MIXED) {
        HIPC((launch_eval<float, float>(ctx->sy, tables_of(ctx, ctx->tf), R, ctx->W, ctx->dExpo + row0,
                                        ctx->dER + row0, ctx->dEI + row0, ctx->dO32 + row0 * P,
                                        ctx->dOther + row0 * 9, nullptr, nullptr, nullptr,
                                        rho ? reinterpret_cast<const float*>(ctx->dRho) : nullptr, ctx->stream)));
    } else {
        HIPC((launch_eval<double, double>(ctx->sy, tables_of(ctx, ctx->td), R, ctx->W, ctx->dExpo + row0,
                                          ctx->dER + row0, ctx->dEI + row0, ctx->dO64 + row0 * P,
                                          ctx->dOther + row0 * 9, nullptr, nullptr, nullptr,
                                          rho ? ctx->dRho : nullptr, ctx->stream)));
    }
    return prof_end(ctx, "eval", ev);
}

int enqueue_moments(tdvmc_ctx* ctx, int prec, int64_t n)
{
    const int P = P_of(ctx);
    hipEvent_t ev{};
    int rc = prof_begin(ctx, "moments", &ev);
    if (rc) return rc;
    // the two slots of a k / -k pair hold the same O: one computed column each
    MomentsWork wk = ctx->mw;
    wk.cols = ctx->dMomCols;
    wk.outmap = ctx->dMomOut;
    wk.P_cols = (int)ctx->mom_cols.size();
    wk.mean_cols = ctx->dMeanCols;
    if (prec == TDVMC_MIXED) {
        HIPC(launch_moments<float>(ctx->dO32, ctx->dER, ctx->dEI, n, P, 1, wk, ctx->dMean, ctx->dS, ctx->dF,
                                   ctx->stream));
    } else {
        HIPC(launch_moments<double>(ctx->dO64, ctx->dER, ctx->dEI, n, P, 0, wk, ctx->dMean, ctx->dS, ctx->dF,
                                    ctx->stream));
    }
    return prof_end(ctx, "moments", ev);
}

int enqueue_solve(tdvmc_ctx* ctx, double* xdot)
{
    SolveArgs a;
    a.P = P_of(ctx);
    a.imag_time = ctx->imag_time;
    a.precondition = ctx->precondition;
    // type 0: S += eps I (0.001 in src/TDVMC.cpp:1737, configurable here);
    // type 1: 0.002 only together with the scaling (:1768-1772)
    a.eps = ctx->solver_type == 1 ? (ctx->precondition ? 0.002 : 0.0) : ctx->reg_eps;
    a.mean_sub = ctx->solver_type == 1;
    a.mean = ctx->dMean;
    a.S = ctx->dS;
    a.F = ctx->dF;
    a.work = ctx->dWork;
    a.xdot = xdot;
    a.status = ctx->dStatus;
    hipEvent_t ev{};
    int rc = prof_begin(ctx, "solve", &ev);
    if (rc) return rc;
    HIPC(launch_solve(a, ctx->stream));
    return prof_end(ctx, "solve", ev);
}

// one sampling stage with the current device tables: moves, measurements, moments
// Grid and wave-vector shells of the pair distribution and S(k)
// (BosonsBulk::InitSystem, src/PhysicalSystems/BosonsBulk.cpp:55, :61-72,
// :124-155): the g(r) grid ends at the last knot ((P_s - 1) rc) / (P_s - 1)
// with GR_BIN_COUNT bins requested and count = (int)(gmax / spacing) kept
// (src/Observables/Grid.cpp:17-49), bin volumes as in
// src/Observables/ObservableVsOnGridWithScaling.cpp:19-42; S(k) on the first
// 50 shells of integer vectors with non-negative components, grouped by
// |n|^2, lexicographic within a shell (the layout of config/kVectors<D>D.json).
constexpr int kObsShells = 50;
constexpr int kObsMaxBins = 8192;

int ensure_obs(tdvmc_ctx* ctx)
{
    auto& o = ctx->obs;
    if (o.ready) return TDVMC_OK;
    if (!ctx->sy.periodic) {
        ctx->err = "g(r) and S(k) need a periodic box (BosonsBulk)";
        return TDVMC_ERR_BAD_CONFIG;
    }
    const int bins = ctx->cfg.GR_BIN_COUNT;
    if (bins <= 0) {
        ctx->err = "GR_BIN_COUNT must be positive for g(r)";
        return TDVMC_ERR_BAD_CONFIG;
    }
    const int D = ctx->sy.D, Ps = ctx->sy.P_s, N = ctx->sy.N;
    o.gmax = ((double)(Ps - 1) * ctx->sy.rc) / ((double)Ps - 1);
    o.spacing = o.gmax / bins;
    o.count = (int)((o.gmax - 0.0) / o.spacing);
    if (o.count <= 0 || o.count > kObsMaxBins) {
        ctx->err = "g(r) grid of " + std::to_string(o.count) + " bins (1.." + std::to_string(kObsMaxBins) + ")";
        return TDVMC_ERR_BAD_CONFIG;
    }
    o.centres.assign(o.count, 0.0);
    {
        double val = 0.0, prev = 0.0;
        for (int i = 0; i < o.count; i++) {
            if (i > 0) o.centres[i - 1] = (val + prev) / 2.0;
            prev = val;
            val += o.spacing;
        }
        o.centres[o.count - 1] = (o.gmax + prev) / 2.0;
    }
    std::vector<double> scaling(o.count);
    for (int i = 0; i < o.count; i++) {
        if (D == 1)
            scaling[i] = o.spacing;
        else if (D == 3)
            scaling[i] = 4.0 * M_PI * std::pow(o.spacing * (i + 1), 3.0) / 3.0;
        else
            scaling[i] = M_PI * std::pow(o.spacing * (i + 1), 2.0);
    }
    if (D != 1)
        for (int i = o.count - 1; i > 0; i--) scaling[i] = scaling[i] - scaling[i - 1];
    const double weight = N > 1 ? 1.0 / ((double)(N - 1)) * D : 0.0;
    o.scale.resize(o.count);
    for (int i = 0; i < o.count; i++) o.scale[i] = weight / scaling[i];

    std::vector<int> vec, off(1, 0);
    o.kgrid.clear();
    o.nmax = 0;
    for (int q = 1; (int)o.kgrid.size() < kObsShells; q++) {
        int lim = 1;
        while (lim * lim <= q) lim++;
        const int l2 = D > 1 ? lim : 1, l3 = D > 2 ? lim : 1;
        int c = 0;
        for (int x = 0; x < lim; x++)
            for (int y = 0; y < l2; y++)
                for (int z = 0; z < l3; z++) {
                    if (x * x + y * y + z * z != q) continue;
                    vec.insert(vec.end(), {x, y, z, 0});
                    o.nmax = std::max(o.nmax, std::max(x, std::max(y, z)));
                    c++;
                }
        if (c == 0) continue;
        off.push_back(off.back() + c);
        o.kgrid.push_back(std::sqrt((double)q) * (2 * M_PI / ctx->sy.L));
    }
    o.nshell = kObsShells;
    o.V = off.back();
    if (o.V > 4 * 256) {
        ctx->err = "too many S(k) wave vectors";
        return TDVMC_ERR_BAD_CONFIG;
    }
    HIPC(dalloc(&o.dVec, vec.size()));
    HIPC(dalloc(&o.dOff, off.size()));
    HIPC(dalloc(&o.dScale, (size_t)o.count));
    HIPC(dalloc(&o.dTot, (size_t)o.count));
    HIPC(dalloc(&o.dSkAcc, (size_t)o.nshell));
    HIPC(hipMemcpyAsync(o.dVec, vec.data(), vec.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(o.dOff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(o.dScale, o.scale.data(), o.count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    o.ready = true;
    return TDVMC_OK;
}

// g(r) counts of W walkers at dR added to obs.dTot, per-walker g(r) into gr_w
// (may be null) and S(k) into sk_w [W][nshell]
int enqueue_obs(tdvmc_ctx* ctx, int prec, const double* dR, int64_t W, double* gr_w, double* sk_w)
{
    auto& o = ctx->obs;
    const int D = ctx->sy.D, N = ctx->sy.N;
    const bool f32 = prec == TDVMC_MIXED;
    const size_t cx = f32 ? 2 * sizeof(float) : 2 * sizeof(double);
    const int chunk = (int)std::min<size_t>(64, std::max<size_t>(1, 24576 / ((size_t)D * (o.nmax + 1) * cx)));
    const size_t lds = obs_lds_bytes(N, D, o.V, chunk, o.nmax, o.count, f32);
    if (lds > 65536) {
        ctx->err = "g(r) / S(k) need " + std::to_string(lds) + " bytes of LDS per walker (at most 65536)";
        return TDVMC_ERR_BAD_ARG;
    }
    ObsArgs a{};
    a.R = dR;
    a.W = W;
    a.N = N;
    a.L = ctx->sy.L;
    a.Lr = 1.0 / ctx->sy.L;
    a.gmax = o.gmax;
    a.spacing = o.spacing;
    a.count = o.count;
    a.scale = o.dScale;
    a.gr_tot = o.dTot;
    a.gr_w = gr_w;
    a.kvec = o.dVec;
    a.shell_off = o.dOff;
    a.V = o.V;
    a.nshell = o.nshell;
    a.nmax = o.nmax;
    a.chunk = chunk;
    a.gr_segs = std::max(1, std::min((N - 1) / 2 + 1, (512 + N - 1) / N));
    a.sk_w = sk_w;
    hipEvent_t ev{};
    int rc = prof_begin(ctx, "observables", &ev);
    if (rc) return rc;
    if (f32) {
        HIPC(launch_obs<float>(a, D, lds, ctx->stream));
    } else {
        HIPC(launch_obs<double>(a, D, lds, ctx->stream));
    }
    return prof_end(ctx, "observables", ev);
}

// zero the g(r) totals and the S(k) accumulator (fill kernels: both may sit in a captured stream)
int clear_obs(tdvmc_ctx* ctx)
{
    auto& o = ctx->obs;
    HIPC(launch_fill_u32(reinterpret_cast<uint32_t*>(o.dTot), 0u, 2 * (int64_t)o.count, ctx->stream));
    HIPC(launch_fill_u32(reinterpret_cast<uint32_t*>(o.dSkAcc), 0u, 2 * (int64_t)o.nshell, ctx->stream));
    return TDVMC_OK;
}

// keep: copy each sample's configurations for later reuse (more than one
// sample per walker; with one, the walkers themselves are the samples)
int enqueue_stage_samples(tdvmc_ctx* ctx, int prec, uint64_t seed, bool keep = false)
{
    int rc = enqueue_sweep(ctx, prec, ctx->init_moves, seed, true);
    if (rc) return rc;
    const size_t per = (size_t)ctx->W * ctx->sy.N * ctx->sy.D;
    for (int s = 0; s < ctx->spw; s++) {
        rc = enqueue_sweep(ctx, prec, ctx->therm_moves, seed, true);
        if (rc) return rc;
        const bool swept = ctx->therm_moves > 0 || (s == 0 && ctx->init_moves > 0);
        rc = enqueue_eval(ctx, prec, (int64_t)s * ctx->W, swept);
        if (rc) return rc;
        if (keep && ctx->spw > 1)
            HIPC(hipMemcpyAsync(ctx->dSamp + s * per, ctx->dR, per * sizeof(double), hipMemcpyDeviceToDevice,
                                ctx->stream));
    }
    return enqueue_moments(ctx, prec, (int64_t)ctx->spw * ctx->W);
}

// UpdateExpectationValuesForGivenSamples (src/TDVMC.cpp:1222-1303): the kept
// samples evaluated at the current parameters, unit weights (the reweighting
// there is commented out), then the moments.  One sample per walker: the
// walkers, with the rho_k their last sweep left when there was one.
int enqueue_reuse_samples(tdvmc_ctx* ctx, int prec)
{
    const size_t per = (size_t)ctx->W * ctx->sy.N * ctx->sy.D;
    for (int s = 0; s < ctx->spw; s++) {
        int rc;
        if (ctx->spw == 1)
            rc = enqueue_eval(ctx, prec, 0, ctx->therm_moves > 0 || ctx->init_moves > 0);
        else
            rc = enqueue_eval(ctx, prec, (int64_t)s * ctx->W, false, ctx->dSamp + s * per);
        if (rc) return rc;
    }
    return enqueue_moments(ctx, prec, (int64_t)ctx->spw * ctx->W);
}

int enqueue_prep(tdvmc_ctx* ctx, int prec)
{
    if (prec == TDVMC_MIXED) return ensure_tables(ctx, ctx->tf, true);
    return ensure_tables(ctx, ctx->td, true);
}

// f(x): parameters x (packed, device) -> xdot.  reuse: evaluate the kept
// samples instead of sampling; keep: keep this stage's samples for reuse
int enqueue_rhs(tdvmc_ctx* ctx, int prec, const double* x, double* xdot, bool reuse = false, bool keep = false)
{
    const int P = P_of(ctx);
    HIPC(launch_unpack_params(P, x, ctx->dUR, ctx->dUI, ctx->stream));
    int rc = enqueue_prep(ctx, prec);
    if (rc) return rc;
    rc = reuse ? enqueue_reuse_samples(ctx, prec) : enqueue_stage_samples(ctx, prec, ctx->seed, keep);
    if (rc) return rc;
    return enqueue_solve(ctx, xdot);
}

int enqueue_tdvp_step(tdvmc_ctx* ctx, int prec, int ode, double dt)
{
    const int P = P_of(ctx);
    int rc;
    // the solves of this step add their non-positive pivots to dStatus
    HIPC(launch_fill_u32(reinterpret_cast<uint32_t*>(ctx->dStatus), 0u, 1, ctx->stream));
    if (ode == 0) {
        rc = enqueue_rhs(ctx, prec, ctx->dX, ctx->dK[0]);
        if (rc) return rc;
        HIPC(launch_axpy_params(P, ctx->dX, ctx->dK[0], dt, ctx->dX, ctx->stream));
    } else if (ode == 2) {
        rc = enqueue_rhs(ctx, prec, ctx->dX, ctx->dK[0]);
        if (rc) return rc;
        HIPC(launch_axpy_params(P, ctx->dX, ctx->dK[0], dt / 2.0, ctx->dXtmp, ctx->stream));
        rc = enqueue_rhs(ctx, prec, ctx->dXtmp, ctx->dK[1]);
        if (rc) return rc;
        HIPC(launch_axpy_params(P, ctx->dX, ctx->dK[1], dt / 2.0, ctx->dXtmp, ctx->stream));
        rc = enqueue_rhs(ctx, prec, ctx->dXtmp, ctx->dK[2]);
        if (rc) return rc;
        HIPC(launch_axpy_params(P, ctx->dX, ctx->dK[2], dt, ctx->dXtmp, ctx->stream));
        rc = enqueue_rhs(ctx, prec, ctx->dXtmp, ctx->dK[3]);
        if (rc) return rc;
        HIPC(launch_rk4_combine(P, ctx->dX, ctx->dK[0], ctx->dK[1], ctx->dK[2], ctx->dK[3], dt, ctx->stream));
    } else if (ode == 3) {
        // CalculateNextParametersRK4ReuseSamples (src/TDVMC.cpp:2037-2103):
        // fresh samples for k1, the same samples for k2..k4
        rc = enqueue_rhs(ctx, prec, ctx->dX, ctx->dK[0], false, true);
        if (rc) return rc;
        HIPC(launch_axpy_params(P, ctx->dX, ctx->dK[0], dt / 2.0, ctx->dXtmp, ctx->stream));
        rc = enqueue_rhs(ctx, prec, ctx->dXtmp, ctx->dK[1], true);
        if (rc) return rc;
        HIPC(launch_axpy_params(P, ctx->dX, ctx->dK[1], dt / 2.0, ctx->dXtmp, ctx->stream));
        rc = enqueue_rhs(ctx, prec, ctx->dXtmp, ctx->dK[2], true);
        if (rc) return rc;
        HIPC(launch_axpy_params(P, ctx->dX, ctx->dK[2], dt, ctx->dXtmp, ctx->stream));
        rc = enqueue_rhs(ctx, prec, ctx->dXtmp, ctx->dK[3], true);
        if (rc) return rc;
        HIPC(launch_rk4_combine(P, ctx->dX, ctx->dK[0], ctx->dK[1], ctx->dK[2], ctx->dK[3], dt, ctx->stream));
    } else if (ode == 1 || ode == 5) {
        // CalculateNextParametersPC (src/TDVMC.cpp:1855-1910): predictor
        // x + dt k1, one corrector x + dt/2 (k1 + k2) on fresh samples;
        // ...PCReuseSamples (:1912-1967): six correctors on the samples of k1
        const bool reuse = ode == 5;
        const int passes = reuse ? 6 : 1;
        rc = enqueue_rhs(ctx, prec, ctx->dX, ctx->dK[0], false, reuse);
        if (rc) return rc;
        HIPC(launch_axpy_params(P, ctx->dX, ctx->dK[0], dt, ctx->dXtmp, ctx->stream));
        for (int c = 0; c < passes; c++) {
            rc = enqueue_rhs(ctx, prec, ctx->dXtmp, ctx->dK[1], reuse);
            if (rc) return rc;
            HIPC(launch_pc_combine(P, ctx->dX, ctx->dK[0], ctx->dK[1], dt, ctx->dXtmp, ctx->stream));
        }
        HIPC(hipMemcpyAsync(ctx->dX, ctx->dXtmp, (2 * (size_t)P + 2) * sizeof(double), hipMemcpyDeviceToDevice,
                            ctx->stream));
    } else {
        ctx->err = "ODE_SOLVER_TYPE " + std::to_string(ode) +
                   " is not implemented (0 Euler, 1 predictor-corrector, 2 RK4, 3 RK4 reusing samples, "
                   "5 predictor-corrector reusing samples)";
        return TDVMC_ERR_BAD_ARG;
    }
    // src/TDVMC.cpp:3757-3775: phi R := -(exponent of the root rank's walker,
    // here walker 0 at the last sample), then the gauge shift
    if (ctx->normalize_wf || ctx->adjust_params)
        HIPC(launch_post_step(P, ctx->dX, ctx->dExpo + (int64_t)(ctx->spw - 1) * ctx->W, ctx->normalize_wf,
                              ctx->adjust_params, ctx->sy.off_pair, ctx->sy.P_s, ctx->stream));
    HIPC(launch_unpack_params(P, ctx->dX, ctx->dUR, ctx->dUI, ctx->stream));
    return TDVMC_OK;
}

// AcceptNewParams (src/TDVMC.cpp:2585-2640): type 0 rejects only a step whose
// Cholesky met a non-positive pivot; type 1 also non-finite parameters;
// type 2 also relative changes above 0.1.  The checks on the parameters apply
// from the second step on (:2599, :2613).
bool accept_step(const tdvmc_ctx* ctx, int n_bad, const std::vector<double>& x, const std::vector<double>& x0)
{
    if (n_bad > 0) return false;
    if (ctx->accept_type == 0 || ctx->step < 1) return true;
    const int P = ctx->sy.P;
    for (int p = 0; p < P; p++) {
        if (!std::isfinite(x[p])) return false;
        if (ctx->accept_type == 2 && std::fabs(1.0 - x[p] / x0[p]) > 0.1) return false;
    }
    return true;
}

int fill_step_data(tdvmc_ctx* ctx, int64_t n, tdvmc_step_data* out)
{
    if (!out) return TDVMC_OK;
    const int P = P_of(ctx);
    std::vector<double> mean(P + 2), S((size_t)P * P), F(2 * (size_t)P);
    HIPC(hipMemcpyAsync(mean.data(), ctx->dMean, mean.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipMemcpyAsync(S.data(), ctx->dS, S.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipMemcpyAsync(F.data(), ctx->dF, F.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> other;
    if (out->otherExpectationValues) {
        other.resize((size_t)n * 9);
        HIPC(hipMemcpyAsync(other.data(), ctx->dOther, other.size() * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    }
    HIPC(hipStreamSynchronize(ctx->stream));
    out->time = ctx->time;
    out->step = ctx->step;
    out->n_samples = n;
    out->localEnergyR = mean[P];
    out->localEnergyI = mean[P + 1];
    for (int i = 0; i < P; i++) {
        if (out->localOperators) out->localOperators[i] = mean[i];
        if (out->FR) out->FR[i] = F[i];
        if (out->FI) out->FI[i] = F[P + i];
        if (out->localOperatorlocalEnergyR) out->localOperatorlocalEnergyR[i] = F[i] + mean[i] * mean[P];
        if (out->localOperatorlocalEnergyI) out->localOperatorlocalEnergyI[i] = F[P + i] + mean[i] * mean[P + 1];
        for (int j = 0; j < P; j++) {
            const double s = S[(size_t)i * P + j];
            if (out->S) out->S[(size_t)i * P + j] = s;
            if (out->localOperatorsMatrix) out->localOperatorsMatrix[(size_t)i * P + j] = s + mean[i] * mean[j];
        }
    }
    if (out->otherExpectationValues) {
        for (int k = 0; k < 9; k++) {
            double s = 0.0;
            for (int64_t w = 0; w < n; w++) s += other[(size_t)w * 9 + k];
            out->otherExpectationValues[k] = s / (double)n;
        }
    }
    return TDVMC_OK;
}

// The device address of this device's copy of `scratch`, looked up once per
// device with that device current (the context's device at creation).
int init_scratch(tdvmc_ctx* ctx)
{
    if (ctx->device < 0 || ctx->device >= kMaxDevices) {
        ctx->err = "device index " + std::to_string(ctx->device) + " beyond the scratch table";
        return TDVMC_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[ctx->device]) return TDVMC_OK;
    void* p = nullptr;
    HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[ctx->device] = static_cast<double*>(p);
    return TDVMC_OK;
}

// Device-pointer entry points run on the context stream, ordered against the
// caller's stream st: the context stream first waits for what st holds
// (join_in), and st then waits for the context's work (join_out); no host
// synchronisation.  A null st (or the context stream itself) needs neither.
int join_in(tdvmc_ctx* ctx, hipStream_t st)
{
    if (!st || st == ctx->stream) return TDVMC_OK;
    if (!ctx->ev_join) HIPC(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    HIPC(hipEventRecord(ctx->ev_join, st));
    HIPC(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    return TDVMC_OK;
}

int join_out(tdvmc_ctx* ctx, hipStream_t st)
{
    if (!st || st == ctx->stream) return TDVMC_OK;
    HIPC(hipEventRecord(ctx->ev_join, ctx->stream));
    HIPC(hipStreamWaitEvent(st, ctx->ev_join, 0));
    return TDVMC_OK;
}

// mean [P+2], S [P][P], F [2][P] of the last moments into caller device buffers
int copy_moments_dev(tdvmc_ctx* ctx, double* mean_dev, double* S_dev, double* F_dev)
{
    const size_t P = (size_t)P_of(ctx);
    if (mean_dev)
        HIPC(hipMemcpyAsync(mean_dev, ctx->dMean, (P + 2) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if (S_dev) HIPC(hipMemcpyAsync(S_dev, ctx->dS, P * P * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if (F_dev) HIPC(hipMemcpyAsync(F_dev, ctx->dF, 2 * P * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return TDVMC_OK;
}

int alloc_ctx(tdvmc_ctx* ctx)
{
    const int P = P_of(ctx);
    HIPC(hipGetDevice(&ctx->device));
    if (int rc = init_scratch(ctx)) return rc;
    HIPC(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIPC(dalloc(&ctx->dUR, P + 1));
    HIPC(dalloc(&ctx->dUI, P + 1));
    HIPC(dalloc(&ctx->dX, 2 * P + 2));
    HIPC(dalloc(&ctx->dXtmp, 2 * P + 2));
    HIPC(dalloc(&ctx->dXprev, 2 * P + 2));
    for (int i = 0; i < 4; i++) HIPC(dalloc(&ctx->dK[i], 2 * P + 2));
    HIPC(dalloc(&ctx->dEpoch, 1));
    HIPC(launch_fill_u32(ctx->dEpoch, 0u, 1, ctx->stream));
    const int Ps = ctx->sy.P_s, K = std::max(ctx->sy.Kh, 1);
    HIPC(dalloc(&ctx->dMomCols, ctx->mom_cols.size()));
    HIPC(dalloc(&ctx->dMomOut, ctx->mom_outmap.size()));
    HIPC(dalloc(&ctx->dMeanCols, ctx->mom_cols.size() + 2));
    HIPC(hipMemcpyAsync(ctx->dMomCols, ctx->mom_cols.data(), ctx->mom_cols.size() * sizeof(int),
                        hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(ctx->dMomOut, ctx->mom_outmap.data(), ctx->mom_outmap.size() * sizeof(int),
                        hipMemcpyHostToDevice, ctx->stream));
    HIPC(dalloc(&ctx->dKmap, ctx->kmap.size()));
    HIPC(hipMemcpyAsync(ctx->dKmap, ctx->kmap.data(), ctx->kmap.size() * sizeof(int), hipMemcpyHostToDevice,
                        ctx->stream));
    HIPC(dalloc(&ctx->tf.spl, (size_t)Ps * 8));
    if (ctx->sy.nu) {
        HIPC(dalloc(&ctx->dNuCell, ctx->knots.cells.size()));
        HIPC(dalloc(&ctx->dNuBeta, ctx->knots.beta.size()));
        HIPC(hipMemcpyAsync(ctx->dNuCell, ctx->knots.cells.data(), ctx->knots.cells.size() * sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream));
        HIPC(hipMemcpyAsync(ctx->dNuBeta, ctx->knots.beta.data(), ctx->knots.beta.size() * sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream));
    }
    HIPC(dalloc(&ctx->tf.dspl, (size_t)Ps * kDsplStride));
    HIPC(dalloc(&ctx->td.dspl, (size_t)Ps * kDsplStride));
    HIPC(dalloc(&ctx->tf.kv, 4 * (size_t)K));
    HIPC(dalloc(&ctx->tf.alpha, 2 * (size_t)K));
    HIPC(dalloc(&ctx->tf.kag, 8 * (size_t)K));
    HIPC(dalloc(&ctx->td.spl, (size_t)Ps * 8));
    HIPC(dalloc(&ctx->td.kv, 4 * (size_t)K));
    HIPC(dalloc(&ctx->td.alpha, 2 * (size_t)K));
    HIPC(dalloc(&ctx->td.kag, 8 * (size_t)K));
    HIPC(hipMemcpyAsync(ctx->tf.kv, ctx->kv_f.data(), ctx->kv_f.size() * sizeof(float), hipMemcpyHostToDevice,
                        ctx->stream));
    HIPC(hipMemcpyAsync(ctx->td.kv, ctx->kv_d.data(), ctx->kv_d.size() * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
    HIPC(dalloc(&ctx->dMean, P + 2));
    HIPC(dalloc(&ctx->dS, (size_t)P * P));
    HIPC(dalloc(&ctx->dF, 2 * (size_t)P));
    HIPC(dalloc(&ctx->dWork, (size_t)P * P + 4 * (size_t)P));
    HIPC(dalloc(&ctx->dStatus, 1));
    int rc = upload_params(ctx);
    if (rc) return rc;
    HIPC(hipStreamSynchronize(ctx->stream));
    return TDVMC_OK;
}

void release(tdvmc_ctx* ctx)
{
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    dfree(ctx->obs.dVec);
    dfree(ctx->obs.dOff);
    dfree(ctx->obs.dScale);
    dfree(ctx->obs.dTot);
    dfree(ctx->obs.dSkAcc);
    dfree(ctx->obs.dSkW);
    for (auto& p : ctx->pending) {
        if (p.owned) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
    }
    ctx->pending.clear();
    drop_graph(ctx);
    dfree(ctx->dR);
    dfree(ctx->dSamp);
    dfree(ctx->dRho);
    dfree(ctx->dAcc);
    dfree(ctx->dEpoch);
    dfree(ctx->dUR);
    dfree(ctx->dUI);
    dfree(ctx->dX);
    dfree(ctx->dXtmp);
    dfree(ctx->dXprev);
    for (auto& k : ctx->dK) dfree(k);
    dfree(ctx->tf.spl);
    dfree(ctx->dNuCell);
    dfree(ctx->dNuBeta);
    dfree(ctx->dMomCols);
    dfree(ctx->dMomOut);
    dfree(ctx->dMeanCols);
    dfree(ctx->dKmap);
    dfree(ctx->tf.dspl);
    dfree(ctx->td.dspl);
    dfree(ctx->tf.kv);
    dfree(ctx->tf.alpha);
    dfree(ctx->tf.kag);
    dfree(ctx->td.spl);
    dfree(ctx->td.kv);
    dfree(ctx->td.alpha);
    dfree(ctx->td.kag);
    dfree(ctx->dO64);
    dfree(ctx->dO32);
    dfree(ctx->dER);
    dfree(ctx->dEI);
    dfree(ctx->dExpo);
    dfree(ctx->dOther);
    dfree(ctx->dMean);
    dfree(ctx->dS);
    dfree(ctx->dF);
    if (ctx->mw.xt) (void)hipFree(ctx->mw.xt);
    dfree(ctx->mw.partial);
    dfree(ctx->mw.colsum);
    dfree(ctx->dWork);
    dfree(ctx->dStatus);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

}  // namespace


// ---------------------------------------------------------------------------
// Batches of independent systems (each a full context for its own buffers and
// tables) sampled, measured and advanced together: one sweep launch and one
// measurement launch for all systems, the moments in two launches (chunks of
// walkers of every system, then a reduction per system), one solve launch
// with a workgroup per system, the RK arithmetic on the concatenated
// parameters.
struct tdvmc_batch {
    std::vector<tdvmc_ctx*> sys;
    std::string err;
    hipStream_t stream = nullptr;
    int D = 0, periodic = 0, kh_max = 0, P_max = 0, N_max = 0;
    int lanes = 0;                    // lanes per walker of the batched sweep (0: default)
    std::vector<int> eval_wpw;        // per system: wavefronts per walker of the measurement (0: default)
    std::vector<int> sweep_lanes;     // per system: lanes per walker of the sweep (0: default; no k-space terms)
    int tune_gen = 0;                 // changes whenever eval_wpw does
    std::vector<int64_t> xoff;
    int64_t xtotal = 0;
    uint64_t seed = 1;
    KSys* dSys = nullptr;
    KTables<float>* dTabF = nullptr;
    KTables<double>* dTabD = nullptr;
    BatchParams<float>* dBpF = nullptr;
    BatchParams<double>* dBpD = nullptr;
    int64_t* dSweepStart = nullptr;
    int64_t* dEvalStart = nullptr;
    SweepArgs* dSweepArgs = nullptr;
    EvalBatchItem<float>* dEvalF = nullptr;
    EvalBatchItem<double>* dEvalD = nullptr;
    MomentsItem* dMom = nullptr;
    double* dMomPart = nullptr;       // partial sums of the batched moments
    int64_t mom_part_cap = 0;         // doubles allocated at dMomPart
    std::vector<int64_t> desc_W;      // walker count of each system the descriptors and the graph were built for
    int mom_nchunk = 0;
    size_t mom_lds = 0;
    SolveArgs* dSolve[4] = {nullptr, nullptr, nullptr, nullptr};
    double* dX = nullptr;
    double* dXtmp = nullptr;
    double* dK[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t* dEpoch = nullptr;
    int64_t sweep_blocks = 0, eval_blocks = 0;
    size_t sweep_lds = 0, eval_lds_f = 0, eval_lds_d = 0;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;
    std::string graph_key;
    std::string desc_key;             // configuration the descriptor arrays hold
};

namespace {

#define BHIPC(expr)                                                                 \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            b->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
            return TDVMC_ERR_HIP;                                                   \
        }                                                                           \
    } while (0)

// Descriptor arrays are allocated once per batch (their size is fixed by the
// number of systems) and overwritten in place, so a captured step graph
// never refers to freed memory.
template <typename T>
int upload_array(tdvmc_batch* b, T** dst, const std::vector<T>& src)
{
    if (!*dst) BHIPC(dalloc(dst, src.size()));
    BHIPC(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, b->stream));
    return TDVMC_OK;
}

// Lanes per walker of the batched sweep shared by every system, or 0 when
// each system has its own (batches without k-space terms, no override).
int batch_lanes(const tdvmc_batch* b)
{
    if (b->lanes) return b->lanes;
    if (b->kh_max > 0) return b->N_max <= 48 ? 16 : sweep_default_lanes(b->N_max, b->kh_max);
    return 0;
}

// lanes per walker of system s: shared, tuned, or the smallest of 8, 16, 32
// that covers N (32 above)
int batch_sys_lanes(const tdvmc_batch* b, int s)
{
    const int u = batch_lanes(b);
    if (u) return u;
    if (s < (int)b->sweep_lanes.size() && b->sweep_lanes[s]) return b->sweep_lanes[s];
    const int N = b->sys[s]->sy.N;
    return N <= 8 ? 8 : (N <= 16 ? 16 : 32);
}

uint64_t batch_system_seed(uint64_t seed, int s) { return seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(s + 1)); }

// wavefronts per walker of system s's measurement: tuned or default, raised
// until the walkers of one workgroup fit the LDS
int batch_eval_wpw(const tdvmc_batch* b, int s)
{
    const tdvmc_ctx* c = b->sys[s];
    int wpw = (s < (int)b->eval_wpw.size() && b->eval_wpw[s]) ? b->eval_wpw[s] : eval_default_wpw(c->sy.N);
    auto slot = [&](int v) {
        return std::max(eval_lds_bytes<float>(c->sy, false, v), eval_lds_bytes<double>(c->sy, false, v));
    };
    while (wpw < 4 && (size_t)(4 / wpw) * slot(wpw) > 160 * 1024) wpw *= 2;
    return wpw;
}

void batch_drop_graph(tdvmc_batch* b);

// Descriptor arrays of a batched sweep with n_moves per system (<= 0: each
// system's MC_NTHERMSTEPS) and of the measurement, moments and solves.
int batch_upload(tdvmc_batch* b, int64_t n_moves, uint64_t seed, bool want_rho)
{
    // the members' generations cover what the descriptors copy from them: the
    // walker count and the buffers sized by it, and the pair potential in force
    uint64_t h = 1469598103934665603ull;
    for (const auto* c : b->sys) h = (h ^ c->gen) * 1099511628211ull;
    char key[160];
    std::snprintf(key, sizeof(key), "%lld/%llu/%d/%d/%d/%llx", (long long)n_moves, (unsigned long long)seed,
                  want_rho ? 1 : 0, b->lanes, b->tune_gen, (unsigned long long)h);
    if (b->desc_key == key) return TDVMC_OK;
    const int n = (int)b->sys.size();
    // a changed walker count changes the launch shapes baked into a captured
    // step and the size of the moments' partial sums
    std::vector<int64_t> Ws(n);
    for (int s = 0; s < n; s++) Ws[s] = b->sys[s]->W;
    if (Ws != b->desc_W) batch_drop_graph(b);
    std::vector<KSys> sy(n);
    std::vector<KTables<float>> tf(n);
    std::vector<KTables<double>> td(n);
    std::vector<BatchParams<float>> bpf(n);
    std::vector<BatchParams<double>> bpd(n);
    std::vector<int64_t> sstart(n), estart(n);
    std::vector<SweepArgs> sa(n);
    std::vector<EvalBatchItem<float>> ef(n);
    std::vector<EvalBatchItem<double>> ed(n);
    std::vector<MomentsItem> mi(n);
    std::vector<SolveArgs> sv[4];
    int64_t sb = 0, eb = 0, part_len = 0;
    size_t slds = 0, elf = 0, eld = 0;
    for (int s = 0; s < n; s++) {
        tdvmc_ctx* c = b->sys[s];
        sy[s] = c->sy;
        tf[s] = tables_of(c, c->tf);
        td[s] = tables_of(c, c->td);
        bpf[s] = BatchParams<float>{c->sy, c->dUR, c->dUI, c->dKmap, c->tf.kv, c->tf.spl, c->tf.dspl, c->tf.alpha,
                                    c->tf.kag, c->dX,
                                    b->xoff[s]};
        bpd[s] = BatchParams<double>{c->sy, c->dUR, c->dUI, c->dKmap, c->td.kv, c->td.spl, c->td.dspl, c->td.alpha,
                                     c->td.kag, c->dX,
                                     b->xoff[s]};
        const int lanes = batch_sys_lanes(b, s);
        const int wpb = sweep_walkers_per_block(lanes);
        sstart[s] = sb;
        sb += (c->W + wpb - 1) / wpb;
        const int wpw = batch_eval_wpw(b, s);
        const size_t slot_f = eval_lds_bytes<float>(c->sy, false, wpw);
        const size_t slot_d = eval_lds_bytes<double>(c->sy, false, wpw);
        estart[s] = eb;
        eb += (c->W + 4 / wpw - 1) / (4 / wpw);
        slds = std::max(slds, std::max(sweep_lds_bytes<float>(c->sy, lanes), sweep_lds_bytes<double>(c->sy, lanes)));
        elf = std::max(elf, (size_t)(4 / wpw) * slot_f);
        eld = std::max(eld, (size_t)(4 / wpw) * slot_d);
        SweepArgs& a = sa[s];
        a.R = c->dR;
        a.W = c->W;
        a.n_moves = n_moves > 0 ? n_moves : c->therm_moves;
        a.seed = batch_system_seed(seed, s);
        a.epoch = b->dEpoch;
        a.acc = c->dAcc;
        a.lanes_per_walker = lanes;
        a.rho_out = (want_rho && c->sy.Kh > 0) ? static_cast<void*>(c->dRho) : nullptr;
        const bool rho = want_rho && c->sy.Kh > 0;
        ef[s] = EvalBatchItem<float>{c->dR, c->W, c->dExpo, c->dER, c->dEI, c->dO64, c->dOther, nullptr,
                                     rho ? reinterpret_cast<const float*>(c->dRho) : nullptr, wpw, (int64_t)slot_f};
        ed[s] = EvalBatchItem<double>{c->dR, c->W, c->dExpo, c->dER, c->dEI, c->dO64, c->dOther, nullptr,
                                      rho ? c->dRho : nullptr, wpw, (int64_t)slot_d};
        {
            const int cw = moments_chunk_walkers(c->sy.P);
            const int nch = (int)((c->W + cw - 1) / cw);
            mi[s] = MomentsItem{c->sy.P, c->W, c->dO64, c->dER, c->dEI, c->dMean, c->dS, c->dF,
                                reinterpret_cast<double*>((intptr_t)part_len), cw, nch};
            part_len += (int64_t)nch * moments_part_len(c->sy.P);
        }
        for (int k = 0; k < 4; k++) {
            SolveArgs v;
            v.P = c->sy.P;
            v.imag_time = c->imag_time;
            v.precondition = c->precondition;
            v.eps = c->solver_type == 1 ? (c->precondition ? 0.002 : 0.0) : c->reg_eps;
            v.mean = c->dMean;
            v.S = c->dS;
            v.F = c->dF;
            v.work = c->dWork;
            v.xdot = b->dK[k] + b->xoff[s];
            v.status = nullptr;
            v.mean_sub = c->solver_type == 1;
            sv[k].push_back(v);
        }
    }
    // partial-sum scratch of the moments: sized by the systems' walker counts
    // (no captured graph refers to it when it grows: the graph was dropped above)
    if (part_len > b->mom_part_cap) {
        BHIPC(dalloc(&b->dMomPart, (size_t)part_len));
        b->mom_part_cap = part_len;
    }
    b->mom_nchunk = 0;
    b->mom_lds = 0;
    for (auto& m : mi) {
        m.part = b->dMomPart + (intptr_t)m.part;
        b->mom_nchunk = std::max(b->mom_nchunk, m.nchunk);
        b->mom_lds = std::max(b->mom_lds, (size_t)(m.P + 2) * (1 + (size_t)m.cw) * sizeof(double));
    }
    b->sweep_blocks = sb;
    b->eval_blocks = eb;
    b->sweep_lds = slds;
    b->eval_lds_f = elf;
    b->eval_lds_d = eld;
    int rc;
    if ((rc = upload_array(b, &b->dSys, sy)) || (rc = upload_array(b, &b->dTabF, tf)) ||
        (rc = upload_array(b, &b->dTabD, td)) || (rc = upload_array(b, &b->dBpF, bpf)) ||
        (rc = upload_array(b, &b->dBpD, bpd)) || (rc = upload_array(b, &b->dSweepStart, sstart)) ||
        (rc = upload_array(b, &b->dEvalStart, estart)) || (rc = upload_array(b, &b->dSweepArgs, sa)) ||
        (rc = upload_array(b, &b->dEvalF, ef)) || (rc = upload_array(b, &b->dEvalD, ed)) ||
        (rc = upload_array(b, &b->dMom, mi)))
        return rc;
    for (int k = 0; k < 4; k++)
        if ((rc = upload_array(b, &b->dSolve[k], sv[k]))) return rc;
    BHIPC(hipStreamSynchronize(b->stream));
    b->desc_key = key;
    b->desc_W = Ws;
    return TDVMC_OK;
}

int batch_sweep_launch(tdvmc_batch* b, int prec)
{
    const int n = (int)b->sys.size();
    const int lanes = batch_lanes(b);
    if (prec == TDVMC_MIXED) {
        SweepBatch<float> bt{n, b->dSweepStart, b->dSys, b->dTabF, b->dSweepArgs};
        BHIPC(launch_sweep_batch<float>(bt, b->D, b->periodic, lanes, b->kh_max, b->sweep_blocks, b->sweep_lds,
                                        b->stream));
    } else {
        SweepBatch<double> bt{n, b->dSweepStart, b->dSys, b->dTabD, b->dSweepArgs};
        BHIPC(launch_sweep_batch<double>(bt, b->D, b->periodic, lanes, b->kh_max, b->sweep_blocks, b->sweep_lds,
                                         b->stream));
    }
    BHIPC(launch_add_u32(b->dEpoch, 1u, b->stream));
    return TDVMC_OK;
}

int batch_measure_launch(tdvmc_batch* b, int prec)
{
    const int n = (int)b->sys.size();
    if (prec == TDVMC_MIXED) {
        EvalBatch<float> bt{n, b->dEvalStart, b->dSys, b->dTabF, b->dEvalF};
        BHIPC(launch_eval_batch<float>(bt, b->D, b->periodic, b->eval_blocks, b->eval_lds_f, b->stream));
    } else {
        EvalBatch<double> bt{n, b->dEvalStart, b->dSys, b->dTabD, b->dEvalD};
        BHIPC(launch_eval_batch<double>(bt, b->D, b->periodic, b->eval_blocks, b->eval_lds_d, b->stream));
    }
    BHIPC(launch_moments_batch(b->dMom, n, b->mom_nchunk, b->mom_lds, b->P_max, b->stream));
    return TDVMC_OK;
}

int batch_prep(tdvmc_batch* b, int prec)
{
    const int n = (int)b->sys.size();
    if (prec == TDVMC_MIXED)
        BHIPC(launch_prep_tables_batch<float>(b->dBpF, n, b->stream));
    else
        BHIPC(launch_prep_tables_batch<double>(b->dBpD, n, b->stream));
    for (auto* c : b->sys) c->tf.dirty = c->td.dirty = true;
    return TDVMC_OK;
}

// x (concatenated) -> xdot for stage k: parameters, tables, sweeps, measurement, moments, solves
// reuse: measure the walkers of the previous stage again instead of sweeping
// (batches take one sample per walker, so the walkers are the samples)
int batch_rhs(tdvmc_batch* b, int prec, const double* x, int k, bool reuse = false)
{
    const int n = (int)b->sys.size();
    BHIPC(launch_unpack_batch<double>(b->dBpD, n, x, b->stream));
    int rc = batch_prep(b, prec);
    if (rc) return rc;
    if (!reuse) {
        rc = batch_sweep_launch(b, prec);
        if (rc) return rc;
    }
    rc = batch_measure_launch(b, prec);
    if (rc) return rc;
    BHIPC(launch_solve_batch(b->dSolve[k], n, b->P_max, b->stream));
    return TDVMC_OK;
}

int batch_step_enqueue(tdvmc_batch* b, int prec, int ode, double dt)
{
    const int n = (int)b->sys.size();
    const int P2 = (int)((b->xtotal - 2) / 2);   // axpy / rk4 run over 2 P2 + 2 = xtotal entries
    BHIPC(launch_gather_batch<double>(b->dBpD, n, b->dX, 1, b->stream));
    int rc;
    if (ode == 0) {
        if ((rc = batch_rhs(b, prec, b->dX, 0))) return rc;
        BHIPC(launch_axpy_params(P2, b->dX, b->dK[0], dt, b->dX, b->stream));
    } else if (ode == 1 || ode == 5) {
        // predictor-corrector (src/TDVMC.cpp:1855-1910), reusing the samples of
        // k1 with six correctors for type 5 (:1912-1967)
        const bool reuse = ode == 5;
        if ((rc = batch_rhs(b, prec, b->dX, 0))) return rc;
        BHIPC(launch_axpy_params(P2, b->dX, b->dK[0], dt, b->dXtmp, b->stream));
        for (int c = 0; c < (reuse ? 6 : 1); c++) {
            if ((rc = batch_rhs(b, prec, b->dXtmp, 1, reuse))) return rc;
            BHIPC(launch_pc_combine(P2, b->dX, b->dK[0], b->dK[1], dt, b->dXtmp, b->stream));
        }
        BHIPC(hipMemcpyAsync(b->dX, b->dXtmp, (size_t)b->xtotal * sizeof(double), hipMemcpyDeviceToDevice,
                             b->stream));
    } else {
        // RK4 (type 2) or RK4 reusing the samples of k1 (type 3, src/TDVMC.cpp:2037-2103)
        const bool reuse = ode == 3;
        if ((rc = batch_rhs(b, prec, b->dX, 0))) return rc;
        BHIPC(launch_axpy_params(P2, b->dX, b->dK[0], dt / 2.0, b->dXtmp, b->stream));
        if ((rc = batch_rhs(b, prec, b->dXtmp, 1, reuse))) return rc;
        BHIPC(launch_axpy_params(P2, b->dX, b->dK[1], dt / 2.0, b->dXtmp, b->stream));
        if ((rc = batch_rhs(b, prec, b->dXtmp, 2, reuse))) return rc;
        BHIPC(launch_axpy_params(P2, b->dX, b->dK[2], dt, b->dXtmp, b->stream));
        if ((rc = batch_rhs(b, prec, b->dXtmp, 3, reuse))) return rc;
        BHIPC(launch_rk4_combine(P2, b->dX, b->dK[0], b->dK[1], b->dK[2], b->dK[3], dt, b->stream));
    }
    BHIPC(launch_gather_batch<double>(b->dBpD, n, b->dX, 0, b->stream));
    BHIPC(launch_unpack_batch<double>(b->dBpD, n, b->dX, b->stream));
    return TDVMC_OK;
}

void batch_drop_graph(tdvmc_batch* b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graph_src) (void)hipGraphDestroy(b->graph_src);
    b->graph = nullptr;
    b->graph_src = nullptr;
    b->graph_key.clear();
}

void batch_release(tdvmc_batch* b)
{
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    dfree(b->dSys);
    dfree(b->dTabF);
    dfree(b->dTabD);
    dfree(b->dBpF);
    dfree(b->dBpD);
    dfree(b->dSweepStart);
    dfree(b->dEvalStart);
    dfree(b->dSweepArgs);
    dfree(b->dEvalF);
    dfree(b->dEvalD);
    dfree(b->dMom);
    dfree(b->dMomPart);
    for (int k = 0; k < 4; k++) {
        dfree(b->dSolve[k]);
        dfree(b->dK[k]);
    }
    dfree(b->dX);
    dfree(b->dXtmp);
    dfree(b->dEpoch);
    for (auto* c : b->sys) tdvmc_destroy(c);
    b->sys.clear();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->stream = nullptr;
}

}  // namespace

namespace {

// Average time of one sweep (N moves per walker) of system s alone with
// `lanes` lanes per walker, on a scratch copy of its walkers (the chains and
// acceptance counts are left as they are).
template <typename T>
int batch_time_sweep(tdvmc_batch* b, int s, int lanes, const KTables<T>* tabs, double* dRs,
                     unsigned long long* dAccS, float* ms)
{
    tdvmc_ctx* c = b->sys[s];
    SweepArgs a;
    BHIPC(hipMemcpy(&a, b->dSweepArgs + s, sizeof(a), hipMemcpyDeviceToHost));
    a.R = dRs;
    a.acc = dAccS;
    a.lanes_per_walker = lanes;
    a.n_moves = c->sy.N;
    a.rho_out = nullptr;
    SweepArgs* dA = nullptr;
    int64_t* dStart = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TDVMC_OK;
    auto fail_hip = [&](hipError_t e) {
        b->err = std::string("HIP error: ") + hipGetErrorString(e);
        rc = TDVMC_ERR_HIP;
    };
    const int64_t zero = 0;
    const size_t rbytes = (size_t)c->W * c->sy.N * c->sy.D * sizeof(double);
    hipError_t e;
    if ((e = dalloc(&dA, 1)) != hipSuccess || (e = dalloc(&dStart, 1)) != hipSuccess ||
        (e = hipMemcpy(dA, &a, sizeof(a), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dStart, &zero, sizeof(zero), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        fail_hip(e);
    } else {
        const SweepBatch<T> bt{1, dStart, b->dSys + s, tabs + s, dA};
        const int wpb = sweep_walkers_per_block(lanes);
        const int64_t blocks = (c->W + wpb - 1) / wpb;
        const size_t lds = sweep_lds_bytes<T>(c->sy, lanes);
        constexpr int kReps = 2;
        float tot = 0.0f;
        for (int r = 0; r <= kReps && rc == TDVMC_OK; r++) {   // r = 0: warm-up
            float t = 0.0f;
            if ((e = hipMemcpyAsync(dRs, c->dR, rbytes, hipMemcpyDeviceToDevice, b->stream)) != hipSuccess ||
                (e = hipEventRecord(e0, b->stream)) != hipSuccess ||
                (e = launch_sweep_batch<T>(bt, b->D, b->periodic != 0, 0, 0, blocks, lds, b->stream)) != hipSuccess ||
                (e = hipEventRecord(e1, b->stream)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
                (e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess)
                fail_hip(e);
            if (r > 0) tot += t;
        }
        *ms = tot / kReps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(b->stream);
    dfree(dA);
    dfree(dStart);
    return rc;
}

// Average time of system s's measurement alone with `wpw` wavefronts per
// walker: a one-system launch whose descriptors point into the batch's
// arrays, except for the item (a copy with the candidate shape).
template <typename T>
int batch_time_eval(tdvmc_batch* b, int s, int wpw, const EvalBatchItem<T>* items, const KTables<T>* tabs,
                    float* ms)
{
    tdvmc_ctx* c = b->sys[s];
    const size_t slot = eval_lds_bytes<T>(c->sy, false, wpw);
    EvalBatchItem<T> it;
    BHIPC(hipMemcpy(&it, items + s, sizeof(it), hipMemcpyDeviceToHost));
    it.wpw = wpw;
    it.slot_bytes = (int64_t)slot;
    EvalBatchItem<T>* dIt = nullptr;
    int64_t* dStart = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TDVMC_OK;
    auto fail_hip = [&](hipError_t e) {
        b->err = std::string("HIP error: ") + hipGetErrorString(e);
        rc = TDVMC_ERR_HIP;
    };
    const int64_t zero = 0;
    hipError_t e;
    if ((e = dalloc(&dIt, 1)) != hipSuccess || (e = dalloc(&dStart, 1)) != hipSuccess ||
        (e = hipMemcpy(dIt, &it, sizeof(it), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dStart, &zero, sizeof(zero), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        fail_hip(e);
    } else {
        const EvalBatch<T> bt{1, dStart, b->dSys + s, tabs + s, dIt};
        const int64_t blocks = (c->W + 4 / wpw - 1) / (4 / wpw);
        const size_t lds = (size_t)(4 / wpw) * slot;
        constexpr int kReps = 3;
        if ((e = launch_eval_batch<T>(bt, b->D, b->periodic != 0, blocks, lds, b->stream)) != hipSuccess ||
            (e = hipEventRecord(e0, b->stream)) != hipSuccess) {
            fail_hip(e);
        } else {
            for (int r = 0; r < kReps && rc == TDVMC_OK; r++)
                if ((e = launch_eval_batch<T>(bt, b->D, b->periodic != 0, blocks, lds, b->stream)) != hipSuccess)
                    fail_hip(e);
            float t = 0.0f;
            if (rc == TDVMC_OK) {
                if ((e = hipEventRecord(e1, b->stream)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
                    (e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess)
                    fail_hip(e);
                *ms = t / kReps;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(b->stream);
    dfree(dIt);
    dfree(dStart);
    return rc;
}

}  // namespace

extern "C" {

int tdvmc_create(const char* config, tdvmc_ctx** out)
{
    if (!config || !out) {
        g_create_error = "tdvmc_create: null argument";
        return TDVMC_ERR_BAD_ARG;
    }
    *out = nullptr;
    tdvmc_ctx* ctx = new tdvmc_ctx();
    try {
        ctx->cfg = parse_config(read_config_text(config));
    } catch (const std::exception& e) {
        g_create_error = e.what();
        delete ctx;
        return TDVMC_ERR_BAD_CONFIG;
    }
    int rc = build_system(ctx);
    if (rc == TDVMC_OK) rc = alloc_ctx(ctx);
    if (rc == TDVMC_OK) {
        const int64_t W = ctx->cfg.N_WALKERS > 0 ? ctx->cfg.N_WALKERS : 1;
        rc = tdvmc_set_walker_count(ctx, W);
        if (rc == TDVMC_OK) rc = tdvmc_init_walkers(ctx, ctx->seed);
    }
    if (rc != TDVMC_OK) {
        g_create_error = ctx->err;
        release(ctx);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return TDVMC_OK;
}

int tdvmc_parse_config(const char* config, tdvmc_info* info)
{
    if (!config || !info) {
        g_create_error = "tdvmc_parse_config: null argument";
        return TDVMC_ERR_BAD_ARG;
    }
    tdvmc_ctx tmp;
    try {
        tmp.cfg = parse_config(read_config_text(config));
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return TDVMC_ERR_BAD_CONFIG;
    }
    int rc = build_system(&tmp);
    if (rc != TDVMC_OK) {
        g_create_error = tmp.err;
        return rc;
    }
    tmp.W = tmp.cfg.N_WALKERS > 0 ? tmp.cfg.N_WALKERS : 1;
    return tdvmc_get_info(&tmp, info);
}

int tdvmc_destroy(tdvmc_ctx* ctx)
{
    if (!ctx) return TDVMC_OK;
    release(ctx);
    delete ctx;
    return TDVMC_OK;
}

const char* tdvmc_last_error(const tdvmc_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int tdvmc_get_info(tdvmc_ctx* ctx, tdvmc_info* info)
{
    if (!ctx || !info) return TDVMC_ERR_BAD_ARG;
    std::memset(info, 0, sizeof(*info));
    info->N = ctx->sy.N;
    info->DIM = ctx->sy.D;
    info->P = ctx->sy.P;
    info->P_pair = ctx->sy.P_s;
    info->K = ctx->sy.K;
    info->one_body = ctx->sy.one_body;
    info->n_walkers = ctx->W;
    info->L = ctx->sy.L;
    info->rc = ctx->sy.rc;
    info->hbar2_2m = ctx->sy.hbar2_2m;
    info->imaginary_time = ctx->imag_time;
    info->ode_solver_type = ctx->ode_type;
    info->precision = ctx->precision;
    info->samples_per_walker = ctx->spw;
    info->timestep = ctx->dt;
    info->mc_step = ctx->sy.step;
    info->therm_moves = ctx->therm_moves;
    info->init_moves = ctx->init_moves;
    info->time = ctx->time;
    info->reg_eps = ctx->reg_eps;
    info->precondition = ctx->precondition;
    info->K_terms = ctx->sy.Kh;
    info->gr_bin_count = ctx->cfg.GR_BIN_COUNT;
    info->use_nurbs = ctx->sy.nu;
    info->bc_origin = ctx->bc_family[0];
    info->bc_cutoff = ctx->bc_family[1];
    return TDVMC_OK;
}

int tdvmc_set_walker_count(tdvmc_ctx* ctx, int64_t W)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(W >= 1 && W <= (int64_t)1 << 31, "walker count out of range");
    if (W == ctx->W) return TDVMC_OK;
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(dalloc(&ctx->dR, (size_t)W * ctx->sy.N * ctx->sy.D));
    HIPC(dalloc(&ctx->dAcc, (size_t)W));
    HIPC(dalloc(&ctx->dRho, (size_t)W * 2 * std::max(ctx->sy.Kh, 1)));
    HIPC(hipMemsetAsync(ctx->dAcc, 0, (size_t)W * sizeof(unsigned long long), ctx->stream));
    ctx->W = W;
    ctx->gen++;
    ctx->trials_total = 0;
    int rc = ensure_samples(ctx, W * ctx->spw);
    if (rc) return rc;
    rc = ensure_moments(ctx, W * ctx->spw);
    if (rc) return rc;
    rc = prof_flush(ctx);
    if (rc) return rc;
    drop_graph(ctx);
    return TDVMC_OK;
}

int tdvmc_set_params(tdvmc_ctx* ctx, const double* uR, const double* uI, double phiR, double phiI, int P)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(P == P_of(ctx), "parameter count does not match the configuration (P = " + std::to_string(P_of(ctx)) + ")");
    CHECK_ARG(uR != nullptr, "uR is null");
    for (int i = 0; i < P; i++) {
        ctx->uR[i] = uR[i];
        ctx->uI[i] = uI ? uI[i] : 0.0;
    }
    ctx->phiR = phiR;
    ctx->phiI = phiI;
    return upload_params(ctx);
}

int tdvmc_get_params(tdvmc_ctx* ctx, double* uR, double* uI, double* phiR, double* phiI, int P)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(P == P_of(ctx), "parameter count does not match the configuration");
    int rc = pull_params(ctx);
    if (rc) return rc;
    for (int i = 0; i < P; i++) {
        if (uR) uR[i] = ctx->uR[i];
        if (uI) uI[i] = ctx->uI[i];
    }
    if (phiR) *phiR = ctx->phiR;
    if (phiI) *phiI = ctx->phiI;
    return TDVMC_OK;
}

int tdvmc_set_time(tdvmc_ctx* ctx, double t)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    ctx->time = t;
    resolve_potential(ctx);
    return TDVMC_OK;
}

int tdvmc_set_walkers(tdvmc_ctx* ctx, const double* R, int64_t W)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R != nullptr, "R is null");
    int rc = tdvmc_set_walker_count(ctx, W);
    if (rc) return rc;
    HIPC(hipMemcpyAsync(ctx->dR, R, (size_t)W * ctx->sy.N * ctx->sy.D * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return TDVMC_OK;
}

int tdvmc_get_walkers(tdvmc_ctx* ctx, double* R, int64_t W)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R != nullptr && W == ctx->W, "R is null or W does not match the walker count");
    HIPC(hipMemcpyAsync(R, ctx->dR, (size_t)W * ctx->sy.N * ctx->sy.D * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return TDVMC_OK;
}

int tdvmc_init_walkers(tdvmc_ctx* ctx, uint64_t seed)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    const int N = ctx->sy.N, D = ctx->sy.D;
    const int64_t W = ctx->W;
    std::vector<double> R((size_t)W * N * D, 0.0);
    if (ctx->sy.periodic) {
        // jittered simple lattice of InitCoordinateConfiguration (src/TDVMC.cpp:713-760):
        // spacing l = L / n, jitter (u - 0.5) l / 10; n = ceil(N^(1/D)) so every
        // particle gets a site when N is not a perfect power
        int n = (int)std::lround(std::pow((double)N, 1.0 / D));
        while (std::pow((double)n, D) < N) n++;
        const double l = ctx->sy.L / n;
        for (int64_t w = 0; w < W; w++) {
            std::mt19937_64 gen(seed * 1000003ull + (uint64_t)w);
            std::uniform_real_distribution<double> u(0.0, 1.0);
            double* Rw = R.data() + (size_t)w * N * D;
            for (int p = 0; p < N; p++) {
                int idx = p;
                int c[3] = {0, 0, 0};
                for (int a = D - 1; a >= 0; a--) {
                    c[a] = idx % n;
                    idx /= n;
                }
                for (int a = 0; a < D; a++) {
                    double x = c[a] * l + (u(gen) - 0.5) * l / 10.0;
                    if (D == 1) x -= ctx->sy.L / 2.0;
                    Rw[(size_t)p * D + a] = x;
                }
            }
        }
    } else {
        // open space: Gaussian cloud of the one-body width (or unit width)
        double sigma = 1.0;
        if (ctx->sy.one_body && ctx->uR[0] > 0) sigma = std::sqrt(1.0 / (4.0 * ctx->uR[0]));
        for (int64_t w = 0; w < W; w++) {
            std::mt19937_64 gen(seed * 1000003ull + (uint64_t)w);
            std::normal_distribution<double> g(0.0, sigma);
            for (size_t i = 0; i < (size_t)N * D; i++) R[(size_t)w * N * D + i] = g(gen);
        }
    }
    HIPC(hipMemcpyAsync(ctx->dR, R.data(), R.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return TDVMC_OK;
}

int tdvmc_sweep(tdvmc_ctx* ctx, int64_t n_moves, uint64_t seed, int precision)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(n_moves >= 0, "n_moves must be >= 0");
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
    if (rc) return rc;
    rc = enqueue_sweep(ctx, prec, n_moves, seed);
    if (rc) return rc;
    HIPC(hipStreamSynchronize(ctx->stream));
    return prof_flush(ctx);
}

int tdvmc_get_acceptance(tdvmc_ctx* ctx, int64_t* accepted, int64_t* trials)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    std::vector<unsigned long long> acc((size_t)ctx->W);
    HIPC(hipMemcpyAsync(acc.data(), ctx->dAcc, acc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    int64_t s = 0;
    for (auto v : acc) s += (int64_t)v;
    if (accepted) *accepted = s;
    if (trials) *trials = ctx->trials_total;
    return TDVMC_OK;
}

int tdvmc_sample(tdvmc_ctx* ctx, uint64_t seed, int precision, tdvmc_step_data* out)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
    if (rc) return rc;
    rc = enqueue_stage_samples(ctx, prec, seed);
    if (rc) return rc;
    rc = fill_step_data(ctx, (int64_t)ctx->spw * ctx->W, out);
    if (rc) return rc;
    HIPC(hipStreamSynchronize(ctx->stream));
    return prof_flush(ctx);
}

int tdvmc_measure(tdvmc_ctx* ctx, int precision, tdvmc_step_data* out)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
    if (rc) return rc;
    rc = enqueue_eval(ctx, prec, 0);
    if (rc) return rc;
    rc = enqueue_moments(ctx, prec, ctx->W);
    if (rc) return rc;
    rc = fill_step_data(ctx, ctx->W, out);
    if (rc) return rc;
    return prof_flush(ctx);
}

int tdvmc_sample_dev(tdvmc_ctx* ctx, uint64_t seed, int precision, double* mean_dev, double* S_dev, double* F_dev,
                     void* stream)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if ((rc = join_in(ctx, st))) return rc;
    rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
    if (rc) return rc;
    if ((rc = enqueue_stage_samples(ctx, prec, seed))) return rc;
    if ((rc = copy_moments_dev(ctx, mean_dev, S_dev, F_dev))) return rc;
    return join_out(ctx, st);
}

int tdvmc_measure_dev(tdvmc_ctx* ctx, int precision, double* mean_dev, double* S_dev, double* F_dev, void* stream)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if ((rc = join_in(ctx, st))) return rc;
    rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
    if (rc) return rc;
    if ((rc = enqueue_eval(ctx, prec, 0))) return rc;
    if ((rc = enqueue_moments(ctx, prec, ctx->W))) return rc;
    if ((rc = copy_moments_dev(ctx, mean_dev, S_dev, F_dev))) return rc;
    return join_out(ctx, st);
}

int tdvmc_set_walkers_dev(tdvmc_ctx* ctx, const double* R_dev, int64_t W, void* stream)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R_dev != nullptr && W == ctx->W,
              "R_dev is null or W differs from the walker count (change it with tdvmc_set_walker_count)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = join_in(ctx, st);
    if (rc) return rc;
    HIPC(hipMemcpyAsync(ctx->dR, R_dev, (size_t)W * ctx->sy.N * ctx->sy.D * sizeof(double),
                        hipMemcpyDeviceToDevice, ctx->stream));
    return join_out(ctx, st);
}

int tdvmc_get_walkers_dev(tdvmc_ctx* ctx, double* R_dev, int64_t W, void* stream)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R_dev != nullptr && W == ctx->W, "R_dev is null or W differs from the walker count");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = join_in(ctx, st);
    if (rc) return rc;
    HIPC(hipMemcpyAsync(R_dev, ctx->dR, (size_t)W * ctx->sy.N * ctx->sy.D * sizeof(double),
                        hipMemcpyDeviceToDevice, ctx->stream));
    return join_out(ctx, st);
}

int tdvmc_eval_configs_dev(tdvmc_ctx* ctx, const double* R_dev, int64_t W, int precision, double* expo_dev,
                           double* ER_dev, double* EI_dev, double* O_dev, double* grad_dev, double* lap_dev,
                           void* stream)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R_dev && expo_dev && ER_dev && EI_dev && O_dev && W >= 1, "null device pointer or W < 1");
    CHECK_ARG((grad_dev == nullptr) == (lap_dev == nullptr), "grad and lap must be both set or both null");
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (st != ctx->stream) HIPC(hipStreamSynchronize(ctx->stream));
    hipStream_t keep = ctx->stream;
    ctx->stream = st;
    rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
    ctx->stream = keep;
    if (rc) return rc;
    if (prec == TDVMC_MIXED) {
        HIPC((launch_eval<float, double>(ctx->sy, tables_of(ctx, ctx->tf), R_dev, W, expo_dev, ER_dev, EI_dev, O_dev,
                                         nullptr, grad_dev, lap_dev, nullptr, nullptr, st)));
    } else {
        HIPC((launch_eval<double, double>(ctx->sy, tables_of(ctx, ctx->td), R_dev, W, expo_dev, ER_dev, EI_dev, O_dev,
                                          nullptr, grad_dev, lap_dev, nullptr, nullptr, st)));
    }
    return TDVMC_OK;
}

namespace {

// Fed configurations from host memory: tdvmc_eval_configs, and (bsum != null)
// the basis sums of tdvmc_basis_sums.
int eval_configs_host(tdvmc_ctx* ctx, const double* R, int64_t W, int precision, double* expo, double* ER,
                      double* EI, double* O, double* grad, double* lap, double* other, double* bsum)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R != nullptr && W >= 1, "R is null or W < 1");
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    const int N = ctx->sy.N, D = ctx->sy.D, P = P_of(ctx), S = ctx->sy.P_s + 2;
    const bool pp = grad != nullptr || lap != nullptr;
    double *dR = nullptr, *dEx = nullptr, *dER = nullptr, *dEI = nullptr, *dO = nullptr, *dG = nullptr,
           *dL = nullptr, *dOt = nullptr, *dB = nullptr;
    auto cleanup = [&]() {
        dfree(dR); dfree(dEx); dfree(dER); dfree(dEI); dfree(dO); dfree(dG); dfree(dL); dfree(dOt); dfree(dB);
    };
    auto hipc = [&](hipError_t e, const char* what) {
        if (e != hipSuccess) {
            ctx->err = std::string(what) + ": " + hipGetErrorString(e);
            return false;
        }
        return true;
    };
    bool ok = hipc(dalloc(&dR, (size_t)W * N * D), "hipMalloc") && hipc(dalloc(&dEx, (size_t)W), "hipMalloc") &&
              hipc(dalloc(&dER, (size_t)W), "hipMalloc") && hipc(dalloc(&dEI, (size_t)W), "hipMalloc") &&
              hipc(dalloc(&dO, (size_t)W * P), "hipMalloc") && hipc(dalloc(&dOt, (size_t)W * 9), "hipMalloc");
    if (ok && pp) ok = hipc(dalloc(&dG, (size_t)W * N * D * 2), "hipMalloc") && hipc(dalloc(&dL, (size_t)W * N * 2), "hipMalloc");
    if (ok && bsum) ok = hipc(dalloc(&dB, (size_t)W * S), "hipMalloc");
    if (ok) ok = hipc(hipMemcpyAsync(dR, R, (size_t)W * N * D * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "hipMemcpy");
    if (ok) {
        rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
        if (rc) {
            cleanup();
            return rc;
        }
        hipEvent_t ev{};
        prof_begin(ctx, "eval", &ev);
        if (prec == TDVMC_MIXED)
            ok = hipc((launch_eval<float, double>(ctx->sy, tables_of(ctx, ctx->tf), dR, W, dEx, dER, dEI, dO, dOt, dG,
                                                  dL, nullptr, nullptr, ctx->stream, dB)), "eval kernel");
        else
            ok = hipc((launch_eval<double, double>(ctx->sy, tables_of(ctx, ctx->td), dR, W, dEx, dER, dEI, dO, dOt,
                                                   dG, dL, nullptr, nullptr, ctx->stream, dB)), "eval kernel");
        prof_end(ctx, "eval", ev);
    }
    auto down = [&](double* h, double* d, size_t n) {
        if (ok && h) ok = hipc(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy");
    };
    down(expo, dEx, (size_t)W);
    down(ER, dER, (size_t)W);
    down(EI, dEI, (size_t)W);
    down(O, dO, (size_t)W * P);
    down(other, dOt, (size_t)W * 9);
    if (pp) {
        down(grad, dG, (size_t)W * N * D * 2);
        down(lap, dL, (size_t)W * N * 2);
    }
    down(bsum, dB, (size_t)W * S);
    if (ok) ok = hipc(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    cleanup();
    if (!ok) return TDVMC_ERR_HIP;
    return prof_flush(ctx);
}

}  // namespace

int tdvmc_eval_configs(tdvmc_ctx* ctx, const double* R, int64_t W, int precision, double* expo, double* ER,
                       double* EI, double* O, double* grad, double* lap, double* other)
{
    return eval_configs_host(ctx, R, W, precision, expo, ER, EI, O, grad, lap, other, nullptr);
}

int tdvmc_basis_sums(tdvmc_ctx* ctx, const double* R, int64_t W, int precision, double* B)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(B != nullptr, "tdvmc_basis_sums: B is null");
    return eval_configs_host(ctx, R, W, precision, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, B);
}

int tdvmc_spline_bc(tdvmc_ctx* ctx, double* start, int* n_start, double* end, int* n_end)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(start && n_start && end && n_end, "tdvmc_spline_bc: null output");
    const KSys& s = ctx->sy;
    for (int i = 0; i < 3; i++)
        for (int m = 0; m < 3; m++) {
            start[i * 3 + m] = i < s.bc_n0 ? s.bc_f0[i][m] : 0.0;
            end[i * 3 + m] = i < s.bc_n1 ? s.bc_f1[i][m] : 0.0;
        }
    *n_start = s.bc_n0;
    *n_end = s.bc_n1;
    return TDVMC_OK;
}

int tdvmc_observables_info(tdvmc_ctx* ctx, int* n_gr, int* n_sk, double* gr_grid, double* sk_grid)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int rc = ensure_obs(ctx);
    if (rc) return rc;
    const auto& o = ctx->obs;
    if (n_gr) *n_gr = o.count;
    if (n_sk) *n_sk = o.nshell;
    if (gr_grid) std::copy(o.centres.begin(), o.centres.end(), gr_grid);
    if (sk_grid) std::copy(o.kgrid.begin(), o.kgrid.end(), sk_grid);
    return TDVMC_OK;
}

int tdvmc_observables_configs(tdvmc_ctx* ctx, const double* R, int64_t W, int precision, double* gr, double* sk)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R != nullptr && W >= 1 && (gr || sk), "R is null, W < 1, or both outputs are null");
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    rc = ensure_obs(ctx);
    if (rc) return rc;
    const auto& o = ctx->obs;
    const int N = ctx->sy.N, D = ctx->sy.D;
    double *dR = nullptr, *dG = nullptr, *dS = nullptr;
    auto cleanup = [&]() {
        dfree(dR);
        dfree(dG);
        dfree(dS);
    };
    auto fail = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return false;
        ctx->err = std::string(what) + ": " + hipGetErrorString(e);
        cleanup();
        return true;
    };
    if (fail(dalloc(&dR, (size_t)W * N * D), "hipMalloc") || fail(dalloc(&dS, (size_t)W * o.nshell), "hipMalloc") ||
        (gr && fail(dalloc(&dG, (size_t)W * o.count), "hipMalloc")) ||
        fail(hipMemcpyAsync(dR, R, (size_t)W * N * D * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
             "hipMemcpy"))
        return TDVMC_ERR_HIP;
    rc = clear_obs(ctx);
    if (!rc) rc = enqueue_obs(ctx, prec, dR, W, dG, dS);
    if (rc) {
        cleanup();
        return rc;
    }
    if ((gr && fail(hipMemcpyAsync(gr, dG, (size_t)W * o.count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                    "hipMemcpy")) ||
        (sk && fail(hipMemcpyAsync(sk, dS, (size_t)W * o.nshell * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                    "hipMemcpy")) ||
        fail(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize"))
        return TDVMC_ERR_HIP;
    cleanup();
    return prof_flush(ctx);
}

int tdvmc_additional_properties(tdvmc_ctx* ctx, int64_t n_samples, uint64_t seed, int precision, double* gr,
                                double* sk)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    const int64_t steps = n_samples > 0 ? n_samples : (int64_t)ctx->cfg.MC_NADDITIONALSTEPS;
    if (steps <= 0) {
        ctx->err = "MC_NADDITIONALSTEPS must be positive (or pass n_samples > 0)";
        return TDVMC_ERR_BAD_CONFIG;
    }
    CHECK_ARG(ctx->W >= 1, "no walkers");
    rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
    if (rc) return rc;
    rc = ensure_obs(ctx);
    if (rc) return rc;
    auto& o = ctx->obs;
    if (o.skw_rows < ctx->W) {
        HIPC(dalloc(&o.dSkW, (size_t)ctx->W * o.nshell));
        o.skw_rows = ctx->W;
    }
    rc = clear_obs(ctx);
    if (rc) return rc;
    rc = enqueue_sweep(ctx, prec, (int64_t)ctx->cfg.MC_NADDITIONALINITIALIZATIONSTEPS, seed);
    for (int64_t i = 0; !rc && i < steps; i++) {
        rc = enqueue_sweep(ctx, prec, (int64_t)ctx->cfg.MC_NADDITIONALTHERMSTEPS, seed);
        if (!rc) rc = enqueue_obs(ctx, prec, ctx->dR, ctx->W, nullptr, o.dSkW);
        if (!rc) HIPC(launch_obs_reduce(o.dSkW, ctx->W, o.nshell, o.dSkAcc, ctx->stream));
    }
    if (rc) return rc;
    std::vector<unsigned long long> tot(o.count);
    std::vector<double> acc(o.nshell);
    HIPC(hipMemcpyAsync(tot.data(), o.dTot, tot.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        ctx->stream));
    HIPC(hipMemcpyAsync(acc.data(), o.dSkAcc, acc.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    const double n = (double)ctx->W * (double)steps;
    if (gr)
        for (int b = 0; b < o.count; b++) gr[b] = (double)tot[b] * o.scale[b] / n;
    if (sk)
        for (int k = 0; k < o.nshell; k++) sk[k] = acc[k] / n;
    return prof_flush(ctx);
}

int tdvmc_moments(tdvmc_ctx* ctx, const double* O, const double* ER, const double* EI, int64_t W, int P,
                  int precision, double* mean, double* S, double* F)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(O && ER && EI && W >= 1 && P >= 1, "bad moments arguments");
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    double *dO = nullptr, *dE = nullptr, *dM = nullptr, *dS = nullptr, *dF = nullptr, *dPart = nullptr, *dCol = nullptr;
    void* dXt = nullptr;
    size_t xb, pb, cb;
    moments_work_bytes(W, P, prec == TDVMC_MIXED ? 1 : 0, &xb, &pb, &cb);
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t x) {
        if (e == hipSuccess) e = x;
    };
    step(dalloc(&dO, (size_t)W * P));
    step(dalloc(&dE, 2 * (size_t)W));
    step(dalloc(&dM, (size_t)P + 2));
    step(dalloc(&dS, (size_t)P * P));
    step(dalloc(&dF, 2 * (size_t)P));
    step(dalloc(&dPart, (pb + 7) / 8));
    step(dalloc(&dCol, (cb + 7) / 8));
    if (e == hipSuccess) e = hipMalloc(&dXt, xb);
    step(hipMemcpyAsync(dO, O, (size_t)W * P * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    step(hipMemcpyAsync(dE, ER, (size_t)W * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    step(hipMemcpyAsync(dE + W, EI, (size_t)W * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (e == hipSuccess) {
        MomentsWork wk{dXt, dPart, dCol, xb, pb, cb};
        e = launch_moments<double>(dO, dE, dE + W, W, P, prec == TDVMC_MIXED ? 1 : 0, wk, dM, dS, dF, ctx->stream);
    }
    if (mean) step(hipMemcpyAsync(mean, dM, ((size_t)P + 2) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (S) step(hipMemcpyAsync(S, dS, (size_t)P * P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (F) step(hipMemcpyAsync(F, dF, 2 * (size_t)P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    step(hipStreamSynchronize(ctx->stream));
    dfree(dO); dfree(dE); dfree(dM); dfree(dS); dfree(dF); dfree(dPart); dfree(dCol);
    if (dXt) (void)hipFree(dXt);
    if (e != hipSuccess) {
        ctx->err = std::string("moments: ") + hipGetErrorString(e);
        return TDVMC_ERR_HIP;
    }
    return TDVMC_OK;
}

int tdvmc_solve(tdvmc_ctx* ctx, const double* mean, const double* S, const double* F, int P, int imag_time,
                int precondition, double eps, int solver_type, double* xdot, int* n_bad)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(mean && S && F && xdot && P >= 1, "bad solve arguments");
    CHECK_ARG(solver_type == 0 || solver_type == 1, "solver type must be 0 or 1");
    CHECK_ARG(P <= solve_max_params(solver_type),
              "P = " + std::to_string(P) + " exceeds the solver's limit of " + std::to_string(solve_max_params(solver_type)));
    double *dM = nullptr, *dS = nullptr, *dF = nullptr, *dW = nullptr, *dX = nullptr;
    int* dSt = nullptr;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t x) {
        if (e == hipSuccess) e = x;
    };
    step(dalloc(&dM, (size_t)P + 2));
    step(dalloc(&dS, (size_t)P * P));
    step(dalloc(&dF, 2 * (size_t)P));
    step(dalloc(&dW, (size_t)P * P + 4 * (size_t)P));
    step(dalloc(&dX, 2 * (size_t)P + 2));
    step(dalloc(&dSt, 1));
    step(hipMemcpyAsync(dM, mean, ((size_t)P + 2) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    step(hipMemcpyAsync(dS, S, (size_t)P * P * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    step(hipMemcpyAsync(dF, F, 2 * (size_t)P * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int bad = 0;
    if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t*>(dSt), 0u, 1, ctx->stream);
    if (e == hipSuccess) {
        // type 1 regularises only together with the scaling (src/TDVMC.cpp:1766-1771)
        const double e_eff = solver_type == 1 ? (precondition ? 0.002 : 0.0) : eps;
        SolveArgs a{P, imag_time, precondition, e_eff, dM, dS, dF, dW, dX, dSt, solver_type == 1 ? 1 : 0};
        e = launch_solve(a, ctx->stream);
    }
    step(hipMemcpyAsync(xdot, dX, (2 * (size_t)P + 2) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    step(hipMemcpyAsync(&bad, dSt, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    step(hipStreamSynchronize(ctx->stream));
    dfree(dM); dfree(dS); dfree(dF); dfree(dW); dfree(dX); dfree(dSt);
    if (e != hipSuccess) {
        ctx->err = std::string("solve: ") + hipGetErrorString(e);
        return TDVMC_ERR_HIP;
    }
    if (n_bad) *n_bad = bad;
    if (bad > 0) {
        ctx->err = "non-positive pivot in the Cholesky factorisation";
        return TDVMC_ERR_NOT_PD;
    }
    return TDVMC_OK;
}

int tdvmc_tdvp_step(tdvmc_ctx* ctx, double dt, int ode_type, int precision, int use_graph)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    const int ode = ode_type < 0 ? ctx->ode_type : ode_type;
    CHECK_ARG(ode == 0 || ode == 1 || ode == 2 || ode == 3 || ode == 5,
              "ODE type must be 0 (Euler), 1 (predictor-corrector), 2 (RK4), 3 (RK4 reusing samples) or 5 "
              "(predictor-corrector reusing samples)");
    if ((ode == 3 || ode == 5) && ctx->spw > 1) {
        // every sample's configurations are kept for the stages that reuse them
        const int64_t need = (int64_t)ctx->spw * ctx->W * ctx->sy.N * ctx->sy.D;
        if (ctx->samp_elems < need) {
            drop_graph(ctx);   // a captured graph holds the old buffer's address
            HIPC(dalloc(&ctx->dSamp, (size_t)need));
            ctx->samp_elems = need;
        }
    }
    CHECK_ARG(P_of(ctx) <= solve_max_params(ctx->solver_type),
              "P = " + std::to_string(P_of(ctx)) + " exceeds the solver's limit of " +
                  std::to_string(solve_max_params(ctx->solver_type)));
    if (use_graph) {
        char key[256];
        std::snprintf(key, sizeof(key), "%d/%d/%.17g/%lld/%.17g/%.17g/%d/%d/%d/%d", prec, ode, dt, (long long)ctx->W,
                      ctx->sy.pot_a, ctx->sy.pot_b, ctx->prof ? 1 : 0, ctx->lanes, ctx->lanes_tuned[0],
                      ctx->lanes_tuned[1]);
        if (!ctx->graph || ctx->graph_key != key) {
            rc = prof_flush(ctx);
            if (rc) return rc;
            drop_graph(ctx);
            if (ctx->prof) {
                // segmented capture with event-record nodes between the segments
                HIPC(hipGraphCreate(&ctx->seg_top, 0));
                ctx->seg_last = nullptr;
                ctx->capturing = true;
                ctx->in_capture = true;
                rc = seg_begin(ctx);
                if (!rc) rc = enqueue_tdvp_step(ctx, prec, ode, dt);
                if (!rc) rc = seg_cut(ctx);
                ctx->capturing = false;
                ctx->in_capture = false;
                if (ctx->cap_active) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(ctx->stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    ctx->cap_active = false;
                }
                hipGraph_t top = ctx->seg_top;
                ctx->seg_top = nullptr;
                ctx->seg_last = nullptr;
                if (rc) {
                    (void)hipGraphDestroy(top);
                    drop_graph(ctx);
                    return rc;
                }
                ctx->graph_src = top;
            } else {
                HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                ctx->in_capture = true;
                rc = enqueue_tdvp_step(ctx, prec, ode, dt);
                ctx->in_capture = false;
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(ctx->stream, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    drop_graph(ctx);
                    return rc;
                }
                HIPC(e);
                ctx->graph_src = g;
            }
            HIPC(hipGraphInstantiate(&ctx->graph, ctx->graph_src, nullptr, nullptr, 0));
            ctx->graph_key = key;
        }
        if (ctx->accept_check)
            HIPC(hipMemcpyAsync(ctx->dXprev, ctx->dX, (2 * (size_t)P_of(ctx) + 2) * sizeof(double),
                                hipMemcpyDeviceToDevice, ctx->stream));
        HIPC(hipGraphLaunch(ctx->graph, ctx->stream));
        // the graph's event nodes were re-recorded by this replay
        for (const auto& p : ctx->graph_ev) ctx->pending.push_back(p);
        ctx->trials_total += fresh_stages(ode) * (ctx->init_moves + ctx->spw * ctx->therm_moves) * ctx->W;
    } else {
        if (ctx->accept_check)
            HIPC(hipMemcpyAsync(ctx->dXprev, ctx->dX, (2 * (size_t)P_of(ctx) + 2) * sizeof(double),
                                hipMemcpyDeviceToDevice, ctx->stream));
        rc = enqueue_tdvp_step(ctx, prec, ode, dt);
        if (rc) return rc;
    }
    ctx->tf.dirty = ctx->td.dirty = true;
    ctx->host_params_stale = true;
    rc = prof_flush(ctx);
    if (rc) return rc;
    if (ctx->accept_check) {
        const size_t n = 2 * (size_t)P_of(ctx) + 2;
        std::vector<double> x(n), x0(n);
        int bad = 0;
        HIPC(hipMemcpyAsync(x.data(), ctx->dX, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(x0.data(), ctx->dXprev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(&bad, ctx->dStatus, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipStreamSynchronize(ctx->stream));
        if (!accept_step(ctx, bad, x, x0)) {
            // roll back to the parameters before the step; time does not advance
            HIPC(hipMemcpyAsync(ctx->dX, ctx->dXprev, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            HIPC(launch_unpack_params(P_of(ctx), ctx->dX, ctx->dUR, ctx->dUI, ctx->stream));
            HIPC(hipStreamSynchronize(ctx->stream));
            ctx->err = bad > 0 ? "step rejected: non-positive pivot in the Cholesky factorisation"
                               : "step rejected by the parameter acceptance check";
            return TDVMC_ERR_REJECTED;
        }
    }
    ctx->time += dt;
    ctx->step += 1;
    resolve_potential(ctx);
    return TDVMC_OK;
}

int tdvmc_get_step_status(tdvmc_ctx* ctx, int* n_bad)
{
    if (!ctx || !n_bad) return TDVMC_ERR_BAD_ARG;
    int bad = 0;
    HIPC(hipMemcpyAsync(&bad, ctx->dStatus, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    *n_bad = bad;
    return TDVMC_OK;
}

int tdvmc_adjoint(tdvmc_ctx* ctx, const double* R, int64_t W, const double* gbar, int precision, double* dR,
                  double* dalpha)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R && gbar && W >= 1, "bad adjoint arguments");
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    const int N = ctx->sy.N, D = ctx->sy.D, P = P_of(ctx);
    double *dRin = nullptr, *dEx = nullptr, *dER = nullptr, *dEI = nullptr, *dO = nullptr, *dG = nullptr,
           *dL = nullptr, *dGb = nullptr, *dOutR = nullptr, *dOutA = nullptr, *dOuter = nullptr;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t x) {
        if (e == hipSuccess) e = x;
    };
    step(dalloc(&dRin, (size_t)W * N * D));
    step(dalloc(&dEx, (size_t)W));
    step(dalloc(&dER, (size_t)W));
    step(dalloc(&dEI, (size_t)W));
    step(dalloc(&dO, (size_t)W * P));
    step(dalloc(&dG, (size_t)W * N * D * 2));
    step(dalloc(&dL, (size_t)W * N * 2));
    step(dalloc(&dGb, (size_t)W));
    step(dalloc(&dOutR, (size_t)W * N * D));
    step(dalloc(&dOutA, (size_t)P));
    step(dalloc(&dOuter, (size_t)W));
    step(hipMemcpyAsync(dRin, R, (size_t)W * N * D * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    step(hipMemcpyAsync(dGb, gbar, (size_t)W * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (e == hipSuccess) {
        rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
        if (rc == TDVMC_OK) {
            if (prec == TDVMC_MIXED)
                e = launch_eval<float, double>(ctx->sy, tables_of(ctx, ctx->tf), dRin, W, dEx, dER, dEI, dO, nullptr,
                                               dG, dL, dOuter, nullptr, ctx->stream);
            else
                e = launch_eval<double, double>(ctx->sy, tables_of(ctx, ctx->td), dRin, W, dEx, dER, dEI, dO, nullptr,
                                                dG, dL, dOuter, nullptr, ctx->stream);
        }
    }
    if (e == hipSuccess && rc == TDVMC_OK) {
        e = launch_adjoint_scale(dG, dGb, W, N * D, dOutR, ctx->stream);
        if (e == hipSuccess) e = launch_adjoint_reduce(dO, dGb, W, P, dOutA, ctx->stream);
        if (e == hipSuccess) e = launch_adjoint_outer(dOuter, dGb, W, ctx->sy.outer_idx, dOutA, ctx->stream);
    }
    if (dR) step(hipMemcpyAsync(dR, dOutR, (size_t)W * N * D * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (dalpha) step(hipMemcpyAsync(dalpha, dOutA, (size_t)P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    step(hipStreamSynchronize(ctx->stream));
    dfree(dRin); dfree(dEx); dfree(dER); dfree(dEI); dfree(dO); dfree(dG); dfree(dL); dfree(dGb); dfree(dOutR);
    dfree(dOutA);
    dfree(dOuter);
    if (rc) return rc;
    if (e != hipSuccess) {
        ctx->err = std::string("adjoint: ") + hipGetErrorString(e);
        return TDVMC_ERR_HIP;
    }
    return TDVMC_OK;
}

int tdvmc_adjoint_noalloc(tdvmc_ctx* ctx, const double* R_dev, int64_t W, const double* gbar_dev, int precision,
                          double* dR_dev, double* dalpha_dev, void* stream)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(R_dev && gbar_dev && W >= 1, "tdvmc_adjoint_noalloc: null device pointer or W < 1");
    CHECK_ARG(dR_dev || dalpha_dev, "tdvmc_adjoint_noalloc: no output requested");
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    const int N = ctx->sy.N, D = ctx->sy.D, P = P_of(ctx);
    // scratch per walker: O [P], grad [N D 2], lap [N 2], ln psi, E_R, E_I, outer count
    const size_t per = (size_t)P + 2 * (size_t)N * D + 2 * (size_t)N + 4;
    const int64_t wmax = (int64_t)(kScratchBytes / sizeof(double) / per);
    if (W > wmax) {
        ctx->err = "tdvmc_adjoint_noalloc: W = " + std::to_string(W) + " exceeds the " +
                   std::to_string(kScratchBytes >> 20) + " MiB scratch, which holds " + std::to_string(wmax) +
                   " walkers of this system";
        return TDVMC_ERR_BAD_ARG;
    }
    double* base = g_scratch[ctx->device];
    if (!base) {
        ctx->err = "tdvmc_adjoint_noalloc: scratch not set up for this device";
        return TDVMC_ERR_STATE;
    }
    // every scratch entry read below is written first by this call's
    // evaluation, so nothing needs clearing between calls
    double* dO = base;
    double* dG = dO + (size_t)W * P;
    double* dL = dG + (size_t)W * N * D * 2;
    double* dEx = dL + (size_t)W * N * 2;
    double* dER = dEx + W;
    double* dEI = dER + W;
    double* dOuter = dEI + W;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (st != ctx->stream) HIPC(hipStreamSynchronize(ctx->stream));
    hipStream_t keep = ctx->stream;
    ctx->stream = st;
    rc = prec == TDVMC_MIXED ? ensure_tables(ctx, ctx->tf, false) : ensure_tables(ctx, ctx->td, false);
    ctx->stream = keep;
    if (rc) return rc;
    if (prec == TDVMC_MIXED) {
        HIPC((launch_eval<float, double>(ctx->sy, tables_of(ctx, ctx->tf), R_dev, W, dEx, dER, dEI, dO, nullptr, dG,
                                         dL, dOuter, nullptr, st)));
    } else {
        HIPC((launch_eval<double, double>(ctx->sy, tables_of(ctx, ctx->td), R_dev, W, dEx, dER, dEI, dO, nullptr, dG,
                                          dL, dOuter, nullptr, st)));
    }
    if (dR_dev) HIPC(launch_adjoint_scale(dG, gbar_dev, W, N * D, dR_dev, st));
    if (dalpha_dev) {
        HIPC(launch_adjoint_reduce(dO, gbar_dev, W, P, dalpha_dev, st));
        HIPC(launch_adjoint_outer(dOuter, gbar_dev, W, ctx->sy.outer_idx, dalpha_dev, st));
    }
    return TDVMC_OK;
}

int tdvmc_get_rng_state(tdvmc_ctx* ctx, uint64_t* seed, uint32_t* epoch)
{
    if (!ctx || !seed || !epoch) return TDVMC_ERR_BAD_ARG;
    uint32_t e = 0;
    HIPC(hipMemcpyAsync(&e, ctx->dEpoch, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    *seed = ctx->seed;
    *epoch = e;
    return TDVMC_OK;
}

int tdvmc_set_rng_state(tdvmc_ctx* ctx, uint64_t seed, uint32_t epoch)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(launch_fill_u32(ctx->dEpoch, epoch, 1, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    if (seed != ctx->seed) drop_graph(ctx);   // a captured step holds the seed in its launch arguments
    ctx->seed = seed;
    return TDVMC_OK;
}

int tdvmc_profile_enable(tdvmc_ctx* ctx, int enable)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int rc = prof_flush(ctx);
    if (rc) return rc;
    ctx->prof = enable != 0;
    ctx->prof_acc.clear();
    return TDVMC_OK;
}

int tdvmc_profile_get(tdvmc_ctx* ctx, const char* name, double* total_ms, int64_t* launches)
{
    if (!ctx || !name) return TDVMC_ERR_BAD_ARG;
    int rc = prof_flush(ctx);
    if (rc) return rc;
    auto it = ctx->prof_acc.find(name);
    if (total_ms) *total_ms = it == ctx->prof_acc.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == ctx->prof_acc.end() ? 0 : it->second.launches;
    return TDVMC_OK;
}

int tdvmc_set_sweep_lanes(tdvmc_ctx* ctx, int lanes_per_walker)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    CHECK_ARG(lanes_per_walker == 0 || lanes_per_walker == 16 || lanes_per_walker == 32 || lanes_per_walker == 64,
              "lanes per walker must be 0, 16, 32 or 64");
    if (lanes_per_walker) {
        const int kpl = (ctx->sy.Kh + lanes_per_walker - 1) / lanes_per_walker;
        CHECK_ARG(kpl <= 8, "too many k-vectors for this lane count");
    }
    ctx->lanes = lanes_per_walker;
    return TDVMC_OK;
}

// Times one sweep (N moves) of a scratch copy of up to 16384 walkers for each
// admissible lane count and keeps the fastest for this precision.  The
// walkers, acceptance counts and random stream of the context are untouched.
int tdvmc_autotune(tdvmc_ctx* ctx, int precision, int* lanes_out)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    int prec, rc = resolve_precision(ctx, precision, &prec);
    if (rc) return rc;
    rc = prof_flush(ctx);
    if (rc) return rc;
    if (prec == TDVMC_MIXED) {
        rc = ensure_tables(ctx, ctx->tf, false);
    } else {
        rc = ensure_tables(ctx, ctx->td, false);
    }
    if (rc) return rc;
    const int N = ctx->sy.N, D = ctx->sy.D;
    const int64_t Wt = std::min<int64_t>(ctx->W, 16384);
    double* dRt = nullptr;
    HIPC(dalloc(&dRt, (size_t)Wt * N * D));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int best = 0;
    float best_ms = 0.f;
    for (int lanes : {16, 32, 64}) {
        if (e != hipSuccess) break;
        if ((ctx->sy.Kh + lanes - 1) / lanes > 8) continue;
        SweepArgs a;
        a.R = dRt;
        a.W = Wt;
        a.n_moves = N;
        a.seed = 0x9e3779b97f4a7c15ull;
        a.epoch = ctx->dEpoch;
        a.acc = nullptr;
        a.lanes_per_walker = lanes;
        a.rho_out = nullptr;
        float ms = 0.f;
        for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
            e = hipMemcpyAsync(dRt, ctx->dR, (size_t)Wt * N * D * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
            if (e == hipSuccess)
                e = prec == TDVMC_MIXED ? launch_sweep<float>(ctx->sy, tables_of(ctx, ctx->tf), a, ctx->stream)
                                        : launch_sweep<double>(ctx->sy, tables_of(ctx, ctx->td), a, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (rep > 0) ms += t;   // the first launch warms up
        }
        if (e == hipSuccess && (best == 0 || ms < best_ms)) {
            best = lanes;
            best_ms = ms;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(dRt);
    if (e != hipSuccess) {
        ctx->err = std::string("autotune: ") + hipGetErrorString(e);
        return TDVMC_ERR_HIP;
    }
    ctx->lanes_tuned[prec == TDVMC_MIXED ? 1 : 0] = best;
    if (lanes_out) *lanes_out = best;
    return TDVMC_OK;
}

int tdvmc_synchronize(tdvmc_ctx* ctx)
{
    if (!ctx) return TDVMC_ERR_BAD_ARG;
    HIPC(hipStreamSynchronize(ctx->stream));
    return prof_flush(ctx);
}

// ---------------------------------------------------------------- batches

int tdvmc_batch_create(const char* const* configs, int n_systems, tdvmc_batch** out)
{
    if (!configs || !out || n_systems < 1) {
        g_create_error = "tdvmc_batch_create: need at least one config";
        return TDVMC_ERR_BAD_ARG;
    }
    *out = nullptr;
    tdvmc_batch* b = new tdvmc_batch();
    auto fail = [&](int rc, const std::string& msg) {
        g_create_error = msg;
        batch_release(b);
        delete b;
        return rc;
    };
    for (int s = 0; s < n_systems; s++) {
        tdvmc_ctx* c = nullptr;
        const int rc = tdvmc_create(configs[s], &c);
        if (rc) return fail(rc, "system " + std::to_string(s) + ": " + g_create_error);
        b->sys.push_back(c);
        if (s == 0) {
            b->D = c->sy.D;
            b->periodic = c->sy.periodic;
            b->seed = c->seed;
        } else if (c->sy.D != b->D || c->sy.periodic != b->periodic) {
            return fail(TDVMC_ERR_BAD_CONFIG, "all systems of a batch need the same DIM and box kind");
        }
        if (c->normalize_wf || c->adjust_params || c->accept_check)
            return fail(TDVMC_ERR_BAD_CONFIG, "USE_NORMALIZE_WF, USE_ADJUST_PARAMETERS and the parameter acceptance "
                                              "check are not available for batches");
        if (c->spw != 1) return fail(TDVMC_ERR_BAD_CONFIG, "batches take one sample per walker and stage");
        if (c->sy.nu) return fail(TDVMC_ERR_BAD_CONFIG, "USE_NURBS (non-uniform knots) is not available for batches");
        // the batched solve sizes its LDS for the Cholesky panel of the largest system
        if (c->sy.P > solve_max_params(0))
            return fail(TDVMC_ERR_BAD_CONFIG, "system " + std::to_string(s) + ": P = " + std::to_string(c->sy.P) +
                                                  " exceeds the batched solver's limit of " +
                                                  std::to_string(solve_max_params(0)));
        b->kh_max = std::max(b->kh_max, c->sy.Kh);
        b->P_max = std::max(b->P_max, c->sy.P);
        b->N_max = std::max(b->N_max, c->sy.N);
        b->xoff.push_back(b->xtotal);
        b->xtotal += 2 * (int64_t)c->sy.P + 2;
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(TDVMC_ERR_HIP, "hipStreamCreate failed");
    if (dalloc(&b->dX, (size_t)b->xtotal) != hipSuccess || dalloc(&b->dXtmp, (size_t)b->xtotal) != hipSuccess ||
        dalloc(&b->dEpoch, 1) != hipSuccess)
        return fail(TDVMC_ERR_HIP, "out of device memory");
    for (int k = 0; k < 4; k++)
        if (dalloc(&b->dK[k], (size_t)b->xtotal) != hipSuccess) return fail(TDVMC_ERR_HIP, "out of device memory");
    if (hipMemsetAsync(b->dX, 0, (size_t)b->xtotal * sizeof(double), b->stream) != hipSuccess ||
        launch_fill_u32(b->dEpoch, 0u, 1, b->stream) != hipSuccess)
        return fail(TDVMC_ERR_HIP, "initialisation failed");
    for (auto* c : b->sys)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(TDVMC_ERR_HIP, "initialisation failed");
    const int rc = batch_upload(b, 0, b->seed, true);
    if (rc) return fail(rc, b->err);
    *out = b;
    return TDVMC_OK;
}

int tdvmc_batch_destroy(tdvmc_batch* b)
{
    if (!b) return TDVMC_OK;
    batch_release(b);
    delete b;
    return TDVMC_OK;
}

const char* tdvmc_batch_last_error(const tdvmc_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int tdvmc_batch_size(tdvmc_batch* b, int* n_systems)
{
    if (!b || !n_systems) return TDVMC_ERR_BAD_ARG;
    *n_systems = (int)b->sys.size();
    return TDVMC_OK;
}

tdvmc_ctx* tdvmc_batch_system(tdvmc_batch* b, int s)
{
    if (!b || s < 0 || s >= (int)b->sys.size()) return nullptr;
    return b->sys[s];
}

// the member contexts' own streams must be idle before batch work touches their buffers
static int batch_sync_members(tdvmc_batch* b)
{
    for (auto* c : b->sys) BHIPC(hipStreamSynchronize(c->stream));
    return TDVMC_OK;
}

static int batch_precision(tdvmc_batch* b, int precision, int* prec)
{
    int p = precision;
    if (p < 0) p = b->sys[0]->precision;
    if (p != TDVMC_FP64 && p != TDVMC_MIXED) {
        b->err = "precision must be TDVMC_FP64 or TDVMC_MIXED";
        return TDVMC_ERR_BAD_ARG;
    }
    *prec = p;
    return TDVMC_OK;
}

int tdvmc_batch_set_sweep_lanes(tdvmc_batch* b, int lanes)
{
    if (!b) return TDVMC_ERR_BAD_ARG;
    if (!(lanes == 0 || lanes == 16 || lanes == 32 || lanes == 64) ||
        (lanes && (b->kh_max + lanes - 1) / lanes > 8)) {
        b->err = "lanes per walker must be 0, 16, 32 or 64 (at most 8 k-space terms per lane)";
        return TDVMC_ERR_BAD_ARG;
    }
    b->lanes = lanes;
    batch_drop_graph(b);
    return TDVMC_OK;
}

int tdvmc_batch_autotune(tdvmc_batch* b, int precision, int* waves_out, int* lanes_out)
{
    if (!b) return TDVMC_ERR_BAD_ARG;
    int prec, rc = batch_precision(b, precision, &prec);
    if (rc) return rc;
    if ((rc = batch_sync_members(b))) return rc;
    if ((rc = batch_upload(b, 0, b->seed, true))) return rc;
    if ((rc = batch_prep(b, prec))) return rc;
    const int n = (int)b->sys.size();
    // measurement: wavefronts per walker
    std::vector<int> best(n, 4);
    for (int s = 0; s < n; s++) {
        const tdvmc_ctx* c = b->sys[s];
        float tb = 0.0f;
        bool have = false;
        for (int wpw : {1, 2, 4}) {
            const size_t slot =
                std::max(eval_lds_bytes<float>(c->sy, false, wpw), eval_lds_bytes<double>(c->sy, false, wpw));
            if ((size_t)(4 / wpw) * slot > 160 * 1024) continue;
            float t = 0.0f;
            rc = prec == TDVMC_MIXED ? batch_time_eval<float>(b, s, wpw, b->dEvalF, b->dTabF, &t)
                                     : batch_time_eval<double>(b, s, wpw, b->dEvalD, b->dTabD, &t);
            if (rc) return rc;
            if (!have || t < tb) {
                tb = t;
                best[s] = wpw;
                have = true;
            }
        }
    }
    // sweep: lanes per walker of each system (batches without k-space terms
    // and without a shared setting), timed on scratch copies of the walkers
    std::vector<int> lanes(n, 0);
    if (batch_lanes(b) == 0) {
        size_t rmax = 0, wmax = 0;
        for (auto* c : b->sys) {
            rmax = std::max(rmax, (size_t)c->W * c->sy.N * c->sy.D);
            wmax = std::max(wmax, (size_t)c->W);
        }
        double* dRs = nullptr;
        unsigned long long* dAccS = nullptr;
        if (dalloc(&dRs, rmax) != hipSuccess || dalloc(&dAccS, wmax) != hipSuccess) {
            dfree(dRs);
            dfree(dAccS);
            b->err = "out of device memory";
            return TDVMC_ERR_HIP;
        }
        for (int s = 0; s < n && rc == TDVMC_OK; s++) {
            float tb = 0.0f;
            bool have = false;
            for (int l : {8, 16, 32, 64}) {
                const size_t lds = std::max(sweep_lds_bytes<float>(b->sys[s]->sy, l),
                                            sweep_lds_bytes<double>(b->sys[s]->sy, l));
                if (lds > 160 * 1024) continue;
                float t = 0.0f;
                rc = prec == TDVMC_MIXED ? batch_time_sweep<float>(b, s, l, b->dTabF, dRs, dAccS, &t)
                                         : batch_time_sweep<double>(b, s, l, b->dTabD, dRs, dAccS, &t);
                if (rc) break;
                if (!have || t < tb) {
                    tb = t;
                    lanes[s] = l;
                    have = true;
                }
            }
        }
        dfree(dRs);
        dfree(dAccS);
        if (rc) return rc;
        b->sweep_lanes = lanes;
    }
    b->eval_wpw = best;
    b->tune_gen++;
    batch_drop_graph(b);
    if ((rc = batch_upload(b, 0, b->seed, true))) return rc;
    for (int s = 0; s < n; s++) {
        if (waves_out) waves_out[s] = batch_eval_wpw(b, s);
        if (lanes_out) lanes_out[s] = batch_sys_lanes(b, s);
    }
    return TDVMC_OK;
}

int tdvmc_batch_sweep(tdvmc_batch* b, int64_t n_moves, uint64_t seed, int precision)
{
    if (!b) return TDVMC_ERR_BAD_ARG;
    int prec, rc = batch_precision(b, precision, &prec);
    if (rc) return rc;
    if ((rc = batch_sync_members(b))) return rc;
    if ((rc = batch_upload(b, n_moves, seed, true))) return rc;
    if ((rc = batch_prep(b, prec))) return rc;
    if ((rc = batch_sweep_launch(b, prec))) return rc;
    BHIPC(hipStreamSynchronize(b->stream));
    for (auto* c : b->sys) {
        const int64_t m = n_moves > 0 ? n_moves : c->therm_moves;
        c->trials_total += m * c->W;
    }
    return TDVMC_OK;
}

int tdvmc_batch_measure(tdvmc_batch* b, int precision, tdvmc_step_data* out)
{
    if (!b) return TDVMC_ERR_BAD_ARG;
    int prec, rc = batch_precision(b, precision, &prec);
    if (rc) return rc;
    if ((rc = batch_sync_members(b))) return rc;
    if ((rc = batch_upload(b, 0, b->seed, false))) return rc;
    if ((rc = batch_prep(b, prec))) return rc;
    if ((rc = batch_measure_launch(b, prec))) return rc;
    BHIPC(hipStreamSynchronize(b->stream));
    if (out)
        for (size_t s = 0; s < b->sys.size(); s++) {
            rc = fill_step_data(b->sys[s], b->sys[s]->W, &out[s]);
            if (rc) {
                b->err = b->sys[s]->err;
                return rc;
            }
        }
    return TDVMC_OK;
}

int tdvmc_batch_tdvp_step(tdvmc_batch* b, double dt, int ode_type, int precision, int use_graph)
{
    if (!b) return TDVMC_ERR_BAD_ARG;
    int prec, rc = batch_precision(b, precision, &prec);
    if (rc) return rc;
    const int ode = ode_type < 0 ? b->sys[0]->ode_type : ode_type;
    if (ode != 0 && ode != 1 && ode != 2 && ode != 3 && ode != 5) {
        b->err = "ODE type must be 0 (Euler), 1 (predictor-corrector), 2 (RK4), 3 (RK4 reusing samples) or 5 "
                 "(predictor-corrector reusing samples)";
        return TDVMC_ERR_BAD_ARG;
    }
    // parameters set through the member contexts live in their packed vectors,
    // which the step gathers first
    if ((rc = batch_sync_members(b))) return rc;
    // descriptors current with the members (a direct sweep or measurement, a
    // new walker count or a quench since the last step); drops a captured
    // step whose launch shapes no longer fit
    if ((rc = batch_upload(b, 0, b->seed, true))) return rc;
    if (use_graph) {
        char key[160];
        std::snprintf(key, sizeof(key), "%d/%d/%.17g/%d", prec, ode, dt, b->lanes);
        if (!b->graph || b->graph_key != key) {
            batch_drop_graph(b);
            BHIPC(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = batch_step_enqueue(b, prec, ode, dt);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            BHIPC(e);
            b->graph_src = g;
            BHIPC(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
            b->graph_key = key;
        }
        BHIPC(hipGraphLaunch(b->graph, b->stream));
    } else {
        if ((rc = batch_step_enqueue(b, prec, ode, dt))) return rc;
    }
    BHIPC(hipStreamSynchronize(b->stream));
    for (auto* c : b->sys) {
        c->tf.dirty = c->td.dirty = true;
        c->host_params_stale = true;
        c->time += dt;
        c->step += 1;
        c->trials_total += fresh_stages(ode) * c->therm_moves * c->W;
        resolve_potential(c);
    }
    return TDVMC_OK;
}

}  // extern "C"
