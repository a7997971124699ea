/*
 * tdvmc.h -- C ABI of the MI355X (gfx950) TDVMC hot path: batched Metropolis
 * sampling of spline-Jastrow (+ k-space, + one-body) bosonic wavefunctions,
 * local energies, log-derivatives O_k, TDVP moments S / F, the FP64
 * regularised solve and the ODE step.
 *
 * The reference exposes this path as the in-process C++ plug-in interface
 * PhysicalSystems::IPhysicalSystem (src/PhysicalSystems/IPhysicalSystem.h:98-122)
 * driven per walker by DoMetropolisStep (src/TDVMC.cpp:858-916) and
 * UpdateExpectationValues (src/TDVMC.cpp:1038-1150); its per-move granularity
 * is too fine for an ABI crossing, so this boundary sits at the batch level:
 * one call sweeps / measures every walker of a context.  Each entry point
 * below names the reference interface it replaces.
 *
 * Conventions
 *   - every function returns TDVMC_OK (0) or an error code; the message of the
 *     last failure is tdvmc_last_error(ctx) (ctx may be NULL for failures of
 *     tdvmc_create);
 *   - host arrays are caller-owned, row-major, FP64;  R is [W][N][DIM];
 *     O is [W][P];  S is [P][P];
 *   - parameter vector layout: [one-body (0 or 1)][pair spline N_PARAM][k-space K];
 *   - one context owns its device buffers and one HIP stream; calls on one
 *     context are not thread-safe, distinct contexts may be used concurrently;
 *     calls returning host data synchronise the context's stream;
 *   - precision: TDVMC_FP64 (plain form) or TDVMC_MIXED (FP32 pair / spline /
 *     phase arithmetic, split-bf16 (bf16x6) MFMA moments, FP64 solve); TDVMC_PRECISION_DEFAULT
 *     selects the context's configured precision.
 */
#ifndef TDVMC_H
#define TDVMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDVMC_OK 0
#define TDVMC_ERR_BAD_ARG 1
#define TDVMC_ERR_BAD_CONFIG 2
#define TDVMC_ERR_HIP 3
#define TDVMC_ERR_NOT_PD 4
#define TDVMC_ERR_NONFINITE 5
#define TDVMC_ERR_STATE 6
#define TDVMC_ERR_REJECTED 7

#define TDVMC_PRECISION_DEFAULT (-1)
#define TDVMC_FP64 0
#define TDVMC_MIXED 1

typedef struct tdvmc_ctx tdvmc_ctx;

typedef struct {
    int N;                 /* particles per walker */
    int DIM;
    int P;                 /* total parameters */
    int P_pair;            /* pair-spline parameters (config N_PARAM) */
    int K;                 /* k-space terms */
    int one_body;          /* 1 if parameter 0 is the one-body width parameter */
    int64_t n_walkers;
    double L;              /* box length (0: open space) */
    double rc;             /* spline cutoff */
    double hbar2_2m;
    int imaginary_time;
    int ode_solver_type;
    int precision;
    int samples_per_walker;
    double timestep;
    double mc_step;
    int64_t therm_moves;   /* MC_NTHERMSTEPS: trial moves between measurements */
    int64_t init_moves;    /* MC_NINITIALIZATIONSTEPS: trial moves at the start of a stage */
    double time;
    double reg_eps;
    int precondition;
    int K_terms;           /* k-space terms the kernels evaluate: one per +-k pair, plus unpaired k */
    int gr_bin_count;      /* GR_BIN_COUNT: g(r) bins requested */
    int use_nurbs;         /* USE_NURBS: pair spline on the knots of NURBS_GRID (rc = last knot) */
    int bc_origin;         /* SPLINE_BC_ORIGIN: boundary-condition family at r = 0 (1..4) */
    int bc_cutoff;         /* SPLINE_BC_CUTOFF: boundary-condition family at r = rc (1..4) */
} tdvmc_info;

/* Moments of one sampling stage; fields mirror SimulationStepData
 * (src/SimulationStepData.h:7-31).  Array pointers are caller-owned and may be
 * NULL when not wanted. */
typedef struct {
    double time;
    int step;
    int64_t n_samples;
    double localEnergyR;                 /* <E_R> */
    double localEnergyI;                 /* <E_I> */
    double* localOperators;              /* [P]   <O_k> */
    double* localOperatorsMatrix;        /* [P*P] <O_k O_j> */
    double* localOperatorlocalEnergyR;   /* [P]   <O_k E_R> */
    double* localOperatorlocalEnergyI;   /* [P]   <O_k E_I> */
    double* S;                           /* [P*P] <O O^T> - <O><O>^T */
    double* FR;                          /* [P]   <O E_R> - <O><E_R> */
    double* FI;                          /* [P]   <O E_I> - <O><E_I> */
    double* otherExpectationValues;      /* [9]   order of BosonsBulk.cpp:449-457 */
} tdvmc_step_data;

/* Replaces ReadConfig + InitializePhysicalSystem + IPhysicalSystem::InitSystem
 * (src/TDVMC.cpp:351-375, :2786-2998; src/PhysicalSystems/BosonsBulk.cpp:51-156).
 * config: the text of a config/NAME.config JSON document or a path to one. */
int tdvmc_create(const char* config, tdvmc_ctx** out);
int tdvmc_destroy(tdvmc_ctx* ctx);
/* Host-only: parse and validate a config without touching the GPU (the
 * checks of ReadConfig / BosonsBulk::InitSystem, src/TDVMC.cpp:351-375 and
 * src/PhysicalSystems/BosonsBulk.cpp:85-91); info receives the derived shape. */
int tdvmc_parse_config(const char* config, tdvmc_info* info);
const char* tdvmc_last_error(const tdvmc_ctx* ctx);
int tdvmc_get_info(tdvmc_ctx* ctx, tdvmc_info* info);

/* Number of walkers (one walker per MPI rank in the reference, src/TDVMC.cpp:3082). */
int tdvmc_set_walker_count(tdvmc_ctx* ctx, int64_t W);

/* Replaces BroadcastNewParameters (src/TDVMC.cpp:506-512): alpha in. */
int tdvmc_set_params(tdvmc_ctx* ctx, const double* uR, const double* uI, double phiR, double phiI, int P);
int tdvmc_get_params(tdvmc_ctx* ctx, double* uR, double* uI, double* phiR, double* phiI, int P);

/* Replaces IPhysicalSystem::SetTime (src/PhysicalSystems/IPhysicalSystem.h:77):
 * selects the quenched pair potential (src/PhysicalSystems/BosonsBulk.cpp:236-243). */
int tdvmc_set_time(tdvmc_ctx* ctx, double t);

/* Walker coordinates [W][N][DIM] in / out (replaces the per-rank R of mainMPI). */
int tdvmc_set_walkers(tdvmc_ctx* ctx, const double* R, int64_t W);
int tdvmc_get_walkers(tdvmc_ctx* ctx, double* R, int64_t W);
/* Replaces InitCoordinateConfiguration (src/TDVMC.cpp:672-785): jittered
 * lattice in the box, or a Gaussian cloud for open systems. */
int tdvmc_init_walkers(tdvmc_ctx* ctx, uint64_t seed);

/* Replaces DoMetropolisSteps (src/TDVMC.cpp:918-924) for every walker:
 * n_moves single-particle trial moves per walker. */
int tdvmc_sweep(tdvmc_ctx* ctx, int64_t n_moves, uint64_t seed, int precision);
int tdvmc_get_acceptance(tdvmc_ctx* ctx, int64_t* accepted, int64_t* trials);

/* Replaces UpdateExpectationValues + ParallelUpdateExpectationValues
 * (src/TDVMC.cpp:1038-1220): init_moves, then samples_per_walker x
 * (therm_moves + one measurement) for every walker, then the moments. */
int tdvmc_sample(tdvmc_ctx* ctx, uint64_t seed, int precision, tdvmc_step_data* out);

/* Measurement of the current walkers only (IPhysicalSystem::CalculateExpectationValues,
 * src/PhysicalSystems/BosonsBulk.cpp:460-466, batched) followed by the moments. */
int tdvmc_measure(tdvmc_ctx* ctx, int precision, tdvmc_step_data* out);

/* Parity entry on fed configurations (IPhysicalSystem::CalculateWavefunction +
 * CalculateExpectationValues, src/PhysicalSystems/BosonsBulk.cpp:179-539).
 * Outputs may be NULL: expo [W], ER [W], EI [W], O [W][P], grad [W][N][DIM][2]
 * (Re, Im of grad ln psi), lap [W][N][2], other [W][9]. */
int tdvmc_eval_configs(tdvmc_ctx* ctx, const double* R, int64_t W, int precision, double* expo, double* ER,
                       double* EI, double* O, double* grad, double* lap, double* other);

/* Basis sums B_s = sum over pairs r_ij <= rc of B_s(r_ij), s = 0 .. P_pair + 1
 * (the cubic B-splines on the knot vector, before the boundary-condition map
 * that gives O), of W fed configurations R [W][N][DIM]; B [W][P_pair + 2]
 * (host memory).  BosonsBulk::splineSums after CalculateLocalOperators
 * (src/PhysicalSystems/BosonsBulk.cpp:179-218); tdvmc_splined_function.hpp
 * exposes them as WFParts::SplinedFunction::splineSums. */
int tdvmc_basis_sums(tdvmc_ctx* ctx, const double* R, int64_t W, int precision, double* B);

/* Boundary-condition factors of the pair spline: WFParts::SplinedFunction::
 * bcFactorsStart / bcFactorsEnd (src/WFParts/SplinedFunction.h:10-40) as
 * BosonsBulk::InitSystem fills them (src/PhysicalSystems/BosonsBulk.cpp:75-82,
 * SetBoundaryConditions3_1D_OR_k / _CO_k of src/SplineFactory.cpp:377-656 with
 * uniform = true; k = SPLINE_BC_ORIGIN / SPLINE_BC_CUTOFF, 2 by default).
 * start and end receive [rows][3] (room for 9 doubles each), n_start and
 * n_end the row counts; O_i = start[i] . (B_0, B_1, B_2) for i < n_start,
 * O_{P_pair-n_end+i} = end[i] . (B_{S-3}, B_{S-2}, B_{S-1}), S = P_pair + 2. */
int tdvmc_spline_bc(tdvmc_ctx* ctx, double* start, int* n_start, double* end, int* n_end);

/* Device-pointer variant of tdvmc_eval_configs (no host copies; enqueued on
 * `stream` (a hipStream_t) or the context stream when NULL; returns without
 * synchronising).  O_dev is FP64 [W][P]. */
int tdvmc_eval_configs_dev(tdvmc_ctx* ctx, const double* R_dev, int64_t W, int precision, double* expo_dev,
                           double* ER_dev, double* EI_dev, double* O_dev, double* grad_dev, double* lap_dev,
                           void* stream);

/* Device-pointer variants of tdvmc_sample / tdvmc_measure and of the walker
 * exchange, for zero-copy use from device frameworks (torch tensors):
 *   tdvmc_sample_dev / tdvmc_measure_dev - the same stage as tdvmc_sample /
 *       tdvmc_measure (UpdateExpectationValues, src/TDVMC.cpp:1038-1220),
 *       leaving mean_dev [P+2] = (<O>, <E_R>, <E_I>), S_dev [P][P] and
 *       F_dev [2][P] = (cov(O, E_R), cov(O, E_I)) in device memory (any may
 *       be NULL);
 *   tdvmc_set_walkers_dev / tdvmc_get_walkers_dev - R_dev [W][N][DIM] in
 *       device memory, W equal to the context's walker count (no reallocation).
 * The work runs on the context stream, ordered after what `stream` (a
 * hipStream_t; NULL: none) holds at the call, and `stream` then waits for it;
 * the calls return without synchronising the host. */
int tdvmc_sample_dev(tdvmc_ctx* ctx, uint64_t seed, int precision, double* mean_dev, double* S_dev, double* F_dev,
                     void* stream);
int tdvmc_measure_dev(tdvmc_ctx* ctx, int precision, double* mean_dev, double* S_dev, double* F_dev, void* stream);
int tdvmc_set_walkers_dev(tdvmc_ctx* ctx, const double* R_dev, int64_t W, void* stream);
int tdvmc_get_walkers_dev(tdvmc_ctx* ctx, double* R_dev, int64_t W, void* stream);

/* Pair distribution g(r) and structure factor S(k), the additional
 * observables of BosonsBulk (periodic boxes only):
 *   grid     - src/PhysicalSystems/BosonsBulk.cpp:142-151 with Grid::Init and
 *              InitGridCenterPoints (src/Observables/Grid.cpp:17-49): GR_BIN_COUNT
 *              bins requested up to the last knot, n_gr = the grid's count;
 *              S(k) on the first 50 shells of config/kVectors<D>D.json
 *              (:55, :124-137), sk_grid = |k| of each shell;
 *   per configuration - BosonsBulk::CalculateAdditionalSystemProperties
 *              (src/PhysicalSystems/BosonsBulk.cpp:474-520);
 *   averaged - CalculateAdditionalSystemProperties + ParallelCalculateAdditional-
 *              SystemProperties (src/TDVMC.cpp:1332-1365, :1438-1444): the
 *              context's walkers do MC_NADDITIONALINITIALIZATIONSTEPS moves,
 *              then n_samples x (MC_NADDITIONALTHERMSTEPS moves + one
 *              observation); g(r) and S(k) averaged over samples and walkers.
 *              n_samples <= 0 takes MC_NADDITIONALSTEPS (the driver's end-of-run
 *              call multiplies it by MC_NFINALSTEPS_MULTIPLICATOR, :3997-3999).
 * Outputs: gr_grid [n_gr] bin centres, sk_grid [n_sk]; gr [W][n_gr], sk [W][n_sk]
 * per configuration; gr [n_gr], sk [n_sk] averaged.  Any output may be NULL. */
int tdvmc_observables_info(tdvmc_ctx* ctx, int* n_gr, int* n_sk, double* gr_grid, double* sk_grid);
int tdvmc_observables_configs(tdvmc_ctx* ctx, const double* R, int64_t W, int precision, double* gr, double* sk);
int tdvmc_additional_properties(tdvmc_ctx* ctx, int64_t n_samples, uint64_t seed, int precision, double* gr,
                                double* sk);

/* Output files of the reference driver, byte for byte in its format (one
 * ".dat" file per series in out_dir, which ends with the path separator as
 * OUT_DIR does; numbers in scientific notation, 15 digits, width 24:
 * src/Utils.h:25-28, WriteDataToFile / AppendDataToFile, src/Utils.cpp:818-1016).
 * Host only; no context needed.
 *   tdvmc_files_init          - ParametersR0 / ParametersI0 with the start
 *                               parameters and the series files with their headers
 *                               (src/TDVMC.cpp:3117-3129);
 *   tdvmc_files_append_step   - one line each for timesSystem, LocalEnergyR,
 *                               LocalEnergyI, LocalOperators [P], OtherExpectationValues
 *                               [9], ParametersR [P], ParametersI [P] (src/TDVMC.cpp:3833-3839);
 *   tdvmc_files_write_grids   - gr_grid / sk_grid (src/TDVMC.cpp:3151-3157);
 *   tdvmc_files_append_observables - timesAdditional, gr, sk (src/TDVMC.cpp:3475, :3496-3500). */
int tdvmc_files_init(const char* out_dir, const double* uR, const double* uI, int P);
int tdvmc_files_append_step(const char* out_dir, double time, const tdvmc_step_data* d, const double* uR,
                            const double* uI, int P);
int tdvmc_files_write_grids(const char* out_dir, const double* gr_grid, int n_gr, const double* sk_grid, int n_sk);
int tdvmc_files_append_observables(const char* out_dir, double time, const double* gr, int n_gr, const double* sk,
                                   int n_sk);
/* Resume files of a successive run.
 *   tdvmc_files_write_coordinates - the walkers' coordinates as the driver
 *       leaves them for the next run (src/TDVMC.cpp:4007-4022): moved into
 *       the first cell when L > 0 (AlignCoordinates, :2569-2582, with
 *       GetCoordinateNIC), one file coords/particleconfiguration_<N>_<DIM>D_<w>.dat
 *       per walker w (the reference writes one per MPI rank) holding one line
 *       of N*DIM values in the .dat number format (WriteParticleInputFile,
 *       :493-504), and AAFinish_particleconfiguration_<N>_<DIM>D.dat with
 *       walker 0 (the root rank's copy); out_dir/coords is created;
 *   tdvmc_files_read_coordinates - LoadLastPositionsFromFile (:640-670): the
 *       last line of `path`, comma-separated (the config/particleconfiguration_*.csv
 *       files), N*DIM coordinates into R [N][DIM];
 *   tdvmc_files_write_rng_state / tdvmc_files_read_rng_state - the random
 *       stream's state in random/state_philox_<rank>.dat.  The reference
 *       saves its mt19937_64 and distribution states there
 *       (WriteRandomGeneratorStatesToFile, src/Utils.cpp:64-78, :3956-3964);
 *       here the whole stream is fixed by (seed, epoch) of the counter-based
 *       Philox generator (tdvmc_get_rng_state / tdvmc_set_rng_state), so that
 *       pair is what the file holds.
 * Host only; no context needed. */
int tdvmc_files_write_coordinates(const char* out_dir, const double* R, int64_t W, int N, int D, double L);
int tdvmc_files_read_coordinates(const char* path, double* R, int N, int D);
int tdvmc_files_write_rng_state(const char* out_dir, uint64_t seed, uint32_t epoch, int rank);
int tdvmc_files_read_rng_state(const char* path, uint64_t* seed, uint32_t* epoch);

/* State of the context's random stream: the seed its TDVP steps draw with
 * (SEED) and the call epoch (advanced by every sweep launch).  Setting both
 * continues a stream exactly where a saved one stopped. */
int tdvmc_get_rng_state(tdvmc_ctx* ctx, uint64_t* seed, uint32_t* epoch);
int tdvmc_set_rng_state(tdvmc_ctx* ctx, uint64_t seed, uint32_t epoch);

/* The config for a successive run, as WriteConfig writes AAFinish_tVMC.config
 * (src/TDVMC.cpp:387-410, :4026-4034): the items of `config` (text or path)
 * with the final parameters in PARAMS_REAL / PARAMS_IMAGINARY [P] and
 * PARAM_PHIR / PARAM_PHII, in the reference's order and value format, then
 * this library's keys (which the reference's reader skips).  Host only. */
int tdvmc_write_final_config(const char* config, const double* uR, const double* uI, double phiR, double phiI, int P,
                             const char* path);

/* TDVP moments of given samples (the accumulation of src/TDVMC.cpp:1103-1109
 * and the centring of :1506-1537).  mean [P+2] = (<O>, <E_R>, <E_I>),
 * S [P][P], F [2][P] = (cov(O, E_R), cov(O, E_I)). */
int tdvmc_moments(tdvmc_ctx* ctx, const double* O, const double* ER, const double* EI, int64_t W, int P,
                  int precision, double* mean, double* S, double* F);

/* SolveForParametersDot (src/TDVMC.cpp:1713-1828): solver type 0 (scaling,
 * S += eps I, Cholesky, :1733-1762) or type 1 (:1763-1822: S += 0.002 I only
 * with the scaling, solve, subtract the mean of each solution; eps is
 * ignored).  xdot [2P+2] = (alpha-dot R, alpha-dot I, phi-dot R, phi-dot I);
 * *n_bad = count of non-positive Cholesky pivots (TDVMC_ERR_NOT_PD if > 0).
 * Sizes: P <= 1129 for type 0, P <= 19200 for type 1 (TDVMC_ERR_BAD_ARG
 * beyond; tdvmc_tdvp_step checks the configured type the same way, and a
 * batch takes systems up to P = 1129). */
int tdvmc_solve(tdvmc_ctx* ctx, const double* mean, const double* S, const double* F, int P, int imag_time,
                int precondition, double eps, int solver_type, double* xdot, int* n_bad);

/* One TDVP time step on the device (CalculateNextParameters, src/TDVMC.cpp:2454-2512):
 *   type 0  CalculateNextParametersEuler (:1834-1853);
 *   type 1  CalculateNextParametersPC (:1855-1910): predictor x + dt k1, one
 *           corrector x + dt/2 (k1 + k2) on fresh samples (chain continued);
 *   type 2  CalculateNextParametersRK4 (:1969-2035, fresh samples per stage,
 *           chain continued);
 *   type 3  CalculateNextParametersRK4ReuseSamples (:2037-2103): k2..k4 on the
 *           samples of k1 (UpdateExpectationValuesForGivenSamples, :1222-1303,
 *           unit weights as there);
 *   type 5  CalculateNextParametersPCReuseSamples (:1912-1967): six correctors
 *           on the samples of k1.
 * Types 4 and 6 (implicit Euler, Crank-Nicolson) return TDVMC_ERR_BAD_ARG.
 * ode_type < 0 uses the configured ODE_SOLVER_TYPE.  With use_graph != 0 the
 * step is captured once into a hipGraph and replayed.  After the step
 * USE_NORMALIZE_WF sets phi R = -exponent of walker 0's last sample and
 * USE_ADJUST_PARAMETERS subtracts the last pair-spline parameter from the
 * N_PARAM pair-spline parameters, real and imaginary parts alike (one-body and
 * k-space parameters are left as they are; src/TDVMC.cpp:3757-3775, :2525-2567). */
int tdvmc_tdvp_step(tdvmc_ctx* ctx, double dt, int ode_type, int precision, int use_graph);

/* Non-positive Cholesky pivots met by the solves of the last TDVP step (the
 * doNotAcceptStep flag of src/TDVMC.cpp:1580-1590).  With
 * USE_PARAMETER_ACCEPTANCE_CHECK = 1 tdvmc_tdvp_step itself applies
 * AcceptNewParams (src/TDVMC.cpp:2585-2640) and on rejection restores the
 * parameters, leaves the time unchanged and returns TDVMC_ERR_REJECTED. */
int tdvmc_get_step_status(tdvmc_ctx* ctx, int* n_bad);

/* Adjoint of Re ln psi over fed configurations: given gbar [W],
 * dalpha [P] = sum_w gbar_w d(Re ln psi_w)/d(alpha)  (= O_w, plus the count of
 * pairs beyond the cutoff for the last pair parameter, whose outer-pair term
 * src/PhysicalSystems/BosonsBulk.cpp:530-534 leaves out of O) and
 * dR [W][N][DIM] = gbar_w Re grad ln psi_w.
 * No counterpart in the reference (which has no differentiation API); it is
 * the backward of a torch.autograd.Function test wrapper. */
int tdvmc_adjoint(tdvmc_ctx* ctx, const double* R, int64_t W, const double* gbar, int precision, double* dR,
                  double* dalpha);

/* Allocation-free, device-pointer variant of tdvmc_adjoint: the same kernels,
 * with R_dev [W][N][DIM], gbar_dev [W], dR_dev [W][N][DIM] and dalpha_dev [P]
 * in device memory (either output may be NULL), enqueued on `stream` (a
 * hipStream_t; NULL: the context stream), returning without synchronising.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  It calls
 * no hipMalloc / hipFree; its scratch is one fixed 64 MiB device array per GPU,
 * so W * (P + 2 N DIM + 2 N + 4) * 8 bytes must fit in 64 MiB (larger W:
 * TDVMC_ERR_BAD_ARG).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other contexts, other threads) race. */
int tdvmc_adjoint_noalloc(tdvmc_ctx* ctx, const double* R_dev, int64_t W, const double* gbar_dev, int precision,
                          double* dR_dev, double* dalpha_dev, void* stream);

/* Kernel timing with HIP events on the context stream (Timer, src/Timer.h).
 * names: "sweep", "eval", "moments", "solve".  Accumulated since enable. */
int tdvmc_profile_enable(tdvmc_ctx* ctx, int enable);
int tdvmc_profile_get(tdvmc_ctx* ctx, const char* name, double* total_ms, int64_t* launches);

/* Lanes per walker of the sweep kernel (16, 32, 64; 0 = automatic). */
int tdvmc_set_sweep_lanes(tdvmc_ctx* ctx, int lanes_per_walker);

int tdvmc_synchronize(tdvmc_ctx* ctx);

/* Autotuner of the sweep's walker-block shape: times one sweep of a scratch
 * copy of up to 16384 walkers for 16, 32 and 64 lanes per walker (4, 8 or 16
 * walkers per 256-thread workgroup) and keeps the fastest for this precision
 * (the context's walkers and random stream are untouched).  A lane count set
 * with tdvmc_set_sweep_lanes takes precedence.  No counterpart in the
 * reference (its sampler is one walker per MPI rank). */
int tdvmc_autotune(tdvmc_ctx* ctx, int precision, int* lanes_out);

/* ---- Batches of independent systems (no counterpart in the reference,
 * which runs one system per process).  Each config describes one system with
 * its own particle count, parameters and walkers; all systems of a batch share
 * DIM and the box kind.  Per stage: one sweep launch and one measurement
 * launch for all systems, one moments and one solve launch with a workgroup
 * per system (FP64 moments), the RK arithmetic over all parameters at once.
 * Member contexts (tdvmc_batch_system) stay usable for parameters, walkers,
 * info and per-system evaluation; they are owned by the batch. */
typedef struct tdvmc_batch tdvmc_batch;
int tdvmc_batch_create(const char* const* configs, int n_systems, tdvmc_batch** out);
int tdvmc_batch_destroy(tdvmc_batch* b);
const char* tdvmc_batch_last_error(const tdvmc_batch* b);
int tdvmc_batch_size(tdvmc_batch* b, int* n_systems);
tdvmc_ctx* tdvmc_batch_system(tdvmc_batch* b, int s);
/* lanes per walker of the batched sweep, shared by every system: 16, 32, 64,
 * or 0 for the default (per system in batches without k-space terms) */
int tdvmc_batch_set_sweep_lanes(tdvmc_batch* b, int lanes);
/* Times each system's measurement with 1, 2 and 4 wavefronts per walker (4,
 * 2 or 1 walkers per workgroup) and, in batches without k-space terms and
 * without a shared lane setting, each system's sweep with 8, 16, 32 and 64
 * lanes per walker (one sweep of a scratch copy of its walkers); keeps the
 * fastest per system.  waves_out, lanes_out [n_systems] (may be NULL) receive
 * the choices.  No counterpart in the reference (the reference evaluates one
 * walker at a time). */
int tdvmc_batch_autotune(tdvmc_batch* b, int precision, int* waves_out, int* lanes_out);
/* n_moves <= 0: each system's MC_NTHERMSTEPS; system s draws from seed
 * mixed with s (walker w of system s follows the same random stream as walker
 * w of a single context seeded with that mix). */
int tdvmc_batch_sweep(tdvmc_batch* b, int64_t n_moves, uint64_t seed, int precision);
/* measurement + moments of every system; out [n_systems] (may be NULL) */
int tdvmc_batch_measure(tdvmc_batch* b, int precision, tdvmc_step_data* out);
/* one TDVP step of every system with its own samples: ODE types 0, 1, 2, 3
 * and 5 as in tdvmc_tdvp_step (types 3 and 5 measure the samples of k1
 * again in their later stages) */
int tdvmc_batch_tdvp_step(tdvmc_batch* b, double dt, int ode_type, int precision, int use_graph);

#ifdef __cplusplus
}
#endif

#endif /* TDVMC_H */
