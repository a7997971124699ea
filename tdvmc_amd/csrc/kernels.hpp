// Host-side launchers of the TDVMC kernels (all enqueue on the given stream).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.hpp"

namespace tdvmc {

// ---- parameter-derived tables (kernels_misc.hip)
template <typename T>
hipError_t launch_prep_tables(const KSys& sy, const double* uR, const double* uI, const int* kmap, const T* kv, T* spl,
                              T* dspl, T* alpha, T* kag, const double* nubeta, hipStream_t st);

// ---- evaluation (kernels_eval.hip); grad/lap non-null selects the per-particle variant
// LDS of one walker with wpw wavefronts (single contexts: 4)
template <typename T>
size_t eval_lds_bytes(const KSys& sy, bool per_particle, int wpw = 4);
// rho_in: [W][K][2] rho_k of the walkers as left by the sweep (skips the
// rho_k pass), or null to compute it from the coordinates
template <typename T, typename TO>
hipError_t launch_eval(const KSys& sy, const KTables<T>& tb, const double* R, int64_t W, double* expo, double* ER,
                       double* EI, TO* O, double* other, double* grad, double* lap, double* outer,
                       const T* rho_in, hipStream_t st, double* bsum = nullptr);

// ---- Metropolis sweeps (kernels_sweep.hip)
struct SweepArgs {
    double* R;                 // [W][N][D] in/out
    int64_t W;
    int64_t n_moves;           // single-particle trial moves per walker
    uint64_t seed;
    const uint32_t* epoch;     // device counter selecting the random stream of this call
    unsigned long long* acc;   // [W] accepted moves (accumulated)
    int lanes_per_walker;      // 16, 32 or 64 (0 = choose); 8 also in batches without k-space terms
    void* rho_out;             // [W][K][2] (T) final rho_k of every walker, or null
};
template <typename T>
hipError_t launch_sweep(const KSys& sy, const KTables<T>& tb, const SweepArgs& a, hipStream_t st);

// A batch of independent systems in one launch (device arrays, n systems;
// block_start [n] is the first workgroup of each system, ascending).  All
// systems share DIM and the box kind.  lanes > 0: every system uses that many
// lanes per walker; lanes = 0 (batches without k-space terms): each system's
// SweepArgs::lanes_per_walker (8, 16, 32 or 64).
template <typename T>
struct SweepBatch {
    int n;
    const int64_t* block_start;
    const KSys* sys;
    const KTables<T>* tab;
    const SweepArgs* args;
};
template <typename T>
hipError_t launch_sweep_batch(const SweepBatch<T>& bt, int D, bool periodic, int lanes, int kh_max, int64_t blocks,
                              size_t lds, hipStream_t st);
int sweep_walkers_per_block(int lanes);

// Measurement of a batch of systems in one launch (1, 2 or 4 wavefronts per
// walker, chosen per system); O is written in FP64, no per-particle outputs.
template <typename T>
struct EvalBatchItem {
    const double* R;           // [W][N][D]
    int64_t W;
    double* expo;
    double* ER;
    double* EI;
    double* O;                 // [W][P]
    double* other;             // [W][9] or null
    double* outer;             // [W] or null
    const T* rho;              // rho_k handed over by the sweep, or null
    int wpw;                   // wavefronts per walker: 1, 2 or 4 (4 / wpw walkers per workgroup)
    int64_t slot_bytes;        // LDS per walker (eval_lds_bytes)
};
// wavefronts per walker of the measurement without tuning
int eval_default_wpw(int N);
template <typename T>
struct EvalBatch {
    int n;
    const int64_t* block_start;
    const KSys* sys;
    const KTables<T>* tab;
    const EvalBatchItem<T>* item;
};
template <typename T>
hipError_t launch_eval_batch(const EvalBatch<T>& bt, int D, bool periodic, int64_t blocks, size_t lds, hipStream_t st);
int sweep_default_lanes(int N, int Kh);
template <typename T>
size_t sweep_lds_bytes(const KSys& sy, int lanes_per_walker);

// ---- moments (kernels_moments.hip)
// O [W][P] (float or double), E_R, E_I [W] (double) ->
//   mean[P+2] = (<O>, <E_R>, <E_I>), S [P][P] = cov(O, O), F [2][P] = cov(O, E_R), cov(O, E_I)
// mode 0: FP64 (MFMA f64); mode 1: split-bf16 (bf16x6) products on MFMA with FP32
// accumulation over 256-walker slices and FP64 reduction across slices.
struct MomentsWork {
    void* xt;                  // packed centred operands (size from moments_work_bytes)
    double* partial;           // per-slice partial tiles
    double* colsum;            // per-block column sums
    size_t xt_bytes, partial_bytes, colsum_bytes;
    // optional fold of identical O columns (null cols: every column computed):
    // computed column c reads O column cols[c] (c < P_cols) and its results
    // go to the full indices outmap[2c], outmap[2c + 1] (-1: none)
    const int* cols = nullptr;
    const int* outmap = nullptr;
    int P_cols = 0;
    double* mean_cols = nullptr;   // [P_cols + 2] means of the computed columns
};
void moments_work_bytes(int64_t W, int P, int mode, size_t* xt, size_t* partial, size_t* colsum);
template <typename TO>
hipError_t launch_moments(const TO* O, const double* ER, const double* EI, int64_t W, int P, int mode,
                          const MomentsWork& wk, double* mean, double* S, double* F, hipStream_t st);

// ---- TDVP linear system, Cholesky and parameter update (kernels_solve.hip)
struct SolveArgs {
    int P;
    int imag_time;
    int precondition;
    double eps;
    const double* mean;        // [P+2]
    const double* S;           // [P][P]
    const double* F;           // [2][P]
    double* work;              // [P][P] scratch (factor), + 4P
    double* xdot;              // [2P+2] out: alpha-dot R, alpha-dot I, phi-dot R, phi-dot I
    int* status;               // [1] count of non-positive pivots (added to)
    int mean_sub;              // 1: subtract the mean of each solution vector (solver type 1)
};
hipError_t launch_solve(const SolveArgs& a, hipStream_t st);
// largest P the solve kernels take: Cholesky (solver type 0; LDS panel of
// NB = 16 columns) or full-pivot QR (type 1)
int solve_max_params(int solver_type);
// one workgroup per system; args_dev [n] in device memory
hipError_t launch_solve_batch(const SolveArgs* args_dev, int n, int P_max, hipStream_t st);

// ---- batches of systems (kernels_misc.hip)
template <typename T>
struct BatchParams {
    KSys sy;
    double* uR;                // [P + 1] this system's parameters, then phi R (device)
    double* uI;
    const int* kmap;
    const T* kv;
    T* spl;
    T* dspl;
    T* alpha;
    T* kag;
    double* xsys;              // [2P+2] this system's packed parameters
    int64_t xoff;              // offset of this system in the concatenated vector
};
template <typename T>
hipError_t launch_prep_tables_batch(const BatchParams<T>* bp_dev, int n, hipStream_t st);
template <typename T>
hipError_t launch_unpack_batch(const BatchParams<T>* bp_dev, int n, const double* x, hipStream_t st);
template <typename T>
hipError_t launch_gather_batch(const BatchParams<T>* bp_dev, int n, double* x, int to_batch, hipStream_t st);
struct MomentsItem {
    int P;
    int64_t W;
    const double* O;           // [W][P]
    const double* ER;
    const double* EI;
    double* mean;              // [P+2]
    double* S;                 // [P][P]
    double* F;                 // [2][P]
    double* part;              // [nchunk][moments_part_len(P)] partial sums (scratch)
    int cw;                    // walkers per chunk
    int nchunk;
};
// walkers per chunk of the batched moments (the chunk's [O | E_R | E_I] rows fit 48 KB of LDS)
int moments_chunk_walkers(int P);
// doubles of one chunk's partial sums: P + 2 column sums, P (P + 1) / 2 + 2 P products
__host__ __device__ inline int64_t moments_part_len(int P)
{
    return (int64_t)(P + 2) + (int64_t)P * (P + 1) / 2 + 2 * (int64_t)P;
}
// lds: largest (P + 2) (1 + cw) * 8 over the systems
hipError_t launch_moments_batch(const MomentsItem* items_dev, int n, int nchunk_max, size_t lds, int P_max,
                                hipStream_t st);

// after a step: alpha[off .. off+n) -= alpha[off+n-1] in both parts (adjust: the
// pair-spline block), phi R = -exponent (normalize)
hipError_t launch_post_step(int P, double* x, const double* expo, int normalize, int adjust, int off, int n,
                            hipStream_t st);
// y = x0 + c * xdot  over the 2P+2 entries (alpha R, alpha I, phi R, phi I)
hipError_t launch_axpy_params(int P, const double* x0, const double* xdot, double c, double* y, hipStream_t st);
// x += dt/6 (k1 + 2 k2 + 2 k3 + k4)
hipError_t launch_rk4_combine(int P, double* x, const double* k1, const double* k2, const double* k3,
                              const double* k4, double dt, hipStream_t st);
// params (2P+2 packed) -> uR, uI arrays used by the tables
// y = x + (k1 + k2) dt / 2 on the packed parameter vector [2P+2]
hipError_t launch_pc_combine(int P, const double* x, const double* k1, const double* k2, double dt, double* y,
                             hipStream_t st);
hipError_t launch_unpack_params(int P, const double* packed, double* uR, double* uI, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);
hipError_t launch_add_u32(uint32_t* p, uint32_t v, hipStream_t st);

// ---- adjoint (kernels_misc.hip): dalpha[P] = sum_w gbar_w O_w ; dR = gbar_w grad_R ln psi
hipError_t launch_adjoint_reduce(const double* O, const double* gbar, int64_t W, int P, double* dalpha,
                                 hipStream_t st);
hipError_t launch_adjoint_scale(const double* grad, const double* gbar, int64_t W, int ND, double* dR,
                                hipStream_t st);
// dalpha[idx] += sum_w gbar_w outer_w (the outer-pair term of the last pair parameter)
hipError_t launch_adjoint_outer(const double* outer, const double* gbar, int64_t W, int idx, double* dalpha,
                                hipStream_t st);

// Pair distribution and structure factor of every walker (kernels_obs.hip).
struct ObsArgs {
    const double* R;                // [W][N][D] FP64 walkers
    int64_t W;
    int N;
    double L, Lr;                   // box length and 1 / L
    double gmax, spacing;           // g(r) grid end and bin width
    int count;                      // g(r) bins
    const double* scale;            // [count] (D / (N-1)) / scaling[bin]
    unsigned long long* gr_tot;     // [count] pair counts summed over walkers (atomics)
    double* gr_w;                   // [W][count] per-walker g(r), or null
    const int* kvec;                // [V][4] integer wave vectors n (components 0..D-1)
    const int* shell_off;           // [nshell + 1] first vector of each shell
    int V, nshell, nmax;            // vectors, shells, largest component
    int chunk;                      // particles per LDS phase table
    int gr_segs;                    // segments per row of the pair loop
    double* sk_w;                   // [W][nshell] per-walker S(k)
};
size_t obs_lds_bytes(int N, int D, int V, int chunk, int nmax, int count, bool fp32);
template <typename T>
hipError_t launch_obs(const ObsArgs& a, int D, size_t lds, hipStream_t st);
// acc[s] += sum_w sk_w[w][s] (fixed order)
hipError_t launch_obs_reduce(const double* sk_w, int64_t W, int nshell, double* acc, hipStream_t st);

}  // namespace tdvmc
