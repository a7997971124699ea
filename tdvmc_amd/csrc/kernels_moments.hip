// TDVP moments of one sampling stage:
//   <O>, <E_R>, <E_I>, S = <O O^T> - <O><O>^T, F = <O E> - <O><E>
// (the sums of src/TDVMC.cpp:1103-1109 combined as in :1506-1537).
//
// The walker-by-parameter matrix X = [O | E_R | E_I] (W x Pa, Pa = P + 2) is
// centred column-wise (two passes: column means in FP64, then centring while
// packing), and S_aug = Xc^T Xc / W is a dense contraction over walkers:
//   mode 0 (FP64):  v_mfma_f64_16x16x4 on FP64 operands;
//   mode 1 (mixed): every centred entry (rounded to FP32) is split into
//                   three bf16 terms x = x0 + x1 + x2 (exact: 24 significant
//                   bits) and each product is formed from the split terms on
//                   v_mfma_f32_32x32x16_bf16: x0*y0 + x0*y1 + x1*y0 (the
//                   "bf16x3" products) + x1*y1 + x0*y2 + x2*y0 (together
//                   "bf16x6", Henry, Tang, Heinecke, ARITH 2019), leaving out
//                   terms of relative size 2^-24 and below.  With two-term
//                   splits (16 bits) the per-entry residual of ~2^-18 left
//                   F = <O E> - <O><E> 1.4e-6 off FP64 on sampled N = 343
//                   data, where O and E_L correlate weakly.  FP32
//                   accumulation is folded into FP64 registers every 256
//                   walkers, and the walker slices are reduced in FP64.
// Only tiles on or above the diagonal are computed; F is column Pa-2 / Pa-1.
//
// Folded columns (MomentsWork::cols): the contraction runs over a subset of
// O's columns (one column per group of identical ones, such as the two slots
// of a k / -k pair that share |rho_k|^2), and the results are written to
// every full index of the group (MomentsWork::outmap).  Rows and columns of
// such a group are then bit-identical in S and F, so the directions along
// which S is singular by construction stay exactly singular in both
// precisions (the Tikhonov term then holds them at eps).

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double double4v __attribute__((ext_vector_type(4)));

constexpr int kColBlock = 256;
constexpr int kChunkRows = 64;     // walkers per column-sum chunk (fewer only past kMaxChunks chunks)
constexpr int kMaxChunks = 1024;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct Geometry {
    int P, Pa, tile, Pp, nb, ntiles, groups;
    int64_t W, Wp, ksteps;   // Wp: padded walkers; ksteps: MFMA k-steps over Wp
    int chunks;
};

Geometry geometry(int64_t W, int P, int mode)
{
    Geometry g;
    g.P = P;
    g.Pa = P + 2;
    g.tile = mode == 0 ? 16 : 32;
    g.Pp = (int)ceil_div(g.Pa, g.tile) * g.tile;
    g.nb = g.Pp / g.tile;
    g.ntiles = g.nb * (g.nb + 1) / 2;
    const int kstep = mode == 0 ? 4 : 16;
    const int fold = 256;                               // walkers per FP32 fold (mode 1)
    g.W = W;
    g.Wp = ceil_div(W, fold) * fold;
    g.ksteps = g.Wp / kstep;
    int64_t groups = g.Wp / fold;
    const int64_t cap = mode == 0 ? 64 : 32;
    if (groups > cap) groups = cap;
    if (groups < 1) groups = 1;
    g.groups = (int)groups;
    int64_t ch = ceil_div(W, kChunkRows);
    g.chunks = (int)(ch > kMaxChunks ? kMaxChunks : (ch < 1 ? 1 : ch));
    return g;
}

// O column of computed column c (c < P): cols[c], or c; ld is O's row stride
struct ColSrc {
    const int* cols;
    int ld;
    __device__ __forceinline__ int operator()(int c) const { return cols ? cols[c] : c; }
};

// ---- pass 1: column sums of X in FP64, per row chunk
template <typename TO>
__global__ void colsum_kernel(const TO* __restrict__ O, const double* __restrict__ ER, const double* __restrict__ EI,
                              int64_t W, int P, ColSrc cs, int chunks, double* __restrict__ part)
{
    const int Pa = P + 2;
    const int c = blockIdx.x * kColBlock + threadIdx.x;
    const int ch = blockIdx.y;
    const int64_t rows = (W + chunks - 1) / chunks;
    const int64_t w0 = ch * rows, w1 = min(W, w0 + rows);
    if (c >= Pa) return;
    double s = 0.0;
    if (c < P) {
        const int oc = cs(c);
        // independent loads in flight; the sum keeps walker order
#pragma unroll 8
        for (int64_t w = w0; w < w1; w++) s += (double)O[w * cs.ld + oc];
    } else if (c == P)
        for (int64_t w = w0; w < w1; w++) s += ER[w];
    else
        for (int64_t w = w0; w < w1; w++) s += EI[w];
    part[(size_t)ch * Pa + c] = s;
}

// one workgroup of 256 threads per column: strided partial sums, then a fixed
// reduction tree (wavefront butterflies, then the 4 wavefronts in order)
__global__ __launch_bounds__(256) void colmean_kernel(const double* __restrict__ part, int chunks, int Pa, int64_t W,
                                                      double* __restrict__ mean, const int* __restrict__ outmap,
                                                      int P_full, double* __restrict__ mean_full)
{
    __shared__ double red[4];
    const int c = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int ch = tid; ch < chunks; ch += 256) s += part[(size_t)ch * Pa + c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        const double m = (((red[0] + red[1]) + red[2]) + red[3]) / (double)W;
        mean[c] = m;
        if (outmap) {
            // computed column c -> its full indices; E_R, E_I -> P_full, P_full + 1
            const int Pc = Pa - 2;
            if (c < Pc) {
                for (int k = 0; k < 2; k++)
                    if (outmap[2 * c + k] >= 0) mean_full[outmap[2 * c + k]] = m;
            } else {
                mean_full[P_full + (c - Pc)] = m;
            }
        }
    }
}

__device__ __forceinline__ double xval(const float* O, const double* ER, const double* EI, int64_t W, int P,
                                       ColSrc cs, const double* mean, int64_t w, int c)
{
    if (w >= W || c >= P + 2) return 0.0;
    const double v = c < P ? (double)O[w * cs.ld + cs(c)] : (c == P ? ER[w] : EI[w]);
    return v - mean[c];
}
__device__ __forceinline__ double xval(const double* O, const double* ER, const double* EI, int64_t W, int P,
                                       ColSrc cs, const double* mean, int64_t w, int c)
{
    if (w >= W || c >= P + 2) return 0.0;
    const double v = c < P ? O[w * cs.ld + cs(c)] : (c == P ? ER[w] : EI[w]);
    return v - mean[c];
}

// ---- pass 2a (mode 0): pack centred FP64 operands as [cb][s][kk][pp]
//      column c = 16 cb + pp, walker w = 4 s + kk
template <typename TO>
__global__ void pack64_kernel(const TO* __restrict__ O, const double* __restrict__ ER, const double* __restrict__ EI,
                              int64_t W, int P, ColSrc cs, const double* __restrict__ mean, int64_t Wp, int Pp,
                              double* __restrict__ xt)
{
    const int64_t total = Wp * Pp;
    const int64_t S4 = Wp / 4;
    for (int64_t idx = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; idx < total;
         idx += (int64_t)blockDim.x * gridDim.x) {
        // idx enumerates the packed layout; recover (w, c)
        const int pp = (int)(idx & 15);
        const int kk = (int)((idx >> 4) & 3);
        const int64_t rest = idx >> 6;
        const int64_t s = rest % S4;
        const int cb = (int)(rest / S4);
        const int64_t w = 4 * s + kk;
        const int c = 16 * cb + pp;
        xt[idx] = xval(O, ER, EI, W, P, cs, mean, w, c);
    }
}

// ---- pass 2b (mode 1): pack centred operands split into three bf16 terms as
//      [cb][s][part][h][l32][j], column c = 32 cb + l32, walker w = 16 s + 8 h + j
constexpr int kSplit = 3;
template <typename TO>
__global__ void pack_bf16_kernel(const TO* __restrict__ O, const double* __restrict__ ER, const double* __restrict__ EI,
                                 int64_t W, int P, ColSrc cs, const double* __restrict__ mean, int64_t Wp, int Pp,
                                 __bf16* __restrict__ xt)
{
    // one thread per (cb, s, h, l32): 8 walkers of one column
    const int64_t S16 = Wp / 16;
    const int64_t total = (int64_t)(Pp / 32) * S16 * 64;
    for (int64_t idx = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; idx < total;
         idx += (int64_t)blockDim.x * gridDim.x) {
        const int l32 = (int)(idx & 31);
        const int h = (int)((idx >> 5) & 1);
        const int64_t rest = idx >> 6;
        const int64_t s = rest % S16;
        const int cb = (int)(rest / S16);
        const int c = 32 * cb + l32;
        __bf16* base = xt + ((((size_t)cb * S16 + s) * kSplit) * 64) * 8;
        // 8 walkers of each term: one 16-byte store per term
        bf16x8 v[kSplit];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int64_t w = 16 * s + 8 * h + j;
            float x = (float)xval(O, ER, EI, W, P, cs, mean, w, c);
#pragma unroll
            for (int q = 0; q < kSplit; q++) {
                const __bf16 xq = (__bf16)x;
                v[q][j] = xq;
                x -= (float)xq;   // exact: the remainder has at most 16 significant bits
            }
        }
#pragma unroll
        for (int q = 0; q < kSplit; q++)
            *reinterpret_cast<bf16x8*>(base + ((size_t)q * 64 + h * 32 + l32) * 8) = v[q];
    }
}

__device__ __forceinline__ void tile_of(int t, int nb, int* I, int* J)
{
    // t enumerates the upper triangle row by row
    int i = 0;
    while (t >= nb - i) {
        t -= nb - i;
        i++;
    }
    *I = i;
    *J = i + t;
}

// Workgroup -> (tile t, walker group g).  Consecutive workgroups go to the 8
// XCDs in turn (workgroup b runs on XCD b mod 8), each with its own L2: the
// blocks of XCD x are given a contiguous range of the group-major order, so
// every tile of a walker group reads that group's rows through one L2 and
// the rows come from HBM about once instead of once per XCD.
__device__ __forceinline__ void tile_group_of(int b, int ntiles, int groups, int* t, int* g)
{
    const int nblk = ntiles * groups;
    const int x = b & 7, k = b >> 3;
    const int per = nblk >> 3, extra = nblk & 7;
    const int lin = x * per + min(x, extra) + k;
    *g = lin / ntiles;
    *t = lin - *g * ntiles;
}

// ---- pass 3a (mode 0): FP64 MFMA over walker slices, one wave per (tile, group)
__global__ __launch_bounds__(64) void gemm64_kernel(const double* __restrict__ xt, int64_t ksteps, int nb, int groups,
                                                    double* __restrict__ partial)
{
    const int lane = threadIdx.x, ntiles = nb * (nb + 1) / 2;
    int t, g, I, J;
    tile_group_of(blockIdx.x, ntiles, groups, &t, &g);
    tile_of(t, nb, &I, &J);
    const int64_t per = (ksteps + groups - 1) / groups;
    const int64_t s0 = g * per, s1 = min(ksteps, s0 + per);
    const double* A = xt + (size_t)I * ksteps * 64;
    const double* B = xt + (size_t)J * ksteps * 64;
    double4v acc = {0.0, 0.0, 0.0, 0.0};
    for (int64_t s = s0; s < s1; s++) {
        const double a = A[s * 64 + lane];
        const double b = B[s * 64 + lane];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    // D[i][j]: j = lane % 16, i = 4 v + lane / 16
    double* out = partial + (((size_t)g * ntiles + t) * 256);
#pragma unroll
    for (int v = 0; v < 4; v++) out[(4 * v + lane / 16) * 16 + (lane % 16)] = acc[v];
}

// ---- pass 3b (mode 1): split-bf16 MFMA, six products per k-step
__global__ __launch_bounds__(64) void gemm_bf16x6_kernel(const __bf16* __restrict__ xt, int64_t ksteps, int nb,
                                                         int groups, double* __restrict__ partial)
{
    const int lane = threadIdx.x, ntiles = nb * (nb + 1) / 2;
    int t, g, I, J;
    tile_group_of(blockIdx.x, ntiles, groups, &t, &g);
    tile_of(t, nb, &I, &J);
    // per k-step block: kSplit parts x 64 lanes x 8 bf16
    const size_t step_elems = kSplit * 64 * 8;
    const bf16x8* A = reinterpret_cast<const bf16x8*>(xt + (size_t)I * ksteps * step_elems);
    const bf16x8* B = reinterpret_cast<const bf16x8*>(xt + (size_t)J * ksteps * step_elems);
    // groups split the k-steps in whole folds of 16 steps (256 walkers)
    const int64_t folds = ksteps / 16;
    const int64_t per = (folds + groups - 1) / groups;
    const int64_t f0 = g * per, f1 = min(folds, f0 + per);
    double acc64[16];
#pragma unroll
    for (int v = 0; v < 16; v++) acc64[v] = 0.0;
    // The leading term x0*y0 carries nearly all of each product: its 16-walker
    // MFMA sums go to FP64 at every k-step, so FP32 rounds only one MFMA's sum
    // of it (a 256-walker FP32 accumulator left S errors of ~1e-4 absolute on
    // entries ~1e5, enough to make S + eps I indefinite at eps = 1e-4).  The
    // five smaller terms (2^-8 of it and below) share an FP32 accumulator
    // folded every 256 walkers.
    const floatx16 zero = {};
    for (int64_t f = f0; f < f1; f++) {
        floatx16 acc = zero;
        for (int64_t s = f * 16; s < f * 16 + 16; s++) {
            const bf16x8* a = A + s * (kSplit * 64) + lane;
            const bf16x8* b = B + s * (kSplit * 64) + lane;
            const bf16x8 a0 = a[0], a1 = a[64], a2 = a[128];
            const bf16x8 b0 = b[0], b1 = b[64], b2 = b[128];
            const floatx16 lead = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, zero, 0, 0, 0);
            // smallest terms first
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b2, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, b0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 16; v++) acc64[v] += (double)lead[v];
        }
#pragma unroll
        for (int v = 0; v < 16; v++) acc64[v] += (double)acc[v];
    }
    // D[i][j]: j = lane % 32, i = 8 (v / 4) + 4 (lane / 32) + (v % 4)
    double* out = partial + (((size_t)g * ntiles + t) * 1024);
#pragma unroll
    for (int v = 0; v < 16; v++) out[(8 * (v / 4) + 4 * (lane / 32) + (v % 4)) * 32 + (lane % 32)] = acc64[v];
}

// ---- pass 4: sum slices, scale 1/W, scatter to S (both triangles) and F
__global__ void reduce_tiles_kernel(const double* __restrict__ partial, int ntiles, int nb, int tile, int groups,
                                    int P, int64_t W, const int* __restrict__ outmap, int P_full,
                                    double* __restrict__ S, double* __restrict__ F)
{
    const int t = blockIdx.x;
    int I, J;
    tile_of(t, nb, &I, &J);
    const int tt = tile * tile;
    const double inv = 1.0 / (double)W;
    for (int e = threadIdx.x; e < tt; e += blockDim.x) {
        // diagonal tiles: take the upper element for both triangles
        if (I == J && e / tile > e % tile) continue;
        double s = 0.0;
        for (int g = 0; g < groups; g++) s += partial[((size_t)g * ntiles + t) * tt + e];
        s *= inv;
        const int i = I * tile + e / tile, j = J * tile + e % tile;
        // full indices of computed columns i and j (one, or two for a folded group)
        int fi[2] = {i, -1}, fj[2] = {j, -1};
        if (outmap) {
            if (i < P) fi[0] = outmap[2 * i], fi[1] = outmap[2 * i + 1];
            if (j < P) fj[0] = outmap[2 * j], fj[1] = outmap[2 * j + 1];
        }
        if (i < P && j < P) {
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    if (fi[a] >= 0 && fj[b] >= 0) {
                        S[(size_t)fi[a] * P_full + fj[b]] = s;
                        S[(size_t)fj[b] * P_full + fi[a]] = s;
                    }
        } else if (i < P && (j == P || j == P + 1)) {
            for (int a = 0; a < 2; a++)
                if (fi[a] >= 0) F[(size_t)(j - P) * P_full + fi[a]] = s;
        }
    }
}

}  // namespace

void moments_work_bytes(int64_t W, int P, int mode, size_t* xt, size_t* partial, size_t* colsum)
{
    Geometry g = geometry(W, P, mode);
    if (mode == 0)
        *xt = (size_t)g.Wp * g.Pp * sizeof(double);
    else
        *xt = (size_t)g.Wp * g.Pp * kSplit * sizeof(__bf16);
    *partial = (size_t)g.groups * g.ntiles * g.tile * g.tile * sizeof(double);
    *colsum = (size_t)g.chunks * g.Pa * sizeof(double);
}

template <typename TO>
hipError_t launch_moments(const TO* O, const double* ER, const double* EI, int64_t W, int P, int mode,
                          const MomentsWork& wk, double* mean, double* S, double* F, hipStream_t st)
{
    if (W <= 0) return hipErrorInvalidValue;
    // computed columns: all P of O, or the wk.P_cols of a fold
    const bool fold = wk.cols != nullptr;
    if (fold && (!wk.outmap || !wk.mean_cols || wk.P_cols < 1 || wk.P_cols > P)) return hipErrorInvalidValue;
    const int Pc = fold ? wk.P_cols : P;
    const ColSrc cs{wk.cols, P};
    double* mc = fold ? wk.mean_cols : mean;   // means of the computed columns
    Geometry g = geometry(W, Pc, mode);
    size_t xb, pb, cb;
    moments_work_bytes(W, Pc, mode, &xb, &pb, &cb);
    if (wk.xt_bytes < xb || wk.partial_bytes < pb || wk.colsum_bytes < cb) return hipErrorInvalidValue;

    hipLaunchKernelGGL(colsum_kernel<TO>, dim3((unsigned)ceil_div(g.Pa, kColBlock), g.chunks), dim3(kColBlock), 0, st,
                       O, ER, EI, W, Pc, cs, g.chunks, wk.colsum);
    hipLaunchKernelGGL(colmean_kernel, dim3((unsigned)g.Pa), dim3(256), 0, st, wk.colsum, g.chunks, g.Pa, W, mc,
                       wk.outmap, P, mean);
    if (mode == 0) {
        const int64_t total = g.Wp * g.Pp;
        int64_t blocks = ceil_div(total, 256);
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(pack64_kernel<TO>, dim3((unsigned)blocks), dim3(256), 0, st, O, ER, EI, W, Pc, cs, mc,
                           g.Wp, g.Pp, static_cast<double*>(wk.xt));
        hipLaunchKernelGGL(gemm64_kernel, dim3(g.ntiles * g.groups), dim3(64), 0, st, static_cast<const double*>(wk.xt),
                           g.ksteps, g.nb, g.groups, wk.partial);
    } else {
        const int64_t total = (int64_t)(g.Pp / 32) * (g.Wp / 16) * 64;
        int64_t blocks = ceil_div(total, 256);
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(pack_bf16_kernel<TO>, dim3((unsigned)blocks), dim3(256), 0, st, O, ER, EI, W, Pc, cs, mc,
                           g.Wp, g.Pp, static_cast<__bf16*>(wk.xt));
        hipLaunchKernelGGL(gemm_bf16x6_kernel, dim3(g.ntiles * g.groups), dim3(64), 0, st,
                           static_cast<const __bf16*>(wk.xt), g.ksteps, g.nb, g.groups, wk.partial);
    }
    hipLaunchKernelGGL(reduce_tiles_kernel, dim3(g.ntiles), dim3(256), 0, st, wk.partial, g.ntiles, g.nb, g.tile,
                       g.groups, Pc, W, wk.outmap, P, S, F);
    return hipGetLastError();
}

template hipError_t launch_moments<float>(const float*, const double*, const double*, int64_t, int, int,
                                          const MomentsWork&, double*, double*, double*, hipStream_t);
template hipError_t launch_moments<double>(const double*, const double*, const double*, int64_t, int, int,
                                           const MomentsWork&, double*, double*, double*, hipStream_t);

}  // namespace tdvmc
