// Small device kernels: parameter-derived tables, RK stage arithmetic on the
// packed parameter vector, fills, and the adjoint contractions.

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

// Effective spline coefficient c_s = (T^T u)_s of the boundary-condition map
// (src/SplineFactory.cpp:377-656, uniform branch, assembled as in
// src/PhysicalSystems/BosonsBulk.cpp:158-177): splines 0..2 take the origin
// rows, splines S-3..S-1 (S = P_s + 2) the cutoff rows, spline s in between
// parameter bc_n0 + s - 3.  With BosonsBulk's zero-slope pair: c_0 = c_2 = u_1,
// c_1 = u_0, c_s = u_{s-1}, c_{P_s-1} = c_{P_s+1} = u_{P_s-2}, c_{P_s} = u_{P_s-1}.
__device__ double spline_coeff(const KSys& sy, const double* u, int s)
{
    const int Ps = sy.P_s, S = Ps + 2;
    if (s < 3) {
        double c = 0.0;
        for (int i = 0; i < sy.bc_n0; i++) c += sy.bc_f0[i][s] * u[i];
        return c;
    }
    if (s < S - 3) return u[sy.bc_n0 + s - 3];
    double c = 0.0;
    for (int i = 0; i < sy.bc_n1; i++) c += sy.bc_f1[i][s - (S - 3)] * u[Ps - sy.bc_n1 + i];
    return c;
}

// Non-uniform knots: piece m of interval j is sum_p beta[j][m][p] s^p in
// s = r - g_j, so u = sum_p A_p s^p with A_p = sum_m c_{j+m} beta[j][m][p]
// (c: the extended coefficients of spline_coeff); the derivatives are taken
// in s directly (u' = A1 + 2 A2 s + 3 A3 s^2, u'' = 2 A2 + 6 A3 s).  Row P_s-1
// of spl holds the constant beyond rc, row P_s-1 of dspl u'(rc), u''(rc).
template <typename T>
__device__ void prep_nu_interval(const KSys& sy, const double* uR, const double* uI, const double* beta, T* spl,
                                 T* dspl, int j)
{
    const int Ps = sy.P_s;
    for (int part = 0; part < 2; part++) {
        const double* u = (part == 0 ? uR : uI) + sy.off_pair;
        double A[4] = {0.0, 0.0, 0.0, 0.0};
        for (int m = 0; m < 4; m++) {
            const double c = spline_coeff(sy, u, j + m);
            for (int p = 0; p < 4; p++) A[p] += c * beta[(j * 4 + m) * 4 + p];
        }
        for (int p = 0; p < 4; p++) spl[j * 8 + part * 4 + p] = (T)A[p];
        T* dr = dspl + j * kDsplStride + part * 4;
        dr[0] = (T)A[1];
        dr[1] = (T)(2.0 * A[2]);
        dr[2] = (T)(3.0 * A[3]);
        dr[3] = (T)(6.0 * A[3]);
        if (j == Ps - 2) {
            // r = rc: the last interval at its end, s = g_{P-1} - g_{P-2}; beyond
            // rc the outer-pair value
            T* de = dspl + (Ps - 1) * kDsplStride + part * 4;
            const double hl = sy.hlast;
            de[0] = (T)(A[1] + hl * (2.0 * A[2] + hl * 3.0 * A[3]));
            de[1] = (T)(2.0 * A[2] + 6.0 * A[3] * hl);
            de[2] = (T)0;
            de[3] = (T)0;
            T* so = spl + (Ps - 1) * 8 + part * 4;
            so[0] = (T)u[Ps - 1];
            so[1] = so[2] = so[3] = (T)0;
        }
    }
}

template <typename T>
__device__ void prep_tables_body(const KSys& sy, const double* __restrict__ uR, const double* __restrict__ uI,
                                 const int* __restrict__ kmap, const T* __restrict__ kv, T* __restrict__ spl,
                                 T* __restrict__ dspl, T* __restrict__ alpha, T* __restrict__ kag,
                                 const double* __restrict__ nubeta, int j0, int jstep)
{
    const int Ps = sy.P_s;
    for (int j = j0; j < Ps - 1 + sy.Kh; j += jstep) {
        if (j < Ps - 1 && sy.nu) {
            prep_nu_interval(sy, uR, uI, nubeta, spl, dspl, j);
        } else if (j < Ps - 1) {
            for (int part = 0; part < 2; part++) {
                const double* u = (part == 0 ? uR : uI) + sy.off_pair;
                const double c0 = spline_coeff(sy, u, j), c1 = spline_coeff(sy, u, j + 1);
                const double c2 = spline_coeff(sy, u, j + 2), c3 = spline_coeff(sy, u, j + 3);
                // u(t) on knot interval j for the four uniform cubic B-spline pieces
                const double A1 = (c2 - c0) / 2.0, A2 = (c0 - 2.0 * c1 + c2) / 2.0;
                const double A3 = (-c0 + 3.0 * c1 - 3.0 * c2 + c3) / 6.0;
                spl[j * 8 + part * 4 + 0] = (T)((c0 + 4.0 * c1 + c2) / 6.0);
                spl[j * 8 + part * 4 + 1] = (T)A1;
                spl[j * 8 + part * 4 + 2] = (T)A2;
                spl[j * 8 + part * 4 + 3] = (T)A3;
                // derivatives per unit length, t = (r - knot_j) / h:
                // u' = d1 + t (d2 + t d3), u'' = hinv (d2 + t e3)
                const double hi = sy.hinv;
                T* dr = dspl + j * kDsplStride + part * 4;
                dr[0] = (T)(A1 * hi);
                dr[1] = (T)(2.0 * A2 * hi);
                dr[2] = (T)(3.0 * A3 * hi);
                dr[3] = (T)(6.0 * A3 * hi);
                if (j == Ps - 2) {
                    // extra row for r = rc exactly (interval P_s-1, t = 0): the
                    // values of interval P_s-2 at t = 1
                    T* de = dspl + (Ps - 1) * kDsplStride + part * 4;
                    de[0] = (T)((A1 + 2.0 * A2 + 3.0 * A3) * hi);
                    de[1] = (T)((2.0 * A2 + 6.0 * A3) * hi);
                    de[2] = (T)0;
                    de[3] = (T)0;
                }
            }
        } else {
            // amplitude of |rho_k|^2: alpha_k, plus alpha_{-k} when -k is in the set
            const int k = j - (Ps - 1);
            const int m0 = kmap[2 * k], m1 = kmap[2 * k + 1];
            double aR = uR[sy.off_k + m0], aI = uI[sy.off_k + m0];
            if (m1 >= 0) {
                aR += uR[sy.off_k + m1];
                aI += uI[sy.off_k + m1];
            }
            alpha[2 * k] = (T)aR;
            alpha[2 * k + 1] = (T)aI;
            // amplitude times (n / L, |k|^2) for the k-space gradient and Laplacian
            for (int m = 0; m < 4; m++) {
                kag[8 * k + m] = (T)(aR * (double)kv[4 * k + m]);
                kag[8 * k + 4 + m] = (T)(aI * (double)kv[4 * k + m]);
            }
        }
    }
}

template <typename T>
__global__ void prep_tables_kernel(KSys sy, const double* __restrict__ uR, const double* __restrict__ uI,
                                   const int* __restrict__ kmap, const T* __restrict__ kv, T* __restrict__ spl,
                                   T* __restrict__ dspl, T* __restrict__ alpha, T* __restrict__ kag,
                                   const double* __restrict__ nubeta)
{
    prep_tables_body(sy, uR, uI, kmap, kv, spl, dspl, alpha, kag, nubeta, threadIdx.x + blockIdx.x * blockDim.x,
                     blockDim.x * gridDim.x);
}

// ---- batches of systems: one workgroup per system
template <typename T>
__global__ void prep_tables_batch_kernel(const BatchParams<T>* __restrict__ bp)
{
    const BatchParams<T> b = bp[blockIdx.x];
    prep_tables_body(b.sy, b.uR, b.uI, b.kmap, b.kv, b.spl, b.dspl, b.alpha, b.kag, nullptr, threadIdx.x,
                     blockDim.x);
}

// x (concatenated packed parameters) -> per-system uR, uI
template <typename T>
__global__ void unpack_batch_kernel(const BatchParams<T>* __restrict__ bp, const double* __restrict__ x)
{
    const BatchParams<T> b = bp[blockIdx.x];
    const double* xs = x + b.xoff;
    for (int i = threadIdx.x; i <= b.sy.P; i += blockDim.x) {
        // slot P: phi R, phi I (packed at 2P, 2P + 1)
        b.uR[i] = i < b.sy.P ? xs[i] : xs[2 * b.sy.P];
        b.uI[i] = i < b.sy.P ? xs[b.sy.P + i] : xs[2 * b.sy.P + 1];
    }
}

// per-system packed parameters <-> the concatenated vector
template <typename T>
__global__ void gather_batch_kernel(const BatchParams<T>* __restrict__ bp, double* __restrict__ x, int to_batch)
{
    const BatchParams<T> b = bp[blockIdx.x];
    const int n = 2 * b.sy.P + 2;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        if (to_batch)
            x[b.xoff + i] = b.xsys[i];
        else
            b.xsys[i] = x[b.xoff + i];
    }
}

// FP64 moments of a batch of systems in two launches.  Walkers are shifted
// by the values of the system's walker 0 (x = v - v_0: the sums of x and x x^T
// carry no large common offset), split into chunks of cw walkers, and every
// (chunk, system) workgroup writes its partial sums; a second launch adds
// the chunks of each system in a fixed order and centres:
// cov(a, b) = sum(x_a x_b) / W - mean(x_a) mean(x_b).
__device__ __forceinline__ double moments_col(const MomentsItem& it, int64_t w, int c)
{
    const int P = it.P;
    return c < P ? it.O[w * P + c] : (c == P ? it.ER[w] : it.EI[w]);
}

// entry e of the product list: S upper triangle (i <= j < P), then F (O_i
// with E_R, O_i with E_I) as columns P and P + 1
__device__ __forceinline__ void moments_entry(int P, int e, int* i, int* j)
{
    const int nS = P * (P + 1) / 2;
    if (e < nS) {
        int a = 0, r = e;
        while (r >= P - a) {
            r -= P - a;
            a++;
        }
        *i = a;
        *j = a + r;
    } else {
        *i = (e - nS) % P;
        *j = P + (e - nS) / P;
    }
}

__global__ __launch_bounds__(256) void moments_partial_kernel(const MomentsItem* __restrict__ items)
{
    const MomentsItem it = items[blockIdx.y];
    const int chunk = blockIdx.x;
    if ((int64_t)chunk * it.cw >= it.W) return;
    const int P = it.P, C = P + 2, tid = threadIdx.x;
    const int64_t w0 = (int64_t)chunk * it.cw;
    const int nw = (int)((it.W - w0) < it.cw ? (it.W - w0) : it.cw);
    extern __shared__ __align__(16) unsigned char smem[];
    double* sh = reinterpret_cast<double*>(smem);   // [C] shift
    double* X = sh + C;                              // [cw][C] shifted values
    for (int c = tid; c < C; c += blockDim.x) sh[c] = moments_col(it, 0, c);
    __syncthreads();
    for (int e = tid; e < nw * C; e += blockDim.x) {
        const int ww = e / C, c = e - ww * C;
        X[e] = moments_col(it, w0 + ww, c) - sh[c];
    }
    __syncthreads();
    const int64_t len = moments_part_len(P);
    double* out = it.part + (int64_t)chunk * len;
    for (int64_t e = tid; e < len; e += blockDim.x) {
        double acc = 0.0;
        if (e < C) {
            for (int ww = 0; ww < nw; ww++) acc += X[ww * C + e];
        } else {
            int i, j;
            moments_entry(P, (int)(e - C), &i, &j);
            for (int ww = 0; ww < nw; ww++) acc += X[ww * C + i] * X[ww * C + j];
        }
        out[e] = acc;
    }
}

__global__ __launch_bounds__(256) void moments_reduce_kernel(const MomentsItem* __restrict__ items)
{
    const MomentsItem it = items[blockIdx.x];
    const int P = it.P, C = P + 2, tid = threadIdx.x;
    const int64_t len = moments_part_len(P);
    const double invW = 1.0 / (double)it.W;
    extern __shared__ __align__(16) unsigned char smem[];
    double* m = reinterpret_cast<double*>(smem);   // [C] means of the shifted values
    for (int c = tid; c < C; c += blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < it.nchunk; k++) s += it.part[(int64_t)k * len + c];
        m[c] = s * invW;
        it.mean[c] = moments_col(it, 0, c) + m[c];
    }
    __syncthreads();
    for (int64_t e = C + tid; e < len; e += blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < it.nchunk; k++) s += it.part[(int64_t)k * len + e];
        int i, j;
        moments_entry(P, (int)(e - C), &i, &j);
        const double v = s * invW - m[i] * m[j];
        if (j < P) {
            it.S[(size_t)i * P + j] = v;
            it.S[(size_t)j * P + i] = v;
        } else {
            it.F[(size_t)(j - P) * P + i] = v;
        }
    }
}

__global__ void axpy_params_kernel(int n, const double* x0, const double* xdot, double c, double* y)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < n; i += blockDim.x * gridDim.x)
        y[i] = x0[i] + c * xdot[i];
}

__global__ void rk4_combine_kernel(int n, double* x, const double* k1, const double* k2, const double* k3,
                                   const double* k4, double dt)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < n; i += blockDim.x * gridDim.x)
        x[i] = x[i] + ((k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0) * dt;
}

// predictor-corrector combination of src/TDVMC.cpp:1899-1902: y = x + (k1 + k2) dt/2
__global__ void pc_combine_kernel(int n, const double* x, const double* k1, const double* k2, double dt_2, double* y)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < n; i += blockDim.x * gridDim.x)
        y[i] = x[i] + (k1[i] + k2[i]) * dt_2;
}

__global__ void unpack_params_kernel(int P, const double* packed, double* uR, double* uI)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i <= P; i += blockDim.x * gridDim.x) {
        // slot P: phi R, phi I (packed at 2P, 2P + 1)
        uR[i] = i < P ? packed[i] : packed[2 * P];
        uI[i] = i < P ? packed[P + i] : packed[2 * P + 1];
    }
}

// After a step (src/TDVMC.cpp:3757-3775): phi R := -exponent (NormalizeWavefunction,
// :2525-2529) and the gauge shift of AdjustParameters (:2531-2567), which
// subtracts the last pair-spline parameter from the pair-spline block
// [off, off + n) only.  With the partition of unity of the B-spline basis (and
// the outer-pair term of that same last parameter) this adds a constant to
// ln psi; the one-body and k-space parameters are left as they are.
__global__ void post_step_kernel(int P, double* x, const double* expo, int normalize, int adjust, int off, int n)
{
    const double lastR = x[off + n - 1], lastI = x[P + off + n - 1];
    const double e = normalize ? expo[0] : 0.0;
    __syncthreads();
    if (adjust) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            x[off + i] -= lastR;
            x[P + off + i] -= lastI;
        }
    }
    if (normalize && threadIdx.x == 0) x[2 * P] = -e;
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n)
{
    for (int64_t i = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; i < n; i += (int64_t)blockDim.x * gridDim.x)
        p[i] = v;
}

__global__ void add_u32_kernel(uint32_t* p, uint32_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] += v;
}

// dalpha[k] = sum_w gbar[w] O[w][k]; one block per 64 parameters, FP64
__global__ void adjoint_reduce_kernel(const double* __restrict__ O, const double* __restrict__ gbar, int64_t W,
                                      int P, double* __restrict__ dalpha)
{
    __shared__ double part[4][64];
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int row = threadIdx.x >> 6;
    double s = 0.0;
    if (k < P)
        for (int64_t w = row; w < W; w += 4) s += gbar[w] * O[w * P + k];
    part[row][threadIdx.x & 63] = s;
    __syncthreads();
    if (row == 0 && k < P) dalpha[k] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

// dR[w][i] = gbar[w] * Re grad[w][i]  (grad interleaves Re, Im)
__global__ void adjoint_scale_kernel(const double* __restrict__ grad, const double* __restrict__ gbar, int64_t W,
                                     int ND, double* __restrict__ dR)
{
    const int64_t n = W * (int64_t)ND;
    for (int64_t i = threadIdx.x + (int64_t)blockIdx.x * blockDim.x; i < n; i += (int64_t)blockDim.x * gridDim.x)
        dR[i] = gbar[i / ND] * grad[2 * i];
}

__global__ void adjoint_outer_kernel(const double* __restrict__ outer, const double* __restrict__ gbar, int64_t W,
                                     int idx, double* __restrict__ dalpha)
{
    __shared__ double part[256];
    double s = 0.0;
    for (int64_t w = threadIdx.x; w < W; w += 256) s += gbar[w] * outer[w];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) dalpha[idx] += part[0];
}

}  // namespace

hipError_t launch_adjoint_outer(const double* outer, const double* gbar, int64_t W, int idx, double* dalpha,
                                hipStream_t st)
{
    hipLaunchKernelGGL(adjoint_outer_kernel, dim3(1), dim3(256), 0, st, outer, gbar, W, idx, dalpha);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_prep_tables(const KSys& sy, const double* uR, const double* uI, const int* kmap, const T* kv, T* spl,
                              T* dspl, T* alpha, T* kag, const double* nubeta, hipStream_t st)
{
    if (sy.nu && !nubeta) return hipErrorInvalidValue;
    const int n = sy.P_s - 1 + sy.Kh;
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(prep_tables_kernel<T>, dim3(blocks), dim3(256), 0, st, sy, uR, uI, kmap, kv, spl, dspl, alpha,
                       kag, nubeta);
    return hipGetLastError();
}
template hipError_t launch_prep_tables<float>(const KSys&, const double*, const double*, const int*, const float*,
                                              float*, float*, float*, float*, const double*, hipStream_t);
template hipError_t launch_prep_tables<double>(const KSys&, const double*, const double*, const int*, const double*,
                                               double*, double*, double*, double*, const double*, hipStream_t);

hipError_t launch_axpy_params(int P, const double* x0, const double* xdot, double c, double* y, hipStream_t st)
{
    const int n = 2 * P + 2;
    hipLaunchKernelGGL(axpy_params_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, x0, xdot, c, y);
    return hipGetLastError();
}

hipError_t launch_rk4_combine(int P, double* x, const double* k1, const double* k2, const double* k3,
                              const double* k4, double dt, hipStream_t st)
{
    const int n = 2 * P + 2;
    hipLaunchKernelGGL(rk4_combine_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, x, k1, k2, k3, k4, dt);
    return hipGetLastError();
}

hipError_t launch_pc_combine(int P, const double* x, const double* k1, const double* k2, double dt, double* y,
                             hipStream_t st)
{
    const int n = 2 * P + 2;
    hipLaunchKernelGGL(pc_combine_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, x, k1, k2, dt / 2.0, y);
    return hipGetLastError();
}

hipError_t launch_unpack_params(int P, const double* packed, double* uR, double* uI, hipStream_t st)
{
    hipLaunchKernelGGL(unpack_params_kernel, dim3((P + 256) / 256), dim3(256), 0, st, P, packed, uR, uI);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_prep_tables_batch(const BatchParams<T>* bp_dev, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(prep_tables_batch_kernel<T>, dim3(n), dim3(256), 0, st, bp_dev);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_unpack_batch(const BatchParams<T>* bp_dev, int n, const double* x, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(unpack_batch_kernel<T>, dim3(n), dim3(256), 0, st, bp_dev, x);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_gather_batch(const BatchParams<T>* bp_dev, int n, double* x, int to_batch, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_batch_kernel<T>, dim3(n), dim3(256), 0, st, bp_dev, x, to_batch);
    return hipGetLastError();
}
template hipError_t launch_prep_tables_batch<float>(const BatchParams<float>*, int, hipStream_t);
template hipError_t launch_prep_tables_batch<double>(const BatchParams<double>*, int, hipStream_t);
template hipError_t launch_unpack_batch<double>(const BatchParams<double>*, int, const double*, hipStream_t);
template hipError_t launch_gather_batch<double>(const BatchParams<double>*, int, double*, int, hipStream_t);

int moments_chunk_walkers(int P)
{
    int cw = 49152 / (8 * (P + 2)) - 1;
    cw = cw > 256 ? 256 : cw;
    return cw < 1 ? 1 : cw;
}

hipError_t launch_moments_batch(const MomentsItem* items_dev, int n, int nchunk_max, size_t lds, int P_max,
                                hipStream_t st)
{
    if (n <= 0 || nchunk_max <= 0) return hipSuccess;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)moments_partial_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(moments_partial_kernel, dim3((unsigned)nchunk_max, (unsigned)n), dim3(256), lds, st, items_dev);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(moments_reduce_kernel, dim3((unsigned)n), dim3(256), (size_t)(P_max + 2) * sizeof(double), st,
                       items_dev);
    return hipGetLastError();
}

hipError_t launch_post_step(int P, double* x, const double* expo, int normalize, int adjust, int off, int n,
                            hipStream_t st)
{
    hipLaunchKernelGGL(post_step_kernel, dim3(1), dim3(256), 0, st, P, x, expo, normalize, adjust, off, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_add_u32(uint32_t* p, uint32_t v, hipStream_t st)
{
    hipLaunchKernelGGL(add_u32_kernel, dim3(1), dim3(64), 0, st, p, v);
    return hipGetLastError();
}

hipError_t launch_adjoint_reduce(const double* O, const double* gbar, int64_t W, int P, double* dalpha, hipStream_t st)
{
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((P + 63) / 64), dim3(256), 0, st, O, gbar, W, P, dalpha);
    return hipGetLastError();
}

hipError_t launch_adjoint_scale(const double* grad, const double* gbar, int64_t W, int ND, double* dR, hipStream_t st)
{
    const int64_t n = W * (int64_t)ND;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(adjoint_scale_kernel, dim3((unsigned)blocks), dim3(256), 0, st, grad, gbar, W, ND, dR);
    return hipGetLastError();
}

}  // namespace tdvmc
