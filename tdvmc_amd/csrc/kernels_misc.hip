// Small device kernels: parameter-derived tables, RK stage arithmetic on the
// packed parameter vector, fills, and the adjoint contractions.

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

// Effective spline coefficient c_s = (T^T u)_s of the zero-slope boundary
// maps (src/SplineFactory.cpp:423-456 and :563-596, uniform branch, as
// assembled in src/PhysicalSystems/BosonsBulk.cpp:158-177).
__device__ double spline_coeff(const double* u, int Ps, int s)
{
    if (s == 0 || s == 2) return u[1];
    if (s == 1) return u[0];
    if (s <= Ps - 2) return u[s - 1];
    if (s == Ps - 1 || s == Ps + 1) return u[Ps - 2];
    return u[Ps - 1];   // s == Ps
}

template <typename T>
__global__ void prep_tables_kernel(KSys sy, const double* __restrict__ uR, const double* __restrict__ uI,
                                   const int* __restrict__ kmap, T* __restrict__ spl, T* __restrict__ dspl,
                                   T* __restrict__ alpha)
{
    const int Ps = sy.P_s;
    for (int j = threadIdx.x + blockIdx.x * blockDim.x; j < Ps - 1 + sy.Kh; j += blockDim.x * gridDim.x) {
        if (j < Ps - 1) {
            for (int part = 0; part < 2; part++) {
                const double* u = (part == 0 ? uR : uI) + sy.off_pair;
                const double c0 = spline_coeff(u, Ps, j), c1 = spline_coeff(u, Ps, j + 1);
                const double c2 = spline_coeff(u, Ps, j + 2), c3 = spline_coeff(u, Ps, j + 3);
                // u(t) on knot interval j for the four uniform cubic B-spline pieces
                const double A1 = (c2 - c0) / 2.0, A2 = (c0 - 2.0 * c1 + c2) / 2.0;
                const double A3 = (-c0 + 3.0 * c1 - 3.0 * c2 + c3) / 6.0;
                spl[j * 8 + part * 4 + 0] = (T)((c0 + 4.0 * c1 + c2) / 6.0);
                spl[j * 8 + part * 4 + 1] = (T)A1;
                spl[j * 8 + part * 4 + 2] = (T)A2;
                spl[j * 8 + part * 4 + 3] = (T)A3;
                // derivatives per unit length, t = (r - knot_j) / h:
                // u' = d1 + t (d2 + t d3), u'' = hinv d2 + t e3
                const double hi = sy.hinv;
                dspl[j * 8 + part * 4 + 0] = (T)(A1 * hi);
                dspl[j * 8 + part * 4 + 1] = (T)(2.0 * A2 * hi);
                dspl[j * 8 + part * 4 + 2] = (T)(3.0 * A3 * hi);
                dspl[j * 8 + part * 4 + 3] = (T)(6.0 * A3 * hi * hi);
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

__global__ void unpack_params_kernel(int P, const double* packed, double* uR, double* uI)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < P; i += blockDim.x * gridDim.x) {
        uR[i] = packed[i];
        uI[i] = packed[P + i];
    }
}

// AdjustParameters (src/TDVMC.cpp:2531-2567: u_i -= u_{P-1}, both parts) and
// NormalizeWavefunction (:2525-2529 with the exponent of :3763: phi R = -exponent)
__global__ void post_step_kernel(int P, double* x, const double* expo, int normalize, int adjust)
{
    const double lastR = x[P - 1], lastI = x[2 * P - 1];
    const double e = normalize ? expo[0] : 0.0;
    __syncthreads();
    if (adjust) {
        for (int i = threadIdx.x; i < P; i += blockDim.x) {
            x[i] -= lastR;
            x[P + i] -= lastI;
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
hipError_t launch_prep_tables(const KSys& sy, const double* uR, const double* uI, const int* kmap, T* spl, T* dspl,
                              T* alpha, hipStream_t st)
{
    const int n = sy.P_s - 1 + sy.Kh;
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(prep_tables_kernel<T>, dim3(blocks), dim3(256), 0, st, sy, uR, uI, kmap, spl, dspl, alpha);
    return hipGetLastError();
}
template hipError_t launch_prep_tables<float>(const KSys&, const double*, const double*, const int*, float*, float*,
                                              float*, hipStream_t);
template hipError_t launch_prep_tables<double>(const KSys&, const double*, const double*, const int*, double*,
                                               double*, double*, hipStream_t);

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

hipError_t launch_unpack_params(int P, const double* packed, double* uR, double* uI, hipStream_t st)
{
    hipLaunchKernelGGL(unpack_params_kernel, dim3((P + 255) / 256), dim3(256), 0, st, P, packed, uR, uI);
    return hipGetLastError();
}

hipError_t launch_post_step(int P, double* x, const double* expo, int normalize, int adjust, hipStream_t st)
{
    hipLaunchKernelGGL(post_step_kernel, dim3(1), dim3(256), 0, st, P, x, expo, normalize, adjust);
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
