// FP64 TDVP solve in one workgroup: S alpha-dot = b for the real and the
// imaginary right-hand side, phi-dot.
//
// Restates src/TDVMC.cpp:1506-1537 (right-hand sides from the moments, with
// the real/imaginary-time conventions), :1684-1701 (optional diagonal
// scaling), :1703-1711 (S += eps I), :1560-1592 (lower Cholesky; a
// non-positive pivot is counted and leaves the diagonal entry unreduced),
// :1594-1622 (forward / back substitution), :1658-1682 (phi-dot, computed on
// the scaled solution before the scaling is undone, as in :1749-1758).
//
// The factor is stored column-major in the work buffer so that the
// column-oriented loops read contiguous memory across threads.

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

constexpr int kThreads = 1024;

__device__ double block_reduce(double v, double* red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; w++) s += red[w];
    return s;
}

__global__ __launch_bounds__(kThreads) void solve_kernel(SolveArgs a)
{
    const int P = a.P, tid = threadIdx.x;
    double* Lm = a.work;                 // column-major P x P: Lm[j * P + i] = L[i][j]
    double* bR = a.work + (size_t)P * P;
    double* bI = bR + P;
    double* sc = bI + P;
    double* diag = sc + P;
    __shared__ double red[kThreads / 64];
    __shared__ double piv;
    __shared__ int nbad;
    const double* mO = a.mean;
    const double eR = a.mean[P], eI = a.mean[P + 1];
    const double* FR = a.F;
    const double* FI = a.F + P;

    if (tid == 0) nbad = 0;
    for (int i = tid; i < P; i += kThreads) {
        if (a.imag_time == 0) {
            bR[i] = FI[i];
            bI[i] = -FR[i];
        } else {
            bR[i] = -FR[i];
            bI[i] = -(FI[i] + mO[i] * eI);
        }
        sc[i] = a.precondition ? sqrt(a.S[(size_t)i * P + i]) : 1.0;
    }
    __syncthreads();
    for (size_t e = tid; e < (size_t)P * P; e += kThreads) {
        const int i = (int)(e % P), j = (int)(e / P);
        double v = a.S[(size_t)i * P + j];
        if (a.precondition) v /= (sc[i] * sc[j]);
        if (i == j) v += a.eps;
        Lm[e] = v;
    }
    for (int i = tid; i < P; i += kThreads) {
        if (a.precondition) {
            bR[i] /= sc[i];
            bI[i] /= sc[i];
        }
        diag[i] = 0.0;
    }
    __syncthreads();

    // Cholesky, column by column: v_i = A_ij - sum_{k<j} L_ik L_jk for i >= j
    for (int j = 0; j < P; j++) {
        for (int i = j + tid; i < P; i += kThreads) {
            double s = Lm[(size_t)j * P + i];
            for (int k = 0; k < j; k++) s -= Lm[(size_t)k * P + i] * Lm[(size_t)k * P + j];
            if (i == j) {
                if (s > 0.0) {
                    piv = sqrt(s);
                } else {
                    piv = Lm[(size_t)j * P + j];   // unreduced diagonal is kept
                    nbad++;
                }
            } else {
                diag[i] = s;                       // staged; divided once the pivot is known
            }
        }
        __syncthreads();
        const double pj = piv;
        for (int i = j + tid; i < P; i += kThreads) {
            if (i == j)
                Lm[(size_t)j * P + j] = pj;
            else
                Lm[(size_t)j * P + i] = diag[i] / pj;
        }
        __syncthreads();
    }

    // forward substitution L t = b (both right-hand sides, in place)
    for (int j = 0; j < P; j++) {
        const double ljj = Lm[(size_t)j * P + j];
        const double tR = bR[j] / ljj, tI = bI[j] / ljj;
        __syncthreads();
        for (int i = j + 1 + tid; i < P; i += kThreads) {
            const double l = Lm[(size_t)j * P + i];
            bR[i] -= l * tR;
            bI[i] -= l * tI;
        }
        if (tid == 0) {
            bR[j] = tR;
            bI[j] = tI;
        }
        __syncthreads();
    }
    // back substitution L^T x = t
    for (int i = P - 1; i >= 0; i--) {
        const double lii = Lm[(size_t)i * P + i];
        const double xR = bR[i] / lii, xI = bI[i] / lii;
        __syncthreads();
        for (int k = tid; k < i; k += kThreads) {
            const double l = Lm[(size_t)k * P + i];   // L[i][k]
            bR[k] -= l * xR;
            bI[k] -= l * xI;
        }
        if (tid == 0) {
            bR[i] = xR;
            bI[i] = xI;
        }
        __syncthreads();
    }

    // phi-dot on the (scaled) solution, then undo the scaling
    double pr = 0.0, pi = 0.0;
    for (int i = tid; i < P; i += kThreads) {
        pr -= mO[i] * bR[i];
        pi -= mO[i] * bI[i];
    }
    pr = block_reduce(pr, red);
    pi = block_reduce(pi, red);
    for (int i = tid; i < P; i += kThreads) {
        a.xdot[i] = bR[i] / sc[i];
        a.xdot[P + i] = bI[i] / sc[i];
    }
    if (tid == 0) {
        if (a.imag_time == 0)
            pi -= eR;
        else
            pr -= eR;
        a.xdot[2 * P] = pr;
        a.xdot[2 * P + 1] = pi;
        if (a.status) a.status[0] = nbad;
    }
}

}  // namespace

hipError_t launch_solve(const SolveArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(solve_kernel, dim3(1), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

}  // namespace tdvmc
