// FP64 TDVP solve in one workgroup: S alpha-dot = b for the real and the
// imaginary right-hand side, phi-dot.
//
// Restates src/TDVMC.cpp:1506-1537 (right-hand sides from the moments, with
// the real/imaginary-time conventions), :1684-1701 (optional diagonal
// scaling), :1703-1711 (S += eps I), :1560-1592 (lower Cholesky; a
// non-positive pivot is counted and leaves the diagonal entry at its
// unreduced value), :1594-1622 (forward / back substitution), :1658-1682
// (phi-dot, computed on the scaled solution before the scaling is undone, as
// in :1749-1758).
//
// Blocked right-looking Cholesky: panels of NB columns are staged in LDS
// (transposed, so consecutive threads read consecutive rows), factored there
// with two barriers per column, written back, and the trailing lower
// triangle is updated by all 1024 threads (a rank-NB update with the panel
// read from LDS).  The triangular solves are blocked the same way: one
// wavefront solves each NB x NB diagonal block with lane exchanges, the rest
// of the right-hand side is updated by the whole workgroup.  The factor lives
// column-major in the work buffer (L2-resident for P <= 600).

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

template <int NB>
__global__ __launch_bounds__(kThreads) void solve_kernel(SolveArgs a)
{
    const int P = a.P, tid = threadIdx.x;
    double* Lm = a.work;                 // column-major P x P: Lm[j * P + i] = L[i][j]
    double* bR = a.work + (size_t)P * P;
    double* bI = bR + P;
    double* sc = bI + P;
    double* dorig = sc + P;              // diagonal after scaling and the eps shift
    extern __shared__ __align__(16) unsigned char smem[];
    double* pan = reinterpret_cast<double*>(smem);   // [NB][rows] panel, transposed
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
        if (i == j) {
            v += a.eps;
            dorig[i] = v;
        }
        Lm[e] = v;
    }
    for (int i = tid; i < P; i += kThreads) {
        if (a.precondition) {
            bR[i] /= sc[i];
            bI[i] /= sc[i];
        }
    }
    __syncthreads();

    // ---------------------------------------------------------------- Cholesky
    for (int p0 = 0; p0 < P; p0 += NB) {
        const int nb = min(NB, P - p0);
        const int rows = P - p0;
        for (int e = tid; e < nb * rows; e += kThreads) {
            const int c = e / rows, r = e % rows;
            pan[c * rows + r] = Lm[(size_t)(p0 + c) * P + p0 + r];
        }
        __syncthreads();
        for (int c = 0; c < nb; c++) {
            double* col = pan + c * rows;
            if (tid == 0) {
                const double d = col[c];
                if (d > 0.0) {
                    piv = sqrt(d);
                } else {
                    piv = dorig[p0 + c];
                    nbad++;
                }
            }
            __syncthreads();
            const double pj = piv, ipj = 1.0 / pj;
            // update the panel's later columns with the (still unscaled) column c
            const int ncols = nb - c - 1;
            if (ncols > 0) {
                const int span = rows - c - 1;     // rows below the pivot
                for (int e = tid; e < ncols * span; e += kThreads) {
                    const int cc = c + 1 + e / span;
                    const int r = c + 1 + e % span;
                    if (r >= cc) pan[cc * rows + r] -= (col[r] * ipj) * (col[cc] * ipj);
                }
            }
            __syncthreads();
            for (int r = c + 1 + tid; r < rows; r += kThreads) col[r] *= ipj;
            if (tid == 0) col[c] = pj;
            // (the next column's pivot thread reads only its own diagonal entry)
        }
        __syncthreads();
        for (int e = tid; e < nb * rows; e += kThreads) {
            const int c = e / rows, r = e % rows;
            if (r >= c) Lm[(size_t)(p0 + c) * P + p0 + r] = pan[c * rows + r];
        }
        // trailing update of the lower triangle: A[i][k] -= sum_c L[i][c] L[k][c]
        const int T = P - p0 - nb;
        if (T > 0) {
            for (int e = tid; e < T * T; e += kThreads) {
                const int kk = e / T, ii = e % T;
                if (ii < kk) continue;
                const int rk = nb + kk, ri = nb + ii;    // row offsets inside the panel
                double s = 0.0;
#pragma unroll 8
                for (int c = 0; c < nb; c++) s += pan[c * rows + ri] * pan[c * rows + rk];
                Lm[(size_t)(p0 + nb + kk) * P + p0 + nb + ii] -= s;
            }
        }
        __syncthreads();
    }

    // ---------------------------------------------------------------- L t = b
    const int lane = tid & 63, wv = tid >> 6;
    for (int p0 = 0; p0 < P; p0 += NB) {
        const int nb = min(NB, P - p0);
        if (wv == 0) {
            // diagonal block by one wavefront: lane l owns row p0 + l
            double tR = lane < nb ? bR[p0 + lane] : 0.0, tI = lane < nb ? bI[p0 + lane] : 0.0;
            for (int j = 0; j < nb; j++) {
                const double ljj = Lm[(size_t)(p0 + j) * P + p0 + j];
                const double xR = __shfl(tR, j, 64) / ljj, xI = __shfl(tI, j, 64) / ljj;
                if (lane == j) {
                    tR = xR;
                    tI = xI;
                } else if (lane > j && lane < nb) {
                    const double l = Lm[(size_t)(p0 + j) * P + p0 + lane];
                    tR -= l * xR;
                    tI -= l * xI;
                }
            }
            if (lane < nb) {
                bR[p0 + lane] = tR;
                bI[p0 + lane] = tI;
            }
        }
        __syncthreads();
        for (int i = p0 + nb + tid; i < P; i += kThreads) {
            double sR = 0.0, sI = 0.0;
            for (int j = 0; j < nb; j++) {
                const double l = Lm[(size_t)(p0 + j) * P + i];
                sR += l * bR[p0 + j];
                sI += l * bI[p0 + j];
            }
            bR[i] -= sR;
            bI[i] -= sI;
        }
        __syncthreads();
    }

    // ---------------------------------------------------------------- L^T x = t
    const int npan = (P + NB - 1) / NB;
    for (int pb = npan - 1; pb >= 0; pb--) {
        const int p0 = pb * NB, nb = min(NB, P - p0), pe = p0 + nb;
        // t_i -= sum_{j >= pe} L[j][i] x_j for the panel's rows (one wave per row group)
        for (int i = p0 + wv; i < pe; i += kThreads / 64) {
            double sR = 0.0, sI = 0.0;
            for (int j = pe + lane; j < P; j += 64) {
                const double l = Lm[(size_t)i * P + j];
                sR += l * bR[j];
                sI += l * bI[j];
            }
            for (int o = 32; o > 0; o >>= 1) {
                sR += __shfl_xor(sR, o, 64);
                sI += __shfl_xor(sI, o, 64);
            }
            if (lane == 0) {
                bR[i] -= sR;
                bI[i] -= sI;
            }
        }
        __syncthreads();
        if (wv == 0) {
            double tR = lane < nb ? bR[p0 + lane] : 0.0, tI = lane < nb ? bI[p0 + lane] : 0.0;
            for (int j = nb - 1; j >= 0; j--) {
                const double ljj = Lm[(size_t)(p0 + j) * P + p0 + j];
                const double xR = __shfl(tR, j, 64) / ljj, xI = __shfl(tI, j, 64) / ljj;
                if (lane == j) {
                    tR = xR;
                    tI = xI;
                } else if (lane < j) {
                    const double l = Lm[(size_t)(p0 + lane) * P + p0 + j];   // L[j][lane]
                    tR -= l * xR;
                    tI -= l * xI;
                }
            }
            if (lane < nb) {
                bR[p0 + lane] = tR;
                bI[p0 + lane] = tI;
            }
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

template <int NB>
hipError_t launch_nb(const SolveArgs& a, hipStream_t st)
{
    const size_t lds = (size_t)NB * a.P * sizeof(double);
    auto kern = solve_kernel<NB>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(1), dim3(kThreads), lds, st, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_solve(const SolveArgs& a, hipStream_t st)
{
    if (a.P <= 0) return hipErrorInvalidValue;
    if ((size_t)32 * a.P * sizeof(double) <= 150 * 1024) return launch_nb<32>(a, st);
    if ((size_t)16 * a.P * sizeof(double) <= 150 * 1024) return launch_nb<16>(a, st);
    return hipErrorInvalidValue;
}

}  // namespace tdvmc
