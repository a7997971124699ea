// FP64 TDVP solve in one workgroup: S alpha-dot = b for the real and the
// imaginary right-hand side, phi-dot.
//
// Restates src/TDVMC.cpp:1506-1537 (right-hand sides from the moments, with
// the real/imaginary-time conventions), :1684-1701 (optional diagonal
// scaling), :1703-1711 (S += eps I), :1560-1592 (lower Cholesky; a
// non-positive pivot is counted and leaves the diagonal entry at its
// unreduced value), :1594-1622 (forward / back substitution), :1658-1682
// (phi-dot, computed on the scaled solution before the scaling is undone, as
// in :1749-1758).  Solver type 1 is the full-pivot QR kernel further down.
//
// Blocked right-looking Cholesky: the rows below each NB-column diagonal
// block are staged in LDS (row-major at a padded stride).  One wavefront
// factors the NB x NB diagonal block (wave-level ordering, no workgroup
// barrier), one thread per row solves the rows below it against that block,
// the panel is written back, and the trailing lower triangle is updated by
// all 1024 threads in 16 x 16 tiles on FP64 MFMA (a rank-NB update with the
// panel read from LDS; every tile's loads are issued before its stores).  The
// triangular solves are blocked the same way: each NB x NB diagonal block is
// staged in LDS and solved by one wavefront with v_readlane exchanges, the
// rest of the right-hand side is updated by the whole workgroup.  The factor
// lives column-major in the work buffer (L2-resident for the sizes the LDS
// panel allows), with 1 / L[c][c] on its diagonal so that the solves only
// multiply.  NB is 16 for P <= 340 (shorter serial chains in the diagonal
// blocks and row solves) and 32 above (fewer panels and barriers).

#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

constexpr int kThreads = 1024;

typedef double double4v __attribute__((ext_vector_type(4)));

// (r, c) of the elements e = e0, e0 + kThreads, e0 + 2 kThreads, ... of a
// column-major range with `rows` rows, stepped without a division per element
struct Walk2 {
    int r, c, dr, dc, rows;
    __device__ Walk2(int e0, int rows_) : rows(rows_)
    {
        c = e0 / rows;
        r = e0 - c * rows;
        dc = kThreads / rows;
        dr = kThreads - dc * rows;
    }
    __device__ void next()
    {
        r += dr;
        c += dc;
        if (r >= rows) {
            r -= rows;
            c++;
        }
    }
};

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
__device__ __forceinline__ void solve_body(const SolveArgs& a)
{
    const int P = a.P, tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    double* Lm = a.work;                 // column-major P x P: Lm[j * P + i] = L[i][j]
    double* bR = a.work + (size_t)P * P;
    double* bI = bR + P;
    double* sc = bI + P;
    double* dorig = sc + P;              // diagonal after scaling and the eps shift
    extern __shared__ __align__(16) unsigned char smem[];
    double* pan = reinterpret_cast<double*>(smem);   // [rows below][NB + 1] panel; [NB][NB] blocks in the solves
    __shared__ double red[kThreads / 64];
    __shared__ double ipiv[NB];          // 1 / L[c][c]
    __shared__ double ldc[NB * NB];      // diagonal block, column-major: ldc[c * NB + k] = L[k][c]
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
    // S is symmetric: read it column by column (coalesced), four entries in flight per thread
    const size_t PP = (size_t)P * P;
    Walk2 wk(tid, P);   // (row, column) of e0, stepped along with it
    for (size_t e0 = tid; e0 < PP; e0 += 4 * kThreads) {
        double v[4];
        int ri[4], ci[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const size_t e = e0 + (size_t)q * kThreads;
            v[q] = e < PP ? a.S[e] : 0.0;
            ri[q] = wk.r;
            ci[q] = wk.c;
            wk.next();
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const size_t e = e0 + (size_t)q * kThreads;
            if (e >= PP) continue;
            const int i = ri[q], j = ci[q];
            double x = v[q];
            if (a.precondition) x /= (sc[i] * sc[j]);
            if (i == j) {
                x += a.eps;
                dorig[i] = x;
            }
            Lm[e] = x;
        }
    }
    for (int i = tid; i < P; i += kThreads) {
        if (a.precondition) {
            bR[i] /= sc[i];
            bI[i] /= sc[i];
        }
    }
    __syncthreads();

    // ---------------------------------------------------------------- Cholesky
    // The rows below the diagonal block sit in LDS row-major with a padded
    // stride (pan[rb * LDP + c], LDP = NB + 1): a thread's row is read with
    // immediate offsets, and only threads rb and rb + 32 share a bank.  The
    // diagonal block is staged on its own, column-major and padded with the
    // identity to NB x NB (ldc[c * NB + l] = A[l][c]): padded columns have
    // pivot 1 and zero off-diagonal entries and leave the real rows unchanged,
    // so no step below has a run-time bound on the block.  Only the last
    // block can be partial, and it has no rows below it.
    constexpr int LDP = NB + 1;
    for (int p0 = 0; p0 < P; p0 += NB) {
        const int nb = min(NB, P - p0);
        const int below = P - p0 - nb;
        if (below > 0) {
            Walk2 wk2(tid, below);
            for (int e = tid; e < nb * below; e += kThreads) {
                pan[wk2.r * LDP + wk2.c] = Lm[(size_t)(p0 + wk2.c) * P + p0 + nb + wk2.r];
                wk2.next();
            }
        }
        for (int e = tid; e < NB * NB; e += kThreads) {
            const int c = e / NB, l = e % NB;
            ldc[e] = (c < nb && l < nb) ? Lm[(size_t)(p0 + c) * P + p0 + l] : (c == l ? 1.0 : 0.0);
        }
        __syncthreads();
        // diagonal block by one wavefront in registers: lane l holds row
        // l mod NB, column entries reach the other lanes by v_readlane.  Every
        // lane updates its whole row (the entries above the diagonal are
        // never read), and the diagonal gets 1 / L[c][c] from ipiv at the end.
        if (wv == 0) {
            const int l = lane & (NB - 1);
            double av[NB];
#pragma unroll
            for (int k = 0; k < NB; k++) av[k] = ldc[k * NB + l];
#pragma unroll
            for (int c = 0; c < NB; c++) {
                const double d = readlane_f(av[c], c);
                double ipj;
                if (d > 0.0) {
                    // 1 / sqrt(d) from the hardware estimate and two Newton
                    // steps (each squares the relative error): a shorter
                    // dependent chain than a correctly rounded sqrt and a
                    // divide, within an ulp or two of their quotient
                    double y = __builtin_amdgcn_rsq(d);
                    const double hd = 0.5 * d;
                    y = y * (1.5 - hd * y * y);
                    y = y * (1.5 - hd * y * y);
                    ipj = y;
                } else {
                    // the unreduced diagonal entry stands in as the pivot
                    ipj = 1.0 / dorig[p0 + c];
                    if (lane == 0) nbad++;
                }
                av[c] *= ipj;
                if (lane == 0) ipiv[c] = ipj;
#pragma unroll
                for (int k = c + 1; k < NB; k++) av[k] -= av[c] * readlane_f(av[c], k);
                // one column's lane reads at a time (scalar registers)
                __builtin_amdgcn_sched_barrier(0);
            }
            if (lane < NB) {
#pragma unroll
                for (int k = 0; k < NB; k++) ldc[k * NB + lane] = av[k];
                // the factor keeps 1 / L[c][c] on its diagonal: the triangular
                // solves multiply by it (same wavefront, after the writes above
                // and lane 0's ipiv writes)
                ldc[lane * NB + lane] = ipiv[lane];
            }
        }
        __syncthreads();
        // rows below the block, one thread per row, right-looking within the
        // row: L[r][c] = v_c / L[c][c], then v_k -= L[r][c] L[k][c] for k > c
        for (int rb = tid; rb < below; rb += kThreads) {
            double* pr = pan + rb * LDP;
            double v[NB];
#pragma unroll
            for (int k = 0; k < NB; k++) v[k] = pr[k];
#pragma unroll
            for (int c = 0; c < NB; c++) {
                const double xc = v[c] * ipiv[c];
                pr[c] = xc;
                const double* lcol = ldc + c * NB;   // L[k][c]
#pragma unroll
                for (int k = c + 1; k < NB; k++) v[k] -= xc * lcol[k];
                // keeps the 32 columns' table reads from being hoisted together (register spills)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        __syncthreads();
        {
            Walk2 wk3(tid, nb + below);
            for (int e = tid; e < nb * (nb + below); e += kThreads) {
                const int c = wk3.c, r = wk3.r;
                if (r >= c)
                    Lm[(size_t)(p0 + c) * P + p0 + r] = r < nb ? ldc[c * NB + r] : pan[(r - nb) * LDP + c];
                wk3.next();
            }
        }
        // trailing update of the lower triangle, A[i][k] -= sum_c L[i][c] L[k][c],
        // on FP64 MFMA in 16 x 16 tiles spread over the wavefronts: lane 16 kk + pp
        // holds L[16 K + pp][4 s + kk] as the A operand and L[16 I + pp][4 s + kk]
        // as the B operand, so acc[v] sums row i = 16 I + pp, column
        // k = 16 K + 4 v + kk, and consecutive lanes update consecutive rows of the
        // column-major factor.  The old entries are loaded before the MFMA chain.
        // Rows past the end of the panel are read (inside the LDS allocation) but
        // their sums are not stored.
        if (below > 0) {
            const int n16 = (below + 15) / 16;
            const int ntiles = n16 * (n16 + 1) / 2;
            const size_t q0 = (size_t)(p0 + nb);   // first trailing row / column
            const int pp = lane & 15, kk = lane >> 4;
            for (int t = wv; t < ntiles; t += kThreads / 64) {
                int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
                while (ti * (ti + 1) / 2 > t) ti--;
                while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
                const int tk = t - ti * (ti + 1) / 2;   // tile (ti, tk), tk <= ti
                const int i = 16 * ti + pp;
                double old[4];
                bool ok[4];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int k = 16 * tk + 4 * v + kk;
                    ok[v] = i < below && k <= i;
                    old[v] = ok[v] ? Lm[(q0 + k) * P + q0 + i] : 0.0;
                }
                const double* pa = pan + (16 * tk + pp) * LDP + kk;   // rows k
                const double* pb = pan + (16 * ti + pp) * LDP + kk;   // rows i
                double4v acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s4 = 0; s4 < NB / 4; s4++)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[4 * s4], pb[4 * s4], acc, 0, 0, 0);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int k = 16 * tk + 4 * v + kk;
                    if (ok[v]) Lm[(q0 + k) * P + q0 + i] = old[v] - acc[v];
                }
            }
        }
        __syncthreads();
    }

    // ---------------------------------------------------------------- L t = b
    for (int p0 = 0; p0 < P; p0 += NB) {
        const int nb = min(NB, P - p0);
        // diagonal block into LDS: pan[j * NB + l] = L[p0 + l][p0 + j]
        for (int e = tid; e < nb * nb; e += kThreads) {
            const int j = e / nb, l = e % nb;
            pan[j * NB + l] = Lm[(size_t)(p0 + j) * P + p0 + l];
        }
        __syncthreads();
        if (wv == 0) {
            // lane l owns row p0 + l; the diagonal holds 1 / L[j][j]
            double tR = lane < nb ? bR[p0 + lane] : 0.0, tI = lane < nb ? bI[p0 + lane] : 0.0;
            for (int j = 0; j < nb; j++) {
                const double rjj = pan[j * NB + j];
                const double xR = readlane_f(tR, j) * rjj, xI = readlane_f(tI, j) * rjj;
                if (lane == j) {
                    tR = xR;
                    tI = xI;
                } else if (lane > j && lane < nb) {
                    const double l = pan[j * NB + lane];
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
#pragma unroll 8
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
        // diagonal block into LDS: pan[j * NB + l] = L[p0 + l][p0 + j]
        for (int e = tid; e < nb * nb; e += kThreads) {
            const int j = e / nb, l = e % nb;
            pan[j * NB + l] = Lm[(size_t)(p0 + j) * P + p0 + l];
        }
        __syncthreads();
        if (wv == 0) {
            double tR = lane < nb ? bR[p0 + lane] : 0.0, tI = lane < nb ? bI[p0 + lane] : 0.0;
            for (int j = nb - 1; j >= 0; j--) {
                const double rjj = pan[j * NB + j];
                const double xR = readlane_f(tR, j) * rjj, xI = readlane_f(tI, j) * rjj;
                if (lane == j) {
                    tR = xR;
                    tI = xI;
                } else if (lane < j) {
                    const double l = pan[lane * NB + j];   // L[j][lane]
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
        if (a.status) a.status[0] += nbad;
    }
}

// ---------------------------------------------------------------------------
// Solver type 1 (src/TDVMC.cpp:1763-1822): Householder QR with full pivoting,
// the decomposition of Eigen::FullPivHouseholderQR, with its conventions
// (pivot = largest magnitude of the trailing block, first in column-major
// order on ties; stop when it falls to n eps of the first pivot; trailing-row
// and whole-column swaps; v = (1, tail / (c0 - beta)), tau = (beta - c0) /
// beta, beta = -sign(c0) |x|; rank from n eps max |R_kk|), then the mean of
// each solution vector is subtracted.  For an S made singular by the exactly
// equal O of k and -k this picks the same solution as the reference's
// decomposition; Cholesky would not.  The matrix is held column-major in the
// work buffer, one thread per column for the reflector updates.

__device__ void argmax_reduce(double& v, int& e, double* rv, int* re)
{
    // larger value wins, equal values: smaller column-major index
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o, 64);
        const int e2 = __shfl_xor(e, o, 64);
        if (v2 > v || (v2 == v && e2 < e)) {
            v = v2;
            e = e2;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        rv[wv] = v;
        re[wv] = e;
    }
    __syncthreads();
    v = rv[0];
    e = re[0];
    for (int w = 1; w < kThreads / 64; w++)
        if (rv[w] > v || (rv[w] == v && re[w] < e)) {
            v = rv[w];
            e = re[w];
        }
}

__device__ __forceinline__ void solve_qr_body(const SolveArgs& a)
{
    const int P = a.P, tid = threadIdx.x;
    double* Q = a.work;                  // column-major: Q[j * P + i] = A[i][j]
    double* bR = a.work + (size_t)P * P;
    double* bI = bR + P;
    double* sc = bI + P;
    double* tau = sc + P;
    extern __shared__ __align__(16) unsigned char smem[];
    int* rt = reinterpret_cast<int*>(smem);   // row transpositions [P]
    int* ct = rt + P;                          // column transpositions [P]
    __shared__ double red[kThreads / 64];
    __shared__ double rv[kThreads / 64];
    __shared__ int re[kThreads / 64];
    const double* mO = a.mean;
    const double eR = a.mean[P], eI = a.mean[P + 1];
    const double* FR = a.F;
    const double* FI = a.F + P;
    const double prec = P * 2.220446049250313e-16;

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
    Walk2 wk(tid, P);
    for (size_t e = tid; e < (size_t)P * P; e += kThreads) {
        const int i = wk.r, j = wk.c;
        wk.next();
        double v = a.S[(size_t)i * P + j];
        if (a.precondition) v /= (sc[i] * sc[j]);
        if (i == j) v += a.eps;
        Q[e] = v;
    }
    for (int i = tid; i < P; i += kThreads) {
        if (a.precondition) {
            bR[i] /= sc[i];
            bI[i] /= sc[i];
        }
    }
    __syncthreads();

    double biggest = 0.0, maxpivot = 0.0;
    int nz = P;
    for (int k = 0; k < P; k++) {
        const int m = P - k;
        double bv = -1.0;
        int be = 0x7fffffff;
        for (int e = tid; e < m * m; e += kThreads) {
            const int jj = e / m, ii = e - jj * m;   // e increases in column-major order
            const double v = fabs(Q[(size_t)(k + jj) * P + k + ii]);
            if (v > bv) {
                bv = v;
                be = e;
            }
        }
        argmax_reduce(bv, be, rv, re);
        if (k == 0) biggest = bv;
        if (bv <= biggest * prec) {
            nz = k;
            break;
        }
        const int bj = k + be / m, bi = k + be % m;
        if (tid == 0) {
            rt[k] = bi;
            ct[k] = bj;
        }
        if (bi != k)
            for (int j = k + tid; j < P; j += kThreads) {
                const double t = Q[(size_t)j * P + k];
                Q[(size_t)j * P + k] = Q[(size_t)j * P + bi];
                Q[(size_t)j * P + bi] = t;
            }
        __syncthreads();
        if (bj != k)
            for (int i = tid; i < P; i += kThreads) {
                const double t = Q[(size_t)k * P + i];
                Q[(size_t)k * P + i] = Q[(size_t)bj * P + i];
                Q[(size_t)bj * P + i] = t;
            }
        __syncthreads();
        double* col = Q + (size_t)k * P;
        double tl = 0.0;
        for (int i = k + 1 + tid; i < P; i += kThreads) tl += col[i] * col[i];
        tl = block_reduce(tl, red);
        const double c0 = col[k];
        double beta, tk;
        if (tl <= 2.2250738585072014e-308) {
            tk = 0.0;
            beta = c0;
            for (int i = k + 1 + tid; i < P; i += kThreads) col[i] = 0.0;
        } else {
            beta = sqrt(c0 * c0 + tl);
            if (c0 >= 0.0) beta = -beta;
            const double den = c0 - beta;
            for (int i = k + 1 + tid; i < P; i += kThreads) col[i] /= den;
            tk = (beta - c0) / beta;
        }
        maxpivot = fmax(maxpivot, fabs(beta));
        __syncthreads();
        if (tid == 0) {
            col[k] = beta;
            tau[k] = tk;
        }
        // H applied to the trailing columns, one thread per column
        for (int j = k + 1 + tid; j < P; j += kThreads) {
            double* cj = Q + (size_t)j * P;
            double w = cj[k];
            for (int i = k + 1; i < P; i++) w += col[i] * cj[i];
            cj[k] -= tk * w;
            for (int i = k + 1; i < P; i++) cj[i] -= (tk * col[i]) * w;
        }
        __syncthreads();
    }
    for (int k = nz + tid; k < P; k += kThreads) {
        rt[k] = k;
        ct[k] = k;
    }
    __syncthreads();
    int rank = 0;
    for (int i = 0; i < nz; i++) rank += fabs(Q[(size_t)i * P + i]) > maxpivot * prec;

    // c = Q^T (row-swapped) b for the first rank reflectors
    for (int k = 0; k < rank; k++) {
        if (tid == 0 && rt[k] != k) {
            double t = bR[k]; bR[k] = bR[rt[k]]; bR[rt[k]] = t;
            t = bI[k]; bI[k] = bI[rt[k]]; bI[rt[k]] = t;
        }
        __syncthreads();
        const double* col = Q + (size_t)k * P;
        const double ck_r = bR[k], ck_i = bI[k];   // read before the barriers of the reductions
        double wr = 0.0, wi = 0.0;
        for (int i = k + 1 + tid; i < P; i += kThreads) {
            wr += col[i] * bR[i];
            wi += col[i] * bI[i];
        }
        wr = ck_r + block_reduce(wr, red);
        wi = ck_i + block_reduce(wi, red);
        const double tk = tau[k];
        for (int i = k + 1 + tid; i < P; i += kThreads) {
            bR[i] -= (tk * col[i]) * wr;
            bI[i] -= (tk * col[i]) * wi;
        }
        if (tid == 0) {
            bR[k] -= tk * wr;
            bI[k] -= tk * wi;
        }
        __syncthreads();
    }
    // R z = c on the leading rank x rank block (column-oriented), zero beyond
    for (int i = rank - 1; i >= 0; i--) {
        const double rii = Q[(size_t)i * P + i];
        const double zr = bR[i] / rii, zi = bI[i] / rii;
        __syncthreads();
        for (int l = tid; l < i; l += kThreads) {
            const double r = Q[(size_t)i * P + l];
            bR[l] -= r * zr;
            bI[l] -= r * zi;
        }
        if (tid == 0) {
            bR[i] = zr;
            bI[i] = zi;
        }
        __syncthreads();
    }
    for (int i = rank + tid; i < P; i += kThreads) {
        bR[i] = 0.0;
        bI[i] = 0.0;
    }
    __syncthreads();
    if (tid == 0)
        for (int k = P - 1; k >= 0; k--) {
            const int c = ct[k];
            if (c != k) {
                double t = bR[k]; bR[k] = bR[c]; bR[c] = t;
                t = bI[k]; bI[k] = bI[c]; bI[c] = t;
            }
        }
    __syncthreads();

    // mean subtraction, phi-dot on the scaled solution, unscaling
    double sr = 0.0, si = 0.0;
    for (int i = tid; i < P; i += kThreads) {
        sr += bR[i];
        si += bI[i];
    }
    const double mr = block_reduce(sr, red) / P;
    const double mi = block_reduce(si, red) / P;
    double pr = 0.0, pi = 0.0;
    for (int i = tid; i < P; i += kThreads) {
        const double xr = bR[i] - mr, xi = bI[i] - mi;
        pr -= mO[i] * xr;
        pi -= mO[i] * xi;
        a.xdot[i] = xr / sc[i];
        a.xdot[P + i] = xi / sc[i];
    }
    pr = block_reduce(pr, red);
    pi = block_reduce(pi, red);
    if (tid == 0) {
        if (a.imag_time == 0)
            pi -= eR;
        else
            pr -= eR;
        a.xdot[2 * P] = pr;
        a.xdot[2 * P + 1] = pi;
    }
}

template <int NB>
__global__ __launch_bounds__(kThreads) void solve_kernel(SolveArgs a)
{
    solve_body<NB>(a);
}
__global__ __launch_bounds__(kThreads) void solve_qr_kernel(SolveArgs a)
{
    solve_qr_body(a);
}
// one workgroup per system of a batch
template <int NB>
__global__ __launch_bounds__(kThreads) void solve_batch_kernel(const SolveArgs* __restrict__ args)
{
    const SolveArgs a = args[blockIdx.x];
    if (a.mean_sub)
        solve_qr_body(a);
    else
        solve_body<NB>(a);
}

// dynamic LDS of the Cholesky: the rows below a block at stride NB + 1 (also
// holds the NB x NB blocks of the triangular solves)
size_t solve_lds_bytes(int NB, int P) { return (size_t)(NB + 1) * std::max(P, NB) * sizeof(double); }
constexpr size_t kMaxPanelLds = 150 * 1024;   // beside the static LDS (NB x NB block, reductions)
// Block width: 16 up to P = 340, 32 above (measured crossover near P = 345:
// at P = 296, 373 against 395 us; at P = 420, 738 against 705 us), 16 again
// where the NB = 32 panel no longer fits in LDS
constexpr int kNb16MaxP = 340;

template <int NB>
hipError_t launch_nb(const SolveArgs& a, hipStream_t st)
{
    const size_t lds = solve_lds_bytes(NB, a.P);
    auto kern = solve_kernel<NB>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(1), dim3(kThreads), lds, st, a);
    return hipGetLastError();
}

}  // namespace

template <int NB>
hipError_t launch_batch_nb(const SolveArgs* args_dev, int n, int P_max, hipStream_t st)
{
    // LDS: Cholesky panels or the QR transpositions (2 P ints)
    const size_t lds = std::max(solve_lds_bytes(NB, P_max), 2 * (size_t)P_max * sizeof(int));
    auto kern = solve_batch_kernel<NB>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(n), dim3(kThreads), lds, st, args_dev);
    return hipGetLastError();
}

hipError_t launch_solve_batch(const SolveArgs* args_dev, int n, int P_max, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (P_max > kNb16MaxP && solve_lds_bytes(32, P_max) <= kMaxPanelLds)
        return launch_batch_nb<32>(args_dev, n, P_max, st);
    if (solve_lds_bytes(16, P_max) <= kMaxPanelLds) return launch_batch_nb<16>(args_dev, n, P_max, st);
    return hipErrorInvalidValue;
}

int solve_max_params(int solver_type)
{
    if (solver_type == 1) return (int)(kMaxPanelLds / (2 * sizeof(int)));
    return (int)(kMaxPanelLds / (17 * sizeof(double)));
}

hipError_t launch_solve(const SolveArgs& a, hipStream_t st)
{
    if (a.P <= 0) return hipErrorInvalidValue;
    if (a.mean_sub) {
        hipLaunchKernelGGL(solve_qr_kernel, dim3(1), dim3(kThreads), 2 * (size_t)a.P * sizeof(int), st, a);
        return hipGetLastError();
    }
    if (a.P > kNb16MaxP && solve_lds_bytes(32, a.P) <= kMaxPanelLds) return launch_nb<32>(a, st);
    if (solve_lds_bytes(16, a.P) <= kMaxPanelLds) return launch_nb<16>(a, st);
    return hipErrorInvalidValue;
}

}  // namespace tdvmc
