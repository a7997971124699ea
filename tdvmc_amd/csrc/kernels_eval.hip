// Wavefunction evaluation of walker configurations: ln psi, O_k = d ln psi /
// d alpha_k, local energy E_L, and (optionally) the per-particle gradient and
// Laplacian of ln psi.  One workgroup of 4 wavefronts evaluates one walker.
//
// Restates src/PhysicalSystems/BosonsBulk.cpp:179-218 (basis sums O_k and the
// outer-pair count), :220-336 (pair potential, gradient and Laplacian pieces),
// :349-458 (E_L and the nine 'other' values) and :522-539 (exponent), adds the
// k-space Jastrow sum_k alpha_k |rho_k|^2 and a one-body Gaussian factor.
//
// Work split (N particles, K k-vectors, 256 threads):
//   pairs    : every unordered pair once, on a half ring: lane = row n of a
//              64-row block, partner i = (n + s) mod N, the s range split over
//              the 4 waves; own-row sums stay in registers, partner sums go
//              to LDS with conflict-free atomics (distinct i per lane);
//   rho_k    : thread = k, loop over particles (LDS broadcast reads);
//   k terms  : lane = row n, the k range split over the 4 waves;
//   E_L      : thread = particle, FP64 block reduction.
// The spline basis sums go to an LDS histogram [bin][16] (4 lanes per column).
// Coordinates are staged in LDS as structure-of-arrays (bank-conflict-free
// reads of consecutive particles); every LDS accumulator is FP64 in both
// precisions (FP64 LDS atomics run at full rate, FP32 ones stall the LDS
// pipeline on this part, and FP64 sums keep O_k accurate in the mixed form).

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kHistCols = 16;

template <typename TO>
struct EvalOut {
    double* expo;     // [W]
    double* ER;       // [W]
    double* EI;       // [W]
    TO* O;            // [W][P]
    double* other;    // [W][9] or null
    double* grad;     // [W][N][D][2] or null (PP)
    double* lap;      // [W][N][2] or null (PP)
    double* outer;    // [W] count of pairs beyond the cutoff, or null
    double phiR;
    __device__ __forceinline__ void O_store(int64_t w, int P, int k, double v) const { O[w * P + k] = (TO)v; }
};

__device__ __forceinline__ void lds_add(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double block_sum(double v, double* red)
{
    // wave reduction, then across the 4 waves through LDS
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) s += red[w];
    return s;
}

template <typename T, int D, bool PP, bool PER, typename TO>
__global__ __launch_bounds__(kThreads) void eval_kernel(KSys sy, KTables<T> tb, const double* __restrict__ Rg,
                                                         int64_t W, EvalOut<TO> out, const T* __restrict__ rho_in)
{
    const int64_t w = blockIdx.x;
    if (w >= W) return;
    const int N = sy.N, Ps = sy.P_s, K = sy.K, S = Ps + 2;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (scalar registers)

    extern __shared__ __align__(16) unsigned char smem[];
    // LDS carve-up (sizes mirrored by eval_lds_bytes)
    double* red = reinterpret_cast<double*>(smem);             // 8 doubles
    double* gsum = red + 8;                                    // N * 2D : (Re g, Im g)
    double* lsum = gsum + (size_t)N * 2 * D;                   // N * 2 (PP only)
    double* hist = lsum + (PP ? (size_t)N * 2 : 0);            // S * 16
    T* pos = reinterpret_cast<T*>(hist + (size_t)S * kHistCols);   // D * N (structure of arrays)
    T* rho = pos + (size_t)N * D;                              // 2K
    T* spl = rho + 2 * (size_t)K;                              // (Ps-1) * 8
    const T* kv = tb.kv;                                       // K * 4 (global, read with uniform k)

    // periodic box: coordinates in box units (x / L, wrapped into the primary
    // cell -- every quantity is invariant under it), so the minimum image is
    // d - rint(d); lengths below are scaled to match
    const double scale = PER ? sy.Linv : 1.0;
    const double* Rw = Rg + w * (size_t)N * D;
    for (int i = tid; i < N; i += kThreads) {
#pragma unroll
        for (int a = 0; a < D; a++) {
            double x = Rw[(size_t)i * D + a] * scale;
            if (PER) x = x - rint(x);
            pos[a * N + i] = (T)x;
        }
#pragma unroll
        for (int c = 0; c < 2 * D; c++) gsum[i * 2 * D + c] = 0.0;
        if (PP) {
            lsum[2 * i] = 0.0;
            lsum[2 * i + 1] = 0.0;
        }
    }
    for (int i = tid; i < S * kHistCols; i += kThreads) hist[i] = 0.0;
    for (int i = tid; i < (Ps - 1) * 8; i += kThreads) spl[i] = tb.spl[i];
    __syncthreads();

    // ------------------------------------------------------------ pairs
    const T hinv = (T)sy.hinv, hsc = (T)(sy.hinv / scale), rc = (T)(sy.rc * scale);
    const T pcut = (T)(sy.pot_cutoff * scale);
    const T pa = (T)(sy.pot_a * scale), pb = (T)sy.pot_b;
    const T lapsc = (T)((sy.D - 1) * scale);   // (D-1)/r = (D-1) scale / r_box
    const int shalf = (N - 1) / 2;
    const int smax = (N % 2 == 0) ? N / 2 : shalf;    // s = N/2 only for n < N/2 (even N)
    const int chunk = (smax + kWaves - 1) / kWaves;
    const int s_lo = 1 + wave * chunk;
    const int s_hi = min(smax, (wave + 1) * chunk);
    double vpot = 0.0, outer = 0.0, lap_tot_R = 0.0, lap_tot_I = 0.0;

    for (int rb = 0; rb < N; rb += 64) {
        const int n = rb + lane;
        if (n < N) {
            T rn[D];
#pragma unroll
            for (int a = 0; a < D; a++) rn[a] = pos[a * N + n];
            T gR[D], gI[D];
#pragma unroll
            for (int a = 0; a < D; a++) gR[a] = gI[a] = (T)0;
            T lR = (T)0, lI = (T)0;
            int i = (n + s_lo) % N;
            for (int s = s_lo; s <= s_hi; s++) {
                const bool live = !(2 * s == N && n >= N / 2);
                if (live) {
                    T d[D];
                    T r2 = (T)0;
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        d[a] = rn[a] - pos[a * N + i];
                        if (PER) d[a] -= Prec<T>::rint_(d[a]);
                        r2 += d[a] * d[a];
                    }
                    const T r = Prec<T>::sqrt_(r2);
                    if (sy.pot_type != 0 && r <= pcut) {
                        T v;
                        if (sy.pot_type == 1) {
                            v = (r < pa) ? pb : (T)0;
                        } else if (sy.pot_type == 2) {
                            T x = r / pa;
                            v = pb * Prec<T>::exp_((T)-0.5 * x * x);
                        } else {
                            T q = pa * pa * Prec<T>::rcp_(r2);
                            T q6 = q * q * q;
                            v = (T)4 * pb * q6 * (q6 - (T)1);
                        }
                        vpot += (double)v;
                    }
                    if (r <= rc) {
                        T t;
                        const int j = knot_interval(r, hsc, Ps, &t);
                        const T u = (T)1 - t;
                        const T t2 = t * t, t3 = t2 * t;
                        // the four non-zero uniform cubic B-splines; the third
                        // follows from the partition of unity
                        const T b0 = u * u * u * (T)(1.0 / 6.0);
                        const T b1 = ((T)3 * t3 - (T)6 * t2 + (T)4) * (T)(1.0 / 6.0);
                        const T b3 = t3 * (T)(1.0 / 6.0);
                        const T b2 = (T)1 - b0 - b1 - b3;
                        double* hb = hist + (size_t)j * kHistCols + (lane & (kHistCols - 1));
                        lds_add(hb, (double)b0);
                        lds_add(hb + kHistCols, (double)b1);
                        lds_add(hb + 2 * kHistCols, (double)b2);
                        lds_add(hb + 3 * kHistCols, (double)b3);
                        const T* c = spl + j * 8;
                        // u'(r) = (A1 + 2 A2 t + 3 A3 t^2) / h ; u''(r) = (2 A2 + 6 A3 t) / h^2
                        const T upR = (c[1] + t * ((T)2 * c[2] + (T)3 * c[3] * t)) * hinv;
                        const T upI = (c[5] + t * ((T)2 * c[6] + (T)3 * c[7] * t)) * hinv;
                        const T uppR = ((T)2 * c[2] + (T)6 * c[3] * t) * hinv * hinv;
                        const T uppI = ((T)2 * c[6] + (T)6 * c[7] * t) * hinv * hinv;
                        const T rinv = Prec<T>::rcp_(r);
                        const T fR = upR * rinv, fI = upI * rinv;
                        double* gp = gsum + (size_t)i * 2 * D;
#pragma unroll
                        for (int a = 0; a < D; a++) {
                            gR[a] += fR * d[a];
                            gI[a] += fI * d[a];
                            lds_add(gp + a, (double)(-fR * d[a]));
                            lds_add(gp + D + a, (double)(-fI * d[a]));
                        }
                        const T lfR = uppR + lapsc * fR;
                        const T lfI = uppI + lapsc * fI;
                        if (PP) {
                            lR += lfR;
                            lI += lfI;
                            lds_add(lsum + 2 * i, (double)lfR);
                            lds_add(lsum + 2 * i + 1, (double)lfI);
                        } else {
                            lap_tot_R += 2.0 * (double)lfR;
                            lap_tot_I += 2.0 * (double)lfI;
                        }
                    } else {
                        outer += 1.0;
                    }
                }
                i = (i + 1 == N) ? 0 : i + 1;
            }
            double* gp = gsum + (size_t)n * 2 * D;
#pragma unroll
            for (int a = 0; a < D; a++) {
                lds_add(gp + a, (double)gR[a]);
                lds_add(gp + D + a, (double)gI[a]);
            }
            if (PP) {
                lds_add(lsum + 2 * n, (double)lR);
                lds_add(lsum + 2 * n + 1, (double)lI);
            }
        }
    }

    // ------------------------------------------------------------ rho_k
    const T ksc = (T)(1.0 / scale);   // phase in revolutions = (kv / scale) . s
    double ex_k = 0.0;
    for (int k = tid; k < K; k += kThreads) {
        T re = (T)0, im = (T)0;
        if (rho_in) {
            re = rho_in[((size_t)w * K + k) * 2];
            im = rho_in[((size_t)w * K + k) * 2 + 1];
        } else {
            T q[D];
#pragma unroll
            for (int a = 0; a < D; a++) q[a] = kv[k * 4 + a] * ksc;
            for (int n = 0; n < N; n++) {
                T ph = (T)0;
#pragma unroll
                for (int a = 0; a < D; a++) ph += q[a] * pos[a * N + n];
                T sn, cs;
                Prec<T>::sincos_rev(ph, &sn, &cs);
                re += cs;
                im += sn;
            }
        }
        rho[2 * k] = re;
        rho[2 * k + 1] = im;
        const double ok = (double)re * (double)re + (double)im * (double)im;
        ex_k += tb.uR[sy.off_k + k] * ok;
        out.O_store(w, sy.P, sy.off_k + k, ok);
    }
    __syncthreads();

    // ------------------------------------------------------------ k-space gradient / Laplacian
    if (K > 0) {
        const int kchunk = (K + kWaves - 1) / kWaves;
        const int k_lo = wave * kchunk, k_hi = min(K, (wave + 1) * kchunk);
        for (int rb = 0; rb < N; rb += 64) {
            const int n = rb + lane;
            if (n < N) {
                T rn[D];
#pragma unroll
                for (int a = 0; a < D; a++) rn[a] = pos[a * N + n];
                T gR[D], gI[D];
#pragma unroll
                for (int a = 0; a < D; a++) gR[a] = gI[a] = (T)0;
                T lR = (T)0, lI = (T)0;
                // grad |rho_k|^2 = -2 k Im(e^{ikr} conj rho), k = 2 pi kv;
                // lap |rho_k|^2 = 2 |k|^2 (1 - Re(e^{ikr} conj rho)); the
                // constant factors are applied after the loop
                for (int k = k_lo; k < k_hi; k++) {
                    const T* q = tb.kv + k * 4;
                    T ph = (T)0;
#pragma unroll
                    for (int a = 0; a < D; a++) ph += (q[a] * ksc) * rn[a];
                    T sn, cs;
                    Prec<T>::sincos_rev(ph, &sn, &cs);
                    const T rr = rho[2 * k], ri = rho[2 * k + 1];
                    const T pre = cs * rr + sn * ri;    // Re(e^{ikr} conj(rho))
                    const T pim = sn * rr - cs * ri;    // Im(e^{ikr} conj(rho))
                    const T aR = tb.alpha[2 * k], aI = tb.alpha[2 * k + 1];
                    const T sR = aR * pim, sI = aI * pim;
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        gR[a] += sR * q[a];
                        gI[a] += sI * q[a];
                    }
                    const T lk = q[3] * ((T)1 - pre);
                    lR += aR * lk;
                    lI += aI * lk;
                }
#pragma unroll
                for (int a = 0; a < D; a++) {
                    gR[a] *= (T)(-2.0 * 6.283185307179586);
                    gI[a] *= (T)(-2.0 * 6.283185307179586);
                }
                lR *= (T)2;
                lI *= (T)2;
                double* gp = gsum + (size_t)n * 2 * D;
#pragma unroll
                for (int a = 0; a < D; a++) {
                    lds_add(gp + a, (double)gR[a]);
                    lds_add(gp + D + a, (double)gI[a]);
                }
                if (PP) {
                    lds_add(lsum + 2 * n, (double)lR);
                    lds_add(lsum + 2 * n + 1, (double)lI);
                } else {
                    lap_tot_R += (double)lR;
                    lap_tot_I += (double)lI;
                }
            }
        }
    }
    __syncthreads();

    // ------------------------------------------------------------ assembly
    const double obR = sy.one_body ? tb.uR[0] : 0.0, obI = sy.one_body ? tb.uI[0] : 0.0;
    double sR1 = 0, sI1 = 0, sR1I1 = 0, sR2 = lap_tot_R, sI2 = lap_tot_I, sr2 = 0, vext = 0;
    for (int n = tid; n < N; n += kThreads) {
        double gr[D], gi[D];
        double r2n = 0.0;
#pragma unroll
        for (int a = 0; a < D; a++) {
            // one-body factor uses the raw (unwrapped, open-space) coordinate
            const double x = Rw[(size_t)n * D + a];
            gr[a] = gsum[n * 2 * D + a] - 2.0 * obR * x;
            gi[a] = gsum[n * 2 * D + D + a] - 2.0 * obI * x;
            r2n += x * x;
        }
        sr2 += r2n;
        double lr = 0.0, li = 0.0;
        if (PP) {
            lr = lsum[2 * n] - 2.0 * D * obR;
            li = lsum[2 * n + 1] - 2.0 * D * obI;
            sR2 += lr;
            sI2 += li;
        } else {
            sR2 += -2.0 * D * obR;
            sI2 += -2.0 * D * obI;
        }
        double dRI = 0, nR = 0, nI = 0;
#pragma unroll
        for (int a = 0; a < D; a++) {
            dRI += gr[a] * gi[a];
            nR += gr[a] * gr[a];
            nI += gi[a] * gi[a];
        }
        sR1I1 += 2.0 * dRI;
        sR1 += nR;
        sI1 += nI;
        if (PP) {
#pragma unroll
            for (int a = 0; a < D; a++) {
                out.grad[(((size_t)w * N + n) * D + a) * 2] = gr[a];
                out.grad[(((size_t)w * N + n) * D + a) * 2 + 1] = gi[a];
            }
            out.lap[((size_t)w * N + n) * 2] = lr;
            out.lap[((size_t)w * N + n) * 2 + 1] = li;
        }
    }
    if (sy.trap_omega != 0.0) vext = 0.5 * sy.trap_omega * sy.trap_omega * sr2;

    // spline basis sums -> O_pair (BC map), exponent contribution
    double ex_p = 0.0;
    for (int k = tid; k < Ps; k += kThreads) {
        auto B = [&](int s) {
            double v = 0.0;
#pragma unroll
            for (int c = 0; c < kHistCols; c++) v += hist[s * kHistCols + c];
            return v;
        };
        double o;
        if (k == 0) o = B(1);
        else if (k == 1) o = B(0) + B(2);
        else if (k <= Ps - 3) o = B(k + 1);
        else if (k == Ps - 2) o = B(Ps - 1) + B(Ps + 1);
        else o = B(Ps);
        ex_p += tb.uR[sy.off_pair + k] * o;
        out.O_store(w, sy.P, sy.off_pair + k, o);
    }

    sR1 = block_sum(sR1, red);
    sI1 = block_sum(sI1, red);
    sR1I1 = block_sum(sR1I1, red);
    sR2 = block_sum(sR2, red);
    sI2 = block_sum(sI2, red);
    sr2 = block_sum(sr2, red);
    vext = block_sum(vext, red);
    vpot = block_sum(vpot, red);
    outer = block_sum(outer, red);
    double ex = block_sum(ex_k + ex_p, red);
    if (tid == 0) {
        if (sy.one_body) {
            out.O_store(w, sy.P, 0, -sr2);
            ex += tb.uR[0] * (-sr2);
        }
        ex += tb.uR[sy.outer_idx] * outer;
        const double kinR = -(sR1 - sI1 + sR2) * sy.hbar2_2m;
        const double kinI = -(sR1I1 + sI2) * sy.hbar2_2m;
        out.expo[w] = ex;
        if (out.outer) out.outer[w] = outer;
        out.ER[w] = kinR + vpot + vext;
        out.EI[w] = kinI;
        if (out.other) {
            double* o = out.other + w * 9;
            o[0] = kinR;
            o[1] = vpot;
            o[2] = exp(ex + out.phiR);
            o[3] = ex;
            o[4] = sR1;
            o[5] = sI1;
            o[6] = sR2;
            o[7] = sI2;
            o[8] = sR1I1;
        }
    }
}

}  // namespace

template <typename T>
size_t eval_lds_bytes(const KSys& sy, bool per_particle)
{
    const size_t nd = (size_t)sy.N * 2 * sy.D + (per_particle ? 2 * (size_t)sy.N : 0) + (size_t)(sy.P_s + 2) * kHistCols;
    const size_t nt = (size_t)sy.N * sy.D + 2 * (size_t)sy.K + (size_t)(sy.P_s - 1) * 8;
    return 64 + nd * sizeof(double) + nt * sizeof(T);
}

template <typename T, int D, bool PP, bool PER, typename TO>
static hipError_t launch_eval_k(const KSys& sy, const KTables<T>& tb, const double* R, int64_t W,
                                const EvalOut<TO>& o, const T* rho_in, hipStream_t st)
{
    const size_t lds = eval_lds_bytes<T>(sy, PP);
    auto kern = eval_kernel<T, D, PP, PER, TO>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)W), dim3(kThreads), lds, st, sy, tb, R, W, o, rho_in);
    return hipGetLastError();
}

template <typename T, int D, typename TO>
static hipError_t launch_eval_d(const KSys& sy, const KTables<T>& tb, const double* R, int64_t W,
                                const EvalOut<TO>& o, const T* rho_in, bool pp, hipStream_t st)
{
    if (sy.periodic)
        return pp ? launch_eval_k<T, D, true, true, TO>(sy, tb, R, W, o, rho_in, st)
                  : launch_eval_k<T, D, false, true, TO>(sy, tb, R, W, o, rho_in, st);
    return pp ? launch_eval_k<T, D, true, false, TO>(sy, tb, R, W, o, rho_in, st)
              : launch_eval_k<T, D, false, false, TO>(sy, tb, R, W, o, rho_in, st);
}

template <typename T, typename TO>
hipError_t launch_eval(const KSys& sy, const KTables<T>& tb, const double* R, int64_t W, double* expo, double* ER,
                       double* EI, TO* O, double* other, double* grad, double* lap, double* outer, double phiR,
                       const T* rho_in, hipStream_t st)
{
    if (W <= 0) return hipSuccess;
    EvalOut<TO> o{expo, ER, EI, O, other, grad, lap, outer, phiR};
    const bool pp = grad != nullptr;
    switch (sy.D) {
    case 1: return launch_eval_d<T, 1, TO>(sy, tb, R, W, o, rho_in, pp, st);
    case 2: return launch_eval_d<T, 2, TO>(sy, tb, R, W, o, rho_in, pp, st);
    case 3: return launch_eval_d<T, 3, TO>(sy, tb, R, W, o, rho_in, pp, st);
    default: return hipErrorInvalidValue;
    }
}

template size_t eval_lds_bytes<float>(const KSys&, bool);
template size_t eval_lds_bytes<double>(const KSys&, bool);
template hipError_t launch_eval<double, double>(const KSys&, const KTables<double>&, const double*, int64_t, double*,
                                                double*, double*, double*, double*, double*, double*, double*, double,
                                                const double*, hipStream_t);
template hipError_t launch_eval<float, float>(const KSys&, const KTables<float>&, const double*, int64_t, double*,
                                              double*, double*, float*, double*, double*, double*, double*, double,
                                              const float*, hipStream_t);
template hipError_t launch_eval<float, double>(const KSys&, const KTables<float>&, const double*, int64_t, double*,
                                               double*, double*, double*, double*, double*, double*, double*, double,
                                               const float*, hipStream_t);

}  // namespace tdvmc
