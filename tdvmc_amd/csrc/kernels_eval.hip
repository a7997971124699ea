// Wavefunction evaluation of walker configurations: ln psi, O_k = d ln psi /
// d alpha_k, local energy E_L, and (optionally) the per-particle gradient and
// Laplacian of ln psi.  One workgroup of 4 wavefronts evaluates one walker.
//
// Restates src/PhysicalSystems/BosonsBulk.cpp:179-218 (basis sums O_k and the
// outer-pair count), :220-336 (pair potential, gradient and Laplacian pieces),
// :349-458 (E_L and the nine 'other' values) and :522-539 (exponent), adds the
// k-space Jastrow sum_k alpha_k |rho_k|^2 (k and -k share |rho_k|^2 and are
// handled once with the summed amplitude) and a one-body Gaussian factor.
//
// Work split (N particles, K k-vectors, 256 threads):
//   pairs    : every unordered pair once, on a half ring: lane = row n of a
//              64-row block, partner i = (n + s) mod N, the s range split over
//              the 4 waves; own-row sums stay in registers, partner sums ride
//              a DPP register ring and reach LDS once per particle and step;
//   rho_k    : thread = k, loop over particles (LDS broadcast reads);
//   k terms  : lane = row n, the k range split over the 4 waves;
//   E_L      : thread = particle, FP64 block reduction.
// The pair's position t in its knot interval goes to an LDS histogram of the
// moments 1, t, t^2, t^3 per interval ([interval][4][8], 8 lanes per column);
// the B-spline basis sums are rebuilt from these moments at the end.
// Coordinates and per-particle sums are staged in LDS as structure-of-arrays
// (bank-conflict-free access to consecutive particles).  Every LDS
// accumulator is FP64 in both precisions: FP64 LDS atomics run several
// times faster than FP32 ones on this part, and FP64 sums keep O_k and the
// gradients accurate in the mixed form.

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kHistCols = 8;    // histogram columns (LDS footprint vs bank conflicts: occupancy wins)

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

template <typename V>
__device__ __forceinline__ void lds_add(V* p, V v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Whole-wavefront rotation by one lane towards lane 0 (lane l receives lane
// l + 1, lane 63 receives lane 0): DPP wave_rol:1 (GFX9 DPP control 0x134),
// one in-place v_mov_b32_dpp per dword.
constexpr int kDppWaveRol1 = 0x134;
__device__ __forceinline__ float wave_rotate(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kDppWaveRol1, 0xF, 0xF, false));
}
__device__ __forceinline__ double wave_rotate(double v)
{
    const int2 x = *reinterpret_cast<int2*>(&v);
    int2 y;
    y.x = __builtin_amdgcn_mov_dpp(x.x, kDppWaveRol1, 0xF, 0xF, false);
    y.y = __builtin_amdgcn_mov_dpp(x.y, kDppWaveRol1, 0xF, 0xF, false);
    return *reinterpret_cast<double*>(&y);
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

// One workgroup evaluates walker w of the system described by (sy, tb).
template <typename T, int D, bool PP, bool PER, typename TO>
__device__ __forceinline__ void eval_body(const KSys& sy, const KTables<T>& tb, const double* __restrict__ Rg,
                                          int64_t W, const EvalOut<TO>& out, const T* __restrict__ rho_in, int64_t w)
{
    if (w >= W) return;
    const int N = sy.N, Ps = sy.P_s, K = sy.Kh;   // K: +-k representatives
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (scalar registers)

    extern __shared__ __align__(16) unsigned char smem[];
    // LDS carve-up (sizes mirrored by eval_lds_bytes); per-particle arrays are
    // structure-of-arrays so consecutive particles sit in consecutive banks
    double* red = reinterpret_cast<double*>(smem);             // 8 doubles
    double* hist = red + 8;                                    // (Ps-1) * 4 * kHistCols
    double* gsum = hist + (size_t)(Ps - 1) * 4 * kHistCols;    // [2D][N]: Re g, Im g
    double* lsum = gsum + (size_t)N * 2 * D;                   // [2][N] (PP only)
    T* pos = reinterpret_cast<T*>(lsum + (PP ? (size_t)N * 2 : 0));   // [D][N]
    T* rho = pos + (size_t)N * D;                              // 2K
    // derivative table (Ps-1) * 8, 16-byte aligned for 128-bit reads
    T* dsp = reinterpret_cast<T*>(smem + ((reinterpret_cast<unsigned char*>(rho + 2 * (size_t)K) - smem + 15) &
                                          ~(size_t)15));
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
        for (int c = 0; c < 2 * D; c++) gsum[c * N + i] = 0.0;
        if (PP) {
            lsum[i] = 0.0;
            lsum[N + i] = 0.0;
        }
    }
    for (int i = tid; i < (Ps - 1) * 4 * kHistCols; i += kThreads) hist[i] = 0.0;
    for (int i = tid; i < (Ps - 1) * 8; i += kThreads) dsp[i] = tb.dspl[i];
    __syncthreads();

    // ------------------------------------------------------------ pairs
    const T hsc = (T)(sy.hinv / scale), rc = (T)(sy.rc * scale);
    const T pcut = (T)(sy.pot_cutoff * scale);
    const T pa = (T)(sy.pot_a * scale), pb = (T)sy.pot_b;
    const T lapsc = (T)((sy.D - 1) * scale);   // (D-1)/r = (D-1) scale / r_box
    double vpot = 0.0, outer = 0.0, lap_tot_R = 0.0, lap_tot_I = 0.0;
    const T hinvT = (T)sy.hinv;

    // Pairs in 64 x 64 tiles (A, B), A <= B, of particle blocks.  Lane l owns
    // row particle 64 A + l; column particles ride a register ring that
    // rotates one lane per step (wave_rol), so at step s lane l meets column
    // 64 B + ((l + s) mod 64) and after 64 steps every row has met every
    // column.  Diagonal tiles take s = 1..32 (s = 32 by lanes < 32 only), so
    // every unordered pair occurs once.  Row sums stay in registers, column
    // sums rotate with the column coordinates; both reach LDS once per tile
    // segment.  The steps of all tiles are split evenly over the 4 waves.
    const int nblk = (N + 63) / 64;
    const int total = nblk * (nblk - 1) / 2 * 64 + nblk * 32;
    const int g0 = total * wave / kWaves, g1 = total * (wave + 1) / kWaves;
    int off = 0;
    for (int A = 0; A < nblk; A++) {
        for (int Bk = A; Bk < nblk; Bk++) {
            const bool diag = A == Bk;
            const int cnt = diag ? 32 : 64;
            const int lo = max(g0, off), hi = min(g1, off + cnt);
            off += cnt;
            if (lo >= hi) continue;
            const int s_first = lo - (off - cnt) + (diag ? 1 : 0);
            const int nsteps = hi - lo;
            const int ra = A * 64 + lane;
            const bool rval = ra < N;
            T rn[D], cp[D];
            int cidx = Bk * 64 + ((lane + s_first) & 63);
#pragma unroll
            for (int a = 0; a < D; a++) {
                rn[a] = rval ? pos[a * N + ra] : (T)0;
                cp[a] = cidx < N ? pos[a * N + cidx] : (T)0;
            }
            T gR[D], gI[D], pR[D], pI[D];
#pragma unroll
            for (int a = 0; a < D; a++) gR[a] = gI[a] = pR[a] = pI[a] = (T)0;
            T lR = (T)0, lI = (T)0, plR = (T)0, plI = (T)0;
            for (int u = 0; u < nsteps; u++) {
                const int s = s_first + u;
                cidx = Bk * 64 + ((lane + s) & 63);
                const bool live = rval && cidx < N && (!diag || s < 32 || lane < 32);
                if (live) {
                    T d[D];
                    T r2 = (T)0;
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        d[a] = rn[a] - cp[a];
                        if (PER) d[a] -= Prec<T>::rint_(d[a]);
                        r2 += d[a] * d[a];
                    }
                    // FP32: one v_rsq gives 1/r, and r = r^2 / r; FP64: sqrt and a divide
                    T r, rinv;
                    Prec<T>::r_and_rinv(r2, &r, &rinv);
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
                        // moments 1, t, t^2, t^3 of the pair's position in knot
                        // interval j (FP64); the basis sums B_s are rebuilt from
                        // them at the end
                        const double td = (double)t, t2 = td * td, t3 = t2 * td;
                        double* hb = hist + j * 4 * kHistCols + (lane & (kHistCols - 1));
                        lds_add(hb, 1.0);
                        lds_add(hb + kHistCols, td);
                        lds_add(hb + 2 * kHistCols, t2);
                        lds_add(hb + 3 * kHistCols, t3);
                        // u'(r), u''(r) per unit length on interval j; the
                        // 24-bit multiply keeps the LDS offset a full-rate op
                        const T* c = reinterpret_cast<const T*>(reinterpret_cast<const unsigned char*>(dsp) +
                                                                __umul24((unsigned)j, 8 * sizeof(T)));
                        const T upR = c[0] + t * (c[1] + t * c[2]);
                        const T uppR = hinvT * c[1] + t * c[3];
                        const T upI = c[4] + t * (c[5] + t * c[6]);
                        const T uppI = hinvT * c[5] + t * c[7];
                        const T fR = upR * rinv, fI = upI * rinv;
#pragma unroll
                        for (int a = 0; a < D; a++) {
                            gR[a] += fR * d[a];
                            gI[a] += fI * d[a];
                            pR[a] -= fR * d[a];
                            pI[a] -= fI * d[a];
                        }
                        const T lfR = uppR + lapsc * fR;
                        const T lfI = uppI + lapsc * fI;
                        lR += lfR;
                        lI += lfI;
                        if (PP) {
                            plR += lfR;
                            plI += lfI;
                        }
                    } else {
                        outer += 1.0;
                    }
                }
#pragma unroll
                for (int a = 0; a < D; a++) {
                    cp[a] = wave_rotate(cp[a]);
                    pR[a] = wave_rotate(pR[a]);
                    pI[a] = wave_rotate(pI[a]);
                }
                if (PP) {
                    plR = wave_rotate(plR);
                    plI = wave_rotate(plI);
                }
            }
            // column sums: lane l now holds column 64 B + ((l + s_first + nsteps) mod 64)
            cidx = Bk * 64 + ((lane + s_first + nsteps) & 63);
            if (cidx < N) {
#pragma unroll
                for (int a = 0; a < D; a++) {
                    lds_add(gsum + a * N + cidx, (double)pR[a]);
                    lds_add(gsum + (D + a) * N + cidx, (double)pI[a]);
                }
                if (PP) {
                    lds_add(lsum + cidx, (double)plR);
                    lds_add(lsum + N + cidx, (double)plI);
                }
            }
            if (rval) {
#pragma unroll
                for (int a = 0; a < D; a++) {
                    lds_add(gsum + a * N + ra, (double)gR[a]);
                    lds_add(gsum + (D + a) * N + ra, (double)gI[a]);
                }
                if (PP) {
                    lds_add(lsum + ra, (double)lR);
                    lds_add(lsum + N + ra, (double)lI);
                } else {
                    // every pair term enters the Laplacians of both particles
                    lap_tot_R += 2.0 * (double)lR;
                    lap_tot_I += 2.0 * (double)lI;
                }
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
        // O = |rho_k|^2 for k and for -k (rho_{-k} = conj rho_k)
        const double ok = (double)re * (double)re + (double)im * (double)im;
        const int m0 = tb.kmap[2 * k], m1 = tb.kmap[2 * k + 1];
        double a = tb.uR[sy.off_k + m0];
        out.O_store(w, sy.P, sy.off_k + m0, ok);
        if (m1 >= 0) {
            a += tb.uR[sy.off_k + m1];
            out.O_store(w, sy.P, sy.off_k + m1, ok);
        }
        ex_k += a * ok;
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
#pragma unroll
                for (int a = 0; a < D; a++) {
                    lds_add(gsum + a * N + n, (double)gR[a]);
                    lds_add(gsum + (D + a) * N + n, (double)gI[a]);
                }
                if (PP) {
                    lds_add(lsum + n, (double)lR);
                    lds_add(lsum + N + n, (double)lI);
                } else {
                    lap_tot_R += (double)lR;
                    lap_tot_I += (double)lI;
                }
            }
        }
    }
    __syncthreads();

    // histogram columns -> column 0 (interval moments M[j][p])
    for (int e = tid; e < (Ps - 1) * 4; e += kThreads) {
        double m = 0.0;
        for (int c = 0; c < kHistCols; c++) m += hist[e * kHistCols + c];
        hist[e * kHistCols] = m;
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
            gr[a] = (double)gsum[a * N + n] - 2.0 * obR * x;
            gi[a] = (double)gsum[(D + a) * N + n] - 2.0 * obI * x;
            r2n += x * x;
        }
        sr2 += r2n;
        double lr = 0.0, li = 0.0;
        if (PP) {
            lr = (double)lsum[n] - 2.0 * D * obR;
            li = (double)lsum[N + n] - 2.0 * D * obI;
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

    // spline basis sums -> O_pair (BC map), exponent contribution.  Knot
    // interval j carries the pieces of B_j .. B_{j+3}; piece m is the cubic
    // sum_p beta[m][p] t^p, so B_s = sum_m sum_p beta[m][p] M[s-m][p] with the
    // interval moments M[j][p] = sum over pairs in j of t^p.
    double ex_p = 0.0;
    for (int k = tid; k < Ps; k += kThreads) {
        constexpr double beta[4][4] = {{1.0 / 6, -3.0 / 6, 3.0 / 6, -1.0 / 6},
                                       {4.0 / 6, 0.0, -6.0 / 6, 3.0 / 6},
                                       {1.0 / 6, 3.0 / 6, 3.0 / 6, -3.0 / 6},
                                       {0.0, 0.0, 0.0, 1.0 / 6}};
        auto B = [&](int s) {
            double v = 0.0;
            for (int m = 0; m < 4; m++) {
                const int j = s - m;
                if (j < 0 || j > Ps - 2) continue;
#pragma unroll
                for (int p = 0; p < 4; p++) v += beta[m][p] * hist[(j * 4 + p) * kHistCols];
            }
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

template <typename T, int D, bool PP, bool PER, typename TO>
__global__ __launch_bounds__(kThreads, 4) void eval_kernel(KSys sy, KTables<T> tb, const double* __restrict__ Rg,
                                                         int64_t W, EvalOut<TO> out, const T* __restrict__ rho_in)
{
    eval_body<T, D, PP, PER, TO>(sy, tb, Rg, W, out, rho_in, blockIdx.x);
}

// Batch of independent systems in one launch, one workgroup per walker.
template <typename T, int D, bool PER>
__global__ __launch_bounds__(kThreads, 4) void eval_batch_kernel(EvalBatch<T> bt)
{
    int lo = 0, hi = bt.n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bt.block_start[mid] <= (int64_t)blockIdx.x)
            lo = mid;
        else
            hi = mid;
    }
    const EvalBatchItem<T>& it = bt.item[lo];
    EvalOut<double> o{it.expo, it.ER, it.EI, it.O, it.other, nullptr, nullptr, it.outer, it.phiR};
    eval_body<T, D, false, PER, double>(bt.sys[lo], bt.tab[lo], it.R, it.W, o, it.rho,
                                        (int64_t)blockIdx.x - bt.block_start[lo]);
}

}  // namespace

template <typename T>
size_t eval_lds_bytes(const KSys& sy, bool per_particle)
{
    const size_t nd = (size_t)(sy.P_s - 1) * 4 * kHistCols + (size_t)sy.N * 2 * sy.D +
                      (per_particle ? 2 * (size_t)sy.N : 0);
    const size_t nt = (size_t)sy.N * sy.D + 2 * (size_t)sy.Kh;
    // the derivative table (8 entries per interval) starts 16-byte aligned
    const size_t head = (64 + nd * sizeof(double) + nt * sizeof(T) + 15) & ~(size_t)15;
    return head + (size_t)(sy.P_s - 1) * 8 * sizeof(T);
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

template <typename T, int D, bool PER>
static hipError_t launch_eval_batch_k(const EvalBatch<T>& bt, int64_t blocks, size_t lds, hipStream_t st)
{
    auto kern = eval_batch_kernel<T, D, PER>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kThreads), lds, st, bt);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_eval_batch(const EvalBatch<T>& bt, int D, bool periodic, int64_t blocks, size_t lds, hipStream_t st)
{
    if (blocks <= 0) return hipSuccess;
    switch (D) {
    case 1: return periodic ? launch_eval_batch_k<T, 1, true>(bt, blocks, lds, st)
                            : launch_eval_batch_k<T, 1, false>(bt, blocks, lds, st);
    case 2: return periodic ? launch_eval_batch_k<T, 2, true>(bt, blocks, lds, st)
                            : launch_eval_batch_k<T, 2, false>(bt, blocks, lds, st);
    case 3: return periodic ? launch_eval_batch_k<T, 3, true>(bt, blocks, lds, st)
                            : launch_eval_batch_k<T, 3, false>(bt, blocks, lds, st);
    default: return hipErrorInvalidValue;
    }
}
template hipError_t launch_eval_batch<float>(const EvalBatch<float>&, int, bool, int64_t, size_t, hipStream_t);
template hipError_t launch_eval_batch<double>(const EvalBatch<double>&, int, bool, int64_t, size_t, hipStream_t);

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
