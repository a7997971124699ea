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
// Work split (N particles, K k-vectors, 256 threads), in this order:
//   rho_k    : thread = k, loop over particles (LDS broadcast reads), or the
//              values the sweep handed over;
//   pairs    : every unordered pair once, on a half ring: lane = row n of a
//              64-row block, partner i = (n + s) mod N, the s range split over
//              the walker's waves; own-row sums stay in registers, partner sums ride
//              a DPP register ring and reach LDS once per particle and step;
//   k terms  : lane = row n, the k range split over the walker's waves;
//   E_L      : thread = particle, FP64 block reduction.
// The pair's position t in its knot interval goes to an LDS histogram of the
// moments 1, t, t^2, t^3 per interval ([interval][4][16 or 8 columns]);
// the B-spline basis sums are rebuilt from these moments at the end, and the
// count of pairs beyond the cutoff follows from the histogram's total.
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
// Histogram columns: 16 for four-wavefront walkers with real parameters (no
// bank conflicts among the 16 lanes of one LDS pass; the imaginary gradient
// sums would take that LDS otherwise), 4 for one-wavefront walkers (four of
// them share a workgroup, and the smaller slots let more workgroups share a
// CU), 8 otherwise (smaller walkers spend more on clearing and reducing the
// histogram than on its conflicts).
constexpr int kHistColsC = 8;
__host__ __device__ constexpr int hist_cols(bool cplx, int nw)
{
    return nw == 1 ? kHistColsC / 2 : ((cplx || nw < 4) ? kHistColsC : 2 * kHistColsC);
}

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
    double* bsum = nullptr;   // [W][P_s + 2] basis sums B_s before the BC map, or null
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
__device__ __forceinline__ uint32_t wave_rotate(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, kDppWaveRol1, 0xF, 0xF, false);
}
__device__ __forceinline__ double wave_rotate(double v)
{
    const int2 x = *reinterpret_cast<int2*>(&v);
    int2 y;
    y.x = __builtin_amdgcn_mov_dpp(x.x, kDppWaveRol1, 0xF, 0xF, false);
    y.y = __builtin_amdgcn_mov_dpp(x.y, kDppWaveRol1, 0xF, 0xF, false);
    return *reinterpret_cast<double*>(&y);
}

// Sums of NV values over the WPW wavefronts of one walker: wave reductions,
// then (WPW > 1) across the walker's waves through LDS.  With WPW = 1 no
// barrier is needed.
template <int WPW, int NV>
__device__ __forceinline__ void walker_sum(double (&v)[NV], double* red, int wave, int lane)
{
    // wave sums by DPP steps and lane reads (no LDS round trips)
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = group_sum<64>(v[i]);
    if constexpr (WPW > 1) {
        // (red is written only here, once per walker: no barrier before the writes)
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) red[wave * NV + i] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NV; i++) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < WPW; w++) s += red[w * NV + i];
            v[i] = s;
        }
    }
}
constexpr int kRedDoubles = 40;   // >= 4 waves x 10 values

// Pair potential of one pair in coordinate-difference units (constants from
// PairConsts); POT: 1 square well, 2 Gaussian, 3 Lennard-Jones.
template <typename T>
struct PairConsts {
    T a;    // square well: range
    T b;    // square well: depth; Gaussian: amplitude
    T g;    // Gaussian: exponent factor of r^2 (base 2 in FP32, base e in FP64)
    T a2;   // Lennard-Jones: sigma^2
    T c4;   // Lennard-Jones: 4 epsilon
};

template <typename T, int POT>
__device__ __forceinline__ T pair_pot(const PairConsts<T>& pc, T r, T r2, T rinv)
{
    if constexpr (POT == 1) {
        return r < pc.a ? pc.b : (T)0;
    } else if constexpr (POT == 2) {
        if constexpr (sizeof(T) == 4)
            return pc.b * Prec<T>::exp2_(r2 * pc.g);
        else
            return pc.b * Prec<T>::exp_(r2 * pc.g);
    } else if constexpr (POT == 3) {
        const T q = pc.a2 * (rinv * rinv);   // (sigma / r)^2
        const T q6 = q * q * q;
        return pc.c4 * (q6 * q6 - q6);
    } else {
        return (T)0;
    }
}

// All pairs of one walker (see the work split above).  Coordinates are
// Coord<T, PER> values; lengths below are in units of one coordinate
// difference (2^-32 box lengths for FP32 in a periodic box, box lengths for
// FP64 in a periodic box, real units in open space).  The potential of a
// periodic box acts within the cutoff only (the same test as the spline); in
// open space it acts on every pair.  Pairs beyond the cutoff are not counted
// here: their number is N (N - 1) / 2 minus the histogram count.  CPLX false:
// every imaginary parameter is zero, so the imaginary parts (exact zeros) are
// not computed.
template <typename T, int D, bool PP, bool PER, int POT, bool CPLX, int NW, bool NU>
__device__ __forceinline__ void eval_pairs(const KSys& sy, const typename Coord<T, PER>::S* __restrict__ pos,
                                           const T* __restrict__ dsp, double* __restrict__ hist,
                                           double* __restrict__ gsum, double* __restrict__ lsum, int wave,
                                           int lane, double& vpot, double& lap_tot_R, double& lap_tot_I)
{
    using C = Coord<T, PER>;
    using S = typename C::S;
    const int N = sy.N;
    const double unit = C::kUnit / (PER ? sy.Linv : 1.0);   // length of one coordinate-difference unit
    const T hsc = (T)(sy.hinv * unit), rc = (T)(sy.rc / unit);
    // u'' = hl (d2 + t e3): t in knot units (uniform knots), s in real length
    // with the derivatives taken in s (non-uniform knots)
    const double hl = NU ? 1.0 : sy.hinv;
    const T hinvT = (T)hl;
    // non-uniform knots: lookup cells after the derivative table (coordinate-
    // difference units; see knots.hpp)
    const T* ncel = dsp + (size_t)sy.P_s * kDsplStride;
    const T cin = (T)(sy.cinv * unit), unitR = (T)unit;
    const int cmax = sy.ncell - 1;
    // Cutoff test of the FP32 form in a periodic box.  O_{P-1} and O_{P-2}
    // jump by 2/3 and 1/3 where a pair crosses rc, so a pair whose FP32 r^2
    // (a few ulp off) lands on the other side of rc than the exact distance
    // shifts O by a whole basis value.  Outside a band of relative width
    // 2^-18 around rc^2 the FP32 r^2 decides; inside it, the exact integer
    // coordinate differences do (in FP64, rounding 2^-53).
    constexpr bool kExactCut = sizeof(T) == 4 && PER;
    const double rc2x = (sy.rc / unit) * (sy.rc / unit);
    const T rc2lo = (T)(rc2x * (1.0 - 0x1p-18)), rc2hi = (T)(rc2x * (1.0 + 0x1p-18));
    const double lapscD = (sy.D - 1) / unit;   // (D-1)/r = lapsc * rinv
    const T lapsc = (T)lapscD;
    PairConsts<T> pc{};
    if constexpr (POT == 1) {
        pc.a = (T)(sy.pot_a / unit);
        pc.b = (T)sy.pot_b;
    } else if constexpr (POT == 2) {
        const double au = sy.pot_a / unit;
        pc.b = (T)sy.pot_b;
        pc.g = (T)((sizeof(T) == 4 ? -0.5 * 1.4426950408889634 : -0.5) / (au * au));
    } else if constexpr (POT == 3) {
        const double au = sy.pot_a / unit;
        pc.a2 = (T)(au * au);
        pc.c4 = (T)(4.0 * sy.pot_b);
    }
    constexpr int HC = hist_cols(CPLX, NW);
    double* const hbase = hist + (lane & (HC - 1));

    // Pairs in 64 x 64 tiles (A, B), A <= B, of particle blocks.  Lane l owns
    // row particle 64 A + l; column particles ride a register ring that
    // rotates one lane per step (wave_rol), so at step s lane l meets ring
    // slot (l + s) mod 64 and after 64 steps every row has met every column.
    // Diagonal tiles take s = 1..32 (s = 32 by lanes < 32 only), so every
    // unordered pair occurs once.  A last block of at most 32 particles is
    // held twice in the ring (slot q: particle q mod 32) and met in 32 steps
    // s = 0..31: 32 consecutive slots hold every residue mod 32 once.  A last
    // block of r <= 32 particles meets itself in s = 1 .. r - 1 only (lane l
    // meets slot l + s, which is a particle of the block only for l + s < r
    // when s <= 32).  Row sums stay in registers, column sums rotate with the
    // column coordinates; both reach LDS once per tile segment.  The steps of
    // all tiles are split evenly over the walker's NW waves.  Off-diagonal
    // tiles of full blocks run without per-step validity tests.
    const int nfull = N / 64, rem = N - 64 * nfull;
    const int nblk = nfull + (rem > 0 ? 1 : 0);
    const bool fold = rem > 0 && rem <= 32 && nfull > 0;
    const int dlast = (rem > 0 && rem <= 32) ? rem - 1 : 32;   // steps of the last block's diagonal tile
    const int total = nfull * (nfull - 1) / 2 * 64 + nfull * 32 + (rem > 0 ? dlast + nfull * (fold ? 32 : 64) : 0);
    const int g0 = total * wave / NW, g1 = total * (wave + 1) / NW;
    int off = 0;
    for (int A = 0; A < nblk; A++) {
        for (int Bk = A; Bk < nblk; Bk++) {
            const bool diag = A == Bk;
            const bool folded = !diag && fold && Bk == nblk - 1;
            const int cnt = diag ? (rem > 0 && A == nblk - 1 ? dlast : 32) : (folded ? 32 : 64);
            const int lo = max(g0, off), hi = min(g1, off + cnt);
            off += cnt;
            if (lo >= hi) continue;
            const int s_first = lo - (off - cnt) + (diag ? 1 : 0);
            const int nsteps = hi - lo;
            const int cmask = folded ? 31 : 63;   // ring slot -> column offset in block Bk
            const int ra = A * 64 + lane;
            const bool rval = ra < N;
            S rn[D], cp[D];
            int cidx = Bk * 64 + ((lane + s_first) & cmask);
#pragma unroll
            for (int a = 0; a < D; a++) {
                rn[a] = rval ? pos[a * N + ra] : (S)0;
                cp[a] = cidx < N ? pos[a * N + cidx] : (S)0;
            }
            T gR[D], gI[D], pR[D], pI[D];
#pragma unroll
            for (int a = 0; a < D; a++) gR[a] = gI[a] = pR[a] = pI[a] = (T)0;
            // Laplacian pieces: mR / mI = sum u'/r (without per-particle output;
            // sum u'' is linear in the interval moments and follows from the
            // histogram at the end), or the full per-pair term lR / lI (with it)
            T lR = (T)0, lI = (T)0, mR = (T)0, mI = (T)0, plR = (T)0, plI = (T)0, vs = (T)0;

            auto pair = [&]() {
                T d[D];
                T r2 = (T)0;
#pragma unroll
                for (int a = 0; a < D; a++) {
                    d[a] = C::diff(rn[a], cp[a]);
                    r2 += d[a] * d[a];
                }
                // FP32: one v_rsq gives 1/r, and r = r^2 / r; FP64: sqrt and a divide
                T r, rinv;
                Prec<T>::r_and_rinv(r2, &r, &rinv);
                if constexpr (!PER && POT != 0) vs += pair_pot<T, POT>(pc, r, r2, rinv);
                bool inside;
                if constexpr (kExactCut) {
                    // below the band's top: inside unless in the band, and
                    // when no lane of the wave is in the band that is all
                    const bool below = r2 <= rc2hi;
                    const bool band = below && r2 > rc2lo;
                    inside = below;
                    if (__builtin_amdgcn_ballot_w64(band)) {   // rare: some lane near rc
                        // keeps the exact test behind the branch (the compiler
                        // would otherwise evaluate it for every pair)
                        asm volatile("" ::: "memory");
                        double e2 = 0.0;
#pragma unroll
                        for (int a = 0; a < D; a++) {
                            const double de = C::diff_exact(rn[a], cp[a]);
                            e2 = fma(de, de, e2);
                        }
                        inside = below && (!band || e2 <= rc2x);
                    }
                } else {
                    inside = r <= rc;
                }
                if (inside) {
                    if constexpr (PER && POT != 0) vs += pair_pot<T, POT>(pc, r, r2, rinv);
                    // knot interval j in [0, P_s-1] (P_s-1 only at r = rc, t = 0)
                    int j;
                    T t;
                    if constexpr (NU) {
                        int c = (int)(r * cin);
                        c = c < cmax ? c : cmax;
                        const T* e = ncel + 4 * c;
                        const bool right = r > e[0];
                        j = (int)e[1] + (right ? 1 : 0);
                        t = (r - (right ? e[3] : e[2])) * unitR;
                    } else {
                        const T x = r * hsc;
                        j = (int)x;
                        t = Prec<T>::fract_(x);
                    }
                    // moments 1, t, t^2, t^3 of the pair's position in knot
                    // interval j (FP64); the basis sums B_s are rebuilt from
                    // them at the end
                    const double td = (double)t, t2 = td * td, t3 = t2 * td;
                    double* hb = hbase + j * (4 * HC);
                    lds_add(hb, 1.0);
                    lds_add(hb + HC, td);
                    lds_add(hb + 2 * HC, t2);
                    lds_add(hb + 3 * HC, t3);
                    // u'(r), u''(r) per unit length on interval j
                    const T* c = dsp + j * kDsplStride;
                    T c0, c1, c2, c3;
                    if constexpr (sizeof(T) == 4) {
                        // one 16-byte LDS read (rows are 16-byte aligned)
                        const float4 cv = *reinterpret_cast<const float4*>(c);
                        c0 = cv.x;
                        c1 = cv.y;
                        c2 = cv.z;
                        c3 = cv.w;
                    } else {
                        c0 = c[0];
                        c1 = c[1];
                        c2 = c[2];
                        c3 = c[3];
                    }
                    const T upR = c0 + t * (c1 + t * c2);
                    const T fR = upR * rinv;
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        gR[a] += fR * d[a];
                        pR[a] -= fR * d[a];
                    }
                    if constexpr (PP) {
                        const T lfR = hinvT * (c1 + t * c3) + lapsc * fR;
                        lR += lfR;
                        plR += lfR;
                    } else {
                        mR += fR;
                        // u'' is not needed per pair here (it follows from the
                        // histogram), but with three entries used the compiler
                        // narrows the read to 12 bytes, which measured slower
                        // in this loop: the empty operand keeps it 16 bytes
                        if constexpr (sizeof(T) == 4) asm volatile("" ::"v"(c3));
                    }
                    if constexpr (CPLX) {
                        const T upI = c[4] + t * (c[5] + t * c[6]);
                        const T fI = upI * rinv;
#pragma unroll
                        for (int a = 0; a < D; a++) {
                            gI[a] += fI * d[a];
                            pI[a] -= fI * d[a];
                        }
                        if constexpr (PP) {
                            const T lfI = hinvT * (c[5] + t * c[7]) + lapsc * fI;
                            lI += lfI;
                            plI += lfI;
                        } else {
                            mI += fI;
                        }
                    }
                }
            };
            auto rotate = [&]() {
#pragma unroll
                for (int a = 0; a < D; a++) {
                    cp[a] = wave_rotate(cp[a]);
                    pR[a] = wave_rotate(pR[a]);
                    if (CPLX) pI[a] = wave_rotate(pI[a]);
                }
                if (PP) {
                    plR = wave_rotate(plR);
                    if (CPLX) plI = wave_rotate(plI);
                }
            };
            if (!diag && !folded && (Bk + 1) * 64 <= N) {
                // both blocks full, no pair repeats: every lane is live at every step
                for (int u = 0; u < nsteps; u++) {
                    pair();
                    rotate();
                }
            } else if (diag && (A + 1) * 64 <= N) {
                // full diagonal block: every lane is live at the steps s < 32,
                // lanes < 32 only at s = 32 (s runs from s_first >= 1 up to 32)
                const int n1 = min(nsteps, max(0, 32 - s_first));
                for (int u = 0; u < n1; u++) {
                    pair();
                    rotate();
                }
                for (int u = n1; u < nsteps; u++) {
                    if (lane < 32) pair();
                    rotate();
                }
            } else {
                for (int u = 0; u < nsteps; u++) {
                    const int s = s_first + u;
                    const int ci = Bk * 64 + ((lane + s) & cmask);
                    if (rval && ci < N && (!diag || s < 32 || lane < 32)) pair();
                    rotate();
                }
            }
            vpot += (double)vs;
            // column sums: lane l now holds ring slot (l + s_first + nsteps) mod 64
            cidx = Bk * 64 + ((lane + s_first + nsteps) & cmask);
            if (cidx < N) {
#pragma unroll
                for (int a = 0; a < D; a++) {
                    lds_add(gsum + a * N + cidx, (double)pR[a]);
                    if (CPLX) lds_add(gsum + (D + a) * N + cidx, (double)pI[a]);
                }
                if (PP) {
                    lds_add(lsum + cidx, (double)plR);
                    if (CPLX) lds_add(lsum + N + cidx, (double)plI);
                }
            }
            if (rval) {
#pragma unroll
                for (int a = 0; a < D; a++) {
                    lds_add(gsum + a * N + ra, (double)gR[a]);
                    if (CPLX) lds_add(gsum + (D + a) * N + ra, (double)gI[a]);
                }
                if (PP) {
                    lds_add(lsum + ra, (double)lR);
                    if (CPLX) lds_add(lsum + N + ra, (double)lI);
                } else {
                    // every pair term enters the Laplacians of both particles
                    // (the u'' part is added from the histogram)
                    lap_tot_R += 2.0 * lapscD * (double)mR;
                    if (CPLX) lap_tot_I += 2.0 * lapscD * (double)mI;
                }
            }
        }
    }
}

// k-space gradient and Laplacian of sum_k alpha_k |rho_k|^2 for every particle
// (lane = particle, the k range split over the walker's NW waves).
//   grad_n |rho_k|^2 = -2 k Im(e^{ikr} conj rho),  k = 2 pi kv,
//   lap_n  |rho_k|^2 = 2 |k|^2 (1 - Re(e^{ikr} conj rho)),
// with the amplitudes folded into kag (alpha kv, alpha |k|^2): per (particle,
// k) one phase, one sincos, the two products with conj rho and four FMAs; the
// constant 2 sum_k alpha |k|^2 of the Laplacian is added once.
template <typename T, int D, bool PP, bool PER, bool CPLX, int NW>
__device__ __forceinline__ void eval_kgrad(const KSys& sy, const KTables<T>& tb, const typename Coord<T, PER>::S* pos,
                                           const T* rho, const T* kq, double* gsum, double* lsum, int wave, int lane,
                                           double& lap_tot_R, double& lap_tot_I)
{
    using C = Coord<T, PER>;
    const int N = sy.N, K = sy.Kh;
    const int kchunk = (K + NW - 1) / NW;
    const int k_lo = wave * kchunk, k_hi = min(K, (wave + 1) * kchunk);
    // the amplitude table is written before the launch and only read here:
    // read through the constant address space, its wave-uniform entries
    // become scalar loads (no vector memory instruction per k)
    const __attribute__((address_space(4))) T* kag = (const __attribute__((address_space(4))) T*)tb.kag;
    // sum over this wave's k of alpha |k|^2 (the constant part of the Laplacian)
    T cR = (T)0, cI = (T)0;
    for (int k = k_lo; k < k_hi; k++) {
        cR += kag[8 * k + 3];
        if (CPLX) cI += kag[8 * k + 7];
    }
    // a last round of at most 32 particles runs below, in half the k steps
    const int nround = (N % 64 != 0 && N % 64 <= 32) ? N - N % 64 : N;
    for (int rb = 0; rb < nround; rb += 64) {
        const int n = rb + lane;
        if (n >= N) continue;
        T rn[D];
#pragma unroll
        for (int a = 0; a < D; a++) rn[a] = C::raw(pos[a * N + n]);
        T gR[D], gI[D];
#pragma unroll
        for (int a = 0; a < D; a++) gR[a] = gI[a] = (T)0;
        T lR = (T)0, lI = (T)0;   // - sum alpha |k|^2 Re(e^{ikr} conj rho)
        // four k per pass: the LDS reads take immediate offsets, and the
        // scalar loads of four amplitudes are in flight together
#pragma unroll 4
        for (int k = k_lo; k < k_hi; k++) {
            const __attribute__((address_space(4))) T* ka = kag + 8 * k;
            const T* qs = kq + k * 4;
            T ph = (T)0;
#pragma unroll
            for (int a = 0; a < D; a++) ph += qs[a] * rn[a];
            T sn, cs;
            Prec<T>::sincos_small(ph, &sn, &cs);
            const T rr = rho[2 * k], ri = rho[2 * k + 1];
            const T pre = cs * rr + sn * ri;    // Re(e^{ikr} conj(rho))
            const T pim = sn * rr - cs * ri;    // Im(e^{ikr} conj(rho))
#pragma unroll
            for (int a = 0; a < D; a++) gR[a] += pim * ka[a];
            lR -= pre * ka[3];
            if constexpr (CPLX) {
#pragma unroll
                for (int a = 0; a < D; a++) gI[a] += pim * ka[4 + a];
                lI -= pre * ka[7];
            }
        }
#pragma unroll
        for (int a = 0; a < D; a++) {
            lds_add(gsum + a * N + n, (double)(gR[a] * (T)(-2.0 * 6.283185307179586)));
            if (CPLX) lds_add(gsum + (D + a) * N + n, (double)(gI[a] * (T)(-2.0 * 6.283185307179586)));
        }
        const double lapR = 2.0 * ((double)cR + (double)lR), lapI = 2.0 * ((double)cI + (double)lI);
        if (PP) {
            lds_add(lsum + n, lapR);
            if (CPLX) lds_add(lsum + N + n, lapI);
        } else {
            lap_tot_R += lapR;
            if (CPLX) lap_tot_I += lapI;
        }
    }
    if (nround < N) {
        // last round of rem <= 32 particles: lanes l and l + 32 hold particle
        // nround + l, the lower half-wave takes the first h of this wave's k,
        // the upper half the rest.  Both amplitude rows are scalar loads (k
        // stays wave-uniform per half); each lane selects its own.
        const int rem = N - nround, hl = lane & 31;
        const bool upper = lane >= 32;
        const bool act = hl < rem;
        const int n = nround + (act ? hl : 0);
        const int h = (k_hi - k_lo + 1) / 2;
        T rn[D];
#pragma unroll
        for (int a = 0; a < D; a++) rn[a] = C::raw(pos[a * N + n]);
        T gR[D], gI[D];
#pragma unroll
        for (int a = 0; a < D; a++) gR[a] = gI[a] = (T)0;
        T lR = (T)0, lI = (T)0;
#pragma unroll 2
        for (int j = 0; j < h; j++) {
            const int kA = k_lo + j, kB = k_lo + h + j;   // wave-uniform
            const bool hasB = kB < k_hi;
            const __attribute__((address_space(4))) T* kaA = kag + 8 * kA;
            const __attribute__((address_space(4))) T* kaB = kag + 8 * (hasB ? kB : kA);
            const int k = upper ? kB : kA;
            if (act && (!upper || hasB)) {
                const T* qs = kq + k * 4;
                T ph = (T)0;
#pragma unroll
                for (int a = 0; a < D; a++) ph += qs[a] * rn[a];
                T sn, cs;
                Prec<T>::sincos_small(ph, &sn, &cs);
                const T rr = rho[2 * k], ri = rho[2 * k + 1];
                const T pre = cs * rr + sn * ri;
                const T pim = sn * rr - cs * ri;
#pragma unroll
                for (int a = 0; a < D; a++) gR[a] += pim * (upper ? kaB[a] : kaA[a]);
                lR -= pre * (upper ? kaB[3] : kaA[3]);
                if constexpr (CPLX) {
#pragma unroll
                    for (int a = 0; a < D; a++) gI[a] += pim * (upper ? kaB[4 + a] : kaA[4 + a]);
                    lI -= pre * (upper ? kaB[7] : kaA[7]);
                }
            }
        }
        if (act) {
#pragma unroll
            for (int a = 0; a < D; a++) {
                lds_add(gsum + a * N + n, (double)(gR[a] * (T)(-2.0 * 6.283185307179586)));
                if (CPLX) lds_add(gsum + (D + a) * N + n, (double)(gI[a] * (T)(-2.0 * 6.283185307179586)));
            }
            // the constant part once per particle (lower half)
            const double lapR = 2.0 * ((upper ? 0.0 : (double)cR) + (double)lR);
            const double lapI = 2.0 * ((upper ? 0.0 : (double)cI) + (double)lI);
            if (PP) {
                lds_add(lsum + n, lapR);
                if (CPLX) lds_add(lsum + N + n, lapI);
            } else {
                lap_tot_R += lapR;
                if (CPLX) lap_tot_I += lapI;
            }
        }
    }
}

// Pairs and the k-space gradient for one value of CPLX, the potential type
// resolved to a template argument.
template <typename T, int D, bool PP, bool PER, bool CPLX, int NW>
__device__ __forceinline__ void eval_forces(const KSys& sy, const KTables<T>& tb,
                                            const typename Coord<T, PER>::S* pos, const T* dsp, double* hist,
                                            double* gsum, double* lsum, int wave, int lane, double& vpot,
                                            double& lap_tot_R, double& lap_tot_I)
{
    if constexpr (PER) {
        if (sy.nu) {
            // non-uniform knots (periodic box only)
            switch (sy.pot_type) {
            case 1: eval_pairs<T, D, PP, PER, 1, CPLX, NW, true>(sy, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R, lap_tot_I); break;
            case 2: eval_pairs<T, D, PP, PER, 2, CPLX, NW, true>(sy, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R, lap_tot_I); break;
            case 3: eval_pairs<T, D, PP, PER, 3, CPLX, NW, true>(sy, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R, lap_tot_I); break;
            default: eval_pairs<T, D, PP, PER, 0, CPLX, NW, true>(sy, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R, lap_tot_I); break;
            }
            return;
        }
    }
    switch (sy.pot_type) {
    case 1: eval_pairs<T, D, PP, PER, 1, CPLX, NW, false>(sy, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R, lap_tot_I); break;
    case 2: eval_pairs<T, D, PP, PER, 2, CPLX, NW, false>(sy, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R, lap_tot_I); break;
    case 3: eval_pairs<T, D, PP, PER, 3, CPLX, NW, false>(sy, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R, lap_tot_I); break;
    default: eval_pairs<T, D, PP, PER, 0, CPLX, NW, false>(sy, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R, lap_tot_I); break;
    }
}

// WPW wavefronts evaluate walker w of the system described by (sy, tb): the
// workgroup's walker slot `slot` (threads slot * 64 WPW ..), with its own
// LDS region of slot_bytes.  A slot past the last walker repeats the last
// walker's work and stores nothing, so every barrier is met by the whole
// workgroup.
template <typename T, int D, bool PP, bool PER, typename TO, int WPW>
__device__ __forceinline__ void eval_body(const KSys& sy, const KTables<T>& tb, const double* __restrict__ Rg,
                                          int64_t W, const EvalOut<TO>& out, const T* __restrict__ rho_in,
                                          int64_t w_slot, int slot, size_t slot_bytes)
{
    using C = Coord<T, PER>;
    using S = typename C::S;
    constexpr int NT = 64 * WPW;   // threads per walker
    const bool active = w_slot < W;
    const int64_t w = active ? w_slot : W - 1;
    const int N = sy.N, Ps = sy.P_s, K = sy.Kh;   // K: +-k representatives
    const int tid = threadIdx.x - slot * NT, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (scalar registers)

    // loaded ahead, so that their latency overlaps the rest: the raw
    // coordinates of this thread's first particle (staged below and used
    // again by the one-body factor), its first pair parameter of the O
    // assembly, and the parameters of the exponent
    const double* Rw = Rg + w * (size_t)N * D;
    double xr0[D];
#pragma unroll
    for (int a = 0; a < D; a++) xr0[a] = tid < N ? Rw[(size_t)tid * D + a] : 0.0;
    const double ur0 = tid < Ps ? tb.uR[sy.off_pair + tid] : 0.0;
    // (wave-uniform: through the constant address space they become scalar
    // loads into scalar registers; the tables are written before the launch)
    typedef const __attribute__((address_space(4))) double* CDP;
    const double u_outer = ((CDP)tb.uR)[sy.outer_idx], u_phi = ((CDP)tb.uR)[sy.P];
    const double obR = sy.one_body ? ((CDP)tb.uR)[0] : 0.0, obI = sy.one_body ? ((CDP)tb.uI)[0] : 0.0;

    // any nonzero imaginary parameter?  If none, the imaginary parts of the
    // pair and k-space sums are exact zeros and are skipped (and the LDS
    // layout trades the imaginary gradient sums for histogram columns)
    int nz = 0;
    if (!PP)
        for (int i = tid; i < sy.P; i += NT) nz |= tb.uI[i] != 0.0;
    const bool cplx = PP || __syncthreads_or(nz) != 0;
    const int hc = hist_cols(cplx, WPW);

    extern __shared__ __align__(16) unsigned char smem_all[];
    unsigned char* smem = smem_all + (size_t)slot * slot_bytes;
    // LDS carve-up (sizes mirrored by eval_lds_bytes); per-particle arrays are
    // structure-of-arrays so consecutive particles sit in consecutive banks
    double* red = reinterpret_cast<double*>(smem);             // kRedDoubles
    double* hist = red + kRedDoubles;                          // P_s * 4 * hc
    double* gsum = hist + (size_t)Ps * 4 * hc;                 // [D][N] Re g, then (cplx) [D][N] Im g
    double* lsum = gsum + (size_t)N * (cplx ? 2 : 1) * D;      // [2][N] (PP only)
    S* pos = reinterpret_cast<S*>(lsum + (PP ? (size_t)N * 2 : 0));   // [D][N]
    T* rho = reinterpret_cast<T*>(pos + (size_t)N * D);        // 2K
    // k-vectors in revolutions per coordinate unit ([K][4]), then the
    // derivative table (P_s * kDsplStride), both 16-byte aligned
    T* kq = reinterpret_cast<T*>(smem + ((reinterpret_cast<unsigned char*>(rho + 2 * (size_t)K) - smem + 15) &
                                         ~(size_t)15));
    T* dsp = kq + 4 * (size_t)K;
    const T* kv = tb.kv;                                       // K * 4 (global, read with uniform k)

    // periodic box: coordinates in box units wrapped into the primary cell
    // (every quantity is invariant under it), stored as Coord values
    const double scale = PER ? sy.Linv : 1.0;
    for (int i = tid; i < N; i += NT) {
#pragma unroll
        for (int a = 0; a < D; a++) pos[a * N + i] = C::load((i == tid ? xr0[a] : Rw[(size_t)i * D + a]) * scale);
        for (int c = 0; c < (cplx ? 2 * D : D); c++) gsum[c * N + i] = 0.0;
        if (PP) {
            lsum[i] = 0.0;
            lsum[N + i] = 0.0;
        }
    }
    // hc is even and hist 16-byte aligned: cleared two doubles per store
    for (int i = tid; i < Ps * 2 * hc; i += NT) reinterpret_cast<double2*>(hist)[i] = make_double2(0.0, 0.0);
    for (int i = tid; i < Ps * kDsplStride; i += NT) dsp[i] = tb.dspl[i];
    if (PER && sy.nu) {
        // lookup cells in coordinate-difference units (after the derivative table)
        const double ul = C::kUnit / scale;
        T* ncel = dsp + (size_t)Ps * kDsplStride;
        for (int i = tid; i < sy.ncell; i += NT) {
            const double* e = tb.nucell + 4 * i;
            ncel[4 * i] = (T)(e[0] / ul);
            ncel[4 * i + 1] = (T)e[1];
            ncel[4 * i + 2] = (T)(e[2] / ul);
            ncel[4 * i + 3] = (T)(e[3] / ul);
        }
    }
    for (int k = tid; k < K; k += NT) {
        // periodic: kv L is the integer lattice vector n (rounded to remove
        // the FP32 storage error of n / L)
#pragma unroll
        for (int a = 0; a < D; a++) {
            const double n = PER ? rint((double)kv[k * 4 + a] / scale) : (double)kv[k * 4 + a];
            kq[k * 4 + a] = (T)(n * C::kUnit);
        }
        for (int a = D; a < 4; a++) kq[k * 4 + a] = (T)0;
    }
    __syncthreads();

    // ------------------------------------------------------------ rho_k
    // phase in revolutions = kq . raw coordinate, |phase| <= sum_a |n_a| / 2
    double ex_k = 0.0;
    for (int k = tid; k < K; k += NT) {
        T re = (T)0, im = (T)0;
        if (rho_in) {
            re = rho_in[((size_t)w * K + k) * 2];
            im = rho_in[((size_t)w * K + k) * 2 + 1];
        } else {
            T q[D];
#pragma unroll
            for (int a = 0; a < D; a++) q[a] = kq[k * 4 + a];
            for (int n = 0; n < N; n++) {
                T ph = (T)0;
#pragma unroll
                for (int a = 0; a < D; a++) ph += q[a] * C::raw(pos[a * N + n]);
                T sn, cs;
                Prec<T>::sincos_small(ph, &sn, &cs);
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
        if (active) out.O_store(w, sy.P, sy.off_k + m0, ok);
        if (m1 >= 0) {
            a += tb.uR[sy.off_k + m1];
            if (active) out.O_store(w, sy.P, sy.off_k + m1, ok);
        }
        ex_k += a * ok;
    }
    __syncthreads();

    // ------------------------------------------------------------ pairs, k-space gradient / Laplacian
    double vpot = 0.0, lap_tot_R = 0.0, lap_tot_I = 0.0;
    if (PP || cplx) {
        eval_forces<T, D, PP, PER, true, WPW>(sy, tb, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R,
                                              lap_tot_I);
        if (K > 0)
            eval_kgrad<T, D, PP, PER, true, WPW>(sy, tb, pos, rho, kq, gsum, lsum, wave, lane, lap_tot_R, lap_tot_I);
    } else if constexpr (!PP) {
        eval_forces<T, D, PP, PER, false, WPW>(sy, tb, pos, dsp, hist, gsum, lsum, wave, lane, vpot, lap_tot_R,
                                               lap_tot_I);
        if (K > 0)
            eval_kgrad<T, D, PP, PER, false, WPW>(sy, tb, pos, rho, kq, gsum, lsum, wave, lane, lap_tot_R, lap_tot_I);
    }
    __syncthreads();

    // histogram columns -> column 0 (interval moments M[j][p])
    for (int e = tid; e < Ps * 4; e += NT) {
        const double2* h2 = reinterpret_cast<const double2*>(hist + e * hc);
        double m = 0.0;
        for (int c = 0; c < hc / 2; c++) {
            const double2 v = h2[c];
            m += v.x;
            m += v.y;
        }
        hist[e * hc] = m;
    }
    __syncthreads();
    // pairs in the histogram (exact integer counts: any order of the sum gives
    // the same bits); the pairs beyond the cutoff follow after the walker sum
    double inner = 0.0;
    for (int j = tid; j < Ps; j += NT) inner += hist[j * 4 * hc];

    // ------------------------------------------------------------ assembly
    double sR1 = 0, sI1 = 0, sR1I1 = 0, sR2 = lap_tot_R, sI2 = lap_tot_I, sr2 = 0, vext = 0;
    for (int n = tid; n < N; n += NT) {
        double gr[D], gi[D];
        double r2n = 0.0;
#pragma unroll
        for (int a = 0; a < D; a++) {
            // one-body factor uses the raw (unwrapped, open-space) coordinate
            const double x = n == tid ? xr0[a] : Rw[(size_t)n * D + a];
            gr[a] = (double)gsum[a * N + n] - 2.0 * obR * x;
            gi[a] = (cplx ? (double)gsum[(D + a) * N + n] : 0.0) - 2.0 * obI * x;
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
        if (PP && active) {
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
    constexpr double beta[4][4] = {{1.0 / 6, -3.0 / 6, 3.0 / 6, -1.0 / 6},
                                   {4.0 / 6, 0.0, -6.0 / 6, 3.0 / 6},
                                   {1.0 / 6, 3.0 / 6, 3.0 / 6, -3.0 / 6},
                                   {0.0, 0.0, 0.0, 1.0 / 6}};
    auto B = [&](int s) {
            double v = 0.0;
            for (int m = 0; m < 4; m++) {
                const int j = s - m;
                if (j < 0 || j > Ps - 1) continue;
                if (sy.nu) {
                    // pieces of interval j in s (non-uniform knots)
                    const double* bj = tb.nubeta + (j * 4 + m) * 4;
#pragma unroll
                    for (int p = 0; p < 4; p++) v += bj[p] * hist[(j * 4 + p) * hc];
                } else {
#pragma unroll
                    for (int p = 0; p < 4; p++) v += beta[m][p] * hist[(j * 4 + p) * hc];
                }
            }
            return v;
    };
    // sum over pairs of u''(r) = hl (d2_j + t e3_j) (uniform knots: hl = 1/h,
    // t in knot units; non-uniform: hl = 1, t = s) is sum_j hl (d2_j M[j][0] +
    // e3_j M[j][1]); each pair counts for both of its particles
    const double hl = (PER && sy.nu) ? 1.0 : sy.hinv;
    for (int j = tid; !PP && j < Ps; j += NT) {
        const double m0 = hist[j * 4 * hc], m1 = hist[(j * 4 + 1) * hc];
        const T* c = dsp + j * kDsplStride;
        sR2 += 2.0 * hl * ((double)c[1] * m0 + (double)c[3] * m1);
        if (cplx) sI2 += 2.0 * hl * ((double)c[5] * m0 + (double)c[7] * m1);
    }
    // each B_s once, into column 1 of histogram row s (the columns past 0
    // are free after the reduction above; B reads column 0 only)
    for (int s = tid; s < Ps + 2; s += NT) hist[s * hc + 1] = B(s);
    __syncthreads();
    auto Bs = [&](int s) { return hist[s * hc + 1]; };
    // boundary-condition map O = T B: origin rows over B_0..B_2, then
    // B_3, B_4, ..., then cutoff rows over B_{P_s-1}..B_{P_s+1}
    for (int k = tid; k < Ps; k += NT) {
        double o;
        if (k < sy.bc_n0) {
            o = sy.bc_f0[k][0] * Bs(0) + sy.bc_f0[k][1] * Bs(1) + sy.bc_f0[k][2] * Bs(2);
        } else if (k < Ps - sy.bc_n1) {
            o = Bs(3 + k - sy.bc_n0);
        } else {
            const double* f = sy.bc_f1[k - (Ps - sy.bc_n1)];
            o = f[0] * Bs(Ps - 1) + f[1] * Bs(Ps) + f[2] * Bs(Ps + 1);
        }
        ex_p += (k == tid ? ur0 : tb.uR[sy.off_pair + k]) * o;
        if (active) out.O_store(w, sy.P, sy.off_pair + k, o);
    }
    if (out.bsum && active)
        for (int s = tid; s < Ps + 2; s += NT) out.bsum[w * (Ps + 2) + s] = Bs(s);

    double sums[10] = {sR1, sI1, sR1I1, sR2, sI2, sr2, vext, vpot, ex_k + ex_p, inner};
    walker_sum<WPW>(sums, red, wave, lane);
    sR1 = sums[0];
    sI1 = sums[1];
    sR1I1 = sums[2];
    sR2 = sums[3];
    sI2 = sums[4];
    sr2 = sums[5];
    vext = sums[6];
    vpot = sums[7];
    double ex = sums[8];
    // pairs beyond the cutoff: all pairs minus those in the histogram
    const double outer = 0.5 * (double)N * (double)(N - 1) - sums[9];
    if (tid == 0 && active) {
        if (sy.one_body) {
            out.O_store(w, sy.P, 0, -sr2);
            ex += obR * (-sr2);
        }
        ex += u_outer * outer;
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
            o[2] = exp(ex + u_phi);   // phi R of the parameters in force (device copy)
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
    eval_body<T, D, PP, PER, TO, 4>(sy, tb, Rg, W, out, rho_in, blockIdx.x, 0, 0);
}

// Batch of independent systems in one launch.  Each system has its own
// walker shape: 1, 2 or 4 wavefronts per walker (4, 2 or 1 walkers per
// workgroup); a workgroup belongs to one system.  Small walkers are bound by
// latency more than by issue, so the FP32 form is compiled for 6 waves per
// SIMD (80 registers) instead of 4; the FP64 form keeps 4 (it would spill).
template <typename T, int D, bool PER>
__global__ __launch_bounds__(kThreads, sizeof(T) == 4 ? 6 : 4) void eval_batch_kernel(EvalBatch<T> bt)
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
    EvalOut<double> o{it.expo, it.ER, it.EI, it.O, it.other, nullptr, nullptr, it.outer};
    const int64_t b = (int64_t)blockIdx.x - bt.block_start[lo];
    const size_t sb = (size_t)it.slot_bytes;
    switch (it.wpw) {
    case 1: {
        const int slot = threadIdx.x >> 6;
        eval_body<T, D, false, PER, double, 1>(bt.sys[lo], bt.tab[lo], it.R, it.W, o, it.rho, b * 4 + slot, slot, sb);
        break;
    }
    case 2: {
        const int slot = threadIdx.x >> 7;
        eval_body<T, D, false, PER, double, 2>(bt.sys[lo], bt.tab[lo], it.R, it.W, o, it.rho, b * 2 + slot, slot, sb);
        break;
    }
    default:
        eval_body<T, D, false, PER, double, 4>(bt.sys[lo], bt.tab[lo], it.R, it.W, o, it.rho, b, 0, sb);
        break;
    }
}

}  // namespace

// One wavefront per walker up to N = 32 (a single 64-row pair tile), two up
// to 64, four above.
int eval_default_wpw(int N) { return N <= 32 ? 1 : (N <= 64 ? 2 : 4); }

template <typename T>
size_t eval_lds_bytes(const KSys& sy, bool per_particle, int wpw)
{
    // the larger of the two layouts: histogram + real gradient sums (real
    // parameters), narrower histogram + real and imaginary sums (otherwise)
    const size_t nd_real = (size_t)sy.P_s * 4 * hist_cols(false, wpw) + (size_t)sy.N * sy.D;
    const size_t nd_cplx = (size_t)sy.P_s * 4 * hist_cols(true, wpw) + (size_t)sy.N * 2 * sy.D;
    const size_t nd = (per_particle ? nd_cplx + 2 * (size_t)sy.N : std::max(nd_real, nd_cplx));
    // coordinates (Coord storage has the size of T) and rho
    const size_t nt = (size_t)sy.N * sy.D + 2 * (size_t)sy.Kh;
    // the k-vector table and the derivative table start 16-byte aligned
    const size_t head = (kRedDoubles * sizeof(double) + nd * sizeof(double) + nt * sizeof(T) + 15) & ~(size_t)15;
    // non-uniform knots: the lookup cells follow the derivative table
    const size_t cells = sy.nu ? 4 * (size_t)sy.ncell : 0;
    return head + (4 * (size_t)sy.Kh + (size_t)sy.P_s * kDsplStride + cells) * sizeof(T);
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
                       double* EI, TO* O, double* other, double* grad, double* lap, double* outer,
                       const T* rho_in, hipStream_t st, double* bsum)
{
    if (W <= 0) return hipSuccess;
    EvalOut<TO> o{expo, ER, EI, O, other, grad, lap, outer, bsum};
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

template size_t eval_lds_bytes<float>(const KSys&, bool, int);
template size_t eval_lds_bytes<double>(const KSys&, bool, int);
template hipError_t launch_eval<double, double>(const KSys&, const KTables<double>&, const double*, int64_t, double*,
                                                double*, double*, double*, double*, double*, double*, double*,
                                                const double*, hipStream_t, double*);
template hipError_t launch_eval<float, float>(const KSys&, const KTables<float>&, const double*, int64_t, double*,
                                              double*, double*, float*, double*, double*, double*, double*,
                                              const float*, hipStream_t, double*);
template hipError_t launch_eval<float, double>(const KSys&, const KTables<float>&, const double*, int64_t, double*,
                                               double*, double*, double*, double*, double*, double*, double*,
                                               const float*, hipStream_t, double*);

}  // namespace tdvmc
