// Persistent Metropolis sweeps: every walker performs n_moves single-particle
// trial moves inside one launch (a sweep is N moves).
//
// Restates the move of src/TDVMC.cpp:858-916 (DoMetropolisStep: random
// particle, Gaussian displacement of width MC_STEP in every component, accept
// with probability min(1, |psi_new / psi_old|^2), restore on reject) with the
// incremental ratio of src/PhysicalSystems/BosonsBulk.cpp:553-657
// (CalculateWFChange: old and new pair terms of the moved particle, pairs
// beyond the cutoff contribute the last pair parameter), plus the k-space
// term (incremental rho_k update, one representative per +-k pair with the
// summed amplitude) and the one-body factor.
//
// Mapping: a walker is a group of LPW lanes (LPW = 16, 32 or 64) inside a
// wavefront; a 256-thread workgroup holds 256/LPW walkers.  The walker's
// coordinates live in LDS (structure of arrays) for the whole launch; lane g
// of the group handles partners i = g, g + LPW, ... and k-vectors
// k = g + m LPW (their rho_k stay in that lane's registers).  The change of
// ln|psi| is summed with DPP row steps whose result is bitwise identical in
// every lane of the group, so all lanes take the same accept decision without
// further communication.  Random numbers come from Philox4x32-10 (counter =
// trial, call epoch, walker): each lane draws the numbers of one future
// trial, and trial t of a batch reads them from lane t (v_readlane when the
// group is the whole wavefront).  The pair terms are branch-free.  In a
// periodic box the coordinates are held in box units (x / L): as 32-bit
// fractions of the box in the FP32 form (the wrapping integer difference is
// the minimum image), as d - rint(d) in FP64; the phase in revolutions is
// n . s, and the spline table runs out to the largest minimum-image distance
// with the outer-pair value past the cutoff.

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

constexpr int kThreads = 256;

template <typename T>
struct Vec4 { T x, y, z, w; };

template <int LPW, typename T>
__device__ __forceinline__ T bcast(T v, int src)
{
    if (LPW == 64) return readlane_f(v, src);
    return __shfl(v, src, LPW);
}
template <int LPW>
__device__ __forceinline__ int bcast_i(int v, int src)
{
    if (LPW == 64) return readlane_i(v, src);
    return __shfl(v, src, LPW);
}
template <int LPW>
__device__ __forceinline__ uint32_t bcast(uint32_t v, int src)
{
    return (uint32_t)bcast_i<LPW>((int)v, src);
}

// Metropolis decision 2 (sum over the group of d - us) >= lu, the same in
// every lane of each aligned group of LPW lanes.  Groups of at most 16 lanes
// sum within their row (every lane gets the same bits).  Wider groups add the
// row sums with row broadcasts: rows 1 and 3 receive r0 + r1 and r2 + r3
// (LPW = 32), row 3 then r0 + r1 + r2 + r3 (LPW = 64); the last lane of the
// group decides and a ballot hands its decision to the group.
constexpr int kDppRowBcast15 = 0x142;   // row r + 1 <- lane 15 of row r
constexpr int kDppRowBcast31 = 0x143;   // rows 2, 3 <- lane 31
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_rows(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_rows(double v)
{
    const int2 x = *reinterpret_cast<int2*>(&v);
    int2 y;
    y.x = __builtin_amdgcn_update_dpp(0, x.x, CTRL, ROWS, 0xF, false);
    y.y = __builtin_amdgcn_update_dpp(0, x.y, CTRL, ROWS, 0xF, false);
    return *reinterpret_cast<double*>(&y);
}
template <int LPW, typename T>
__device__ __forceinline__ bool group_accept(T d, T us, T lu)
{
    if constexpr (LPW <= 16) {
        return (T)2 * (group_sum<LPW>(d) - us) >= lu;
    } else {
        d = group_sum<16>(d);
        d += dpp_rows<kDppRowBcast15, 0xA>(d);
        if constexpr (LPW == 64) d += dpp_rows<kDppRowBcast31, 0x8>(d);
        const uint64_t m = __builtin_amdgcn_ballot_w64((T)2 * (d - us) >= lu);
        if constexpr (LPW == 64) return (m >> 63) & 1;
        return ((threadIdx.x & 63) < 32 ? (m >> 31) : (m >> 63)) & 1;
    }
}

// u(r) from the spline table.  Periodic box: the table extends over every
// possible minimum-image distance, entries past the cutoff hold the
// outer-pair value, so the lookup needs no clamp or select.  Open space:
// clamp to the last interval and select the outer value past the cutoff.
template <typename T, bool PER>
__device__ __forceinline__ T pair_u(T r, T rc, T hinv, int Ps, const Vec4<T>* spl, T u_out)
{
    const T x = r * hinv;
    if (PER) {
        const int j = (int)x;
        const T t = Prec<T>::fract_(x);
        const Vec4<T> cf = spl[j];
        return cf.x + t * (cf.y + t * (cf.z + t * cf.w));
    }
    int j = (int)x;
    j = j > Ps - 2 ? Ps - 2 : j;
    const T t = x - (T)j;
    const Vec4<T> cf = spl[j];
    const T u = cf.x + t * (cf.y + t * (cf.z + t * cf.w));
    return r <= rc ? u : u_out;
}

// u(r) on non-uniform knots (periodic box): cell c = min((int)(r cin), ncell-1)
// gives the interval (j, or j + 1 past the cell's knot b) and its left knot;
// the piece is a cubic in s = (r - g) unitR (real length).  Row P_s - 1 of
// spl is the constant outer-pair value beyond rc.
template <typename T>
__device__ __forceinline__ T pair_u_nu(T r, T cin, int cmax, T unitR, const Vec4<T>* cells, const Vec4<T>* spl)
{
    int c = (int)(r * cin);
    c = c < cmax ? c : cmax;
    const Vec4<T> e = cells[c];
    const bool right = r > e.x;
    const int j = (int)e.y + (right ? 1 : 0);
    const T s = (r - (right ? e.w : e.z)) * unitR;
    const Vec4<T> cf = spl[j];
    return cf.x + s * (cf.y + s * (cf.z + s * cf.w));
}

// Rows of the spline table in LDS (uniform knots: intervals, extended in a
// periodic box to the largest minimum-image distance; non-uniform knots: P_s
// interval rows, then the lookup cells).
__host__ __device__ inline int sweep_table_len(const KSys& sy)
{
    if (sy.nu) return sy.P_s + sy.ncell;
    if (!sy.periodic) return sy.P_s - 1;
    // largest minimum-image distance sqrt(D)/2 L, in knot intervals, + margin
    const double xmax = 0.5 * sqrt((double)sy.D) * sy.L * sy.hinv;
    const int j = (int)xmax + 2;
    return j > sy.P_s - 1 ? j : sy.P_s - 1;
}

// One workgroup's share of the sweep: walkers blk * WPB .. (blk + 1) * WPB - 1
// of the system described by (sy, tb, a).
template <typename T, int D, int LPW, int KPL, bool PER, bool NU = false>
__device__ __forceinline__ void sweep_body(const KSys& sy, const KTables<T>& tb, const SweepArgs& a, int64_t blk)
{
    using C = Coord<T, PER>;
    using S = typename C::S;
    using Inc = typename C::Inc;
    constexpr int WPB = kThreads / LPW;   // walkers per block
    const int N = sy.N, Ps = sy.P_s, K = sy.Kh;   // k-space: one representative per +-k pair
    const int NT = sweep_table_len(sy);
    const int tid = threadIdx.x;
    const int grp = tid / LPW, gl = tid % LPW;
    const int64_t w = blk * WPB + grp;

    extern __shared__ __align__(16) unsigned char smem[];
    Vec4<T>* spl = reinterpret_cast<Vec4<T>*>(smem);                      // NT x (A0..A3) of Re u
    Vec4<T>* kvl = spl + NT;                                               // K x (n/L, |k|^2)
    T* all = reinterpret_cast<T*>(kvl + K);                                // K x Re alpha
    S* posb = reinterpret_cast<S*>(all + ((K + 3) & ~3));                  // WPB x N x D (particle-major)
    S* pos = posb + (size_t)grp * N * D;

    const T u_out = (T)tb.uR[sy.outer_idx];
    // real length per coordinate-difference unit
    const double scale = PER ? sy.Linv : 1.0;
    const double unit_len = C::kUnit / scale;
    if constexpr (NU) {
        // P_s rows (the last one the outer constant), then the cells in
        // coordinate-difference units
        for (int i = tid; i < Ps; i += kThreads) {
            spl[i].x = tb.spl[i * 8 + 0];
            spl[i].y = tb.spl[i * 8 + 1];
            spl[i].z = tb.spl[i * 8 + 2];
            spl[i].w = tb.spl[i * 8 + 3];
        }
        for (int i = tid; i < sy.ncell; i += kThreads) {
            const double* e = tb.nucell + 4 * i;
            spl[Ps + i].x = (T)(e[0] / unit_len);
            spl[Ps + i].y = (T)e[1];
            spl[Ps + i].z = (T)(e[2] / unit_len);
            spl[Ps + i].w = (T)(e[3] / unit_len);
        }
    }
    for (int i = NU ? NT : tid; i < NT; i += kThreads) {
        if (i < Ps - 1) {
            spl[i].x = tb.spl[i * 8 + 0];
            spl[i].y = tb.spl[i * 8 + 1];
            spl[i].z = tb.spl[i * 8 + 2];
            spl[i].w = tb.spl[i * 8 + 3];
        } else {
            spl[i].x = u_out;
            spl[i].y = spl[i].z = spl[i].w = (T)0;
        }
    }
    for (int i = tid; i < K; i += kThreads) {
        kvl[i].x = tb.kv[i * 4 + 0];
        kvl[i].y = tb.kv[i * 4 + 1];
        kvl[i].z = tb.kv[i * 4 + 2];
        kvl[i].w = tb.kv[i * 4 + 3];
        all[i] = tb.alpha[2 * i];
    }
    // lengths in box units when periodic (scale = 1/L), unscaled otherwise;
    // diff() and raw() return multiples of C::kUnit of that unit
    const bool active = w < a.W;
    double* Rw = a.R + (active ? w : 0) * (size_t)N * D;
    if (active) {
        for (int e = gl; e < N * D; e += LPW) pos[e] = C::load(Rw[e] * scale);
    }
    __syncthreads();
    if (!active) return;

    const T hinv = (T)(sy.hinv / scale * C::kUnit), rc = (T)(sy.rc * scale / C::kUnit);
    const T cin = (T)(sy.cinv * unit_len), unitR = (T)unit_len;
    const Vec4<T>* cells = spl + Ps;
    auto uof = [&](T r) -> T {
        if constexpr (NU)
            return pair_u_nu<T>(r, cin, sy.ncell - 1, unitR, cells, spl);
        else
            return pair_u<T, PER>(r, rc, hinv, Ps, spl, u_out);
    };
    const T step = (T)(sy.step * scale);
    // one-body factor on raw coordinates: -ob sum raw^2
    const T ob = sy.one_body ? (T)(tb.uR[0] / (scale * scale) * C::kUnit * C::kUnit) : (T)0;

    // this lane's k-vectors (revolutions per raw coordinate unit) and
    // amplitudes, and twice their rho_k (sR + i sI = 2 rho_k: scaling by 2 is
    // exact, and it saves a multiply per k and trial)
    T sR[KPL > 0 ? KPL : 1], sI[KPL > 0 ? KPL : 1];
    T qx[KPL > 0 ? KPL : 1], qy[KPL > 0 ? KPL : 1], qz[KPL > 0 ? KPL : 1], qa[KPL > 0 ? KPL : 1];
#pragma unroll
    for (int m = 0; m < KPL; m++) {
        sR[m] = sI[m] = (T)0;
        const int k = gl + m * LPW;
        qx[m] = qy[m] = qz[m] = qa[m] = (T)0;   // lanes past K: zero amplitude, no effect
        if (k < K) {
            // periodic: kv L is the integer lattice vector n (rounded to
            // remove the storage error of n / L)
            const Vec4<T> q = kvl[k];
            auto qs = [&](T v) { return (T)((PER ? rint((double)v / scale) : (double)v) * C::kUnit); };
            qx[m] = qs(q.x);
            qy[m] = qs(q.y);
            qz[m] = qs(q.z);
            qa[m] = all[k];
            for (int n = 0; n < N; n++) {
                T ph = qx[m] * C::raw(pos[n * D]);
                if (D > 1) ph += qy[m] * C::raw(pos[n * D + 1]);
                if (D > 2) ph += qz[m] * C::raw(pos[n * D + 2]);
                T s, c;
                Prec<T>::sincos_small(ph, &s, &c);
                sR[m] += c;
                sI[m] += s;
            }
            sR[m] *= (T)2;
            sI[m] *= (T)2;
        }
    }
    // u at distance 0: the moved particle meets itself in the partner loop
    const T u_zero = uof((T)0);

    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const uint32_t ep = a.epoch ? *a.epoch : 0u;
    const int64_t ntr = a.n_moves;
    unsigned long long nacc = 0;
    int seq = 0;   // sequential order: particle of trial base + gl

    for (int64_t base = 0; base < ntr; base += LPW) {
        // lane gl prepares trial base + gl: particle, displacement increments,
        // log u, and the self term u(|move|) - u(0) that the partner loop adds
        const uint32_t tr = (uint32_t)(base + gl);
        Philox4 x = philox4x32_10(tr, ep, (uint32_t)w, 0u, k0, k1);
        int pm;
        if (sy.move_order == 0) {
            pm = (int)__umulhi(x.v[0], (uint32_t)N);
        } else {
            pm = seq + gl;
            while (pm >= N) pm -= N;
            seq += LPW;
            while (seq >= N) seq -= N;
        }
        Inc dl[3];
        T logu, uself;
        {
            // Box-Muller on (v1, v2) -> two normals; D = 3 draws a second block
            T z[3];
            const float u1 = u01f(x.v[1]), u2 = u01f(x.v[2]);
            const float rr = sqrtf(-2.0f * __logf(u1));
            float sn, cs;
            Prec<float>::sincos_rev(u2, &sn, &cs);
            z[0] = (T)(rr * cs);
            z[1] = (T)(rr * sn);
            if (D == 3) {
                Philox4 y = philox4x32_10(tr, ep, (uint32_t)w, 1u, k0, k1);
                const float v1 = u01f(y.v[0]), v2 = u01f(y.v[1]);
                const float r2 = sqrtf(-2.0f * __logf(v1));
                float s2, c2;
                Prec<float>::sincos_rev(v2, &s2, &c2);
                z[2] = (T)(r2 * c2);
            } else {
                z[2] = (T)0;
            }
            logu = (T)__logf(u01f(x.v[3]));
            T s2 = (T)0;
#pragma unroll
            for (int c = 0; c < D; c++) {
                dl[c] = C::inc(step * z[c]);
                const T dm = C::diff((S)0, C::add((S)0, dl[c]));
                s2 += dm * dm;
            }
            uself = uof(Prec<T>::sqrt_(s2)) - u_zero;
        }
        const int nb = (int)((ntr - base) < LPW ? (ntr - base) : LPW);
        for (int tt = 0; tt < nb; tt++) {
            const int p = bcast_i<LPW>(pm, tt);
            S* const pp = pos + __umul24((uint32_t)p, (uint32_t)D);   // the moved particle's slot
            S old[D], nw[D];
#pragma unroll
            for (int c = 0; c < D; c++) {
                old[c] = pp[c];
                nw[c] = C::add(old[c], bcast<LPW>(dl[c], tt));
            }
            const T lu = bcast<LPW>(logu, tt);
            const T us = bcast<LPW>(uself, tt);
            T delta = (T)0;
            auto partner = [&](int i) {
                T ro2 = (T)0, rn2 = (T)0;
#pragma unroll
                for (int c = 0; c < D; c++) {
                    const S pi = pos[i * D + c];
                    const T dOld = C::diff(pi, old[c]), dNew = C::diff(pi, nw[c]);
                    ro2 += dOld * dOld;
                    rn2 += dNew * dNew;
                }
                const T uo = uof(Prec<T>::sqrt_(ro2));
                const T un = uof(Prec<T>::sqrt_(rn2));
                delta += un - uo;
            };
            // rounds in which every lane of the group has a partner run under a
            // bound common to the group (loop control on scalar registers), the
            // last, partial round under the lane's own test; same partners, same
            // order as one loop over i = gl, gl + LPW, ... < N
            const int nfull = N / LPW;
#pragma unroll 1
            for (int m = 0; m < nfull; m++) partner(gl + m * LPW);
            if (gl + nfull * LPW < N) partner(gl + nfull * LPW);
            T oc[D], ncd[D];
#pragma unroll
            for (int c = 0; c < D; c++) {
                oc[c] = C::raw(old[c]);
                ncd[c] = C::raw(nw[c]);
            }
            T dR[KPL > 0 ? KPL : 1], dI[KPL > 0 ? KPL : 1];
#pragma unroll
            for (int m = 0; m < KPL; m++) {
                // |phase| <= sum |n_a| / 2 revolutions: no range reduction needed
                T po = qx[m] * oc[0], pn = qx[m] * ncd[0];
                if (D > 1) {
                    po += qy[m] * oc[1];
                    pn += qy[m] * ncd[1];
                }
                if (D > 2) {
                    po += qz[m] * oc[2];
                    pn += qz[m] * ncd[2];
                }
                T so, co, sn, cn;
                Prec<T>::sincos_small(po, &so, &co);
                Prec<T>::sincos_small(pn, &sn, &cn);
                dR[m] = cn - co;
                dI[m] = sn - so;
                // |rho + d|^2 - |rho|^2 = 2 Re(conj(rho) d) + |d|^2 = Re((2 rho + d) conj(d))
                delta += qa[m] * ((sR[m] + dR[m]) * dR[m] + (sI[m] + dI[m]) * dI[m]);
            }
            if (sy.one_body) {
                // a branch on the (uniform) flag, not a select: systems without
                // the one-body factor skip these instructions
                asm volatile("" ::: "memory");
                T ro2 = (T)0, rn2 = (T)0;
#pragma unroll
                for (int c = 0; c < D; c++) {
                    ro2 += oc[c] * oc[c];
                    rn2 += ncd[c] * ncd[c];
                }
                if (gl == 0) delta += -ob * (rn2 - ro2);
            }
            const bool accept = group_accept<LPW>(delta, us, lu);
            if (accept) {
                if (gl == 0) {
#pragma unroll
                    for (int c = 0; c < D; c++) pp[c] = nw[c];
                }
#pragma unroll
                for (int m = 0; m < KPL; m++) {
                    sR[m] += (T)2 * dR[m];
                    sI[m] += (T)2 * dI[m];
                }
                nacc++;
            }
            // the moved particle's LDS slot is read by every lane of the group next trial
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }

    for (int e = gl; e < N * D; e += LPW) Rw[e] = C::store(pos[e]) * (PER ? sy.L : 1.0);
    if (a.rho_out) {
        T* ro = static_cast<T*>(a.rho_out) + (size_t)w * K * 2;
#pragma unroll
        for (int m = 0; m < KPL; m++) {
            const int k = gl + m * LPW;
            if (k < K) {
                ro[2 * k] = (T)0.5 * sR[m];
                ro[2 * k + 1] = (T)0.5 * sI[m];
            }
        }
    }
    if (gl == 0 && a.acc) a.acc[w] += nacc;
}

template <typename T, int D, int LPW, int KPL, bool PER, bool NU>
__global__ __launch_bounds__(kThreads) void sweep_kernel(KSys sy, KTables<T> tb, SweepArgs a)
{
    sweep_body<T, D, LPW, KPL, PER, NU>(sy, tb, a, blockIdx.x);
}

// Batch of independent systems in one launch: workgroup b belongs to system
// s with block_start[s] <= b < block_start[s + 1].
template <typename T, int D, int LPW, int KPL, bool PER>
__global__ __launch_bounds__(kThreads) void sweep_batch_kernel(SweepBatch<T> bt)
{
    int lo = 0, hi = bt.n;   // last s with block_start[s] <= blockIdx.x
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bt.block_start[mid] <= (int64_t)blockIdx.x)
            lo = mid;
        else
            hi = mid;
    }
    sweep_body<T, D, LPW, KPL, PER>(bt.sys[lo], bt.tab[lo], bt.args[lo], (int64_t)blockIdx.x - bt.block_start[lo]);
}

// Default lanes per walker without tuning: 32 from N = 24 on (two walkers
// per wavefront share the per-trial overhead and waste fewer partner slots
// than one walker per wavefront), 16 below; more when the k-space terms would
// need more than 8 per lane.
int pick_lpw(int N, int K)
{
    int l = N >= 24 ? 32 : 16;
    while (l < 64 && K > 8 * l) l *= 2;
    return l;
}

}  // namespace

template <typename T>
size_t sweep_lds_bytes(const KSys& sy, int lpw)
{
    if (lpw <= 0) lpw = pick_lpw(sy.N, sy.Kh);
    const size_t wpb = kThreads / lpw;
    // coordinates: uint32 fractions (FP32, periodic) or T
    const size_t cs = (sizeof(T) == 4 && sy.periodic) ? sizeof(uint32_t) : sizeof(T);
    return sizeof(T) * (4 * (size_t)sweep_table_len(sy) + 4 * (size_t)sy.Kh + (((size_t)sy.Kh + 3) & ~(size_t)3)) +
           cs * wpb * sy.N * sy.D;
}

template <typename T, int D, int LPW, int KPL, bool PER>
static hipError_t launch_sweep_k(const KSys& sy, const KTables<T>& tb, const SweepArgs& a, hipStream_t st)
{
    const size_t lds = sweep_lds_bytes<T>(sy, LPW);
    if (sy.nu && (!PER || !tb.nucell)) return hipErrorInvalidValue;
    auto kern = sy.nu ? sweep_kernel<T, D, LPW, KPL, PER, PER> : sweep_kernel<T, D, LPW, KPL, PER, false>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int wpb = kThreads / LPW;
    const int64_t blocks = (a.W + wpb - 1) / wpb;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kThreads), lds, st, sy, tb, a);
    return hipGetLastError();
}

template <typename T, int D, int LPW, bool PER>
static hipError_t launch_sweep_l(const KSys& sy, const KTables<T>& tb, const SweepArgs& a, hipStream_t st)
{
    const int kpl = (sy.Kh + LPW - 1) / LPW;
    if (kpl == 0) return launch_sweep_k<T, D, LPW, 0, PER>(sy, tb, a, st);
    if constexpr (!PER) {
        return hipErrorInvalidValue;   // k-space terms need the periodic box
    } else {
        if (kpl <= 2) return launch_sweep_k<T, D, LPW, 2, PER>(sy, tb, a, st);
        if (kpl <= 4) return launch_sweep_k<T, D, LPW, 4, PER>(sy, tb, a, st);
        if (kpl <= 8) return launch_sweep_k<T, D, LPW, 8, PER>(sy, tb, a, st);
        return hipErrorInvalidValue;
    }
}

template <typename T, int D>
static hipError_t launch_sweep_d(const KSys& sy, const KTables<T>& tb, const SweepArgs& a, hipStream_t st)
{
    const int lpw = a.lanes_per_walker > 0 ? a.lanes_per_walker : pick_lpw(sy.N, sy.Kh);
    const bool per = sy.periodic != 0;
    switch (lpw) {
    case 16: return per ? launch_sweep_l<T, D, 16, true>(sy, tb, a, st) : launch_sweep_l<T, D, 16, false>(sy, tb, a, st);
    case 32: return per ? launch_sweep_l<T, D, 32, true>(sy, tb, a, st) : launch_sweep_l<T, D, 32, false>(sy, tb, a, st);
    case 64: return per ? launch_sweep_l<T, D, 64, true>(sy, tb, a, st) : launch_sweep_l<T, D, 64, false>(sy, tb, a, st);
    default: return hipErrorInvalidValue;
    }
}

// Batch without k-space terms, each system with its own lanes per walker
// (8, 16, 32 or 64, from its SweepArgs); a workgroup belongs to one system.
template <typename T, int D, bool PER>
__global__ __launch_bounds__(kThreads) void sweep_batch_lanes_kernel(SweepBatch<T> bt)
{
    int lo = 0, hi = bt.n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bt.block_start[mid] <= (int64_t)blockIdx.x)
            lo = mid;
        else
            hi = mid;
    }
    const int64_t blk = (int64_t)blockIdx.x - bt.block_start[lo];
    switch (bt.args[lo].lanes_per_walker) {
    case 8: sweep_body<T, D, 8, 0, PER>(bt.sys[lo], bt.tab[lo], bt.args[lo], blk); break;
    case 16: sweep_body<T, D, 16, 0, PER>(bt.sys[lo], bt.tab[lo], bt.args[lo], blk); break;
    case 64: sweep_body<T, D, 64, 0, PER>(bt.sys[lo], bt.tab[lo], bt.args[lo], blk); break;
    default: sweep_body<T, D, 32, 0, PER>(bt.sys[lo], bt.tab[lo], bt.args[lo], blk); break;
    }
}

template <typename T, int D, bool PER>
static hipError_t launch_sweep_batch_lanes(const SweepBatch<T>& bt, int64_t blocks, size_t lds, hipStream_t st)
{
    auto kern = sweep_batch_lanes_kernel<T, D, PER>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kThreads), lds, st, bt);
    return hipGetLastError();
}

template <typename T, int D, int LPW, int KPL, bool PER>
static hipError_t launch_sweep_batch_k(const SweepBatch<T>& bt, int64_t blocks, size_t lds, hipStream_t st)
{
    auto kern = sweep_batch_kernel<T, D, LPW, KPL, PER>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kThreads), lds, st, bt);
    return hipGetLastError();
}

template <typename T, int D, int LPW, bool PER>
static hipError_t launch_sweep_batch_l(const SweepBatch<T>& bt, int kh_max, int64_t blocks, size_t lds,
                                       hipStream_t st)
{
    const int kpl = (kh_max + LPW - 1) / LPW;
    if (kpl == 0) return launch_sweep_batch_k<T, D, LPW, 0, PER>(bt, blocks, lds, st);
    if constexpr (!PER) {
        return hipErrorInvalidValue;
    } else {
        if (kpl <= 2) return launch_sweep_batch_k<T, D, LPW, 2, PER>(bt, blocks, lds, st);
        if (kpl <= 4) return launch_sweep_batch_k<T, D, LPW, 4, PER>(bt, blocks, lds, st);
        if (kpl <= 8) return launch_sweep_batch_k<T, D, LPW, 8, PER>(bt, blocks, lds, st);
        return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_sweep_batch(const SweepBatch<T>& bt, int D, bool periodic, int lanes, int kh_max, int64_t blocks,
                              size_t lds, hipStream_t st)
{
    if (blocks <= 0) return hipSuccess;
    if (lanes == 0) {
        // per-system lanes (no k-space terms)
        if (kh_max > 0) return hipErrorInvalidValue;
        switch (D) {
        case 1: return periodic ? launch_sweep_batch_lanes<T, 1, true>(bt, blocks, lds, st)
                                : launch_sweep_batch_lanes<T, 1, false>(bt, blocks, lds, st);
        case 2: return periodic ? launch_sweep_batch_lanes<T, 2, true>(bt, blocks, lds, st)
                                : launch_sweep_batch_lanes<T, 2, false>(bt, blocks, lds, st);
        case 3: return periodic ? launch_sweep_batch_lanes<T, 3, true>(bt, blocks, lds, st)
                                : launch_sweep_batch_lanes<T, 3, false>(bt, blocks, lds, st);
        default: return hipErrorInvalidValue;
        }
    }
#define TDVMC_SWEEP_BATCH_D(DD)                                                                              \
    switch (lanes) {                                                                                         \
    case 16: return periodic ? launch_sweep_batch_l<T, DD, 16, true>(bt, kh_max, blocks, lds, st)             \
                             : launch_sweep_batch_l<T, DD, 16, false>(bt, kh_max, blocks, lds, st);           \
    case 32: return periodic ? launch_sweep_batch_l<T, DD, 32, true>(bt, kh_max, blocks, lds, st)             \
                             : launch_sweep_batch_l<T, DD, 32, false>(bt, kh_max, blocks, lds, st);           \
    case 64: return periodic ? launch_sweep_batch_l<T, DD, 64, true>(bt, kh_max, blocks, lds, st)             \
                             : launch_sweep_batch_l<T, DD, 64, false>(bt, kh_max, blocks, lds, st);           \
    default: return hipErrorInvalidValue;                                                                    \
    }
    switch (D) {
    case 1: TDVMC_SWEEP_BATCH_D(1)
    case 2: TDVMC_SWEEP_BATCH_D(2)
    case 3: TDVMC_SWEEP_BATCH_D(3)
    default: return hipErrorInvalidValue;
    }
#undef TDVMC_SWEEP_BATCH_D
}

int sweep_walkers_per_block(int lanes) { return kThreads / lanes; }
int sweep_default_lanes(int N, int Kh) { return pick_lpw(N, Kh); }

template <typename T>
hipError_t launch_sweep(const KSys& sy, const KTables<T>& tb, const SweepArgs& a, hipStream_t st)
{
    if (a.W <= 0 || a.n_moves <= 0) return hipSuccess;
    switch (sy.D) {
    case 1: return launch_sweep_d<T, 1>(sy, tb, a, st);
    case 2: return launch_sweep_d<T, 2>(sy, tb, a, st);
    case 3: return launch_sweep_d<T, 3>(sy, tb, a, st);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_sweep_batch<float>(const SweepBatch<float>&, int, bool, int, int, int64_t, size_t,
                                              hipStream_t);
template hipError_t launch_sweep_batch<double>(const SweepBatch<double>&, int, bool, int, int, int64_t, size_t,
                                               hipStream_t);
template hipError_t launch_sweep<float>(const KSys&, const KTables<float>&, const SweepArgs&, hipStream_t);
template hipError_t launch_sweep<double>(const KSys&, const KTables<double>&, const SweepArgs&, hipStream_t);
template size_t sweep_lds_bytes<float>(const KSys&, int);
template size_t sweep_lds_bytes<double>(const KSys&, int);

}  // namespace tdvmc
