// Device-side shared definitions for the TDVMC hot-path kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tdvmc {

// Launch-time description of one system (passed by value to every kernel).
// Parameter layout: [one-body (0 or 1)][pair spline P_s][k-space K].
struct KSys {
    int N, D;
    int P_s, K, P;
    int Kh;                // k-space terms after pairing k with -k (see KTables)
    int off_pair, off_k, one_body;
    int periodic;          // 1: minimum image in a box of length L
    int pot_type;          // 0 none, 1 square well, 2 Gaussian, 3 Lennard-Jones
    int outer_idx;         // parameter multiplying the count of pairs beyond rc
    int move_order;        // 0 random particle (reference), 1 sequential sweep
    double L, Linv;
    double rc, h, hinv;
    double hbar2_2m;
    double trap_omega;
    double pot_cutoff, pot_a, pot_b;   // pair potential resolved for the current time
    double step;           // Gaussian Metropolis step (MC_STEP)
    // non-uniform knots (USE_NURBS): interval lookup by cells, pieces in
    // s = r - g_j (KTables::nucell / nubeta); rc is then the last knot
    int nu;                // 0 uniform knots, 1 NURBS_GRID
    int ncell;             // lookup cells, the last one for every r beyond rc
    double cinv;           // cells per unit length
    double hlast;          // last knot gap g_{P_s-1} - g_{P_s-2}
    // boundary conditions of the pair spline: the first bc_n0 parameters
    // weight splines 0..2 by the rows of bc_f0, the last bc_n1 weight splines
    // P_s-1..P_s+1 by the rows of bc_f1, the others are splines 3, 4, ...
    // (bc_n0 + bc_n1 = 4; BosonsBulk's zero-slope pair: 2 + 2)
    int bc_n0, bc_n1;
    double bc_f0[3][3], bc_f1[3][3];
};

// Per-precision device tables derived from the parameters (written by
// prep_tables_kernel, so parameter updates stay on the device).
// k-space terms come in pairs k, -k whenever both are in the set.  For real
// coordinates rho_{-k} = conj(rho_k), so both share |rho_k|^2, its gradient
// and Laplacian; the kernels work on one representative per pair (Kh of them)
// with the amplitude sum alpha_k + alpha_{-k}, and write O = |rho_k|^2 to both
// parameter slots.
template <typename T>
struct KTables {
    const T* spl;     // [P_s][8]: cubic-in-t coefficients A0..A3 of Re u, then of Im u (non-uniform
                      //   knots: cubics in s = r - g_j; row P_s-1: the constant beyond rc)
    const T* dspl;    // [P_s][kDsplStride]: derivatives per unit length on knot interval j,
                      //   u' = d1 + t (d2 + t d3), u'' = hinv (d2 + t e3), e3 = 2 d3, stored as
                      //   (d1, d2, d3, e3) of Re u, then of Im u; row P_s-1 holds u'(rc)
                      //   and u''(rc) / hinv as d1, d2 (used at t = 0 only)
    const T* kv;      // [Kh][4]: n/L of the representative (phase in revolutions = kv . r), |k|^2
    const T* alpha;   // [Kh][2]: (Re, Im) of alpha_k (+ alpha_{-k} when paired)
    const T* kag;     // [Kh][8]: Re alpha_k times (kv, |k|^2), then Im alpha_k times the same
    const int* kmap;  // [Kh][2]: parameter slots (relative to off_k) of k and of -k (-1: unpaired)
    const double* uR; // [P + 1] parameters (FP64 master copy), then phi R
    const double* uI; // [P + 1] parameters, then phi I
    // non-uniform knots only (null otherwise); lengths in real units
    const double* nucell;   // [ncell][4]: knot b, interval j, g_j, g_{j+1}: r <= b -> j, else j + 1
    const double* nubeta;   // [P_s][4][4]: piece m of interval j as sum_p beta[j][m][p] s^p
};

constexpr int kMaxD = 3;
constexpr int kDsplStride = 8;    // entries per knot interval of KTables::dspl

template <int D> struct PadDim { static constexpr int value = (D == 3) ? 4 : D; };

// ------------------------------------------------------------- Philox4x32-10
// Counter-based generator of Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11): 10 rounds of the 4x32 Philox
// bijection with the published multipliers and Weyl key increments.
struct Philox4 {
    uint32_t v[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                  uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    Philox4 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
}

// uniform in (0, 1]
__device__ __forceinline__ float u01f(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }
__device__ __forceinline__ double u01d(uint32_t hi, uint32_t lo)
{
    uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
    m &= ((1ull << 53) - 1ull);
    return (double)(m + 1ull) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------- precision traits
template <typename T> struct Prec;

template <> struct Prec<float> {
    __device__ static __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static __forceinline__ float rint_(float x) { return __builtin_rintf(x); }
    __device__ static __forceinline__ float floor_(float x) { return __builtin_floorf(x); }
    __device__ static __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
    // r = sqrt(r2) and 1/r from one reciprocal square root
    __device__ static __forceinline__ void r_and_rinv(float r2, float* r, float* rinv)
    {
        const float q = __builtin_amdgcn_rsqf(r2);
        *rinv = q;
        *r = r2 * q;
    }
    // sin / cos of 2*pi*t, t in revolutions, |t| <= 0.5 after reduction
    __device__ static __forceinline__ void sincos_rev(float t, float* s, float* c)
    {
        t = t - __builtin_rintf(t);
        *s = __builtin_amdgcn_sinf(t);
        *c = __builtin_amdgcn_cosf(t);
    }
    // sin / cos of 2*pi*t for |t| <= 256 revolutions (v_sin / v_cos reduce
    // the argument themselves; within 1 ulp of the reduced form)
    __device__ static __forceinline__ void sincos_small(float t, float* s, float* c)
    {
        *s = __builtin_amdgcn_sinf(t);
        *c = __builtin_amdgcn_cosf(t);
    }
    __device__ static __forceinline__ float fract_(float x) { return __builtin_amdgcn_fractf(x); }
    __device__ static __forceinline__ float exp_(float x) { return __expf(x); }
    __device__ static __forceinline__ float exp2_(float x) { return __builtin_amdgcn_exp2f(x); }
    __device__ static __forceinline__ float log_(float x) { return __logf(x); }
};

template <> struct Prec<double> {
    __device__ static __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
    __device__ static __forceinline__ double rint_(double x) { return __builtin_rint(x); }
    __device__ static __forceinline__ double floor_(double x) { return __builtin_floor(x); }
    __device__ static __forceinline__ double rcp_(double x) { return 1.0 / x; }
    __device__ static __forceinline__ void r_and_rinv(double r2, double* r, double* rinv)
    {
        *r = __builtin_sqrt(r2);
        *rinv = 1.0 / *r;
    }
    __device__ static __forceinline__ void sincos_rev(double t, double* s, double* c)
    {
        t = t - __builtin_rint(t);
        sincospi(2.0 * t, s, c);
    }
    __device__ static __forceinline__ void sincos_small(double t, double* s, double* c) { sincos_rev(t, s, c); }
    __device__ static __forceinline__ double fract_(double x) { return x - __builtin_floor(x); }
    __device__ static __forceinline__ double exp_(double x) { return exp(x); }
    __device__ static __forceinline__ double exp2_(double x) { return exp2(x); }
    __device__ static __forceinline__ double log_(double x) { return log(x); }
};

// Storage of one coordinate in LDS.  FP32 in a periodic box: the position as
// an unsigned 32-bit fraction of the box, so a wrapping integer difference is
// the minimum image (one subtract and one convert per component, exact);
// distances come out in units of 2^-32 box lengths.  Otherwise the compute
// type in box units (periodic, minimum image by rint) or in real units.
template <typename T, bool PER>
struct Coord {
    using S = T;
    static constexpr double kUnit = 1.0;                 // length unit of diff() in box (or real) units
    __device__ static S load(double x) { return (T)(PER ? x - rint(x) : x); }
    __device__ static double store(S p) { return (double)p; }
    __device__ static T diff(S a, S b)
    {
        T d = a - b;
        if (PER) d -= Prec<T>::rint_(d);
        return d;
    }
    // a displacement dz (box or real units) as an increment, and its application
    using Inc = T;
    __device__ static Inc inc(T dz) { return dz; }
    __device__ static S add(S p, Inc d)
    {
        T v = p + d;
        if (PER) v -= Prec<T>::rint_(v);
        return v;
    }
    __device__ static S move(S p, T dz) { return add(p, inc(dz)); }
    __device__ static T unit_coord(S p) { return p; }   // in box (or real) units
    __device__ static T raw(S p) { return p; }           // value in units of kUnit
};
template <>
struct Coord<float, true> {
    using S = uint32_t;
    static constexpr double kUnit = 2.3283064365386963e-10;   // 2^-32
    __device__ static S load(double x)
    {
        const double f = x - floor(x);   // [0, 1)
        return (uint32_t)(uint64_t)(f * 4294967296.0);
    }
    __device__ static double store(S p) { return (double)(int32_t)p * kUnit; }   // [-1/2, 1/2)
    __device__ static float diff(S a, S b) { return (float)(int32_t)(a - b); }
    // the same minimum-image difference without rounding (32 significant bits)
    __device__ static double diff_exact(S a, S b) { return (double)(int32_t)(a - b); }
    using Inc = uint32_t;
    __device__ static Inc inc(float dz)
    {
        // dz in box units; only its image in [-1/2, 1/2) matters
        const float w = dz - __builtin_rintf(dz);
        return (uint32_t)__float2int_rn(w * 4294967296.0f);
    }
    __device__ static S add(S p, Inc d) { return p + d; }   // wraps: the periodic image
    __device__ static S move(S p, float dz) { return add(p, inc(dz)); }
    __device__ static float unit_coord(S p) { return (float)(int32_t)p * (float)kUnit; }
    __device__ static float raw(S p) { return (float)(int32_t)p; }
};

// minimum-image component (round to nearest multiple of L)
template <typename T>
__device__ __forceinline__ T min_image(T d, T L, T Linv, int periodic)
{
    return periodic ? d - L * Prec<T>::rint_(d * Linv) : d;
}

// Knot interval j in [0, P_s-2] and local coordinate t in [0,1] of distance r.
template <typename T>
__device__ __forceinline__ int knot_interval(T r, T hinv, int Ps, T* t)
{
    T x = r * hinv;
    int j = (int)x;
    j = j > Ps - 2 ? Ps - 2 : j;
    *t = x - (T)j;
    return j;
}

// ------------------------------------------------------------- lane exchange
// DPP controls (GFX9 encoding)
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppRowMirror = 0x140;  // lane l <- 15 - l within a row of 16
constexpr int kDppHalfMirror = 0x141; // lane l <- 7 - l within a half row of 8

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f(double v)
{
    const int2 x = *reinterpret_cast<int2*>(&v);
    int2 y;
    y.x = __builtin_amdgcn_update_dpp(0, x.x, CTRL, 0xF, 0xF, false);
    y.y = __builtin_amdgcn_update_dpp(0, x.y, CTRL, 0xF, 0xF, false);
    return *reinterpret_cast<double*>(&y);
}

__device__ __forceinline__ float readlane_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double readlane_f(double v, int l)
{
    const int2 x = *reinterpret_cast<int2*>(&v);
    int2 y;
    y.x = __builtin_amdgcn_readlane(x.x, l);
    y.y = __builtin_amdgcn_readlane(x.y, l);
    return *reinterpret_cast<double*>(&y);
}
__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Sum over each aligned group of G lanes (G = 16, 32 or 64).  Every lane of a
// group receives the same bits: the in-row steps pair lanes symmetrically
// (xor within quads, then mirrors), so each add sees the same two operands in
// either order, and the rows are combined in a fixed order.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v)
{
    v += dpp_f<kDppXor1>(v);
    v += dpp_f<kDppXor2>(v);
    if (G == 4) return v;
    v += dpp_f<kDppHalfMirror>(v);
    if (G == 8) return v;
    v += dpp_f<kDppRowMirror>(v);
    if (G == 16) return v;
    const T r0 = readlane_f(v, 0), r1 = readlane_f(v, 16), r2 = readlane_f(v, 32), r3 = readlane_f(v, 48);
    if (G == 64) return (r0 + r1) + (r2 + r3);
    return ((threadIdx.x & 63) < 32) ? (r0 + r1) : (r2 + r3);
}

}  // namespace tdvmc
