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
};

// Per-precision device tables derived from the parameters (written by
// prep_tables_kernel, so parameter updates stay on the device).
template <typename T>
struct KTables {
    const T* spl;     // [(P_s-1)][8]: cubic-in-t coefficients A0..A3 of Re u, then of Im u
    const T* kv;      // [K][4]: n/L (phase in revolutions = kv . r), |k|^2
    const T* alpha;   // [K][2]: (Re, Im) of the k-space amplitudes
    const double* uR; // [P] parameters (FP64 master copy)
    const double* uI;
};

constexpr int kMaxD = 3;

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
    // sin / cos of 2*pi*t, t in revolutions, |t| <= 0.5 after reduction
    __device__ static __forceinline__ void sincos_rev(float t, float* s, float* c)
    {
        t = t - __builtin_rintf(t);
        *s = __builtin_amdgcn_sinf(t);
        *c = __builtin_amdgcn_cosf(t);
    }
    __device__ static __forceinline__ float exp_(float x) { return __expf(x); }
    __device__ static __forceinline__ float log_(float x) { return __logf(x); }
};

template <> struct Prec<double> {
    __device__ static __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
    __device__ static __forceinline__ double rint_(double x) { return __builtin_rint(x); }
    __device__ static __forceinline__ double floor_(double x) { return __builtin_floor(x); }
    __device__ static __forceinline__ double rcp_(double x) { return 1.0 / x; }
    __device__ static __forceinline__ void sincos_rev(double t, double* s, double* c)
    {
        t = t - __builtin_rint(t);
        sincospi(2.0 * t, s, c);
    }
    __device__ static __forceinline__ double exp_(double x) { return exp(x); }
    __device__ static __forceinline__ double log_(double x) { return log(x); }
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

}  // namespace tdvmc
