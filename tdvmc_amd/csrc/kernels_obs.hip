// Additional observables of every walker: the pair distribution g(r) and the
// structure factor S(k) of BosonsBulk::CalculateAdditionalSystemProperties
// (src/PhysicalSystems/BosonsBulk.cpp:474-520).
//
// One 256-thread workgroup per walker, coordinates staged in LDS as FP64
// (padded to 4 components; 32-bit box fractions as well in the mixed form).
//   g(r): every unordered pair once (row i meets partners i + 1 .. i + M
//   mod N, M = (N-1)/2, rows i < N/2 also i + N/2 for even N; a row is split
//   into segments over threads); the pair's bin is counted in an LDS u32 histogram, and the counts
//   reach a u64 total per bin with one global atomic each (exact and
//   order-free).  The value of a bin is count x (D/(N-1)) / scaling[bin]
//   (ObservableVsOnGridWithScaling::AddToHistogram, :45-50).
//   S(k): the shells' wave vectors are integer n with non-negative
//   components, k = 2 pi n / L, so exp(i k.r) = prod_a e_a^{n_a} with
//   e_a = exp(2 pi i x_a / L).  Per chunk of particles the powers e_a^m,
//   m = 0..nmax, are tabulated in LDS (particle index fastest, so a thread
//   walks three fixed rows with immediate offsets); a thread then forms the products for
//   its wave vectors (D-1 complex multiplies per particle and vector) and
//   adds them up (the vectors past the last multiple of 256 are split over
//   the threads by particle residue).  One thread per shell sums its
//   vectors' totals in order:
//   S = (C^2 + S^2) / (N n_k).
// FP64 form: distances by the reference's minimum image (src/Utils.cpp:266-281)
// and FP64 division for the bin, powers by complex recurrence from one
// sincospi.  Mixed form: FP32 distances from the wrapping difference of 32-bit
// box fractions (exact minimum image), FP32 powers each from its own reduced
// phase, FP32 sums per chunk added in FP64.

#include "common.hpp"
#include "kernels.hpp"

namespace tdvmc {

namespace {

constexpr int kObsThreads = 256;

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// src/Utils.cpp:266-281: round half away from zero, an exact +-L/2 nudged inwards
__device__ __forceinline__ double nic_coord_ref(double x, double L, double Lr)
{
    const int k = (int)(x * Lr + ((x >= 0.0) ? 0.5 : -0.5));
    double res = x - k * L;
    if (res == L / 2.0)
        res -= 1e-10;
    else if (res == -L / 2.0)
        res += 1e-10;
    return res;
}

template <typename T, int D>
__global__ __launch_bounds__(kObsThreads) void obs_kernel(ObsArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = blockIdx.x;
    const int tid = threadIdx.x;
    const int N = a.N;
    const int nm = a.nmax + 1;
    double* pos = reinterpret_cast<double*>(smem);                           // [N][4] (x, y, z, 0)
    uint4* fix = reinterpret_cast<uint4*>(pos + 4 * (size_t)N);               // [N] box fractions (mixed form)
    // S(k) work split (below): R rounds of 256 vectors, Vr = V mod 256 vectors
    // left over, split into G items each
    const int R = a.V >> 8, Vr = a.V & 255;
    const int G = Vr ? kObsThreads / Vr : 1;
    double* part = reinterpret_cast<double*>(fix + (sizeof(T) == 4 ? N : 0));   // [256 R + Vr G][2] totals
    // phase tables [D][nm][chunk + 1]: the particle index runs fastest, and the
    // padded row keeps the rows of different powers in different banks
    const int CS = a.chunk + 1;
    Cx<T>* tab = reinterpret_cast<Cx<T>*>(part + 2 * (size_t)(256 * R + Vr * G));
    unsigned* cnt = reinterpret_cast<unsigned*>(tab + (size_t)D * nm * CS);   // [count]

    const double* Rw = a.R + w * (int64_t)N * D;
    for (int e = tid; e < N * 4; e += kObsThreads) {
        const int i = e >> 2, c = e & 3;
        pos[e] = c < D ? Rw[i * D + c] : 0.0;
    }
    if constexpr (sizeof(T) == 4) {
        // 32-bit fractions of the box: the wrapping difference is the minimum image
        for (int i = tid; i < N; i += kObsThreads) {
            unsigned u[3] = {0u, 0u, 0u};
#pragma unroll
            for (int c = 0; c < D; c++) {
                double f = Rw[i * D + c] * a.Lr;
                f -= __builtin_floor(f);
                u[c] = (unsigned)(unsigned long long)(f * 4294967296.0);
            }
            fix[i] = make_uint4(u[0], u[1], u[2], 0u);
        }
    }
    for (int b = tid; b < a.count; b += kObsThreads) cnt[b] = 0u;
    __syncthreads();

    // ------------------------------------------------------------ g(r)
    // Work item (row i, segment): particle i stays in registers and meets the
    // partners j = i + 1 + o (mod N) for the segment's offsets o.  Every row
    // has M = (N-1)/2 offsets, rows i < N/2 one more (o = N/2 - 1, j = i + N/2)
    // when N is even, so every unordered pair occurs once.  Consecutive
    // threads hold consecutive rows and read consecutive partners.
    {
        const int M = (N - 1) / 2;
        const int even = (N % 2 == 0) ? 1 : 0;
        const int S = a.gr_segs;
        const int seg = (M + even + S - 1) / S;
        const float sc = (float)(a.L * (1.0 / 4294967296.0));
        for (int t = tid; t < N * S; t += kObsThreads) {
            const int i = t / S, sg = t - i * S;
            const int Mi = M + ((even && i < N / 2) ? 1 : 0);
            const int o0 = sg * seg, o1 = min(Mi, o0 + seg);
            int j = i + 1 + o0;
            if (j >= N) j -= N;
            if constexpr (sizeof(T) == 8) {
                // the reference's minimum image and sum of squares, without fused multiply-adds
#pragma clang fp contract(off)
                const double4 ri = *reinterpret_cast<const double4*>(pos + 4 * i);
                for (int o = o0; o < o1; o++) {
                    const double4 rj = *reinterpret_cast<const double4*>(pos + 4 * j);
                    double d = nic_coord_ref(ri.x - rj.x, a.L, a.Lr);
                    double s = d * d;
                    if constexpr (D > 1) {
                        d = nic_coord_ref(ri.y - rj.y, a.L, a.Lr);
                        s += d * d;
                    }
                    if constexpr (D > 2) {
                        d = nic_coord_ref(ri.z - rj.z, a.L, a.Lr);
                        s += d * d;
                    }
                    const double r = __builtin_sqrt(s);
                    if (r < a.gmax) {
                        const int bin = (int)__builtin_floor(r / a.spacing);
                        if (bin < a.count) atomicAdd(cnt + bin, 1u);
                    }
                    if (++j == N) j = 0;
                }
            } else {
                const uint4 ui = fix[i];
                const float gmax = (float)a.gmax, isp = (float)(1.0 / a.spacing);
                for (int o = o0; o < o1; o++) {
                    const uint4 uj = fix[j];
                    float d = (float)(int)(ui.x - uj.x);
                    float s = d * d;
                    if constexpr (D > 1) {
                        d = (float)(int)(ui.y - uj.y);
                        s += d * d;
                    }
                    if constexpr (D > 2) {
                        d = (float)(int)(ui.z - uj.z);
                        s += d * d;
                    }
                    const float r = __builtin_amdgcn_sqrtf(s) * sc;
                    if (r < gmax) {
                        const int bin = (int)(r * isp);
                        if (bin < a.count) atomicAdd(cnt + bin, 1u);
                    }
                    if (++j == N) j = 0;
                }
            }
        }
    }

    // ------------------------------------------------------------ S(k)
    // Vectors v < 256 R (R = V / 256) go one per thread and round over all
    // particles.  The Vr = V mod 256 left over are split by particle residue:
    // item (r, g), g < G = 256 / Vr, sums vector 256 R + r over the particles
    // i = g mod G, so every thread has at most one such item.  Per-vector and
    // per-item totals meet in LDS and are added per shell in a fixed order.
    constexpr int kRounds = 4;
    const bool has_item = tid < Vr * G;
    const int ir = has_item ? tid % Vr : 0, ig = has_item ? tid / Vr : 0;
    int nv[kRounds + 1][3];
    double accR[kRounds + 1], accI[kRounds + 1];
#pragma unroll
    for (int q = 0; q <= kRounds; q++) {
        const int v = q < kRounds ? tid + q * kObsThreads : 256 * R + ir;
        const bool ok = q < kRounds ? q < R : has_item;
#pragma unroll
        for (int c = 0; c < 3; c++) nv[q][c] = (ok && c < D) ? a.kvec[v * 4 + c] : 0;
        accR[q] = accI[q] = 0.0;
    }
    // this thread's table rows: power n[c] of axis c, all particles of the chunk
    auto rows = [&](const int* n, const Cx<T>** r) {
#pragma unroll
        for (int c = 0; c < D; c++) r[c] = tab + ((size_t)c * nm + n[c]) * CS;
    };
    auto term = [&](const Cx<T>* const* r, int il) {
        Cx<T> z = r[0][il];
        if constexpr (D > 1) z = cmul(z, r[1][il]);
        if constexpr (D > 2) z = cmul(z, r[2][il]);
        return z;
    };
    for (int c0 = 0; c0 < N; c0 += a.chunk) {
        const int nc = min(a.chunk, N - c0);
        __syncthreads();   // previous chunk's table reads are done
        for (int e = tid; e < nc * D; e += kObsThreads) {
            const int il = e / D, c = e - il * D;
            double t = pos[(c0 + il) * 4 + c] * a.Lr;
            t -= __builtin_floor(t);
            Cx<T>* col = tab + (size_t)c * nm * CS + il;   // power m at col[m * CS]
            if constexpr (sizeof(T) == 8) {
                double sn, cs;
                sincospi(2.0 * t, &sn, &cs);
                const Cx<T> e1{cs, sn};
                Cx<T> pw{1.0, 0.0};
                for (int m = 0; m < nm; m++) {
                    col[(size_t)m * CS] = pw;
                    pw = cmul(pw, e1);
                }
            } else {
                for (int m = 0; m < nm; m++) {
                    double u = (double)m * t;
                    u -= __builtin_floor(u);
                    float sn, cs;
                    Prec<float>::sincos_small((float)u, &sn, &cs);
                    col[(size_t)m * CS] = Cx<T>{cs, sn};
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kRounds; q++) {
            if (q >= R) break;
            const Cx<T>* r[D];
            rows(nv[q], r);
            T sr = (T)0, si = (T)0;
#pragma unroll 4
            for (int il = 0; il < nc; il++) {
                const Cx<T> z = term(r, il);
                sr += z.re;
                si += z.im;
            }
            accR[q] += (double)sr;
            accI[q] += (double)si;
        }
        if (has_item) {
            const Cx<T>* r[D];
            rows(nv[kRounds], r);
            T sr = (T)0, si = (T)0;
            for (int il = (ig - c0 % G + G) % G; il < nc; il += G) {
                const Cx<T> z = term(r, il);
                sr += z.re;
                si += z.im;
            }
            accR[kRounds] += (double)sr;
            accI[kRounds] += (double)si;
        }
    }
#pragma unroll
    for (int q = 0; q < kRounds; q++) {
        if (q < R) {
            const int v = tid + q * kObsThreads;
            part[2 * v] = accR[q];
            part[2 * v + 1] = accI[q];
        }
    }
    double* rest = part + 2 * 256 * R;   // [Vr][G][2]
    if (has_item) {
        rest[2 * (ir * G + ig)] = accR[kRounds];
        rest[2 * (ir * G + ig) + 1] = accI[kRounds];
    }
    __syncthreads();
    for (int s = tid; s < a.nshell; s += kObsThreads) {
        double cs = 0.0, sn = 0.0;
        const int v0 = a.shell_off[s], v1 = a.shell_off[s + 1];
        for (int v = v0; v < v1; v++) {
            if (v < 256 * R) {
                cs += part[2 * v];
                sn += part[2 * v + 1];
            } else {
                const double* p = rest + 2 * (v - 256 * R) * G;
                double pc = 0.0, ps = 0.0;
                for (int g = 0; g < G; g++) {
                    pc += p[2 * g];
                    ps += p[2 * g + 1];
                }
                cs += pc;
                sn += ps;
            }
        }
        a.sk_w[w * a.nshell + s] = (cs * cs + sn * sn) / ((double)N * (double)(v1 - v0));
    }
    for (int b = tid; b < a.count; b += kObsThreads) {
        const unsigned c = cnt[b];
        if (c) atomicAdd(a.gr_tot + b, (unsigned long long)c);
        if (a.gr_w) a.gr_w[w * a.count + b] = (double)c * a.scale[b];
    }
}

// acc[s] += sum over walkers of sk_w[w][s], in a fixed order: one workgroup
// per shell, a strided sum per thread, then a tree.
__global__ __launch_bounds__(kObsThreads) void obs_reduce_kernel(const double* __restrict__ sk_w, int64_t W,
                                                                  int nshell, double* __restrict__ acc)
{
    __shared__ double red[kObsThreads];
    const int s = blockIdx.x, tid = threadIdx.x;
    double v = 0.0;
    for (int64_t w = tid; w < W; w += kObsThreads) v += sk_w[w * nshell + s];
    red[tid] = v;
    __syncthreads();
    for (int o = kObsThreads / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) acc[s] += red[0];
}

}  // namespace

size_t obs_lds_bytes(int N, int D, int V, int chunk, int nmax, int count, bool fp32)
{
    const size_t cx = fp32 ? 2 * sizeof(float) : 2 * sizeof(double);
    const int R = V >> 8, Vr = V & 255, G = Vr ? 256 / Vr : 0;
    return (size_t)N * 4 * sizeof(double) + (fp32 ? (size_t)N * 16 : 0) + 2 * (size_t)(256 * R + Vr * G) * sizeof(double) +
           (size_t)(chunk + 1) * D * (nmax + 1) * cx + (size_t)count * sizeof(unsigned);
}

template <typename T>
hipError_t launch_obs(const ObsArgs& a, int D, size_t lds, hipStream_t st)
{
    if (a.W <= 0) return hipSuccess;
    const dim3 grid((unsigned)a.W), block(kObsThreads);
    switch (D) {
    case 1: hipLaunchKernelGGL((obs_kernel<T, 1>), grid, block, lds, st, a); break;
    case 2: hipLaunchKernelGGL((obs_kernel<T, 2>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((obs_kernel<T, 3>), grid, block, lds, st, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_obs_reduce(const double* sk_w, int64_t W, int nshell, double* acc, hipStream_t st)
{
    if (W <= 0 || nshell <= 0) return hipSuccess;
    hipLaunchKernelGGL(obs_reduce_kernel, dim3(nshell), dim3(kObsThreads), 0, st, sk_w, W, nshell, acc);
    return hipGetLastError();
}

template hipError_t launch_obs<float>(const ObsArgs&, int, size_t, hipStream_t);
template hipError_t launch_obs<double>(const ObsArgs&, int, size_t, hipStream_t);

}  // namespace tdvmc
