// Additional observables of every walker: the pair distribution g(r) and the
// structure factor S(k) of BosonsBulk::CalculateAdditionalSystemProperties
// (src/PhysicalSystems/BosonsBulk.cpp:474-520).
//
// One 256-thread workgroup per walker, coordinates staged in LDS as FP64.
//   g(r): every unordered pair once (pair p -> particle i = p / M and partner
//   i + 1 + p mod M, M = (N-1)/2, plus the N/2 pairs at offset N/2 for even
//   N); the pair's bin is counted in an LDS u32 histogram, and the counts
//   reach a u64 total per bin with one global atomic each (exact and
//   order-free).  The value of a bin is count x (D/(N-1)) / scaling[bin]
//   (ObservableVsOnGridWithScaling::AddToHistogram, :45-50).
//   S(k): the shells' wave vectors are integer n with non-negative
//   components, k = 2 pi n / L, so exp(i k.r) = prod_a e_a^{n_a} with
//   e_a = exp(2 pi i x_a / L).  Per chunk of particles the powers e_a^m,
//   m = 0..nmax, are tabulated in LDS; a thread then forms the products for
//   its wave vectors (D-1 complex multiplies per particle and vector) and
//   adds them up.  One thread per shell sums its vectors' totals in order:
//   S = (C^2 + S^2) / (N n_k).
// FP64 form: distances by the reference's minimum image (src/Utils.cpp:266-281)
// and FP64 division for the bin, powers by complex recurrence from one
// sincospi.  Mixed form: FP32 distances in box units (d - rint(d)), FP32
// powers each from its own reduced phase, FP32 sums per chunk added in FP64.

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
    double* pos = reinterpret_cast<double*>(smem);                           // [N][D]
    double* part = pos + (size_t)N * D;                                        // [V][2] vector totals
    Cx<T>* tab = reinterpret_cast<Cx<T>*>(part + 2 * (size_t)a.V);            // [chunk][D][nm]
    unsigned* cnt = reinterpret_cast<unsigned*>(tab + (size_t)a.chunk * D * nm);   // [count]

    const double* Rw = a.R + w * (int64_t)N * D;
    for (int e = tid; e < N * D; e += kObsThreads) pos[e] = Rw[e];
    for (int b = tid; b < a.count; b += kObsThreads) cnt[b] = 0u;
    __syncthreads();

    // ------------------------------------------------------------ g(r)
    {
        const int M = (N - 1) / 2;
        const int nrot = N * M;
        const int ptot = nrot + ((N % 2 == 0) ? N / 2 : 0);
        const float invM = M > 0 ? 1.0f / (float)M : 0.0f;
        for (int p = tid; p < ptot; p += kObsThreads) {
            int i, j;
            if (p < nrot) {
                i = (int)(((float)p + 0.5f) * invM);
                if (i * M > p) i--;
                else if ((i + 1) * M <= p) i++;
                j = i + 1 + (p - i * M);
                if (j >= N) j -= N;
            } else {
                i = p - nrot;
                j = i + N / 2;
            }
            int bin = -1;
            if constexpr (sizeof(T) == 8) {
                // the reference's sum of squares, without fused multiply-adds
#pragma clang fp contract(off)
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < D; c++) {
                    const double d = nic_coord_ref(pos[i * D + c] - pos[j * D + c], a.L, a.Lr);
                    s += d * d;
                }
                const double r = __builtin_sqrt(s);
                if (r < a.gmax) bin = (int)__builtin_floor(r / a.spacing);
            } else {
                float s = 0.0f;
#pragma unroll
                for (int c = 0; c < D; c++) {
                    float d = (float)((pos[i * D + c] - pos[j * D + c]) * a.Lr);
                    d -= __builtin_rintf(d);
                    s += d * d;
                }
                const float r = __builtin_amdgcn_sqrtf(s) * (float)a.L;
                if (r < (float)a.gmax) bin = (int)(r * (float)(1.0 / a.spacing));
            }
            if (bin >= 0 && bin < a.count) atomicAdd(cnt + bin, 1u);
        }
    }

    // ------------------------------------------------------------ S(k)
    // this thread's wave vectors v = tid + q * 256, q < vpt
    constexpr int kVPT = 4;
    int nv[kVPT][3];
    double accR[kVPT], accI[kVPT];
#pragma unroll
    for (int q = 0; q < kVPT; q++) {
        const int v = tid + q * kObsThreads;
        const bool ok = v < a.V;
#pragma unroll
        for (int c = 0; c < 3; c++) nv[q][c] = (ok && c < D) ? a.kvec[v * 4 + c] : 0;
        accR[q] = accI[q] = 0.0;
    }
    for (int c0 = 0; c0 < N; c0 += a.chunk) {
        const int nc = min(a.chunk, N - c0);
        __syncthreads();   // previous chunk's table reads are done
        for (int e = tid; e < nc * D; e += kObsThreads) {
            const int il = e / D, c = e - il * D;
            double t = pos[(c0 + il) * D + c] * a.Lr;
            t -= __builtin_floor(t);
            Cx<T>* row = tab + ((size_t)il * D + c) * nm;
            if constexpr (sizeof(T) == 8) {
                double sn, cs;
                sincospi(2.0 * t, &sn, &cs);
                const Cx<T> e1{cs, sn};
                Cx<T> pw{1.0, 0.0};
                for (int m = 0; m < nm; m++) {
                    row[m] = pw;
                    pw = cmul(pw, e1);
                }
            } else {
                for (int m = 0; m < nm; m++) {
                    double u = (double)m * t;
                    u -= __builtin_floor(u);
                    float sn, cs;
                    Prec<float>::sincos_small((float)u, &sn, &cs);
                    row[m] = Cx<T>{cs, sn};
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kVPT; q++) {
            if (q >= a.vpt) break;
            T sr = (T)0, si = (T)0;
            for (int il = 0; il < nc; il++) {
                const Cx<T>* row = tab + (size_t)il * D * nm;
                Cx<T> z = row[nv[q][0]];
                if constexpr (D > 1) z = cmul(z, row[nm + nv[q][1]]);
                if constexpr (D > 2) z = cmul(z, row[2 * nm + nv[q][2]]);
                sr += z.re;
                si += z.im;
            }
            accR[q] += (double)sr;
            accI[q] += (double)si;
        }
    }
#pragma unroll
    for (int q = 0; q < kVPT; q++) {
        const int v = tid + q * kObsThreads;
        if (v < a.V) {
            part[2 * v] = accR[q];
            part[2 * v + 1] = accI[q];
        }
    }
    __syncthreads();
    for (int s = tid; s < a.nshell; s += kObsThreads) {
        double cs = 0.0, sn = 0.0;
        const int v0 = a.shell_off[s], v1 = a.shell_off[s + 1];
        for (int v = v0; v < v1; v++) {
            cs += part[2 * v];
            sn += part[2 * v + 1];
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
    return (size_t)N * D * sizeof(double) + 2 * (size_t)V * sizeof(double) + (size_t)chunk * D * (nmax + 1) * cx +
           (size_t)count * sizeof(unsigned);
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
