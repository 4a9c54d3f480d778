// Krylov-basis kernels for the batched Lippmann-Schwinger GMRES: block dot
// products h[r][i] = sum_n conj(V[r][i][n]) w[r][n] and block updates
// w[r] += s sum_i h[r][i] V[r][i], for r < k right-hand sides and i < nv
// basis vectors of length n.  The reference has no solver (its examples use
// IterativeSolvers' gmres on the host, docs/src/usage.md:370-391); these are
// the reductions that a library GEMM handles badly at this shape (two tiny
// dimensions and one of millions).
//
// Dots: every workgroup reduces one chunk of one right-hand side for all nv
// vectors (w read once per vector from L2), accumulating in FP64; the chunk
// partials are summed in a fixed order by a second kernel, so h does not
// depend on how many right-hand sides share the launch.
#include <hip/hip_runtime.h>

#include "gla_fft.hpp"
#include "gla_krylov.hpp"

namespace gla {
namespace {

constexpr int KD_THREADS = 256;

template <typename R>
__global__ __launch_bounds__(KD_THREADS) void k_dots(const cpx<R>* __restrict__ V, int64_t vr, int64_t vv,
                                                     const cpx<R>* __restrict__ w, int64_t wr, int64_t n,
                                                     int nv, int64_t chunk, int nchunk,
                                                     cpx<double>* __restrict__ part)
{
    __shared__ double red[2][KD_THREADS / 64];
    const int r = blockIdx.y, c = blockIdx.x;
    const int64_t lo = (int64_t)c * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    const cpx<R>* wp = w + r * wr;
    const cpx<R>* vp = V + r * vr;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = 0; i < nv; ++i) {
        double ar = 0.0, ai = 0.0;
        for (int64_t q = lo + threadIdx.x; q < hi; q += KD_THREADS) {
            const cpx<R> a = vp[i * vv + q], b = wp[q];
            // conj(a) * b
            ar += (double)a.x * b.x + (double)a.y * b.y;
            ai += (double)a.x * b.y - (double)a.y * b.x;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            ar += __shfl_xor(ar, o);
            ai += __shfl_xor(ai, o);
        }
        if (lane == 0) {
            red[0][wv] = ar;
            red[1][wv] = ai;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int u = 0; u < KD_THREADS / 64; ++u) {
                sr += red[0][u];
                si += red[1][u];
            }
            part[((int64_t)r * nv + i) * nchunk + c] = cpx<double>{sr, si};
        }
        __syncthreads();
    }
}

__global__ void k_dots_sum(const cpx<double>* __restrict__ part, int nchunk, int64_t total, cpx<double>* __restrict__ h)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= total) return;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < nchunk; ++c) {
        sr += part[j * nchunk + c].x;
        si += part[j * nchunk + c].y;
    }
    h[j] = cpx<double>{sr, si};
}

template <typename R>
__global__ __launch_bounds__(256) void k_axpy(const cpx<R>* __restrict__ V, int64_t vr, int64_t vv,
                                              const cpx<double>* __restrict__ h, int64_t hr,
                                              cpx<R>* __restrict__ w, int64_t wr, int64_t n, int nv, double s)
{
    const int r = blockIdx.y;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const cpx<R>* vp = V + r * vr + q;
    const cpx<double>* hp = h + r * hr;
    double ar = 0.0, ai = 0.0;
    for (int i = 0; i < nv; ++i) {
        const cpx<R> a = vp[i * vv];
        const cpx<double> c = hp[i];
        ar += c.x * a.x - c.y * a.y;
        ai += c.x * a.y + c.y * a.x;
    }
    cpx<R>& o = w[r * wr + q];
    o = cpx<R>{R((double)o.x + s * ar), R((double)o.y + s * ai)};
}

}  // namespace

int64_t krylov_chunk(int64_t n)
{
    // about 2048 workgroups per right-hand side for long vectors, at least
    // 4096 elements per workgroup
    int64_t c = (n + 2047) / 2048;
    if (c < 4096) c = 4096;
    return c;
}

template <typename R>
hipError_t run_krylov_dots(int64_t k, int nv, int64_t n, const void* V, int64_t vr, int64_t vv, const void* w,
                           int64_t wr, void* part, void* h, hipStream_t st)
{
    const int64_t chunk = krylov_chunk(n);
    const int nchunk = (int)((n + chunk - 1) / chunk);
    dim3 grid(nchunk, (unsigned)k);
    hipLaunchKernelGGL(k_dots<R>, grid, dim3(KD_THREADS), 0, st, (const cpx<R>*)V, vr, vv, (const cpx<R>*)w, wr, n,
                       nv, chunk, nchunk, (cpx<double>*)part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t total = k * nv;
    hipLaunchKernelGGL(k_dots_sum, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       (const cpx<double>*)part, nchunk, total, (cpx<double>*)h);
    return hipGetLastError();
}

template <typename R>
hipError_t run_krylov_axpy(int64_t k, int nv, int64_t n, const void* V, int64_t vr, int64_t vv, const void* h,
                           int64_t hr, void* w, int64_t wr, double s, hipStream_t st)
{
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)k);
    hipLaunchKernelGGL(k_axpy<R>, grid, dim3(256), 0, st, (const cpx<R>*)V, vr, vv, (const cpx<double>*)h, hr,
                       (cpx<R>*)w, wr, n, nv, s);
    return hipGetLastError();
}

template hipError_t run_krylov_dots<float>(int64_t, int, int64_t, const void*, int64_t, int64_t, const void*,
                                           int64_t, void*, void*, hipStream_t);
template hipError_t run_krylov_dots<double>(int64_t, int, int64_t, const void*, int64_t, int64_t, const void*,
                                            int64_t, void*, void*, hipStream_t);
template hipError_t run_krylov_axpy<float>(int64_t, int, int64_t, const void*, int64_t, int64_t, const void*,
                                           int64_t, void*, int64_t, double, hipStream_t);
template hipError_t run_krylov_axpy<double>(int64_t, int, int64_t, const void*, int64_t, int64_t, const void*,
                                            int64_t, void*, int64_t, double, hipStream_t);

}  // namespace gla
