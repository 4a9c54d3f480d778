// Krylov-basis kernels for the batched Lippmann-Schwinger GMRES: block dot
// products h[r][i] = sum_n conj(V[r][i][n]) w[r][n], block updates
// w[r] += s sum_i h[r][i] V[r][i] (with the squared norm of the result), and
// the row gather / scaled scatter between a basis slab and a batch, for r < k
// right-hand sides and i < nv basis vectors of length n.  A right-hand side
// may use fewer vectors (ncol[r] < nv: a request in the middle of its own
// restart cycle); the kernels then read only its first ncol[r] vectors and
// its coefficients beyond them are zero.  The reference has
// no solver (its examples use IterativeSolvers' gmres on the host,
// docs/src/usage.md:370-391); these are the reductions a library GEMM handles
// badly at this shape (two tiny dimensions and one of millions).
//
// Every reduction is two-stage: workgroups reduce fixed chunks of one
// right-hand side (FP64 accumulators, a fixed shuffle tree), and one wave per
// output sums the chunk partials in a fixed order.  A right-hand side's
// results therefore do not depend on how many share the launch.
#include <hip/hip_runtime.h>

#include "gla_fft.hpp"
#include "gla_krylov.hpp"

namespace gla {
namespace {

constexpr int KT = 256;          // threads per chunk workgroup
constexpr int KW = KT / 64;      // waves per chunk workgroup
constexpr int KG = 8;            // basis vectors per sweep over a chunk

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// elements per thread and sweep: every load of a sweep (the right-hand side
// and up to KG basis vectors at E elements) is issued before the first use,
// so a wave keeps (KG + 1) E loads in flight instead of one
template <typename R> struct KE { static constexpr int value = 2; };
template <> struct KE<double> { static constexpr int value = 1; };

// one sweep over a chunk for G basis vectors i0 .. i0 + G - 1 (G = 0: the
// norm of w alone).  A thread visits q = lo + t, lo + t + KT, ... in
// ascending order whatever E is; out-of-chunk elements load from a valid
// address and contribute zero.
template <typename R, int G>
__device__ __forceinline__ void dots_sweep(const cpx<R>* __restrict__ vp, int64_t vv, int i0,
                                           const cpx<R>* __restrict__ wp, int64_t lo, int64_t hi, bool first,
                                           double (&acc)[2 * KG + 1])
{
    constexpr int E = KE<R>::value;
    const cpx<R>* vg[G > 0 ? G : 1];
#pragma unroll
    for (int g = 0; g < G; ++g) vg[g] = vp + (int64_t)(i0 + g) * vv;
    for (int64_t q0 = lo + threadIdx.x; q0 < hi; q0 += (int64_t)KT * E) {
        int64_t qe[E];
        bool ok[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            ok[e] = q0 + (int64_t)e * KT < hi;
            qe[e] = ok[e] ? q0 + (int64_t)e * KT : q0;
        }
        cpx<R> b[E];
        cpx<R> a[G > 0 ? G : 1][E];
#pragma unroll
        for (int e = 0; e < E; ++e) b[e] = wp[qe[e]];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int e = 0; e < E; ++e) a[g][e] = vg[g][qe[e]];
        __builtin_amdgcn_sched_barrier(0);   // keep the sweep's loads ahead of the arithmetic
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (!ok[e]) b[e] = cpx<R>{R(0), R(0)};
            if (first) acc[2 * KG] += (double)b[e].x * b[e].x + (double)b[e].y * b[e].y;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                acc[2 * g] += (double)a[g][e].x * b[e].x + (double)a[g][e].y * b[e].y;      // conj(a) b
                acc[2 * g + 1] += (double)a[g][e].x * b[e].y - (double)a[g][e].y * b[e].x;
            }
        }
    }
}

// part layout: [r][slot][chunk]; slots 0..nv-1 the dots, slot nv the squared
// norm of w (when requested)
template <typename R>
__global__ __launch_bounds__(KT) void k_dots(const cpx<R>* __restrict__ V, int64_t vr, int64_t vv,
                                             const cpx<R>* __restrict__ w, int64_t wr, int64_t n, int nv,
                                             const int64_t* __restrict__ ncol, int64_t chunk, int nchunk, int nslot,
                                             int self, double* __restrict__ part)
{
    __shared__ double red[2 * KG + 1][KW];
    const int r = blockIdx.y, c = blockIdx.x;
    const int nvr = ncol ? (int)(ncol[r] < 0 ? 0 : ncol[r] < nv ? ncol[r] : nv) : nv;
    const int64_t lo = (int64_t)c * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    const cpx<R>* wp = w + r * wr;
    const cpx<R>* vp = V + r * vr;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double* pr = part + (int64_t)r * nslot * 2 * nchunk;
    // coefficients of the vectors this right-hand side does not use
    for (int t = threadIdx.x; t < 2 * (nv - nvr); t += KT) pr[((int64_t)nvr * 2 + t) * nchunk + c] = 0.0;
    for (int i0 = 0; i0 < nvr || (i0 == 0 && self); i0 += KG) {
        double acc[2 * KG + 1];
#pragma unroll
        for (int g = 0; g < 2 * KG + 1; ++g) acc[g] = 0.0;
        const bool first = i0 == 0 && self;
        switch (nvr - i0 < KG ? nvr - i0 : KG) {
        case 8: dots_sweep<R, 8>(vp, vv, i0, wp, lo, hi, first, acc); break;
        case 7: dots_sweep<R, 7>(vp, vv, i0, wp, lo, hi, first, acc); break;
        case 6: dots_sweep<R, 6>(vp, vv, i0, wp, lo, hi, first, acc); break;
        case 5: dots_sweep<R, 5>(vp, vv, i0, wp, lo, hi, first, acc); break;
        case 4: dots_sweep<R, 4>(vp, vv, i0, wp, lo, hi, first, acc); break;
        case 3: dots_sweep<R, 3>(vp, vv, i0, wp, lo, hi, first, acc); break;
        case 2: dots_sweep<R, 2>(vp, vv, i0, wp, lo, hi, first, acc); break;
        case 1: dots_sweep<R, 1>(vp, vv, i0, wp, lo, hi, first, acc); break;
        default: dots_sweep<R, 0>(vp, vv, i0, wp, lo, hi, first, acc); break;
        }
#pragma unroll
        for (int g = 0; g < 2 * KG + 1; ++g) {
            const double v = wave_sum(acc[g]);
            if (lane == 0) red[g][wv] = v;
        }
        __syncthreads();
        if (threadIdx.x < 2 * KG + 1) {
            const int g = threadIdx.x;
            double sum = 0.0;
#pragma unroll
            for (int u = 0; u < KW; ++u) sum += red[g][u];
            if (g < 2 * KG) {
                const int i = i0 + g / 2;
                if (i < nvr) pr[((int64_t)i * 2 + (g & 1)) * nchunk + c] = sum;
            } else if (first) {
                pr[((int64_t)nv * 2) * nchunk + c] = sum;
            }
        }
        __syncthreads();
    }
}

// basis vectors per step of the update loop (their loads issued together)
constexpr int KU = 4;

// w[r] += s sum_i h[r][i] V[r][i] over one chunk; the chunk's squared norm
// of the result goes to part[r][0][chunk] when norm is set.  Each element
// sums its nv terms in order i = 0, 1, ...; a thread visits its elements in
// ascending order (E at a time).
template <typename R>
__global__ __launch_bounds__(KT) void k_axpy(const cpx<R>* __restrict__ V, int64_t vr, int64_t vv,
                                             const cpx<double>* __restrict__ h, int64_t hr,
                                             cpx<R>* __restrict__ w, int64_t wr, int64_t n, int nv,
                                             const int64_t* __restrict__ ncol, double s, int64_t chunk, int nchunk,
                                             int norm, double* __restrict__ part)
{
    constexpr int E = KE<R>::value * 2;
    __shared__ double red[KW];
    const int r = blockIdx.y, c = blockIdx.x;
    if (ncol) nv = (int)(ncol[r] < 0 ? 0 : ncol[r] < nv ? ncol[r] : nv);
    const int64_t lo = (int64_t)c * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    const cpx<R>* vp = V + r * vr;
    const cpx<double>* hp = h + r * hr;
    cpx<R>* wp = w + r * wr;
    double nrm = 0.0;
    for (int64_t q0 = lo + threadIdx.x; q0 < hi; q0 += (int64_t)KT * E) {
        int64_t qe[E];
        bool ok[E];
        double ar[E], ai[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            ok[e] = q0 + (int64_t)e * KT < hi;
            qe[e] = ok[e] ? q0 + (int64_t)e * KT : q0;
            ar[e] = 0.0;
            ai[e] = 0.0;
        }
        cpx<R> o[E];
#pragma unroll
        for (int e = 0; e < E; ++e) o[e] = wp[qe[e]];
        int i = 0;
        for (; i + KU <= nv; i += KU) {
            cpx<R> a[KU][E];
#pragma unroll
            for (int u = 0; u < KU; ++u)
#pragma unroll
                for (int e = 0; e < E; ++e) a[u][e] = vp[(int64_t)(i + u) * vv + qe[e]];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < KU; ++u) {
                const cpx<double> cc = hp[i + u];
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    ar[e] += cc.x * a[u][e].x - cc.y * a[u][e].y;
                    ai[e] += cc.x * a[u][e].y + cc.y * a[u][e].x;
                }
            }
        }
        for (; i < nv; ++i) {
            cpx<R> a[E];
#pragma unroll
            for (int e = 0; e < E; ++e) a[e] = vp[(int64_t)i * vv + qe[e]];
            const cpx<double> cc = hp[i];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                ar[e] += cc.x * a[e].x - cc.y * a[e].y;
                ai[e] += cc.x * a[e].y + cc.y * a[e].x;
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (!ok[e]) continue;
            const cpx<R> res{R((double)o[e].x + s * ar[e]), R((double)o[e].y + s * ai[e])};
            wp[qe[e]] = res;
            nrm += (double)res.x * res.x + (double)res.y * res.y;
        }
    }
    if (!norm) return;   // uniform over the launch
    nrm = wave_sum(nrm);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
#pragma unroll
        for (int u = 0; u < KW; ++u) sum += red[u];
        part[(int64_t)r * nchunk + c] = sum;
    }
}

// out[j] = sum over chunks of part[j][chunk], one wave per output: lane l sums
// chunks l, l+64, ... in order, then a fixed shuffle tree
__global__ __launch_bounds__(64) void k_chunk_sum(const double* __restrict__ part, int nchunk, int64_t total,
                                                  double* __restrict__ out)
{
    const int64_t j = blockIdx.x;
    if (j >= total) return;
    double v = 0.0;
    for (int c = threadIdx.x; c < nchunk; c += 64) v += part[j * nchunk + c];
    v = wave_sum(v);
    if (threadIdx.x == 0) out[j] = v;
}

// out[r] = V[r][idx[r]] (row gather); with scale: V[r][idx[r]] = out[r] / scale[r]
template <typename R, bool SCATTER>
__global__ __launch_bounds__(256) void k_rows(cpx<R>* __restrict__ V, int64_t vr, int64_t vv,
                                              const int64_t* __restrict__ idx, cpx<R>* __restrict__ b,
                                              int64_t br, const double* __restrict__ scale, int64_t n)
{
    const int r = blockIdx.y;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    cpx<R>* row = V + r * vr + idx[r] * vv;
    if constexpr (SCATTER) {
        const double inv = 1.0 / scale[r];
        const cpx<R> v = b[r * br + q];
        row[q] = cpx<R>{R(v.x * inv), R(v.y * inv)};
    } else {
        b[r * br + q] = row[q];
    }
}

// the next basis vector: V[r][idx[r]] = w[r] / max(sqrt(w2[r]), 1e-30), and
// the same into stage[r] (the next matvec's input) when stage is not null
template <typename R>
__global__ __launch_bounds__(256) void k_next(cpx<R>* __restrict__ V, int64_t vr, int64_t vv,
                                              const int64_t* __restrict__ idx, const cpx<R>* __restrict__ w,
                                              int64_t wr, const double* __restrict__ w2, cpx<R>* __restrict__ stage,
                                              int64_t sr, int64_t n)
{
    const int r = blockIdx.y;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double nrm = sqrt(w2[r]);
    nrm = nrm > 1e-30 ? nrm : 1e-30;
    const double inv = 1.0 / nrm;
    const cpx<R> v = w[r * wr + q];
    const cpx<R> o{R(v.x * inv), R(v.y * inv)};
    V[r * vr + idx[r] * vv + q] = o;
    if (stage) stage[r * sr + q] = o;
}

}  // namespace

int64_t krylov_chunk(int64_t n)
{
    // about 1024 workgroups per right-hand side for long vectors, at least
    // 4096 elements per workgroup
    int64_t c = (n + 1023) / 1024;
    if (c < 4096) c = 4096;
    return c;
}

int64_t krylov_part_doubles(int64_t k, int64_t nv, int64_t n)
{
    // chunk partials of 2 (nv + 1) rows per right-hand side, then their sums
    const int64_t chunk = krylov_chunk(n);
    const int64_t nchunk = (n + chunk - 1) / chunk;
    return k * 2 * (nv + 1) * (nchunk + 1);
}

template <typename R>
hipError_t run_krylov_dots(int64_t k, int nv, int64_t n, const void* V, int64_t vr, int64_t vv, const void* w,
                           int64_t wr, void* part, void* h, double* w2, const int64_t* ncol, hipStream_t st)
{
    const int64_t chunk = krylov_chunk(n);
    const int nchunk = (int)((n + chunk - 1) / chunk);
    const int self = w2 != nullptr;
    const int nslot = nv + 1;
    hipLaunchKernelGGL(k_dots<R>, dim3(nchunk, (unsigned)k), dim3(KT), 0, st, (const cpx<R>*)V, vr, vv,
                       (const cpx<R>*)w, wr, n, nv, ncol, chunk, nchunk, nslot, self, (double*)part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // every (r, slot, re/im) row of part sums into h (re/im interleaved,
    // h[r][i]) -- the slot-nv rows (norms) into a scratch tail of part
    const int64_t rows = k * nslot * 2;
    double* sums = (double*)part + rows * nchunk;
    hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)rows), dim3(64), 0, st, (const double*)part, nchunk, rows, sums);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // unpack: sums[(r*nslot + i)*2 + {0,1}] -> h[r*nv + i], norms -> w2[r]
    if (nv > 0) {
        e = hipMemcpy2DAsync(h, (size_t)nv * 16, sums, (size_t)nslot * 16, (size_t)nv * 16, (size_t)k,
                             hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    if (self)
        e = hipMemcpy2DAsync(w2, sizeof(double), sums + (int64_t)nv * 2, (size_t)nslot * 16, sizeof(double),
                             (size_t)k, hipMemcpyDeviceToDevice, st);
    return e;
}

template <typename R>
hipError_t run_krylov_axpy(int64_t k, int nv, int64_t n, const void* V, int64_t vr, int64_t vv, const void* h,
                           int64_t hr, void* w, int64_t wr, double s, void* part, double* w2, const int64_t* ncol,
                           hipStream_t st)
{
    const int64_t chunk = krylov_chunk(n);
    const int nchunk = (int)((n + chunk - 1) / chunk);
    const int norm = w2 != nullptr;
    hipLaunchKernelGGL(k_axpy<R>, dim3(nchunk, (unsigned)k), dim3(KT), 0, st, (const cpx<R>*)V, vr, vv,
                       (const cpx<double>*)h, hr, (cpx<R>*)w, wr, n, nv, ncol, s, chunk, nchunk, norm,
                       (double*)part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !norm) return e;
    hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)k), dim3(64), 0, st, (const double*)part, nchunk, k, w2);
    return hipGetLastError();
}

template <typename R>
hipError_t run_krylov_rows(int64_t k, int64_t n, void* V, int64_t vr, int64_t vv, const int64_t* idx, void* b,
                           int64_t br, const double* scale, hipStream_t st)
{
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)k);
    if (scale)
        hipLaunchKernelGGL((k_rows<R, true>), grid, dim3(256), 0, st, (cpx<R>*)V, vr, vv, idx, (cpx<R>*)b, br,
                           scale, n);
    else
        hipLaunchKernelGGL((k_rows<R, false>), grid, dim3(256), 0, st, (cpx<R>*)V, vr, vv, idx, (cpx<R>*)b, br,
                           scale, n);
    return hipGetLastError();
}

template <typename R>
hipError_t run_krylov_next(int64_t k, int64_t n, void* V, int64_t vr, int64_t vv, const int64_t* idx, const void* w,
                           int64_t wr, const double* w2, void* stage, int64_t sr, hipStream_t st)
{
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)k);
    hipLaunchKernelGGL(k_next<R>, grid, dim3(256), 0, st, (cpx<R>*)V, vr, vv, idx, (const cpx<R>*)w, wr, w2,
                       (cpx<R>*)stage, sr, n);
    return hipGetLastError();
}

template hipError_t run_krylov_dots<float>(int64_t, int, int64_t, const void*, int64_t, int64_t, const void*,
                                           int64_t, void*, void*, double*, const int64_t*, hipStream_t);
template hipError_t run_krylov_dots<double>(int64_t, int, int64_t, const void*, int64_t, int64_t, const void*,
                                            int64_t, void*, void*, double*, const int64_t*, hipStream_t);
template hipError_t run_krylov_axpy<float>(int64_t, int, int64_t, const void*, int64_t, int64_t, const void*,
                                           int64_t, void*, int64_t, double, void*, double*, const int64_t*,
                                           hipStream_t);
template hipError_t run_krylov_axpy<double>(int64_t, int, int64_t, const void*, int64_t, int64_t, const void*,
                                            int64_t, void*, int64_t, double, void*, double*, const int64_t*,
                                            hipStream_t);
template hipError_t run_krylov_next<float>(int64_t, int64_t, void*, int64_t, int64_t, const int64_t*, const void*,
                                           int64_t, const double*, void*, int64_t, hipStream_t);
template hipError_t run_krylov_next<double>(int64_t, int64_t, void*, int64_t, int64_t, const int64_t*, const void*,
                                            int64_t, const double*, void*, int64_t, hipStream_t);
template hipError_t run_krylov_rows<float>(int64_t, int64_t, void*, int64_t, int64_t, const int64_t*, void*,
                                           int64_t, const double*, hipStream_t);
template hipError_t run_krylov_rows<double>(int64_t, int64_t, void*, int64_t, int64_t, const int64_t*, void*,
                                            int64_t, const double*, hipStream_t);

}  // namespace gla
