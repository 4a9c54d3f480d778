// Batched Krylov-basis reductions (gla_krylov.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gla {

// elements of one right-hand side reduced by one workgroup of the dot kernel
int64_t krylov_chunk(int64_t n);

// h[r*nv + i] = sum_q conj(V[r*vr + i*vv + q]) w[r*wr + q] (FP64 sums);
// part holds k * nv * ceil(n / krylov_chunk(n)) complex doubles
template <typename R>
hipError_t run_krylov_dots(int64_t k, int nv, int64_t n, const void* V, int64_t vr, int64_t vv, const void* w,
                           int64_t wr, void* part, void* h, hipStream_t st);

// w[r*wr + q] += s * sum_i h[r*hr + i] V[r*vr + i*vv + q]   (h complex double)
template <typename R>
hipError_t run_krylov_axpy(int64_t k, int nv, int64_t n, const void* V, int64_t vr, int64_t vv, const void* h,
                           int64_t hr, void* w, int64_t wr, double s, hipStream_t st);

}  // namespace gla
