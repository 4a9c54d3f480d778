// Batched Krylov-basis reductions (gla_krylov.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gla {

// elements of one right-hand side reduced by one workgroup
int64_t krylov_chunk(int64_t n);

// doubles of the partial-sum workspace for k right-hand sides and nv vectors
int64_t krylov_part_doubles(int64_t k, int64_t nv, int64_t n);

// h[r*nv + i] = sum_q conj(V[r*vr + i*vv + q]) w[r*wr + q] (complex double,
// FP64 sums); w2[r] = ||w[r]||^2 when w2 is not null.  With ncol, right-hand
// side r uses its first ncol[r] vectors (h zero beyond)
template <typename R>
hipError_t run_krylov_dots(int64_t k, int nv, int64_t n, const void* V, int64_t vr, int64_t vv, const void* w,
                           int64_t wr, void* part, void* h, double* w2, const int64_t* ncol, hipStream_t st);

// w[r*wr + q] += s * sum_i h[r*hr + i] V[r*vr + i*vv + q] (h complex double,
// i < ncol[r] with ncol); w2[r] = ||w[r]||^2 afterwards when w2 is not null
template <typename R>
hipError_t run_krylov_axpy(int64_t k, int nv, int64_t n, const void* V, int64_t vr, int64_t vv, const void* h,
                           int64_t hr, void* w, int64_t wr, double s, void* part, double* w2, const int64_t* ncol,
                           hipStream_t st);

// V[r*vr + idx[r]*vv + q] = w[r*wr + q] / max(sqrt(w2[r]), 1e-30), also into
// stage[r*sr + q] when stage is not null
template <typename R>
hipError_t run_krylov_next(int64_t k, int64_t n, void* V, int64_t vr, int64_t vv, const int64_t* idx, const void* w,
                           int64_t wr, const double* w2, void* stage, int64_t sr, hipStream_t st);

// scale == null: b[r*br + q] = V[r*vr + idx[r]*vv + q]; otherwise
// V[r*vr + idx[r]*vv + q] = b[r*br + q] / scale[r]
template <typename R>
hipError_t run_krylov_rows(int64_t k, int64_t n, void* V, int64_t vr, int64_t vv, const int64_t* idx, void* b,
                           int64_t br, const double* scale, hipStream_t st);

}  // namespace gla
