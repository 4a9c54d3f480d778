/*
 * gila_amd.h -- C ABI of the MI355X Green-function operator library
 * (libgila_amd.so).  Plain pointers and sizes only.
 *
 * Arrays use the reference's Julia layout: column-major, axis 1 fastest,
 * then axis 2, axis 3, field component (1..3), then right-hand side, with
 * interleaved (re, im) complex values that are layout-compatible with Julia
 * ComplexF32 / ComplexF64 and C float/double _Complex.
 *
 * Citations are to the reference repository GilaElectromagnetics.jl.
 *
 * Error convention: every call returns a gla_status; GLA_OK is 0.  The text
 * of the last error raised on the calling thread is returned by
 * gla_last_error().  One status per row of the reference's error table:
 *   GLA_ERR_SIZE        wrong input size           (src/glaAct.jl:45, ErrorException)
 *   GLA_ERR_NONFINITE   NaN/Inf in the Green data  (src/glaMemSup.jl:94,111,162)
 *   GLA_ERR_NO_DEVICE   no usable GPU              (src/glaMemSup.jl:23)
 *   GLA_ERR_OVERLAP     overlapping volumes        (src/glaGen.jl:78)
 *   GLA_ERR_SCALE       incompatible cell scales   (src/glaMem.jl:141,194,208)
 *   GLA_ERR_TYPE        element-type mismatch      (src/glaLinAlg.jl:35, AssertionError)
 *   GLA_ERR_ARG         invalid argument / non-Complex set type (src/glaMemSup.jl:311,342)
 *   GLA_ERR_UNSUPPORTED a configuration this library does not implement
 *   GLA_ERR_HIP         a HIP runtime failure (out of memory, launch failure)
 */
#ifndef GILA_AMD_H
#define GILA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GLA_OK = 0,
    GLA_ERR_SIZE = 1,
    GLA_ERR_NONFINITE = 2,
    GLA_ERR_NO_DEVICE = 3,
    GLA_ERR_OVERLAP = 4,
    GLA_ERR_SCALE = 5,
    GLA_ERR_TYPE = 6,
    GLA_ERR_ARG = 7,
    GLA_ERR_UNSUPPORTED = 8,
    GLA_ERR_HIP = 9
} gla_status;

/* element types: the reference's setTyp */
typedef enum { GLA_C64 = 1 /* ComplexF32 */, GLA_C128 = 2 /* ComplexF64 */ } gla_dtype;

/* A cuboid volume, GlaVol (src/glaMem.jl:12-19, ctor :136-147): cell counts,
 * cell size and centre as exact rationals (numerator / denominator), in
 * wavelengths. */
typedef struct {
    int64_t cel[3];
    int64_t scl_num[3], scl_den[3];
    int64_t org_num[3], org_den[3];
} gla_vol;

/* Operator options, GlaKerOpt (src/glaMem.jl:62-70).  frq_phz and int_ord
 * only affect Green-tensor generation; adj_mod is carried as a flag exactly
 * as the reference carries it (the matvec never reads it,
 * src/glaLinAlg.jl:16-29); device selects the HIP device. */
typedef struct {
    double frq_phz_re, frq_phz_im;
    int32_t int_ord;
    int32_t adj_mod;
    int32_t dtype;      /* gla_dtype: the operator's setTyp */
    int32_t device;     /* HIP device ordinal */
} gla_opts;

typedef struct gla_opr_s* gla_opr_t;

/* Geometry of an operator as the matvec sees it: the partition sizes
 * src_cel / trg_cel and divisors src_div / trg_div of GlaExtInf
 * (src/glaMem.jl:190-216; one partition per side when the cell scales are
 * equal), branch lengths B = (S+T)/2 of a partition pair (src/glaAct.jl:68),
 * the stored Green extents U (src/glaMemSup.jl:121-134) and the per-axis
 * symmetric-storage flags.  A whole volume has cel * div cells per axis; the
 * Green data holds one block per (source, target) partition pair. */
typedef struct {
    int64_t src_cel[3], trg_cel[3];
    int64_t brn[3];
    int64_t fur_ext[3];
    int32_t sym[3];
    int32_t dtype;
    int32_t adj_mod;
    int32_t self_opr;
    int64_t dev_brn[3];   /* branch lengths actually transformed on the GPU */
    int64_t src_div[3], trg_div[3];
} gla_dims;

/* ---------------------------------------------------------------------- */

/* Text of the last error on the calling thread ("" if none). */
const char* gla_last_error(void);

/* Library / device check.  Returns GLA_ERR_NO_DEVICE when no HIP device is
 * usable (the reference's "CUDA not functional", src/glaMemSup.jl:22-24). */
gla_status gla_init(int32_t device);

/* Geometry that gla_opr_create_from_fur expects for (trg, src): the reference
 * derives it in GlaOprMem (src/glaMemSup.jl:16-156); dev_brn gives the lengths
 * the GPU will transform.  src == NULL means a
 * self operator.  An external pair whose cell counts sum to an odd number on
 * some axis is first regenerated with gla_vol_even_gen, as GlaOprMem does
 * (src/glaMemSup.jl:34-40).  Cell scales that differ by an integer factor
 * per axis give partitions (GlaExtInf, src/glaMem.jl:190-216): every volume
 * must then divide into whole partitions, the partition cells of source and
 * target must sum to an even number per axis (GLA_ERR_SIZE otherwise), and
 * scales that share no common grid fail with GLA_ERR_SCALE. */
gla_status gla_opr_dims_for(const gla_vol* trg, const gla_vol* src, gla_dims* out);

/* glaVolEveGen (src/glaMem.jl:152-176): a volume with an odd cell count on
 * some axis gets one more cell there, the cell (and grid) size shrinking so
 * that the extent is kept; an all-even volume is copied unchanged. */
gla_status gla_vol_even_gen(const gla_vol* in, gla_vol* out);

/* Operator from precomputed Fourier Green data: the egoFur= re-entry of
 * GlaOprMem (src/glaMemSup.jl:16-20, 54, 158-164).  fur holds the 8 branches
 * back to back, each (U1,U2,U3,6,Ps,Pt) column-major with components
 * [xx,yy,zz,xy,xz,yz] and Ps / Pt the source / target partition counts
 * (1 and 1 for equal cell scales); fur_dtype gives its element type; fur_len
 * is the number of complex elements (8*U1*U2*U3*6*Ps*Pt).  Non-finite values
 * fail with GLA_ERR_NONFINITE as the reference's check does.  Any branch
 * length B <= 512 is accepted: data at a length without a compiled transform
 * is re-expressed in FP64 at the next fast length (the circulant it encodes,
 * recovered by an inverse transform, is re-embedded on the longer grid and
 * transformed there), so gla_opr_get_dims then reports that length in
 * dev_brn and gla_opr_export_fur returns the data at it. */
gla_status gla_opr_create_from_fur(const gla_opts* opts, const gla_vol* trg, const gla_vol* src,
                                   const void* fur, int32_t fur_dtype, int64_t fur_len,
                                   gla_opr_t* out);

/* Operator from geometry: the full GlaOprMem constructor
 * (src/glaMemSup.jl:16-171) -- FP64 Green-tensor generation on the GPU
 * (genEgoSlf!/genEgoExt!, src/glaGen.jl:17-48), circulant embedding, FFT,
 * branch extraction and truncation, then rounding to opts->dtype. */
gla_status gla_opr_create(const gla_opts* opts, const gla_vol* trg, const gla_vol* src, gla_opr_t* out);

/* Copy out the Fourier Green data in the reference's layout (for caching, the
 * preload-file equivalent of examples/GilaOperators.jl:28-55). buf receives
 * 8*U1*U2*U3*6*Ps*Pt elements of the operator's dtype; len is its capacity. */
gla_status gla_opr_export_fur(gla_opr_t h, void* buf, int64_t len);

gla_status gla_opr_get_dims(gla_opr_t h, gla_dims* out);

/* The adjoint wrapper of the reference (src/glaLinAlg.jl:16-29): a new
 * handle that shares the Green data, with adj_mod flipped and target /
 * source volumes swapped.  The matvec is unchanged, as in the reference. */
gla_status gla_opr_adjoint(gla_opr_t h, gla_opr_t* out);

gla_status gla_opr_destroy(gla_opr_t h);

/* Matvec, egoOpr! (src/glaAct.jl:33-49) on device memory.
 * x: (S1,S2,S3,3,nrhs), y: (T1,T2,T3,3,nrhs) over the whole source and
 * target volumes (partition cells times divisors; a partitioned operator
 * gathers the source partitions, sums the leaves of every partition pair and
 * scatters the target partitions, genPrt / mrgPrt, src/glaActSup.jl:4-18,
 * 939-953, src/glaAct.jl:158-164), both device pointers of the
 * operator's dtype, caller-owned; x is not modified.  nx / ny give the
 * number of complex elements of each buffer (a mismatch is GLA_ERR_SIZE, the
 * reference's size check).  The call is asynchronous on `stream` (a
 * hipStream_t, NULL for the default stream); the reference blocks, so its
 * binding synchronizes the stream before returning.  Work buffers belong to
 * the handle: calls that overlap on one handle must be serialized by the
 * caller (use separate handles for concurrent streams). */
gla_status gla_opr_apply(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs,
                         void* stream);

/* Size of the fixed per-device scratch of gla_opr_apply_noalloc: 256 MiB
 * (for instance one right-hand side of a 96^3 ComplexF32 or a 64^3
 * ComplexF64 self operator). */
#define GLA_NOALLOC_SCRATCH_BYTES (256LL << 20)

/* The matvec of gla_opr_apply without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * two work planes of the launch sequence live in one fixed array of
 * GLA_NOALLOC_SCRATCH_BYTES per device (held from the first gla_init on that
 * device until the process ends) instead of the handle's workspace; a call
 * whose planes do not fit fails with GLA_ERR_SIZE before launching anything.
 * Arguments, layout and results are those of gla_opr_apply.  Every call on
 * one GPU shares that array: calls that overlap on the GPU (different
 * streams, or different handles on one device) race and must be serialized
 * by the caller. */
gla_status gla_opr_apply_noalloc(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs,
                                 void* stream);

/* mul!(y, G, x, alpha, beta) (src/glaLinAlg.jl:61-69): y = alpha*G*x + beta*y,
 * fused into the last pass.  beta == 0 never reads y. */
gla_status gla_opr_apply_axpby(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs,
                               double alpha_re, double alpha_im, double beta_re, double beta_im,
                               void* stream);

/* Lippmann-Schwinger operator y = x - chi .* (G x) of the reference's
 * examples (LippmannSchwinger, examples/GilaOperators.jl:57-96), with the
 * medium multiply and the subtraction fused into the last kernel.  Self
 * operators only.  chi: one complex value per cell, (T1,T2,T3), in the
 * operator's dtype; chi_stride = 0 shares it across the nrhs right-hand
 * sides, otherwise it is the element stride between per-RHS media; nchi is
 * the number of elements of the chi buffer.  x and y must not alias.  No
 * counterpart entry point in the reference (its LippmannSchwinger is user
 * code over egoOpr!). */
gla_status gla_ls_apply(gla_opr_t h, const void* chi, int64_t nchi, int64_t chi_stride, const void* x, int64_t nx,
                        void* y, int64_t ny, int64_t nrhs, void* stream);

/* G * x on host memory (src/glaLinAlg.jl:33-43 with the host-to-device copy
 * of :37-40): copies in, applies, copies out and synchronizes.  The device
 * copies of x and y are kept by the handle (grown to the largest call) until
 * gla_opr_release. */
gla_status gla_opr_apply_host(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs);

/* Per-kernel timing of the matvec with HIP events on `stream`: after one
 * full matvec, each of the 5 passes (0 x-forward, 1 y-forward, 2 z fused
 * forward/multiply/inverse, 3 y-inverse, 4 x-inverse) is launched `reps`
 * times back to back between two events; ms[p] receives its average launch
 * duration and bytes[p] its algorithmic HBM bytes (compulsory reads and
 * writes of that pass; pass 2 includes the stored Green data).  y holds a
 * valid matvec on return.  No counterpart in the reference (which has no
 * profiling); for benchmarks. */
gla_status gla_opr_profile(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs, void* stream,
                           int32_t reps, double* ms, int64_t* bytes);

/* The same figures measured inside the real sequence: `reps` matvecs issued
 * exactly as gla_opr_apply issues them, with an event at every pass
 * boundary; ms[p] is pass p's average duration between its boundary events,
 * bytes[p] its algorithmic HBM bytes and *total_ms the average duration of
 * one whole matvec over the run.  Synchronizes; y holds a valid matvec on
 * return.  No counterpart in the reference; for benchmarks. */
gla_status gla_opr_profile_seq(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs,
                               void* stream, int32_t reps, double* ms, int64_t* bytes, double* total_ms);

/* The operator's target and source volumes, and its options, as created. */
gla_status gla_opr_get_vols(gla_opr_t h, gla_vol* trg, gla_vol* src);
gla_status gla_opr_get_opts(gla_opr_t h, gla_opts* opts);

/* Green-data files: replace the reference's preload cache, which stores
 * GlaOprMem with Julia serialization (examples/GilaOperators.jl:28-55,
 * test/runtests.jl:38-49) -- a format only Julia can read.  Little-endian:
 * a 512-byte header (magic "GILAFUR1", version, element type, options,
 * volumes, branch / device lengths, stored extents, symmetric axes, element
 * count, FNV-1a-64 of the payload), then the eight branches as
 * gla_opr_export_fur returns them.  Loading checks the layout against the
 * volumes, the size and the checksum; dtype 0 keeps the stored type,
 * GLA_C64 / GLA_C128 converts. */
gla_status gla_opr_save(gla_opr_t h, const char* path);
gla_status gla_opr_load(const char* path, int32_t dtype, int32_t device, gla_opr_t* out);

/* Grow the handle's work buffers to nrhs right-hand sides now.  An apply on
 * a stream under HIP graph capture must not allocate, so it fails with
 * GLA_ERR_ARG instead of growing them: reserve before capturing.  No
 * counterpart in the reference (its device path allocates per call). */
gla_status gla_opr_reserve(gla_opr_t h, int64_t nrhs);

/* ---- Recorded batches of operators ------------------------------------
 * No counterpart in the reference (it applies one operator at a time).  A
 * fixed batch -- operator ops[i] applied to x[i] -> y[i] with nrhs[i] right-
 * hand sides (nrhs may be NULL for 1 each), device pointers -- is recorded
 * once into a HIP graph whose work is dealt over `lanes` parallel branches,
 * so independent operators too small to fill the GPU run side by side; each
 * gla_batch_launch replays the whole batch with one graph launch on `stream`
 * (asynchronous).  The operators' work buffers are grown before recording and
 * pinned while the batch exists (an apply that would grow them fails); the
 * operators and vectors must outlive the batch, and an operator listed twice
 * runs its entries one after the other (they share its work buffers).  All
 * operators must be on one device. */
typedef struct gla_batch_s* gla_batch_t;
gla_status gla_batch_create(const gla_opr_t* ops, const void* const* x, void* const* y, const int64_t* nrhs,
                            int32_t n, int32_t lanes, gla_batch_t* out);
gla_status gla_batch_launch(gla_batch_t b, void* stream);
gla_status gla_batch_destroy(gla_batch_t b);

/* Bytes of device work memory the handle holds for nrhs right-hand sides. */
int64_t gla_opr_workspace_bytes(gla_opr_t h, int64_t nrhs);

/* ---- Batched Krylov-basis reductions (GMRES building blocks) ------------
 * No counterpart in the reference: it has no solver, and its examples pass
 * the operator to IterativeSolvers' gmres / bicgstabl on the host
 * (examples/scattering_guide.jl:65-68, docs/src/usage.md:370-391).  For
 * callers running batched Gram-Schmidt on device-resident bases.  k right-
 * hand sides, nv basis vectors of n complex elements each (dtype GLA_C64 or
 * GLA_C128); strides in elements.  h is complex double, h[r*nv + i]
 * (gla_krylov_dots) or h[r*h_rhs_stride + i] (gla_krylov_axpy).  Work goes to
 * `stream` on the current device. */

/* Bytes of the partial-sum workspace of gla_krylov_dots for (k, nv, n); with
 * nv = 0 it is the size a norm in gla_krylov_axpy needs. */
int64_t gla_krylov_dots_workspace(int64_t k, int64_t nv, int64_t n);

/* h[r][i] = sum_q conj(V[r][i][q]) w[r][q], and w2[r] = ||w[r]||^2 when w2 is
 * not NULL (device array of k doubles); FP64 sums in a fixed order (the
 * result for one right-hand side does not depend on k).  ncol (device int64,
 * k entries, or NULL for nv each): right-hand side r uses only its first
 * ncol[r] vectors, and h[r][i] = 0 for i >= ncol[r]. */
gla_status gla_krylov_dots(int32_t dtype, int64_t k, int64_t nv, int64_t n, const void* V, int64_t v_rhs_stride,
                           int64_t v_vec_stride, const void* w, int64_t w_rhs_stride, void* part,
                           int64_t part_bytes, void* h, double* w2, const int64_t* ncol, void* stream);

/* w[r][q] += s * sum_i h[r][i] V[r][i][q] over i < ncol[r] (i < nv when ncol
 * is NULL); then w2[r] = ||w[r]||^2 when w2 is not NULL (part is then
 * required, part may be NULL otherwise). */
gla_status gla_krylov_axpy(int32_t dtype, int64_t k, int64_t nv, int64_t n, const void* V, int64_t v_rhs_stride,
                           int64_t v_vec_stride, const void* h, int64_t h_rhs_stride, void* w,
                           int64_t w_rhs_stride, double s, void* part, int64_t part_bytes, double* w2,
                           const int64_t* ncol, void* stream);

/* Row moves between a basis slab and a batch, idx[r] (device int64) the
 * vector of right-hand side r: scale == NULL gathers b[r] = V[r][idx[r]];
 * otherwise scatters V[r][idx[r]] = b[r] / scale[r] (device doubles). */
gla_status gla_krylov_rows(int32_t dtype, int64_t k, int64_t n, void* V, int64_t v_rhs_stride, int64_t v_vec_stride,
                           const int64_t* idx, void* b, int64_t b_rhs_stride, const double* scale, void* stream);

/* Host routine (no device work): the Givens update of k restarted-GMRES
 * slots after one Arnoldi step.  Complex values are interleaved (re, im)
 * doubles in row-major host arrays: hcol (k, ldh) the new Hessenberg columns
 * (entries 0..js[s]), hn (k) the subdiagonal norms, cs (k, m) real and
 * sn (k, m) complex the slots' rotations, H (k, m+1, m) the rotated
 * Hessenberg matrices and g (k, m+1) the rotated residual vectors.  Slot s
 * at step j = js[s] applies rotations 0..j-1 to its column, forms rotation j
 * (zeroing the subdiagonal), stores the column into H[s][0..j+1][j] and
 * rotates g[s][j], g[s][j+1].  The reference has no solver; this is the
 * least-squares update of the GMRES that its usage guide calls
 * (docs/src/usage.md:370-391). */
gla_status gla_krylov_givens(int64_t k, int64_t m, const int64_t* js, const double* hcol, int64_t ldh,
                             const double* hn, double* cs, double* sn, double* H, double* g);

/* The next Arnoldi vector from the device-resident squared norm:
 * V[r][idx[r]] = w[r] / max(sqrt(w2[r]), 1e-30), written also to stage[r]
 * (the next matvec's input) when stage is not NULL.  Saves the host round
 * trip of the norm and the gather of the next step. */
gla_status gla_krylov_next(int32_t dtype, int64_t k, int64_t n, void* V, int64_t v_rhs_stride, int64_t v_vec_stride,
                           const int64_t* idx, const void* w, int64_t w_rhs_stride, const double* w2, void* stage,
                           int64_t stage_rhs_stride, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GILA_AMD_H */
