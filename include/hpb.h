/*
 * hpb.h -- C ABI of the MI355X (gfx950) compensated BLAS1 + Krylov library.
 *
 * The drop-in boundary for the compensated BLAS1 hot path of
 * evstigneevnm/high_precision_blas1.  The reference exposes that path as C++
 * templates (VectorOperations concept, gpu_reduction_ogita, csr::gpu_matrix,
 * the lin_solvers); this header is the flat layer those templates sit on here.
 * Every entry point names the reference interface it replaces.
 *
 * Conventions
 *   - Vector arguments are DEVICE pointers (hipMalloc / hpb_malloc memory);
 *     sizes are element counts (size_t).  Arguments named out_* are HOST
 *     pointers unless the name says dev.
 *   - Every call returns int: 0 = success, > 0 a hipError_t value, < 0 an
 *     HPB_ERR_* code.  hpb_last_error() gives the message of the last failure
 *     on the calling thread (file:line like the reference's CUDA_SAFE_CALL,
 *     utils/cuda_safe_call.h:143-147).
 *   - A handle owns one HIP stream; all work of the handle runs on it in order.
 *     Calls returning host scalars synchronise that stream (the reference's
 *     reductions are host-synchronous too, gpu_reduction_ogita_impl.cuh:1273-1274);
 *     *_async calls write results to device memory and do not synchronise.
 *   - A handle is not thread-safe; use one handle per host thread / stream.
 *   - Compensated reductions report out[4] = { s, c, fl(s + c), sqrt(fl(s + c)) }:
 *     the unrounded Ogita-Rump-Oishi pair, the rounded result and (for
 *     nrm2-type use) its square root.  Plain reductions report c = 0.
 *   - Reductions are deterministic: for a given (n, pointer alignment, tuning)
 *     the summation order is fixed; no floating-point atomics are used.
 */
#ifndef HPB_H
#define HPB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPB_OK 0
#define HPB_ERR_INVALID (-1)     /* bad argument (null pointer, size, range) */
#define HPB_ERR_UNSUPPORTED (-2) /* valid request this build does not support */
#define HPB_ERR_NOT_CONVERGED (-3)
#define HPB_ERR_BREAKDOWN (-4)

typedef struct hpb_handle_s* hpb_handle_t;

/* ---- status ------------------------------------------------------------ */
/* utils/cuda_safe_call.h:143-147 (error text with file:line) */
const char* hpb_last_error(void);
const char* hpb_version(void);

/* ---- handle ------------------------------------------------------------ */
/* Replaces the per-object state of gpu_vector_operations (gpu_vector_operations.h:98-127)
 * and gpu_reduction_ogita's N-sized workspaces (gpu_reduction_ogita.h:44-52):
 * one handle owns an O(grid) workspace.  stream may be NULL (a new stream is
 * created and owned) or an existing hipStream_t. */
int hpb_create(hpb_handle_t* h, void* stream);
int hpb_destroy(hpb_handle_t h);
int hpb_get_stream(hpb_handle_t h, void** stream);
int hpb_synchronize(hpb_handle_t h);
int hpb_device_info(hpb_handle_t h, int* num_cu, int* device);

/* ---- memory (utils/cuda_support.h:111-209: device_allocate,
 *      host_2_device_cpy, device_2_host_cpy; device_deallocate) ------------ */
int hpb_malloc(void** dev_ptr, size_t bytes);
int hpb_free(void* dev_ptr);
int hpb_memcpy_h2d(hpb_handle_t h, void* dev_dst, const void* host_src, size_t bytes);
int hpb_memcpy_d2h(hpb_handle_t h, void* host_dst, const void* dev_src, size_t bytes);
int hpb_memcpy_d2d(hpb_handle_t h, void* dev_dst, const void* dev_src, size_t bytes);

/* ---- tuning / introspection -------------------------------------------- */
/* kind: 0 two-input reductions (dot), 1 elementwise, 2 one-input reductions
 * (sum, asum, nrm2), 3 fused multi-dot (block 256; unroll = basis columns per chunk,
 * 4 or 8; grid_max = blocks over all chunks); dtype: 0 f64, 1 f32.  The autotuner
 * (high_precision_blas1_amd/autotune.py) writes these; BLOCK_SIZE=1024 is a fixed
 * template parameter in the reference (gpu_reduction_ogita.h:36). */
#define HPB_TUNE_NONTEMPORAL 1 /* nontemporal loads (and, elementwise, stores) for the streamed tiles */
int hpb_set_tuning(hpb_handle_t h, int kind, int dtype, int block, int unroll, int grid_max, int flags);
int hpb_get_tuning(hpb_handle_t h, int kind, int dtype, int* block, int* unroll, int* grid_max, int* flags);
/* Launch shape a reduction of n elements would use (binary: two-input dot, else
 * one-input sum/asum/nrm2): cfg = {block, unroll, grid,
 * vector width (1 if the pointers are not 16-byte aligned)}. */
int hpb_reduce_config(hpb_handle_t h, int dtype, size_t n, int aligned16, int binary, int* cfg4);

/* ---- timing (HIP events on the handle's stream) ------------------------- */
int hpb_timer_start(hpb_handle_t h);
int hpb_timer_stop(hpb_handle_t h, float* out_ms);

/* ---- compensated (Ogita-Rump-Oishi) reductions -------------------------- */
/* Dot2: gpu_reduction_ogita::dot (gpu_reduction_ogita.h:85-89,
 *   gpu_reduction_ogita_impl.cuh:471-676,1227-1282) and
 *   gpu_vector_operations::scalar_prod in high precision (gpu_vector_operations.h:257-269). */
int hpb_dot2_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out4);
int hpb_dot2_f32(hpb_handle_t h, size_t n, const float* x, const float* y, float* out4);
int hpb_dot2_f64_async(hpb_handle_t h, size_t n, const double* x, const double* y, double* dev_out4);
int hpb_dot2_f32_async(hpb_handle_t h, size_t n, const float* x, const float* y, float* dev_out4);
/* Sum2: gpu_reduction_ogita::sum (gpu_reduction_ogita.h:74-78, impl.cuh:248-468,1169-1223). */
int hpb_sum2_f64(hpb_handle_t h, size_t n, const double* x, double* out4);
int hpb_sum2_f32(hpb_handle_t h, size_t n, const float* x, float* out4);
/* Compensated sum of |x_i|: gpu_reduction_ogita::asum (gpu_reduction_ogita.h:79-84).
 * Computes the true sum of absolute values (the reference's real-T kernel sums x). */
int hpb_asum2_f64(hpb_handle_t h, size_t n, const double* x, double* out4);
int hpb_asum2_f32(hpb_handle_t h, size_t n, const float* x, float* out4);
/* Compensated 2-norm: gpu_reduction_ogita::norm = sqrt(Dot2(x,x)) (gpu_reduction_ogita.h:91-96),
 * reading x once; out4[3] = sqrt(fl(s + c)).  Safe against overflow and underflow
 * like the reference's standard-precision norm (cuBLAS nrm2, gpu_vector_operations.h:
 * 315-327): while M = max |x_i| lies in [2^-400, 2^400] (FP64) / [2^-30, 2^30] (FP32)
 * the squares are summed unscaled; otherwise a second pass sums (x_i 2^k)^2 with
 * the power of two that puts M in [1/2, 1) and out4 = { s 2^-2k, c 2^-2k,
 * fl(s + c) 2^-2k, sqrt(fl(s + c)) 2^-k } (the first three are the sum of squares,
 * which may itself overflow or underflow; out4[3] is the norm).  An inf entry gives
 * inf, a NaN entry NaN.  The blocking call adds the second pass only when needed;
 * the _async call always queues it (it returns at once when not needed).  The
 * plain hpb_nrm2_* follow the same rule. */
int hpb_nrm2c_f64(hpb_handle_t h, size_t n, const double* x, double* out4);
int hpb_nrm2c_f32(hpb_handle_t h, size_t n, const float* x, float* out4);
int hpb_nrm2c_f64_async(hpb_handle_t h, size_t n, const double* x, double* dev_out4);
int hpb_nrm2c_f32_async(hpb_handle_t h, size_t n, const float* x, float* dev_out4);

/* ---- complex dot products (interleaved re, im; n complex elements) -------- */
/* x^H y = sum conj(x_i) y_i: the complex instantiations of gpu_reduction_ogita::dot
 * (gpu_reduction_ogita_impl_functions.cuh:126-195: four TwoProds, a TwoSum per
 * component, error terms (e1 + e2) + t).  out6 = { s_re, c_re, s_im, c_im,
 * fl(s_re + c_re), fl(s_im + c_im) }.  The plain form accumulates with fma.
 * The compensated 2-norm of a complex vector is hpb_nrm2c_* over its 2n reals. */
int hpb_zdot2_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out6);
int hpb_zdot2_f32(hpb_handle_t h, size_t n, const float* x, const float* y, float* out6);
int hpb_zdot_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out6);
int hpb_zdot_f32(hpb_handle_t h, size_t n, const float* x, const float* y, float* out6);
int hpb_zdot2_f64_async(hpb_handle_t h, size_t n, const double* x, const double* y, double* dev_out6);

/* ---- plain (uncompensated) reductions ----------------------------------- */
/* gpu_reduction::{dot,sum,asum} (gpu_reduction.h:71-90) and the cuBLAS
 * standard-precision calls of gpu_vector_operations (cublas_wrap.h:525-591).
 * dot accumulates s = fma(x_i, y_i, s). */
int hpb_dot_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out4);
int hpb_dot_f32(hpb_handle_t h, size_t n, const float* x, const float* y, float* out4);
int hpb_dot_f64_async(hpb_handle_t h, size_t n, const double* x, const double* y, double* dev_out4);
int hpb_dot_f32_async(hpb_handle_t h, size_t n, const float* x, const float* y, float* dev_out4);
int hpb_sum_f64(hpb_handle_t h, size_t n, const double* x, double* out4);
int hpb_sum_f32(hpb_handle_t h, size_t n, const float* x, float* out4);
int hpb_asum_f64(hpb_handle_t h, size_t n, const double* x, double* out4);
int hpb_asum_f32(hpb_handle_t h, size_t n, const float* x, float* out4);
int hpb_nrm2_f64(hpb_handle_t h, size_t n, const double* x, double* out4);
int hpb_nrm2_f32(hpb_handle_t h, size_t n, const float* x, float* out4);
int hpb_nrm2_f64_async(hpb_handle_t h, size_t n, const double* x, double* dev_out4);
int hpb_nrm2_f32_async(hpb_handle_t h, size_t n, const float* x, float* dev_out4);
/* min and max: gpu_reduction::min_max (gpu_reduction.h:61-68, gpu_reduction_impl.cuh:
 * 82-264).  NaN entries are skipped (as the reference's comparisons skip them);
 * -0 counts below +0.  out2 = { min, max }; { +inf, -inf } for n = 0 (the
 * reference starts from the integer 0x7f800000 converted to T, 2139095040, rather
 * than infinity).  Blocking. */
int hpb_minmax_f64(hpb_handle_t h, size_t n, const double* x, double* out2);
int hpb_minmax_f32(hpb_handle_t h, size_t n, const float* x, float* out2);

/* ---- elementwise (gpu_vector_operations_impl.cuh:226-528) ---------------- */
/* Each *_f64 has an *_f32 twin with float scalars and pointers. */
/* assign_scalar: x := a                                  (:226-241) */
int hpb_fill_f64(hpb_handle_t h, size_t n, double a, double* x);
int hpb_fill_f32(hpb_handle_t h, size_t n, float a, float* x);
/* add_mul_scalar: x := mul_x*x + scalar (one fma)          (:243-256) */
int hpb_add_mul_scalar_f64(hpb_handle_t h, size_t n, double scalar, double mul_x, double* x);
int hpb_add_mul_scalar_f32(hpb_handle_t h, size_t n, float scalar, float mul_x, float* x);
/* scale: x := a*x (cublasDscal, gpu_vector_operations.h:433-436) */
int hpb_scal_f64(hpb_handle_t h, size_t n, double a, double* x);
int hpb_scal_f32(hpb_handle_t h, size_t n, float a, float* x);
/* assign: y := x (cublasDcopy, gpu_vector_operations.h:442-446) */
int hpb_copy_f64(hpb_handle_t h, size_t n, const double* x, double* y);
int hpb_copy_f32(hpb_handle_t h, size_t n, const float* x, float* y);
/* swap: x <-> y (cublasDswap, gpu_vector_operations.h:448-451) */
int hpb_swap_f64(hpb_handle_t h, size_t n, double* x, double* y);
int hpb_swap_f32(hpb_handle_t h, size_t n, float* x, float* y);
/* make_abs_copy / make_abs: y := |x| (x == y allowed)     (:305-332) */
int hpb_abs_f64(hpb_handle_t h, size_t n, const double* x, double* y);
int hpb_abs_f32(hpb_handle_t h, size_t n, const float* x, float* y);
/* assign_mul(a,x,y): y := a*x                              (:258-272) */
int hpb_assign_mul_f64(hpb_handle_t h, size_t n, double a, const double* x, double* y);
int hpb_assign_mul_f32(hpb_handle_t h, size_t n, float a, const float* x, float* y);
/* add_mul(a,x,y): y := y + a*x, one fma (cublasDaxpy, gpu_vector_operations.h:459-462) */
int hpb_axpy_f64(hpb_handle_t h, size_t n, double a, const double* x, double* y);
int hpb_axpy_f32(hpb_handle_t h, size_t n, float a, const float* x, float* y);
/* assign_mul(a,x,b,y,z): z := a*x + b*y = fma(a, x, b*y)   (:334-349);
 * add_mul(a,x,b,y) is the case z == y                     (:352-365) */
int hpb_axpby_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double* z);
int hpb_axpby_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float* z);
/* add_mul(a,x,b,y,c,z): z := a*x + b*y + c*z = fma(c, z, fma(a, x, b*y))  (:368-383) */
int hpb_axpbypcz_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y,
                     double c, double* z);
int hpb_axpbypcz_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y,
                     float c, float* z);
/* add_mul(a,x,b,y,c,w,d,z): z := a*x + b*y + c*w + d*z     (:451-466) */
int hpb_axpbypcwpdz_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y,
                        double c, const double* w, double d, double* z);
int hpb_axpbypcwpdz_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y,
                        float c, const float* w, float d, float* z);
/* assign_mul(a,x,b,y,c,v,d,w,z): z := a*x + b*y + c*v + d*w (:468-484) */
int hpb_assign_mul4_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y,
                        double c, const double* v, double d, const double* w, double* z);
int hpb_assign_mul4_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y,
                        float c, const float* v, float d, const float* w, float* z);
/* mul_pointwise(a,x,b,y,z): z := (a*x)*(b*y); mul_pointwise(x,b,y): x := x*(b*y)  (:386-416) */
int hpb_mul_pointwise_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double* z);
int hpb_mul_pointwise_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float* z);
int hpb_mul_pointwise_inplace_f64(hpb_handle_t h, size_t n, double* x, double b, const double* y);
int hpb_mul_pointwise_inplace_f32(hpb_handle_t h, size_t n, float* x, float b, const float* y);
/* div_pointwise(a,x,b,y,z): z := (a*x)/(b*y); div_pointwise(x,b,y): x := x/(b*y)  (:419-449) */
int hpb_div_pointwise_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double* z);
int hpb_div_pointwise_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float* z);
int hpb_div_pointwise_inplace_f64(hpb_handle_t h, size_t n, double* x, double b, const double* y);
int hpb_div_pointwise_inplace_f32(hpb_handle_t h, size_t n, float* x, float b, const float* y);

/* ---- compensated elementwise (new: no reference counterpart; the nearest is
 *      add_mul, gpu_vector_operations.h:459-470, impl.cuh:352-383) ---------- */
/* z := a*x + b*y from the error-free split a x + b y = s + t + e1 + e2 (TwoProd,
 * TwoProd, TwoSum); z = s + ((e1 + e2) + t), where the correction term is rounded
 * twice before the final addition: z is faithful (one of the two floating-point
 * numbers around the exact value), not always correctly rounded.  z == y or
 * z == x allowed. */
int hpb_axpbyc_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double* z);
int hpb_axpbyc_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float* z);

/* ---- misc vector ops ----------------------------------------------------- */
/* check_is_valid_number (impl.cuh:151-224): out_valid = 1 iff all x_i finite. */
int hpb_valid_f64(hpb_handle_t h, size_t n, const double* x, int* out_valid);
int hpb_valid_f32(hpb_handle_t h, size_t n, const float* x, int* out_valid);
/* set_value_at_point / get_value_at_point (impl.cuh:486-528) */
int hpb_set_value_f64(hpb_handle_t h, size_t n, size_t at, double val, double* x);
int hpb_set_value_f32(hpb_handle_t h, size_t n, size_t at, float val, float* x);
int hpb_get_value_f64(hpb_handle_t h, size_t n, size_t at, const double* x, double* out_val);
int hpb_get_value_f32(hpb_handle_t h, size_t n, size_t at, const float* x, float* out_val);
/* assign_random(x, a, b) (gpu_vector_operations.h:496-523): x_i = a + (b-a) u_i,
 * u_i in (0, 1] from Philox4x32-10 with key = seed, counter = (i, stream_id);
 * seeded and reproducible (the reference seeds cuRAND from std::random_device). */
int hpb_random_uniform_f64(hpb_handle_t h, size_t n, uint64_t seed, uint32_t stream_id,
                           double a, double b, double* x);
int hpb_random_uniform_f32(hpb_handle_t h, size_t n, uint64_t seed, uint32_t stream_id,
                           float a, float b, float* x);
/* Element conversions (generate_vector_pair_helper_impl.cuh:110-120 and the
 * FP32/FP64 boundary of the refined solver). */
int hpb_convert_f32_to_f64(hpb_handle_t h, size_t n, const float* x, double* y);
int hpb_convert_f64_to_f32(hpb_handle_t h, size_t n, const double* x, float* y);

/* ---- fused multi-vector kernels (new: the reference's Gram-Schmidt is one
 *      dot + one axpy per basis vector, gmres.h:426-452) ------------------- */
/* out4k[4j..4j+3] = {s, c, fl(s+c), fl(s+c)} of V_j . w for j < k <= 64, reading w
 * once per chunk of 8 columns; V is a HOST array of k device pointers. */
int hpb_multidot_f64(hpb_handle_t h, int k, const double* const* V, size_t n, const double* w, int compensated,
                     double* out4k);
int hpb_multidot_f32(hpb_handle_t h, int k, const float* const* V, size_t n, const float* w, int compensated,
                     float* out4k);
int hpb_multidot_f64_async(hpb_handle_t h, int k, const double* const* V, size_t n, const double* w, int compensated,
                           double* dev_out4k);
int hpb_multidot_f32_async(hpb_handle_t h, int k, const float* const* V, size_t n, const float* w, int compensated,
                           float* dev_out4k);
/* No counterpart in the reference: hpb_multidot_f64_async for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No handle: the work goes on `stream` (a hipStream_t, may be
 * null for the default stream), with the launch shape of a new handle's default
 * tuning, and the per-block partials go to a file-scope device array of the library
 * (256 KiB per device, 16384 (s, c) pairs, looked up once per device: at hpb_create,
 * or at the first call on that device) instead of a handle's workspace; nothing is
 * allocated.  Same results as hpb_multidot_f64_async with default tuning on a device
 * where the array's cap does not bind (any k on MI355X).  Rejects k outside 1..64
 * (HPB_ERR_INVALID).  All calls on one device share the array: calls that overlap
 * on one device race. */
int hpb_multidot_f64_noalloc(void* stream, int k, const double* const* V, size_t n, const double* w,
                             int compensated, double* dev_out4k);
/* y := y + sum_{j<k} a_j V_j in one pass (k <= 64); compensated evaluates every
 * element as a Dot2 over its k + 1 terms.  a is a host array (the _dev form
 * reads the k coefficients from device memory). */
int hpb_multiaxpy_f64(hpb_handle_t h, int k, const double* a, const double* const* V, size_t n, double* y,
                      int compensated);
int hpb_multiaxpy_f32(hpb_handle_t h, int k, const float* a, const float* const* V, size_t n, float* y,
                      int compensated);
int hpb_multiaxpy_dev_f64(hpb_handle_t h, int k, const double* dev_a, const double* const* V, size_t n, double* y,
                          int compensated);
int hpb_multiaxpy_dev_f32(hpb_handle_t h, int k, const float* dev_a, const float* const* V, size_t n, float* y,
                          int compensated);

/* ---- CSR sparse matrices (csr::matrix, source/common/csr/matrix.h:91-158:
 *      zero-based int32 row_ptr[rows+1], col_ind[nnz], values[nnz]) ---------- */
/* y := alpha*A*x + beta*y; beta == 0 does not read y.  Replaces
 * csr::gpu_matrix::axpy (gpu_matrix.h:322-375, cusparseSpMV) as called by
 * csr::linear_operator::apply (csr/linear_operator.h:39-42, alpha 1, beta 0).
 * Plain rows accumulate s = fma(a_ij, x_j, s). */
int hpb_csr_spmv_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                     const double* val, double alpha, const double* x, double beta, double* y);
int hpb_csr_spmv_f32(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                     const float* val, float alpha, const float* x, float beta, float* y);
/* Compensated-row SpMV (new, no reference counterpart): every row product is an
 * Ogita-Rump-Oishi Dot2, so y_i is accurate to u|row_i.x| + gamma^2 |row_i|.|x|. */
int hpb_csr_spmvc_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                      const double* val, double alpha, const double* x, double beta, double* y);
int hpb_csr_spmvc_f32(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                      const float* val, float alpha, const float* x, float beta, float* y);
/* r := b - A*x in one pass (the solvers' calc_residual_, gmres.h:122-130, fused);
 * the c variant enters b_i as the first term of a compensated row sum. */
int hpb_csr_residual_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                         const double* val, const double* b, const double* x, double* r);
int hpb_csr_residualc_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                          const double* val, const double* b, const double* x, double* r);
/* Same as hpb_csr_spmv(c)_f64 with an explicit number of lanes per row
 * (1, 2, ..., 64; 0 = chosen from the mean row length) -- for tuning and tests. */
int hpb_csr_spmv_lpr_f64(hpb_handle_t h, int lanes_per_row, int compensated, int rows, int cols, int64_t nnz,
                         const int* row_ptr, const int* col_ind, const double* val, double alpha, const double* x,
                         double beta, double* y);

/* ---- MatrixMarket files (host side; csr::matrix_market_reader,
 *      source/common/csr/matrix_market_reader.h:246-430) ------------------- */
/* Reads a coordinate real/integer/pattern file (general, symmetric or
 * skew-symmetric) into zero-based CSR host arrays allocated by the library
 * (release each with hpb_mm_free).  Entries are stably sorted by (row, column);
 * duplicates are kept in file order, as in the reference. */
int hpb_mm_read_csr(const char* path, int* out_rows, int* out_cols, int64_t* out_nnz, int** out_row_ptr,
                    int** out_col_ind, double** out_val);
void hpb_mm_free(void* host_ptr);
/* Writes host CSR arrays as "coordinate real general" with 17 significant digits. */
int hpb_mm_write_csr(const char* path, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                     const double* val);

/* ---- structured test matrices (the benchmark problems; the reference reads
 *      MatrixMarket files instead, csr/matrix_market_reader.h:246) ---------- */
/* rows and non-zeros of an nx x ny x nz 7-point (nz == 1: 5-point) stencil with
 * Dirichlet boundary neighbours dropped. */
int hpb_stencil_nnz(int nx, int ny, int nz, int64_t* out_rows, int64_t* out_nnz);
/* Constant-coefficient stencil, rows idx = i + nx*(j + ny*k), columns ascending;
 * coef7 (host) = { z-, y-, x-, diagonal, x+, y+, z+ }.  Poisson: {-1,-1,-1,6,-1,-1,-1}. */
int hpb_csr_stencil7_f64(hpb_handle_t h, int nx, int ny, int nz, const double* coef7, int* row_ptr, int* col_ind,
                         double* val);
/* kappa_c = exp(sigma g_c), g_c ~ N(0,1) by Box-Muller from Philox (key seed, counter c). */
int hpb_lognormal_field_f64(hpb_handle_t h, int64_t ncell, double sigma, uint64_t seed, double* kappa);
/* 5-point diffusion on m x m cells: face coefficient = harmonic mean 2ab/(a+b)
 * of the two cells (the cell's own kappa on boundary faces); off-diagonal -face,
 * diagonal = sum of the four faces ((s + w) + e) + n. */
int hpb_csr_diffusion2d_f64(hpb_handle_t h, int m, const double* kappa, int* row_ptr, int* col_ind, double* val);

/* ---- CSR transpose ---------------------------------------------------------- */
/* B = A^T as CSR (cols x rows) into caller arrays t_row_ptr[cols + 1],
 * t_col_ind[nnz], t_val[nnz], for the transposed products of csr::gpu_matrix::axpy
 * (transpose = true, common/csr/gpu_matrix.h:322-333, cuSPARSE
 * CUSPARSE_OPERATION_TRANSPOSE) and the transposed triangular solves: the ordinary
 * row kernels then run on B.  Deterministic (rows of B sorted by column, then by
 * value bits for duplicate entries); no floating-point atomics.  A setup call:
 * allocates a temporary of about 4 (cols + 1) bytes and synchronises the stream.
 * Requires nnz < 2^31 and column indices in [0, cols) (else HPB_ERR_INVALID). */
int hpb_csr_transpose_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                          const double* val, int* t_row_ptr, int* t_col_ind, double* t_val);
int hpb_csr_transpose_f32(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                          const float* val, int* t_row_ptr, int* t_col_ind, float* t_val);

/* ---- sparse triangular solves and ILU(0) ---------------------------------- */
/* csr::gpu_matrix::triangular_solve_lower / _upper (common/csr/gpu_matrix.h:142-320,
 * cuSPARSE csrsv2 with an analysis and a cudaMalloc per call).  The analysis is done
 * once per pattern (levels of the dependency graph; the caller's row_ptr / col_ind
 * must stay valid and unchanged while the analysis is used; columns sorted and
 * unique in each row):
 *   uplo 0: L x = alpha b with the strictly lower triangle and a unit diagonal
 *           (the stored diagonal is ignored, CUSPARSE_DIAG_TYPE_UNIT);
 *   uplo 1: U x = alpha b with the upper triangle and its stored diagonal (every
 *           row must store one);
 *   uplo 2: lower triangle with its stored diagonal; uplo 3: strictly upper triangle
 *           with a unit diagonal.  On the transpose B = A^T (hpb_csr_transpose_*)
 *           these are the reference's transposed solves L^T x = alpha b (uplo 3 on B)
 *           and U^T x = alpha b (uplo 2 on B), gpu_matrix.h:142-154,232-244.
 * One kernel per level; compensated = 1 evaluates each row's sum with Dot2. x may
 * alias b.  The analysis copies the pattern to the host through a pinned buffer that
 * the handle keeps, (rows + 1 + nnz) * 4 bytes at the largest pattern analysed, until
 * hpb_destroy. */
typedef struct hpb_trsv_s* hpb_trsv_t;
int hpb_csr_trsv_analysis(hpb_handle_t h, int uplo, int rows, int64_t nnz, const int* row_ptr, const int* col_ind,
                          hpb_trsv_t* out);
int hpb_csr_trsv_levels(hpb_trsv_t t, int* nlevels);
int hpb_trsv_destroy(hpb_trsv_t t);
int hpb_csr_trsv_f64(hpb_handle_t h, hpb_trsv_t t, const double* val, double alpha, const double* b, double* x,
                     int compensated);
int hpb_csr_trsv_f32(hpb_handle_t h, hpb_trsv_t t, const float* val, float alpha, const float* b, float* x,
                     int compensated);
/* ILU(0) in place on the pattern of A, using the uplo 0 analysis of that pattern:
 * afterwards the strictly lower part holds L (unit diagonal implied) and the rest U,
 * with (L U)_ij = a_ij on the pattern.  The reference's gpu_ilu0_preconditioner
 * (common/csr/gpu_ilu0_preconditioner.h:62-72) does not factor; this one does.
 * Every row must store its diagonal entry. */
int hpb_csr_ilu0_f64(hpb_handle_t h, hpb_trsv_t lower, double* val);
int hpb_csr_ilu0_f32(hpb_handle_t h, hpb_trsv_t lower, float* val);

/* ---- Krylov solvers on CSR matrices ---------------------------------------- */
/* The solver templates (gmres.h:65-495, bicgstabl.h:68-275 of the reference's
 * numerical_algos/lin_solvers, restated in high_precision_blas1_amd/cpp) over
 * hip_vector_operations and csr::hip_matrix, unpreconditioned, x0 = x on entry. */
#define HPB_SOLVER_GMRES 0
#define HPB_SOLVER_BICGSTABL 1
#define HPB_SOLVER_BICGSTAB 2 /* plain BiCGStab, numerical_algos/lin_solvers/bicgstab.h:134-253 */
#define HPB_PRECOND_NONE 0
#define HPB_PRECOND_ILU0 1
typedef struct hpb_solver_params {
    int solver;              /* HPB_SOLVER_GMRES, HPB_SOLVER_BICGSTABL or HPB_SOLVER_BICGSTAB */
    int restart;             /* GMRES(m): m, 1 <= m < 256 (gmres.h:355-371) */
    int basis;               /* BiCGStab(L): L, 1..30 (bicgstabl.h:142-147) */
    int max_iters, min_iters;
    double rel_tol, abs_tol; /* monitor: ||r|| <= abs_tol + rel_tol ||b|| (default_monitor.h:159) */
    int high_precision;      /* 1: compensated Dot2 / nrm2 in the solver (use_high_precision) */
    int compensated_spmv;    /* 1: compensated-row SpMV inside the solver */
    int gram_schmidt;        /* GMRES: 0 MGS (the reference's), 1 CGS2 with fused kernels, 2 CGS with
                                the second pass only when the first shrank w by more than 1/sqrt(2)
                                (Daniel-Gragg-Kaufman-Stewart selective reorthogonalisation) */
    int early_exit;          /* GMRES: leave a cycle when the Givens estimate meets the tolerance */
    int reorthogonalization; /* GMRES MGS: one extra correction per basis vector */
    int resid_recalc_freq;   /* BiCGStab(L) / BiCGStab: recompute b - Ax every k cycles / steps (0 never;
                                the library default is 1, the reference's 0).  With 0 the monitor, and so
                                `converged`, sees the recursive residual, as in the reference; the
                                result's true_rel_resid is the check on b - Ax. */
    int save_history;        /* keep (iteration, ||r||/||b||) pairs */
    int verbose;             /* 0 silent, 1 errors, 2 per-check info */
    int device_resident;     /* 1: iterations run on the device (scalars in device memory, one captured
                                hipGraph per GMRES restart cycle / BiCGStab(L) cycle, no host round trip
                                per dot); applies to CGS2 GMRES and BiCGStab(L <= 8), other settings run
                                the host-driven templates.  Same recurrences and monitor rules. */
    int preconditioner;      /* HPB_PRECOND_NONE or HPB_PRECOND_ILU0 (left preconditioning with an
                                ILU(0) of A, csr/preconditioner.h; host-driven templates) */
    int use_precond_resid;   /* with a preconditioner: the monitor tests ||P r|| against ||P b||
                                (set_use_precond_resid, gmres.h:396-405); 0: unpreconditioned residual */
    int refinement;          /* hpb_csr_solve_ir_f64 with a device-resident BiCGStab(L) inner solver:
                                1 (default) reliable updates -- one FP32 iteration that is never
                                restarted; each time its recursive residual has dropped by the inner
                                rel_tol, x64 += x32 and r = b - A x is recomputed in FP64 and fed back;
                                a breakdown of the FP32 iteration also triggers that update, followed by a
                                restart from the true residual (at most 16 times);
                                0 restarted refinement (a fresh inner solve per sweep) */
} hpb_solver_params;
typedef struct hpb_solver_result {
    int converged;           /* monitor converged (true residual for GMRES restarts) */
    int iters;               /* iterations (inner iterations summed for refinement) */
    int outer_iters;         /* refinement sweeps (1 for a plain solve) */
    int breakdown;           /* BiCGStab(L): rho = 0 */
    int history_len;
    double resid_norm;       /* last residual norm seen by the monitor */
    double rhs_norm;
    double true_rel_resid;   /* ||b - A x|| / ||b||, compensated FP64, after the solve */
    double seconds;          /* wall time of the solve (stream synchronised at both ends) */
} hpb_solver_result;
void hpb_solver_params_default(hpb_solver_params* p);
/* FP64 solve.  history (host, optional) receives up to history_cap pairs. */
int hpb_csr_solve_f64(hpb_handle_t h, int rows, int64_t nnz, const int* row_ptr, const int* col_ind,
                      const double* val, const double* b, double* x, const hpb_solver_params* p,
                      hpb_solver_result* res, double* history, int history_cap);
/* Mixed precision (new): FP32 copies of the values and FP32 Krylov vectors with
 * compensated FP32 dots inside, FP64 iterative refinement outside:
 *   r = b - A x (compensated FP64 rows); stop when ||r|| <= rel_tol ||b||;
 *   solve A d = r/||r|| in FP32 with `inner` (its rel_tol is the inner tolerance);
 *   x += ||r|| d in FP64.
 * With a BiCGStab(L) inner solver and inner->refinement = 1 the FP32 iteration is
 * not restarted between updates (reliable updates); res->outer_iters then counts
 * the FP64 updates and inner->max_iters caps the FP32 iterations. */
int hpb_csr_solve_ir_f64(hpb_handle_t h, int rows, int64_t nnz, const int* row_ptr, const int* col_ind,
                         const double* val, const double* b, double* x, const hpb_solver_params* inner,
                         double rel_tol, int max_outer, hpb_solver_result* res);
/* helpers of the refinement: y32 := (float)(s x64); y64 := fma(a, (double)x32, y64) */
int hpb_scale_convert_f64_to_f32(hpb_handle_t h, size_t n, double s, const double* x, float* y);
int hpb_axpy_f32_to_f64(hpb_handle_t h, size_t n, double a, const float* x, double* y);

/* ---- batched solves (new: the reference solves one system at a time) -------- */
/* Many independent 2-D log-normal diffusion systems (side m_s, cells exp(sigma g),
 * Philox key seed + s * 0x9E3779B97F4A7C15, harmonic face means as in
 * hpb_csr_diffusion2d_f64) held as one block-diagonal CSR matrix, b = A x* with
 * x* uniform in [-1, 1] (rhs_seed).  Solved together by BiCGStab(2) with one
 * captured hipGraph per cycle (bicgstabl.h:151-274 recurrences, L = 2; the true
 * residual confirms convergence, and replaces the recursive one as often as
 * hpb_batch_set_resid_recalc_freq asks). */
typedef struct hpb_batch_s* hpb_batch_t;
int hpb_batch_create_diffusion2d(hpb_handle_t h, int nsys, const int* sides, double sigma, uint64_t seed,
                                 uint64_t rhs_seed, hpb_batch_t* out);
int hpb_batch_info(hpb_batch_t b, int64_t* rows, int64_t* nnz, int* ntiles);
/* Solve all systems to ||b_s - A_s x_s|| <= rel_tol ||b_s||.  use_graph: 1 replays
 * the captured cycle, reading the active count every check_every cycles; 0 launches
 * the same kernels directly; 2 runs every cycle in ONE persistent cooperative launch
 * (hipLaunchCooperativeKernel, as many 256-thread blocks as fit resident, at most 8
 * per CU) whose phases are separated by grid barriers and which stops by itself when
 * no system is active (check_every unused); the three modes give the same iterates
 * bit for bit.  On MI355X the graph replay is the fastest (the 512-system solve:
 * 18.6 s replayed, 21.2 s persistent).  Every kernel of a cycle runs on a fixed grid that
 * strides over the list of tiles of systems still active (rebuilt on the device at
 * the end of each cycle), so converged systems cost nothing.  Per-system outputs (host, nsys entries, each
 * optional): iterations, true relative residual, converged flag; plus wall seconds
 * and cycles run. */
int hpb_batch_solve_bicgstab2(hpb_batch_t b, double rel_tol, int max_iters, int use_graph, int check_every,
                              int* out_iters, double* out_true_rel_resid, int* out_converged, double* out_seconds,
                              int* out_cycles);
/* on != 0: Jacobi right preconditioning, A D^-1 (D x) = b with D = diag(A); the
 * residual tested is still b - A x of the original system. */
int hpb_batch_set_jacobi(hpb_batch_t b, int on);
/* The true residual b - A x replaces the recursive one in cycles freq, 2 freq, ...
 * (freq >= 1; 1: every cycle), as resid_recalc_freq does in the reference's
 * BiCGStab(L) (bicgstabl.h:247-248; 0, the default there and here: never).  In the
 * other cycles the convergence test uses the recursive residual of the MR update,
 * and a system that meets the tolerance gets its true residual computed (replacing
 * the recursive one) and tested in the same cycle: a system is only ever declared
 * converged on its true residual.  Saves one of the five SpMVs of such cycles. */
int hpb_batch_set_resid_recalc_freq(hpb_batch_t b, int freq);
/* verbose != 0: print the count of active systems to stderr at every check
 * (the role of the reference monitor's log output, default_monitor.h:204-214) */
int hpb_batch_set_verbose(hpb_batch_t b, int verbose);
/* host copies of the solution and of x* (rows entries each, optional) */
int hpb_batch_solution(hpb_batch_t b, double* host_x, double* host_xstar);
/* device pointers of the block-diagonal CSR and the right-hand side */
int hpb_batch_csr(hpb_batch_t b, const int** row_ptr, const int** col_ind, const double** val, const double** rhs);
int hpb_batch_destroy(hpb_batch_t b);

/* ---- exact dot product / sum -------------------------------------------- */
/* Replaces the exact oracles dot_product_gmp::dot_exact (dot_product_gmp.hpp:108-130,
 * host GMP) and its CUMP GPU form (dot_product_cump.hpp:134-151,
 * cump_blas_kernels.cuh:381-444), which return the 512-1024-bit value truncated to
 * one double.  Here the value is exact (fixed-point long accumulator, bit 0 =
 * 2^-1074) and returned as out2 = { hi, lo }, hi = round(exact) to nearest even,
 * lo = round(exact - hi).  Deterministic for any launch shape.
 * Domain: finite inputs (else HPB_ERR_INVALID); for FP64 dots every nonzero
 * product satisfies 2^-969 <= |x_i y_i| < 2^960, for sums |x_i| < 2^960 (else
 * HPB_ERR_UNSUPPORTED); n < 2^38.  The FP32 dot forms the products exactly in FP64. */
int hpb_exact_dot_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out2);
int hpb_exact_dot_f32(hpb_handle_t h, size_t n, const float* x, const float* y, double* out2);
int hpb_exact_sum_f64(hpb_handle_t h, size_t n, const double* x, double* out2);
/* dev_out3 = { hi, lo, flags } in device memory; flags 1 = non-finite input,
 * 2 = outside the exact range (hi, lo are NaN when flags != 0). */
int hpb_exact_dot_f64_async(hpb_handle_t h, size_t n, const double* x, const double* y, double* dev_out3);

/* ---- ill-conditioned pair generator -------------------------------------- */
/* Device-side, seeded GenDot pair (Ogita-Rump-Oishi 2005 Alg. 6.1) replacing
 * generate_vector_pair::generate (generate_vector_pair.hpp:119-182, host-driven,
 * std::random_device seeded, O(N_fix N) Dot2 calls).  Head n - n_tail entries:
 * x_i, y_i = m 2^e with e uniform in [0, log2(cond)/2]; tail: exponents descending
 * to 0, y chosen to cancel the running dot.  The exact dot is tracked in 256-bit
 * fixed point (LSB 2^-106): out_exact4 receives its two's complement limbs
 * (least significant first), out3 = { round(x.y), round(sum |x_i y_i|), kappa }.
 * Bit-identical to the host restatement for the same arguments.  Requires
 * 2 <= n, 1 <= n_tail < n, cond >= 1, log2(cond) + log2(n) <= 143. Blocking. */
int hpb_gendot_f64(hpb_handle_t h, size_t n, double log2cond, uint64_t seed, size_t n_tail, double* x, double* y,
                   uint64_t* out_exact4, double* out3);

#ifdef __cplusplus
}
#endif

#endif /* HPB_H */
