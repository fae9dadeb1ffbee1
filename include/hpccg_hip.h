/*
 * libhpccg_hip -- MI355X (gfx950) conjugate-gradient path of HPCCG-F90, C ABI.
 *
 * Drop-in boundary for the reference's Fortran procedures (RRZE-HPC/HPCCG-F90, MPI/ variant):
 * the Fortran modules in hpccg-f90_amd/fortran keep the reference's module names, procedure
 * names and argument lists and forward to the functions below through ISO_C_BINDING.
 *
 * Conventions
 *   - CSR arrays are passed exactly as generate_matrix builds them: 1-based rowindex[nrow+1]
 *     (rowindex[0] == 1), 1-based colindex[nnz], FP64 values[nnz]  (SparseMatrix.F90:44-64).
 *   - All counts are int32_t, as in the reference (default INTEGER).
 *   - Functions without a _dev suffix take HOST pointers (Fortran arrays) and are synchronous.
 *     Functions with a _dev suffix take DEVICE pointers (from hpccg_dev_malloc) and are
 *     synchronous too unless stated otherwise.
 *   - Every function returns 0 on success or a negative HPCCG_ERR_* code; hpccg_last_error()
 *     returns the message of the most recent failure on the calling thread.
 *   - One context = one device + one HIP stream; calls on one context are not re-entrant.
 */
#ifndef HPCCG_HIP_H
#define HPCCG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpccg_ctx hpccg_ctx;

enum {
    HPCCG_OK = 0,
    HPCCG_ERR_ARG = -1,       /* invalid argument (size mismatch, null pointer, bad mode)     */
    HPCCG_ERR_HIP = -2,       /* a HIP runtime call failed                                     */
    HPCCG_ERR_STATE = -3,     /* no matrix uploaded / wrong call order                          */
    HPCCG_ERR_NOMEM = -4,     /* device allocation failed                                       */
    HPCCG_WARN_BREAKDOWN = 1  /* cg finished but p.Ap was zero or non-finite in some iteration  */
};

/* Dot-product summation order used by the FP64 path.
 *   SEQUENTIAL: one left-to-right FP64 chain from 0.0, exactly the order of the reference's
 *               serial ddot/ddot1 (vector_operations.F90:54-59,86-91); cg then reproduces the
 *               reference's iterates bit for bit.  Slow (a serial chain on one wavefront).
 *   TREE:       fixed-shape wavefront/workgroup/grid tree (deterministic, run-to-run identical);
 *               the default fast path. */
enum { HPCCG_DOT_SEQUENTIAL = 0, HPCCG_DOT_TREE = 1 };

/* ---- lifecycle ------------------------------------------------------------------------- */
int hpccg_create(hpccg_ctx **ctx, int device);
int hpccg_destroy(hpccg_ctx *ctx);
const char *hpccg_last_error(void);
int hpccg_set_dot_mode(hpccg_ctx *ctx, int mode);
int hpccg_get_dot_mode(hpccg_ctx *ctx, int *mode);
/* Device facts used for launch shapes: CU count, XCD count, gfx arch number (e.g. 950). */
int hpccg_device_info(hpccg_ctx *ctx, int *num_cu, int *num_xcd, int *gfx);

/* ---- device memory (for callers that keep vectors resident) ----------------------------- */
int hpccg_dev_malloc(hpccg_ctx *ctx, void **ptr, int64_t bytes);
int hpccg_dev_free(hpccg_ctx *ctx, void *ptr);
int hpccg_memcpy_h2d(hpccg_ctx *ctx, void *dst, const void *src, int64_t bytes);
int hpccg_memcpy_d2h(hpccg_ctx *ctx, void *dst, const void *src, int64_t bytes);
int hpccg_memcpy_d2d(hpccg_ctx *ctx, void *dst, const void *src, int64_t bytes);

/* ---- matrix ------------------------------------------------------------------------------ */
/* Upload the CSR built by generate_matrix (SparseMatrix.F90:68-185); replaces any previous
 * matrix.  ncol == nrow on a single rank (SparseMatrix.F90:183).  The arrays are copied. */
int hpccg_matrix_upload(hpccg_ctx *ctx, int32_t nrow, int32_t ncol, const int32_t *rowindex,
                        const int32_t *colindex, const double *values);
/* Build generate_matrix's CSR for an nx*ny*nz grid directly in device memory (same bytes as
 * the host build; SparseMatrix.F90:138-175).  Any of the host output pointers may be NULL;
 * non-NULL ones receive a copy (rowindex[N+1], colindex/values[nnz], x/b/xexact[N]).
 * *nnz_out receives the true nnz. */
int hpccg_generate_matrix(hpccg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, int32_t *rowindex,
                          int32_t *colindex, double *values, double *x, double *b, double *xexact,
                          int64_t *nnz_out);
int hpccg_matrix_info(hpccg_ctx *ctx, int32_t *nrow, int64_t *nnz, int32_t *max_row_nnz);
/* Device format of the uploaded matrix (no reference counterpart: an inspection call).
 * *ncodes: size of the column-offset table when the sliced copies store one-byte column codes
 * (column = row + table[code]; used when the matrix has at most 256 distinct offsets col - row,
 * 27 for the 27-point stencil), 0 when they store int32 columns.  *copy_bytes: device bytes of
 * the FP64 sliced copy (values, columns or codes, slice pointers). */
int hpccg_matrix_format(hpccg_ctx *ctx, int32_t *ncodes, int64_t *copy_bytes);

/* ---- kernels, host arrays (drop-ins for the reference procedures) ---------------------- */
/* spmv(A, x, y)             SparseMatrix.F90:213-232 */
int hpccg_spmv(hpccg_ctx *ctx, const double *x, double *y);
/* ddot(n, x, y, r)          vector_operations.F90:40-71 */
int hpccg_ddot(hpccg_ctx *ctx, int32_t n, const double *x, const double *y, double *r);
/* ddot1(n, x, r)            vector_operations.F90:73-103 */
int hpccg_ddot1(hpccg_ctx *ctx, int32_t n, const double *x, double *r);
/* waxpby(n, alpha, x, beta, y, w): w = alpha*x + beta*y     vector_operations.F90:105-121 */
int hpccg_waxpby(hpccg_ctx *ctx, int32_t n, double alpha, const double *x, double beta,
                 const double *y, double *w);
/* wwpby(n, beta, y, w): w = w + beta*y                      vector_operations.F90:123-137 */
int hpccg_wwpby(hpccg_ctx *ctx, int32_t n, double beta, const double *y, double *w);
/* wxbw(n, x, beta, w): w = x + beta*w                       vector_operations.F90:139-153 */
int hpccg_wxbw(hpccg_ctx *ctx, int32_t n, const double *x, double beta, double *w);
/* vcopy(n, x, y): y = x                                     vector_operations.F90:155-169 */
int hpccg_vcopy(hpccg_ctx *ctx, int32_t n, const double *x, double *y);

/* compute_residual(n, v1, v2, residual): max_i |v1_i - v2_i|   numerics.F90:164-198
 * Exact (a maximum is order-free).  Unlike the reference, whose local_residual keeps its value
 * between calls (initialised in its declaration, numerics.F90:177, hence implicitly SAVEd), every
 * call starts from 0; the result equals the reference's first call.  NaN differences are skipped. */
int hpccg_compute_residual(hpccg_ctx *ctx, int32_t n, const double *v1, const double *v2,
                           double *residual);

/* cg(A, b, x, max_iter, tolerance, niters, normr)           numerics.F90:41-162
 * Unpreconditioned CG from the x passed in (updated in place), same operation order, same
 * lagging convergence test (exit at the top of iteration k when ||r_{k-2}|| < tolerance),
 * niters = last iteration run, normr = ||r_{niters-1}||.  hist (may be NULL) receives ||r_j||
 * for j = 0..max_iter (max_iter+1 doubles; entries never formed are -1.0; hist[max_iter] is
 * ||r_max_iter||, which the reference computes but never norms).  Returns HPCCG_WARN_BREAKDOWN
 * if p.Ap was zero or non-finite in some iteration (the reference continues with NaNs, and so
 * does this). */
int hpccg_cg(hpccg_ctx *ctx, const double *b, double *x, int32_t max_iter, double tolerance,
             int32_t *niters, double *normr, double *hist);

/* cg (host arrays, as hpccg_cg) with the reference's timer regions measured on the device
 * (numerics.F90:91-156, TIMER_* levels of MPI/includes/timer.h:40-58).  Times in seconds.
 *   level 1:  the graph-replayed solve; *t_total = the iteration loop with its final x update
 *             (the "Total" region, numerics.F90:106,149), *t_prepare = Prepare (:91-100).
 *   level >= 2: iterations launched one at a time with the stopping test on the host before
 *             each (numerics.F90:109): t_iter[k-1] = iteration k ("Iteration", :113,142),
 *             t_spmv[k-1] = its SpMV ("spmv", :132-134), for k = 1..niters (arrays of max_iter,
 *             may be NULL); *t_total, *t_prepare as above.
 * roctx ranges named "Prepare", "Total", "Iteration" (level >= 2) and "spmv" (level >= 2)
 * bracket the same device work (rocprofv3 --marker-trace).  Results are hpccg_cg's bit for bit. */
int hpccg_cg_timed(hpccg_ctx *ctx, const double *b, double *x, int32_t max_iter, double tolerance,
                   int32_t level, int32_t *niters, double *normr, double *hist, double *t_total,
                   double *t_prepare, double *t_iter, double *t_spmv);

/* ---- kernels, device arrays --------------------------------------------------------------- */
int hpccg_spmv_dev(hpccg_ctx *ctx, const double *x, double *y);
int hpccg_ddot_dev(hpccg_ctx *ctx, int32_t n, const double *x, const double *y, double *r);
int hpccg_waxpby_dev(hpccg_ctx *ctx, int32_t n, double alpha, const double *x, double beta,
                     const double *y, double *w);
int hpccg_wwpby_dev(hpccg_ctx *ctx, int32_t n, double beta, const double *y, double *w);
int hpccg_wxbw_dev(hpccg_ctx *ctx, int32_t n, const double *x, double beta, double *w);
int hpccg_compute_residual_dev(hpccg_ctx *ctx, int32_t n, const double *v1, const double *v2,
                               double *residual);
/* cg with b and x resident on the device.  *elapsed_ms (may be NULL) receives the device time
 * of the whole solve -- Prepare, iteration loop and final x update (HIP events on the
 * context's stream; the reference's "Total" timer, numerics.F90:106, covers the loop only). */
int hpccg_cg_dev(hpccg_ctx *ctx, const double *b, double *x, int32_t max_iter, double tolerance,
                 int32_t *niters, double *normr, double *hist, double *elapsed_ms);

/* No counterpart in the reference.  hpccg_cg_dev for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid allocation): same kernels in the same order, so the same
 * results bit for bit, but the workspace (r, p, Ap, control block, residual history, partial
 * sums) is a static array that each device holds for the life of the process (50.6 MB), its
 * kernels are launched directly (no graph is captured), and the call allocates nothing.  Limits:
 * nrow <= HPCCG_NOALLOC_MAX_ROWS (128^3) and 0 <= max_iter <= HPCCG_NOALLOC_MAX_ITER; larger
 * sizes return HPCCG_ERR_ARG.  Every call on one device shares that device's workspace: calls
 * that overlap on one device (two contexts or threads) race.  Not reentrant. */
enum { HPCCG_NOALLOC_MAX_ROWS = 2097152, HPCCG_NOALLOC_MAX_ITER = 4094 };
int hpccg_cg_dev_noalloc(hpccg_ctx *ctx, const double *b, double *x, int32_t max_iter, double tolerance,
                         int32_t *niters, double *normr, double *hist, double *elapsed_ms);

/* ---- mixed precision (no reference counterpart: extends cg, numerics.F90:41-162) ----------
 * FP16 matrix values (exact for generate_matrix's 27 / -1; other matrices are rejected unless
 * every value is an FP16 number), int32 columns, FP32 r/p/Ap with FP64 arithmetic inside the
 * kernels, FP64 dots for alpha/beta, FP64 x, and replacement of the recursive residual by the
 * FP64 true residual b - A x every replace_every iterations (0 = no periodic replacement; the
 * true residual is still formed whenever the recursive norm falls below the tolerance, at
 * max_iter, and once the recursive r.r has fallen 1e10-fold since the last true one).
 * Stops when the TRUE residual norm is below `tolerance` (absolute).  niters = iterations run;
 * *true_normr = ||b - A x|| at exit; hist_rec[j] (j = 1..max_iter) the recursive ||r_j||;
 * hist_true[j] (j = 0..max_iter) the true ||b - A x_j|| where formed, else -1. */
int hpccg_cg_mixed(hpccg_ctx *ctx, const double *b, double *x, int32_t max_iter, double tolerance,
                   int32_t replace_every, int32_t *niters, double *true_normr, double *hist_rec,
                   double *hist_true);
int hpccg_cg_mixed_dev(hpccg_ctx *ctx, const double *b, double *x, int32_t max_iter,
                       double tolerance, int32_t replace_every, int32_t *niters, double *true_normr,
                       double *hist_rec, double *hist_true, double *elapsed_ms);
/* y = A x with FP16 values and FP32 vectors (FP64 accumulation); device arrays. */
int hpccg_spmv_mixed_dev(hpccg_ctx *ctx, const float *x, float *y);
int hpccg_time_spmv_mixed(hpccg_ctx *ctx, const float *x, float *y, int32_t reps, double *avg_ms);

/* ---- batched solves (no reference counterpart: many independent generate_matrix grids,
 * SparseMatrix.F90:68-185, each solved by the mixed-precision cg above) -------------------
 * hpccg_batch_setup builds the block-diagonal CSR of `count` grids (dims = nx,ny,nz per
 * system) on the device.  Vectors are concatenated in system order (row offsets from
 * hpccg_batch_info, count+1 entries).  A solve is one launch of one captured graph. */
int hpccg_batch_setup(hpccg_ctx *ctx, int32_t count, const int32_t *dims);
int hpccg_batch_info(hpccg_ctx *ctx, int32_t *count, int64_t *total_rows, int64_t *total_nnz,
                     int64_t *row_off);
int hpccg_batch_spmv_dev(hpccg_ctx *ctx, const float *x, float *y);
/* compiled launch variants: rows per wave (64 per slice) and waves per workgroup */
int hpccg_batch_variants(int32_t *count, int32_t *rows, int32_t *waves);
/* autotuner: times every variant on this batch (x restored) and keeps the fastest for solves
 * with variant = -1.  variant_ms[v] = ms of probe_iters working iterations + ms of probe_iters
 * iterations after every system has converged (a solve's graph holds max_iter iterations). */
int hpccg_batch_autotune(hpccg_ctx *ctx, const double *b, double *x, int32_t probe_iters,
                         int32_t *best, double *variant_ms);
/* every system to ||b_s - A_s x_s|| <= rel_tol ||b_s|| (FP64 true residual) or max_iter;
 * per-system niters / relres / converged (arrays of `count`) */
int hpccg_batch_solve_dev(hpccg_ctx *ctx, const double *b, double *x, double rel_tol,
                          int32_t max_iter, int32_t replace_every, int32_t variant,
                          int32_t *niters, double *relres, int32_t *converged, double *elapsed_ms);

/* ---- seeded synthetic inputs (no reference counterpart: the reference hard-codes x* = 1,
 * SparseMatrix.F90:170-172).  Host arrays; deterministic SplitMix64 streams. --------------- */
/* x* = sum of 8 random low-frequency 3-D sines with random phases + U(-0.05,0.05) noise */
int hpccg_xstar_sines(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double *xstar);
/* x*_i ~ U(-1,1) */
int hpccg_xstar_uniform(int64_t n, uint64_t seed, double *xstar);
/* 3*count extents ~ U{lo..hi}: (nx,ny,nz) of each batched system */
int hpccg_batch_dims(int32_t count, int32_t lo, int32_t hi, uint64_t seed, int32_t *dims);

/* ---- learned preconditioner (no reference counterpart: the reference's cg is unpreconditioned,
 * numerics.F90:41-162).  A 3-layer bias-free 3x3x3 convolutional network u -> z
 * (1 -> 16 -> 16 -> 1 channels, ReLU after layers 1 and 2, zero padding at the grid faces),
 * BF16 activations and weights, FP32 accumulation on MFMA.  Weights are FP32 masters,
 * hpccg_cnn_num_weights() of them: W1[16][27], W2[16][16][27], W3[16][27], tap index
 * t = (dz+1)*9 + (dy+1)*3 + (dx+1), voxel index (z*ny + y)*nx + x. -------------------------- */
int hpccg_cnn_num_weights(void);
/* Weights uniform in +-1/sqrt(fan_in) from a SplitMix64 stream; Adam state reset. */
int hpccg_cnn_init(hpccg_ctx *ctx, uint64_t seed);
int hpccg_cnn_set_weights(hpccg_ctx *ctx, const float *w);
int hpccg_cnn_get_weights(hpccg_ctx *ctx, float *w);
/* Forward pass on host arrays: u[batch][nvox] (rounded to BF16), h1/h2 = FP32 copies of the
 * BF16 hidden activations [batch][nvox][16] (may be NULL), z[batch][nvox]. */
int hpccg_cnn_forward(hpccg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, int32_t batch,
                      const float *u, float *h1, float *h2, float *z);
/* Forward + backward on host arrays for a given dL/dz: dW[hpccg_cnn_num_weights()]. */
int hpccg_cnn_backward(hpccg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, int32_t batch,
                       const float *u, const float *dz, float *dW);
/* Training set: npairs (r_k, e_k) pairs from plain FP64 CG runs on n^3 grids (seeded
 * x* = hpccg_xstar_sines), each a random crop^3 window scaled by 1/rms(r); per axis the window
 * lies against the low domain face, the high one, or anywhere, with probability 1/3 each.
 * Replaces the context's matrix. */
int hpccg_cnn_make_dataset(hpccg_ctx *ctx, int32_t n, int32_t npairs, int32_t crop, uint64_t seed);
/* Pair i of the training set (crop^3 floats each); margins6 (may be NULL) = loss-mask margins. */
int hpccg_cnn_dataset_pair(hpccg_ctx *ctx, int32_t i, float *r, float *e, int32_t *margins6);
/* `steps` Adam steps (warm-up, cosine decay) of the residual loss
 * ||mask.A(z - e)||^2 / ||mask.A e||^2 (A e = r on the mask, so this is ||A z - r||^2 / ||r||^2)
 * on batches of random reflected crops; one captured graph
 * per step.  loss_hist[steps] may be NULL; elapsed_ms = device time of all steps. */
int hpccg_cnn_train(hpccg_ctx *ctx, int32_t steps, int32_t batch, float lr, uint64_t seed,
                    float *loss_hist, double *elapsed_ms);
/* z = rms(r) * net(r / rms(r)) on device arrays of one grid; inv_scale = 1/rms(r). */
int hpccg_cnn_apply_dev(hpccg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, const float *r,
                        float inv_scale, float *z);
/* Flexible CG preconditioned by the network (FP16 matrix values, FP32 r/p/Ap/z, FP64 dots and
 * x, true-residual replacement every replace_every iterations and after every 1e10-fold fall
 * of the recursive r.r, as in hpccg_cg_mixed; stop on the FP64 true residual
 * ||b - A x|| <= tolerance).  Every `window` iterations ||r|| must have dropped by min_factor,
 * else (or on a breakdown) the current x is handed to FP64 CG (hpccg_cg_dev semantics) for the
 * rest of the solve.  (nx,ny,nz) is the matrix's grid.  hist_rec[0..max_iter] (may be NULL)
 * receives the recurrence ||r|| per FCG iteration, -1 past the last; elapsed_ms covers the
 * whole solve. */
int hpccg_fcg_learned_dev(hpccg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, const double *b,
                          double *x, int32_t max_iter, double tolerance, int32_t replace_every,
                          int32_t window, double min_factor, int32_t *fcg_iters,
                          int32_t *cg_iters, int32_t *fell_back, double *true_normr,
                          double *hist_rec, double *elapsed_ms);

/* ---- measurement -------------------------------------------------------------------------- */
/* Run `reps` back-to-back FP64 SpMVs y = A x on device arrays -- the kernel the CG loop
 * launches, with its p.Ap epilogue -- and return the mean kernel time in ms (HIP events on the
 * context's stream, the stream the kernel is launched on). */
int hpccg_time_spmv(hpccg_ctx *ctx, const double *x, double *y, int32_t reps, double *avg_ms);
/* No reference counterpart (the reference always multiplies by its stored CSR): the
 * matrix-free form of generate_matrix's operator (SparseMatrix.F90:138-175) on an nx x ny x nz
 * grid, for comparison with the CSR SpMV.  y = A x from the grid extents alone (27 on the
 * diagonal, -1 per in-grid neighbour), each row summed in generate_matrix's column order, so y
 * is bitwise hpccg_spmv_dev's on that matrix; no matrix needs to be resident.  Device arrays. */
int hpccg_spmv_stencil_dev(hpccg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, const double *x, double *y);
/* Mean time in ms of `reps` back-to-back matrix-free SpMVs (HIP events on the context's stream). */
int hpccg_time_spmv_stencil(hpccg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, const double *x, double *y,
                            int32_t reps, double *avg_ms);
/* A plain device copy of `bytes` bytes, `reps` times: the HBM calibration kernel. */
int hpccg_time_copy(hpccg_ctx *ctx, void *dst, const void *src, int64_t bytes, int32_t reps,
                    double *avg_ms);
/* No reference counterpart (the learned preconditioner is an addition): run `reps` back-to-back
 * applications of the fused network inference z = net(r) (hpccg_cnn_apply_dev with scale 1) on
 * device arrays of one nx*ny*nz field and return the mean kernel time in ms (HIP events on the
 * context's stream, the stream the kernel is launched on). */
int hpccg_time_cnn_apply(hpccg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, const float *r, float *z,
                         int32_t reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
