/*
 * CPU restatement of the reference CG path (RRZE-HPC/HPCCG-F90, MPI/ variant, single rank).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as the checker.  The product (libhpccg_hip) never links it.
 *
 * Conventions follow the reference: CSR is 1-based (rowindex[0] == 1, columns 1..N), FP64
 * values, int32 indices, rows ordered ix fastest then iy then iz.
 */
#ifndef HPCCG_ORACLE_H
#define HPCCG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* generate_matrix (MPI/src/SparseMatrix.F90:68-185).  rowindex[N+1], colindex/values sized
 * 27*N.  Returns the true nnz (rowindex[N]-1). */
int64_t oracle_generate_matrix(int nx, int ny, int nz, int32_t *rowindex, int32_t *colindex,
                               double *values, double *x, double *b, double *xexact);

/* spmv (MPI/src/SparseMatrix.F90:213-232): y_i = sum_k values_k * x_{colindex_k}, ascending k
 * from 0.0, products rounded before each add. */
void oracle_spmv(int n, const int32_t *rowindex, const int32_t *colindex, const double *values,
                 const double *x, double *y);

/* ddot / ddot1 (MPI/src/vector_operations.F90:40-103): one sequential chain from 0.0. */
double oracle_ddot(int n, const double *x, const double *y);
double oracle_ddot1(int n, const double *x);
/* Dot products rounded once from a double-double accumulation (near-exact); not a reference
 * function: the yardstick for reordered (tree) reductions. */
double oracle_ddot_exact(int n, const double *x, const double *y);

/* waxpby family (MPI/src/vector_operations.F90:105-169). */
void oracle_waxpby(int n, double alpha, const double *x, double beta, const double *y, double *w);
void oracle_wwpby(int n, double beta, const double *y, double *w);
void oracle_wxbw(int n, const double *x, double beta, double *w);
void oracle_vcopy(int n, const double *x, double *y);

/* cg (MPI/src/numerics.F90:41-162).  hist (may be NULL) receives ||r_j|| for j = 0..max_iter
 * (max_iter+1 entries; entries after early exit are set to -1).  hist[j] for j < max_iter is
 * exactly the "normr" the reference computes at the top of iteration j+1; hist[max_iter] is
 * ||r_max_iter|| (formed here with the same sequential ddot1, never formed by the reference).
 * exact_dots != 0 swaps both dots for oracle_ddot_exact.  Returns 0, or 1 on breakdown
 * (p.Ap == 0). */
int oracle_cg(int n, const int32_t *rowindex, const int32_t *colindex, const double *values,
              const double *b, double *x, int max_iter, double tolerance, int *niters,
              double *normr, double *hist, int exact_dots);

/* compute_residual (MPI/src/numerics.F90:164-198): max_i |v1_i - v2_i| (without the implicit
 * SAVE of local_residual that makes the reference correct only on its first call). */
double oracle_compute_residual(int n, const double *v1, const double *v2);

/* Seeded synthetic inputs (no reference counterpart; the reference hard-codes x* = 1).
 *   xstar_sines: sum of 8 low-frequency 3-D sines with random phases plus U(-0.05,0.05) noise
 *   xstar_uniform: U(-1,1)
 *   batch_dims: 3*count extents drawn from U{lo..hi}
 * All draws use SplitMix64; the formula is fixed in hpccg_oracle.c and mirrored bit-for-bit
 * by the product's generator. */
void oracle_xstar_sines(int nx, int ny, int nz, uint64_t seed, double *xstar);
void oracle_xstar_uniform(int64_t n, uint64_t seed, double *xstar);
void oracle_batch_dims(int count, int lo, int hi, uint64_t seed, int32_t *dims);

/* ---- learned preconditioner (no reference counterpart) -----------------------------------
 * 3-layer bias-free 3x3x3 CNN on an nx*ny*nz grid (ix fastest), zero padding outside the grid:
 *   h1 = relu(conv(u; W1[16][27])), h2 = relu(conv(h1; W2[16][16][27])), z = conv(h2; W3[16][27])
 * conv(in; W)[co](v) = sum_{ci,t} W[co][ci][t] * in[ci](v + off(t)),
 * off(t) = (t%3-1, (t/3)%3-1, t/9-1) in (x, y, z).  Activations are channel-last
 * ([batch][voxel][16]).  Sums are formed in double.  quant != 0 applies the GPU's storage
 * precisions: u, weights, h1, h2 (and dz, dh2, dh1 in the backward pass) rounded to bfloat16
 * (round to nearest even); z and weight gradients are float. */
float oracle_bf16(float v);
void oracle_cnn_forward(int nx, int ny, int nz, int batch, const float *W1, const float *W2,
                        const float *W3, const float *u, float *h1, float *h2, float *z, int quant);
/* Given dz = dL/dz, the weight gradients dW1, dW2, dW3 (same layouts as W). */
void oracle_cnn_backward(int nx, int ny, int nz, int batch, const float *W1, const float *W2,
                         const float *W3, const float *u, const float *h1, const float *h2,
                         const float *dz, float *dW1, float *dW2, float *dW3, int quant);

#ifdef __cplusplus
}
#endif
#endif
