/*
 * CPU restatement of the reference CG path -- TEST INFRASTRUCTURE ONLY (see hpccg_oracle.h).
 *
 * Every loop keeps the reference's operation order; the file is compiled with
 * -ffp-contract=off so that no multiply-add is fused (the reference's serial amdflang -O3
 * build performs none: products round before each add).
 */
#include "hpccg_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>

/* MPI/src/SparseMatrix.F90:68-185, single rank (start_row = 1, total_nrow = N). */
int64_t oracle_generate_matrix(int nx, int ny, int nz, int32_t *rowindex, int32_t *colindex,
                               double *values, double *x, double *b, double *xexact)
{
    const int64_t total_nrow = (int64_t)nx * ny * nz;
    int64_t cursor = 1; /* 1-based, SparseMatrix.F90:102 */
    int64_t rowcursor = 1;
    for (int iz = 1; iz <= nz; ++iz)
        for (int iy = 1; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix) {
                /* SparseMatrix.F90:141-142 */
                const int64_t curlocalrow = (int64_t)(iz - 1) * nx * ny + (int64_t)(iy - 1) * nx + (ix - 1);
                const int64_t currow = 1 + curlocalrow;
                int nnzrow = 0;
                rowindex[rowcursor - 1] = (int32_t)cursor;
                /* SparseMatrix.F90:146-167: sz, sy, sx each -1..1, sx fastest */
                for (int sz = -1; sz <= 1; ++sz)
                    for (int sy = -1; sy <= 1; ++sy)
                        for (int sx = -1; sx <= 1; ++sx) {
                            const int64_t curcol = currow + (int64_t)sz * nx * ny + (int64_t)sy * nx + sx;
                            if (ix + sx >= 1 && ix + sx <= nx && iy + sy >= 1 && iy + sy <= ny &&
                                curcol >= 1 && curcol <= total_nrow) {
                                values[cursor - 1] = (curcol == currow) ? 27.0 : -1.0;
                                colindex[cursor - 1] = (int32_t)curcol;
                                ++cursor;
                                ++nnzrow;
                            }
                        }
                ++rowcursor;
                /* SparseMatrix.F90:170-172 */
                x[curlocalrow] = 0.0;
                b[curlocalrow] = 27.0 - (double)(nnzrow - 1);
                xexact[curlocalrow] = 1.0;
            }
    rowindex[rowcursor - 1] = (int32_t)cursor; /* SparseMatrix.F90:177 */
    return cursor - 1;
}

/* MPI/src/SparseMatrix.F90:222-231 */
void oracle_spmv(int n, const int32_t *rowindex, const int32_t *colindex, const double *values,
                 const double *x, double *y)
{
    for (int i = 0; i < n; ++i) {
        double rowsum = 0.0;
        for (int32_t k = rowindex[i]; k <= rowindex[i + 1] - 1; ++k) {
            const double prod = values[k - 1] * x[colindex[k - 1] - 1];
            rowsum = rowsum + prod;
        }
        y[i] = rowsum;
    }
}

/* MPI/src/vector_operations.F90:54-59 */
double oracle_ddot(int n, const double *x, const double *y)
{
    double r = 0.0;
    for (int i = 0; i < n; ++i) {
        const double prod = x[i] * y[i];
        r = r + prod;
    }
    return r;
}

/* MPI/src/vector_operations.F90:86-91 */
double oracle_ddot1(int n, const double *x)
{
    double r = 0.0;
    for (int i = 0; i < n; ++i) {
        const double prod = x[i] * x[i];
        r = r + prod;
    }
    return r;
}

/* double-double accumulation: TwoProd via fma (exact product), TwoSum (Knuth). */
double oracle_ddot_exact(int n, const double *x, const double *y)
{
    double hi = 0.0, lo = 0.0;
    for (int i = 0; i < n; ++i) {
        const double p = x[i] * y[i];
        const double pe = fma(x[i], y[i], -p);
        const double s = hi + p;
        const double bb = s - hi;
        const double se = (hi - (s - bb)) + (p - bb);
        hi = s;
        lo = lo + (se + pe);
    }
    return hi + lo;
}

/* MPI/src/vector_operations.F90:116-119: w = alpha*x + beta*y */
void oracle_waxpby(int n, double alpha, const double *x, double beta, const double *y, double *w)
{
    for (int i = 0; i < n; ++i) {
        const double t1 = alpha * x[i];
        const double t2 = beta * y[i];
        w[i] = t1 + t2;
    }
}

/* MPI/src/vector_operations.F90:132-135: w = w + beta*y */
void oracle_wwpby(int n, double beta, const double *y, double *w)
{
    for (int i = 0; i < n; ++i) {
        const double t = beta * y[i];
        w[i] = w[i] + t;
    }
}

/* MPI/src/vector_operations.F90:148-151: w = x + beta*w */
void oracle_wxbw(int n, const double *x, double beta, double *w)
{
    for (int i = 0; i < n; ++i) {
        const double t = beta * w[i];
        w[i] = x[i] + t;
    }
}

/* MPI/src/vector_operations.F90:164-167 */
void oracle_vcopy(int n, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] = x[i];
}

/* MPI/src/numerics.F90:41-162 (single rank: exchange_externals drops out; ncol == nrow). */
int oracle_cg(int n, const int32_t *rowindex, const int32_t *colindex, const double *values,
              const double *b, double *x, int max_iter, double tolerance, int *niters,
              double *normr_out, double *hist, int exact_dots)
{
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    double *Ap = (double *)malloc(sizeof(double) * (size_t)n);
    int status = 0;
    if (!r || !p || !Ap) {
        free(r); free(p); free(Ap);
        return 2;
    }
    if (hist)
        for (int j = 0; j <= max_iter; ++j) hist[j] = -1.0;

    double rtrans, oldrtrans, alpha, beta, normr;
    *niters = 0;
    /* Prepare, numerics.F90:92-99 */
    oracle_vcopy(n, x, p);
    oracle_spmv(n, rowindex, colindex, values, p, Ap);
    oracle_waxpby(n, 1.0, b, -1.0, Ap, r);
    rtrans = exact_dots ? oracle_ddot_exact(n, r, r) : oracle_ddot1(n, r);
    normr = sqrt(rtrans);

    int k;
    for (k = 1; k <= max_iter; ++k) {
        if (normr < tolerance) break; /* numerics.F90:109-112 (lags one iteration) */
        if (k == 1) {
            oracle_vcopy(n, r, p);
        } else {
            oldrtrans = rtrans;
            rtrans = exact_dots ? oracle_ddot_exact(n, r, r) : oracle_ddot1(n, r);
            beta = rtrans / oldrtrans;
            oracle_wxbw(n, r, beta, p);
        }
        normr = sqrt(rtrans); /* ||r_{k-1}|| */
        if (hist) hist[k - 1] = normr;
        oracle_spmv(n, rowindex, colindex, values, p, Ap);
        alpha = exact_dots ? oracle_ddot_exact(n, p, Ap) : oracle_ddot(n, p, Ap);
        if (alpha == 0.0) status = 1;
        alpha = rtrans / alpha;
        oracle_wwpby(n, -alpha, Ap, r);
        oracle_wwpby(n, alpha, p, x);
        *niters = k;
    }
    if (hist && k > max_iter) {
        const double rr = exact_dots ? oracle_ddot_exact(n, r, r) : oracle_ddot1(n, r);
        hist[max_iter] = sqrt(rr);
    }
    *normr_out = normr;
    free(r); free(p); free(Ap);
    return status;
}

/* MPI/src/numerics.F90:176-186 */
double oracle_compute_residual(int n, const double *v1, const double *v2)
{
    double local_residual = 0.0;
    for (int i = 0; i < n; ++i) {
        const double diff = fabs(v1[i] - v2[i]);
        if (diff > local_residual) local_residual = diff;
    }
    return local_residual;
}

/* ------------------------------------------------------------------------------------------ */
/* Seeded inputs.  SplitMix64 (Steele, Lea, Flood 2014): state += golden; mix(state).        */
static const uint64_t GOLDEN = 0x9E3779B97F4A7C15ULL;

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static uint64_t next64(uint64_t *s)
{
    *s += GOLDEN;
    return mix64(*s);
}
static double unit53(uint64_t z) { return (double)(z >> 11) * 0x1.0p-53; }

void oracle_xstar_sines(int nx, int ny, int nz, uint64_t seed, double *xstar)
{
    const int dims[3] = {nx, ny, nz};
    double amp[8];
    int kk[8][3];
    double phi[8][3];
    uint64_t s = seed;
    for (int m = 0; m < 8; ++m) {
        amp[m] = 2.0 * unit53(next64(&s)) - 1.0;
        for (int d = 0; d < 3; ++d) {
            kk[m][d] = 1 + (int)((next64(&s) >> 33) % 4u);
            phi[m][d] = 6.283185307179586 * unit53(next64(&s));
        }
    }
    /* separable factor tables S[m][d][i] = sin(pi*k*(i+1)/(n_d+1) + phi) */
    double *tab[8][3];
    for (int m = 0; m < 8; ++m)
        for (int d = 0; d < 3; ++d) {
            tab[m][d] = (double *)malloc(sizeof(double) * (size_t)dims[d]);
            for (int i = 0; i < dims[d]; ++i) {
                double t = 3.141592653589793 * (double)kk[m][d];
                t = t * (double)(i + 1);
                t = t / (double)(dims[d] + 1);
                t = t + phi[m][d];
                tab[m][d][i] = sin(t);
            }
        }
    const uint64_t nseed = seed ^ 0xD1B54A32D192ED03ULL;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int64_t v = (int64_t)ix + (int64_t)nx * ((int64_t)iy + (int64_t)ny * iz);
                double acc = 0.0;
                for (int m = 0; m < 8; ++m) {
                    double t = amp[m] * tab[m][0][ix];
                    t = t * tab[m][1][iy];
                    t = t * tab[m][2][iz];
                    acc = acc + t;
                }
                const double noise = 0.1 * unit53(mix64(nseed + (uint64_t)(v + 1) * GOLDEN)) - 0.05;
                xstar[v] = acc + noise;
            }
    for (int m = 0; m < 8; ++m)
        for (int d = 0; d < 3; ++d) free(tab[m][d]);
}

void oracle_xstar_uniform(int64_t n, uint64_t seed, double *xstar)
{
    for (int64_t i = 0; i < n; ++i)
        xstar[i] = 2.0 * unit53(mix64(seed + (uint64_t)(i + 1) * GOLDEN)) - 1.0;
}

void oracle_batch_dims(int count, int lo, int hi, uint64_t seed, int32_t *dims)
{
    uint64_t s = seed;
    const uint64_t span = (uint64_t)(hi - lo + 1);
    for (int j = 0; j < 3 * count; ++j) dims[j] = lo + (int32_t)((next64(&s) >> 33) % span);
}

/* ------------------------------------------------------------------------------------------ */
/* Learned preconditioner CNN (no reference counterpart).                                     */
#include <string.h>

float oracle_bf16(float v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v; /* inf / nan unchanged */
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

static float q_(float v, int quant) { return quant ? oracle_bf16(v) : v; }

/* out[b][v][co] = sum_{ci,t} W[co][ci][t] in[b][v+off(t)][ci], channel-last, zero padding;
 * flip != 0 applies the transposed-flipped kernel W[ci][co][26-t] (backward data). */
static void conv3(int nx, int ny, int nz, int batch, int cin, int cout, const float *W, int flip,
                  const float *in, double *out)
{
    const int64_t N = (int64_t)nx * ny * nz;
    for (int b = 0; b < batch; ++b)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const int64_t v = ((int64_t)z * ny + y) * nx + x;
                    double acc[16] = {0};
                    for (int t = 0; t < 27; ++t) {
                        const int xx = x + t % 3 - 1, yy = y + (t / 3) % 3 - 1, zz = z + t / 9 - 1;
                        if (xx < 0 || xx >= nx || yy < 0 || yy >= ny || zz < 0 || zz >= nz) continue;
                        const float *iv = in + ((int64_t)b * N + ((int64_t)zz * ny + yy) * nx + xx) * cin;
                        for (int co = 0; co < cout; ++co)
                            for (int ci = 0; ci < cin; ++ci) {
                                const float w = flip ? W[((int64_t)ci * cout + co) * 27 + (26 - t)]
                                                     : W[((int64_t)co * cin + ci) * 27 + t];
                                acc[co] += (double)w * (double)iv[ci];
                            }
                    }
                    for (int co = 0; co < cout; ++co) out[((int64_t)b * N + v) * cout + co] = acc[co];
                }
}

/* dW[co][ci][t] += sum_v dout[v][co] * in[v+off(t)][ci] */
static void conv3_wgrad(int nx, int ny, int nz, int batch, int cin, int cout, const float *dout,
                        const float *in, double *dW)
{
    const int64_t N = (int64_t)nx * ny * nz;
    for (int b = 0; b < batch; ++b)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const int64_t v = ((int64_t)z * ny + y) * nx + x;
                    const float *dv = dout + ((int64_t)b * N + v) * cout;
                    for (int t = 0; t < 27; ++t) {
                        const int xx = x + t % 3 - 1, yy = y + (t / 3) % 3 - 1, zz = z + t / 9 - 1;
                        if (xx < 0 || xx >= nx || yy < 0 || yy >= ny || zz < 0 || zz >= nz) continue;
                        const float *iv = in + ((int64_t)b * N + ((int64_t)zz * ny + yy) * nx + xx) * cin;
                        for (int co = 0; co < cout; ++co)
                            for (int ci = 0; ci < cin; ++ci)
                                dW[((int64_t)co * cin + ci) * 27 + t] += (double)dv[co] * (double)iv[ci];
                    }
                }
}

static float *qcopy(const float *a, int64_t n, int quant)
{
    float *o = (float *)malloc(sizeof(float) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) o[i] = q_(a[i], quant);
    return o;
}

void oracle_cnn_forward(int nx, int ny, int nz, int batch, const float *W1, const float *W2,
                        const float *W3, const float *u, float *h1, float *h2, float *z, int quant)
{
    const int64_t M = (int64_t)nx * ny * nz * batch;
    float *w1 = qcopy(W1, 16 * 27, quant), *w2 = qcopy(W2, 16 * 16 * 27, quant), *w3 = qcopy(W3, 16 * 27, quant);
    float *uq = qcopy(u, M, quant);
    double *acc = (double *)malloc(sizeof(double) * (size_t)M * 16);
    conv3(nx, ny, nz, batch, 1, 16, w1, 0, uq, acc);
    for (int64_t i = 0; i < M * 16; ++i) h1[i] = q_(acc[i] > 0 ? (float)acc[i] : 0.0f, quant);
    conv3(nx, ny, nz, batch, 16, 16, w2, 0, h1, acc);
    for (int64_t i = 0; i < M * 16; ++i) h2[i] = q_(acc[i] > 0 ? (float)acc[i] : 0.0f, quant);
    conv3(nx, ny, nz, batch, 16, 1, w3, 0, h2, acc);
    for (int64_t i = 0; i < M; ++i) z[i] = (float)acc[i];
    free(acc); free(uq); free(w1); free(w2); free(w3);
}

void oracle_cnn_backward(int nx, int ny, int nz, int batch, const float *W1, const float *W2,
                         const float *W3, const float *u, const float *h1, const float *h2,
                         const float *dz, float *dW1, float *dW2, float *dW3, int quant)
{
    const int64_t M = (int64_t)nx * ny * nz * batch;
    float *w2 = qcopy(W2, 16 * 16 * 27, quant), *w3 = qcopy(W3, 16 * 27, quant);
    float *uq = qcopy(u, M, quant), *dzq = qcopy(dz, M, quant);
    double *acc = (double *)malloc(sizeof(double) * (size_t)M * 16);
    float *dh2 = (float *)malloc(sizeof(float) * (size_t)M * 16);
    float *dh1 = (float *)malloc(sizeof(float) * (size_t)M * 16);
    double g3[16 * 27] = {0}, g2[16 * 16 * 27] = {0}, g1[16 * 27] = {0};
    conv3_wgrad(nx, ny, nz, batch, 16, 1, dzq, h2, g3);
    conv3(nx, ny, nz, batch, 1, 16, w3, 1, dzq, acc); /* W3 as [co=1][ci=16] -> flipped [16][1] */
    for (int64_t i = 0; i < M * 16; ++i) dh2[i] = h2[i] > 0 ? q_((float)acc[i], quant) : 0.0f;
    conv3_wgrad(nx, ny, nz, batch, 16, 16, dh2, h1, g2);
    conv3(nx, ny, nz, batch, 16, 16, w2, 1, dh2, acc);
    for (int64_t i = 0; i < M * 16; ++i) dh1[i] = h1[i] > 0 ? q_((float)acc[i], quant) : 0.0f;
    conv3_wgrad(nx, ny, nz, batch, 1, 16, dh1, uq, g1);
    for (int i = 0; i < 16 * 27; ++i) dW3[i] = (float)g3[i];
    for (int i = 0; i < 16 * 16 * 27; ++i) dW2[i] = (float)g2[i];
    for (int i = 0; i < 16 * 27; ++i) dW1[i] = (float)g1[i];
    free(acc); free(dh1); free(dh2); free(uq); free(dzq); free(w2); free(w3);
}
