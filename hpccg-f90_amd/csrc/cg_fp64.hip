// FP64 CG path of HPCCG-F90 on MI355X: CSR SpMV, dot products, vector updates, and the CG
// iteration captured as a HIP graph.  Reference: MPI/src/numerics.F90:41-162,
// MPI/src/SparseMatrix.F90:213-232, MPI/src/vector_operations.F90:40-169.
//
// Rounding contract.  The file is compiled with -ffp-contract=off: every product rounds
// before it is added, exactly as the reference's serial loops do.
//   * SpMV keeps the reference's per-row order (ascending k, from 0.0) in both dot modes, so
//     y = A x is bitwise equal to the reference spmv.
//   * waxpby / wwpby / wxbw are elementwise with the reference's operand order: bitwise.
//   * Dots: HPCCG_DOT_SEQUENTIAL reproduces the reference's single left-to-right chain;
//     HPCCG_DOT_TREE uses a fixed wave/workgroup/grid tree (fused multiply-adds inside).

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "context.hpp"
#include "spmv.hpp"

namespace hpccg {

static thread_local char g_err[512] = "";
static std::atomic<uint64_t> g_buffer_gen{1};

uint64_t buffer_generation() { return g_buffer_gen.load(std::memory_order_acquire); }

hipError_t dmalloc_bytes(void **p, size_t bytes)
{
    g_buffer_gen.fetch_add(1, std::memory_order_acq_rel);
    return hipMalloc(p, bytes);
}

hipError_t dfree(void *p)
{
    if (!p) return hipSuccess;
    g_buffer_gen.fetch_add(1, std::memory_order_acq_rel);
    return hipFree(p);
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}


// FP64 SpMV of this thread's row of a one-wave-per-slice launch: y[row] = (A x)[row],
// returning x[row] * y[row] (the p.Ap term when x = p), or 0 past the last row.
__device__ __forceinline__ double sell_spmv_f64(const SellRef<double> &A, const double *__restrict__ x,
                                                double *__restrict__ y, bool want_dot)
{
    int64_t row;
    bool valid;
    const double sum = sell_rowsum_of_thread<kGroupF64, double, double, kSellWaves, true>(A, x, row, valid);
    if (!valid) return 0.0;
    __builtin_nontemporal_store(sum, y + row);  // streamed out: keeps x's lines in L2
    return want_dot ? x[row] * sum : 0.0;
}

// y = A x  (standalone spmv, and Prepare's Ap = A x with p = x, numerics.F90:92-96)
__global__ __launch_bounds__(kSellThreads) void k_spmv(const SellRef<double> A, const double *__restrict__ x,
                                                       double *__restrict__ y)
{
    sell_spmv_f64(A, x, y, false);
}

// Prepare's Ap = A x, skipped when k_nonzero found x = 0 (k_cg_prep_r then reads Ap as +0.0)
__global__ __launch_bounds__(kSellThreads) void k_spmv_prep(const SellRef<double> A, const double *__restrict__ x,
                                                            double *__restrict__ y, const CgState *st)
{
    if (!st->x_nonzero) return;
    sell_spmv_f64(A, x, y, false);
}

// ---- CG control logic shared by the tree and sequential variants ---------------------------
__device__ __forceinline__ bool cg_stopped(const CgState *st)
{
    return st->k >= st->stop_at;
}

// after p.Ap of iteration k: alpha = rtrans / (p.Ap)          numerics.F90:135-138
__device__ __forceinline__ void cg_after_pap(CgState *st, const double *rr_hist, int k, double pap)
{
    if (!(pap != 0.0) || !isfinite(pap)) st->breakdown = 1;
    st->alpha = rr_hist[k - 1] / pap;
}

// after r.r of iteration k (k = 0: Prepare).  Records ||r_k||^2, applies the reference's
// lagging test (the top of iteration k+2 compares ||r_k|| with the tolerance; Prepare's
// ||r_0|| is tested at the top of iteration 1), forms beta for iteration k+1 and advances k.
__device__ __forceinline__ void cg_after_rr(CgState *st, double *rr_hist, int k, double rr)
{
    rr_hist[k] = rr;
    st->rr = rr;
    const double normr = sqrt(rr);
    if (k == 0) {
        if (normr < st->tol) st->stop_at = 1;
        st->k = 1;
        return;
    }
    if (normr < st->tol) st->stop_at = min(st->stop_at, k + 2);
    st->beta = rr / rr_hist[k - 1];  // numerics.F90:117-119
    st->k = k + 1;
}

// ---- CG kernels ----------------------------------------------------------------------------
// Ap = A p, and (TREE) the workgroup's share of p.Ap     numerics.F90:133-136
template <bool TREE>
__global__ __launch_bounds__(kSellThreads) void k_cg_spmv(const SellRef<double> A, const double *__restrict__ p,
                                                          double *__restrict__ Ap, const CgState *st,
                                                          double *red_part)
{
    __shared__ double red[kSellThreads / kWave];
    if (st->k >= st->stop_at) return;
    const double dot = sell_spmv_f64(A, p, Ap, TREE);
    if (TREE) store_partial(block_sum<kSellThreads>(dot, red), red_part);
}

// p.Ap from the SpMV's partials -> alpha                 numerics.F90:136-138
__global__ __launch_bounds__(kFinishThreads) void k_cg_pap_finish(const double *__restrict__ red_part, int nparts,
                                                                  CgState *st, const double *rr_hist)
{
    __shared__ double red[kFinishThreads / kWave];
    const int k = st->k;
    if (k >= st->stop_at) return;
    const double pap = sum_partials(red_part, nparts, red);
    if (threadIdx.x == 0) cg_after_pap(st, rr_hist, k, pap);
}

// Prepare: r = 1*b + (-1)*Ap and (TREE) the workgroup's r.r      numerics.F90:97-98
template <bool TREE>
__global__ __launch_bounds__(kVecThreads) void k_cg_prep_r(const double *__restrict__ b,
                                                           const double *__restrict__ Ap,
                                                           double *__restrict__ r, int n, double *partials,
                                                           const CgState *st)
{
    __shared__ double red[kVecThreads / kWave];
    const bool x0 = !st->x_nonzero;  // x = 0: Ap = A x is +0.0 in every row and was not formed
    int64_t i0, i1;
    int lb;
    elem_range(n, i0, i1, lb);
    double dot = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) {
        const double t1 = 1.0 * b[i];
        const double t2 = -1.0 * (x0 ? 0.0 : Ap[i]);
        const double ri = t1 + t2;
        r[i] = ri;
        dot = fma(ri, ri, dot);
    }
    if (TREE) store_partial(block_sum<kVecThreads>(dot, red), partials);
}

// r = r + (-alpha)*Ap and (TREE) the workgroup's r.r      numerics.F90:139, :117
template <bool TREE>
__global__ __launch_bounds__(kVecThreads) void k_cg_update_r(double *__restrict__ r,
                                                             const double *__restrict__ Ap, int n,
                                                             const CgState *st, double *partials)
{
    __shared__ double red[kVecThreads / kWave];
    if (st->k >= st->stop_at) return;
    const double nalpha = -st->alpha;
    int64_t i0, i1;
    int lb;
    elem_range(n, i0, i1, lb);
    double dot = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) {
        const double t = nalpha * __builtin_nontemporal_load(Ap + i);
        const double ri = r[i] + t;
        r[i] = ri;
        dot = fma(ri, ri, dot);
    }
    if (TREE) store_partial(block_sum<kVecThreads>(dot, red), partials);
}

// r.r from the partials -> beta, test, k++ (PREP: rtrans of Prepare)   numerics.F90:98-99, :117-119
template <bool PREP>
__global__ __launch_bounds__(kFinishThreads) void k_cg_rr_finish(const double *__restrict__ partials, int nparts,
                                                                 CgState *st, double *rr_hist)
{
    __shared__ double red[kFinishThreads / kWave];
    const int k = PREP ? 0 : st->k;
    if (!PREP && k >= st->stop_at) return;
    const double rr = sum_partials(partials, nparts, red);
    if (threadIdx.x == 0) cg_after_rr(st, rr_hist, k, rr);
}

// Top of iteration k: k == 1: p = r (numerics.F90:114-115); else x = x + alpha_{k-1} p_{k-1}
// (the x update of iteration k-1, numerics.F90:140, deferred to here so that p is read once)
// and p = r + beta p (numerics.F90:120).
__global__ __launch_bounds__(kVecThreads) void k_cg_update_p(const double *__restrict__ r,
                                                             double *__restrict__ p,
                                                             double *__restrict__ x, int n,
                                                             const CgState *st)
{
    const int k = st->k;
    if (k >= st->stop_at) return;
    int64_t i0, i1;
    int lb;
    elem_range(n, i0, i1, lb);
    // r and x are streamed non-temporally: p, which the SpMV gathers next, then stays in the
    // last-level cache (measured: the following SpMV runs 1.06 instead of 1.11 ms at 256^3)
    if (k == 1) {
        for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) p[i] = __builtin_nontemporal_load(r + i);
        return;
    }
    const double alpha = st->alpha, beta = st->beta;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) {
        const double pi = p[i];
        const double tx = alpha * pi;
        __builtin_nontemporal_store(__builtin_nontemporal_load(x + i) + tx, x + i);
        const double tp = beta * pi;
        p[i] = __builtin_nontemporal_load(r + i) + tp;
    }
}

// After the loop: the x update of the last iteration that ran.
__global__ __launch_bounds__(kVecThreads) void k_cg_finish_x(const double *__restrict__ p,
                                                             double *__restrict__ x, int n,
                                                             const CgState *st)
{
    if (st->k - 1 < 1) return;
    const double alpha = st->alpha;
    int64_t i0, i1;
    int lb;
    elem_range(n, i0, i1, lb);
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) {
        const double tx = alpha * p[i];
        x[i] = x[i] + tx;
    }
}

// ---- sequential dots: the reference's single left-to-right chain --------------------------
// One wavefront.  Lanes load 4x64 consecutive elements and form the rounded products; the
// running sum then walks them in index order (v_readlane broadcasts), one rounded add each.
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
    return __longlong_as_double(((uint64_t)hi << 32) | lo);
}

template <bool SELF>
__device__ double serial_dot(const double *__restrict__ x, const double *__restrict__ y, int n)
{
    const int lane = threadIdx.x;
    double s = 0.0;
    int base = 0;
    for (; base + 4 * kWave <= n; base += 4 * kWave) {
        double pr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double a = x[base + u * kWave + lane];
            const double b = SELF ? a : y[base + u * kWave + lane];
            pr[u] = a * b;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int j = 0; j < kWave; ++j) s = s + readlane_f64(pr[u], j);
    }
    for (; base < n; base += kWave) {
        const int i = base + lane;
        double pr = 0.0;
        if (i < n) {
            const double a = x[i];
            const double b = SELF ? a : y[i];
            pr = a * b;
        }
        const int cnt = min(kWave, n - base);
        for (int j = 0; j < cnt; ++j) s = s + readlane_f64(pr, j);
    }
    return s;
}

__global__ __launch_bounds__(kWave) void k_dot_serial(const double *x, const double *y, int n,
                                                      double *out)
{
    const double s = serial_dot<false>(x, y, n);
    if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(kWave) void k_cg_serial_pap(const double *p, const double *Ap, int n,
                                                         CgState *st, const double *rr_hist)
{
    const int k = st->k;
    if (k >= st->stop_at) return;
    const double s = serial_dot<false>(p, Ap, n);
    if (threadIdx.x == 0) cg_after_pap(st, rr_hist, k, s);
}

__global__ __launch_bounds__(kWave) void k_cg_serial_rr(const double *r, int n, CgState *st,
                                                        double *rr_hist, int prep)
{
    const int k = prep ? 0 : st->k;
    if (!prep && k >= st->stop_at) return;
    const double s = serial_dot<true>(r, r, n);
    if (threadIdx.x == 0) cg_after_rr(st, rr_hist, k, s);
}

// ---- standalone tree dot and vector ops -----------------------------------------------------
__global__ __launch_bounds__(kVecThreads) void k_dot_tree(const double *__restrict__ x,
                                                          const double *__restrict__ y, int n, double *partials)
{
    __shared__ double red[kVecThreads / kWave];
    int64_t i0, i1;
    int lb;
    elem_range(n, i0, i1, lb);
    double dot = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) dot = fma(x[i], y[i], dot);
    store_partial(block_sum<kVecThreads>(dot, red), partials);
}
__global__ __launch_bounds__(kFinishThreads) void k_dot_finish(const double *__restrict__ partials, int nparts,
                                                               double *out)
{
    __shared__ double red[kFinishThreads / kWave];
    const double s = sum_partials(partials, nparts, red);
    if (threadIdx.x == 0) *out = s;
}

// max_i |v1_i - v2_i| (numerics.F90:178-183): per-thread maxima, a wavefront/workgroup max, and
// one atomic max per workgroup on the bit pattern of the non-negative result (bit patterns of
// non-negative doubles order like the values), so the result is exact and order-free
__global__ __launch_bounds__(kVecThreads) void k_maxabs_diff(const double *__restrict__ v1,
                                                             const double *__restrict__ v2, int n,
                                                             unsigned long long *out)
{
    __shared__ double red[kVecThreads / kWave];
    double m = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        m = fmax(m, fabs(v1[i] - v2[i]));
    for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kVecThreads / kWave; ++w) m = fmax(m, red[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// w = alpha*x + beta*y      vector_operations.F90:116-119
__global__ void k_waxpby(int n, double alpha, const double *__restrict__ x, double beta,
                         const double *__restrict__ y, double *__restrict__ w)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double t1 = alpha * x[i];
        const double t2 = beta * y[i];
        w[i] = t1 + t2;
    }
}
// w = w + beta*y            vector_operations.F90:132-135
__global__ void k_wwpby(int n, double beta, const double *__restrict__ y, double *__restrict__ w)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double t = beta * y[i];
        w[i] = w[i] + t;
    }
}
// w = x + beta*w            vector_operations.F90:148-151
__global__ void k_wxbw(int n, const double *__restrict__ x, double beta, double *__restrict__ w)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double t = beta * w[i];
        w[i] = x[i] + t;
    }
}

// plain copy: the HBM calibration kernel (16-byte accesses)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy16(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, int64_t n16)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = __builtin_nontemporal_load(src + i);
}

// ---- device generate_matrix ---------------------------------------------------------------
// nnz of all rows before coordinate m along one axis of extent n (1-D neighbour counts are 3,
// except 2 at each end, and 1 when n == 1).
__device__ __forceinline__ int64_t axis_prefix(int64_t m, int64_t n)
{
    return 3 * m - (m >= 1 ? 1 : 0) - (m >= n ? 1 : 0);
}
__device__ __forceinline__ int64_t axis_count(int64_t i, int64_t n)
{
    return axis_prefix(i + 1, n) - axis_prefix(i, n);
}

// One thread per row; same row order, column order and values as SparseMatrix.F90:138-175.
__global__ void k_generate(int nx, int ny, int nz, int32_t *__restrict__ rowindex,
                           int32_t *__restrict__ colindex, double *__restrict__ values,
                           double *__restrict__ x, double *__restrict__ b, double *__restrict__ xexact)
{
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (row >= N) return;
    const int ix = (int)(row % nx), iy = (int)((row / nx) % ny), iz = (int)(row / ((int64_t)nx * ny));
    const int64_t cx = 3 * (int64_t)nx - 2, cy = 3 * (int64_t)ny - 2;
    const int64_t cursor = axis_prefix(iz, nz) * cy * cx + axis_count(iz, nz) * axis_prefix(iy, ny) * cx +
                           axis_count(iz, nz) * axis_count(iy, ny) * axis_prefix(ix, nx);
    rowindex[row] = (int32_t)(cursor + 1);
    int64_t c = cursor;
    for (int sz = -1; sz <= 1; ++sz) {
        if (iz + sz < 0 || iz + sz >= nz) continue;
        for (int sy = -1; sy <= 1; ++sy) {
            if (iy + sy < 0 || iy + sy >= ny) continue;
            for (int sx = -1; sx <= 1; ++sx) {
                if (ix + sx < 0 || ix + sx >= nx) continue;
                const int64_t curcol = row + (int64_t)sz * nx * ny + (int64_t)sy * nx + sx;
                values[c] = (curcol == row) ? 27.0 : -1.0;
                colindex[c] = (int32_t)(curcol + 1);
                ++c;
            }
        }
    }
    const int nnzrow = (int)(c - cursor);
    x[row] = 0.0;
    b[row] = 27.0 - (double)(nnzrow - 1);
    xexact[row] = 1.0;
    if (row == N - 1) rowindex[N] = (int32_t)(c + 1);
}

// ============================================================================================
// host side
// ============================================================================================
static int vec_grid(const hpccg_ctx *c, int64_t n) { return vec_grid_for(c->info.num_cu, n); }

static int ensure_workspace(hpccg_ctx *c, int64_t n)
{
    if (c->ws_n >= n && c->r) return 0;
    dfree(c->r);
    dfree(c->p);
    dfree(c->Ap);
    c->r = c->p = c->Ap = nullptr;
    c->ws_n = 0;
    if (dmalloc(&c->r, n * sizeof(double)) != hipSuccess || dmalloc(&c->p, n * sizeof(double)) != hipSuccess ||
        dmalloc(&c->Ap, n * sizeof(double)) != hipSuccess) {
        set_error("ensure_workspace: hipMalloc of %lld doubles failed", (long long)(3 * n));
        return HPCCG_ERR_NOMEM;
    }
    c->ws_n = n;
    return 0;
}

static int ensure_stage(hpccg_ctx *c, int64_t n)
{
    if (c->stage_n >= n && c->s0) return 0;
    dfree(c->s0);
    dfree(c->s1);
    dfree(c->s2);
    c->s0 = c->s1 = c->s2 = nullptr;
    c->stage_n = 0;
    if (dmalloc(&c->s0, n * sizeof(double)) != hipSuccess || dmalloc(&c->s1, n * sizeof(double)) != hipSuccess ||
        dmalloc(&c->s2, n * sizeof(double)) != hipSuccess) {
        set_error("ensure_stage: hipMalloc failed");
        return HPCCG_ERR_NOMEM;
    }
    c->stage_n = n;
    return 0;
}

static int ensure_hist(hpccg_ctx *c, int need)
{
    if (c->hist_cap >= need) return 0;
    dfree(c->rr_hist);
    c->rr_hist = nullptr;
    c->hist_cap = 0;
    if (dmalloc(&c->rr_hist, (size_t)need * sizeof(double)) != hipSuccess) {
        set_error("ensure_hist: hipMalloc failed");
        return HPCCG_ERR_NOMEM;
    }
    c->hist_cap = need;
    return 0;
}

static int matrix_required(const hpccg_ctx *c)
{
    if (!c->A.rowindex) {
        set_error("no matrix: call hpccg_matrix_upload or hpccg_generate_matrix first");
        return HPCCG_ERR_STATE;
    }
    return 0;
}

int launch_spmv(hpccg_ctx *c, const double *x, double *y)
{
    if (c->A.nslices == 0) return 0;
    hipLaunchKernelGGL(k_spmv, dim3(sell_grid(c->A)), dim3(kSellThreads), 0, c->stream, sell64(c->A), x, y);
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

static int launch_dot(hpccg_ctx *c, int n, const double *x, const double *y, double *out_dev)
{
    if (c->dot_mode == HPCCG_DOT_SEQUENTIAL) {
        hipLaunchKernelGGL(k_dot_serial, dim3(1), dim3(kWave), 0, c->stream, x, y, n, out_dev);
    } else {
        const int G = vec_grid(c, n);
        hipLaunchKernelGGL(k_dot_tree, dim3(G), dim3(kVecThreads), 0, c->stream, x, y, n, c->partials);
        hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(kFinishThreads), 0, c->stream, c->partials, G, out_dev);
    }
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

// Workspace of one FP64 CG solve: the vectors r, p, Ap, the control block, the r.r history
// (max_iter + 2 entries) and the partial sums of the SpMV (sell_grid(A) entries) and of the
// vector kernels (kMaxGrid entries).
struct CgWs {
    double *r, *p, *Ap;
    CgState *state;
    double *rr_hist;
    double *vec_part;
    double *spmv_part;
};
static CgWs ctx_ws(hpccg_ctx *c) { return CgWs{c->r, c->p, c->Ap, c->state, c->rr_hist, c->partials, c->A.red_part}; }

// Ap = A p, and in TREE mode p.Ap -> alpha, exactly as the CG loop launches it
static void launch_cg_spmv(hpccg_ctx *c, const CgWs &w, bool tree, const double *p, double *Ap)
{
    const Matrix &A = c->A;
    const int G = sell_grid(A);
    if (tree) {
        hipLaunchKernelGGL(k_cg_spmv<true>, dim3(G), dim3(kSellThreads), 0, c->stream, sell64(A), p, Ap, w.state,
                           w.spmv_part);
        hipLaunchKernelGGL(k_cg_pap_finish, dim3(1), dim3(kFinishThreads), 0, c->stream, w.spmv_part, G, w.state,
                           w.rr_hist);
    } else {
        hipLaunchKernelGGL(k_cg_spmv<false>, dim3(G), dim3(kSellThreads), 0, c->stream, sell64(A), p, Ap, w.state,
                           w.spmv_part);
    }
}

// Enqueue `count` CG iterations (each skips itself on the device once the loop has ended).
static int enqueue_iterations(hpccg_ctx *c, const CgWs &w, double *x, int count)
{
    const int n = c->A.nrow;
    const int Gv = vec_grid(c, n);
    const bool tree = c->dot_mode == HPCCG_DOT_TREE;
    for (int it = 0; it < count; ++it) {
        hipLaunchKernelGGL(k_cg_update_p, dim3(Gv), dim3(kVecThreads), 0, c->stream, w.r, w.p, x, n, w.state);
        launch_cg_spmv(c, w, tree, w.p, w.Ap);
        if (tree) {
            hipLaunchKernelGGL(k_cg_update_r<true>, dim3(Gv), dim3(kVecThreads), 0, c->stream, w.r, w.Ap, n, w.state,
                               w.vec_part);
            hipLaunchKernelGGL(k_cg_rr_finish<false>, dim3(1), dim3(kFinishThreads), 0, c->stream, w.vec_part, Gv,
                               w.state, w.rr_hist);
        } else {
            hipLaunchKernelGGL(k_cg_serial_pap, dim3(1), dim3(kWave), 0, c->stream, w.p, w.Ap, n, w.state, w.rr_hist);
            hipLaunchKernelGGL(k_cg_update_r<false>, dim3(Gv), dim3(kVecThreads), 0, c->stream, w.r, w.Ap, n,
                               w.state, w.vec_part);
            hipLaunchKernelGGL(k_cg_serial_rr, dim3(1), dim3(kWave), 0, c->stream, w.r, n, w.state, w.rr_hist, 0);
        }
    }
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

// Prepare (numerics.F90:92-99): Ap = A x; r = b - Ap; rtrans = r.r (p is formed at k = 1)
static int enqueue_prepare(hpccg_ctx *c, const CgWs &w, const double *b, const double *x)
{
    const int n = c->A.nrow, Gv = vec_grid(c, n);
    // x = 0 (the usual start) needs no A x: k_nonzero sets the control block's flag otherwise
    hipLaunchKernelGGL(k_nonzero<>, dim3(Gv), dim3(kVecThreads), 0, c->stream, x, (int64_t)n, &w.state->x_nonzero);
    if (c->A.nslices > 0)
        hipLaunchKernelGGL(k_spmv_prep, dim3(sell_grid(c->A)), dim3(kSellThreads), 0, c->stream, sell64(c->A), x, w.Ap,
                           w.state);
    if (c->dot_mode == HPCCG_DOT_TREE) {
        hipLaunchKernelGGL(k_cg_prep_r<true>, dim3(Gv), dim3(kVecThreads), 0, c->stream, b, w.Ap, w.r, n, w.vec_part,
                           w.state);
        hipLaunchKernelGGL(k_cg_rr_finish<true>, dim3(1), dim3(kFinishThreads), 0, c->stream, w.vec_part, Gv, w.state,
                           w.rr_hist);
    } else {
        hipLaunchKernelGGL(k_cg_prep_r<false>, dim3(Gv), dim3(kVecThreads), 0, c->stream, b, w.Ap, w.r, n,
                           w.vec_part, w.state);
        hipLaunchKernelGGL(k_cg_serial_rr, dim3(1), dim3(kWave), 0, c->stream, w.r, n, w.state, w.rr_hist, 1);
    }
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

static int get_chunk_graph(hpccg_ctx *c, const double *b, double *x, int chunk, hipGraphExec_t *out)
{
    const CgWs w = ctx_ws(c);
    return cached_graph(c->graphs, c->stream, GraphKey{b, x, c->dot_mode, chunk},
                        [&]() { return enqueue_iterations(c, w, x, chunk); }, out);
}

// The solve on device-resident b and x with workspace w (allocated by the caller).  GRAPHS:
// the iterations replay captured chunk graphs (cached per b, x); otherwise every kernel is
// launched directly, so the solve allocates nothing at all.  *elapsed_ms: Prepare + loop +
// final x update; *loop_ms: the iteration loop with its final x update only (the span of the
// reference's "Total" timer, numerics.F90:106-150).
static int cg_run(hpccg_ctx *c, const CgWs &w, bool graphs, const double *b, double *x, int max_iter, double tol,
                  int *niters, double *normr, double *hist, double *elapsed_ms, double *loop_ms,
                  bool ranges = false)
{
    const int n = c->A.nrow;
    const bool tree = c->dot_mode == HPCCG_DOT_TREE;
    CgState init{};
    init.tol = tol;
    init.k = 0;
    init.stop_at = max_iter + 1;
    init.max_iter = max_iter;
    *c->state_h = init;
    HPCCG_HIP_TRY(hipMemcpyAsync(w.state, c->state_h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));

    const int cap = tree ? 16 : 4;
    auto get = [&](int size, hipGraphExec_t *ge) { return get_chunk_graph(c, b, x, size, ge); };
    if (graphs) HPCCG_TRY(capture_iteration_graphs(cap, max_iter, get));
    auto launch = [&](int m) { return graphs ? launch_iterations(c->stream, m, get) : enqueue_iterations(c, w, x, m); };

    if (ranges) roctxRangePushA("Prepare");
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HPCCG_TRY(enqueue_prepare(c, w, b, x));
    HPCCG_HIP_TRY(hipEventRecord(c->evp, c->stream));
    if (ranges) {
        // the host ranges bracket the device work: wait for Prepare before closing its range
        const hipError_t e = hipEventSynchronize(c->evp);
        roctxRangePop();
        roctxRangePushA("Total");
        HPCCG_HIP_TRY(e);
    }

    // Iteration loop ("Total" timer, numerics.F90:106-150): chunks of up to 16 (tree) or 4
    // (sequential dots) iterations.  With a tolerance the host reads the state after each
    // launch and sizes the next one; the reference's lagging test runs one iteration after
    // ||r|| < tol.  Without one, all max_iter iterations are enqueued at once.
    const bool can_stop_early = tol > 0.0 || std::isnan(tol);
    ChunkPlanner plan{cap, tol, 1};
    int launched = 0, m = std::min(cap, max_iter);
    while (launched < max_iter) {
        HPCCG_TRY(launch(m));
        launched += m;
        if (launched >= max_iter) break;
        if (can_stop_early) {
            HPCCG_HIP_TRY(hipMemcpyAsync(c->state_h, w.state, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
            HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
            const CgState &sh = *c->state_h;
            if (sh.k >= sh.stop_at) break;
            m = plan.next(sh.rr, sh.k - 1, max_iter - launched);
        } else {
            m = std::min(cap, max_iter - launched);
        }
    }
    hipLaunchKernelGGL(k_cg_finish_x, dim3(vec_grid(c, n)), dim3(kVecThreads), 0, c->stream, w.p, x, n, w.state);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->state_h, w.state, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    const hipError_t esync = hipStreamSynchronize(c->stream);
    if (ranges) roctxRangePop();
    HPCCG_HIP_TRY(esync);
    if (elapsed_ms || loop_ms) {
        float ms = 0.f;
        HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (elapsed_ms) *elapsed_ms = ms;
        HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->evp, c->ev1));
        if (loop_ms) *loop_ms = ms;
    }
    const CgState s = *c->state_h;
    const int nit = s.k - 1;
    if (niters) *niters = nit;
    // normr = ||r_{niters-1}|| (numerics.F90:122); with max_iter == 0 it stays ||r_0||
    const int last = std::max(0, nit - 1);
    double rr_last = 0.0;
    if (hist) {
        HPCCG_HIP_TRY(hipMemcpyAsync(hist, w.rr_hist, (max_iter + 1) * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        rr_last = hist[last];
        for (int j = 0; j <= max_iter; ++j) hist[j] = (j <= nit) ? std::sqrt(hist[j]) : -1.0;
    } else {
        HPCCG_HIP_TRY(hipMemcpyAsync(c->scalar_h, w.rr_hist + last, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        rr_last = *c->scalar_h;
    }
    if (normr) *normr = std::sqrt(rr_last);
    return s.breakdown ? HPCCG_WARN_BREAKDOWN : 0;
}

// cg with the reference's timer regions (numerics.F90:91-156), iterations launched one at a
// time: HIP events bracket Prepare, every iteration and its SpMV, roctx ranges of the same
// names bracket the same device work (the host waits at the end of each), and the host tests
// the stopping rule before each iteration as the reference does (numerics.F90:109).  Same
// kernels in the same order as cg_run, so the same results bit for bit.  Times in seconds;
// t_iter / t_spmv hold max_iter entries (k-th iteration at [k-1]).
static int cg_timed_iterations(hpccg_ctx *c, const double *b, double *x, int max_iter, double tol, int *niters,
                               double *normr, double *hist, double *t_total, double *t_prepare, double *t_iter,
                               double *t_spmv)
{
    const int n = c->A.nrow, Gv = vec_grid(c, n), G = sell_grid(c->A);
    const bool tree = c->dot_mode == HPCCG_DOT_TREE;
    const CgWs w = ctx_ws(c);
    CgState init{};
    init.tol = tol;
    init.stop_at = max_iter + 1;
    init.max_iter = max_iter;
    *c->state_h = init;
    HPCCG_HIP_TRY(hipMemcpyAsync(w.state, c->state_h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    float ms = 0.f;
    roctxRangePushA("Prepare");
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HPCCG_TRY(enqueue_prepare(c, w, b, x));
    HPCCG_HIP_TRY(hipEventRecord(c->evp, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->state_h, w.state, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    hipError_t e = hipStreamSynchronize(c->stream);
    roctxRangePop();
    HPCCG_HIP_TRY(e);
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->evp));
    if (t_prepare) *t_prepare = ms * 1e-3;
    roctxRangePushA("Total");
    for (int k = 1; k <= max_iter && c->state_h->k < c->state_h->stop_at; ++k) {
        roctxRangePushA("Iteration");
        HPCCG_HIP_TRY(hipEventRecord(c->ev2, c->stream));
        hipLaunchKernelGGL(k_cg_update_p, dim3(Gv), dim3(kVecThreads), 0, c->stream, w.r, w.p, x, n, w.state);
        roctxRangePushA("spmv");
        HPCCG_HIP_TRY(hipEventRecord(c->evs0, c->stream));
        if (tree)
            hipLaunchKernelGGL(k_cg_spmv<true>, dim3(G), dim3(kSellThreads), 0, c->stream, sell64(c->A), w.p, w.Ap,
                               w.state, w.spmv_part);
        else
            hipLaunchKernelGGL(k_cg_spmv<false>, dim3(G), dim3(kSellThreads), 0, c->stream, sell64(c->A), w.p, w.Ap,
                               w.state, w.spmv_part);
        HPCCG_HIP_TRY(hipEventRecord(c->evs1, c->stream));
        e = hipEventSynchronize(c->evs1);
        roctxRangePop();
        HPCCG_HIP_TRY(e);
        if (tree) {
            hipLaunchKernelGGL(k_cg_pap_finish, dim3(1), dim3(kFinishThreads), 0, c->stream, w.spmv_part, G, w.state,
                               w.rr_hist);
            hipLaunchKernelGGL(k_cg_update_r<true>, dim3(Gv), dim3(kVecThreads), 0, c->stream, w.r, w.Ap, n, w.state,
                               w.vec_part);
            hipLaunchKernelGGL(k_cg_rr_finish<false>, dim3(1), dim3(kFinishThreads), 0, c->stream, w.vec_part, Gv,
                               w.state, w.rr_hist);
        } else {
            hipLaunchKernelGGL(k_cg_serial_pap, dim3(1), dim3(kWave), 0, c->stream, w.p, w.Ap, n, w.state, w.rr_hist);
            hipLaunchKernelGGL(k_cg_update_r<false>, dim3(Gv), dim3(kVecThreads), 0, c->stream, w.r, w.Ap, n,
                               w.state, w.vec_part);
            hipLaunchKernelGGL(k_cg_serial_rr, dim3(1), dim3(kWave), 0, c->stream, w.r, n, w.state, w.rr_hist, 0);
        }
        HPCCG_HIP_TRY(hipGetLastError());
        HPCCG_HIP_TRY(hipEventRecord(c->ev3, c->stream));
        HPCCG_HIP_TRY(hipMemcpyAsync(c->state_h, w.state, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
        e = hipStreamSynchronize(c->stream);
        roctxRangePop();
        HPCCG_HIP_TRY(e);
        HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev2, c->ev3));
        if (t_iter) t_iter[k - 1] = ms * 1e-3;
        HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->evs0, c->evs1));
        if (t_spmv) t_spmv[k - 1] = ms * 1e-3;
    }
    hipLaunchKernelGGL(k_cg_finish_x, dim3(Gv), dim3(kVecThreads), 0, c->stream, w.p, x, n, w.state);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    e = hipStreamSynchronize(c->stream);
    roctxRangePop();
    HPCCG_HIP_TRY(e);
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->evp, c->ev1));
    if (t_total) *t_total = ms * 1e-3;
    const CgState s = *c->state_h;
    const int nit = s.k - 1;
    if (niters) *niters = nit;
    std::vector<double> rr(max_iter + 1);
    HPCCG_HIP_TRY(hipMemcpy(rr.data(), w.rr_hist, (max_iter + 1) * sizeof(double), hipMemcpyDeviceToHost));
    if (normr) *normr = std::sqrt(rr[std::max(0, nit - 1)]);
    if (hist)
        for (int j = 0; j <= max_iter; ++j) hist[j] = (j <= nit) ? std::sqrt(rr[j]) : -1.0;
    return s.breakdown ? HPCCG_WARN_BREAKDOWN : 0;
}

// The whole solve on device-resident b and x, with the context's workspace.
static int cg_device(hpccg_ctx *c, const double *b, double *x, int max_iter, double tol, int *niters,
                     double *normr, double *hist, double *elapsed_ms, double *loop_ms = nullptr)
{
    HPCCG_TRY(matrix_required(c));
    HPCCG_CHECK_ARG(max_iter >= 0, "max_iter must be >= 0");
    HPCCG_TRY(ensure_workspace(c, c->A.nrow));
    HPCCG_TRY(ensure_hist(c, max_iter + 2));
    return cg_run(c, ctx_ws(c), true, b, x, max_iter, tol, niters, normr, hist, elapsed_ms, loop_ms);
}

// ---- no-allocation solve ----------------------------------------------------------------------
// The workspace of hpccg_cg_dev_noalloc lives in one file-scope device array, a copy of which
// every device that loads this library holds until the process ends.  Layout in doubles, every
// part 256-byte aligned: the control block (32), the r.r history (max iterations + 2), the
// vector kernels' partials (kMaxGrid), the SpMV's partials (sell_grid of the largest matrix),
// then r, p, Ap of HPCCG_NOALLOC_MAX_ROWS entries each: 50,561,280 bytes.  Host code never
// names the array except to ask for its address (scratch_setup); kernels get that address.
constexpr int64_t kNaRows = HPCCG_NOALLOC_MAX_ROWS;
constexpr int64_t kNaOffHist = 32;
constexpr int64_t kNaOffVec = kNaOffHist + (HPCCG_NOALLOC_MAX_ITER + 2);
constexpr int64_t kNaOffSpmv = kNaOffVec + kMaxGrid;
constexpr int64_t kNaSpmvPart = (kNaRows / kSlice + kSellWaves - 1) / kSellWaves;
constexpr int64_t kNaOffR = kNaOffSpmv + kNaSpmvPart;
constexpr int64_t kNaDoubles = kNaOffR + 3 * kNaRows;
static_assert(sizeof(CgState) <= 32 * sizeof(double), "control block fits its slot");
static_assert(kNaOffVec % 32 == 0 && kNaOffSpmv % 32 == 0 && kNaOffR % 32 == 0 && kNaRows % 32 == 0,
              "256-byte aligned parts");

static __device__ __align__(256) double scratch[kNaDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

// The device address of `device`'s copy of the workspace, looked up once per device; the
// caller has made `device` current.
static int scratch_setup(int device, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) {
        set_error("no-allocation workspace: device index %d beyond %d", device, kMaxDevices);
        return HPCCG_ERR_ARG;
    }
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("no-allocation workspace: hipGetSymbolAddress -> %s", hipGetErrorString(e));
            return HPCCG_ERR_HIP;
        }
        g_scratch_dev[device] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[device];
    return 0;
}

int cg_device_fp64(hpccg_ctx *c, const double *b, double *x, int max_iter, double tol, int *niters, double *normr,
                   double *hist, double *elapsed_ms)
{
    return cg_device(c, b, x, max_iter, tol, niters, normr, hist, elapsed_ms);
}

int cg_trace(hpccg_ctx *c, const double *b, double *x, int max_iter, double tol, const std::function<int(int)> &hook)
{
    HPCCG_TRY(matrix_required(c));
    const int n = c->A.nrow;
    HPCCG_TRY(ensure_workspace(c, n));
    HPCCG_TRY(ensure_hist(c, max_iter + 2));
    const int Gv = vec_grid(c, n);
    CgState init{};
    init.tol = tol;
    init.stop_at = max_iter + 1;
    init.max_iter = max_iter;
    *c->state_h = init;
    HPCCG_HIP_TRY(hipMemcpyAsync(c->state, c->state_h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    const CgWs w = ctx_ws(c);
    const int mode = c->dot_mode;
    c->dot_mode = HPCCG_DOT_TREE;
    const int prep = enqueue_prepare(c, w, b, x);
    c->dot_mode = mode;
    HPCCG_TRY(prep);
    // p_0 = r_0 and alpha = 0 make the hook's x + alpha p equal x_0
    HPCCG_HIP_TRY(hipMemcpyAsync(c->p, c->r, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    if (hook(0)) return 0;
    for (int it = 1; it <= max_iter; ++it) {
        HPCCG_HIP_TRY(hipMemcpyAsync(c->state_h, c->state, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->state_h->k >= c->state_h->stop_at) break;
        const int k = c->state_h->k;
        hipLaunchKernelGGL(k_cg_update_p, dim3(Gv), dim3(kVecThreads), 0, c->stream, c->r, c->p, x, n, c->state);
        launch_cg_spmv(c, w, true, c->p, c->Ap);
        hipLaunchKernelGGL(k_cg_update_r<true>, dim3(Gv), dim3(kVecThreads), 0, c->stream, c->r, c->Ap, n, c->state,
                           c->partials);
        hipLaunchKernelGGL(k_cg_rr_finish<false>, dim3(1), dim3(kFinishThreads), 0, c->stream, c->partials, Gv,
                           c->state, c->rr_hist);
        HPCCG_HIP_TRY(hipGetLastError());
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        if (hook(k)) break;
    }
    hipLaunchKernelGGL(k_cg_finish_x, dim3(Gv), dim3(kVecThreads), 0, c->stream, c->p, x, n, c->state);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace hpccg

using namespace hpccg;

// ============================================================================================
// C ABI
// ============================================================================================
extern "C" {

const char *hpccg_last_error(void) { return g_err; }

int hpccg_create(hpccg_ctx **out, int device)
{
    HPCCG_CHECK_ARG(out, "null handle pointer");
    *out = nullptr;
    int count = 0;
    HPCCG_HIP_TRY(hipGetDeviceCount(&count));
    HPCCG_CHECK_ARG(device >= 0 && device < count, "device index out of range");
    HPCCG_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HPCCG_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("hpccg_create: device %d is %s; this library is built for gfx950 (MI355X) only", device,
                  prop.gcnArchName);
        return HPCCG_ERR_ARG;
    }
    auto *c = new hpccg_ctx();
    c->device = device;
    c->info.num_cu = prop.multiProcessorCount;
    c->info.gfx = 950;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        dmalloc(&c->state, sizeof(CgState)) != hipSuccess ||
        hipHostMalloc(&c->state_h, sizeof(CgState), hipHostMallocDefault) != hipSuccess ||
        dmalloc(&c->partials, kMaxGrid * sizeof(double)) != hipSuccess ||
        dmalloc(&c->scalar, 8 * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&c->scalar_h, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess ||
        hipEventCreate(&c->evp) != hipSuccess || hipEventCreate(&c->evs0) != hipSuccess ||
        hipEventCreate(&c->evs1) != hipSuccess) {
        set_error("hpccg_create: HIP resource allocation failed");
        hpccg_destroy(c);
        return HPCCG_ERR_HIP;
    }
    const int sst = scratch_setup(device, &c->scratch);
    if (sst) {
        hpccg_destroy(c);
        return sst;
    }
    // the control block starts zeroed (every solve rewrites it before use)
    CgState zero{};
    memcpy(c->state_h, &zero, sizeof(zero));
    if (hipMemcpy(c->state, c->state_h, sizeof(CgState), hipMemcpyHostToDevice) != hipSuccess) {
        hpccg_destroy(c);
        set_error("hpccg_create: state init failed");
        return HPCCG_ERR_HIP;
    }
    *out = c;
    return 0;
}

void hpccg_free_ext(hpccg_ctx *c);  // mixed / batched / learned paths release their buffers
static void free_matrix(hpccg_ctx *c);

int hpccg_destroy(hpccg_ctx *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    free_matrix(c);  // also drops every captured graph
    hpccg_free_ext(c);
    dfree(c->r);
    dfree(c->p);
    dfree(c->Ap);
    dfree(c->state);
    hipHostFree(c->state_h);
    dfree(c->rr_hist);
    dfree(c->partials);
    dfree(c->scalar);
    hipHostFree(c->scalar_h);
    dfree(c->s0);
    dfree(c->s1);
    dfree(c->s2);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->ev2) hipEventDestroy(c->ev2);
    if (c->ev3) hipEventDestroy(c->ev3);
    if (c->evp) hipEventDestroy(c->evp);
    if (c->evs0) hipEventDestroy(c->evs0);
    if (c->evs1) hipEventDestroy(c->evs1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int hpccg_set_dot_mode(hpccg_ctx *c, int mode)
{
    HPCCG_CHECK_ARG(c, "null context");
    HPCCG_CHECK_ARG(mode == HPCCG_DOT_SEQUENTIAL || mode == HPCCG_DOT_TREE, "unknown dot mode");
    c->dot_mode = mode;
    return 0;
}

int hpccg_get_dot_mode(hpccg_ctx *c, int *mode)
{
    HPCCG_CHECK_ARG(c && mode, "null argument");
    *mode = c->dot_mode;
    return 0;
}

int hpccg_device_info(hpccg_ctx *c, int *num_cu, int *num_xcd, int *gfx)
{
    HPCCG_CHECK_ARG(c, "null context");
    if (num_cu) *num_cu = c->info.num_cu;
    if (num_xcd) *num_xcd = kNumXcd;
    if (gfx) *gfx = c->info.gfx;
    return 0;
}

int hpccg_dev_malloc(hpccg_ctx *c, void **ptr, int64_t bytes)
{
    HPCCG_CHECK_ARG(c && ptr && bytes >= 0, "bad argument");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    if (hipMalloc(ptr, std::max<int64_t>(bytes, 1)) != hipSuccess) {
        set_error("hpccg_dev_malloc: %lld bytes failed", (long long)bytes);
        return HPCCG_ERR_NOMEM;
    }
    return 0;
}

int hpccg_dev_free(hpccg_ctx *c, void *ptr)
{
    HPCCG_CHECK_ARG(c, "null context");
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    // caller-owned vectors are part of the graph keys, not of the buffer generation
    HPCCG_HIP_TRY(hipFree(ptr));
    return 0;
}

int hpccg_memcpy_h2d(hpccg_ctx *c, void *dst, const void *src, int64_t bytes)
{
    HPCCG_CHECK_ARG(c && bytes >= 0, "bad argument");
    HPCCG_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_memcpy_d2h(hpccg_ctx *c, void *dst, const void *src, int64_t bytes)
{
    HPCCG_CHECK_ARG(c && bytes >= 0, "bad argument");
    HPCCG_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_memcpy_d2d(hpccg_ctx *c, void *dst, const void *src, int64_t bytes)
{
    HPCCG_CHECK_ARG(c && bytes >= 0, "bad argument");
    HPCCG_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Every captured graph of every path holds the matrix addresses.  Freeing the matrix bumps the
// buffer generation, which already retires them; dropping them here releases their memory.
static void drop_graphs(hpccg_ctx *c)
{
    c->graphs.clear();
    c->mix.graphs.clear();
    invalidate_learned_graphs(c);
}

static void free_matrix(hpccg_ctx *c)
{
    drop_graphs(c);
    Matrix &A = c->A;
    for (void *p : {(void *)A.rowindex, (void *)A.colindex, (void *)A.values, (void *)A.slice_ptr, (void *)A.code,
                    (void *)A.ctab, (void *)A.scol64, (void *)A.sval64, (void *)A.scol16, (void *)A.sval16,
                    (void *)A.red_part})
        dfree(p);
    A = Matrix{};
}

// Distinct offsets (col - 1) - row of a 1-based CSR, collected into an open-addressed set of
// kOffsetSlots slots (initialised to kNoOffset), one thread per row.  *count is the number
// inserted.  Threads stop once it exceeds kMaxCodes (the copy then keeps int32 columns).
// Concurrent threads can insert past kMaxCodes before they see that, so a probe sequence is
// capped at kMaxProbe slots: with at most kMaxCodes entries in kOffsetSlots slots a longer
// one means an overfull set, and it raises *count to kOffsetSlots.  Every thread thus stops
// after a bounded number of loads, however many offsets a general matrix has.
constexpr int kMaxProbe = 64;
__global__ void k_column_offsets(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int nrow,
                                 int32_t *set, int *count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrow) return;
    const int k1 = rowptr[t + 1] - 1;
    for (int k = rowptr[t] - 1; k < k1; ++k) {
        if (__atomic_load_n(count, __ATOMIC_RELAXED) > kMaxCodes) return;
        const int32_t off = (col[k] - 1) - t;
        uint32_t h = ((uint32_t)off * 2654435761u) >> 20;
        int probe = 0;
        for (; probe < kMaxProbe; ++probe, h = (h + 1) & (kOffsetSlots - 1)) {
            int32_t cur = __atomic_load_n(set + h, __ATOMIC_RELAXED);
            if (cur == off) break;
            if (cur != kNoOffset) continue;
            if (__atomic_load_n(count, __ATOMIC_RELAXED) > kMaxCodes) return;
            cur = atomicCAS(set + h, kNoOffset, off);
            if (cur == kNoOffset) {
                atomicAdd(count, 1);
                break;
            }
            if (cur == off) break;
        }
        if (probe == kMaxProbe) {
            atomicMax(count, kOffsetSlots);
            return;
        }
    }
}

// The sorted table of the distinct offsets col - row of the device CSR, when there are at
// most kMaxCodes of them (tab empty otherwise).  Collected on the device in a hash set.
static hipError_t column_offset_table(hpccg_ctx *c, std::vector<int32_t> &tab)
{
    Matrix &A = c->A;
    tab.clear();
    if (A.nrow == 0 || A.nnz == 0) return hipSuccess;
    int32_t *set = nullptr;  // kOffsetSlots slots, then the count
    hipError_t e = dmalloc(&set, (kOffsetSlots + 1) * sizeof(int32_t));
    if (e != hipSuccess) return e;
    std::vector<int32_t> h(kOffsetSlots + 1);
    e = hipMemsetD32Async(set, kNoOffset, kOffsetSlots, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(set + kOffsetSlots, 0, sizeof(int32_t), c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_column_offsets, dim3((unsigned)((A.nrow + 255) / 256)), dim3(256), 0, c->stream,
                           A.rowindex, A.colindex, A.nrow, set, set + kOffsetSlots);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), set, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(set);
    if (e != hipSuccess || h[kOffsetSlots] > kMaxCodes) return e;
    for (int i = 0; i < kOffsetSlots; ++i)
        if (h[i] != kNoOffset) tab.push_back(h[i]);
    std::sort(tab.begin(), tab.end());
    if ((int)tab.size() > kMaxCodes) tab.clear();
    return hipSuccess;
}

// After the CSR is on the device: row statistics, the FP64 sliced-ELL copy every FP64 SpMV
// reads (column-coded when the matrix has at most kMaxCodes distinct offsets col - row), and
// the reduction scratch of the one-wave-per-slice kernels.
static int finish_matrix_setup(hpccg_ctx *c, const int32_t *rowindex_host)
{
    Matrix &A = c->A;
    const int n = A.nrow;
    int32_t mr = 0;
    for (int i = 0; i < n; ++i) mr = std::max(mr, rowindex_host[i + 1] - rowindex_host[i]);
    A.max_row_nnz = mr;
    drop_graphs(c);
    A.nslices = (n + kSlice - 1) / kSlice;
    std::vector<int64_t> sp(A.nslices + 1);
    A.slots = sell_slice_ptr(rowindex_host, n, 0, sp.data());
    std::vector<int32_t> tab;
    hipError_t e = column_offset_table(c, tab);
    if (e != hipSuccess) {
        free_matrix(c);
        set_error("matrix setup: %s", hipGetErrorString(e));
        return HPCCG_ERR_HIP;
    }
    const bool coded = !tab.empty();
    const int64_t cbytes = code_bytes(A.slots, A.nslices);
    const int G = sell_grid(A);
    int *flag = nullptr;
    if (dmalloc(&A.slice_ptr, sp.size() * sizeof(int64_t)) != hipSuccess ||
        (coded && dmalloc(&A.code, std::max<int64_t>(cbytes, 4)) != hipSuccess) ||
        (coded && dmalloc(&A.ctab, kMaxCodes * sizeof(int32_t)) != hipSuccess) ||
        (!coded && dmalloc(&A.scol64, std::max<int64_t>(A.slots, 1) * sizeof(int32_t)) != hipSuccess) ||
        dmalloc(&A.sval64, std::max<int64_t>(A.slots, 1) * sizeof(double)) != hipSuccess ||
        dmalloc(&A.red_part, (size_t)G * sizeof(double)) != hipSuccess || dmalloc(&flag, sizeof(int)) != hipSuccess) {
        dfree(flag);
        free_matrix(c);
        set_error("matrix setup: device allocation of the sliced copy failed (%lld slots)", (long long)sp.back());
        return HPCCG_ERR_NOMEM;
    }
    A.ncode = (int32_t)tab.size();
    e = hipMemcpyAsync(A.slice_ptr, sp.data(), sp.size() * sizeof(int64_t), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(int), c->stream);
    if (e == hipSuccess && coded) e = hipMemsetAsync(A.code, 0, std::max<int64_t>(cbytes, 4), c->stream);
    if (e == hipSuccess && coded)
        e = hipMemcpyAsync(A.ctab, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && A.nslices > 0) {
        const int64_t threads = (int64_t)A.nslices * kSlice;
        hipLaunchKernelGGL((k_csr_to_sell<kGroupF64, double>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                           c->stream, A.rowindex, A.colindex, A.values, A.slice_ptr, A.scol64, A.sval64, (int64_t)n,
                           A.nslices, flag, A.code, A.ctab, A.ncode);
        if (coded)
            hipLaunchKernelGGL(k_mark_uniform<>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c->stream,
                               A.slice_ptr, A.code, A.nslices);
        e = hipGetLastError();
    }
    int bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(flag);
    if (e != hipSuccess || bad) {
        free_matrix(c);
        if (e != hipSuccess) {
            set_error("matrix setup: %s", hipGetErrorString(e));
            return HPCCG_ERR_HIP;
        }
        set_error("matrix setup: a column offset is missing from the code table");
        return HPCCG_ERR_HIP;
    }
    return 0;
}

int hpccg_matrix_upload(hpccg_ctx *c, int32_t nrow, int32_t ncol, const int32_t *rowindex,
                        const int32_t *colindex, const double *values)
{
    HPCCG_CHECK_ARG(c && rowindex && nrow >= 0, "bad argument");
    HPCCG_CHECK_ARG(ncol == nrow, "single-rank matrix: ncol must equal nrow");
    HPCCG_CHECK_ARG(nrow == 0 || rowindex[0] == 1, "rowindex must be 1-based (rowindex[0] == 1)");
    const int64_t nnz = (int64_t)rowindex[nrow] - 1;
    HPCCG_CHECK_ARG(nnz >= 0, "rowindex[nrow] must be >= 1");
    HPCCG_CHECK_ARG(nnz == 0 || (colindex && values), "null colindex/values");
    for (int i = 0; i < nrow; ++i)
        HPCCG_CHECK_ARG(rowindex[i + 1] >= rowindex[i], "rowindex must be non-decreasing");
    for (int64_t k = 0; k < nnz; ++k)
        HPCCG_CHECK_ARG(colindex[k] >= 1 && colindex[k] <= ncol, "colindex out of range 1..ncol");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    free_matrix(c);
    c->A.nrow = nrow;
    c->A.ncol = ncol;
    c->A.nnz = nnz;
    if (dmalloc(&c->A.rowindex, (nrow + 1) * sizeof(int32_t)) != hipSuccess ||
        dmalloc(&c->A.colindex, std::max<int64_t>(nnz, 1) * sizeof(int32_t)) != hipSuccess ||
        dmalloc(&c->A.values, std::max<int64_t>(nnz, 1) * sizeof(double)) != hipSuccess) {
        free_matrix(c);
        set_error("hpccg_matrix_upload: device allocation failed (nnz=%lld)", (long long)nnz);
        return HPCCG_ERR_NOMEM;
    }
    HPCCG_HIP_TRY(hipMemcpyAsync(c->A.rowindex, rowindex, (nrow + 1) * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    if (nnz) {
        HPCCG_HIP_TRY(hipMemcpyAsync(c->A.colindex, colindex, nnz * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        HPCCG_HIP_TRY(hipMemcpyAsync(c->A.values, values, nnz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return finish_matrix_setup(c, rowindex);
}

int hpccg_generate_matrix(hpccg_ctx *c, int32_t nx, int32_t ny, int32_t nz, int32_t *rowindex,
                          int32_t *colindex, double *values, double *x, double *b, double *xexact,
                          int64_t *nnz_out)
{
    HPCCG_CHECK_ARG(c && nx >= 1 && ny >= 1 && nz >= 1, "grid extents must be >= 1");
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t nnz = (int64_t)(3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
    HPCCG_CHECK_ARG(N < INT32_MAX && nnz + 1 < INT32_MAX, "grid too large for int32 CSR");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    free_matrix(c);
    c->A.nrow = c->A.ncol = (int32_t)N;
    c->A.nnz = nnz;
    c->A.grid[0] = nx;
    c->A.grid[1] = ny;
    c->A.grid[2] = nz;
    double *dx = nullptr, *db = nullptr, *dxe = nullptr;
    if (dmalloc(&c->A.rowindex, (N + 1) * sizeof(int32_t)) != hipSuccess ||
        dmalloc(&c->A.colindex, nnz * sizeof(int32_t)) != hipSuccess ||
        dmalloc(&c->A.values, nnz * sizeof(double)) != hipSuccess ||
        dmalloc(&dx, N * sizeof(double)) != hipSuccess || dmalloc(&db, N * sizeof(double)) != hipSuccess ||
        dmalloc(&dxe, N * sizeof(double)) != hipSuccess) {
        dfree(dx);
        dfree(db);
        dfree(dxe);
        free_matrix(c);
        set_error("hpccg_generate_matrix: device allocation failed");
        return HPCCG_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_generate, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->stream, nx, ny, nz,
                       c->A.rowindex, c->A.colindex, c->A.values, dx, db, dxe);
    hipError_t e = hipGetLastError();
    std::vector<int32_t> rowh(N + 1);
    if (e == hipSuccess)
        e = hipMemcpyAsync(rowh.data(), c->A.rowindex, (N + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && colindex)
        e = hipMemcpyAsync(colindex, c->A.colindex, nnz * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && values)
        e = hipMemcpyAsync(values, c->A.values, nnz * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && x) e = hipMemcpyAsync(x, dx, N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && b) e = hipMemcpyAsync(b, db, N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && xexact) e = hipMemcpyAsync(xexact, dxe, N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(dx);
    dfree(db);
    dfree(dxe);
    if (e != hipSuccess) {
        set_error("hpccg_generate_matrix: %s", hipGetErrorString(e));
        return HPCCG_ERR_HIP;
    }
    if (rowindex) memcpy(rowindex, rowh.data(), (N + 1) * sizeof(int32_t));
    if (nnz_out) *nnz_out = nnz;
    return finish_matrix_setup(c, rowh.data());
}

int hpccg_matrix_info(hpccg_ctx *c, int32_t *nrow, int64_t *nnz, int32_t *max_row_nnz)
{
    HPCCG_CHECK_ARG(c, "null context");
    if (nrow) *nrow = c->A.nrow;
    if (nnz) *nnz = c->A.nnz;
    if (max_row_nnz) *max_row_nnz = c->A.max_row_nnz;
    return 0;
}

int hpccg_matrix_format(hpccg_ctx *c, int32_t *ncodes, int64_t *copy_bytes)
{
    HPCCG_CHECK_ARG(c, "null context");
    const Matrix &A = c->A;
    if (ncodes) *ncodes = A.ncode;
    if (copy_bytes)
        *copy_bytes = A.slots * (int64_t)sizeof(double) +
                      (A.code ? code_bytes(A.slots, A.nslices) : A.slots * (int64_t)sizeof(int32_t)) +
                      (A.nrow ? (A.nslices + 1) * (int64_t)sizeof(int64_t) : 0);
    return 0;
}

// ---- device-array kernels ------------------------------------------------------------------
int hpccg_spmv_dev(hpccg_ctx *c, const double *x, double *y)
{
    HPCCG_CHECK_ARG(c && x && y, "null argument");
    HPCCG_TRY(matrix_required(c));
    HPCCG_TRY(launch_spmv(c, x, y));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_ddot_dev(hpccg_ctx *c, int32_t n, const double *x, const double *y, double *r)
{
    HPCCG_CHECK_ARG(c && x && y && r && n >= 0, "bad argument");
    if (n == 0) {
        *r = 0.0;
        return 0;
    }
    HPCCG_TRY(launch_dot(c, n, x, y, c->scalar));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->scalar_h, c->scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    *r = *c->scalar_h;
    return 0;
}

static int vec_launch_grid(hpccg_ctx *c, int n) { return std::max(1, std::min((n + 255) / 256, c->info.num_cu * 8)); }

int hpccg_compute_residual_dev(hpccg_ctx *c, int32_t n, const double *v1, const double *v2, double *residual)
{
    HPCCG_CHECK_ARG(c && residual && n >= 0 && (n == 0 || (v1 && v2)), "bad argument");
    if (n == 0) {
        *residual = 0.0;
        return 0;
    }
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    HPCCG_HIP_TRY(hipMemsetAsync(c->scalar, 0, sizeof(double), c->stream));
    hipLaunchKernelGGL(k_maxabs_diff, dim3(vec_launch_grid(c, n)), dim3(kVecThreads), 0, c->stream, v1, v2, n,
                       (unsigned long long *)c->scalar);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipMemcpyAsync(c->scalar_h, c->scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    *residual = *c->scalar_h;
    return 0;
}

int hpccg_waxpby_dev(hpccg_ctx *c, int32_t n, double alpha, const double *x, double beta, const double *y, double *w)
{
    HPCCG_CHECK_ARG(c && n >= 0 && (n == 0 || (x && y && w)), "bad argument");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_waxpby, dim3(vec_launch_grid(c, n)), dim3(256), 0, c->stream, n, alpha, x, beta, y, w);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_wwpby_dev(hpccg_ctx *c, int32_t n, double beta, const double *y, double *w)
{
    HPCCG_CHECK_ARG(c && n >= 0 && (n == 0 || (y && w)), "bad argument");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_wwpby, dim3(vec_launch_grid(c, n)), dim3(256), 0, c->stream, n, beta, y, w);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_wxbw_dev(hpccg_ctx *c, int32_t n, const double *x, double beta, double *w)
{
    HPCCG_CHECK_ARG(c && n >= 0 && (n == 0 || (x && w)), "bad argument");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_wxbw, dim3(vec_launch_grid(c, n)), dim3(256), 0, c->stream, n, x, beta, w);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_cg_dev(hpccg_ctx *c, const double *b, double *x, int32_t max_iter, double tolerance, int32_t *niters,
                 double *normr, double *hist, double *elapsed_ms)
{
    HPCCG_CHECK_ARG(c && b && x, "null argument");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    return cg_device(c, b, x, max_iter, tolerance, niters, normr, hist, elapsed_ms);
}

// No reference counterpart: hpccg_cg_dev for callers that cannot allocate device memory.  The
// workspace is this device's static array (scratch_setup); the kernels, their order and the
// result are hpccg_cg_dev's, but every kernel is launched directly (no graph is captured or
// instantiated) and nothing is allocated.
int hpccg_cg_dev_noalloc(hpccg_ctx *c, const double *b, double *x, int32_t max_iter, double tolerance,
                         int32_t *niters, double *normr, double *hist, double *elapsed_ms)
{
    HPCCG_CHECK_ARG(c && b && x, "null argument");
    HPCCG_TRY(matrix_required(c));
    if (c->A.nrow > HPCCG_NOALLOC_MAX_ROWS) {
        set_error("hpccg_cg_dev_noalloc: %d rows exceed the static workspace (at most %d rows)", c->A.nrow,
                  (int)HPCCG_NOALLOC_MAX_ROWS);
        return HPCCG_ERR_ARG;
    }
    if (max_iter < 0 || max_iter > HPCCG_NOALLOC_MAX_ITER) {
        set_error("hpccg_cg_dev_noalloc: max_iter %d outside 0..%d (static workspace)", max_iter,
                  (int)HPCCG_NOALLOC_MAX_ITER);
        return HPCCG_ERR_ARG;
    }
    if (!c->scratch) {
        set_error("hpccg_cg_dev_noalloc: no workspace address for device %d", c->device);
        return HPCCG_ERR_STATE;
    }
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    double *s = c->scratch;
    const CgWs w{s + kNaOffR, s + kNaOffR + kNaRows, s + kNaOffR + 2 * kNaRows, reinterpret_cast<CgState *>(s),
                 s + kNaOffHist, s + kNaOffVec, s + kNaOffSpmv};
    return cg_run(c, w, false, b, x, max_iter, tolerance, niters, normr, hist, elapsed_ms, nullptr);
}

// ---- host-array drop-ins ---------------------------------------------------------------------
int hpccg_spmv(hpccg_ctx *c, const double *x, double *y)
{
    HPCCG_CHECK_ARG(c && x && y, "null argument");
    HPCCG_TRY(matrix_required(c));
    const int64_t n = c->A.nrow;
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_TRY(launch_spmv(c, c->s0, c->s1));
    HPCCG_HIP_TRY(hipMemcpyAsync(y, c->s1, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_ddot(hpccg_ctx *c, int32_t n, const double *x, const double *y, double *r)
{
    HPCCG_CHECK_ARG(c && r && n >= 0 && (n == 0 || (x && y)), "bad argument");
    if (n == 0) {
        *r = 0.0;
        return 0;
    }
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s1, y, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return hpccg_ddot_dev(c, n, c->s0, c->s1, r);
}

int hpccg_ddot1(hpccg_ctx *c, int32_t n, const double *x, double *r)
{
    HPCCG_CHECK_ARG(c && r && n >= 0 && (n == 0 || x), "bad argument");
    if (n == 0) {
        *r = 0.0;
        return 0;
    }
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return hpccg_ddot_dev(c, n, c->s0, c->s0, r);
}

int hpccg_compute_residual(hpccg_ctx *c, int32_t n, const double *v1, const double *v2, double *residual)
{
    HPCCG_CHECK_ARG(c && residual && n >= 0 && (n == 0 || (v1 && v2)), "bad argument");
    if (n == 0) {
        *residual = 0.0;
        return 0;
    }
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, v1, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s1, v2, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return hpccg_compute_residual_dev(c, n, c->s0, c->s1, residual);
}

int hpccg_waxpby(hpccg_ctx *c, int32_t n, double alpha, const double *x, double beta, const double *y, double *w)
{
    HPCCG_CHECK_ARG(c && n >= 0 && (n == 0 || (x && y && w)), "bad argument");
    if (n == 0) return 0;
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s1, y, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_TRY(hpccg_waxpby_dev(c, n, alpha, c->s0, beta, c->s1, c->s2));
    return hpccg_memcpy_d2h(c, w, c->s2, n * sizeof(double));
}

int hpccg_wwpby(hpccg_ctx *c, int32_t n, double beta, const double *y, double *w)
{
    HPCCG_CHECK_ARG(c && n >= 0 && (n == 0 || (y && w)), "bad argument");
    if (n == 0) return 0;
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, y, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s2, w, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_TRY(hpccg_wwpby_dev(c, n, beta, c->s0, c->s2));
    return hpccg_memcpy_d2h(c, w, c->s2, n * sizeof(double));
}

int hpccg_wxbw(hpccg_ctx *c, int32_t n, const double *x, double beta, double *w)
{
    HPCCG_CHECK_ARG(c && n >= 0 && (n == 0 || (x && w)), "bad argument");
    if (n == 0) return 0;
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s2, w, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_TRY(hpccg_wxbw_dev(c, n, c->s0, beta, c->s2));
    return hpccg_memcpy_d2h(c, w, c->s2, n * sizeof(double));
}

int hpccg_vcopy(hpccg_ctx *c, int32_t n, const double *x, double *y)
{
    HPCCG_CHECK_ARG(c && n >= 0 && (n == 0 || (x && y)), "bad argument");
    if (n == 0) return 0;
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s1, c->s0, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return hpccg_memcpy_d2h(c, y, c->s1, n * sizeof(double));
}

int hpccg_cg(hpccg_ctx *c, const double *b, double *x, int32_t max_iter, double tolerance, int32_t *niters,
             double *normr, double *hist)
{
    HPCCG_CHECK_ARG(c && b && x, "null argument");
    HPCCG_TRY(matrix_required(c));
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    const int64_t n = c->A.nrow;
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, b, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s1, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const int st = cg_device(c, c->s0, c->s1, max_iter, tolerance, niters, normr, hist, nullptr);
    if (st < 0) return st;
    HPCCG_TRY(hpccg_memcpy_d2h(c, x, c->s1, n * sizeof(double)));
    return st;
}

int hpccg_cg_timed(hpccg_ctx *c, const double *b, double *x, int32_t max_iter, double tolerance, int32_t level,
                   int32_t *niters, double *normr, double *hist, double *t_total, double *t_prepare, double *t_iter,
                   double *t_spmv)
{
    HPCCG_CHECK_ARG(c && b && x, "null argument");
    HPCCG_CHECK_ARG(level >= 1, "level must be >= 1 (the reference times nothing at TIMINGLEV 0)");
    HPCCG_TRY(matrix_required(c));
    HPCCG_CHECK_ARG(max_iter >= 0, "max_iter must be >= 0");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    const int64_t n = c->A.nrow;
    HPCCG_TRY(ensure_stage(c, n));
    HPCCG_TRY(ensure_workspace(c, n));
    HPCCG_TRY(ensure_hist(c, max_iter + 2));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s0, b, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->s1, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int st;
    if (level == 1) {
        double all_ms = 0.0, loop_ms = 0.0;
        st = cg_run(c, ctx_ws(c), true, c->s0, c->s1, max_iter, tolerance, niters, normr, hist, &all_ms, &loop_ms,
                    true);
        if (t_total) *t_total = loop_ms * 1e-3;
        if (t_prepare) *t_prepare = (all_ms - loop_ms) * 1e-3;
    } else {
        st = cg_timed_iterations(c, c->s0, c->s1, max_iter, tolerance, niters, normr, hist, t_total, t_prepare,
                                 t_iter, t_spmv);
    }
    if (st < 0) return st;
    HPCCG_TRY(hpccg_memcpy_d2h(c, x, c->s1, n * sizeof(double)));
    return st;
}

// ---- measurement ----------------------------------------------------------------------------
int hpccg_time_spmv(hpccg_ctx *c, const double *x, double *y, int32_t reps, double *avg_ms)
{
    // Times the SpMV kernel exactly as the CG loop launches it (k_cg_spmv with the p.Ap
    // partials of the tree mode; the one-workgroup finish kernel that forms alpha from them
    // is not part of it), driven by a private control block that never stops.
    HPCCG_CHECK_ARG(c && x && y && reps > 0 && avg_ms, "bad argument");
    HPCCG_TRY(matrix_required(c));
    HPCCG_TRY(ensure_hist(c, 2));
    CgState init{};
    init.k = 1;
    init.stop_at = INT32_MAX;
    init.max_iter = INT32_MAX - 1;
    *c->state_h = init;
    HPCCG_HIP_TRY(hipMemcpyAsync(c->state, c->state_h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    const double one = 1.0;
    HPCCG_HIP_TRY(hipMemcpyAsync(c->rr_hist, &one, sizeof(double), hipMemcpyHostToDevice, c->stream));
    auto launch = [&]() {
        hipLaunchKernelGGL(k_cg_spmv<true>, dim3(sell_grid(c->A)), dim3(kSellThreads), 0, c->stream, sell64(c->A), x,
                           y, c->state, c->A.red_part);
    };
    launch();  // warm
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; ++i) launch();
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / reps;
    return 0;
}

int hpccg_time_copy(hpccg_ctx *c, void *dst, const void *src, int64_t bytes, int32_t reps, double *avg_ms)
{
    HPCCG_CHECK_ARG(c && dst && src && reps > 0 && avg_ms && bytes % 16 == 0, "bad argument (bytes % 16 == 0)");
    const int64_t n16 = bytes / 16;
    const int G = c->info.num_cu * 4;  // 4 workgroups of 256 per CU: the fastest shape measured
    hipLaunchKernelGGL(k_copy16, dim3(G), dim3(256), 0, c->stream, (const u32x4 *)src, (u32x4 *)dst, n16);
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; ++i)
        hipLaunchKernelGGL(k_copy16, dim3(G), dim3(256), 0, c->stream, (const u32x4 *)src, (u32x4 *)dst, n16);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / reps;
    return 0;
}

}  // extern "C"
