// Mixed-precision CG: FP16 matrix values (27 and -1 are exact), int32 columns, FP32 vectors
// r / p / Ap, FP64 dot products for alpha and beta, FP64 solution x, and replacement of the
// recursive residual by the FP64 true residual b - A x every `replace_every` iterations
// (van der Vorst & Ye, "Residual replacement strategies for Krylov subspace iterative
// methods", SIAM J. Sci. Comput. 22, 2000).  No reference counterpart; it extends cg
// (MPI/src/numerics.F90:41-162) with the same operation sequence per iteration.
//
// Arithmetic: every product and sum of the SpMV and of the vector updates is formed in FP64
// registers and rounded once to the FP32 storage type (the kernels are bound by HBM traffic,
// so the wider arithmetic is free); only the stored vectors are narrow.
//
// Termination: the loop stops when the FP64 TRUE residual ||b - A x|| is below the absolute
// tolerance.  A recursive residual below the tolerance only triggers a true-residual check
// (and a replacement if it fails), so a reported convergence is never a rounding artefact of
// the FP32 recurrence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "spmv.hpp"

namespace hpccg {

struct MixState {
    double alpha;
    double beta;
    double rtrans;  // r.r used by the current iteration (recursive or replaced)
    double tol;
    int k;
    int stop_at;
    int max_iter;
    int replace_every;
    int need_true;
    int converged;
    int breakdown;
    int inexact16;
    unsigned counter[4];
};

// ---- FP16 values -------------------------------------------------------------------------
__global__ void k_to_half(const double *__restrict__ v, __half *__restrict__ h, int64_t nnz, int *inexact)
{
    int bad = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const __half q = __float2half((float)v[i]);  // exact whenever v is an FP16 value
        h[i] = q;
        if ((double)__half2float(q) != v[i]) bad = 1;
    }
    if (bad) atomicOr(inexact, 1);
}

// ---- kernels --------------------------------------------------------------------------------
// True residual r = b - A x (FP16 values, FP64 x, FP64 accumulation, FP32 store) and r.r.
// PREP: Prepare (iteration 0).  Otherwise runs only when the previous kernel requested it.
template <int ROWS, bool PREP>
__global__ __launch_bounds__(ROWS) void km_true_residual(const int32_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ col,
                                                         const __half *__restrict__ val,
                                                         const double *__restrict__ x,
                                                         const double *__restrict__ b,
                                                         float *__restrict__ r, int nrow, int ntiles,
                                                         MixState *st, double *hist_true,
                                                         double *partials)
{
    __shared__ double lds[ROWS * 27];
    __shared__ double red[ROWS / kWave];
    const int k = PREP ? 0 : st->k;
    if (!PREP && (k >= st->stop_at || !st->need_true)) return;
    int t0, t1, lb;
    block_range(ntiles, t0, t1, lb);
    double dot = 0.0;
    for (int tile = t0; tile < t1; ++tile) {
        bool valid;
        int row;
        const double ax = spmv_tile_rowsum<ROWS, __half, double>(rowptr, col, val, x, nrow, tile, lds, valid, row);
        if (valid) {
            const float ri = (float)(b[row] - ax);
            r[row] = ri;
            dot = fma((double)ri, (double)ri, dot);
        }
    }
    dot = block_sum<ROWS>(dot, red);
    double rr;
    if (grid_sum<ROWS>(dot, partials, &st->counter[0], gridDim.x, lb, red, &rr) && threadIdx.x == 0) {
        hist_true[k] = rr;
        const bool conv = sqrt(rr) < st->tol;
        if (PREP) {
            st->rtrans = rr;
            if (conv) {
                st->stop_at = 1;
                st->converged = 1;
            }
            st->k = 1;
        } else {
            if (conv) {
                st->stop_at = min(st->stop_at, k + 1);
                st->converged = 1;
            }
            st->beta = rr / st->rtrans;
            st->rtrans = rr;
            st->need_true = 0;
            st->k = k + 1;
        }
    }
}

// Top of iteration k: p = r (k == 1) or p = r + beta p.
__global__ __launch_bounds__(kVecThreads) void km_update_p(const float *__restrict__ r, float *__restrict__ p,
                                                           int n, const MixState *st)
{
    const int k = st->k;
    if (k >= st->stop_at) return;
    int64_t i0, i1;
    int lb;
    elem_range(n, i0, i1, lb);
    if (k == 1) {
        for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) p[i] = r[i];
        return;
    }
    const double beta = st->beta;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) p[i] = (float)fma(beta, (double)p[i], (double)r[i]);
}

// Ap = A p (FP16 values, FP32 p, FP64 accumulation, FP32 store) and p.Ap -> alpha.
template <int ROWS>
__global__ __launch_bounds__(ROWS) void km_spmv(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                const __half *__restrict__ val, const float *__restrict__ p,
                                                float *__restrict__ Ap, int nrow, int ntiles, MixState *st,
                                                double *partials)
{
    __shared__ double lds[ROWS * 27];
    __shared__ double red[ROWS / kWave];
    const int k = st->k;
    if (k >= st->stop_at) return;
    int t0, t1, lb;
    block_range(ntiles, t0, t1, lb);
    double dot = 0.0;
    for (int tile = t0; tile < t1; ++tile) {
        bool valid;
        int row;
        const double s = spmv_tile_rowsum<ROWS, __half, float>(rowptr, col, val, p, nrow, tile, lds, valid, row);
        if (valid) {
            const float a = (float)s;
            Ap[row] = a;
            dot = fma((double)p[row], (double)a, dot);
        }
    }
    dot = block_sum<ROWS>(dot, red);
    double pap;
    if (grid_sum<ROWS>(dot, partials, &st->counter[1], gridDim.x, lb, red, &pap) && threadIdx.x == 0) {
        if (!(pap != 0.0) || !isfinite(pap)) st->breakdown = 1;
        st->alpha = st->rtrans / pap;
    }
}

// x = x + alpha p (FP64), r = r - alpha Ap (FP32 store), recursive r.r; decides whether the
// true residual is formed next (every replace_every iterations, when the recursive norm passes
// the tolerance, and at max_iter).
__global__ __launch_bounds__(kVecThreads) void km_update_xr(double *__restrict__ x, float *__restrict__ r,
                                                            const float *__restrict__ p,
                                                            const float *__restrict__ Ap, int n, MixState *st,
                                                            double *hist_rec, double *partials)
{
    __shared__ double red[kVecThreads / kWave];
    const int k = st->k;
    if (k >= st->stop_at) return;
    const double alpha = st->alpha;
    int64_t i0, i1;
    int lb;
    elem_range(n, i0, i1, lb);
    double dot = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kVecThreads) {
        x[i] = fma(alpha, (double)p[i], x[i]);
        const float ri = (float)fma(-alpha, (double)Ap[i], (double)r[i]);
        r[i] = ri;
        dot = fma((double)ri, (double)ri, dot);
    }
    dot = block_sum<kVecThreads>(dot, red);
    double rr;
    if (grid_sum<kVecThreads>(dot, partials, &st->counter[2], gridDim.x, lb, red, &rr) && threadIdx.x == 0) {
        hist_rec[k] = rr;
        const bool check = (st->replace_every > 0 && k % st->replace_every == 0) || sqrt(rr) < st->tol ||
                           k >= st->max_iter;
        if (check) {
            st->need_true = 1;
        } else {
            st->beta = rr / st->rtrans;
            st->rtrans = rr;
            st->k = k + 1;
        }
    }
}

// ---- host ------------------------------------------------------------------------------------
int ensure_values16(hpccg_ctx *c)
{
    if (c->A.values16) return 0;
    const int64_t nnz = c->A.nnz;
    HPCCG_HIP_TRY(hipMalloc(&c->A.values16, std::max<int64_t>(nnz, 1) * sizeof(__half)));
    int *flag = nullptr;
    HPCCG_HIP_TRY(hipMalloc(&flag, sizeof(int)));
    int zero = 0, bad = 0;
    HPCCG_HIP_TRY(hipMemcpyAsync(flag, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (nnz > 0)
        hipLaunchKernelGGL(k_to_half, dim3(c->info.num_cu * 8), dim3(256), 0, c->stream, c->A.values, c->A.values16, nnz,
                           flag);
    HPCCG_HIP_TRY(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(flag);
    if (bad) {
        hipFree(c->A.values16);
        c->A.values16 = nullptr;
        set_error("mixed precision: some matrix values are not exactly representable in FP16");
        return HPCCG_ERR_ARG;
    }
    return 0;
}

void free_mixed(hpccg_ctx *c)
{
    MixedWs &m = c->mix;
    for (auto &g : m.graphs) hipGraphExecDestroy(g.second);
    m.graphs.clear();
    hipFree(m.r);
    hipFree(m.p);
    hipFree(m.Ap);
    hipFree(m.state);
    hipHostFree(m.state_h);
    hipFree(m.hist);
    m = MixedWs{};
}

static int ensure_mixed_ws(hpccg_ctx *c, int64_t n, int max_iter)
{
    MixedWs &m = c->mix;
    if (!m.state) {
        HPCCG_HIP_TRY(hipMalloc(&m.state, sizeof(MixState)));
        HPCCG_HIP_TRY(hipHostMalloc(&m.state_h, sizeof(MixState), hipHostMallocDefault));
        MixState z{};
        std::memcpy(m.state_h, &z, sizeof(z));
        HPCCG_HIP_TRY(hipMemcpy(m.state, m.state_h, sizeof(MixState), hipMemcpyHostToDevice));
    }
    if (m.n < n) {
        for (auto &g : m.graphs) hipGraphExecDestroy(g.second);
        m.graphs.clear();
        hipFree(m.r);
        hipFree(m.p);
        hipFree(m.Ap);
        m.r = m.p = m.Ap = nullptr;
        m.n = 0;
        if (hipMalloc(&m.r, n * sizeof(float)) != hipSuccess || hipMalloc(&m.p, n * sizeof(float)) != hipSuccess ||
            hipMalloc(&m.Ap, n * sizeof(float)) != hipSuccess) {
            set_error("mixed: workspace allocation failed");
            return HPCCG_ERR_NOMEM;
        }
        m.n = n;
    }
    if (m.hist_cap < max_iter + 2) {
        for (auto &g : m.graphs) hipGraphExecDestroy(g.second);
        m.graphs.clear();
        hipFree(m.hist);
        m.hist = nullptr;
        HPCCG_HIP_TRY(hipMalloc(&m.hist, 2 * (size_t)(max_iter + 2) * sizeof(double)));
        m.hist_cap = max_iter + 2;
    }
    return 0;
}

static int mixed_enqueue(hpccg_ctx *c, const double *b, double *x, int count)
{
    MixedWs &m = c->mix;
    const int n = c->A.nrow, nt = (n + kTileRows - 1) / kTileRows;
    const int Gs = HPCCG_SPMV_GRID(km_spmv<kTileRows>, c->info.num_cu, nt), Gv = vec_grid_for(c->info.num_cu, n);
    MixState *st = (MixState *)m.state;
    double *hrec = m.hist, *htrue = m.hist + m.hist_cap;
    for (int it = 0; it < count; ++it) {
        hipLaunchKernelGGL(km_update_p, dim3(Gv), dim3(kVecThreads), 0, c->stream, m.r, m.p, n, st);
        hipLaunchKernelGGL(km_spmv<kTileRows>, dim3(Gs), dim3(kTileRows), 0, c->stream, c->A.rowindex, c->A.colindex,
                           c->A.values16, m.p, m.Ap, n, nt, st, c->partials);
        hipLaunchKernelGGL(km_update_xr, dim3(Gv), dim3(kVecThreads), 0, c->stream, x, m.r, m.p, m.Ap, n, st, hrec,
                           c->partials);
        hipLaunchKernelGGL((km_true_residual<kTileRows, false>), dim3(Gs), dim3(kTileRows), 0, c->stream,
                           c->A.rowindex, c->A.colindex, c->A.values16, x, b, m.r, n, nt, st, htrue, c->partials);
    }
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

static int mixed_graph(hpccg_ctx *c, const double *b, double *x, int chunk, hipGraphExec_t *out)
{
    GraphKey key{b, x, 100, chunk};
    auto it = c->mix.graphs.find(key);
    if (it != c->mix.graphs.end()) {
        *out = it->second;
        return 0;
    }
    hipGraph_t g;
    HPCCG_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int st = mixed_enqueue(c, b, x, chunk);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (st) return st;
    HPCCG_HIP_TRY(e);
    hipGraphExec_t ge;
    const hipError_t e2 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HPCCG_HIP_TRY(e2);
    c->mix.graphs[key] = ge;
    *out = ge;
    return 0;
}

static int cg_mixed_device(hpccg_ctx *c, const double *b, double *x, int max_iter, double tol, int replace_every,
                           int *niters, double *true_normr, double *hist_rec, double *hist_true, double *elapsed_ms)
{
    if (!c->A.rowindex) {
        set_error("no matrix: call hpccg_matrix_upload or hpccg_generate_matrix first");
        return HPCCG_ERR_STATE;
    }
    HPCCG_CHECK_ARG(max_iter >= 0 && replace_every >= 0, "max_iter and replace_every must be >= 0");
    const int n = c->A.nrow, nt = (n + kTileRows - 1) / kTileRows;
    HPCCG_TRY(ensure_values16(c));
    HPCCG_TRY(ensure_mixed_ws(c, n, max_iter));
    MixedWs &m = c->mix;
    MixState init{};
    init.tol = tol;
    init.stop_at = max_iter + 1;
    init.max_iter = max_iter;
    init.replace_every = replace_every;
    std::memcpy(m.state_h, &init, sizeof(init));
    HPCCG_HIP_TRY(hipMemcpyAsync(m.state, m.state_h, sizeof(MixState), hipMemcpyHostToDevice, c->stream));
    // iterations that form no true residual keep -1 in its history
    std::vector<double> neg(2 * (size_t)m.hist_cap, -1.0);
    HPCCG_HIP_TRY(hipMemcpyAsync(m.hist, neg.data(), neg.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));

    hipGraphExec_t ge = nullptr;
    const int chunk = std::max(1, std::min(max_iter, 16));
    if (max_iter > 0) HPCCG_TRY(mixed_graph(c, b, x, chunk, &ge));

    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL((km_true_residual<kTileRows, true>), dim3(HPCCG_SPMV_GRID((km_true_residual<kTileRows, true>), c->info.num_cu, nt)),
                       dim3(kTileRows), 0, c->stream, c->A.rowindex, c->A.colindex, c->A.values16, x, b, m.r, n, nt, (MixState *)m.state,
                       m.hist + m.hist_cap, c->partials);
    HPCCG_HIP_TRY(hipGetLastError());
    int launched = 0;
    MixState *sh = (MixState *)m.state_h;
    while (max_iter > 0) {
        HPCCG_HIP_TRY(hipGraphLaunch(ge, c->stream));
        launched += chunk;
        HPCCG_HIP_TRY(hipMemcpyAsync(m.state_h, m.state, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        if (sh->k >= sh->stop_at) break;
        if (launched > 4 * max_iter + 64) {
            set_error("mixed cg: iteration counter did not advance");
            return HPCCG_ERR_STATE;
        }
    }
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(m.state_h, m.state, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
    std::vector<double> h(2 * (size_t)m.hist_cap);
    HPCCG_HIP_TRY(hipMemcpyAsync(h.data(), m.hist, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    if (elapsed_ms) {
        float ms = 0.f;
        HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *elapsed_ms = ms;
    }
    const MixState s = *sh;
    const int nit = s.k - 1;
    if (niters) *niters = nit;
    // the last true residual formed: at the last iteration run (max_iter and convergence both
    // force one), or Prepare's
    const double *ht = h.data() + m.hist_cap;
    if (true_normr) *true_normr = ht[std::max(0, nit)] >= 0.0 ? std::sqrt(ht[std::max(0, nit)]) : -1.0;
    if (hist_rec)
        for (int j = 0; j <= max_iter; ++j) hist_rec[j] = (j >= 1 && j <= nit) ? std::sqrt(h[j]) : -1.0;
    if (hist_true)
        for (int j = 0; j <= max_iter; ++j) hist_true[j] = ht[j] >= 0.0 ? std::sqrt(ht[j]) : -1.0;
    return s.breakdown ? HPCCG_WARN_BREAKDOWN : 0;
}

}  // namespace hpccg

using namespace hpccg;

extern "C" {

int hpccg_cg_mixed_dev(hpccg_ctx *c, const double *b, double *x, int32_t max_iter, double tolerance,
                       int32_t replace_every, int32_t *niters, double *true_normr, double *hist_rec,
                       double *hist_true, double *elapsed_ms)
{
    HPCCG_CHECK_ARG(c && b && x, "null argument");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    return cg_mixed_device(c, b, x, max_iter, tolerance, replace_every, niters, true_normr, hist_rec, hist_true,
                           elapsed_ms);
}

int hpccg_cg_mixed(hpccg_ctx *c, const double *b, double *x, int32_t max_iter, double tolerance,
                   int32_t replace_every, int32_t *niters, double *true_normr, double *hist_rec, double *hist_true)
{
    HPCCG_CHECK_ARG(c && b && x, "null argument");
    if (!c->A.rowindex) {
        set_error("no matrix: call hpccg_matrix_upload or hpccg_generate_matrix first");
        return HPCCG_ERR_STATE;
    }
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    const int64_t n = c->A.nrow;
    double *db = nullptr, *dx = nullptr;
    if (hipMalloc(&db, n * sizeof(double)) != hipSuccess || hipMalloc(&dx, n * sizeof(double)) != hipSuccess) {
        hipFree(db);
        set_error("hpccg_cg_mixed: allocation failed");
        return HPCCG_ERR_NOMEM;
    }
    int st = 0;
    if (hipMemcpyAsync(db, b, n * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(dx, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        set_error("hpccg_cg_mixed: upload failed");
        st = HPCCG_ERR_HIP;
    }
    if (st == 0)
        st = cg_mixed_device(c, db, dx, max_iter, tolerance, replace_every, niters, true_normr, hist_rec, hist_true,
                             nullptr);
    if (st >= 0) {
        if (hipMemcpyAsync(x, dx, n * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            set_error("hpccg_cg_mixed: download failed");
            st = HPCCG_ERR_HIP;
        }
    }
    // the cached graphs captured these temporaries' addresses
    for (auto &g : c->mix.graphs) hipGraphExecDestroy(g.second);
    c->mix.graphs.clear();
    hipFree(db);
    hipFree(dx);
    return st;
}

// y = A x with FP16 values and FP32 vectors (FP64 accumulation), device arrays.
int hpccg_spmv_mixed_dev(hpccg_ctx *c, const float *x, float *y)
{
    HPCCG_CHECK_ARG(c && x && y, "null argument");
    if (!c->A.rowindex) {
        set_error("no matrix");
        return HPCCG_ERR_STATE;
    }
    HPCCG_TRY(ensure_values16(c));
    HPCCG_TRY(ensure_mixed_ws(c, c->A.nrow, 1));
    MixState init{};
    init.k = 1;
    init.stop_at = INT32_MAX;
    init.rtrans = 1.0;
    std::memcpy(c->mix.state_h, &init, sizeof(init));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->mix.state, c->mix.state_h, sizeof(MixState), hipMemcpyHostToDevice, c->stream));
    const int n = c->A.nrow, nt = (n + kTileRows - 1) / kTileRows;
    hipLaunchKernelGGL(km_spmv<kTileRows>, dim3(HPCCG_SPMV_GRID(km_spmv<kTileRows>, c->info.num_cu, nt)), dim3(kTileRows), 0, c->stream,
                       c->A.rowindex, c->A.colindex, c->A.values16, x, y, n, nt, (MixState *)c->mix.state,
                       c->partials);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_time_spmv_mixed(hpccg_ctx *c, const float *x, float *y, int32_t reps, double *avg_ms)
{
    HPCCG_CHECK_ARG(c && x && y && reps > 0 && avg_ms, "bad argument");
    HPCCG_TRY(hpccg_spmv_mixed_dev(c, x, y));  // sets up state and warms
    const int n = c->A.nrow, nt = (n + kTileRows - 1) / kTileRows;
    const int G = HPCCG_SPMV_GRID(km_spmv<kTileRows>, c->info.num_cu, nt);
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; ++i)
        hipLaunchKernelGGL(km_spmv<kTileRows>, dim3(G), dim3(kTileRows), 0, c->stream, c->A.rowindex, c->A.colindex,
                           c->A.values16, x, y, n, nt, (MixState *)c->mix.state, c->partials);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / reps;
    return 0;
}

}  // extern "C"
