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
    double rr_true;  // r.r of the last true residual
    int k;
    int stop_at;
    int max_iter;
    int replace_every;
    int need_true;
    int converged;
    int breakdown;
    int inexact16;
    int x_nonzero;  // set by Prepare's k_nonzero when x0 != 0 (else A x0 is skipped)
};

// ---- kernels --------------------------------------------------------------------------------
// True residual r = b - A x (FP16 values, FP64 x, FP64 accumulation, FP32 store) and the
// workgroup's share of r.r.  PREP: Prepare (iteration 0).  Otherwise runs only when the
// previous kernel requested it.
template <bool PREP>
__device__ __forceinline__ bool km_true_due(const MixState *st)
{
    return PREP || (st->k < st->stop_at && st->need_true);
}

template <bool PREP>
__global__ __launch_bounds__(kTrueThreads) void km_true_residual(const SellRef<__half> A, const double *__restrict__ x,
                                                                 const double *__restrict__ b, float *__restrict__ r,
                                                                 const MixState *st, double *red_part)
{
    __shared__ double red[kTrueThreads / kWave];
    if (!km_true_due<PREP>(st)) return;
    const bool x0 = PREP && !st->x_nonzero;  // x = 0: A x is +0.0 in every row
    double dot = 0.0;
    for (int rep = 0; rep < kTrueLoop; ++rep) {
        int64_t row;
        bool valid;
        const int g = blockIdx.x * kTrueLoop + rep;
        double ax = 0.0;
        if (x0) sell_row_of_thread<kTrueWaves>(A.nslices, A.nrow, g, row, valid);
        else ax = sell_rowsum_of_thread<kGroupF16, __half, double, kTrueWaves>(A, x, row, valid, g, rep == 0);
        if (valid) {
            const float ri = (float)(b[row] - ax);
            __builtin_nontemporal_store(ri, r + row);
            dot += (double)ri * (double)ri;
        }
    }
    store_partial(block_sum<kTrueThreads>(dot, red), red_part);
}

// r.r of the true residual: replacement, convergence test on the FP64 true residual, beta.
template <bool PREP>
__global__ __launch_bounds__(kFinishThreads) void km_true_finish(const double *__restrict__ red_part, int nparts,
                                                                 MixState *st, double *hist_true)
{
    __shared__ double red[kFinishThreads / kWave];
    if (!km_true_due<PREP>(st)) return;
    const int k = PREP ? 0 : st->k;
    const double rr = sum_partials(red_part, nparts, red);
    if (threadIdx.x == 0) {
        hist_true[k] = rr;
        st->rr_true = rr;
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
    elem_range4(n, i0, i1, lb);
    const bool first = k == 1;
    const double beta = st->beta;
    auto one = [&](float ri, float pi) { return first ? ri : (float)fma(beta, (double)pi, (double)ri); };
    vec4_loop(
        i0, i1,
        [&](int64_t i) {
            const f32x4v rv = ld4_nt(r + i);  // streamed: keeps p, which the SpMV gathers next, cached
            f32x4v pv = first ? rv : ld4(p + i);
            pv.x = one(rv.x, pv.x);
            pv.y = one(rv.y, pv.y);
            pv.z = one(rv.z, pv.z);
            pv.w = one(rv.w, pv.w);
            st4(p + i, pv);
        },
        [&](int64_t i) { p[i] = one(r[i], first ? 0.f : p[i]); });
}

// Ap = A p (FP16 values, FP32 p, FP64 accumulation, FP32 store) and the workgroup's p.Ap.
__global__ __launch_bounds__(kSellThreads) __attribute__((amdgpu_waves_per_eu(kSellF16WavesPerEU, kSellF16WavesPerEU))) void km_spmv(const SellRef<__half> A, const float *__restrict__ p,
                                                        float *__restrict__ Ap, const MixState *st, double *red_part)
{
    __shared__ double red[kSellThreads / kWave];
    if (st->k >= st->stop_at) return;
    int64_t row;
    bool valid;
    const double s = sell_rowsum_of_thread<kGroupF16, __half, float, kSellWaves, true>(A, p, row, valid);
    double dot = 0.0;
    if (valid) {
        const float a = (float)s;
        __builtin_nontemporal_store(a, Ap + row);
        dot = (double)p[row] * (double)a;
    }
    store_partial(block_sum<kSellThreads>(dot, red), red_part);
}

// p.Ap -> alpha
__global__ __launch_bounds__(kFinishThreads) void km_spmv_finish(const double *__restrict__ red_part, int nparts,
                                                                 MixState *st)
{
    __shared__ double red[kFinishThreads / kWave];
    if (st->k >= st->stop_at) return;
    const double pap = sum_partials(red_part, nparts, red);
    if (threadIdx.x == 0) {
        if (!(pap != 0.0) || !isfinite(pap)) st->breakdown = 1;
        st->alpha = st->rtrans / pap;
    }
}

// x = x + alpha p (FP64), r = r - alpha Ap (FP32 store), and the workgroup's recursive r.r
__global__ __launch_bounds__(kVecThreads) void km_update_xr(double *__restrict__ x, float *__restrict__ r,
                                                            const float *__restrict__ p,
                                                            const float *__restrict__ Ap, int n, const MixState *st,
                                                            double *partials)
{
    __shared__ double red[kVecThreads / kWave];
    if (st->k >= st->stop_at) return;
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
    store_partial(block_sum<kVecThreads>(dot, red), partials);
}

// recursive r.r; decides whether the true residual is formed next (every replace_every
// iterations, after a kReplaceDrop2 fall since the last one, when the recursive norm passes
// the tolerance, and at max_iter)
__global__ __launch_bounds__(kFinishThreads) void km_xr_finish(const double *__restrict__ partials, int nparts,
                                                               MixState *st, double *hist_rec)
{
    __shared__ double red[kFinishThreads / kWave];
    const int k = st->k;
    if (k >= st->stop_at) return;
    const double rr = sum_partials(partials, nparts, red);
    if (threadIdx.x == 0) {
        hist_rec[k] = rr;
        const bool check = (st->replace_every > 0 && k % st->replace_every == 0) || sqrt(rr) < st->tol ||
                           k >= st->max_iter || rr < kReplaceDrop2 * st->rr_true;
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
// FP16 sliced copy of the matrix (built on first use): every value must be an FP16 number.
// A coded matrix shares the FP64 copy's column codes; otherwise the copy gets int32 columns.
int ensure_sell16(hpccg_ctx *c)
{
    Matrix &A = c->A;
    if (A.sval16) return 0;
    // A non-null sval16 means a finished copy: every failure below leaves both pointers null.
    int *flag = nullptr;
    if (dmalloc(&flag, sizeof(int)) != hipSuccess ||
        (!A.code && dmalloc(&A.scol16, std::max<int64_t>(A.slots, 1) * sizeof(int32_t)) != hipSuccess) ||
        dmalloc(&A.sval16, std::max<int64_t>(A.slots, 1) * sizeof(__half)) != hipSuccess) {
        dfree(flag);
        dfree(A.scol16);
        dfree(A.sval16);
        A.scol16 = nullptr;
        A.sval16 = nullptr;
        set_error("mixed precision: allocation of the FP16 sliced copy failed");
        return HPCCG_ERR_NOMEM;
    }
    int bad = 0;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), c->stream);
    if (e == hipSuccess && A.nslices > 0) {
        const int64_t threads = (int64_t)A.nslices * kSlice;
        hipLaunchKernelGGL((k_csr_to_sell<kGroupF16, __half>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                           c->stream, A.rowindex, A.colindex, A.values, A.slice_ptr, A.scol16, A.sval16,
                           (int64_t)A.nrow, A.nslices, flag, nullptr, nullptr, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(flag);
    if (e != hipSuccess || bad) {
        dfree(A.scol16);
        dfree(A.sval16);
        A.scol16 = nullptr;
        A.sval16 = nullptr;
        if (e != hipSuccess) {
            set_error("mixed precision: FP16 copy: %s", hipGetErrorString(e));
            return HPCCG_ERR_HIP;
        }
        set_error("mixed precision: some matrix values are not exactly representable in FP16");
        return HPCCG_ERR_ARG;
    }
    return 0;
}

void free_mixed(hpccg_ctx *c)
{
    MixedWs &m = c->mix;
    m.graphs.clear();
    dfree(m.r);
    dfree(m.p);
    dfree(m.Ap);
    dfree(m.state);
    hipHostFree(m.state_h);
    dfree(m.hist);
    m = MixedWs{};
}

static int ensure_mixed_ws(hpccg_ctx *c, int64_t n, int max_iter)
{
    MixedWs &m = c->mix;
    if (!m.state) {
        HPCCG_HIP_TRY(dmalloc(&m.state, sizeof(MixState)));
        HPCCG_HIP_TRY(hipHostMalloc(&m.state_h, sizeof(MixState), hipHostMallocDefault));
        MixState z{};
        std::memcpy(m.state_h, &z, sizeof(z));
        HPCCG_HIP_TRY(hipMemcpy(m.state, m.state_h, sizeof(MixState), hipMemcpyHostToDevice));
    }
    if (m.n < n) {
        dfree(m.r);
        dfree(m.p);
        dfree(m.Ap);
        m.r = m.p = m.Ap = nullptr;
        m.n = 0;
        if (dmalloc(&m.r, n * sizeof(float)) != hipSuccess || dmalloc(&m.p, n * sizeof(float)) != hipSuccess ||
            dmalloc(&m.Ap, n * sizeof(float)) != hipSuccess) {
            set_error("mixed: workspace allocation failed");
            return HPCCG_ERR_NOMEM;
        }
        m.n = n;
    }
    if (m.hist_cap < max_iter + 2) {
        dfree(m.hist);
        m.hist = nullptr;
        HPCCG_HIP_TRY(dmalloc(&m.hist, 2 * (size_t)(max_iter + 2) * sizeof(double)));
        m.hist_cap = max_iter + 2;
    }
    return 0;
}

// Ap = A p with p.Ap -> alpha; true residual r = b - A x with its epilogue
static void launch_km_spmv(hpccg_ctx *c, const float *p, float *Ap)
{
    const Matrix &A = c->A;
    const int G = sell_grid(A);
    MixState *st = (MixState *)c->mix.state;
    hipLaunchKernelGGL(km_spmv, dim3(G), dim3(kSellThreads), 0, c->stream, sell16(A), p, Ap, st, A.red_part);
    hipLaunchKernelGGL(km_spmv_finish, dim3(1), dim3(kFinishThreads), 0, c->stream, A.red_part, G, st);
}
template <bool PREP>
static void launch_km_true(hpccg_ctx *c, const double *b, const double *x)
{
    const Matrix &A = c->A;
    const int G = true_blocks(A.nslices);
    MixedWs &m = c->mix;
    MixState *st = (MixState *)m.state;
    if (PREP)  // x = 0 (the usual start) needs no A x: the flag is set otherwise
        hipLaunchKernelGGL(k_nonzero<>, dim3(vec_grid_for(c->info.num_cu, A.nrow)), dim3(kVecThreads), 0, c->stream, x,
                           (int64_t)A.nrow, &st->x_nonzero);
    hipLaunchKernelGGL(km_true_residual<PREP>, dim3(G), dim3(kTrueThreads), 0, c->stream, sell16(A), x, b, m.r, st,
                       A.red_part);
    hipLaunchKernelGGL(km_true_finish<PREP>, dim3(1), dim3(kFinishThreads), 0, c->stream, A.red_part, G, st,
                       m.hist + m.hist_cap);
}

static int mixed_enqueue(hpccg_ctx *c, const double *b, double *x, int count)
{
    MixedWs &m = c->mix;
    const int n = c->A.nrow;
    const int Gv = vec_grid_for(c->info.num_cu, n);
    MixState *st = (MixState *)m.state;
    double *hrec = m.hist;
    for (int it = 0; it < count; ++it) {
        hipLaunchKernelGGL(km_update_p, dim3(Gv), dim3(kVecThreads), 0, c->stream, m.r, m.p, n, st);
        launch_km_spmv(c, m.p, m.Ap);
        hipLaunchKernelGGL(km_update_xr, dim3(Gv), dim3(kVecThreads), 0, c->stream, x, m.r, m.p, m.Ap, n, st,
                           c->partials);
        hipLaunchKernelGGL(km_xr_finish, dim3(1), dim3(kFinishThreads), 0, c->stream, c->partials, Gv, st, hrec);
        launch_km_true<false>(c, b, x);
    }
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

static int mixed_graph(hpccg_ctx *c, const double *b, double *x, int chunk, hipGraphExec_t *out)
{
    return cached_graph(c->mix.graphs, c->stream, GraphKey{b, x, 100, chunk},
                        [&]() { return mixed_enqueue(c, b, x, chunk); }, out);
}

static int cg_mixed_device(hpccg_ctx *c, const double *b, double *x, int max_iter, double tol, int replace_every,
                           int *niters, double *true_normr, double *hist_rec, double *hist_true, double *elapsed_ms)
{
    if (!c->A.rowindex) {
        set_error("no matrix: call hpccg_matrix_upload or hpccg_generate_matrix first");
        return HPCCG_ERR_STATE;
    }
    HPCCG_CHECK_ARG(max_iter >= 0 && replace_every >= 0, "max_iter and replace_every must be >= 0");
    const int n = c->A.nrow;
    HPCCG_TRY(ensure_sell16(c));
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

    // graphs of up to 16 iterations, sized by the host from the residual rate (ChunkPlanner)
    auto get = [&](int size, hipGraphExec_t *ge) { return mixed_graph(c, b, x, size, ge); };
    HPCCG_TRY(capture_iteration_graphs(16, max_iter, get));
    ChunkPlanner plan{16, tol, 0};

    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    launch_km_true<true>(c, b, x);
    HPCCG_HIP_TRY(hipGetLastError());
    int launched = 0, mm = std::min(16, max_iter);
    MixState *sh = (MixState *)m.state_h;
    while (max_iter > 0) {
        HPCCG_TRY(launch_iterations(c->stream, mm, get));
        launched += mm;
        HPCCG_HIP_TRY(hipMemcpyAsync(m.state_h, m.state, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        if (sh->k >= sh->stop_at) break;
        if (launched > 4 * max_iter + 64) {
            set_error("mixed cg: iteration counter did not advance");
            return HPCCG_ERR_STATE;
        }
        // iterations whose true residual is still due stay in the count (k does not advance)
        mm = plan.next(sh->rtrans, sh->k - 1, std::max(1, max_iter - (sh->k - 1)));
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
    // -1 marks an iteration that formed no true residual; a formed one is reported even when it
    // is NaN (a NaN or Inf in b or x)
    auto formed = [](double v) { return v != -1.0; };
    if (true_normr) *true_normr = formed(ht[std::max(0, nit)]) ? std::sqrt(ht[std::max(0, nit)]) : -1.0;
    if (hist_rec)
        for (int j = 0; j <= max_iter; ++j) hist_rec[j] = (j >= 1 && j <= nit) ? std::sqrt(h[j]) : -1.0;
    if (hist_true)
        for (int j = 0; j <= max_iter; ++j) hist_true[j] = formed(ht[j]) ? std::sqrt(ht[j]) : -1.0;
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
    if (dmalloc(&db, n * sizeof(double)) != hipSuccess || dmalloc(&dx, n * sizeof(double)) != hipSuccess) {
        dfree(db);
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
    // freeing the temporaries retires the graphs captured with them (buffer generation)
    dfree(db);
    dfree(dx);
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
    HPCCG_TRY(ensure_sell16(c));
    HPCCG_TRY(ensure_mixed_ws(c, c->A.nrow, 1));
    MixState init{};
    init.k = 1;
    init.stop_at = INT32_MAX;
    init.rtrans = 1.0;
    std::memcpy(c->mix.state_h, &init, sizeof(init));
    HPCCG_HIP_TRY(hipMemcpyAsync(c->mix.state, c->mix.state_h, sizeof(MixState), hipMemcpyHostToDevice, c->stream));
    launch_km_spmv(c, x, y);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_time_spmv_mixed(hpccg_ctx *c, const float *x, float *y, int32_t reps, double *avg_ms)
{
    HPCCG_CHECK_ARG(c && x && y && reps > 0 && avg_ms, "bad argument");
    HPCCG_TRY(hpccg_spmv_mixed_dev(c, x, y));  // sets up state and warms
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; ++i)
        hipLaunchKernelGGL(km_spmv, dim3(sell_grid(c->A)), dim3(kSellThreads), 0, c->stream, sell16(c->A), x, y,
                           (const MixState *)c->mix.state, c->A.red_part);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / reps;
    return 0;
}

}  // extern "C"
