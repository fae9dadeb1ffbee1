// Batched CG: many independent 27-point systems of uneven size (each with its own grid,
// right-hand side and convergence), solved together.  No reference counterpart: the reference
// solves one nx*ny*nz grid per rank (MPI/src/main.F90:117-130); each system here is exactly
// such a grid (SparseMatrix.F90:138-175) and each is solved by the mixed-precision CG of
// mixed.hip (FP16 values, FP32 r/p/Ap, FP64 dots and x, FP64 true-residual replacement).
//
// Layout: the systems' CSRs are concatenated into one block-diagonal CSR (global 1-based
// row pointers and columns), vectors are concatenated by row offset.  Every kernel covers all
// systems at once: system s owns a "team" of g_s workgroups (g_s proportional to its row
// tiles), each team reduces its own dots with the single-pass last-block pattern, and a
// system that has converged makes its team's workgroups return at once.  The whole solve --
// Prepare and max_iter iterations -- is captured as ONE graph and replayed with one launch.
//
// Launch shape (rows per workgroup tile = rows per wave x waves, and workgroups per CU) is
// chosen per batch by hpccg_batch_autotune, which times every compiled variant on the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "spmv.hpp"

namespace hpccg {

struct BatchState {
    double alpha, beta, rtrans, tol2;  // tol2 = (rel_tol * ||b||)^2
    double bb;                         // ||b||^2
    double rr_true;                    // last true ||b - A x||^2
    int k;                             // iteration to run next
    int active;                        // 1 while iterating
    int need_true;
    int niters;
    int converged;
    int breakdown;
    int replace_every, max_iter;
    unsigned counter[4];
};

struct BatchLayout {  // device arrays describing systems and teams for one variant
    int *wg_team = nullptr;     // [G] team (system) of each logical workgroup, -1 = idle
    int *team_first = nullptr;  // [S] first logical workgroup of each team
    int *team_size = nullptr;   // [S]
    int G = 0;
    int rows = 0, occ = 0;
};

struct Batch {
    int S = 0;
    std::vector<int32_t> dims;          // 3*S
    std::vector<int64_t> row_off_h;     // S+1
    int64_t E = 0, nnz = 0;
    int32_t *rowindex = nullptr;        // E+1, global 1-based
    int32_t *colindex = nullptr;        // nnz, global 1-based
    __half *values = nullptr;           // nnz (27 / -1 exactly)
    int64_t *row_off = nullptr;         // device copy of row_off_h
    // sliced-ELL copy (spmv.hpp), slices never straddle two systems
    int64_t *slice_ptr = nullptr;       // [total slices + 1] slot offsets
    int32_t *scol = nullptr;            // 0-based global columns
    __half *sval = nullptr;
    int *slice_off = nullptr;           // [S+1] first slice of each system
    int64_t slots = 0;
    float *r = nullptr, *p = nullptr, *Ap = nullptr;
    BatchState *st = nullptr;
    double *partials = nullptr;         // one slot per logical workgroup
    int partials_cap = 0;
    BatchLayout lay[8];
    int best_variant = -1;
    hipGraphExec_t graph = nullptr;
    const void *g_b = nullptr, *g_x = nullptr;
    int g_variant = -1, g_max_iter = -1;
    double g_tol2 = -1.0;
};

// compiled variants: rows per tile (one row per lane) x workgroups per CU
struct Variant {
    int rows, occ;
};
static const Variant kVariants[] = {{64, 8}, {64, 16}, {128, 4}, {128, 8}, {256, 2}, {256, 4}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

static Batch *batch_of(hpccg_ctx *c) { return (Batch *)c->ext[0]; }

// ---- generation: system s rows [off, off+n) of the block-diagonal CSR ----------------------
__device__ __forceinline__ int64_t axp(int64_t m, int64_t n) { return 3 * m - (m >= 1 ? 1 : 0) - (m >= n ? 1 : 0); }
__device__ __forceinline__ int64_t axc(int64_t i, int64_t n) { return axp(i + 1, n) - axp(i, n); }

__global__ void kb_generate(int nx, int ny, int nz, int64_t row_off, int64_t nnz_off, int32_t *__restrict__ rowindex,
                            int32_t *__restrict__ colindex, __half *__restrict__ values)
{
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (row >= N) return;
    const int ix = (int)(row % nx), iy = (int)((row / nx) % ny), iz = (int)(row / ((int64_t)nx * ny));
    const int64_t cx = 3 * (int64_t)nx - 2, cy = 3 * (int64_t)ny - 2;
    const int64_t cursor = axp(iz, nz) * cy * cx + axc(iz, nz) * axp(iy, ny) * cx + axc(iz, nz) * axc(iy, ny) * axp(ix, nx);
    rowindex[row_off + row] = (int32_t)(nnz_off + cursor + 1);
    int64_t c = nnz_off + cursor;
    for (int sz = -1; sz <= 1; ++sz) {
        if (iz + sz < 0 || iz + sz >= nz) continue;
        for (int sy = -1; sy <= 1; ++sy) {
            if (iy + sy < 0 || iy + sy >= ny) continue;
            for (int sx = -1; sx <= 1; ++sx) {
                if (ix + sx < 0 || ix + sx >= nx) continue;
                const int64_t col = row + (int64_t)sz * nx * ny + (int64_t)sy * nx + sx;
                values[c] = __float2half(col == row ? 27.0f : -1.0f);
                colindex[c] = (int32_t)(row_off + col + 1);
                ++c;
            }
        }
    }
}

__global__ void kb_set_last(int32_t *rowindex, int64_t E, int64_t nnz) { rowindex[E] = (int32_t)(nnz + 1); }

// ---- per-team helpers -------------------------------------------------------------------------
struct TeamPos {
    int s, l, g;
    int64_t off, n;
};

__device__ __forceinline__ bool team_pos(const int *wg_team, const int *team_first, const int *team_size,
                                         const int64_t *row_off, TeamPos &tp, int &lb)
{
    lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int s = wg_team[lb];
    if (s < 0) return false;
    tp.s = s;
    tp.l = lb - team_first[s];
    tp.g = team_size[s];
    tp.off = row_off[s];
    tp.n = row_off[s + 1] - row_off[s];
    return true;
}

// The batch's sliced copy as the SpMV kernels take it.
struct BatchSell {
    const int64_t *slice_ptr;
    const int32_t *col;
    const __half *val;
    const int32_t *rowptr;  // global 1-based row pointer
    const int *slice_off;   // first slice of each system
};

static inline BatchSell batch_sell(const Batch *B) { return BatchSell{B->slice_ptr, B->scol, B->sval, B->rowindex, B->slice_off}; }

// A team's workgroup walks tiles of ROWS/64 slices of its system (one slice per wave) over a
// contiguous tile range; f(global row, row sum) runs for every row of the system it covers.
template <int ROWS, typename XT, typename F>
__device__ __forceinline__ void team_rows(const BatchSell &A, const TeamPos &tp, const XT *__restrict__ x, F f)
{
    constexpr int W = ROWS / kSlice;
    const int nsl = (int)((tp.n + kSlice - 1) / kSlice);
    const int ntiles = (nsl + W - 1) / W;
    const int t0 = (int)(((int64_t)tp.l * ntiles) / tp.g), t1 = (int)(((int64_t)(tp.l + 1) * ntiles) / tp.g);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kSlice), lane = threadIdx.x % kSlice;
    const int soff = A.slice_off[tp.s];
    for (int t = t0; t < t1; ++t) {
        const int ls = t * W + w;
        if (ls >= nsl) continue;  // wave-uniform
        const int64_t rl = (int64_t)ls * kSlice + lane;
        const bool valid = rl < tp.n;
        const int64_t gr = tp.off + rl;
        const double sum = sell_rowsum<kGroupF16, __half, XT>(A.slice_ptr, A.col, A.val, A.rowptr, x, soff + ls, lane,
                                                             gr, valid);
        if (valid) f(gr, sum);
    }
}

// ---- kernels ------------------------------------------------------------------------------------
// True residual of every system that asked for one (PREP: all systems, also forms ||b||^2).
template <int ROWS, bool PREP>
__global__ __launch_bounds__(ROWS) void kb_true(const BatchSell A, const double *__restrict__ x,
                                                const double *__restrict__ b, float *__restrict__ r,
                                                const int64_t *__restrict__ row_off, const int *wg_team,
                                                const int *team_first, const int *team_size, BatchState *st,
                                                double *partials, double rel_tol2)
{
    __shared__ double red[ROWS / kWave];
    TeamPos tp;
    int lb;
    if (!team_pos(wg_team, team_first, team_size, row_off, tp, lb)) return;
    BatchState *S = st + tp.s;
    if (!PREP && (!S->active || !S->need_true)) return;
    double rr = 0.0, bb = 0.0;
    team_rows<ROWS>(A, tp, x, [&](int64_t gr, double ax) {
        const double bi = b[gr];
        const float ri = (float)(bi - ax);
        __builtin_nontemporal_store(ri, r + gr);
        rr = fma((double)ri, (double)ri, rr);
        if (PREP) bb = fma(bi, bi, bb);
    });
    rr = block_sum<ROWS>(rr, red);
    if (PREP) bb = block_sum<ROWS>(bb, red);
    __shared__ int is_last;
    if (threadIdx.x == 0) {
        partials[2 * lb] = rr;
        partials[2 * lb + 1] = bb;
        __threadfence();
        is_last = atomicAdd(&S->counter[0], 1u) == (unsigned)tp.g - 1u;
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double a = 0.0, c2 = 0.0;
    const int first = lb - tp.l;
    for (int i = threadIdx.x; i < tp.g; i += ROWS) {
        a += load_agent(partials + 2 * (first + i));
        c2 += load_agent(partials + 2 * (first + i) + 1);
    }
    a = block_sum<ROWS>(a, red);
    c2 = block_sum<ROWS>(c2, red);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&S->counter[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        S->rr_true = a;
        if (PREP) {
            S->bb = c2;
            S->tol2 = rel_tol2 * c2;
            S->rtrans = a;
            S->k = 1;
            S->niters = 0;
            S->active = (a < S->tol2) ? 0 : (S->max_iter > 0 ? 1 : 0);
            S->converged = a < S->tol2;
        } else {
            const int k = S->k;
            S->niters = k;
            if (a < S->tol2) {
                S->converged = 1;
                S->active = 0;
            } else if (k >= S->max_iter) {
                S->active = 0;
            }
            S->beta = a / S->rtrans;
            S->rtrans = a;
            S->need_true = 0;
            S->k = k + 1;
        }
    }
}

template <int ROWS>
__global__ __launch_bounds__(ROWS) void kb_update_p(const float *__restrict__ r, float *__restrict__ p,
                                                    const int64_t *__restrict__ row_off, const int *wg_team,
                                                    const int *team_first, const int *team_size, const BatchState *st)
{
    TeamPos tp;
    int lb;
    if (!team_pos(wg_team, team_first, team_size, row_off, tp, lb)) return;
    const BatchState *S = st + tp.s;
    if (!S->active) return;
    const int64_t i0 = tp.off + (tp.n * tp.l) / tp.g, i1 = tp.off + (tp.n * (tp.l + 1)) / tp.g;
    if (S->k == 1) {
        for (int64_t i = i0 + threadIdx.x; i < i1; i += ROWS) p[i] = r[i];
        return;
    }
    const double beta = S->beta;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += ROWS) p[i] = (float)fma(beta, (double)p[i], (double)r[i]);
}

template <int ROWS>
__global__ __launch_bounds__(ROWS) void kb_spmv(const BatchSell A, const float *__restrict__ p, float *__restrict__ Ap,
                                                const int64_t *__restrict__ row_off, const int *wg_team,
                                                const int *team_first, const int *team_size, BatchState *st,
                                                double *partials)
{
    __shared__ double red[ROWS / kWave];
    TeamPos tp;
    int lb;
    if (!team_pos(wg_team, team_first, team_size, row_off, tp, lb)) return;
    BatchState *S = st + tp.s;
    if (!S->active) return;
    double dot = 0.0;
    team_rows<ROWS>(A, tp, p, [&](int64_t gr, double s) {
        const float a = (float)s;
        __builtin_nontemporal_store(a, Ap + gr);
        dot = fma((double)p[gr], (double)a, dot);
    });
    dot = block_sum<ROWS>(dot, red);
    double pap;
    if (grid_sum<ROWS>(dot, partials + (lb - tp.l), &S->counter[1], tp.g, tp.l, red, &pap) && threadIdx.x == 0) {
        if (!(pap != 0.0) || !isfinite(pap)) S->breakdown = 1;
        S->alpha = S->rtrans / pap;
    }
}

template <int ROWS>
__global__ __launch_bounds__(ROWS) void kb_update_xr(double *__restrict__ x, float *__restrict__ r,
                                                     const float *__restrict__ p, const float *__restrict__ Ap,
                                                     const int64_t *__restrict__ row_off, const int *wg_team,
                                                     const int *team_first, const int *team_size, BatchState *st,
                                                     double *partials)
{
    __shared__ double red[ROWS / kWave];
    TeamPos tp;
    int lb;
    if (!team_pos(wg_team, team_first, team_size, row_off, tp, lb)) return;
    BatchState *S = st + tp.s;
    if (!S->active) return;
    const double alpha = S->alpha;
    const int64_t i0 = tp.off + (tp.n * tp.l) / tp.g, i1 = tp.off + (tp.n * (tp.l + 1)) / tp.g;
    double dot = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += ROWS) {
        x[i] = fma(alpha, (double)p[i], x[i]);
        const float ri = (float)fma(-alpha, (double)Ap[i], (double)r[i]);
        r[i] = ri;
        dot = fma((double)ri, (double)ri, dot);
    }
    dot = block_sum<ROWS>(dot, red);
    double rr;
    if (grid_sum<ROWS>(dot, partials + (lb - tp.l), &S->counter[2], tp.g, tp.l, red, &rr) && threadIdx.x == 0) {
        const int k = S->k;
        const bool check = (S->replace_every > 0 && k % S->replace_every == 0) || rr < S->tol2 ||
                           k >= S->max_iter || rr < kReplaceDrop2 * S->rr_true;
        if (check) {
            S->need_true = 1;
        } else {
            S->niters = k;
            S->beta = rr / S->rtrans;
            S->rtrans = rr;
            S->k = k + 1;
        }
    }
}

// ---- host ---------------------------------------------------------------------------------------
void free_batched(hpccg_ctx *c)
{
    Batch *B = batch_of(c);
    if (!B) return;
    if (B->graph) hipGraphExecDestroy(B->graph);
    hipFree(B->rowindex);
    hipFree(B->colindex);
    hipFree(B->values);
    hipFree(B->row_off);
    hipFree(B->slice_ptr);
    hipFree(B->scol);
    hipFree(B->sval);
    hipFree(B->slice_off);
    hipFree(B->r);
    hipFree(B->p);
    hipFree(B->Ap);
    hipFree(B->st);
    hipFree(B->partials);
    for (auto &L : B->lay) {
        hipFree(L.wg_team);
        hipFree(L.team_first);
        hipFree(L.team_size);
    }
    delete B;
    c->ext[0] = nullptr;
}

static int build_layout(hpccg_ctx *c, Batch *B, int v)
{
    BatchLayout &L = B->lay[v];
    if (L.wg_team) return 0;
    const int rows = kVariants[v].rows, occ = kVariants[v].occ;
    std::vector<int64_t> tiles(B->S);
    int64_t T = 0;
    for (int s = 0; s < B->S; ++s) {
        tiles[s] = (B->row_off_h[s + 1] - B->row_off_h[s] + rows - 1) / rows;
        T += tiles[s];
    }
    const double target = (double)c->info.num_cu * occ;
    std::vector<int> first(B->S), size(B->S);
    int G = 0;
    for (int s = 0; s < B->S; ++s) {
        int g = (int)std::lround(target * (double)tiles[s] / (double)T);
        g = std::max(1, std::min<int>(g, (int)tiles[s]));
        first[s] = G;
        size[s] = g;
        G += g;
    }
    const int Gp = (G + kNumXcd - 1) / kNumXcd * kNumXcd;
    std::vector<int> team(Gp, -1);
    for (int s = 0; s < B->S; ++s)
        for (int i = 0; i < size[s]; ++i) team[first[s] + i] = s;
    HPCCG_HIP_TRY(hipMalloc(&L.wg_team, Gp * sizeof(int)));
    HPCCG_HIP_TRY(hipMalloc(&L.team_first, B->S * sizeof(int)));
    HPCCG_HIP_TRY(hipMalloc(&L.team_size, B->S * sizeof(int)));
    HPCCG_HIP_TRY(hipMemcpy(L.wg_team, team.data(), Gp * sizeof(int), hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(L.team_first, first.data(), B->S * sizeof(int), hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(L.team_size, size.data(), B->S * sizeof(int), hipMemcpyHostToDevice));
    L.G = Gp;
    L.rows = rows;
    L.occ = occ;
    if (2 * Gp > B->partials_cap) {
        hipFree(B->partials);
        B->partials = nullptr;
        HPCCG_HIP_TRY(hipMalloc(&B->partials, 2 * (size_t)Gp * sizeof(double)));
        B->partials_cap = 2 * Gp;
        if (B->graph) hipGraphExecDestroy(B->graph);
        B->graph = nullptr;
    }
    return 0;
}

template <int ROWS>
static void enqueue_batch(hpccg_ctx *c, Batch *B, const BatchLayout &L, const double *b, double *x, int max_iter,
                          double rel_tol2)
{
    const dim3 grid(L.G), blk(ROWS);
    const BatchSell A = batch_sell(B);
    hipLaunchKernelGGL((kb_true<ROWS, true>), grid, blk, 0, c->stream, A, x, b, B->r, B->row_off, L.wg_team,
                       L.team_first, L.team_size, B->st, B->partials, rel_tol2);
    for (int it = 0; it < max_iter; ++it) {
        hipLaunchKernelGGL(kb_update_p<ROWS>, grid, blk, 0, c->stream, B->r, B->p, B->row_off, L.wg_team,
                           L.team_first, L.team_size, B->st);
        hipLaunchKernelGGL(kb_spmv<ROWS>, grid, blk, 0, c->stream, A, B->p, B->Ap, B->row_off, L.wg_team,
                           L.team_first, L.team_size, B->st, B->partials);
        hipLaunchKernelGGL(kb_update_xr<ROWS>, grid, blk, 0, c->stream, x, B->r, B->p, B->Ap, B->row_off, L.wg_team,
                           L.team_first, L.team_size, B->st, B->partials);
        hipLaunchKernelGGL((kb_true<ROWS, false>), grid, blk, 0, c->stream, A, x, b, B->r, B->row_off, L.wg_team,
                           L.team_first, L.team_size, B->st, B->partials, rel_tol2);
    }
}

static int enqueue_variant(hpccg_ctx *c, Batch *B, int v, const double *b, double *x, int max_iter, double rel_tol2)
{
    const BatchLayout &L = B->lay[v];
    switch (kVariants[v].rows) {
    case 64: enqueue_batch<64>(c, B, L, b, x, max_iter, rel_tol2); break;
    case 128: enqueue_batch<128>(c, B, L, b, x, max_iter, rel_tol2); break;
    case 256: enqueue_batch<256>(c, B, L, b, x, max_iter, rel_tol2); break;
    default: set_error("batched: bad variant"); return HPCCG_ERR_ARG;
    }
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

// One solve: state reset (plain copies, outside the graph) + one graph launch.
static int batch_solve(hpccg_ctx *c, Batch *B, const double *b, double *x, double rel_tol, int max_iter,
                       int replace_every, int v, double *ms_out)
{
    HPCCG_TRY(build_layout(c, B, v));
    std::vector<BatchState> init(B->S);
    for (auto &s : init) {
        std::memset(&s, 0, sizeof(s));
        s.replace_every = replace_every;
        s.max_iter = max_iter;
    }
    HPCCG_HIP_TRY(hipMemcpyAsync(B->st, init.data(), B->S * sizeof(BatchState), hipMemcpyHostToDevice, c->stream));
    const double rel_tol2 = rel_tol * rel_tol;
    // the graph bakes pointers, the variant, max_iter and rel_tol2: re-capture when they change
    if (!B->graph || B->g_b != b || B->g_x != x || B->g_variant != v || B->g_max_iter != max_iter ||
        B->g_tol2 != rel_tol2) {
        if (B->graph) hipGraphExecDestroy(B->graph);
        B->graph = nullptr;
        hipGraph_t g;
        HPCCG_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int st = enqueue_variant(c, B, v, b, x, max_iter, rel_tol2);
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (st) return st;
        HPCCG_HIP_TRY(e);
        const hipError_t e2 = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        HPCCG_HIP_TRY(e2);
        B->g_b = b;
        B->g_x = x;
        B->g_variant = v;
        B->g_max_iter = max_iter;
        B->g_tol2 = rel_tol2;
    }
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HPCCG_HIP_TRY(hipGraphLaunch(B->graph, c->stream));
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (ms_out) *ms_out = ms;
    return 0;
}

}  // namespace hpccg

using namespace hpccg;

extern "C" {

int hpccg_batch_setup(hpccg_ctx *c, int32_t count, const int32_t *dims)
{
    HPCCG_CHECK_ARG(c && count >= 1 && dims, "bad argument");
    for (int i = 0; i < 3 * count; ++i) HPCCG_CHECK_ARG(dims[i] >= 1, "grid extents must be >= 1");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    free_batched(c);
    Batch *B = new Batch();
    c->ext[0] = B;
    B->S = count;
    B->dims.assign(dims, dims + 3 * count);
    B->row_off_h.assign(count + 1, 0);
    std::vector<int64_t> nnz_off(count + 1, 0);
    for (int s = 0; s < count; ++s) {
        const int64_t nx = dims[3 * s], ny = dims[3 * s + 1], nz = dims[3 * s + 2];
        B->row_off_h[s + 1] = B->row_off_h[s] + nx * ny * nz;
        nnz_off[s + 1] = nnz_off[s] + (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
    }
    B->E = B->row_off_h[count];
    B->nnz = nnz_off[count];
    if (B->E >= INT32_MAX || B->nnz + 1 >= INT32_MAX) {
        free_batched(c);
        set_error("batch too large for int32 CSR");
        return HPCCG_ERR_ARG;
    }
    if (hipMalloc(&B->rowindex, (B->E + 1) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&B->colindex, B->nnz * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&B->values, B->nnz * sizeof(__half)) != hipSuccess ||
        hipMalloc(&B->row_off, (count + 1) * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&B->r, B->E * sizeof(float)) != hipSuccess || hipMalloc(&B->p, B->E * sizeof(float)) != hipSuccess ||
        hipMalloc(&B->Ap, B->E * sizeof(float)) != hipSuccess ||
        hipMalloc(&B->st, count * sizeof(BatchState)) != hipSuccess) {
        free_batched(c);
        set_error("batched: device allocation failed");
        return HPCCG_ERR_NOMEM;
    }
    HPCCG_HIP_TRY(hipMemcpy(B->row_off, B->row_off_h.data(), (count + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    for (int s = 0; s < count; ++s) {
        const int64_t n = B->row_off_h[s + 1] - B->row_off_h[s];
        hipLaunchKernelGGL(kb_generate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, dims[3 * s],
                           dims[3 * s + 1], dims[3 * s + 2], B->row_off_h[s], nnz_off[s], B->rowindex, B->colindex,
                           B->values);
    }
    hipLaunchKernelGGL(kb_set_last, dim3(1), dim3(1), 0, c->stream, B->rowindex, B->E, B->nnz);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    // sliced copy, system by system (each system's rows start a new slice)
    std::vector<int32_t> rowh(B->E + 1);
    HPCCG_HIP_TRY(hipMemcpy(rowh.data(), B->rowindex, (B->E + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int> soff(count + 1, 0);
    for (int s = 0; s < count; ++s)
        soff[s + 1] = soff[s] + (int)((B->row_off_h[s + 1] - B->row_off_h[s] + kSlice - 1) / kSlice);
    std::vector<int64_t> sp(soff[count] + 1, 0);
    for (int s = 0; s < count; ++s)
        sell_slice_ptr(rowh.data() + B->row_off_h[s], B->row_off_h[s + 1] - B->row_off_h[s], sp[soff[s]],
                       sp.data() + soff[s]);
    B->slots = sp[soff[count]];
    if (hipMalloc(&B->slice_ptr, sp.size() * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&B->scol, std::max<int64_t>(B->slots, 1) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&B->sval, std::max<int64_t>(B->slots, 1) * sizeof(__half)) != hipSuccess ||
        hipMalloc(&B->slice_off, (count + 1) * sizeof(int)) != hipSuccess) {
        free_batched(c);
        set_error("batched: device allocation of the sliced copy failed");
        return HPCCG_ERR_NOMEM;
    }
    HPCCG_HIP_TRY(hipMemcpy(B->slice_ptr, sp.data(), sp.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(B->slice_off, soff.data(), (count + 1) * sizeof(int), hipMemcpyHostToDevice));
    for (int s = 0; s < count; ++s) {
        const int ns = soff[s + 1] - soff[s];
        const int64_t threads = (int64_t)ns * kSlice;
        hipLaunchKernelGGL((k_csr_to_sell<kGroupF16, __half, __half>), dim3((unsigned)((threads + 255) / 256)),
                           dim3(256), 0, c->stream, B->rowindex + B->row_off_h[s], B->colindex, B->values,
                           B->slice_ptr + soff[s], B->scol, B->sval, B->row_off_h[s + 1] - B->row_off_h[s], ns,
                           nullptr);
    }
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_batch_info(hpccg_ctx *c, int32_t *count, int64_t *total_rows, int64_t *total_nnz, int64_t *row_off)
{
    HPCCG_CHECK_ARG(c, "null context");
    Batch *B = batch_of(c);
    if (!B) {
        set_error("no batch: call hpccg_batch_setup first");
        return HPCCG_ERR_STATE;
    }
    if (count) *count = B->S;
    if (total_rows) *total_rows = B->E;
    if (total_nnz) *total_nnz = B->nnz;
    if (row_off) std::memcpy(row_off, B->row_off_h.data(), (B->S + 1) * sizeof(int64_t));
    return 0;
}

// y = A x for the whole batch (FP16 values, FP32 vectors, FP64 accumulation); device arrays.
int hpccg_batch_spmv_dev(hpccg_ctx *c, const float *x, float *y)
{
    HPCCG_CHECK_ARG(c && x && y, "null argument");
    Batch *B = batch_of(c);
    if (!B) {
        set_error("no batch");
        return HPCCG_ERR_STATE;
    }
    HPCCG_TRY(build_layout(c, B, 2));
    const BatchLayout &L = B->lay[2];
    std::vector<BatchState> init(B->S);
    for (auto &s : init) {
        std::memset(&s, 0, sizeof(s));
        s.active = 1;
        s.rtrans = 1.0;
    }
    HPCCG_HIP_TRY(hipMemcpyAsync(B->st, init.data(), B->S * sizeof(BatchState), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(kb_spmv<128>, dim3(L.G), dim3(128), 0, c->stream, batch_sell(B), x, y, B->row_off, L.wg_team,
                       L.team_first, L.team_size, B->st, B->partials);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_batch_variants(int32_t *count, int32_t *rows, int32_t *occ)
{
    if (count) *count = kNumVariants;
    for (int v = 0; v < kNumVariants; ++v) {
        if (rows) rows[v] = kVariants[v].rows;
        if (occ) occ[v] = kVariants[v].occ;
    }
    return 0;
}

// Times every compiled launch variant on this batch (solves of `probe_iters` iterations from
// the given b and x, x restored afterwards), keeps the fastest for variant = -1 solves, and
// returns it with the per-variant times (ms).
int hpccg_batch_autotune(hpccg_ctx *c, const double *b, double *x, int32_t probe_iters, int32_t *best,
                         double *variant_ms)
{
    HPCCG_CHECK_ARG(c && b && x && probe_iters >= 1, "bad argument");
    Batch *B = batch_of(c);
    if (!B) {
        set_error("no batch");
        return HPCCG_ERR_STATE;
    }
    double *x0 = nullptr;
    HPCCG_HIP_TRY(hipMalloc(&x0, B->E * sizeof(double)));
    HPCCG_HIP_TRY(hipMemcpyAsync(x0, x, B->E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    int bi = -1;
    double bt = 1e300;
    for (int v = 0; v < kNumVariants; ++v) {
        double t = 0.0, ms = 0.0;
        for (int rep = 0; rep < 3; ++rep) {
            HPCCG_HIP_TRY(hipMemcpyAsync(x, x0, B->E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            const int st = batch_solve(c, B, b, x, 0.0, probe_iters, 0, v, &ms);
            if (st) {
                hipFree(x0);
                return st;
            }
            if (rep > 0) t += ms;  // first run captures and warms
        }
        t /= 2;
        if (variant_ms) variant_ms[v] = t;
        if (t < bt) {
            bt = t;
            bi = v;
        }
    }
    HPCCG_HIP_TRY(hipMemcpyAsync(x, x0, B->E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(x0);
    if (B->graph) hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
    B->best_variant = bi;
    if (best) *best = bi;
    return 0;
}

// Solve every system to ||b_s - A_s x_s|| <= rel_tol * ||b_s|| (FP64 true residual) or
// max_iter iterations.  variant = -1 uses the autotuned (else the default) launch variant.
// niters[s], relres[s] (true relative residual at exit) and converged[s] are per system.
int hpccg_batch_solve_dev(hpccg_ctx *c, const double *b, double *x, double rel_tol, int32_t max_iter,
                          int32_t replace_every, int32_t variant, int32_t *niters, double *relres,
                          int32_t *converged, double *elapsed_ms)
{
    HPCCG_CHECK_ARG(c && b && x && max_iter >= 0 && replace_every >= 0, "bad argument");
    HPCCG_CHECK_ARG(variant >= -1 && variant < kNumVariants, "variant out of range");
    Batch *B = batch_of(c);
    if (!B) {
        set_error("no batch: call hpccg_batch_setup first");
        return HPCCG_ERR_STATE;
    }
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    const int v = variant >= 0 ? variant : (B->best_variant >= 0 ? B->best_variant : 3);
    HPCCG_TRY(batch_solve(c, B, b, x, rel_tol, max_iter, replace_every, v, elapsed_ms));
    std::vector<BatchState> s(B->S);
    HPCCG_HIP_TRY(hipMemcpy(s.data(), B->st, B->S * sizeof(BatchState), hipMemcpyDeviceToHost));
    int brk = 0;
    for (int i = 0; i < B->S; ++i) {
        if (niters) niters[i] = s[i].niters;
        if (relres) relres[i] = std::sqrt(s[i].rr_true) / std::sqrt(s[i].bb);
        if (converged) converged[i] = s[i].converged;
        brk |= s[i].breakdown;
    }
    return brk ? HPCCG_WARN_BREAKDOWN : 0;
}

}  // extern "C"
