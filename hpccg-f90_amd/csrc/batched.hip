// Batched CG: many independent 27-point systems of uneven size (each with its own grid,
// right-hand side and convergence), solved together.  No reference counterpart: the reference
// solves one nx*ny*nz grid per rank (MPI/src/main.F90:117-130); each system here is exactly
// such a grid (SparseMatrix.F90:138-175) and each is solved by the mixed-precision CG of
// mixed.hip (FP16 values, FP32 r/p/Ap, FP64 dots and x, FP64 true-residual replacement).
//
// Layout: the systems' CSRs are concatenated into one block-diagonal CSR (global 1-based
// row pointers and columns), vectors are concatenated by row offset, and the sliced-ELL copy
// starts every system on a new 64-row slice.  The copy stores one-byte column codes, each an
// index into its system's table of stencil offsets col - row.  Every kernel covers all systems at once and is
// slice-parallel: a wave owns `rows per wave` / 64 consecutive slices of the whole batch and
// writes one dot partial per slice; a per-system finish kernel (one workgroup per system) sums
// its slices' partials in a fixed order and updates that system's scalars.  A converged
// system's slices return at once, and once every system has stopped every launch returns at
// its first instruction.  Iterations run as captured graphs of 1, 2, 4, ... kBatchCap
// iterations, enqueued back to back without a host round trip; the host reads the systems'
// state only between launches that it sized from the batch's worst residual ratio, and a
// repeated solve of the same batch enqueues the iteration count of the previous solve at once
// (one launch sequence, no read until it has drained).
//
// Launch shape (rows per wave and waves per workgroup) is chosen per batch by
// hpccg_batch_autotune, which times every compiled variant on the batch.
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
};

// Where a slice lies, in one 16-byte record (one scalar load per slice instead of a chain of
// system-table loads): its first global row, its system and how many of its 64 rows exist.
struct SliceInfo {
    int64_t row0;
    int32_t sys;
    int32_t nvalid;
};

struct Batch {
    int S = 0;
    std::vector<int32_t> dims;          // 3*S
    std::vector<int64_t> row_off_h;     // S+1
    int64_t E = 0, nnz = 0;
    int32_t *rowindex = nullptr;        // E+1, global 1-based
    int32_t *colindex = nullptr;        // nnz, global 1-based
    __half *values = nullptr;           // nnz (27 / -1 exactly)
    // sliced-ELL copy (spmv.hpp), slices never straddle two systems
    int64_t *slice_ptr = nullptr;       // [total slices + 1] slot offsets
    uint8_t *code = nullptr;            // column codes (spmv.hpp), offsets relative to each system
    int32_t *tabs = nullptr;            // [S][kBatchTab] each system's sorted offset table
    __half *sval = nullptr;
    int *slice_off = nullptr;           // [S+1] first slice of each system
    SliceInfo *sinfo = nullptr;         // [NS] each slice's system, first row and valid rows
    int NS = 0;                         // slices of the whole batch
    int64_t slots = 0;
    float *r = nullptr, *p = nullptr, *Ap = nullptr;
    BatchState *st = nullptr;
    double *partials = nullptr;         // [2 NS]: per-slice dot partials
    int *nactive = nullptr;             // [0] systems still iterating (0: every launch returns at once),
                                        // [1] Prepare's k_nonzero flag (x0 != 0)
    int best_variant = -1;
    BatchState *snap = nullptr;         // pinned: [2][S] state snapshots (after Prepare, latest)
    int hint = 0;                       // iterations the previous solve needed (0: none yet)
    GraphCache graphs;                  // iteration graphs by (b, x, variant, size)
};

// most iterations in one captured graph; launches are split into power-of-two graphs
constexpr int kBatchCap = 64;
// iterations of the first launch when no previous solve gave a count
constexpr int kBatchFirst = 16;

// compiled variants: waves per workgroup x slices per wave (rows per wave = 64 x slices)
struct Variant {
    int waves, spw;
};
static const Variant kVariants[] = {{4, 1}, {8, 1}, {2, 1}, {16, 1}, {4, 2}, {8, 2}, {4, 4}, {8, 4}};
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

// ---- slice helpers -----------------------------------------------------------------------------
// The batch's sliced copy as the kernels take it.  Columns are one-byte codes into the slice's
// system's table of its (at most 27) stencil offsets col - row.
constexpr int kBatchTab = 32;  // table entries per system (27 used, the rest repeat the last)
struct BatchSell {
    const int64_t *slice_ptr;
    const uint8_t *code;
    const int32_t *tabs;
    const __half *val;
    const int32_t *rowptr;  // global 1-based row pointer
    const int *slice_off;   // first slice of each system
    const SliceInfo *sinfo; // [NS] system, first row and valid rows of each slice
    int NS;
};

static inline BatchSell batch_sell(const Batch *B)
{
    return BatchSell{B->slice_ptr, B->code, B->tabs, B->sval, B->rowindex, B->slice_off, B->sinfo, B->NS};
}

// Where slice gs lies: its system, the lane's global row and whether that row exists.
struct SlicePos {
    int s;
    int64_t gr;
    bool valid;
};
__device__ __forceinline__ SlicePos slice_pos(const BatchSell &A, int gs, int lane)
{
    const SliceInfo si = A.sinfo[gs];
    SlicePos q;
    q.s = si.sys;
    q.valid = lane < si.nvalid;
    q.gr = si.row0 + lane;
    return q;
}

// The SPW consecutive slices of this wave: f(gs) for each that exists.
template <int W, int SPW, typename F>
__device__ __forceinline__ void wave_slices(int NS, F f)
{
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kSlice);
    const int g0 = (blockIdx.x * W + w) * SPW;
#pragma unroll
    for (int j = 0; j < SPW; ++j)
        if (g0 + j < NS) f(g0 + j);
}

// Row sum (FP64, the reference's entry order) of this lane's row of slice gs.  The slice's
// system table is held one entry per lane (lane i: entry i % kBatchTab) and each code is looked
// up with a cross-lane read; all 64 lanes of the wave take part.
template <typename XT>
__device__ __forceinline__ double batch_rowsum(const BatchSell &A, const SlicePos &q, const XT *__restrict__ x, int gs,
                                               int lane)
{
    const int tv = A.tabs[kBatchTab * q.s + lane % kBatchTab];
    return sell_rowsum_coded<kGroupF16, __half, XT>(
        A.slice_ptr, A.code, [&](int c) { return __builtin_amdgcn_ds_bpermute(c << 2, tv); }, A.val, A.rowptr, x, gs,
        lane, (int)q.gr, q.valid);
}

// Sum of the slice partials [slice_off[s], slice_off[s+1]) (stride `stride`, offset `o`) in
// a fixed order; every thread of the 256-thread workgroup gets it.
__device__ __forceinline__ double system_sum(const double *__restrict__ part, const int *__restrict__ slice_off, int s,
                                             int stride, int o, double *red)
{
    // eight independent loads in flight per thread (a 96^3 system has 13,824 slices)
    constexpr int U = 8;
    double a[U] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int i1 = slice_off[s + 1];
    int i = slice_off[s] + threadIdx.x;
    for (; i + (U - 1) * 256 < i1; i += U * 256) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = part[(int64_t)stride * (i + u * 256) + o];
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] += v[u];
    }
#pragma unroll
    for (int u = 0; u < U - 1; ++u)
        if (i + u * 256 < i1) a[u] += part[(int64_t)stride * (i + u * 256) + o];
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) t += a[u];
    return block_sum<256>(t, red);
}

// ---- kernels ------------------------------------------------------------------------------------
// True residual r = b - A x of every system that asked for one (PREP: all systems, also
// ||b||^2), per-slice partials (r.r, b.b) at part[2 gs], part[2 gs + 1].
template <int W, int SPW, bool PREP>
__global__ __launch_bounds__(W * 64) void kb_true(const BatchSell A, const double *__restrict__ x,
                                                   const double *__restrict__ b, float *__restrict__ r,
                                                   const BatchState *st, double *part, const int *nactive)
{
    if (!PREP && *nactive == 0) return;
    const bool x0 = PREP && nactive[1] == 0;  // x = 0: A x is +0.0 in every row
    const int lane = threadIdx.x % kSlice;
    wave_slices<W, SPW>(A.NS, [&](int gs) {
        const SlicePos q = slice_pos(A, gs, lane);
        const BatchState *S = st + q.s;
        if (!PREP && (!S->active || !S->need_true)) return;
        const double ax = x0 ? 0.0 : batch_rowsum(A, q, x, gs, lane);
        double rr = 0.0, bb = 0.0;
        if (q.valid) {
            const double bi = b[q.gr];
            const float ri = (float)(bi - ax);
            __builtin_nontemporal_store(ri, r + q.gr);
            rr = (double)ri * (double)ri;
            bb = bi * bi;
        }
        rr = wave_sum(rr);
        if (PREP) bb = wave_sum(bb);
        if (lane == 0) {
            part[2 * (int64_t)gs] = rr;
            part[2 * (int64_t)gs + 1] = bb;
        }
    });
}

// one workgroup per system: the true residual's norm, convergence, the iteration count
template <bool PREP>
__global__ __launch_bounds__(256) void kb_true_finish(const double *__restrict__ part, const int *__restrict__ slice_off,
                                                      BatchState *st, double rel_tol2, int *nactive)
{
    __shared__ double red[4];
    const int s = blockIdx.x;
    BatchState *S = st + s;
    if (!PREP && (!S->active || !S->need_true)) return;
    const double a = system_sum(part, slice_off, s, 2, 0, red);
    const double c2 = PREP ? system_sum(part, slice_off, s, 2, 1, red) : 0.0;
    if (threadIdx.x != 0) return;
    S->rr_true = a;
    if (PREP) {
        S->bb = c2;
        S->tol2 = rel_tol2 * c2;
        S->rtrans = a;
        S->k = 1;
        S->niters = 0;
        S->active = (a < S->tol2) ? 0 : (S->max_iter > 0 ? 1 : 0);
        S->converged = a < S->tol2;
        if (S->active) atomicAdd(nactive, 1);
    } else {
        const int k = S->k;
        S->niters = k;
        if (a < S->tol2) {
            S->converged = 1;
            S->active = 0;
        } else if (k >= S->max_iter) {
            S->active = 0;
        }
        if (!S->active) atomicSub(nactive, 1);
        S->beta = a / S->rtrans;
        S->rtrans = a;
        S->need_true = 0;
        S->k = k + 1;
    }
}

// Top of iteration k: p = r (k == 1) or p = r + beta p.
template <int W, int SPW>
__global__ __launch_bounds__(W * 64) void kb_update_p(const BatchSell A, const float *__restrict__ r,
                                                       float *__restrict__ p, const BatchState *st,
                                                       const int *nactive)
{
    if (*nactive == 0) return;
    const int lane = threadIdx.x % kSlice;
    wave_slices<W, SPW>(A.NS, [&](int gs) {
        const SlicePos q = slice_pos(A, gs, lane);
        const BatchState *S = st + q.s;
        if (!S->active || !q.valid) return;
        const float ri = __builtin_nontemporal_load(r + q.gr);  // keeps p, which the SpMV gathers next, cached
        p[q.gr] = S->k == 1 ? ri : (float)fma(S->beta, (double)p[q.gr], (double)ri);
    });
}

// Ap = A p and the slice's p.Ap partial at part[gs]
template <int W, int SPW>
__global__ __launch_bounds__(W * 64) void kb_spmv(const BatchSell A, const float *__restrict__ p,
                                                   float *__restrict__ Ap, const BatchState *st, double *part,
                                                   const int *nactive)
{
    if (*nactive == 0) return;
    const int lane = threadIdx.x % kSlice;
    wave_slices<W, SPW>(A.NS, [&](int gs) {
        const SlicePos q = slice_pos(A, gs, lane);
        if (!st[q.s].active) return;
        const double sum = batch_rowsum(A, q, p, gs, lane);
        double dot = 0.0;
        if (q.valid) {
            const float a = (float)sum;
            __builtin_nontemporal_store(a, Ap + q.gr);
            dot = (double)p[q.gr] * (double)a;
        }
        dot = wave_sum(dot);
        if (lane == 0) part[gs] = dot;
    });
}

// one workgroup per system: alpha = rtrans / p.Ap.  On a breakdown (p.Ap zero or not finite)
// the system stops here: x keeps the last iterate, and it is reported unconverged.
__global__ __launch_bounds__(256) void kb_pap_finish(const double *__restrict__ part, const int *__restrict__ slice_off,
                                                     BatchState *st, int *nactive)
{
    __shared__ double red[4];
    BatchState *S = st + blockIdx.x;
    if (!S->active) return;
    const double pap = system_sum(part, slice_off, blockIdx.x, 1, 0, red);
    if (threadIdx.x == 0) {
        if (!(pap != 0.0) || !isfinite(pap)) {
            S->breakdown = 1;
            S->active = 0;
            atomicSub(nactive, 1);
            return;
        }
        S->alpha = S->rtrans / pap;
    }
}

// x = x + alpha p (FP64), r = r - alpha Ap (FP32), the slice's r.r partial at part[gs]
template <int W, int SPW>
__global__ __launch_bounds__(W * 64) void kb_update_xr(const BatchSell A, double *__restrict__ x,
                                                        float *__restrict__ r, const float *__restrict__ p,
                                                        const float *__restrict__ Ap, const BatchState *st,
                                                        double *part, const int *nactive)
{
    if (*nactive == 0) return;
    const int lane = threadIdx.x % kSlice;
    wave_slices<W, SPW>(A.NS, [&](int gs) {
        const SlicePos q = slice_pos(A, gs, lane);
        const BatchState *S = st + q.s;
        if (!S->active) return;
        double dot = 0.0;
        if (q.valid) {
            const double alpha = S->alpha;
            x[q.gr] = fma(alpha, (double)p[q.gr], x[q.gr]);
            const float ri = (float)fma(-alpha, (double)Ap[q.gr], (double)r[q.gr]);
            r[q.gr] = ri;
            dot = (double)ri * (double)ri;
        }
        dot = wave_sum(dot);
        if (lane == 0) part[gs] = dot;
    });
}

// one workgroup per system: recursive r.r; a true residual next (every replace_every
// iterations, after a kReplaceDrop2 fall, at the tolerance, at max_iter) or beta and k + 1
__global__ __launch_bounds__(256) void kb_xr_finish(const double *__restrict__ part, const int *__restrict__ slice_off,
                                                    BatchState *st)
{
    __shared__ double red[4];
    BatchState *S = st + blockIdx.x;
    if (!S->active) return;
    const double rr = system_sum(part, slice_off, blockIdx.x, 1, 0, red);
    if (threadIdx.x != 0) return;
    const int k = S->k;
    const bool check = (S->replace_every > 0 && k % S->replace_every == 0) || rr < S->tol2 || k >= S->max_iter ||
                       rr < kReplaceDrop2 * S->rr_true;
    if (check) {
        S->need_true = 1;
    } else {
        S->niters = k;
        S->beta = rr / S->rtrans;
        S->rtrans = rr;
        S->k = k + 1;
    }
}

// ---- host ---------------------------------------------------------------------------------------
void free_batched(hpccg_ctx *c)
{
    Batch *B = batch_of(c);
    if (!B) return;
    B->graphs.clear();
    hipHostFree(B->snap);
    dfree(B->rowindex);
    dfree(B->colindex);
    dfree(B->values);
    dfree(B->slice_ptr);
    dfree(B->code);
    dfree(B->tabs);
    dfree(B->sval);
    dfree(B->slice_off);
    dfree(B->r);
    dfree(B->p);
    dfree(B->Ap);
    dfree(B->st);
    dfree(B->partials);
    dfree(B->sinfo);
    dfree(B->nactive);
    delete B;
    c->ext[0] = nullptr;
}

// Prepare (iters < 0: r = b - A x, ||b||, tolerances, active count) or `iters` iterations.
template <int W, int SPW>
static void enqueue_batch(hpccg_ctx *c, Batch *B, const double *b, double *x, int iters, double rel_tol2)
{
    const dim3 grid((unsigned)((B->NS + W * SPW - 1) / (W * SPW))), blk(W * 64), sys((unsigned)B->S), fin(256);
    const BatchSell A = batch_sell(B);
    double *part = B->partials;
    int *na = B->nactive;
    if (iters < 0) {
        // x = 0 (the usual start) needs no A x: the flag is set otherwise
        hipLaunchKernelGGL(k_nonzero<>, dim3(vec_grid_for(c->info.num_cu, B->E)), dim3(kVecThreads), 0, c->stream, x,
                           B->E, na + 1);
        hipLaunchKernelGGL((kb_true<W, SPW, true>), grid, blk, 0, c->stream, A, x, b, B->r, B->st, part, na);
        hipLaunchKernelGGL(kb_true_finish<true>, sys, fin, 0, c->stream, part, B->slice_off, B->st, rel_tol2, na);
        return;
    }
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL((kb_update_p<W, SPW>), grid, blk, 0, c->stream, A, B->r, B->p, B->st, na);
        hipLaunchKernelGGL((kb_spmv<W, SPW>), grid, blk, 0, c->stream, A, B->p, B->Ap, B->st, part, na);
        hipLaunchKernelGGL(kb_pap_finish, sys, fin, 0, c->stream, part, B->slice_off, B->st, na);
        hipLaunchKernelGGL((kb_update_xr<W, SPW>), grid, blk, 0, c->stream, A, x, B->r, B->p, B->Ap, B->st, part,
                           na);
        hipLaunchKernelGGL(kb_xr_finish, sys, fin, 0, c->stream, part, B->slice_off, B->st);
        hipLaunchKernelGGL((kb_true<W, SPW, false>), grid, blk, 0, c->stream, A, x, b, B->r, B->st, part, na);
        hipLaunchKernelGGL(kb_true_finish<false>, sys, fin, 0, c->stream, part, B->slice_off, B->st, rel_tol2, na);
    }
}

static int enqueue_variant(hpccg_ctx *c, Batch *B, int v, const double *b, double *x, int iters, double rel_tol2)
{
    switch (v) {
    case 0: enqueue_batch<4, 1>(c, B, b, x, iters, rel_tol2); break;
    case 1: enqueue_batch<8, 1>(c, B, b, x, iters, rel_tol2); break;
    case 2: enqueue_batch<2, 1>(c, B, b, x, iters, rel_tol2); break;
    case 3: enqueue_batch<16, 1>(c, B, b, x, iters, rel_tol2); break;
    case 4: enqueue_batch<4, 2>(c, B, b, x, iters, rel_tol2); break;
    case 5: enqueue_batch<8, 2>(c, B, b, x, iters, rel_tol2); break;
    case 6: enqueue_batch<4, 4>(c, B, b, x, iters, rel_tol2); break;
    case 7: enqueue_batch<8, 4>(c, B, b, x, iters, rel_tol2); break;
    default: set_error("batched: bad variant"); return HPCCG_ERR_ARG;
    }
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

// max over the systems still iterating of r.r / (rel_tol ||b||)^2 (0 when none is)
static double worst_ratio(const BatchState *st, int S)
{
    double w = 0.0;
    for (int i = 0; i < S; ++i)
        if (st[i].active && st[i].tol2 > 0.0) w = std::max(w, st[i].rtrans / st[i].tol2);
    return w;
}

// One solve: state reset (plain copies), Prepare, then iteration graphs.  The first launch
// sequence holds the previous solve's iteration count (or kBatchFirst); after it the host reads
// the states and plans the rest from the worst system's residual rate (ChunkPlanner on
// r.r / tol^2, whose tolerance is 1).  Launches past the last convergence cost ~50 us per
// iteration, so the plan aims at the last one.
static int batch_solve(hpccg_ctx *c, Batch *B, const double *b, double *x, double rel_tol, int max_iter,
                       int replace_every, int v, double *ms_out)
{
    const int S = B->S;
    std::vector<BatchState> init(S);
    for (auto &s : init) {
        std::memset(&s, 0, sizeof(s));
        s.replace_every = replace_every;
        s.max_iter = max_iter;
    }
    HPCCG_HIP_TRY(hipMemcpyAsync(B->st, init.data(), S * sizeof(BatchState), hipMemcpyHostToDevice, c->stream));
    static const int zero[2] = {0, 0};  // active count, x0 flag
    HPCCG_HIP_TRY(hipMemcpyAsync(B->nactive, zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
    const double rel_tol2 = rel_tol * rel_tol;
    // iteration graphs bake b, x and the variant (tolerances and limits live in the states)
    auto get = [&](int size, hipGraphExec_t *ge) {
        return cached_graph(B->graphs, c->stream, GraphKey{b, x, 300 + v, size},
                            [&]() { return enqueue_variant(c, B, v, b, x, size, rel_tol2); }, ge);
    };
    HPCCG_TRY(capture_iteration_graphs(kBatchCap, max_iter, get));
    BatchState *prep = B->snap, *now = B->snap + S;
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HPCCG_TRY(enqueue_variant(c, B, v, b, x, -1, rel_tol2));
    HPCCG_HIP_TRY(hipMemcpyAsync(prep, B->st, S * sizeof(BatchState), hipMemcpyDeviceToHost, c->stream));
    ChunkPlanner plan{kBatchCap, 1.0, 1};
    int launched = 0, m = std::min(max_iter, B->hint > 0 ? B->hint : kBatchFirst);
    bool first = true;
    while (m > 0) {
        HPCCG_TRY(launch_iterations(c->stream, m, get));
        launched += m;
        if (launched >= max_iter) break;
        HPCCG_HIP_TRY(hipMemcpyAsync(now, B->st, S * sizeof(BatchState), hipMemcpyDeviceToHost, c->stream));
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        if (first) plan.next(worst_ratio(prep, S), 0, max_iter);
        first = false;
        const double w = worst_ratio(now, S);
        if (w == 0.0) break;  // every system has stopped
        m = plan.next(w, launched, max_iter - launched);
    }
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(now, B->st, S * sizeof(BatchState), hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (ms_out) *ms_out = ms;
    int need = 0;
    for (int i = 0; i < S; ++i) need = std::max(need, now[i].niters);
    B->hint = need;
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
    if (dmalloc(&B->rowindex, (B->E + 1) * sizeof(int32_t)) != hipSuccess ||
        dmalloc(&B->colindex, B->nnz * sizeof(int32_t)) != hipSuccess ||
        dmalloc(&B->values, B->nnz * sizeof(__half)) != hipSuccess ||
        dmalloc(&B->r, B->E * sizeof(float)) != hipSuccess || dmalloc(&B->p, B->E * sizeof(float)) != hipSuccess ||
        dmalloc(&B->Ap, B->E * sizeof(float)) != hipSuccess ||
        dmalloc(&B->st, count * sizeof(BatchState)) != hipSuccess) {
        free_batched(c);
        set_error("batched: device allocation failed");
        return HPCCG_ERR_NOMEM;
    }
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
    if (dmalloc(&B->slice_ptr, sp.size() * sizeof(int64_t)) != hipSuccess ||
        dmalloc(&B->code, std::max<int64_t>(code_bytes(B->slots, soff[count]), 4)) != hipSuccess ||
        dmalloc(&B->tabs, (size_t)count * kBatchTab * sizeof(int32_t)) != hipSuccess ||
        dmalloc(&B->sval, std::max<int64_t>(B->slots, 1) * sizeof(__half)) != hipSuccess ||
        dmalloc(&B->slice_off, (count + 1) * sizeof(int)) != hipSuccess) {
        free_batched(c);
        set_error("batched: device allocation of the sliced copy failed");
        return HPCCG_ERR_NOMEM;
    }
    B->NS = soff[count];
    std::vector<SliceInfo> sinfo(std::max(B->NS, 1), SliceInfo{0, 0, 0});
    for (int s = 0; s < count; ++s)
        for (int i = soff[s]; i < soff[s + 1]; ++i) {
            const int64_t r0 = (int64_t)(i - soff[s]) * kSlice, rows = B->row_off_h[s + 1] - B->row_off_h[s];
            sinfo[i] = SliceInfo{B->row_off_h[s] + r0, s, (int32_t)std::min<int64_t>(kSlice, rows - r0)};
        }
    if (dmalloc(&B->sinfo, sinfo.size() * sizeof(SliceInfo)) != hipSuccess ||
        dmalloc(&B->partials, 2 * sinfo.size() * sizeof(double)) != hipSuccess ||
        dmalloc(&B->nactive, 2 * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&B->snap, 2 * (size_t)count * sizeof(BatchState), hipHostMallocDefault) != hipSuccess) {
        free_batched(c);
        set_error("batched: device allocation failed");
        return HPCCG_ERR_NOMEM;
    }
    HPCCG_HIP_TRY(hipMemcpy(B->sinfo, sinfo.data(), sinfo.size() * sizeof(SliceInfo), hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(B->slice_ptr, sp.data(), sp.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(B->slice_off, soff.data(), (count + 1) * sizeof(int), hipMemcpyHostToDevice));
    // each system's table: the distinct stencil offsets dz*nx*ny + dy*nx + dx, sorted
    std::vector<int32_t> tabs((size_t)count * kBatchTab);
    std::vector<int> ntab(count);
    for (int s = 0; s < count; ++s) {
        const int32_t nx = dims[3 * s], ny = dims[3 * s + 1];
        std::vector<int32_t> t;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) t.push_back(dz * nx * ny + dy * nx + dx);
        std::sort(t.begin(), t.end());
        t.erase(std::unique(t.begin(), t.end()), t.end());
        ntab[s] = (int)t.size();
        for (int i = 0; i < kBatchTab; ++i) tabs[(size_t)s * kBatchTab + i] = t[std::min(i, ntab[s] - 1)];
    }
    int *flag = nullptr;
    HPCCG_HIP_TRY(dmalloc(&flag, sizeof(int)));
    int bad = 0;
    hipError_t e = hipMemcpyAsync(B->tabs, tabs.data(), tabs.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(int), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(B->code, 0, std::max<int64_t>(code_bytes(B->slots, B->NS), 4), c->stream);
    for (int s = 0; s < count && e == hipSuccess; ++s) {
        const int ns = soff[s + 1] - soff[s];
        const int64_t threads = (int64_t)ns * kSlice;
        // codes of system s: offsets relative to its own rows (col_shift = its first row)
        hipLaunchKernelGGL((k_csr_to_sell<kGroupF16, __half, __half>), dim3((unsigned)((threads + 255) / 256)),
                           dim3(256), 0, c->stream, B->rowindex + B->row_off_h[s], B->colindex, B->values,
                           B->slice_ptr + soff[s], (int32_t *)nullptr, B->sval, B->row_off_h[s + 1] - B->row_off_h[s],
                           ns, flag, B->code + kCodePad * soff[s], B->tabs + (size_t)s * kBatchTab, ntab[s],
                           (int32_t)B->row_off_h[s]);
        e = hipGetLastError();
    }
    if (e == hipSuccess && B->NS > 0) {
        hipLaunchKernelGGL(k_mark_uniform<>, dim3((unsigned)(((int64_t)B->NS * kSlice + 255) / 256)), dim3(256), 0,
                           c->stream, B->slice_ptr, B->code, B->NS);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(flag);
    if (e != hipSuccess || bad) {
        free_batched(c);
        set_error(e != hipSuccess ? "batched: sliced copy: %s" : "batched: a column offset is missing from its table%s",
                  e != hipSuccess ? hipGetErrorString(e) : "");
        return HPCCG_ERR_HIP;
    }
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
    std::vector<BatchState> init(B->S);
    for (auto &s : init) {
        std::memset(&s, 0, sizeof(s));
        s.active = 1;
        s.rtrans = 1.0;
    }
    HPCCG_HIP_TRY(hipMemcpyAsync(B->st, init.data(), B->S * sizeof(BatchState), hipMemcpyHostToDevice, c->stream));
    const int all = B->S;
    HPCCG_HIP_TRY(hipMemcpyAsync(B->nactive, &all, sizeof(int), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL((kb_spmv<4, 1>), dim3((unsigned)((B->NS + 3) / 4)), dim3(256), 0, c->stream, batch_sell(B), x, y,
                       B->st, B->partials, B->nactive);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_batch_variants(int32_t *count, int32_t *rows, int32_t *waves)
{
    if (count) *count = kNumVariants;
    for (int v = 0; v < kNumVariants; ++v) {
        if (rows) rows[v] = kSlice * kVariants[v].spw;
        if (waves) waves[v] = kVariants[v].waves;
    }
    return 0;
}

// Times every compiled launch variant on this batch (solves of `probe_iters` iterations from
// the given b and x, x restored afterwards; see the score below), keeps the fastest for
// variant = -1 solves, and returns it with the per-variant scores (ms).
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
    HPCCG_HIP_TRY(dmalloc(&x0, B->E * sizeof(double)));
    HPCCG_HIP_TRY(hipMemcpyAsync(x0, x, B->E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    // Score = time of probe_iters working iterations + time of probe_iters iterations after
    // every system has converged (rel_tol 1e100 stops them all at Prepare).  A solve captures
    // max_iter iterations, so the idle ones past the last convergence count too, and the
    // variants with more workgroups launch them more slowly.
    auto probe = [&](int v, double rel_tol, double *out) -> int {
        double t = 0.0, ms = 0.0;
        for (int rep = 0; rep < 3; ++rep) {
            HPCCG_HIP_TRY(hipMemcpyAsync(x, x0, B->E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            HPCCG_TRY(batch_solve(c, B, b, x, rel_tol, probe_iters, 0, v, &ms));
            if (rep > 0) t += ms;  // the first run captures and warms
        }
        *out = t / 2;
        return 0;
    };
    std::vector<double> score(kNumVariants, 0.0);
    int st = 0;
    for (int v = 0; v < kNumVariants && st == 0; ++v) st = probe(v, 0.0, &score[v]);
    for (int v = 0; v < kNumVariants && st == 0; ++v) {
        double idle = 0.0;
        st = probe(v, 1e100, &idle);
        score[v] += idle;
    }
    if (st) {
        dfree(x0);
        return st;
    }
    int bi = 0;
    for (int v = 0; v < kNumVariants; ++v) {
        if (variant_ms) variant_ms[v] = score[v];
        if (score[v] < score[bi]) bi = v;
    }
    HPCCG_HIP_TRY(hipMemcpyAsync(x, x0, B->E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(x0);
    B->graphs.clear();
    B->hint = 0;  // the probes' counts say nothing about a real solve
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
    const int v = variant >= 0 ? variant : (B->best_variant >= 0 ? B->best_variant : 0);
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
