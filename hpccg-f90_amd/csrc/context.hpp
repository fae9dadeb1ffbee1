// Context object behind the opaque hpccg_ctx handle.
#pragma once

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <tuple>

#include "common.hpp"
#include "spmv.hpp"

namespace hpccg {

// Device-resident CG control block.  The iteration loop reads and updates it entirely on
// the device, so a captured graph needs no host round trip per iteration.
struct CgState {
    double alpha;      // alpha of the last iteration that computed p.Ap
    double beta;       // beta for the next p update
    double tol;        // absolute tolerance on the recursive residual norm
    int k;             // iteration to run next (1-based, numerics.F90:108)
    int stop_at;       // first iteration that does not run (max_iter+1 or lagged test)
    int max_iter;
    int breakdown;     // p.Ap was zero or non-finite
    int x_nonzero;     // set by Prepare's k_nonzero when x0 != 0 (else A x0 is skipped)
    double rr;         // ||r_{k-1}||^2, the last recorded r.r
};

// Sizes the graph launches of an iterative solve.  Each iteration skips itself on the device
// once the solve has stopped, but a skipped iteration still costs its launches (~0.1 ms at
// 256^3); so after every launch the host reads ||r|| and enqueues only the iterations it
// predicts are left, from the geometric rate of the last launch, plus `lag` (iterations that
// run after ||r|| < tol).  A short prediction costs one more small launch.
struct ChunkPlanner {
    int cap;     // most iterations per launch
    double tol;  // absolute tolerance on ||r||; <= 0 or NaN: run to the iteration limit
    int lag;
    double rr0 = -1.0;
    int k0 = 0;
    double rate = 0.0;  // per-iteration factor of ||r||
    // rr = r.r after iteration k; left = iterations before the limit
    int next(double rr, int k, int left)
    {
        if (rr0 > 0.0 && rr > 0.0 && k > k0) rate = std::pow(rr / rr0, 0.5 / (double)(k - k0));
        rr0 = rr;
        k0 = k;
        int m = std::min(cap, left);
        if (tol > 0.0 && rate > 0.0 && rate < 1.0 && rr > 0.0) {
            const double need = std::log(tol / std::sqrt(rr)) / std::log(rate);
            if (need < (double)m) m = std::max(0, (int)std::ceil(need)) + lag;
        }
        return std::max(1, std::min(m, std::min(cap, left)));
    }
};

// Captures the graphs of 1, 2, 4, ... <= min(cap, max_iter) iterations (get(size, &exec)
// returns the cached graph of `size` iterations, capturing it on first use); called before
// the timed region and before anything is queued on the stream.
template <typename Get>
int capture_iteration_graphs(int cap, int max_iter, Get get)
{
    for (int size = 1; size <= cap && size <= max_iter; size *= 2) {
        hipGraphExec_t ge = nullptr;
        HPCCG_TRY(get(size, &ge));
    }
    return 0;
}

// Enqueues m iterations as graphs of power-of-two sizes, all captured beforehand.
template <typename Get>
int launch_iterations(hipStream_t s, int m, Get get)
{
    for (int bit = 30; bit >= 0; --bit) {
        if (!(m & (1 << bit))) continue;
        hipGraphExec_t ge = nullptr;
        HPCCG_TRY(get(1 << bit, &ge));
        HPCCG_HIP_TRY(hipGraphLaunch(ge, s));
    }
    return 0;
}

struct Matrix {
    int32_t nrow = 0, ncol = 0;
    int64_t nnz = 0;
    int32_t max_row_nnz = 0;
    int32_t *rowindex = nullptr;  // 1-based, nrow+1 (the reference's CSR, kept on the device)
    int32_t *colindex = nullptr;  // 1-based, nnz
    double *values = nullptr;
    int32_t grid[3] = {0, 0, 0};  // nx, ny, nz when built by hpccg_generate_matrix

    // sliced-ELL copies read by every SpMV kernel (layout in spmv.hpp)
    int32_t nslices = 0;          // ceil(nrow / 64)
    int64_t slots = 0;            // slots of one copy (>= nnz; padding of short rows)
    int64_t *slice_ptr = nullptr; // nslices+1 slot offsets, shared by both copies
    uint8_t *code = nullptr;      // column codes shared by both copies (null: int32 columns)
    int32_t *ctab = nullptr;      // their sorted offset table (kMaxCodes entries)
    int32_t ncode = 0;            // offsets in the table
    int32_t *scol64 = nullptr;    // FP64 copy: 0-based columns, 2 entries per lane group (uncoded only)
    double *sval64 = nullptr;
    int32_t *scol16 = nullptr;    // FP16 copy (mixed / learned paths, built on first use)
    __half *sval16 = nullptr;
    double *red_part = nullptr;   // per-workgroup partials of the one-wave-per-slice kernels
};

inline SellRef<double> sell64(const Matrix &A)
{
    return SellRef<double>{A.slice_ptr, A.scol64, A.sval64, A.rowindex, A.nrow, A.nslices, A.code, A.ctab, A.ncode};
}
inline SellRef<__half> sell16(const Matrix &A)
{
    return SellRef<__half>{A.slice_ptr, A.scol16, A.sval16, A.rowindex, A.nrow, A.nslices, A.code, A.ctab, A.ncode};
}
// workgroups of a one-wave-per-slice launch (at least one, so reductions always run)
inline int sell_grid(const Matrix &A, int waves = kSellWaves) { return A.nslices > 0 ? sell_blocks(A.nslices, waves) : 1; }

struct GraphKey {
    const void *b, *x;
    int mode, chunk;
    bool operator<(const GraphKey &o) const
    {
        return std::tie(b, x, mode, chunk) < std::tie(o.b, o.x, o.mode, o.chunk);
    }
};

// Instantiated graphs keyed by the caller's vectors and the launch shape, each stamped with the
// buffer generation (common.hpp) at capture.  find() returns a graph only if no library buffer
// has been allocated or freed since it was captured; older graphs are destroyed on the way.
struct GraphCache {
    struct Entry {
        hipGraphExec_t exec;
        uint64_t gen;
    };
    std::map<GraphKey, Entry> map;

    hipGraphExec_t find(const GraphKey &k)
    {
        const uint64_t g = buffer_generation();
        for (auto it = map.begin(); it != map.end();) {
            if (it->second.gen != g) {
                hipGraphExecDestroy(it->second.exec);
                it = map.erase(it);
            } else {
                ++it;
            }
        }
        auto it = map.find(k);
        return it == map.end() ? nullptr : it->second.exec;
    }
    void put(const GraphKey &k, hipGraphExec_t e) { map[k] = Entry{e, buffer_generation()}; }
    void clear()
    {
        for (auto &e : map) hipGraphExecDestroy(e.second.exec);
        map.clear();
    }
};

// Captures what enqueue() puts on stream s into an instantiated graph.  enqueue must not
// allocate (a capture records no allocation, and the graph would be stale at once).
template <typename F>
int capture_graph(hipStream_t s, F enqueue, hipGraphExec_t *out)
{
    const uint64_t g0 = buffer_generation();
    hipGraph_t g;
    HPCCG_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = enqueue();
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (st) {
        if (e == hipSuccess) hipGraphDestroy(g);
        return st;
    }
    HPCCG_HIP_TRY(e);
    const hipError_t e2 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HPCCG_HIP_TRY(e2);
    if (buffer_generation() != g0) {
        hipGraphExecDestroy(*out);
        *out = nullptr;
        set_error("graph capture: a buffer was allocated while capturing");
        return HPCCG_ERR_STATE;
    }
    return 0;
}

// The cached graph of `key`, captured by enqueue() on first use (or after any buffer change).
template <typename F>
int cached_graph(GraphCache &cache, hipStream_t s, const GraphKey &key, F enqueue, hipGraphExec_t *out)
{
    if ((*out = cache.find(key))) return 0;
    HPCCG_TRY(capture_graph(s, enqueue, out));
    cache.put(key, *out);
    return 0;
}

// Workspace of the mixed-precision path (FP16 values, FP32 vectors, FP64 dots and x).
struct MixedWs {
    int64_t n = 0;
    float *r = nullptr, *p = nullptr, *Ap = nullptr;
    void *state = nullptr;    // device MixState
    void *state_h = nullptr;  // pinned host mirror
    double *hist = nullptr;   // [2][hist_cap]: recursive and true r.r per iteration
    int hist_cap = 0;
    GraphCache graphs;
};

}  // namespace hpccg

struct hpccg_ctx {
    int device = 0;
    hpccg::DeviceInfo info;
    hipStream_t stream = nullptr;
    int dot_mode = HPCCG_DOT_TREE;
    hpccg::Matrix A;

    // CG workspace (sized to A.nrow)
    int64_t ws_n = 0;
    double *r = nullptr, *p = nullptr, *Ap = nullptr;
    hpccg::CgState *state = nullptr;    // device
    hpccg::CgState *state_h = nullptr;  // pinned host mirror
    double *rr_hist = nullptr;          // device, hist_cap doubles
    int hist_cap = 0;
    double *partials = nullptr;         // device, kMaxGrid doubles
    double *scalar = nullptr;           // device result slot for standalone dots
    double *scalar_h = nullptr;         // pinned host

    // staging for the host-array API
    int64_t stage_n = 0;
    double *s0 = nullptr, *s1 = nullptr, *s2 = nullptr;

    hpccg::GraphCache graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev2 = nullptr, ev3 = nullptr;  // an outer solve that calls cg_device_fp64 (fall-back)
    hipEvent_t evp = nullptr;                 // end of the FP64 CG's Prepare
    hipEvent_t evs0 = nullptr, evs1 = nullptr;  // the SpMV of one timed CG iteration
    double *scratch = nullptr;  // this device's static workspace of the no-allocation solve

    hpccg::MixedWs mix;
    void *ext[8] = {};  // owned by the batched / learned paths
};

namespace hpccg {
void free_mixed(hpccg_ctx *c);
void free_batched(hpccg_ctx *c);
void free_learned(hpccg_ctx *c);
// drop the learned path's captured graphs (memory only: stale graphs are never replayed)
void invalidate_learned_graphs(hpccg_ctx *c);
int ensure_sell16(hpccg_ctx *c);
int cg_device_fp64(hpccg_ctx *c, const double *b, double *x, int max_iter, double tol, int *niters,
                   double *normr, double *hist, double *elapsed_ms);
// FP64 cg (tree dots) on device arrays without graphs, calling hook(k) after the residual
// update of every iteration k that ran (k = 0 after Prepare).  During hook(k): c->r holds r_k,
// c->p holds p_k, c->state->alpha holds alpha_k, and x holds x_{k-1} (x_k = x + alpha_k p_k;
// the x update is applied at the top of the next iteration).  A nonzero hook return stops.
int cg_trace(hpccg_ctx *c, const double *b, double *x, int max_iter, double tol,
             const std::function<int(int)> &hook);
int launch_spmv(hpccg_ctx *c, const double *x, double *y);
}  // namespace hpccg
