// Shared device/host helpers for libhpccg_hip (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "hpccg_hip.h"

namespace hpccg {

// ------------------------------------------------------------------------------------------
// error reporting: thread-local message + negative status codes
void set_error(const char *fmt, ...);

#define HPCCG_HIP_TRY(call)                                                                    \
    do {                                                                                       \
        hipError_t _e = (call);                                                                \
        if (_e != hipSuccess) {                                                                \
            ::hpccg::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(_e)); \
            return HPCCG_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

#define HPCCG_CHECK_ARG(cond, msg)                                                             \
    do {                                                                                       \
        if (!(cond)) {                                                                         \
            ::hpccg::set_error("%s: %s", __func__, msg);                                       \
            return HPCCG_ERR_ARG;                                                              \
        }                                                                                      \
    } while (0)

#define HPCCG_TRY(call)                                                                        \
    do {                                                                                       \
        int _s = (call);                                                                       \
        if (_s < 0) return _s;                                                                 \
    } while (0)

// Device memory of the library's own buffers (matrix copies, workspaces, control blocks,
// activations).  Every allocation and every free bumps one process-wide generation counter.
// A captured graph holds the addresses of the buffers it was captured with; graph caches
// stamp each graph with the generation at capture and replay it only while the generation is
// unchanged, so no graph can run on a buffer that was freed or reallocated after it was made.
// Caller-owned vectors (hpccg_dev_malloc) are part of the cache keys instead.
uint64_t buffer_generation();
hipError_t dmalloc_bytes(void **p, size_t bytes);
template <typename T>
inline hipError_t dmalloc(T **p, size_t bytes)
{
    return dmalloc_bytes(reinterpret_cast<void **>(p), bytes);
}
hipError_t dfree(void *p);

constexpr int kWave = 64;
constexpr int kNumXcd = 8;  // MI355X: 8 XCDs, each with its own L2; workgroup i -> XCD i % 8

// Mixed-precision solvers keep r in FP32: the recursive residual drifts from b - A x by about
// 1e-7 of the residual norm at the last true-residual replacement, so it cannot show a 1e-10
// reduction by itself.  Besides the periodic replacement, the FP64 true residual is formed
// again whenever the recursive r.r has fallen below this fraction of the last true r.r (a
// 1e5-fold drop in the norm; each true residual lowers the drift floor by the same factor).
constexpr double kReplaceDrop2 = 1e-10;

// Workgroups are dealt to XCDs round-robin.  Renumber them so that XCD x owns the contiguous
// logical range [x*G/8, (x+1)*G/8): neighbouring row tiles (which share x-vector planes)
// then run on the same XCD and hit in that XCD's L2.  G must be a multiple of 8.
__device__ __forceinline__ int xcd_logical_block(int b, int G)
{
    return (b % kNumXcd) * (G / kNumXcd) + b / kNumXcd;
}

// ------------------------------------------------------------------------------------------
// reductions (fixed shape => deterministic)
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide sum; every thread gets the total.  `red` needs blockDim/64 doubles of LDS.
template <int THREADS>
__device__ __forceinline__ double block_sum(double v, double *red)
{
    constexpr int W = THREADS / kWave;
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (W == 1) return v;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < W; ++i) t += red[i];
    return t;
}

// Device-wide launch facts, queried once per context.
struct DeviceInfo {
    int num_cu = 256;
    int gfx = 950;
};

}  // namespace hpccg
