// CSR SpMV building blocks shared by the FP64, mixed-precision and batched paths.
#pragma once

#include <hip/hip_fp16.h>

#include "common.hpp"

namespace hpccg {

constexpr int kTileRows = 64;  // rows per SpMV tile == threads per SpMV workgroup (one wave)
constexpr int kVecThreads = 256;
constexpr int kMaxGrid = 16384;

// FP64 SpMV kernels: LDS (27 FP64 products per row, 13.8 KB per 64-row tile) admits 11
// workgroups per CU, i.e. 3 waves per SIMD; cap the registers so they do not cut that to 2.
#define HPCCG_SPMV_OCC __attribute__((amdgpu_waves_per_eu(3)))

__device__ __forceinline__ double to_f64(double v) { return v; }
__device__ __forceinline__ double to_f64(float v) { return (double)v; }
__device__ __forceinline__ double to_f64(__half v) { return (double)__half2float(v); }

// streaming (non-temporal) loads: matrix values and columns are read once per SpMV
__device__ __forceinline__ double ld_stream(const double *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ int32_t ld_stream(const int32_t *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ __half ld_stream(const __half *p)
{
    return __ushort_as_half(__builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p)));
}

// One tile of a CSR SpMV: rows [tile*ROWS, tile*ROWS+ROWS) of a 1-based CSR.
//
// Phase 1 streams the tile's contiguous (value, column) range with unit-stride (coalesced)
// loads, gathers x, and stages the products, formed and rounded in FP64, in LDS.  Phase 2
// sums each row's products in ascending k from 0.0 (the reference's order,
// SparseMatrix.F90:225-229).  Tiles whose nnz exceed the LDS capacity (rows longer than 27 on
// average; never the case for generate_matrix's CSR) take a direct per-row loop with the same
// order.  VT = stored value type (double, or __half holding {27,-1} exactly), XT = vector
// type.  With VT = double and XT = double the result is bitwise the reference's spmv.
// Returns the row sum (FP64) of this thread's row, or 0 past the last row.
//
// The phase-1 loads are branch-free: a slot past the tile's last element re-reads that last
// element (a valid address) and its product is dropped, so all UNR value and column loads of
// a thread issue back to back.  UNR = 27 covers a full tile of the 27-point stencil in one pass.
template <int ROWS, typename VT, typename XT, int UNR = 27>
__device__ __forceinline__ double spmv_tile_rowsum(const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col,
                                                   const VT *__restrict__ val,
                                                   const XT *__restrict__ x, int nrow, int tile,
                                                   double *lds, bool &valid, int &row)
{
    const int t = threadIdx.x;
    const int r0 = tile * ROWS;
    const int r = r0 + t;
    const int rlast = min(r0 + ROWS, nrow);
    const int base = rowptr[r0] - 1;
    const int end = rowptr[rlast] - 1;
    int ms = 0, me = 0;
    valid = r < nrow;
    row = r;
    if (valid) {
        ms = rowptr[r] - 1;
        me = rowptr[r + 1] - 1;
    }
    double sum = 0.0;
    if (end - base <= ROWS * 27) {
        const int last = end - 1;  // only read when end > base
        for (int k0 = base; k0 < end; k0 += UNR * ROWS) {
            VT v[UNR];
            int c[UNR];
            if (sizeof(VT) == 8) {
                // FP64 values: all column loads first (their gathers can start while the values
                // are in flight; this order also needs fewer registers)
#pragma unroll
                for (int u = 0; u < UNR; ++u) c[u] = ld_stream(col + min(k0 + t + u * ROWS, last));
#pragma unroll
                for (int u = 0; u < UNR; ++u) v[u] = ld_stream(val + min(k0 + t + u * ROWS, last));
            } else {
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    const int kk = min(k0 + t + u * ROWS, last);
                    v[u] = ld_stream(val + kk);
                    c[u] = ld_stream(col + kk);
                }
            }
            // all gathers before the first LDS store: x is a generic pointer the compiler must
            // assume may alias LDS, so an interleaved store would serialise the gathers
            XT xg[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) xg[u] = x[c[u] - 1];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int kk = k0 + t + u * ROWS;
                const double prod = to_f64(v[u]) * to_f64(xg[u]);
                if (kk < end) lds[kk - base] = prod;
            }
        }
        __syncthreads();
        for (int j = ms; j < me; ++j) sum = sum + lds[j - base];
        __syncthreads();
    } else {
        for (int j = ms; j < me; ++j) {
            const double prod = to_f64(val[j]) * to_f64(x[col[j] - 1]);
            sum = sum + prod;
        }
    }
    return sum;
}

// Contiguous tile range of one (XCD-renumbered) workgroup.
__device__ __forceinline__ void block_range(int ntiles, int &t0, int &t1, int &lb)
{
    const int G = gridDim.x;
    lb = xcd_logical_block(blockIdx.x, G);
    t0 = (int)(((int64_t)lb * ntiles) / G);
    t1 = (int)(((int64_t)(lb + 1) * ntiles) / G);
}

// Element range [i0, i1) of one (XCD-renumbered) workgroup in a contiguous split of n.
__device__ __forceinline__ void elem_range(int64_t n, int64_t &i0, int64_t &i1, int &lb)
{
    const int G = gridDim.x;
    lb = xcd_logical_block(blockIdx.x, G);
    i0 = (n * lb) / G;
    i1 = (n * (lb + 1)) / G;
}

// Launch shapes (host).  SpMV kernels run one resident wave of workgroups (CUs x blocks that
// fit per CU, from the occupancy calculator), each walking a contiguous tile range, so there
// is no partial second wave; the count is a multiple of 8 for the XCD renumbering.
inline int spmv_grid_for(int num_cu, int ntiles, int blocks_per_cu)
{
    int G = ntiles < num_cu * blocks_per_cu ? ntiles : num_cu * blocks_per_cu;
    G = (G / kNumXcd) * kNumXcd;
    return G < kNumXcd ? kNumXcd : G;
}
template <typename K>
inline int resident_blocks(K kernel, int block)
{
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, 0) != hipSuccess || occ < 1) occ = 1;
    return occ;
}
// grid for kernel `k` launched with kTileRows threads over ntiles tiles
#define HPCCG_SPMV_GRID(k, num_cu, ntiles) \
    ::hpccg::spmv_grid_for((num_cu), (ntiles), ::hpccg::resident_blocks((k), kTileRows))
inline int vec_grid_for(int num_cu, int64_t n)
{
    int64_t G = (n + kVecThreads - 1) / kVecThreads;
    if (G > (int64_t)num_cu * 4) G = (int64_t)num_cu * 4;
    G = (G / kNumXcd) * kNumXcd;
    return G < kNumXcd ? kNumXcd : (int)G;
}

}  // namespace hpccg
