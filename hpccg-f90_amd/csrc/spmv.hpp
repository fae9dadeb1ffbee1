// CSR SpMV building blocks shared by the FP64, mixed-precision, batched and learned paths.
//
// The reference's CSR (1-based rowindex / colindex / values, SparseMatrix.F90:44-64) is the
// interface; on the device every SpMV reads a sliced-ELL copy of it built once per upload:
//
//   * rows are cut into slices of 64 consecutive rows (one wavefront, one lane per row);
//   * slice s is K_s entries wide (K_s = its longest row) and occupies the slot range
//     [slice_ptr[s], slice_ptr[s] + 64 K_s);
//   * inside a slice the entries are stored "column-major" in groups of G per lane: entry j of
//     the slice's lane l sits at slot  base + ((j / G) * 64 + l) * G + j % G  for the first
//     (K_s / G) * G entries and at  base + KG * 64 + (j - KG) * 64 + l  after them (KG =
//     (K_s / G) * G), so that one load instruction of the wave reads 64 * G consecutive
//     entries (fully coalesced) and each lane gets G consecutive entries of its own row;
//   * columns are stored 0-based; slots past a row's end hold value 0 and a valid column;
//   * slot offsets are multiples of 64, so bit 0 of slice_ptr[s] is free: it flags a "full"
//     slice (64 rows, all exactly K_s long), whose row lengths need no rowindex reads;
//   * column codes: when the matrix has at most kMaxCodes distinct offsets col - row (the
//     27-point stencil has 27), columns are not stored as int32 but as one byte per slot, an
//     index into the sorted offset table (column = row + table[code]).  The codes of slice s
//     start at byte slice_base(slice_ptr[s]) + kCodePad * s; entry j of lane l is byte
//     ((j / 4) * 64 + l) * 4 + j % 4 there, so one 4-byte load per lane fetches four codes.
//     That cuts the index stream from 4 to 1 byte per slot (FP64 slot 12 -> 9 bytes, FP16
//     6 -> 3) and serves the FP64 and FP16 copies alike.  A matrix with more distinct
//     offsets keeps int32 columns.  Bit 1 of slice_ptr[s] flags a "uniform" coded slice (full,
//     and every lane carries lane 0's codes: the interior rows of a stencil), whose lanes
//     all read lane 0's code words, so the other lanes' code lines are never fetched.
//
// Each lane sums its own row in ascending entry order from 0.0 and skips the padding slots
// (the row length comes from rowindex), so y = A x is bitwise the reference's row loop
// (SparseMatrix.F90:225-229) for FP64 values -- and for FP16 values it is the FP64 row sum
// of the exactly widened products.  Compared with a CSR kernel this layout needs no LDS
// staging of the matrix and no per-row scatter: the value and column (or code) streams are
// read with unit stride and the x gathers of neighbouring lanes touch neighbouring rows'
// columns.
#pragma once

#include <hip/hip_fp16.h>

#include "common.hpp"

namespace hpccg {

constexpr int kSlice = 64;            // rows per slice == lanes per wave
constexpr int kSellWaves = 4;         // slices per SpMV workgroup (one per wave)
constexpr int kSellThreads = kSlice * kSellWaves;
// true-residual kernels (they run only every few iterations and exit at once otherwise): 4
// slices per workgroup.  16 made the empty launches cheaper but a real one slower (256^3:
// 0.75 ms against 0.66 ms), and a real one costs more than all empty ones of a solve.
constexpr int kTrueWaves = 4;
constexpr int kTrueThreads = kSlice * kTrueWaves;
// ... and each true-residual workgroup walks kTrueLoop consecutive groups of kTrueWaves
// slices, so that the launches that find nothing due start 4x fewer workgroups (256^3: an
// empty launch 15 -> 4.9 us, a real one 0.44 -> 0.48 ms; mixed CG has ~67 empty ones and 3 real)
constexpr int kTrueLoop = 4;
constexpr int kStencilK = 27;         // slice width of the 27-point stencil: fully unrolled path
constexpr int kGroupF64 = 2;          // FP64 copy: 2 entries per lane per load (16-byte value loads)
// FP16 copy: 1 entry per lane per load.  Groups of 2, 4 and 8 entries per lane (4-, 8- and
// 16-byte loads) measured the same at 256^3: mixed SpMV 0.283-0.288 ms for every group size,
// the same solve times, bitwise the same results.
constexpr int kGroupF16 = 1;
// FP16-copy SpMVs of the CG loops (km_spmv, kf_spmv) are compiled for 7 waves per SIMD (at most
// 72 VGPRs).  Measured at 256^3: 0.293 ms with the compiler's own choice (80 VGPRs, 6 waves),
// 0.279 ms with 7, 0.316 ms with 5 and 0.421 ms with 8.  The FP64 SpMV is fastest at its own 5.
constexpr int kSellF16WavesPerEU = 7;
constexpr int kVecThreads = 256;
constexpr int kMaxGrid = 16384;       // partial slots of the vector kernels' grid reductions

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

typedef int i32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));

// G consecutive entries of one lane (G = 1 or 2; the slot is G-aligned)
template <int G>
__device__ __forceinline__ void ld_group(const int32_t *p, int32_t *o)
{
    if constexpr (G == 1) {
        o[0] = ld_stream(p);
    } else {
        static_assert(G == 2, "group of 1 or 2");
        const i32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const i32x2_t *>(p));
        o[0] = v.x;
        o[1] = v.y;
    }
}
template <int G>
__device__ __forceinline__ void ld_group(const double *p, double *o)
{
    if constexpr (G == 1) {
        o[0] = ld_stream(p);
    } else {
        static_assert(G == 2, "group of 1 or 2");
        const f64x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f64x2_t *>(p));
        o[0] = v.x;
        o[1] = v.y;
    }
}
template <int G>
__device__ __forceinline__ void ld_group(const __half *p, __half *o)
{
    static_assert(G == 1, "FP16 copy uses single-entry groups");
    o[0] = ld_stream(p);
}

// slot of entry j of lane `lane` in a slice of width K starting at slot `base`
template <int G>
__host__ __device__ __forceinline__ int64_t sell_slot(int64_t base, int K, int j, int lane)
{
    const int KG = (K / G) * G;
    if (j < KG) return base + ((int64_t)(j / G) * kSlice + lane) * G + (j % G);
    return base + (int64_t)KG * kSlice + (int64_t)(j - KG) * kSlice + lane;
}

constexpr int64_t kSliceFull = 1;                       // flag bit of slice_ptr[s]
constexpr int64_t kSliceUniform = 2;  // flag bit: coded full slice whose lanes all carry lane 0's codes
__host__ __device__ __forceinline__ int64_t slice_base(int64_t sp) { return sp & ~(int64_t)(kSlice - 1); }

// Row sum (FP64, ascending entry order from 0.0) of global row `row` held by lane `lane` of
// slice `s`; 0 when !valid.  rowptr is the 1-based CSR row pointer (row lengths).
template <int G, typename VT, typename XT>
__device__ __forceinline__ double sell_rowsum(const int64_t *__restrict__ slice_ptr,
                                              const int32_t *__restrict__ scol,
                                              const VT *__restrict__ sval,
                                              const int32_t *__restrict__ rowptr,
                                              const XT *__restrict__ x, int s, int lane, int64_t row,
                                              bool valid)
{
    const int64_t sp0 = slice_ptr[s], base = slice_base(sp0);
    const int K = (int)((slice_base(slice_ptr[s + 1]) - base) / kSlice);
    const int len = !valid ? 0 : (sp0 & kSliceFull) ? K : rowptr[row + 1] - rowptr[row];
    double sum = 0.0;
    if (K == kStencilK) {
        // branch-free: all column and value loads issue back to back, then the gathers
        constexpr int NG = kStencilK / G, KG = NG * G;
        int32_t c[kStencilK];
        VT v[kStencilK];
#pragma unroll
        for (int g = 0; g < NG; ++g) ld_group<G>(scol + base + ((int64_t)g * kSlice + lane) * G, c + g * G);
#pragma unroll
        for (int j = KG; j < kStencilK; ++j) c[j] = ld_stream(scol + base + (int64_t)j * kSlice + lane);
#pragma unroll
        for (int g = 0; g < NG; ++g) ld_group<G>(sval + base + ((int64_t)g * kSlice + lane) * G, v + g * G);
#pragma unroll
        for (int j = KG; j < kStencilK; ++j) v[j] = ld_stream(sval + base + (int64_t)j * kSlice + lane);
#pragma unroll
        for (int j = 0; j < kStencilK; ++j) {
            const double prod = to_f64(v[j]) * to_f64(x[c[j]]);
            sum = (j < len) ? sum + prod : sum;
        }
    } else {
        for (int j = 0; j < K; ++j) {
            const int64_t q = sell_slot<G>(base, K, j, lane);
            const double prod = to_f64(sval[q]) * to_f64(x[scol[q]]);
            sum = (j < len) ? sum + prod : sum;
        }
    }
    return sum;
}

// ---- column codes ----------------------------------------------------------------------------
constexpr int kMaxCodes = 256;              // distinct offsets col - row a coded copy can hold
constexpr int64_t kCodePad = 3 * kSlice;    // code bytes per slice beyond its slots (K rounded up to 4)
constexpr int kOffsetSlots = 4096;          // hash set that collects the offsets
constexpr int32_t kNoOffset = INT32_MIN;    // empty slot (|col - row| < 2^31 - 1)
__host__ __device__ __forceinline__ int64_t code_base(int64_t slot_base, int s) { return slot_base + kCodePad * s; }
__host__ __device__ __forceinline__ int64_t code_bytes(int64_t slots, int nslices) { return slots + kCodePad * nslices; }
__host__ __device__ __forceinline__ int64_t code_byte(int j, int lane) { return ((int64_t)(j >> 2) * kSlice + lane) * 4 + (j & 3); }

// Flags (kSliceUniform) each full slice of a coded copy whose 64 lanes all carry lane 0's
// codes.  One wave per slice; only that wave writes slice_ptr[s], and the flag bits never
// change the slot offsets other waves read from it.
template <typename Unused = void>
__global__ void k_mark_uniform(int64_t *slice_ptr, const uint8_t *__restrict__ code, int nslices)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int s = (int)(t / kSlice), lane = (int)(t % kSlice);
    if (s >= nslices) return;
    const int64_t sp0 = slice_ptr[s], base = slice_base(sp0);
    if (!(sp0 & kSliceFull)) return;
    const int K = (int)((slice_base(slice_ptr[s + 1]) - base) / kSlice);
    const uint32_t *cw = reinterpret_cast<const uint32_t *>(code + code_base(base, s));
    int same = 1;
    for (int g = 0; g < (K + 3) / 4; ++g) same &= cw[g * kSlice + lane] == cw[g * kSlice];
    if (__all(same) && lane == 0) slice_ptr[s] = sp0 | kSliceUniform;
}

// index of `off` in the sorted table tab[0..n) (n >= 1); tab[result] != off if it is absent
__device__ __forceinline__ int column_code(const int32_t *__restrict__ tab, int n, int32_t off)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (tab[mid] < off) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Row sum as sell_rowsum, over a coded copy: the column of a slot is row + lookup(code), the
// offset table entry (from LDS, or from another lane's register).  Every lane of the wave runs
// the lookups.  Rows with no entries gather x[0] for their (skipped) padding.
template <int G, typename VT, typename XT, typename Lookup>
__device__ __forceinline__ double sell_rowsum_coded(const int64_t *__restrict__ slice_ptr,
                                                    const uint8_t *__restrict__ code, Lookup lookup,
                                                    const VT *__restrict__ sval, const int32_t *__restrict__ rowptr,
                                                    const XT *__restrict__ x, int s, int lane, int row, bool valid)
{
    const int64_t sp0 = slice_ptr[s], base = slice_base(sp0);
    const int K = (int)((slice_base(slice_ptr[s + 1]) - base) / kSlice);
    const int len = !valid ? 0 : (sp0 & kSliceFull) ? K : rowptr[row + 1] - rowptr[row];
    // a uniform slice (interior rows of a stencil) reads lane 0's codes only: one address per load
    const uint32_t *cw = reinterpret_cast<const uint32_t *>(code + code_base(base, s)) + ((sp0 & kSliceUniform) ? 0 : lane);
    double sum = 0.0;
    if (K == kStencilK) {
        constexpr int NW = (kStencilK + 3) / 4, NG = kStencilK / G, KG = NG * G;
        uint32_t w[NW];
#pragma unroll
        for (int g = 0; g < NW; ++g) w[g] = __builtin_nontemporal_load(cw + g * kSlice);
        VT v[kStencilK];
#pragma unroll
        for (int g = 0; g < NG; ++g) ld_group<G>(sval + base + ((int64_t)g * kSlice + lane) * G, v + g * G);
#pragma unroll
        for (int j = KG; j < kStencilK; ++j) v[j] = ld_stream(sval + base + (int64_t)j * kSlice + lane);
#pragma unroll
        for (int j = 0; j < kStencilK; ++j) {
            const int off = lookup((w[j >> 2] >> (8 * (j & 3))) & 0xff);
            const int c = len > 0 ? row + off : 0;
            const double prod = to_f64(v[j]) * to_f64(x[c]);
            sum = (j < len) ? sum + prod : sum;
        }
    } else {
        uint32_t wj = 0;
        for (int j = 0; j < K; ++j) {
            if ((j & 3) == 0) wj = __builtin_nontemporal_load(cw + (j >> 2) * kSlice);
            const int off = lookup((wj >> (8 * (j & 3))) & 0xff);
            const int c = len > 0 ? row + off : 0;
            const double prod = to_f64(sval[sell_slot<G>(base, K, j, lane)]) * to_f64(x[c]);
            sum = (j < len) ? sum + prod : sum;
        }
    }
    return sum;
}

// A sliced-ELL copy as kernels take it (by value): int32 columns (col) or, when code is set,
// column codes with their offset table (ntab entries).
template <typename VT>
struct SellRef {
    const int64_t *slice_ptr;  // nslices+1 slot offsets
    const int32_t *col;        // 0-based columns (null in a coded copy)
    const VT *val;
    const int32_t *rowptr;     // the 1-based CSR row pointer (row lengths)
    int nrow, nslices;
    const uint8_t *code;       // column codes, or null
    const int32_t *tab;        // sorted offsets col - row
    int ntab;
};

// One wave per slice, WAVES slices per workgroup: the FP64 row sum of this thread's row (0 past
// the last row or slice), with the row and whether it exists.
// Workgroup -> slice-group numbering.  The hardware deals workgroups to the 8 XCDs round-robin
// (workgroup b runs on XCD b % 8).  Renumbered, each XCD takes runs of kXcdRun consecutive
// slice groups (256 x 4 slices x 64 rows = 65,536 rows: one 256 x 256 plane of a 27-point
// grid), so the x lines a group gathers are mostly those its XCD's L2 fetched for the groups
// just before it.  Workgroups past the last whole span of 8 runs keep their number.  Measured
// on the FP64 SpMV at 256^3: 1.000 -> 0.973 ms (runs of 64: no change; of 2,048 or more: slower).
constexpr int kXcdRun = 256;
__device__ __forceinline__ int sell_block()
{
    const int b = blockIdx.x, span = kNumXcd * kXcdRun;
    if (b >= ((int)gridDim.x / span) * span) return b;
    return (b / span) * span + (b % kNumXcd) * kXcdRun + (b / kNumXcd) % kXcdRun;
}

// The row of this thread in slice group g of a one-wave-per-slice launch (WAVES slices per
// group), whether it exists, and its slice.
template <int WAVES>
__device__ __forceinline__ int sell_row_of_thread(int nslices, int nrow, int g, int64_t &row, bool &valid)
{
    const int s = g * WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x / kSlice);
    row = (int64_t)s * kSlice + threadIdx.x % kSlice;
    valid = s < nslices && row < nrow;
    return s;
}

// Sets *flag when x[0..n) holds an entry other than +-0.0 (a NaN counts).  The flag lives in a
// solver's control block, which the host zeroes before Prepare.  A solve that starts from
// x = 0 then skips Prepare's A x: every row sum of A 0 is +0.0 (products +-0 added to 0.0), so
// r = b - A x is formed from b alone with the same bits.
template <typename Unused = void>
__global__ __launch_bounds__(kVecThreads) void k_nonzero(const double *__restrict__ x, int64_t n, int *flag)
{
    bool nz = false;
    for (int64_t i = blockIdx.x * (int64_t)kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        nz |= !(x[i] == 0.0);
    if (__any(nz) && threadIdx.x % kWave == 0) atomicOr(flag, 1);
}

// XRUN: number the workgroups with sell_block().  Used by the FP64 SpMV and the mixed-precision
// FP16 SpMV (coded copy at 256^3: mixed CG 36.0 -> 35.6 ms); the flexible-CG SpMV, which reads
// the preconditioner's output just written, keeps hardware order (22.2 -> 22.4 ms with it).
template <int G, typename VT, typename XT, int WAVES = kSellWaves, bool XRUN = false>
__device__ __forceinline__ double sell_rowsum_of_thread(const SellRef<VT> &A, const XT *__restrict__ x, int64_t &row,
                                                        bool &valid, int group = -1, bool load_table = true)
{
    // the offset table of a coded copy goes to LDS first (every thread of the workgroup gets
    // here; a workgroup that walks several groups loads it on the first call only)
    __shared__ int32_t tab[kMaxCodes];
    if (A.code && load_table) {
        for (int i = threadIdx.x; i < A.ntab; i += blockDim.x) tab[i] = A.tab[i];
        __syncthreads();
    }
    // slice group: `group` when given, else this workgroup's (renumbered when XRUN)
    const int g = group >= 0 ? group : XRUN ? sell_block() : (int)blockIdx.x;
    const int s = sell_row_of_thread<WAVES>(A.nslices, A.nrow, g, row, valid);
    const int lane = threadIdx.x % kSlice;
    if (s >= A.nslices) return 0.0;
    if (A.code)
        return sell_rowsum_coded<G, VT, XT>(A.slice_ptr, A.code, [&](int c) { return tab[c]; }, A.val, A.rowptr, x, s,
                                            lane, (int)row, valid);
    return sell_rowsum<G, VT, XT>(A.slice_ptr, A.col, A.val, A.rowptr, x, s, lane, row, valid);
}

// CSR -> sliced ELL for the `nslices` slices of rows [0, nrow) of a 1-based CSR (rowptr,
// col, val).  One thread per slice row.  Columns go to scol (0-based) when it is set and as
// codes into `codes` (offset table tab[0..ntab), sorted) when that is set; a code stands for
// the offset (col - 1 - col_shift) - row, so that columns of a sub-block (rows and columns
// shifted by col_shift in a larger matrix) are coded relative to that block.  A VT narrower
// than double sets bit 0 of *flags for every value that does not convert exactly; an offset
// missing from tab sets bit 1.
template <int G, typename VT, typename VS = double>
__global__ void k_csr_to_sell(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                              const VS *__restrict__ val, const int64_t *__restrict__ slice_ptr,
                              int32_t *__restrict__ scol, VT *__restrict__ sval, int64_t nrow, int nslices,
                              int *flags, uint8_t *__restrict__ codes, const int32_t *__restrict__ tab, int ntab,
                              int32_t col_shift = 0)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)nslices * kSlice) return;
    const int s = (int)(t / kSlice), lane = (int)(t % kSlice);
    const int64_t base = slice_base(slice_ptr[s]);
    const int K = (int)((slice_base(slice_ptr[s + 1]) - base) / kSlice);
    uint8_t *cb = codes ? codes + code_base(base, s) : nullptr;
    int len = 0, padc = 0, bad = 0, padcode = 0;
    auto put_code = [&](int j, int c) {
        const int32_t off = c - col_shift - (int32_t)t;
        const int code = column_code(tab, ntab, off);
        if (tab[code] != off) bad |= 2;
        cb[code_byte(j, lane)] = (uint8_t)code;
        return code;
    };
    if (t < nrow) {
        const int64_t k0 = rowptr[t] - 1;
        len = rowptr[t + 1] - rowptr[t];
        for (int j = 0; j < len; ++j) {
            const int64_t q = sell_slot<G>(base, K, j, lane);
            const double v = to_f64(val[k0 + j]);
            if (scol) scol[q] = col[k0 + j] - 1;
            if (cb) padcode = put_code(j, col[k0 + j] - 1);
            const VT w = (VT)v;
            sval[q] = w;
            if (to_f64(w) != v) bad |= 1;
        }
        padc = len > 0 ? col[k0 + len - 1] - 1 : 0;
    }
    for (int j = len; j < K; ++j) {
        const int64_t q = sell_slot<G>(base, K, j, lane);
        if (scol) scol[q] = padc;
        if (cb) cb[code_byte(j, lane)] = (uint8_t)(len > 0 ? padcode : 0);
        sval[q] = (VT)0.0;
    }
    if (bad && flags) atomicOr(flags, bad);
}

// Host: slot offsets (with the full-slice flags) of the slices of rows [0, nrow) of a 1-based
// CSR row pointer; returns the total slot count.  sp must hold ceil(nrow/64)+1 entries and
// sp[0] = base (a multiple of 64) on entry; sp[ns] carries no flag.
inline int64_t sell_slice_ptr(const int32_t *rowptr, int64_t nrow, int64_t base, int64_t *sp)
{
    const int64_t ns = (nrow + kSlice - 1) / kSlice;
    sp[0] = base;
    for (int64_t s = 0; s < ns; ++s) {
        int K = 0, kmin = 1 << 30;
        const int64_t r1 = nrow < (s + 1) * kSlice ? nrow : (s + 1) * kSlice;
        for (int64_t r = s * kSlice; r < r1; ++r) {
            const int len = rowptr[r + 1] - rowptr[r];
            K = K > len ? K : len;
            kmin = kmin < len ? kmin : len;
        }
        sp[s + 1] = slice_base(sp[s]) + (int64_t)K * kSlice;
        if (r1 - s * kSlice == kSlice && kmin == K) sp[s] |= kSliceFull;
    }
    return sp[ns] - base;
}

// Grid reductions of the one-wave-per-slice kernels.  Those launch one workgroup per
// kSellWaves slices (65,536 workgroups at 256^3), too many for a last-workgroup ticket: its
// device-scope release fence writes back the L2 of the workgroup's XCD, and tens of thousands
// of them cost more than the SpMV.  Instead each workgroup stores its total in part[blockIdx.x]
// (plain store; the kernel boundary makes it visible) and a one-workgroup finish kernel sums
// the partials in a fixed order -- deterministic, fence-free.
constexpr int kFinishThreads = 1024;

__device__ __forceinline__ void store_partial(double block_total, double *part)
{
    if (threadIdx.x == 0) part[blockIdx.x] = block_total;
}

// Sum of part[0..n) by the single workgroup of a finish kernel (kFinishThreads threads); every
// thread gets the total.  `red` needs kFinishThreads/64 doubles of LDS.  Each thread sums its
// strided share of partial pairs with eight independent 16-byte loads in flight, in a fixed
// order.  A 65,536-partial sum is bound by one CU's load latency: 7.4-7.8 us at 256^3, against
// 10.6-11.1 us with 8-byte loads; sixteen loads in flight gain nothing more.  part must be
// 16-byte aligned.
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, int n, double *red)
{
    constexpr int U = 8;
    const f64x2_t *p2 = reinterpret_cast<const f64x2_t *>(part);
    const int n2 = n / 2;
    double s[U] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int i = threadIdx.x;
    for (; i + (U - 1) * kFinishThreads < n2; i += U * kFinishThreads) {
        f64x2_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = p2[i + u * kFinishThreads];
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = (s[u] + v[u].x) + v[u].y;
    }
#pragma unroll
    for (int u = 0; u < U - 1; ++u)
        if (i + u * kFinishThreads < n2) {
            const f64x2_t v = p2[i + u * kFinishThreads];
            s[u] = (s[u] + v.x) + v.y;
        }
    if ((n & 1) && threadIdx.x == 0) s[0] += part[n - 1];
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) t += s[u];
    return block_sum<kFinishThreads>(t, red);
}

// Element range [i0, i1) of one (XCD-renumbered) workgroup in a contiguous split of n.
__device__ __forceinline__ void elem_range(int64_t n, int64_t &i0, int64_t &i1, int &lb)
{
    const int G = gridDim.x;
    lb = xcd_logical_block(blockIdx.x, G);
    i0 = (n * lb) / G;
    i1 = (n * (lb + 1)) / G;
}

// Element loops of the FP32 vector kernels: the workgroup's range starts on a multiple of 4
// (the last logical workgroup takes the remainder) so that its body reads and writes whole
// 16-byte float4s; each thread takes every kVecThreads-th group of four, then a scalar tail.
__device__ __forceinline__ void elem_range4(int64_t n, int64_t &i0, int64_t &i1, int &lb)
{
    const int G = gridDim.x;
    lb = xcd_logical_block(blockIdx.x, G);
    const int64_t q = n / 4;
    i0 = 4 * ((q * lb) / G);
    i1 = lb + 1 == G ? n : 4 * ((q * (lb + 1)) / G);
}
template <typename F4, typename F1>
__device__ __forceinline__ void vec4_loop(int64_t i0, int64_t i1, F4 body4, F1 body1)
{
    const int64_t iv = i0 + ((i1 - i0) & ~(int64_t)3);
    for (int64_t i = i0 + 4 * (int64_t)threadIdx.x; i < iv; i += 4 * (int64_t)blockDim.x) body4(i);
    for (int64_t i = iv + threadIdx.x; i < i1; i += blockDim.x) body1(i);
}
typedef float f32x4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f32x4v ld4(const float *p) { return *(const f32x4v *)p; }
// streamed operand (read once here, next used after the matrix has passed through the caches):
// non-temporal, so that it does not displace the vector the next SpMV gathers
__device__ __forceinline__ f32x4v ld4_nt(const float *p) { return __builtin_nontemporal_load((const f32x4v *)p); }
__device__ __forceinline__ void st4(float *p, f32x4v v) { *(f32x4v *)p = v; }

// Launch shapes (host).  SpMV kernels launch one wave per slice, kSellWaves slices per
// workgroup, in slice order (the hardware dispatches workgroups in order, so the waves in
// flight stream a compact window of the matrix); the vector kernels use a resident grid.
inline int sell_blocks(int64_t nslices, int waves = kSellWaves) { return (int)((nslices + waves - 1) / waves); }
inline int true_blocks(int64_t nslices)
{
    return nslices > 0 ? (sell_blocks(nslices, kTrueWaves) + kTrueLoop - 1) / kTrueLoop : 1;
}
inline int vec_grid_for(int num_cu, int64_t n)
{
    int64_t G = (n + kVecThreads - 1) / kVecThreads;
    if (G > (int64_t)num_cu * 4) G = (int64_t)num_cu * 4;
    G = (G / kNumXcd) * kNumXcd;
    return G < kNumXcd ? kNumXcd : (int)G;
}

}  // namespace hpccg
