// Matrix-free form of generate_matrix's operator (MPI/src/SparseMatrix.F90:138-175): y = A x
// for the 27-point stencil on an nx x ny x nz grid -- 27 on the diagonal, -1 for every
// neighbour inside the grid -- computed from the grid extents alone, without the CSR or its
// sliced copy.  No reference counterpart (the reference always multiplies by its stored CSR);
// it is the comparison point for the CSR SpMV: it moves x once and y, 16 bytes per row,
// against the CSR's 12 bytes per nonzero (the coded copy's 9).  Each row is summed from 0.0
// over its neighbours in generate_matrix's column order (z, then y, then x offsets ascending),
// every product value * x rounded before the add (built with -ffp-contract=off), so y is
// bitwise the CSR SpMV's (SparseMatrix.F90:225-229) on that matrix.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "context.hpp"

namespace hpccg {

constexpr int kStencilThreads = 256;

// One thread per row, rows in x-fastest order, so a wave's 27 neighbour loads are 27
// unit-stride runs that the caches serve after the first touch.
__global__ __launch_bounds__(kStencilThreads) void k_stencil_spmv(int nx, int ny, int nz,
                                                                  const double *__restrict__ x,
                                                                  double *__restrict__ y)
{
    const int plane = nx * ny;
    const int64_t n = (int64_t)plane * nz;
    const int64_t r64 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r64 >= n) return;
    const int row = (int)r64;
    const int iz = row / plane, rem = row - iz * plane, iy = rem / nx, ix = rem - iy * nx;
    double sum = 0.0;
#pragma unroll
    for (int sz = -1; sz <= 1; ++sz) {
#pragma unroll
        for (int sy = -1; sy <= 1; ++sy) {
#pragma unroll
            for (int sx = -1; sx <= 1; ++sx) {
                const bool in = iz + sz >= 0 && iz + sz < nz && iy + sy >= 0 && iy + sy < ny && ix + sx >= 0 &&
                                ix + sx < nx;
                const double v = (sz == 0 && sy == 0 && sx == 0) ? 27.0 : -1.0;
                const int64_t col = (int64_t)row + (int64_t)sz * plane + (int64_t)sy * nx + sx;
                const double prod = v * x[in ? col : (int64_t)row];
                sum = in ? sum + prod : sum;
            }
        }
    }
    y[row] = sum;
}

// Grids with nx a multiple of R: a thread owns R consecutive rows of one x-line and reads each
// of its 9 neighbour lines once -- the R values above its rows (16-byte loads) and the one on
// either side -- instead of 27 loads per row.  Each row's sum keeps the order above.  R = 4 is
// used (256^3: 0.110 ms against 0.205 ms for one row per thread and 0.134 ms for R = 8).
template <int R>
__global__ __launch_bounds__(kStencilThreads) void k_stencil_spmv_rows(int nx, int ny, int nz,
                                                                       const double *__restrict__ x,
                                                                       double *__restrict__ y)
{
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const int plane = nx * ny;
    const int64_t nr = (int64_t)plane * nz / R;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nr) return;
    const int row = (int)(R * t);
    const int iz = row / plane, rem = row - iz * plane, iy = rem / nx, ix = rem - iy * nx;
    double sum[R];
#pragma unroll
    for (int j = 0; j < R; ++j) sum[j] = 0.0;
#pragma unroll
    for (int sz = -1; sz <= 1; ++sz) {
#pragma unroll
        for (int sy = -1; sy <= 1; ++sy) {
            const bool line = iz + sz >= 0 && iz + sz < nz && iy + sy >= 0 && iy + sy < ny;
            const int64_t base = line ? (int64_t)row + (int64_t)sz * plane + (int64_t)sy * nx : (int64_t)row;
            // w[1 + i] = x at line column ix + i; the outer neighbours w[0], w[R + 1] are read at
            // clamped, always valid addresses (unused where off the line)
            double w[R + 2];
            w[0] = x[ix > 0 ? base - 1 : base];
            w[R + 1] = x[ix + R < nx ? base + R : base];
#pragma unroll
            for (int i = 0; i < R / 2; ++i) {
                const f64x2 v2 = *reinterpret_cast<const f64x2 *>(x + base + 2 * i);
                w[1 + 2 * i] = v2.x;
                w[2 + 2 * i] = v2.y;
            }
#pragma unroll
            for (int j = 0; j < R; ++j) {
#pragma unroll
                for (int sx = -1; sx <= 1; ++sx) {
                    const bool in = line && ix + j + sx >= 0 && ix + j + sx < nx;
                    const double v = (sz == 0 && sy == 0 && sx == 0) ? 27.0 : -1.0;
                    const double prod = v * w[j + sx + 1];
                    sum[j] = in ? sum[j] + prod : sum[j];
                }
            }
        }
    }
    f64x2 *yo = reinterpret_cast<f64x2 *>(y + row);
#pragma unroll
    for (int i = 0; i < R / 2; ++i) yo[i] = f64x2{sum[2 * i], sum[2 * i + 1]};
}

static void launch_stencil(hpccg_ctx *c, int nx, int ny, int nz, const double *x, double *y)
{
    const int64_t n = (int64_t)nx * ny * nz;
    const bool al = reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0;
    if (al && nx % 4 == 0)
        hipLaunchKernelGGL(k_stencil_spmv_rows<4>, dim3((unsigned)((n / 4 + kStencilThreads - 1) / kStencilThreads)),
                           dim3(kStencilThreads), 0, c->stream, nx, ny, nz, x, y);
    else
        hipLaunchKernelGGL(k_stencil_spmv, dim3((unsigned)((n + kStencilThreads - 1) / kStencilThreads)),
                           dim3(kStencilThreads), 0, c->stream, nx, ny, nz, x, y);
}

}  // namespace hpccg

using namespace hpccg;

extern "C" {

int hpccg_spmv_stencil_dev(hpccg_ctx *c, int32_t nx, int32_t ny, int32_t nz, const double *x, double *y)
{
    HPCCG_CHECK_ARG(c && x && y && nx >= 1 && ny >= 1 && nz >= 1, "bad argument");
    HPCCG_CHECK_ARG((int64_t)nx * ny * nz < INT32_MAX, "grid too large for int32 rows");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    launch_stencil(c, nx, ny, nz, x, y);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_time_spmv_stencil(hpccg_ctx *c, int32_t nx, int32_t ny, int32_t nz, const double *x, double *y,
                            int32_t reps, double *avg_ms)
{
    HPCCG_CHECK_ARG(c && x && y && reps > 0 && avg_ms && nx >= 1 && ny >= 1 && nz >= 1, "bad argument");
    HPCCG_CHECK_ARG((int64_t)nx * ny * nz < INT32_MAX, "grid too large for int32 rows");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    launch_stencil(c, nx, ny, nz, x, y);  // warm
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; ++i) launch_stencil(c, nx, ny, nz, x, y);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / reps;
    return 0;
}

}  // extern "C"
