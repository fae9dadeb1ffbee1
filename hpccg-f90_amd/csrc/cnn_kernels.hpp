// Learned preconditioner: BF16 implicit-GEMM 3x3x3 convolution kernels on CDNA4 matrix cores.
//
// Network (bias-free, so positively homogeneous: net(c*u) = c*net(u) for c > 0):
//   h1 = relu(conv(u; W1[16][27])), h2 = relu(conv(h1; W2[16][16][27])), z = conv(h2; W3[16][27])
// conv(in; W)[co](v) = sum_{ci,t} W[co][ci][t] * in[ci](v + off(t)), off(t) = (t%3-1, t/3%3-1,
// t/9-1) in (x, y, z), zero outside the grid.  Activations are channel-last BF16
// ([batch][voxel][16], 32 bytes per voxel), accumulation is FP32 (MFMA).
//
// Implicit GEMM on v_mfma_f32_16x16x32_bf16 (D[16x16] += A[16x32] * B[32x16]):
//   M = output channels (16), N = 16 voxels, K = 2 taps x 16 input channels per MFMA, or
//   K = 27(+5) taps for a 1-channel input.  16-channel K order ("z-major"): 15 K-steps
//   s = 5 (dz+1) + p, each pairing the in-plane taps 2p and 2p+1 of the 3 x 3 taps of plane
//   offset dz (the tenth in-plane slot is zero), so that the five operands of one input plane
//   serve all three output planes it touches.
//   The one-channel output of layer 3 is formed as three partial sums, one per plane offset
//   dz, in MFMA rows 0..2 (row 1+dz takes only the steps of offset dz), then added in FP32 in
//   the order ((P(-1) + P(0)) + P(+1)).
//   A = weights, pre-packed per lane and held in registers for the whole kernel.
//   B = activations from an LDS halo tile: each lane reads the 8 channels it needs as one
//   16-byte LDS load.
// Operand layout per lane l (wave64): A[m = l%16][k = 8*(l/16) + j], B[k = 8*(l/16) + j][n = l%16],
// D[m = 4*(l/16) + i][n = l%16], j = 0..7, i = 0..3.
//
// Tiles: 8 x 8 x 4 output voxels per 256-thread workgroup (4 waves, one z-plane each, four
// 16-voxel groups per wave), halo 10 x 10 x 6 staged in LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.hpp"

namespace hpccg {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int TX = 8, TY = 8, TZ = 4;
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
constexpr int HV = HX * HY * HZ;  // 600 halo voxels
constexpr int TV = TX * TY * TZ;  // 256 voxels per tile
constexpr int NSTEP16 = 15;       // MFMA K-steps for 16 input channels x 27 taps (z-major)
constexpr int NSTEP3F = 5;        // in-plane tap pairs of one input plane (z-major K order)
// Banded form of the one-channel layer 3 (fused inference and the layer-wise forward alike): one
// MFMA block is L3ROWS = 2 output rows x 16 voxels (the N columns) of one h2 plane, with K = a
// window of 4 h2 rows x 3 x-offsets x 16 channels = 192 = 6 K-steps.  K-step s pairs the window
// (row, x-offset) (s / 3, s % 3) for operand rows k = 0..15 with (s / 3 + 2, s % 3) for k = 16..31.
// MFMA row m = 4 dzi + j (dzi < 3, j < 2) holds the plane-offset dz = dzi - 1 partial of
// output row j, whose taps are window rows j .. j+2 (dy = row - j - 1); the other rows are zero.
// So a lane holds in register j the dz = c - 1 partial of row j (c = lane / 16), and an output
// voxel is ((P(dz=-1) + P(0)) + P(+1)) of three blocks on consecutive h2 planes.
constexpr int NSTEP3B = 6;
constexpr int L3ROWS = 2;
constexpr int NW1 = 16 * 27, NW2 = 16 * 16 * 27, NW3 = 16 * 27;
constexpr int NWALL = NW1 + NW2 + NW3;  // 7776 parameters

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// two floats -> two bfloat16 in one word (low half = a), round to nearest even: one
// v_cvt_pk_bf16_f32 on gfx950
__device__ inline uint32_t pack_bf16(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

// float -> bfloat16, round to nearest even (inf/nan pass through, NaN quieted); the device
// form is the hardware conversion, the host form the same rounding on the bit pattern
__host__ __device__ inline uint16_t f2bf(float f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint16_t)(pack_bf16(f, 0.f) & 0xffffu);
#else
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
#endif
}
__host__ __device__ inline float bf2f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ inline int tap_dx(int t) { return t % 3 - 1; }
__device__ inline int tap_dy(int t) { return (t / 3) % 3 - 1; }
__device__ inline int tap_dz(int t) { return t / 9 - 1; }

// K order of a one-channel input (27 taps in K = 32): lane row c holds the x-triplets of tap
// lines 2c and 2c+1 in slots 0-5 (tap 6c + j), and the last line (taps 24-26) fills slots 6-7
// of rows 0 and 1; -1 marks a zero slot.  A line is (dz, dy) with dx = -1, 0, 1 consecutive, so a
// lane reads whole triplets.
__host__ __device__ inline int k1_tap(int c, int j)
{
    if (j < 6) return 6 * c + j;
    if (c == 0) return 24 + (j - 6);
    if (c == 1 && j == 6) return 26;
    return -1;
}

// Tap of K-step s (z-major order) for the lanes of operand row half `hi` (0: taps 2p, 1:
// taps 2p+1); -1 for the zero slot.
__host__ __device__ inline int zk_tap(int s, int hi)
{
    const int dz1 = s / 5, t9 = 2 * (s % 5) + hi;
    return t9 < 9 ? 9 * dz1 + t9 : -1;
}

__device__ inline f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ReLU of two BF16 values in one word: a BF16 number is negative exactly when its bit pattern
// is negative as an int16, so one packed integer max with 0 zeroes the negatives (-0 included).
// Rounding to BF16 preserves signs, so bf16(max(a, 0)) == max(bf16(a), 0) for every non-NaN a.
typedef short i16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t relu_pk(uint32_t p)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2, p), i16x2{0, 0}));
}
// BF16 then ReLU of the four accumulator rows of a lane, packed as 4 x 16 bits
__device__ inline u32x2 relu_bf16x4(f32x4 a)
{
    u32x2 pk;
    pk[0] = relu_pk(pack_bf16(a[0], a[1]));
    pk[1] = relu_pk(pack_bf16(a[2], a[3]));
    return pk;
}

struct Grid3 {
    int nx, ny, nz;
    int tx, ty, tz;  // tiles per axis
    __host__ __device__ int64_t nvox() const { return (int64_t)nx * ny * nz; }
    __host__ __device__ int ntiles() const { return tx * ty * tz; }
};

inline Grid3 make_grid(int nx, int ny, int nz)
{
    return Grid3{nx, ny, nz, (nx + TX - 1) / TX, (ny + TY - 1) / TY, (nz + TZ - 1) / TZ};
}

// Input of a convolution: BF16 channel-last activations, or (layer 1) an FP32 field multiplied
// by a per-sample scale and rounded to BF16 while it is staged.
enum InKind { IN_BF16_16 = 0, IN_BF16_1 = 1, IN_F32_1 = 2 };
// Epilogue: ReLU -> BF16 [v][16]; masked BF16 (the backward pass through a ReLU whose forward
// output is `mask`).  The one-channel layer 3 of the forward pass has its own kernel (k_conv3).
enum OutKind { OUT_RELU_BF16 = 0, OUT_MASKED_BF16 = 2 };

// Stage the 10x10x6 halo of tile (ox,oy,oz) of sample b into LDS (channel-last, 16 or 1 ch).
template <int IK>
__device__ inline void stage_halo(const void *in, float scale, uint16_t *sm, const Grid3 &g, int b, int ox, int oy,
                                  int oz)
{
    const int64_t N = g.nvox();
    if (IK == IN_BF16_16) {
        const uint16_t *src = (const uint16_t *)in;
        for (int i = threadIdx.x; i < HV * 2; i += blockDim.x) {
            const int hv = i >> 1, ch = i & 1;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = ox - 1 + hx, gy = oy - 1 + hy, gz = oz - 1 + hz;
            u32x4 v = {0u, 0u, 0u, 0u};
            if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz) {
                const int64_t gv = ((int64_t)gz * g.ny + gy) * g.nx + gx;
                v = *(const u32x4 *)(src + ((int64_t)b * N + gv) * 16 + ch * 8);
            }
            *(u32x4 *)(sm + hv * 16 + ch * 8) = v;
        }
    } else {
        for (int hv = threadIdx.x; hv < HV; hv += blockDim.x) {
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = ox - 1 + hx, gy = oy - 1 + hy, gz = oz - 1 + hz;
            uint16_t v = 0;
            if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz) {
                const int64_t gv = (int64_t)b * N + ((int64_t)gz * g.ny + gy) * g.nx + gx;
                v = IK == IN_F32_1 ? f2bf(((const float *)in)[gv] * scale) : ((const uint16_t *)in)[gv];
            }
            sm[hv] = v;
        }
    }
}

// One 3x3x3 convolution layer with 16 output channels, or the backward pass through one.
// Grid: one workgroup per (tile, sample).
template <int IK, int OK>
__global__ __launch_bounds__(256) void k_conv(const void *__restrict__ in, const float *__restrict__ in_scale,
                                              const bf16x8 *__restrict__ wpack, void *__restrict__ out,
                                              const uint16_t *__restrict__ mask, const float *__restrict__ out_scale,
                                              Grid3 g, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int CIN = IK == IN_BF16_16 ? 16 : 1;
    constexpr int NSTEP = CIN == 16 ? NSTEP16 : 1;
    __shared__ __attribute__((aligned(16))) uint16_t sm[HV * CIN];
    const int nt = g.ntiles();
    const int b = blockIdx.x / nt, tile = blockIdx.x % nt;
    const int ox = (tile % g.tx) * TX, oy = ((tile / g.tx) % g.ty) * TY, oz = (tile / (g.tx * g.ty)) * TZ;
    const float scale = (IK == IN_F32_1 && in_scale) ? in_scale[b] : 1.0f;
    stage_halo<IK>(in, scale, sm, g, b, ox, oy, oz);

    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int c = lane / 16, n = lane % 16;
    bf16x8 A[NSTEP];
#pragma unroll
    for (int s = 0; s < NSTEP; ++s) A[s] = wpack[s * 64 + lane];
    __syncthreads();

    const int64_t N = g.nvox();
    const int z = w;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const int x = n % 8, y = 2 * q + n / 8;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        if (CIN == 16) {
#pragma unroll
            for (int s = 0; s < NSTEP; ++s) {
                const int t = zk_tap(s, c / 2);
                u32x4 raw = {0u, 0u, 0u, 0u};
                if (t >= 0) {
                    const int hv = (z + tap_dz(t) + 1) * (HY * HX) + (y + tap_dy(t) + 1) * HX + (x + tap_dx(t) + 1);
                    raw = *(const u32x4 *)(sm + hv * 16 + (c & 1) * 8);
                }
                acc = mfma16(A[s], __builtin_bit_cast(bf16x8, raw), acc);
            }
        } else {
            uint16_t v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int t = k1_tap(c, j);
                v[j] = 0;
                if (t >= 0) {
                    const int hv = (z + tap_dz(t) + 1) * (HY * HX) + (y + tap_dy(t) + 1) * HX + (x + tap_dx(t) + 1);
                    v[j] = sm[hv];
                }
            }
            u32x4 raw;
            raw[0] = v[0] | ((uint32_t)v[1] << 16);
            raw[1] = v[2] | ((uint32_t)v[3] << 16);
            raw[2] = v[4] | ((uint32_t)v[5] << 16);
            raw[3] = v[6] | ((uint32_t)v[7] << 16);
            acc = mfma16(A[0], __builtin_bit_cast(bf16x8, raw), acc);
        }
        // epilogue: lane holds channels 4c..4c+3 of voxel (x, y, z) of this tile
        const int gx = ox + x, gy = oy + y, gz = oz + z;
        if (gx < g.nx && gy < g.ny && gz < g.nz) {
            const int64_t gv = (int64_t)b * N + ((int64_t)gz * g.ny + gy) * g.nx + gx;
            {
                u32x2 pk;
                if (OK == OUT_RELU_BF16) {
                    pk = relu_bf16x4(acc);
                } else {
                    const u32x2 mm = *(const u32x2 *)(mask + gv * 16 + 4 * c);
                    const uint16_t m4[4] = {(uint16_t)(mm[0] & 0xffff), (uint16_t)(mm[0] >> 16),
                                            (uint16_t)(mm[1] & 0xffff), (uint16_t)(mm[1] >> 16)};
                    uint16_t o[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[i] = bf2f(m4[i]) > 0.f ? f2bf(acc[i]) : (uint16_t)0;
                    pk[0] = o[0] | ((uint32_t)o[1] << 16);
                    pk[1] = o[2] | ((uint32_t)o[3] << 16);
                }
                *(u32x2 *)((uint16_t *)out + gv * 16 + 4 * c) = pk;
            }
        }
    }
}

// The one-channel layer 3 of the forward pass, z = conv(h2; W3) times a per-sample scale, in the
// banded form (NSTEP3B): the MFMA columns of a tile are (x, row block, h2 plane) -- 8 x 4 x 6 =
// 192, twelve groups of 16, three per wave -- each six MFMAs on the staged h2 halo; the
// partials go to LDS and every output voxel adds its three in the fused kernel's order,
// ((P(-1) + P(0)) + P(+1)), so this is bitwise the fused inference.
__global__ __launch_bounds__(256) void k_conv3(const uint16_t *__restrict__ in, const bf16x8 *__restrict__ wpack,
                                               float *__restrict__ out, const float *__restrict__ out_scale, Grid3 g,
                                               const int *__restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int NB = TY / L3ROWS;  // row blocks per tile
    __shared__ __attribute__((aligned(16))) uint16_t sm[HV * 16];
    __shared__ float P[HZ][NB][L3ROWS][3][TX];  // partial of (h2 plane, block, row, dz + 1, x)
    const int nt = g.ntiles();
    const int b = blockIdx.x / nt, tile = blockIdx.x % nt;
    const int ox = (tile % g.tx) * TX, oy = ((tile / g.tx) % g.ty) * TY, oz = (tile / (g.tx * g.ty)) * TZ;
    stage_halo<IN_BF16_16>(in, 1.0f, sm, g, b, ox, oy, oz);
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int c = lane / 16, n = lane % 16;
    const bool hi = c >= 2;
    const int half = (c & 1) * 8;
    bf16x8 A[NSTEP3B];
#pragma unroll
    for (int s = 0; s < NSTEP3B; ++s) A[s] = wpack[s * 64 + lane];
    __syncthreads();
    constexpr int GROUPS = TX * NB * HZ / 16;
#pragma unroll 1
    for (int gi = w; gi < GROUPS; gi += 4) {
        const int id = 16 * gi + n;
        const int x = id % TX, bk = (id / TX) % NB, hz = id / (TX * NB);
        const int base = ((hz * HY + L3ROWS * bk + (hi ? L3ROWS : 0)) * HX + x) * 16 + half;
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < NSTEP3B; ++s)
            d = mfma16(A[s], __builtin_bit_cast(bf16x8, *(const u32x4 *)(sm + base + ((s / 3) * HX + s % 3) * 16)), d);
        if (c < 3) {
#pragma unroll
            for (int j = 0; j < L3ROWS; ++j) P[hz][bk][j][c][x] = d[j];
        }
    }
    __syncthreads();
    const int64_t N = g.nvox();
    const float scale = out_scale ? out_scale[b] : 1.0f;
    {
        const int v = threadIdx.x;  // one output voxel of the 8 x 8 x 4 tile per thread
        const int x = v % TX, y = (v / TX) % TY, z = v / (TX * TY);
        const int bk = y / L3ROWS, j = y % L3ROWS;
        const int gx = ox + x, gy = oy + y, gz = oz + z;
        if (gx < g.nx && gy < g.ny && gz < g.nz) {
            const float r = (P[z][bk][j][0][x] + P[z + 1][bk][j][1][x]) + P[z + 2][bk][j][2][x];
            out[(int64_t)b * N + ((int64_t)gz * g.ny + gy) * g.nx + gx] = r * scale;
        }
    }
}

// Weight gradient dW[co][ci][t] = sum_v dout[v][co] * in[v + off(t)][ci] for CO_IN output
// channels of dout (16, or 1 for dL/dz) and CI_IN input channels (16, or 1 for u): an MFMA
// GEMM with M = co, N = ci, K = voxels.  Each workgroup walks a contiguous range of
// (tile, sample) items and writes its FP32 partial [16][16][27] (padded) to partials[wg];
// k_wgrad_reduce sums the partials in a fixed order.
template <int CO_IN, int CI_IN>
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t *__restrict__ dout, const uint16_t *__restrict__ in,
                                               float *__restrict__ partials, Grid3 g, int batch)
{
    __shared__ __attribute__((aligned(16))) uint16_t doutT[16 * TV];           // [co][voxel]
    __shared__ __attribute__((aligned(16))) uint16_t inT[3 * 16 * HZ * HY * 8];  // [shift][ci][hz][hy][x]
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int c = lane / 16, n = lane % 16;
    const int items = g.ntiles() * batch;
    const int i0 = (int)(((int64_t)blockIdx.x * items) / gridDim.x);
    const int i1 = (int)(((int64_t)(blockIdx.x + 1) * items) / gridDim.x);
    const int64_t N = g.nvox();
    f32x4 acc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int it = i0; it < i1; ++it) {
        const int b = it / g.ntiles(), tile = it % g.ntiles();
        const int ox = (tile % g.tx) * TX, oy = ((tile / g.tx) % g.ty) * TY, oz = (tile / (g.tx * g.ty)) * TZ;
        __syncthreads();
        // dout tile, transposed to [co][voxel]; voxels outside the grid contribute zero
        for (int i = threadIdx.x; i < TV * 16; i += blockDim.x) {
            const int v = i / 16, co = i % 16;
            const int x = v % TX, y = (v / TX) % TY, z = v / (TX * TY);
            const int gx = ox + x, gy = oy + y, gz = oz + z;
            uint16_t d = 0;
            if (co < CO_IN && gx < g.nx && gy < g.ny && gz < g.nz) {
                const int64_t gv = (int64_t)b * N + ((int64_t)gz * g.ny + gy) * g.nx + gx;
                d = dout[gv * CO_IN + co];
            }
            doutT[co * TV + v] = d;
        }
        // input halo, transposed to [ci][hz][hy][hx] and stored three times, shifted by
        // 0/1/2 in x, so that every 8-voxel x-run of a tap is one aligned 16-byte load
        for (int i = threadIdx.x; i < 16 * HZ * HY * 8; i += blockDim.x) {
            const int x = i % 8, hy = (i / 8) % HY, hz = (i / (8 * HY)) % HZ, ci = i / (8 * HY * HZ);
            const int gy = oy - 1 + hy, gz = oz - 1 + hz;
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int gx = ox - 1 + x + s;
                uint16_t v = 0;
                if (ci < CI_IN && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz) {
                    const int64_t gv = (int64_t)b * N + ((int64_t)gz * g.ny + gy) * g.nx + gx;
                    v = in[gv * CI_IN + ci];
                }
                inT[((s * 16 + ci) * HZ + hz) * (HY * 8) + hy * 8 + x] = v;
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int ks = 0; ks < TV / 32; ++ks) {
            // K = 32 voxels = 4 x-rows of 8; this lane's 8 voxels: row r = 4*ks + c
            const int row = 4 * ks + c;
            const int y = row % TY, z = row / TY;
            const bf16x8 a = __builtin_bit_cast(bf16x8, *(const u32x4 *)(doutT + n * TV + row * 8));
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int t = w + 4 * k;
                if (t < 27) {
                    const int s = tap_dx(t) + 1, hz = z + tap_dz(t) + 1, hy = y + tap_dy(t) + 1;
                    const bf16x8 bb = __builtin_bit_cast(
                        bf16x8, *(const u32x4 *)(inT + ((s * 16 + n) * HZ + hz) * (HY * 8) + hy * 8));
                    acc[k] = mfma16(a, bb, acc[k]);
                }
            }
        }
    }
    // partial[wg][(co*16 + ci)*27 + t]: lane holds co = 4c + i, ci = n
    float *P = partials + (int64_t)blockIdx.x * (16 * 16 * 27);
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const int t = w + 4 * k;
        if (t < 27)
#pragma unroll
            for (int i = 0; i < 4; ++i) P[((4 * c + i) * 16 + n) * 27 + t] = acc[k][i];
    }
}

// ---- fused inference: the three layers in one kernel -------------------------------------------
// A workgroup of 8 waves (2 per SIMD, up to 256 VGPRs each) owns a 16 x 16 (x, y) column of
// output voxels and a range of z-planes and marches along z, one plane per step.  In step i every
// wave does its share of
//   layer 1: h1 plane i-1 from the input window (planes i-2 .. i; one MFMA per 16 voxels,
//            K = 27 taps; 25 groups);
//   layer 2: reads h1 plane i-2 ONCE -- five operands per 16 voxels, one per in-plane tap
//            pair -- and adds it to the three h2 planes it touches (15 MFMAs), h2 plane i-3
//            being complete afterwards (21 groups);
//   layer 3: one banded block per wave (2 output rows, see NSTEP3B): six MFMAs on h2 plane i-4,
//            whose three plane-offset partials move one 16-lane group per step (a shuffle), so
//            output plane i-5 is complete;
//   input:   stores plane i+1 (rounded to BF16, loaded three steps earlier) into a four-plane
//            ring and requests plane i+4.
// The layers of one step touch different planes -- h1 and h2 are double-buffered in LDS -- so a
// step needs a single barrier, and a step lasts as long as its slowest wave.  The groups are
// dealt by measured cost (a layer-1 group costs well under a layer-2 group, a layer-3 block
// about as much as a layer-2 group) rather than evenly:
//   waves 0-4 ("heavy"): layer-2 groups w, w+8, w+16, layer-1 groups w, w+8, layer-3 block w;
//   waves 5-7 ("light"): layer-2 groups w, w+8, layer-1 groups w, w+8 and three of 16..24,
//                        layer-3 block w.
// Dealing 3 layer-2 groups per wave (three spare slots) and the 25th layer-1 group to wave 0
// took 4-5 % longer.  Each role runs its own copy of the march (a wave-uniform branch before
// it), so that a step of the march's steady state (every layer active: all steps but the first
// seven and the last two) is one basic block and the compiler can overlap a wave's three layers:
// it has no plane-range tests, and lanes with nothing to store write to a scratch slot instead of
// branching.  The h2 partial sums of unfinished planes stay in registers, rotating through three
// slots (the march is unrolled three times so that the slots are compile-time registers).  h1 is computed on a
// 20 x 20 and h2 on an 18 x 18 footprint (the halos the next layer needs).  Every LDS address
// and grid-boundary test that does not depend on the plane is formed once, before the march.
// Each accumulator receives its K-steps in the order of the layer-wise k_conv, so the output is
// bitwise the three-kernel forward pass's.
constexpr int FX = 16, FY = 16;             // output footprint per workgroup
constexpr int F2X = FX + 2, F2Y = FY + 2;   // h2 footprint (324 voxels, 21 groups of 16)
constexpr int F1X = FX + 4, F1Y = FY + 4;   // h1 footprint (400 voxels, 25 groups)
constexpr int F0X = FX + 6, F0Y = FY + 6;   // input footprint (484 voxels)
constexpr int F2V = F2X * F2Y, F1V = F1X * F1Y, F0V = F0X * F0Y;
constexpr int F2G = (F2V + 15) / 16, F1G = F1V / 16, F3G = FX * FY / 16;
constexpr int FWAVES = 8, FTHREADS = 64 * FWAVES;  // waves per fused workgroup (2 per SIMD)
constexpr int G2W = (F2G + FWAVES - 1) / FWAVES;    // layer-2 groups of a heavy wave (3)
constexpr int FHEAVY = F2G - (G2W - 1) * FWAVES;    // heavy waves (5): the rest have G2W - 1 groups
constexpr int FLIGHT = FWAVES - FHEAVY;             // light waves (3)
constexpr int G1H = 2, G1X = 3;                     // layer-1 groups: heavy, extra per light wave
constexpr int G1W = G1H + G1X;                      // layer-1 slots of a light wave (5)
static_assert(F2G == 21 && F1G == G1H * FWAVES + G1X * FLIGHT, "group deal of the 16 x 16 footprint");
constexpr int F3B = FY / L3ROWS;                   // layer 3: row blocks (one per wave)
static_assert(F3B == FWAVES, "one layer-3 row block per wave");
// layer-1 group of slot k of wave w (k < G1H for heavy waves): w + 8k, then the light wave's
// share of groups 16..24 dealt round robin
__host__ __device__ constexpr int fused_g1(int w, int k)
{
    return k < G1H ? w + FWAVES * k : G1H * FWAVES + (w - FHEAVY) + FLIGHT * (k - G1H);
}
// the deal covers every layer-1 group (slots 0..G1H-1 of every wave, G1H..G1W-1 of the light
// waves) and every layer-2 group (w + 8k, k < G2W for heavy waves, k < G2W - 1 for light) once
constexpr bool fused_deal_exact()
{
    int n1[F1G] = {}, n2[F2G] = {};
    for (int w = 0; w < FWAVES; ++w) {
        for (int k = 0; k < (w < FHEAVY ? G1H : G1W); ++k) {
            const int g = fused_g1(w, k);
            if (g < 0 || g >= F1G) return false;
            ++n1[g];
        }
        for (int k = 0; k < (w < FHEAVY ? G2W : G2W - 1); ++k) {
            const int g = w + FWAVES * k;
            if (g >= F2G) return false;
            ++n2[g];
        }
    }
    for (int g = 0; g < F1G; ++g)
        if (n1[g] != 1) return false;
    for (int g = 0; g < F2G; ++g)
        if (n2[g] != 1) return false;
    return true;
}
static_assert(fused_deal_exact(), "every layer-1 and layer-2 group is dealt to exactly one wave slot");
constexpr int TPL = F0Y * F1X;                     // x-triplets per input-window plane (440)
static_assert(64 + TPL <= FTHREADS, "the input loader (one triplet per thread) fits waves 1..7");
constexpr int VS = 16;  // h1 / h2 voxel record in LDS: 16 BF16 channels (32 bytes)
constexpr int H1S = F1V * VS + 8, H2S = F2V * VS + 8;  // plane buffers + 8 zero channels

template <int K>
using Slot = std::integral_constant<int, K>;

__global__ __launch_bounds__(FTHREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_fused_fwd(
    const float *__restrict__ in, const float *__restrict__ in_scale, const bf16x8 *__restrict__ pk1,
    const bf16x8 *__restrict__ pk2, const bf16x8 *__restrict__ pk3f, float *__restrict__ out,
    const float *__restrict__ out_scale, int nx, int ny, int nz, int tiles_x, int tiles_y, int zchunk,
    const int *__restrict__ skip)
{
    if (skip && *skip) return;
    // input ring: plane p in slot p & 3, as x-triplets: T[slot][row * F1X + x] holds the input
    // at window row `row`, columns x, x+1, x+2 (one 8-byte read per tap line of layer 1)
    __shared__ __attribute__((aligned(16))) u32x2 T[4][TPL];
    // h1 / h2 plane p in buffer p & 1, each followed by 8 zero channels (the operand of the
    // empty tenth tap slot)
    __shared__ __attribute__((aligned(16))) uint16_t H1[2][H1S];
    __shared__ __attribute__((aligned(16))) uint16_t H2[2][H2S];
    // lanes with nothing to store write here, so that the stores need no branch
    __shared__ __attribute__((aligned(16))) u32x2 Trash[FTHREADS];
    const int txy = blockIdx.x % (tiles_x * tiles_y), zc = blockIdx.x / (tiles_x * tiles_y);
    const int ox = (txy % tiles_x) * FX, oy = (txy / tiles_x) * FY;
    const int z0 = zc * zchunk, z1 = min(z0 + zchunk, nz);
    const float scale = in_scale ? in_scale[0] : 1.0f;
    const float oscale = out_scale ? out_scale[0] : 1.0f;
    const int lane = threadIdx.x % 64, w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int c = lane / 16, n = lane % 16;
    const bool hi = c >= 2;        // operand rows of the second tap of each pair
    const int half = (c & 1) * 8;  // channel half
    const int hiv = hi ? VS : 0;
    const int64_t plane = (int64_t)nx * ny;
    // layer-1 and layer-2 weights stay in registers; layer 3's (read five times per step) in LDS
    const bf16x8 A1 = pk1[lane];
    bf16x8 A2[NSTEP16];
#pragma unroll
    for (int s = 0; s < NSTEP16; ++s) A2[s] = pk2[s * 64 + lane];
    __shared__ bf16x8 A3s[NSTEP3B * 64];
    for (int t = threadIdx.x; t < NSTEP3B * 64; t += FTHREADS) A3s[t] = pk3f[t];
    if (threadIdx.x < 8) {
        H1[0][F1V * VS + threadIdx.x] = 0;
        H1[1][F1V * VS + threadIdx.x] = 0;
        H2[0][F2V * VS + threadIdx.x] = 0;
        H2[1][F2V * VS + threadIdx.x] = 0;
    }

    // ---- plane-independent per-lane facts ----
    // layer 1: slot k holds group g = fused_g1(w, k), voxel v = 16 g + n of the 20 x 20 footprint
    // (T row = v); tap lines 2c, 2c+1 and 8 of this lane's operand row, window planes lX / 3
    const bool heavy = w < FHEAVY;
    const int v1 = 16 * w + n;
    const int lA = 2 * c, lB = 2 * c + 1;
    const int rA = (lA % 3) * F1X + v1, rB = (lB % 3) * F1X + v1, rC = 2 * F1X + v1;
    const int pA = lA / 3, pB = lB / 3;
    unsigned ok1 = 0;  // bit k: slot k holds a group whose voxel lies inside the grid (x, y)
#pragma unroll
    for (int k = 0; k < G1W; ++k) {
        const int g = k < G1H || !heavy ? fused_g1(w, k) : F1G, v = 16 * g + n;
        const int gx = ox - 2 + v % F1X, gy = oy - 2 + v / F1X;
        if (g < F1G && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ok1 |= 1u << k;
    }
    // layer 2: group k is g = w + 8k (k < 2 for light waves); operand base (hi offset included)
    // and validity
    int b2[G2W];
    unsigned ok2 = 0;
#pragma unroll
    for (int k = 0; k < G2W; ++k) {
        const int g = w + FWAVES * k, v = 16 * g + n, vc = min(v, F2V - 1), hx = vc % F2X, hy = vc / F2X;
        b2[k] = (hy * F1X + hx) * VS + half + hiv;
        const int gx = ox - 1 + hx, gy = oy - 1 + hy;
        if (g < F2G && v < F2V && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ok2 |= 1u << k;
    }
    // layer 3: wave w owns the row block of output rows 2w, 2w+1 (h2 window rows 2w .. 2w+3,
    // operand rows 16..31 two rows lower)
    const int bk = w;
    const int b3 = ((L3ROWS * bk + (hi ? L3ROWS : 0)) * F2X + n) * VS + half;
    const int64_t ob = (int64_t)(oy + L3ROWS * bk) * nx + ox + n;
    unsigned ok3 = 0;  // bit j: output row 2 bk + j of this lane's voxel lies inside the grid
#pragma unroll
    for (int j = 0; j < L3ROWS; ++j)
        if (ox + n < nx && oy + L3ROWS * bk + j < ny) ok3 |= 1u << j;

    // input loader: thread 64 + t, t < TPL, owns triplet t (window row t / F1X, columns t % F1X
    // .. +2) -- waves 1-7, so that wave 0, a heavy wave sharing its SIMD with heavy wave 4, loads
    // nothing (measured 0.381-0.382 against 0.385-0.386 ms with the loader on waves 0-6).  The
    // three loads read clamped, always valid addresses without branches, so they stay in flight
    // until the plane is stored three steps later; the zero padding outside the grid and the BF16
    // rounding are applied then.
    const int tl = (int)threadIdx.x - 64;
    const bool trip = tl >= 0 && tl < TPL;
    const int tlc = trip ? tl : 0;
    const int tgy = oy - 3 + tlc / F1X, tgx0 = ox - 3 + tlc % F1X;
    const bool trow = trip && tgy >= 0 && tgy < ny;
    const bool tx0 = tgx0 >= 0 && tgx0 < nx, tx1 = tgx0 + 1 >= 0 && tgx0 + 1 < nx, tx2 = tgx0 + 2 >= 0 && tgx0 + 2 < nx;
    const float *trowp = in + (int64_t)min(max(tgy, 0), ny - 1) * nx;
    const int cx0 = min(max(tgx0, 0), nx - 1), cx1 = min(max(tgx0 + 1, 0), nx - 1), cx2 = min(max(tgx0 + 2, 0), nx - 1);
    struct F3 {
        float v[3];
    };
    auto fetch = [&](int ip) -> F3 {
        const float *row = trowp + (int64_t)min(max(ip, 0), nz - 1) * plane;
        return F3{{row[cx0], row[cx1], row[cx2]}};
    };
    auto to_triplet = [&](int ip, F3 g) -> u32x2 {
        const float *f = g.v;
        const bool ok = trow && ip >= 0 && ip < nz;
        u32x2 t3;
        t3[0] = pack_bf16(ok && tx0 ? f[0] * scale : 0.f, ok && tx1 ? f[1] * scale : 0.f);
        t3[1] = pack_bf16(ok && tx2 ? f[2] * scale : 0.f, 0.f);
        return t3;
    };
    // five B operands of a voxel group whose operand base (hi offset included) is `base`, in a
    // one-plane buffer H of row length RX whose zero channels start at ZOFF: off(t9) =
    // (t9 / 3) RX + t9 % 3 voxels; hi lanes add 1 voxel (pairs 0, 2, 3) or RX - 2 (pair 1), or
    // read the zero channels (pair 4)
    auto load_pairs = [&](const uint16_t *H, auto RXc, int ZOFF, int base, u32x4 *raw) {
        constexpr int RX = decltype(RXc)::value;
        const uint16_t *pb = H + base;
        raw[0] = *(const u32x4 *)(pb);
        raw[1] = *(const u32x4 *)(pb + (hi ? (RX - 3) * VS : 0) + 2 * VS);
        raw[2] = *(const u32x4 *)(pb + (RX + 1) * VS);
        raw[3] = *(const u32x4 *)(pb + (2 * RX) * VS);
        raw[4] = *(const u32x4 *)(hi ? H + ZOFF : pb + (2 * RX + 2) * VS);
    };

    f32x4 acc2[3][G2W];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int k = 0; k < G2W; ++k) acc2[r][k] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    // layer-3 running sums: lanes 16 dzi + n hold, for the 2 rows of the block, the sum over
    // plane offsets -1 .. dzi - 1 of the output plane those offsets have reached
    f32x4 run3 = {0.f, 0.f, 0.f, 0.f};
    // input planes are requested three steps before they are stored (step i stores plane i+1,
    // loaded during step i-3), so the global latency hides behind three steps of work
    F3 pre0 = fetch(z0 - 3), pre1 = fetch(z0 - 2), pre2 = fetch(z0 - 1);

    // step i; rotating slot R holds the partial sums the step starts (h2 plane i-1, output plane
    // i-3), R+2 those it continues, R+1 those it completes.  HEAVY: the wave's role (see above).
    auto step = [&](auto R, int i, auto steadyc, auto heavyc) __attribute__((always_inline)) {
        constexpr bool HEAVY = decltype(heavyc)::value;
        constexpr int NXT = decltype(R)::value, CUR = (NXT + 2) % 3, PRV = (NXT + 1) % 3;
        constexpr bool STEADY = decltype(steadyc)::value;  // every layer active: no range tests
        *(trip ? &T[(i + 1) & 3][tlc] : &Trash[threadIdx.x]) = to_triplet(i + 1, pre0);
        pre0 = pre1;
        pre1 = pre2;
        pre2 = fetch(i + 4);
        // layer 2 first (most MFMA work; layers 1 and 3 fill its gaps):
        // h1 plane q2 = i-2 feeds h2 planes q2+1 (NXT), q2 (CUR) and q2-1 (PRV, done)
        const int q2 = i - 2;
        if (STEADY || (q2 >= z0 - 2 && q2 <= z1 + 1)) {
            const int pd = q2 - 1;
            const bool keep = pd >= z0 - 1, zin = pd >= 0 && pd < nz;
            const uint16_t *Hq = H1[q2 & 1];
            uint16_t *Hw = H2[pd & 1] + v1 * VS + 4 * c;
#pragma unroll
            for (int k = 0; k < G2W; ++k) {
                if (HEAVY || k < G2W - 1) {
                    u32x4 raw[NSTEP3F];
                    load_pairs(Hq, std::integral_constant<int, F1X>{}, F1V * VS, b2[k], raw);
                    f32x4 an = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int p = 0; p < NSTEP3F; ++p) {
                        const bf16x8 bop = __builtin_bit_cast(bf16x8, raw[p]);
                        an = mfma16(A2[p], bop, an);
                        acc2[CUR][k] = mfma16(A2[5 + p], bop, acc2[CUR][k]);
                        acc2[PRV][k] = mfma16(A2[10 + p], bop, acc2[PRV][k]);
                    }
                    u32x2 pk = relu_bf16x4(acc2[PRV][k]);
                    acc2[NXT][k] = an;
                    const uint32_t m2 = (zin && ((ok2 >> k) & 1u)) ? 0xffffffffu : 0u;
                    pk[0] &= m2;
                    pk[1] &= m2;
                    const bool st2 = keep && 16 * (w + FWAVES * k) + n < F2V;
                    *(st2 ? (u32x2 *)(Hw + 16 * FWAVES * k * VS) : &Trash[threadIdx.x]) = pk;
                }
            }
        }
        // layer 1: h1 plane q1 = i-1 from input planes q1-1, q1, q1+1
        const int q1 = i - 1;
        if (STEADY || (q1 >= z0 - 2 && q1 <= z1 + 1)) {
            const bool zin = q1 >= 0 && q1 < nz;
            const u32x2 *TA = T[(q1 - 1 + pA) & 3] + rA, *TB = T[(q1 - 1 + pB) & 3] + rB, *TC = T[(q1 + 1) & 3] + rC;
            uint16_t *Hw = H1[q1 & 1] + v1 * VS + 4 * c;
#pragma unroll
            for (int k = 0; k < G1W; ++k) {
                if (!HEAVY || k < G1H) {
                    const int o16 = 16 * (fused_g1(w, k) - w);
                    const u32x2 ta = TA[o16], tb = TB[o16], tc = TC[o16];
                    u32x4 raw;
                    raw[0] = ta[0];
                    raw[1] = (ta[1] & 0xffffu) | (tb[0] << 16);
                    raw[2] = (tb[0] >> 16) | (tb[1] << 16);
                    raw[3] = c == 0 ? tc[0] : (c == 1 ? (tc[1] & 0xffffu) : 0u);
                    const f32x4 a = mfma16(A1, __builtin_bit_cast(bf16x8, raw), f32x4{0.f, 0.f, 0.f, 0.f});
                    u32x2 pk = relu_bf16x4(a);
                    const uint32_t m1 = (zin && ((ok1 >> k) & 1u)) ? 0xffffffffu : 0u;
                    pk[0] &= m1;
                    pk[1] &= m1;
                    *(u32x2 *)(Hw + o16 * VS) = pk;
                }
            }
        }
        // layer 3: h2 plane q3 = i-4 feeds output planes q3+1 (dz = -1, lanes 0-15), q3 (lanes
        // 16-31) and q3-1 (lanes 32-47, done).  Each lane group takes the running sums of the
        // group below it, so an output is ((P(-1) + P(0)) + P(+1)); lanes 48-63 hold nothing.
        const int q3 = i - 4;
        if (STEADY || (q3 >= z0 - 1 && q3 <= z1)) {
            const int o = q3 - 1;
            const uint16_t *Hp = H2[q3 & 1] + b3;
            u32x4 raw[NSTEP3B];
#pragma unroll
            for (int s3 = 0; s3 < NSTEP3B; ++s3) raw[s3] = *(const u32x4 *)(Hp + ((s3 / 3) * F2X + s3 % 3) * VS);
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s3 = 0; s3 < NSTEP3B; ++s3) d = mfma16(A3s[s3 * 64 + lane], __builtin_bit_cast(bf16x8, raw[s3]), d);
#pragma unroll
            for (int j = 0; j < L3ROWS; ++j) {
                const float below = __shfl_up(run3[j], 16, 64);
                run3[j] = c == 0 ? d[j] : below + d[j];
            }
            if (c == 2 && o >= z0 && o < z1) {
#pragma unroll
                for (int j = 0; j < L3ROWS; ++j)
                    if ((ok3 >> j) & 1u) out[o * plane + ob + (int64_t)j * nx] = run3[j] * oscale;
            }
        }
        __syncthreads();
    };
    // steps z0-4 .. z1+4; the rotating slot of step i is (i - z0 + 4) % 3.  Steps z0+3 .. z1+2
    // run every layer (the steady state, no range tests); the first 7 and last 2 test ranges.
    // Every wave takes the same steps and barriers; heavy and light waves run their own copy.
    using Tr = std::true_type;
    using Fa = std::false_type;
    auto march = [&](auto heavyc) __attribute__((always_inline)) {
        auto any_step = [&](int i, auto steadyc) __attribute__((always_inline)) {
            switch ((i - z0 + 4) % 3) {
            case 0: step(Slot<0>{}, i, steadyc, heavyc); break;
            case 1: step(Slot<1>{}, i, steadyc, heavyc); break;
            default: step(Slot<2>{}, i, steadyc, heavyc); break;
            }
        };
        int i = z0 - 4;
        for (; i <= z0 + 2; ++i) any_step(i, Fa{});
        for (; i + 2 <= z1 + 2; i += 3) {  // slots 1, 2, 0 (z0+3 is slot 1)
            step(Slot<1>{}, i, Tr{}, heavyc);
            step(Slot<2>{}, i + 1, Tr{}, heavyc);
            step(Slot<0>{}, i + 2, Tr{}, heavyc);
        }
        for (; i <= z1 + 2; ++i) any_step(i, Tr{});
        for (; i <= z1 + 4; ++i) any_step(i, Fa{});
    };
    if (heavy) march(Tr{});
    else march(Fa{});
}

}  // namespace cnn
}  // namespace hpccg
