// Learned preconditioner: BF16 implicit-GEMM 3x3x3 convolution kernels on CDNA4 matrix cores.
//
// Network (bias-free, so positively homogeneous: net(c*u) = c*net(u) for c > 0):
//   h1 = relu(conv(u; W1[16][27])), h2 = relu(conv(h1; W2[16][16][27])), z = conv(h2; W3[16][27])
// conv(in; W)[co](v) = sum_{ci,t} W[co][ci][t] * in[ci](v + off(t)), off(t) = (t%3-1, t/3%3-1,
// t/9-1) in (x, y, z), zero outside the grid.  Activations are channel-last BF16
// ([batch][voxel][16], 32 bytes per voxel), accumulation is FP32 (MFMA).
//
// Implicit GEMM on v_mfma_f32_16x16x32_bf16 (D[16x16] += A[16x32] * B[32x16]):
//   M = output channels (16), N = 16 voxels, K = 2 taps x 16 input channels per MFMA
//   (14 MFMAs for the 27 taps, the 28th tap zero), or K = 27(+5) taps for a 1-channel input.
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
constexpr int NSTEP16 = 14;       // MFMA K-steps for 16 input channels x 27 taps
constexpr int NW1 = 16 * 27, NW2 = 16 * 16 * 27, NW3 = 16 * 27;
constexpr int NWALL = NW1 + NW2 + NW3;  // 7776 parameters

// float -> bfloat16, round to nearest even (inf/nan pass through)
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
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

__device__ inline f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
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
// Epilogue: ReLU -> BF16 [v][16]; FP32 channel 0 times a per-sample scale; masked BF16 (the
// backward pass through a ReLU whose forward output is `mask`).
enum OutKind { OUT_RELU_BF16 = 0, OUT_F32_CH0 = 1, OUT_MASKED_BF16 = 2 };

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

// One 3x3x3 convolution layer with 16 output channels (layer 3 packs its single output channel
// as channel 0 of 16).  Grid: one workgroup per (tile, sample).
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
                const int t = 2 * s + c / 2;
                u32x4 raw = {0u, 0u, 0u, 0u};
                if (t < 27) {
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
            if (OK == OUT_F32_CH0) {
                if (c == 0) ((float *)out)[gv] = acc[0] * (out_scale ? out_scale[b] : 1.0f);
            } else {
                uint16_t o[4];
                if (OK == OUT_RELU_BF16) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[i] = f2bf(acc[i] > 0.f ? acc[i] : 0.f);
                } else {
                    const u32x2 mm = *(const u32x2 *)(mask + gv * 16 + 4 * c);
                    const uint16_t m4[4] = {(uint16_t)(mm[0] & 0xffff), (uint16_t)(mm[0] >> 16),
                                            (uint16_t)(mm[1] & 0xffff), (uint16_t)(mm[1] >> 16)};
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[i] = bf2f(m4[i]) > 0.f ? f2bf(acc[i]) : (uint16_t)0;
                }
                u32x2 pk;
                pk[0] = o[0] | ((uint32_t)o[1] << 16);
                pk[1] = o[2] | ((uint32_t)o[3] << 16);
                *(u32x2 *)((uint16_t *)out + gv * 16 + 4 * c) = pk;
            }
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
// A workgroup owns a 16 x 16 (x, y) column of output voxels and a range of z-planes, and marches
// along z keeping three planes of each layer's input in LDS rings: step i loads input plane i
// (rounded to BF16), computes h1 plane i-1, h2 plane i-2 and the output plane i-3.  h1 is
// computed on a 20 x 20 and h2 on an 18 x 18 footprint (the halos the next layer needs), so the
// activations never leave LDS.  The MFMA operands, their K order and the epilogues are those of
// k_conv, so the output is bitwise the three-kernel forward pass's.
constexpr int FX = 16, FY = 16;             // output footprint per workgroup
constexpr int F2X = FX + 2, F2Y = FY + 2;   // h2 footprint (324 voxels, 21 groups of 16)
constexpr int F1X = FX + 4, F1Y = FY + 4;   // h1 footprint (400 voxels, 25 groups)
constexpr int F0X = FX + 6, F0Y = FY + 6;   // input footprint (484 voxels)
constexpr int F2V = F2X * F2Y, F1V = F1X * F1Y, F0V = F0X * F0Y;
constexpr int F2G = (F2V + 15) / 16, F1G = F1V / 16, F3G = FX * FY / 16;

__device__ inline int ring3(int p) { return ((p % 3) + 3) % 3; }

// Layer 1 reads its taps as three x-triplets per lane (the input window keeps its planes in
// order, so the offsets are fixed per lane); layers 2 and 3 resolve both candidate taps of each K-step at compile time
// (a lane's row of the B operand picks one), load all 14 operands of a voxel group before its
// MFMA chain, and keep h1/h2 in three-plane rings.  Two workgroups fit a CU (78 KB of LDS each).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_fused_fwd(const float *__restrict__ in, const float *__restrict__ in_scale,
                                               const bf16x8 *__restrict__ pk1, const bf16x8 *__restrict__ pk2,
                                               const bf16x8 *__restrict__ pk3, float *__restrict__ out,
                                               const float *__restrict__ out_scale, int nx, int ny, int nz,
                                               int tiles_x, int tiles_y, int zchunk, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    // input window, planes i-2, i-1, i, stored as x-triplets: T[plane][row][x] holds the input at
    // window columns x, x+1, x+2 (one 8-byte read per tap line of layer 1)
    __shared__ __attribute__((aligned(16))) u32x2 T[3][F0Y][F1X];
    __shared__ __attribute__((aligned(16))) uint16_t H1[3][F1V * 16];
    __shared__ __attribute__((aligned(16))) uint16_t H2[3][F2V * 16];
    const int txy = blockIdx.x % (tiles_x * tiles_y), zc = blockIdx.x / (tiles_x * tiles_y);
    const int ox = (txy % tiles_x) * FX, oy = (txy / tiles_x) * FY;
    const int z0 = zc * zchunk, z1 = min(z0 + zchunk, nz);
    const float scale = in_scale ? in_scale[0] : 1.0f;
    const float oscale = out_scale ? out_scale[0] : 1.0f;
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int c = lane / 16, n = lane % 16;
    const bool hi = c >= 2;       // second tap of each K-step pair
    const int half = (c & 1) * 8; // channel half
    bf16x8 A1 = pk1[lane];
    bf16x8 A2[NSTEP16], A3[NSTEP16];
#pragma unroll
    for (int s = 0; s < NSTEP16; ++s) { A2[s] = pk2[s * 64 + lane]; A3[s] = pk3[s * 64 + lane]; }
    const int64_t plane = (int64_t)nx * ny;
    // layer-1 tap lines of this lane's B-operand row (k1_tap order): lines 2c, 2c+1 and line 8
    const int lA = 2 * c, lB = 2 * c + 1;
    const int offA = ((lA / 3) * F0Y + lA % 3) * F1X, offB = ((lB / 3) * F0Y + lB % 3) * F1X;
    const int offC = (2 * F0Y + 2) * F1X;
    const u32x2 *T0 = &T[0][0][0];
    // triplet e of a plane: window row e / F1X, columns e % F1X .. +2 (440 per plane, two per thread)
    auto load_triplet = [&](int ip, int e) -> u32x2 {
        u32x2 t3;
        t3[0] = t3[1] = 0u;
        if (e >= F0Y * F1X || ip < 0 || ip >= nz) return t3;
        const int gy = oy - 3 + e / F1X, gx0 = ox - 3 + e % F1X;
        if (gy < 0 || gy >= ny) return t3;
        const float *row = in + ip * plane + (int64_t)gy * nx;
        uint16_t u[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int gx = gx0 + k;
            u[k] = (gx >= 0 && gx < nx) ? f2bf(row[gx] * scale) : (uint16_t)0;
        }
        t3[0] = u[0] | ((uint32_t)u[1] << 16);
        t3[1] = u[2];
        return t3;
    };
    u32x2 pre[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) pre[e] = load_triplet(z0 - 3, threadIdx.x + 256 * e);
    int r0 = ring3(z0 - 3);  // ring slot of plane z0-3; slots advance by one per step
    for (int i = z0 - 3; i <= z1 + 2; ++i) {
        // three ring slots: plane i -> si, i-1 -> sm1, i-2 -> sm2 (and i-3 -> si, i-4 -> sm1)
        const int si = r0, sm1 = r0 == 0 ? 2 : r0 - 1, sm2 = sm1 == 0 ? 2 : sm1 - 1;
        r0 = r0 == 2 ? 0 : r0 + 1;
        {
            // shift the window (each thread owns the same triplets of all three planes) and
            // store plane i, prefetched during the previous step; then prefetch plane i + 1
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int v = threadIdx.x + 256 * e;
                if (v < F0Y * F1X) {
                    u32x2 *Tv = &T[0][0][0] + v;
                    const u32x2 a1 = Tv[F0Y * F1X], a2 = Tv[2 * F0Y * F1X];
                    Tv[0] = a1;
                    Tv[F0Y * F1X] = a2;
                    Tv[2 * F0Y * F1X] = pre[e];
                }
            }
#pragma unroll
            for (int e = 0; e < 2; ++e) pre[e] = load_triplet(i + 1, threadIdx.x + 256 * e);
        }
        __syncthreads();
        const int q = i - 1;  // h1 plane q uses input planes q-1, q, q+1 = window planes 0, 1, 2
        if (q >= z0 - 2) {
            uint16_t *Hq = H1[sm1];
            const bool zin = q >= 0 && q < nz;
            for (int g = w; g < F1G; g += 4) {
                const int v = 16 * g + n, hx = v % F1X, hy = v / F1X;
                const int gb = hy * F1X + hx;
                const u32x2 ta = T0[offA + gb], tb = T0[offB + gb], tc = T0[offC + gb];
                u32x4 raw;
                raw[0] = ta[0];
                raw[1] = (ta[1] & 0xffffu) | (tb[0] << 16);
                raw[2] = (tb[0] >> 16) | (tb[1] << 16);
                raw[3] = c == 0 ? tc[0] : (c == 1 ? (tc[1] & 0xffffu) : 0u);
                const f32x4 acc = mfma16(A1, __builtin_bit_cast(bf16x8, raw), f32x4{0.f, 0.f, 0.f, 0.f});
                const int gx = ox - 2 + hx, gy = oy - 2 + hy;
                const bool in_grid = zin && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
                u32x2 pk; pk[0] = pk[1] = 0u;
                if (in_grid) {
                    pk[0] = f2bf(acc[0] > 0.f ? acc[0] : 0.f) | ((uint32_t)f2bf(acc[1] > 0.f ? acc[1] : 0.f) << 16);
                    pk[1] = f2bf(acc[2] > 0.f ? acc[2] : 0.f) | ((uint32_t)f2bf(acc[3] > 0.f ? acc[3] : 0.f) << 16);
                }
                *(u32x2 *)(Hq + v * 16 + 4 * c) = pk;
            }
        }
        __syncthreads();
        const int p = i - 2;  // h2 plane p uses H1 planes p-1 (slot of i-3 == si), p (sm2), p+1 (sm1)
        if (p >= z0 - 1) {
            const uint16_t *Hp3[3] = {H1[si], H1[sm2], H1[sm1]};
            uint16_t *Hp = H2[sm2];
            const bool zin = p >= 0 && p < nz;
            for (int g = w; g < F2G; g += 4) {
                const int v = 16 * g + n, vc = min(v, F2V - 1), hx = vc % F2X, hy = vc / F2X;
                const int gb = (hy * F1X + hx) * 16 + half;
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                u32x4 raw[NSTEP16];
#pragma unroll
                for (int s = 0; s < NSTEP16; ++s) {
                    const int tA = 2 * s, tB = 2 * s + 1;
                    const uint16_t *pa = Hp3[tap_dz(tA) + 1] + ((tap_dy(tA) + 1) * F1X + tap_dx(tA) + 1) * 16;
                    if (tB < 27) {
                        const uint16_t *pb = Hp3[tap_dz(tB) + 1] + ((tap_dy(tB) + 1) * F1X + tap_dx(tB) + 1) * 16;
                        raw[s] = *(const u32x4 *)((hi ? pb : pa) + gb);
                    } else {
                        raw[s] = *(const u32x4 *)(pa + gb);
                        if (hi) raw[s] = u32x4{0u, 0u, 0u, 0u};
                    }
                }
#pragma unroll
                for (int s = 0; s < NSTEP16; ++s) acc = mfma16(A2[s], __builtin_bit_cast(bf16x8, raw[s]), acc);
                // schedule: the 14 LDS reads first, then the 14 MFMAs (the default schedule
                // issues each read just before its MFMA and waits on it)
                __builtin_amdgcn_sched_group_barrier(0x100, 14, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 14, 0);
                const int gx = ox - 1 + hx, gy = oy - 1 + hy;
                const bool in_grid = zin && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
                u32x2 pk; pk[0] = pk[1] = 0u;
                if (in_grid) {
                    pk[0] = f2bf(acc[0] > 0.f ? acc[0] : 0.f) | ((uint32_t)f2bf(acc[1] > 0.f ? acc[1] : 0.f) << 16);
                    pk[1] = f2bf(acc[2] > 0.f ? acc[2] : 0.f) | ((uint32_t)f2bf(acc[3] > 0.f ? acc[3] : 0.f) << 16);
                }
                if (v < F2V) *(u32x2 *)(Hp + v * 16 + 4 * c) = pk;
            }
        }
        __syncthreads();
        const int o = i - 3;  // output plane o uses H2 planes o-1 (slot of i-4 == sm1), o (si), o+1 (sm2)
        if (o >= z0 && o < z1) {
            const uint16_t *Ho3[3] = {H2[sm1], H2[si], H2[sm2]};
            for (int g = w; g < F3G; g += 4) {
                const int gb = (g * F2X + n) * 16 + half;
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                u32x4 raw[NSTEP16];
#pragma unroll
                for (int s = 0; s < NSTEP16; ++s) {
                    const int tA = 2 * s, tB = 2 * s + 1;
                    const uint16_t *pa = Ho3[tap_dz(tA) + 1] + ((tap_dy(tA) + 1) * F2X + tap_dx(tA) + 1) * 16;
                    if (tB < 27) {
                        const uint16_t *pb = Ho3[tap_dz(tB) + 1] + ((tap_dy(tB) + 1) * F2X + tap_dx(tB) + 1) * 16;
                        raw[s] = *(const u32x4 *)((hi ? pb : pa) + gb);
                    } else {
                        raw[s] = *(const u32x4 *)(pa + gb);
                        if (hi) raw[s] = u32x4{0u, 0u, 0u, 0u};
                    }
                }
#pragma unroll
                for (int s = 0; s < NSTEP16; ++s) acc = mfma16(A3[s], __builtin_bit_cast(bf16x8, raw[s]), acc);
                // schedule: the 14 LDS reads first, then the 14 MFMAs (the default schedule
                // issues each read just before its MFMA and waits on it)
                __builtin_amdgcn_sched_group_barrier(0x100, 14, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 14, 0);
                const int gx = ox + n, gy = oy + g;
                if (c == 0 && gx < nx && gy < ny) out[o * plane + (int64_t)gy * nx + gx] = acc[0] * oscale;
            }
        }
    }
}

}  // namespace cnn
}  // namespace hpccg
