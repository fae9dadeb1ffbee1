// Learned preconditioner: network state, training on residual/error pairs that plain CG
// generates, and flexible CG with the network as preconditioner.  No reference counterpart;
// it extends cg (MPI/src/numerics.F90:41-162), whose FP64 path is both its fallback and the
// measure of its answers (FP64 true residual).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cnn_kernels.hpp"
#include "context.hpp"
#include "spmv.hpp"

namespace hpccg {
using namespace cnn;

// ---- network state ----------------------------------------------------------------------------
struct TrainData {
    int count = 0, crop = 0;
    float *r = nullptr;    // [count][crop^3] r / s
    float *e = nullptr;    // [count][crop^3] e / s
    int *margin = nullptr; // [count][6] loss-mask margins (lo/hi per axis)
};

struct CnnNet {
    float *W = nullptr;              // FP32 master weights: W1 | W2 | W3
    float *m = nullptr, *v = nullptr;  // Adam moments
    bf16x8 *pk_f1 = nullptr, *pk_f2 = nullptr;                    // forward operands
    bf16x8 *pk_f3f = nullptr;                                     // layer 3, banded form (both forward kernels)
    bf16x8 *pk_b3 = nullptr, *pk_b2 = nullptr;                    // backward-data operands
    int *step = nullptr;             // device training step counter
    TrainData data;
    // activation / scratch buffers (grown on demand)
    int64_t cap_vox = 0;
    uint16_t *h1 = nullptr, *h2 = nullptr, *dh2 = nullptr, *dh1 = nullptr, *dz = nullptr;
    float *in = nullptr, *z = nullptr, *tgt = nullptr, *wbuf = nullptr;
    float *scale_in = nullptr, *scale_out = nullptr;
    float *partials = nullptr;
    int partials_wg = 0;
    float *grad = nullptr;           // [NWALL]
    float *loss = nullptr;           // [steps] per-step mean loss
    int loss_cap = 0;
    double *lpart = nullptr;         // per (sample, slab) loss partials
    int *batch_meta = nullptr;       // [batch][8]: flips, sign, margins
    hipGraphExec_t train_graph = nullptr;
    int g_batch = 0, g_steps = 0;
};

static CnnNet *net_of(hpccg_ctx *c) { return (CnnNet *)c->ext[1]; }

// ---- weight packing into MFMA A-operands -------------------------------------------------------
// pack_f1[l][j] = W1[co=l%16][t=k1_tap(l/16, j)];
// pack_f2[s][l][j] = W2[co=l%16][ci=8((l/16)&1)+j][t=zk_tap(s, (l/16)/2)] (z-major K order);
// pack_f3f[s][l][j] = W3[ci][t] of the banded layer 3 (cnn_kernels.hpp, NSTEP3B): MFMA row
//   m = l%16 = 4 dzi + jr, K slot (window row s/3 + 2 (l/16 >= 2), x-offset s%3, ci =
//   8((l/16)&1)+j), t = 9 dzi + 3 (row - jr) + x-offset where 0 <= row - jr <= 2, else 0;
// pack_b3[l][j] = W3[l%16][26 - k1_tap(l/16, j)];
// pack_b2[s][l][j] = W2[ci'][co'=l%16][26 - t] with ci' = 8((l/16)&1)+j (transposed, flipped).
__global__ void k_pack(const float *__restrict__ W, uint16_t *f1, uint16_t *f2, uint16_t *f3f, uint16_t *b3,
                       uint16_t *b2)
{
    const float *W1 = W, *W2 = W + NW1, *W3 = W + NW1 + NW2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 64 * 8) {  // 1-step packs (one-channel input: K order k1_tap)
        const int l = i / 8, j = i % 8, co = l % 16, t = k1_tap(l / 16, j);
        f1[i] = t >= 0 ? f2bf(W1[co * 27 + t]) : 0;
        b3[i] = t >= 0 ? f2bf(W3[co * 27 + (26 - t)]) : 0;
    }
    if (i < NSTEP16 * 64 * 8) {
        const int s = i / 512, l = (i / 8) % 64, j = i % 8;
        const int co = l % 16, ci = 8 * ((l / 16) & 1) + j, t = zk_tap(s, (l / 16) / 2);
        f2[i] = t >= 0 ? f2bf(W2[(co * 16 + ci) * 27 + t]) : 0;
        b2[i] = t >= 0 ? f2bf(W2[(ci * 16 + co) * 27 + (26 - t)]) : 0;
    }
    if (i < NSTEP3B * 64 * 8) {  // banded layer 3
        const int s = i / 512, l = (i / 8) % 64, j = i % 8;
        const int m = l % 16, cl = l / 16, ci = 8 * (cl & 1) + j;
        const int row = s / 3 + (cl >= 2 ? L3ROWS : 0), dxi = s % 3;  // window row and x-offset + 1 of this K slot
        const int dzi = m / 4, jr = m % 4, dyi = row - jr;  // MFMA row m = 4 dzi + jr
        const bool on = dzi < 3 && jr < L3ROWS && dyi >= 0 && dyi <= 2;
        f3f[i] = on ? f2bf(W3[ci * 27 + 9 * dzi + 3 * dyi + dxi]) : 0;
    }
}

// dW (unpadded layouts) = sum over workgroups of the padded [16][16][27] partials
__global__ void k_wgrad_reduce(const float *__restrict__ partials, int nwg, int layer, float *__restrict__ grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int src;
    if (layer == 1) {  // W1[co][t] = P[co][ci=0][t]
        if (i >= NW1) return;
        src = (i / 27) * 16 * 27 + (i % 27);
    } else if (layer == 2) {
        if (i >= NW2) return;
        src = i;
    } else {  // W3[ci][t] = P[co=0][ci][t]
        if (i >= NW3) return;
        src = i;
    }
    float s = 0.f;
    for (int w = 0; w < nwg; ++w) s += partials[(int64_t)w * (16 * 16 * 27) + src];
    grad[i] = s;
}

// ---- launch helpers -----------------------------------------------------------------------------
template <int IK, int OK>
static void conv(hpccg_ctx *c, const void *in, const float *in_scale, const bf16x8 *pk, void *out,
                 const uint16_t *mask, const float *out_scale, const Grid3 &g, int batch, const int *skip = nullptr)
{
    hipLaunchKernelGGL((k_conv<IK, OK>), dim3(g.ntiles() * batch), dim3(256), 0, c->stream, in, in_scale, pk, out,
                       mask, out_scale, g, skip);
}

static int wgrad_wg(hpccg_ctx *c, const Grid3 &g, int batch)
{
    return std::max(1, std::min(g.ntiles() * batch, c->info.num_cu * 2));
}

static int ensure_net(hpccg_ctx *c)
{
    if (net_of(c)) return 0;
    CnnNet *N = new CnnNet();
    c->ext[1] = N;
    if (dmalloc(&N->W, NWALL * 4) != hipSuccess || dmalloc(&N->m, NWALL * 4) != hipSuccess ||
        dmalloc(&N->v, NWALL * 4) != hipSuccess || dmalloc(&N->grad, NWALL * 4) != hipSuccess ||
        dmalloc(&N->pk_f1, 64 * 16) != hipSuccess || dmalloc(&N->pk_b3, 64 * 16) != hipSuccess ||
        dmalloc(&N->pk_f2, NSTEP16 * 64 * 16) != hipSuccess ||
        dmalloc(&N->pk_f3f, NSTEP3B * 64 * 16) != hipSuccess ||
        dmalloc(&N->pk_b2, NSTEP16 * 64 * 16) != hipSuccess || dmalloc(&N->step, sizeof(int)) != hipSuccess ||
        dmalloc(&N->scale_in, 4096 * 4) != hipSuccess || dmalloc(&N->scale_out, 4096 * 4) != hipSuccess) {
        set_error("cnn: allocation failed");
        return HPCCG_ERR_NOMEM;
    }
    std::vector<float> zero(NWALL, 0.f);
    HPCCG_HIP_TRY(hipMemcpy(N->W, zero.data(), NWALL * 4, hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(N->m, zero.data(), NWALL * 4, hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(N->v, zero.data(), NWALL * 4, hipMemcpyHostToDevice));
    int z0 = 0;
    HPCCG_HIP_TRY(hipMemcpy(N->step, &z0, sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

static int repack(hpccg_ctx *c)
{
    CnnNet *N = net_of(c);
    hipLaunchKernelGGL(k_pack, dim3((NSTEP16 * 512 + 255) / 256), dim3(256), 0, c->stream, N->W, (uint16_t *)N->pk_f1,
                       (uint16_t *)N->pk_f2, (uint16_t *)N->pk_f3f, (uint16_t *)N->pk_b3,
                       (uint16_t *)N->pk_b2);
    HPCCG_HIP_TRY(hipGetLastError());
    return 0;
}

static int ensure_act(hpccg_ctx *c, int64_t vox, int wg)
{
    CnnNet *N = net_of(c);
    if (vox > N->cap_vox) {
        dfree(N->h1);
        dfree(N->h2);
        dfree(N->dh1);
        dfree(N->dh2);
        dfree(N->dz);
        dfree(N->in);
        dfree(N->z);
        dfree(N->tgt);
        dfree(N->wbuf);
        N->cap_vox = 0;
        if (dmalloc(&N->h1, vox * 32) != hipSuccess || dmalloc(&N->h2, vox * 32) != hipSuccess ||
            dmalloc(&N->dh1, vox * 32) != hipSuccess || dmalloc(&N->dh2, vox * 32) != hipSuccess ||
            dmalloc(&N->dz, vox * 2) != hipSuccess || dmalloc(&N->in, vox * 4) != hipSuccess ||
            dmalloc(&N->z, vox * 4) != hipSuccess || dmalloc(&N->tgt, vox * 4) != hipSuccess ||
            dmalloc(&N->wbuf, vox * 4) != hipSuccess) {
            set_error("cnn: activation allocation failed (%lld voxels)", (long long)vox);
            return HPCCG_ERR_NOMEM;
        }
        N->cap_vox = vox;
        if (N->train_graph) hipGraphExecDestroy(N->train_graph);
        N->train_graph = nullptr;
        invalidate_learned_graphs(c);  // graphs captured with the old activations (already stale)
    }
    if (wg > N->partials_wg) {
        dfree(N->partials);
        N->partials = nullptr;
        HPCCG_HIP_TRY(dmalloc(&N->partials, (size_t)wg * 16 * 16 * 27 * 4));
        N->partials_wg = wg;
        if (N->train_graph) hipGraphExecDestroy(N->train_graph);
        N->train_graph = nullptr;
    }
    return 0;
}

// forward: in (FP32, times scale_in[b]) -> h1, h2 -> z (FP32 times scale_out[b])
static void forward(hpccg_ctx *c, const float *in, const float *s_in, float *z, const float *s_out, const Grid3 &g,
                    int batch, const int *skip = nullptr)
{
    CnnNet *N = net_of(c);
    conv<IN_F32_1, OUT_RELU_BF16>(c, in, s_in, N->pk_f1, N->h1, nullptr, nullptr, g, batch, skip);
    conv<IN_BF16_16, OUT_RELU_BF16>(c, N->h1, nullptr, N->pk_f2, N->h2, nullptr, nullptr, g, batch, skip);
    hipLaunchKernelGGL(k_conv3, dim3(g.ntiles() * batch), dim3(256), 0, c->stream, N->h2, N->pk_f3f, z, s_out, g, skip);
}

// inference on one field: the fused kernel (bitwise equal to forward(), no activation traffic)
static void forward_fused(hpccg_ctx *c, const float *in, const float *s_in, float *z, const float *s_out,
                          const Grid3 &g, const int *skip = nullptr)
{
    CnnNet *N = net_of(c);
    const int tx = (g.nx + FX - 1) / FX, ty = (g.ny + FY - 1) / FY;
    // about one workgroup per CU (registers admit one 8-wave workgroup per CU), z-chunks of at
    // least 8 planes (each chunk re-reads 6 planes of halo)
    const int want = std::max(1, (c->info.num_cu + tx * ty - 1) / (tx * ty));
    const int nzc = std::max(1, std::min(want, (g.nz + 7) / 8));
    const int zchunk = (g.nz + nzc - 1) / nzc;
    const int chunks = (g.nz + zchunk - 1) / zchunk;
    hipLaunchKernelGGL(k_fused_fwd, dim3(tx * ty * chunks), dim3(FTHREADS), 0, c->stream, in, s_in, N->pk_f1,
                       N->pk_f2, N->pk_f3f, z, s_out, g.nx, g.ny, g.nz, tx, ty, zchunk, skip);
}

// backward from dz (BF16 [b][v]) and the stored forward activations: grad = dL/dW
static void backward(hpccg_ctx *c, const float *in, const float *s_in, const Grid3 &g, int batch)
{
    CnnNet *N = net_of(c);
    const int wg = wgrad_wg(c, g, batch);
    // layer 3: dW3 from (dz, h2); dh2 = mask(h2) * conv^T(dz)
    hipLaunchKernelGGL((k_wgrad<1, 16>), dim3(wg), dim3(256), 0, c->stream, N->dz, N->h2, N->partials, g, batch);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((NW3 + 255) / 256), dim3(256), 0, c->stream, N->partials, wg, 3,
                       N->grad + NW1 + NW2);
    conv<IN_BF16_1, OUT_MASKED_BF16>(c, N->dz, nullptr, N->pk_b3, N->dh2, N->h2, nullptr, g, batch);
    // layer 2: dW2 from (dh2, h1); dh1 = mask(h1) * conv^T(dh2)
    hipLaunchKernelGGL((k_wgrad<16, 16>), dim3(wg), dim3(256), 0, c->stream, N->dh2, N->h1, N->partials, g, batch);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((NW2 + 255) / 256), dim3(256), 0, c->stream, N->partials, wg, 2,
                       N->grad + NW1);
    conv<IN_BF16_16, OUT_MASKED_BF16>(c, N->dh2, nullptr, N->pk_b2, N->dh1, N->h1, nullptr, g, batch);
    // layer 1: dW1 from (dh1, u) with u = bf16(in * s_in), re-rounded here into dz's buffer? no:
    // the input is staged from FP32 by the wgrad kernel's BF16 view, so round it once first
    (void)in;
    (void)s_in;
    hipLaunchKernelGGL((k_wgrad<16, 1>), dim3(wg), dim3(256), 0, c->stream, N->dh1, (const uint16_t *)N->wbuf,
                       N->partials, g, batch);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((NW1 + 255) / 256), dim3(256), 0, c->stream, N->partials, wg, 1, N->grad);
}

// u = bf16(in * s_in[b]) into a BF16 [b][v] buffer (the exact values layer 1 consumed)
__global__ void k_round_input(const float *__restrict__ in, const float *__restrict__ s_in, uint16_t *__restrict__ u,
                              int64_t nvox, int batch)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nvox * batch;
         i += (int64_t)gridDim.x * blockDim.x)
        u[i] = f2bf(in[i] * (s_in ? s_in[i / nvox] : 1.0f));
}

// ---- training loss: A^2-norm of the error on the crop's valid region -----------------------------
// d = z - e (both scaled by 1/s), w = mask * A d; L_b = |w|^2 / |mask * A e|^2;
// dL/dz = 2 A w / |mask * A e|^2 (A is symmetric on the zero-padded crop).  mask excludes
// margin[] voxels on crop faces that are not domain faces (their z lacks part of its 7^3
// receptive field, and A needs one more layer).
__device__ inline double stencil_apply(const float *f, int x, int y, int z, int C)
{
    double s = 0.0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = x + dx, yy = y + dy, zz = z + dz;
                if (xx < 0 || xx >= C || yy < 0 || yy >= C || zz < 0 || zz >= C) continue;
                s += (double)f[((int64_t)zz * C + yy) * C + xx];
            }
    return 28.0 * (double)f[((int64_t)z * C + y) * C + x] - s;
}

__device__ inline bool in_mask(const int *mg, int x, int y, int z, int C)
{
    return x >= mg[0] && x < C - mg[1] && y >= mg[2] && y < C - mg[3] && z >= mg[4] && z < C - mg[5];
}

// one workgroup per (sample, z-slab); writes w = mask*(A(z - e)) into wbuf and the slab's sums
__global__ __launch_bounds__(256) void k_loss_a(const float *__restrict__ z, const float *__restrict__ e,
                                                float *__restrict__ wbuf, const int *__restrict__ meta, int C,
                                                int slabs, double *__restrict__ lpart)
{
    __shared__ double red[4];
    const int b = blockIdx.x / slabs, sl = blockIdx.x % slabs;
    const int64_t V = (int64_t)C * C * C;
    const float *zb = z + b * V, *eb = e + b * V;
    float *wb = wbuf + b * V;
    const int *mg = meta + b * 8 + 2;
    const int z0 = (int)((int64_t)sl * C / slabs), z1 = (int)((int64_t)(sl + 1) * C / slabs);
    double num = 0.0, den = 0.0;
    for (int i = threadIdx.x; i < (z1 - z0) * C * C; i += 256) {
        const int x = i % C, y = (i / C) % C, zz = z0 + i / (C * C);
        double w = 0.0;
        if (in_mask(mg, x, y, zz, C)) {
            const double az = stencil_apply(zb, x, y, zz, C), ae = stencil_apply(eb, x, y, zz, C);
            w = az - ae;
            num += w * w;
            den += ae * ae;
        }
        wb[((int64_t)zz * C + y) * C + x] = (float)w;
    }
    num = block_sum<256>(num, red);
    den = block_sum<256>(den, red);
    if (threadIdx.x == 0) {
        lpart[2 * blockIdx.x] = num;
        lpart[2 * blockIdx.x + 1] = den;
    }
}

// dz = bf16(2 A w / den_b); also records the batch-mean loss of this step
__global__ __launch_bounds__(256) void k_loss_b(const float *__restrict__ wbuf, uint16_t *__restrict__ dz, int C,
                                                int slabs, int batch, const double *__restrict__ lpart,
                                                float *__restrict__ loss_out, const int *step)
{
    const int b = blockIdx.x / slabs, sl = blockIdx.x % slabs;
    const int64_t V = (int64_t)C * C * C;
    double den = 0.0;
    for (int i = 0; i < slabs; ++i) den += lpart[2 * (b * slabs + i) + 1];
    const double inv = den > 0.0 ? 2.0 / den : 0.0;
    const float *wb = wbuf + b * V;
    const int z0 = (int)((int64_t)sl * C / slabs), z1 = (int)((int64_t)(sl + 1) * C / slabs);
    for (int i = threadIdx.x; i < (z1 - z0) * C * C; i += 256) {
        const int x = i % C, y = (i / C) % C, zz = z0 + i / (C * C);
        dz[b * V + ((int64_t)zz * C + y) * C + x] = f2bf((float)(stencil_apply(wb, x, y, zz, C) * inv / batch));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) {
        double tot = 0.0;
        for (int bb = 0; bb < batch; ++bb) {
            double nn = 0.0, dd = 0.0;
            for (int i = 0; i < slabs; ++i) {
                nn += lpart[2 * (bb * slabs + i)];
                dd += lpart[2 * (bb * slabs + i) + 1];
            }
            tot += dd > 0.0 ? nn / dd : 0.0;
        }
        loss_out[*step] = (float)(tot / batch);
    }
}

// ---- batch assembly (device RNG: SplitMix64 keyed by seed, step, sample) ----------------------
__device__ inline uint64_t mix64d(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// sample b of this step: crop index, sign and axis flips (the 27-point operator is invariant
// under reflections, so a reflected pair is a valid pair); copies the crop into the batch
// buffers and writes its (reflected) loss margins
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ R, const float *__restrict__ E,
                                                const int *__restrict__ margin, int count, int C,
                                                float *__restrict__ in, float *__restrict__ tgt, int *meta,
                                                const int *step, uint64_t seed)
{
    const int b = blockIdx.y;
    const uint64_t h = mix64d(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(1 + *step * 4096 + b));
    const int idx = (int)(h % (uint64_t)count);
    const int fx = (h >> 32) & 1, fy = (h >> 33) & 1, fz = (h >> 34) & 1;
    const float sg = ((h >> 35) & 1) ? -1.f : 1.f;
    const int64_t V = (int64_t)C * C * C;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int *m = margin + idx * 6;
        int *o = meta + b * 8;
        o[0] = idx;
        o[1] = (fx) | (fy << 1) | (fz << 2);
        o[2] = fx ? m[1] : m[0];
        o[3] = fx ? m[0] : m[1];
        o[4] = fy ? m[3] : m[2];
        o[5] = fy ? m[2] : m[3];
        o[6] = fz ? m[5] : m[4];
        o[7] = fz ? m[4] : m[5];
    }
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % C), y = (int)((i / C) % C), z = (int)(i / ((int64_t)C * C));
        const int sx = fx ? C - 1 - x : x, sy = fy ? C - 1 - y : y, sz = fz ? C - 1 - z : z;
        const int64_t src = (int64_t)idx * V + ((int64_t)sz * C + sy) * C + sx;
        in[b * V + i] = sg * R[src];
        tgt[b * V + i] = sg * E[src];
    }
}

// Adam (Kingma & Ba 2015) with warm-up and a cosine learning-rate schedule over `total` steps;
// the weight repack is launched after it.
__global__ void k_adam(float *__restrict__ W, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, int n, float lr0, int total, int *step)
{
    const int t = *step + 1;
    // linear warm-up over the first 5% of the steps, then cosine decay to zero
    const int warm = max(1, total / 20);
    const float lr = t <= warm ? lr0 * (float)t / (float)warm
                               : lr0 * 0.5f * (1.0f + cosf(3.14159265f * (float)(t - warm) / (float)max(1, total - warm)));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        W[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}
__global__ void k_step_inc(int *step)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// ---- training-data generation from plain CG ---------------------------------------------------------
// crop of r_k and e_k = x* - (x + alpha p) at origin (ox,oy,oz), scaled by 1/rms(r_crop)
__global__ __launch_bounds__(256) void k_crop(const double *__restrict__ r, const double *__restrict__ x,
                                              const double *__restrict__ p, const double *__restrict__ xs,
                                              const CgState *st, int n, int C, int ox, int oy, int oz,
                                              float *__restrict__ R, float *__restrict__ E)
{
    __shared__ double red[4];
    const double alpha = st->alpha;
    const int64_t V = (int64_t)C * C * C;
    double ss = 0.0;
    for (int64_t i = threadIdx.x; i < V; i += 256) {
        const int cx = (int)(i % C), cy = (int)((i / C) % C), cz = (int)(i / ((int64_t)C * C));
        const int64_t g = ((int64_t)(oz + cz) * n + (oy + cy)) * n + (ox + cx);
        ss += r[g] * r[g];
    }
    ss = block_sum<256>(ss, red);
    const double inv = ss > 0.0 ? 1.0 / sqrt(ss / (double)V) : 1.0;
    for (int64_t i = threadIdx.x; i < V; i += 256) {
        const int cx = (int)(i % C), cy = (int)((i / C) % C), cz = (int)(i / ((int64_t)C * C));
        const int64_t g = ((int64_t)(oz + cz) * n + (oy + cy)) * n + (ox + cx);
        R[i] = (float)(r[g] * inv);
        E[i] = (float)((xs[g] - (x[g] + alpha * p[g])) * inv);
    }
}

// ---- flexible CG with the network as preconditioner (mixed precision) ----------------------------
struct FcgState {
    double alpha, beta, rz, pap, tol, rr;
    double rr_true;
    int k, stop_at, max_iter, replace_every, need_true, converged, breakdown, done;
    int x_nonzero;      // set by Prepare's k_nonzero when x0 != 0 (else A x0 is skipped)
    float s_in, s_out;  // network input / output scales (1/rms(r), rms(r))
};

// r = b - A x (FP32 store, FP64 arithmetic) and the workgroup's share of r.r
template <bool PREP>
__device__ __forceinline__ bool kf_true_due(const FcgState *st)
{
    return PREP || (st->k < st->stop_at && st->need_true);
}

template <bool PREP>
__global__ __launch_bounds__(kTrueThreads) void kf_true(const SellRef<__half> A, const double *__restrict__ x,
                                                        const double *__restrict__ b, float *__restrict__ r,
                                                        const FcgState *st, double *red_part)
{
    __shared__ double red[kTrueThreads / kWave];
    if (!kf_true_due<PREP>(st)) return;
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

// r.r of the true residual: network scales, stopping test; PREP also decides the first stop
template <bool PREP>
__global__ __launch_bounds__(kFinishThreads) void kf_true_finish(const double *__restrict__ red_part, int nparts,
                                                                 int nrow, FcgState *st, double *hist_true)
{
    __shared__ double red[kFinishThreads / kWave];
    if (!kf_true_due<PREP>(st)) return;
    const int k = PREP ? 0 : st->k;
    const double rr = sum_partials(red_part, nparts, red);
    if (threadIdx.x == 0) {
        hist_true[k] = rr;
        st->rr = rr;
        st->rr_true = rr;
        const double rms = sqrt(rr / (double)nrow);
        st->s_in = rms > 0.0 ? (float)(1.0 / rms) : 1.0f;
        st->s_out = (float)rms;
        st->need_true = 0;
        if (sqrt(rr) < st->tol) {
            st->converged = 1;
            st->stop_at = PREP ? 1 : min(st->stop_at, k + 1);
        } else if (!PREP && k >= st->max_iter) {
            st->stop_at = min(st->stop_at, k + 1);
        }
        if (!PREP) st->k = k + 1;
        if (PREP) st->k = 1;
        st->done = st->k >= st->stop_at;
    }
}

// Ap = A p and the workgroup's share of p.Ap
__global__ __launch_bounds__(kSellThreads) __attribute__((amdgpu_waves_per_eu(kSellF16WavesPerEU, kSellF16WavesPerEU))) void kf_spmv(const SellRef<__half> A, const float *__restrict__ p,
                                                        float *__restrict__ Ap, const FcgState *st, double *red_part)
{
    __shared__ double red[kSellThreads / kWave];
    if (st->k >= st->stop_at) return;
    int64_t row;
    bool valid;
    const double s = sell_rowsum_of_thread<kGroupF16>(A, p, row, valid);
    double dot = 0.0;
    if (valid) {
        const float a = (float)s;
        __builtin_nontemporal_store(a, Ap + row);
        dot = (double)p[row] * (double)a;
    }
    store_partial(block_sum<kSellThreads>(dot, red), red_part);
}

// p.Ap -> alpha = (r.z) / (p.Ap); a breakdown stops before x is touched
__global__ __launch_bounds__(kFinishThreads) void kf_spmv_finish(const double *__restrict__ red_part, int nparts,
                                                                 FcgState *st)
{
    __shared__ double red[kFinishThreads / kWave];
    if (st->k >= st->stop_at) return;
    const double pap = sum_partials(red_part, nparts, red);
    if (threadIdx.x == 0) {
        st->pap = pap;
        st->alpha = st->rz / pap;
        if (!(pap > 0.0) || !isfinite(pap) || !isfinite(st->alpha)) {
            st->breakdown = 1;
            st->stop_at = st->k;
            st->done = 1;
        }
    }
}

// x += alpha p (FP64), r -= alpha Ap (FP32), and the workgroup's r.r
__global__ __launch_bounds__(kVecThreads) void kf_update_xr(double *__restrict__ x, float *__restrict__ r,
                                                            const float *__restrict__ p,
                                                            const float *__restrict__ Ap, int n, const FcgState *st,
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

// r.r: network scales; requests a true residual when due
__global__ __launch_bounds__(kFinishThreads) void kf_xr_finish(const double *__restrict__ partials, int nparts, int n,
                                                               FcgState *st, double *hist_rec)
{
    __shared__ double red[kFinishThreads / kWave];
    const int k = st->k;
    if (k >= st->stop_at) return;
    const double rr = sum_partials(partials, nparts, red);
    if (threadIdx.x == 0) {
        hist_rec[k] = rr;
        st->rr = rr;
        const double rms = sqrt(rr / (double)n);
        st->s_in = rms > 0.0 ? (float)(1.0 / rms) : 1.0f;
        st->s_out = (float)rms;
        const bool check = (st->replace_every > 0 && k % st->replace_every == 0) || sqrt(rr) < st->tol ||
                           k >= st->max_iter || rr < kReplaceDrop2 * st->rr_true;
        if (check) st->need_true = 1;
        else st->k = k + 1;
        st->done = st->k >= st->stop_at;
    }
}

// z.Ap and z.r of the workgroup (two partials per workgroup)
template <bool PREP>
__global__ __launch_bounds__(kVecThreads) void kf_dots(const float *__restrict__ z, const float *__restrict__ Ap,
                                                       const float *__restrict__ r, int n, const FcgState *st,
                                                       double *partials)
{
    __shared__ double red[kVecThreads / kWave];
    if (!PREP && st->k >= st->stop_at) return;
    if (PREP && st->stop_at <= 1) return;
    int64_t i0, i1;
    int lb;
    elem_range4(n, i0, i1, lb);
    double za = 0.0, zr = 0.0;
    auto one = [&](float zi, float ai, float ri) {
        if (!PREP) za = fma((double)zi, (double)ai, za);
        zr = fma((double)zi, (double)ri, zr);
    };
    vec4_loop(
        i0, i1,
        [&](int64_t i) {
            const f32x4v zv = ld4(z + i), rv = ld4_nt(r + i);
            const f32x4v av = PREP ? f32x4v{0.f, 0.f, 0.f, 0.f} : ld4_nt(Ap + i);
            one(zv.x, av.x, rv.x);
            one(zv.y, av.y, rv.y);
            one(zv.z, av.z, rv.z);
            one(zv.w, av.w, rv.w);
        },
        [&](int64_t i) { one(z[i], PREP ? 0.f : Ap[i], r[i]); });
    za = block_sum<kVecThreads>(za, red);
    zr = block_sum<kVecThreads>(zr, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = za;
        partials[2 * blockIdx.x + 1] = zr;
    }
}

// beta and p.  Every workgroup sums the z.Ap / z.r partials of kf_dots itself (two per kf_dots
// workgroup, 2,048 values at 256 CUs; the same fixed order in every workgroup, so all agree) and forms beta = -(z.Ap)/(p.Ap)
// (flexible CG, Notay 2000) -- no separate finish launch.  Workgroup 0 records beta and
// rz = z.r.  A z that is not a descent direction (z.r <= 0 or not finite) stops the loop for
// the host's fallback, and no workgroup touches p.  FIRST (Prepare): rz only, p = z.
template <bool FIRST>
__global__ __launch_bounds__(kVecThreads) void kf_update_p(const float *__restrict__ z, float *__restrict__ p, int n,
                                                           FcgState *st, const double *__restrict__ partials,
                                                           int nparts)
{
    __shared__ double red[kVecThreads / kWave];
    if (!FIRST && st->k >= st->stop_at) return;
    if (FIRST && st->stop_at <= 1) return;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kVecThreads) {
        a += partials[2 * i];
        b += partials[2 * i + 1];
    }
    a = block_sum<kVecThreads>(a, red);
    b = block_sum<kVecThreads>(b, red);
    const double beta = FIRST ? 0.0 : -a / st->pap;
    const bool breakdown = !(b > 0.0) || !isfinite(b) || !isfinite(beta);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->beta = beta;
        st->rz = b;
        if (breakdown) {
            st->breakdown = 1;
            st->stop_at = st->k;
            st->done = 1;
        }
    }
    if (breakdown) return;
    int64_t i0, i1;
    int lb;
    elem_range4(n, i0, i1, lb);
    auto one = [&](float zi, float pi) { return FIRST ? zi : (float)fma(beta, (double)pi, (double)zi); };
    vec4_loop(
        i0, i1,
        [&](int64_t i) {
            const f32x4v zv = ld4_nt(z + i);  // streamed: keeps p, which the SpMV gathers next, cached
            f32x4v pv = FIRST ? zv : ld4(p + i);
            pv.x = one(zv.x, pv.x);
            pv.y = one(zv.y, pv.y);
            pv.z = one(zv.z, pv.z);
            pv.w = one(zv.w, pv.w);
            st4(p + i, pv);
        },
        [&](int64_t i) { p[i] = one(z[i], FIRST ? 0.f : p[i]); });
}

struct FcgWs {
    int64_t n = 0;
    float *r = nullptr, *p = nullptr, *Ap = nullptr, *z = nullptr;
    FcgState *st = nullptr, *st_h = nullptr;
    double *hist = nullptr;
    int hist_cap = 0;
    uint16_t *h1 = nullptr, *h2 = nullptr;
    int gdims[3] = {0, 0, 0};  // grid the cached graphs were captured for
    GraphCache graphs;
};

static FcgWs *fcg_of(hpccg_ctx *c) { return (FcgWs *)c->ext[2]; }

void invalidate_learned_graphs(hpccg_ctx *c)
{
    if (FcgWs *F = fcg_of(c)) F->graphs.clear();
}

void free_learned(hpccg_ctx *c)
{
    if (CnnNet *N = net_of(c)) {
        if (N->train_graph) hipGraphExecDestroy(N->train_graph);
        for (void *p : {(void *)N->W, (void *)N->m, (void *)N->v, (void *)N->pk_f1, (void *)N->pk_f2,
                        (void *)N->pk_f3f, (void *)N->pk_b3, (void *)N->pk_b2, (void *)N->step, (void *)N->h1,
                        (void *)N->h2, (void *)N->dh2, (void *)N->dh1, (void *)N->dz, (void *)N->in, (void *)N->z,
                        (void *)N->tgt, (void *)N->wbuf, (void *)N->scale_in, (void *)N->scale_out,
                        (void *)N->partials, (void *)N->grad, (void *)N->loss, (void *)N->lpart,
                        (void *)N->batch_meta, (void *)N->data.r, (void *)N->data.e, (void *)N->data.margin})
            dfree(p);
        delete N;
        c->ext[1] = nullptr;
    }
    if (FcgWs *F = fcg_of(c)) {
        F->graphs.clear();
        dfree(F->r);
        dfree(F->p);
        dfree(F->Ap);
        dfree(F->z);
        dfree(F->st);
        hipHostFree(F->st_h);
        dfree(F->hist);
        dfree(F->h1);
        dfree(F->h2);
        delete F;
        c->ext[2] = nullptr;
    }
}

}  // namespace hpccg

using namespace hpccg;
using namespace hpccg::cnn;

extern "C" {

int hpccg_cnn_num_weights(void) { return NWALL; }

// Initial weights uniform in +-1/sqrt(fan_in) (fan_in = 27 * input channels; SplitMix64 draws),
// Adam state reset.  Small enough that no ReLU layer starts saturated.
int hpccg_cnn_init(hpccg_ctx *c, uint64_t seed)
{
    HPCCG_CHECK_ARG(c, "null context");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    HPCCG_TRY(ensure_net(c));
    std::vector<float> w(NWALL);
    uint64_t s = seed;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return ((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
    };
    const double b1 = 1.0 / std::sqrt(27.0), b2 = 1.0 / std::sqrt(16.0 * 27.0);
    for (int i = 0; i < NW1; ++i) w[i] = (float)(b1 * (2.0 * next() - 1.0));
    for (int i = 0; i < NW2 + NW3; ++i) w[NW1 + i] = (float)(b2 * (2.0 * next() - 1.0));
    CnnNet *N = net_of(c);
    HPCCG_HIP_TRY(hipMemcpy(N->W, w.data(), NWALL * 4, hipMemcpyHostToDevice));
    std::vector<float> zero(NWALL, 0.f);
    HPCCG_HIP_TRY(hipMemcpy(N->m, zero.data(), NWALL * 4, hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(N->v, zero.data(), NWALL * 4, hipMemcpyHostToDevice));
    int z0 = 0;
    HPCCG_HIP_TRY(hipMemcpy(N->step, &z0, sizeof(int), hipMemcpyHostToDevice));
    HPCCG_TRY(repack(c));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_cnn_set_weights(hpccg_ctx *c, const float *w)
{
    HPCCG_CHECK_ARG(c && w, "null argument");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    HPCCG_TRY(ensure_net(c));
    HPCCG_HIP_TRY(hipMemcpy(net_of(c)->W, w, NWALL * 4, hipMemcpyHostToDevice));
    HPCCG_TRY(repack(c));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_cnn_get_weights(hpccg_ctx *c, float *w)
{
    HPCCG_CHECK_ARG(c && w, "null argument");
    if (!net_of(c)) {
        set_error("no network: call hpccg_cnn_init or hpccg_cnn_set_weights");
        return HPCCG_ERR_STATE;
    }
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    HPCCG_HIP_TRY(hipMemcpy(w, net_of(c)->W, NWALL * 4, hipMemcpyDeviceToHost));
    return 0;
}

// Forward pass on host arrays: u[batch][nvox] (FP32, used as BF16), outputs h1/h2 as FP32
// copies of the BF16 activations ([batch][nvox][16], may be NULL) and z[batch][nvox].
int hpccg_cnn_forward(hpccg_ctx *c, int32_t nx, int32_t ny, int32_t nz, int32_t batch, const float *u, float *h1,
                      float *h2, float *z)
{
    HPCCG_CHECK_ARG(c && u && z && nx >= 1 && ny >= 1 && nz >= 1 && batch >= 1, "bad argument");
    CnnNet *N = net_of(c);
    if (!N) {
        set_error("no network");
        return HPCCG_ERR_STATE;
    }
    const Grid3 g = make_grid(nx, ny, nz);
    const int64_t V = g.nvox() * batch;
    HPCCG_TRY(ensure_act(c, V, wgrad_wg(c, g, batch)));
    HPCCG_HIP_TRY(hipMemcpyAsync(N->in, u, V * 4, hipMemcpyHostToDevice, c->stream));
    forward(c, N->in, nullptr, N->z, nullptr, g, batch);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipMemcpyAsync(z, N->z, V * 4, hipMemcpyDeviceToHost, c->stream));
    std::vector<uint16_t> tmp;
    for (int k = 0; k < 2; ++k) {
        float *dst = k == 0 ? h1 : h2;
        if (!dst) continue;
        tmp.resize(V * 16);
        HPCCG_HIP_TRY(hipMemcpyAsync(tmp.data(), k == 0 ? N->h1 : N->h2, V * 32, hipMemcpyDeviceToHost, c->stream));
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < V * 16; ++i) dst[i] = bf2f(tmp[i]);
    }
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Forward + backward on host arrays for a given dL/dz (FP32, used as BF16): dW[NWALL].
int hpccg_cnn_backward(hpccg_ctx *c, int32_t nx, int32_t ny, int32_t nz, int32_t batch, const float *u,
                       const float *dz, float *dW)
{
    HPCCG_CHECK_ARG(c && u && dz && dW && nx >= 1 && ny >= 1 && nz >= 1 && batch >= 1, "bad argument");
    CnnNet *N = net_of(c);
    if (!N) {
        set_error("no network");
        return HPCCG_ERR_STATE;
    }
    const Grid3 g = make_grid(nx, ny, nz);
    const int64_t V = g.nvox() * batch;
    HPCCG_TRY(ensure_act(c, V, wgrad_wg(c, g, batch)));
    HPCCG_HIP_TRY(hipMemcpyAsync(N->in, u, V * 4, hipMemcpyHostToDevice, c->stream));
    forward(c, N->in, nullptr, N->z, nullptr, g, batch);
    std::vector<uint16_t> dzq(V);
    for (int64_t i = 0; i < V; ++i) dzq[i] = f2bf(dz[i]);
    HPCCG_HIP_TRY(hipMemcpyAsync(N->dz, dzq.data(), V * 2, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_round_input, dim3(c->info.num_cu * 4), dim3(256), 0, c->stream, N->in, nullptr,
                       (uint16_t *)N->wbuf, g.nvox(), batch);
    backward(c, N->in, nullptr, g, batch);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipMemcpyAsync(dW, N->grad, NWALL * 4, hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Training set: `npairs` (r_k, e_k) pairs from plain FP64 CG runs on n^3 grids (x* = 8 random
// low-frequency sines + noise, a fresh seed per run, b = A x*, to relative residual 1e-10),
// each stored as a random crop^3 window scaled by 1/rms(r window).  Uses (and replaces) the
// context's matrix.
int hpccg_cnn_make_dataset(hpccg_ctx *c, int32_t n, int32_t npairs, int32_t crop, uint64_t seed)
{
    HPCCG_CHECK_ARG(c && n >= 8 && npairs >= 1 && crop >= 9 && crop <= n, "bad argument");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    HPCCG_TRY(ensure_net(c));
    CnnNet *N = net_of(c);
    TrainData &D = N->data;
    dfree(D.r);
    dfree(D.e);
    dfree(D.margin);
    D = TrainData{};
    const int64_t V = (int64_t)crop * crop * crop;
    if (dmalloc(&D.r, npairs * V * 4) != hipSuccess || dmalloc(&D.e, npairs * V * 4) != hipSuccess ||
        dmalloc(&D.margin, npairs * 6 * sizeof(int)) != hipSuccess) {
        set_error("make_dataset: allocation failed");
        return HPCCG_ERR_NOMEM;
    }
    D.crop = crop;
    HPCCG_TRY(hpccg_generate_matrix(c, n, n, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    const int64_t NN = (int64_t)n * n * n;
    double *xs = nullptr, *b = nullptr, *x = nullptr;
    if (dmalloc(&xs, NN * 8) != hipSuccess || dmalloc(&b, NN * 8) != hipSuccess ||
        dmalloc(&x, NN * 8) != hipSuccess) {
        dfree(xs);
        dfree(b);
        set_error("make_dataset: allocation failed");
        return HPCCG_ERR_NOMEM;
    }
    std::vector<double> xh(NN), zero(NN, 0.0);
    std::vector<int> margins;
    uint64_t rs = seed ^ 0xA5A5A5A5DEADBEEFULL;
    auto rnd = [&]() {
        rs += 0x9E3779B97F4A7C15ULL;
        uint64_t z = rs;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    int st = 0;
    for (int run = 0; D.count < npairs && st == 0; ++run) {
        hpccg_xstar_sines(n, n, n, seed + 1000003ULL * (uint64_t)run, xh.data());
        if (hipMemcpy(xs, xh.data(), NN * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(x, zero.data(), NN * 8, hipMemcpyHostToDevice) != hipSuccess) {
            st = HPCCG_ERR_HIP;
            break;
        }
        if ((st = launch_spmv(c, xs, b)) != 0) break;
        double bb = 0.0;
        if ((st = hpccg_ddot_dev(c, (int)NN, b, b, &bb)) != 0) break;
        const int mode = c->dot_mode;
        c->dot_mode = HPCCG_DOT_TREE;
        st = cg_trace(c, b, x, 400, 1e-10 * std::sqrt(bb), [&](int k) {
            (void)k;
            if (D.count >= npairs) return 1;
            // window origin per axis: against the low domain face, against the high one, or
            // uniform, one third each.  Rows on domain faces, edges and corners have 18, 12 or 8
            // entries; uniform windows would touch them rarely (2 in 49 per axis at 64^3, crop 16)
            // and leave the network untrained -- and its preconditioner indefinite -- there.
            auto origin = [&]() {
                const uint64_t u = rnd() % 3u;
                return u == 0 ? 0 : u == 1 ? n - crop : (int)(rnd() % (uint64_t)(n - crop + 1));
            };
            const int ox = origin(), oy = origin(), oz = origin();
            hipLaunchKernelGGL(k_crop, dim3(1), dim3(256), 0, c->stream, c->r, x, c->p, xs, c->state, n, crop, ox,
                               oy, oz, D.r + D.count * V, D.e + D.count * V);
            // loss-mask margins: 4 on crop faces inside the domain, 0 on domain faces
            const int mg = 4;
            margins.insert(margins.end(), {ox == 0 ? 0 : mg, ox + crop == n ? 0 : mg, oy == 0 ? 0 : mg,
                                           oy + crop == n ? 0 : mg, oz == 0 ? 0 : mg, oz + crop == n ? 0 : mg});
            ++D.count;
            return 0;
        });
        c->dot_mode = mode;
    }
    if (st == 0 && hipMemcpy(D.margin, margins.data(), margins.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        st = HPCCG_ERR_HIP;
    if (st == 0 && hipStreamSynchronize(c->stream) != hipSuccess) st = HPCCG_ERR_HIP;
    dfree(xs);
    dfree(b);
    dfree(x);
    if (st) {
        if (st == HPCCG_ERR_HIP) set_error("make_dataset: HIP failure");
        return st;
    }
    return 0;
}

// Copy pair i of the training set to host (r/s and e/s, crop^3 each) -- for inspection/tests.
int hpccg_cnn_dataset_pair(hpccg_ctx *c, int32_t i, float *r, float *e, int32_t *margins6)
{
    HPCCG_CHECK_ARG(c && r && e, "null argument");
    CnnNet *N = net_of(c);
    if (!N || i < 0 || i >= N->data.count) {
        set_error("dataset index out of range");
        return HPCCG_ERR_ARG;
    }
    const int64_t V = (int64_t)N->data.crop * N->data.crop * N->data.crop;
    HPCCG_HIP_TRY(hipMemcpy(r, N->data.r + i * V, V * 4, hipMemcpyDeviceToHost));
    HPCCG_HIP_TRY(hipMemcpy(e, N->data.e + i * V, V * 4, hipMemcpyDeviceToHost));
    if (margins6) HPCCG_HIP_TRY(hipMemcpy(margins6, N->data.margin + i * 6, 6 * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

// `steps` Adam steps on batches of `batch` random (reflected, sign-flipped) crops; one step is
// one captured graph (gather, forward, loss, backward, Adam, repack) replayed `steps` times.
// loss_hist[steps] (may be NULL) receives each step's batch-mean loss.
int hpccg_cnn_train(hpccg_ctx *c, int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss_hist,
                    double *elapsed_ms)
{
    HPCCG_CHECK_ARG(c && steps >= 1 && batch >= 1 && batch <= 4096, "bad argument");
    CnnNet *N = net_of(c);
    if (!N || N->data.count == 0) {
        set_error("no training set: call hpccg_cnn_make_dataset");
        return HPCCG_ERR_STATE;
    }
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    const int C = N->data.crop;
    const Grid3 g = make_grid(C, C, C);
    const int64_t V = g.nvox();
    const int slabs = std::max(1, C / 4);
    HPCCG_TRY(ensure_act(c, V * batch, wgrad_wg(c, g, batch)));
    if (N->loss_cap < steps) {
        dfree(N->loss);
        N->loss = nullptr;
        HPCCG_HIP_TRY(dmalloc(&N->loss, (size_t)steps * 4));
        N->loss_cap = steps;
    }
    dfree(N->lpart);
    dfree(N->batch_meta);
    HPCCG_HIP_TRY(dmalloc(&N->lpart, (size_t)2 * batch * slabs * 8));
    HPCCG_HIP_TRY(dmalloc(&N->batch_meta, (size_t)batch * 8 * sizeof(int)));
    int z0 = 0;
    HPCCG_HIP_TRY(hipMemcpy(N->step, &z0, sizeof(int), hipMemcpyHostToDevice));
    std::vector<float> zero(NWALL, 0.f);
    HPCCG_HIP_TRY(hipMemcpy(N->m, zero.data(), NWALL * 4, hipMemcpyHostToDevice));
    HPCCG_HIP_TRY(hipMemcpy(N->v, zero.data(), NWALL * 4, hipMemcpyHostToDevice));

    if (N->train_graph) hipGraphExecDestroy(N->train_graph);
    N->train_graph = nullptr;
    hipGraph_t gr;
    HPCCG_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_gather, dim3(16, batch), dim3(256), 0, c->stream, N->data.r, N->data.e, N->data.margin,
                       N->data.count, C, N->in, N->tgt, N->batch_meta, N->step, seed);
    forward(c, N->in, nullptr, N->z, nullptr, g, batch);
    hipLaunchKernelGGL(k_loss_a, dim3(batch * slabs), dim3(256), 0, c->stream, N->z, N->tgt, N->wbuf, N->batch_meta,
                       C, slabs, N->lpart);
    hipLaunchKernelGGL(k_loss_b, dim3(batch * slabs), dim3(256), 0, c->stream, N->wbuf, N->dz, C, slabs, batch,
                       N->lpart, N->loss, N->step);
    hipLaunchKernelGGL(k_round_input, dim3(c->info.num_cu * 4), dim3(256), 0, c->stream, N->in, nullptr,
                       (uint16_t *)N->wbuf, V, batch);
    backward(c, N->in, nullptr, g, batch);
    hipLaunchKernelGGL(k_adam, dim3((NWALL + 255) / 256), dim3(256), 0, c->stream, N->W, N->m, N->v, N->grad, NWALL,
                       lr, steps, N->step);
    hipLaunchKernelGGL(k_pack, dim3((NSTEP16 * 512 + 255) / 256), dim3(256), 0, c->stream, N->W,
                       (uint16_t *)N->pk_f1, (uint16_t *)N->pk_f2, (uint16_t *)N->pk_f3f,
                       (uint16_t *)N->pk_b3, (uint16_t *)N->pk_b2);
    hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, c->stream, N->step);
    const hipError_t ec = hipStreamEndCapture(c->stream, &gr);
    HPCCG_HIP_TRY(ec);
    const hipError_t ei = hipGraphInstantiate(&N->train_graph, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    HPCCG_HIP_TRY(ei);
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < steps; ++s) HPCCG_HIP_TRY(hipGraphLaunch(N->train_graph, c->stream));
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    if (elapsed_ms) {
        float ms = 0.f;
        HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *elapsed_ms = ms;
    }
    if (loss_hist) HPCCG_HIP_TRY(hipMemcpy(loss_hist, N->loss, (size_t)steps * 4, hipMemcpyDeviceToHost));
    return 0;
}

// z = net(r) on device arrays of one nx*ny*nz field, with the positive-homogeneity scaling
// z = rms(r) * net(r / rms(r)) done inside the kernels (scale = 1/rms(r) given by the caller).
int hpccg_cnn_apply_dev(hpccg_ctx *c, int32_t nx, int32_t ny, int32_t nz, const float *r, float inv_scale, float *z)
{
    HPCCG_CHECK_ARG(c && r && z && inv_scale > 0.f && nx >= 1 && ny >= 1 && nz >= 1, "bad argument");
    CnnNet *N = net_of(c);
    if (!N) {
        set_error("no network");
        return HPCCG_ERR_STATE;
    }
    const Grid3 g = make_grid(nx, ny, nz);
    HPCCG_TRY(ensure_act(c, g.nvox(), 1));
    const float sc[2] = {inv_scale, 1.0f / inv_scale};
    HPCCG_HIP_TRY(hipMemcpyAsync(N->scale_in, &sc[0], 4, hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(N->scale_out, &sc[1], 4, hipMemcpyHostToDevice, c->stream));
    forward_fused(c, r, N->scale_in, z, N->scale_out, g);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int hpccg_time_cnn_apply(hpccg_ctx *c, int32_t nx, int32_t ny, int32_t nz, const float *r, float *z, int32_t reps,
                         double *avg_ms)
{
    HPCCG_CHECK_ARG(c && r && z && reps > 0 && avg_ms && nx >= 1 && ny >= 1 && nz >= 1, "bad argument");
    CnnNet *N = net_of(c);
    if (!N) {
        set_error("no network");
        return HPCCG_ERR_STATE;
    }
    const Grid3 g = make_grid(nx, ny, nz);
    HPCCG_TRY(ensure_act(c, g.nvox(), 1));
    const float sc[2] = {1.0f, 1.0f};
    HPCCG_HIP_TRY(hipMemcpyAsync(N->scale_in, &sc[0], 4, hipMemcpyHostToDevice, c->stream));
    HPCCG_HIP_TRY(hipMemcpyAsync(N->scale_out, &sc[1], 4, hipMemcpyHostToDevice, c->stream));
    forward_fused(c, r, N->scale_in, z, N->scale_out, g);  // warm
    HPCCG_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; ++i) forward_fused(c, r, N->scale_in, z, N->scale_out, g);
    HPCCG_HIP_TRY(hipGetLastError());
    HPCCG_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / reps;
    return 0;
}

// Flexible CG preconditioned by the network (mixed precision as hpccg_cg_mixed_dev: FP16 values,
// FP32 r/p/Ap/z, FP64 dots and x, true-residual replacement every replace_every iterations and
// a TRUE-residual stopping test).  Every `window` iterations the host compares ||r|| with its
// value `window` iterations earlier; a reduction below `min_factor` (or a breakdown) hands the
// current x to plain FP64 CG (tree dots, hpccg_cg_dev semantics) for the rest of the solve.
// grid (nx,ny,nz) must be the matrix's grid (nx*ny*nz == nrow).  Outputs: fcg_iters, cg_iters
// (0 unless it fell back), fell_back, true_normr = ||b - A x|| at exit (FP64).
int hpccg_fcg_learned_dev(hpccg_ctx *c, int32_t nx, int32_t ny, int32_t nz, const double *b, double *x,
                          int32_t max_iter, double tolerance, int32_t replace_every, int32_t window,
                          double min_factor, int32_t *fcg_iters, int32_t *cg_iters, int32_t *fell_back,
                          double *true_normr, double *hist_rec, double *elapsed_ms)
{
    HPCCG_CHECK_ARG(c && b && x && max_iter >= 0 && window >= 1, "bad argument");
    HPCCG_HIP_TRY(hipSetDevice(c->device));
    if (!c->A.rowindex) {
        set_error("no matrix");
        return HPCCG_ERR_STATE;
    }
    HPCCG_CHECK_ARG((int64_t)nx * ny * nz == c->A.nrow, "grid does not match the matrix rows");
    CnnNet *Nt = net_of(c);
    if (!Nt) {
        set_error("no network: call hpccg_cnn_init / hpccg_cnn_set_weights / hpccg_cnn_train");
        return HPCCG_ERR_STATE;
    }
    HPCCG_TRY(ensure_sell16(c));
    const int n = c->A.nrow;
    const Grid3 g = make_grid(nx, ny, nz);
    HPCCG_TRY(ensure_act(c, g.nvox(), 1));
    if (!fcg_of(c)) c->ext[2] = new FcgWs();
    FcgWs *F = fcg_of(c);
    if (F->gdims[0] != nx || F->gdims[1] != ny || F->gdims[2] != nz) {  // the grid is baked into the graphs
        invalidate_learned_graphs(c);
        F->gdims[0] = nx;
        F->gdims[1] = ny;
        F->gdims[2] = nz;
    }
    if (F->n < n) {
        dfree(F->r);
        dfree(F->p);
        dfree(F->Ap);
        dfree(F->z);
        F->n = 0;
        if (dmalloc(&F->r, n * 4) != hipSuccess || dmalloc(&F->p, n * 4) != hipSuccess ||
            dmalloc(&F->Ap, n * 4) != hipSuccess || dmalloc(&F->z, n * 4) != hipSuccess) {
            set_error("fcg: allocation failed");
            return HPCCG_ERR_NOMEM;
        }
        F->n = n;
    }
    if (!F->st) {
        HPCCG_HIP_TRY(dmalloc(&F->st, sizeof(FcgState)));
        HPCCG_HIP_TRY(hipHostMalloc(&F->st_h, sizeof(FcgState), hipHostMallocDefault));
    }
    if (F->hist_cap < max_iter + 2) {
        dfree(F->hist);
        F->hist = nullptr;
        HPCCG_HIP_TRY(dmalloc(&F->hist, 2 * (size_t)(max_iter + 2) * 8));
        F->hist_cap = max_iter + 2;
    }
    FcgState init{};
    init.tol = tolerance;
    init.stop_at = max_iter + 1;
    init.max_iter = max_iter;
    init.replace_every = replace_every;
    *F->st_h = init;
    HPCCG_HIP_TRY(hipMemcpyAsync(F->st, F->st_h, sizeof(FcgState), hipMemcpyHostToDevice, c->stream));
    std::vector<double> neg(2 * (size_t)F->hist_cap, -1.0);
    HPCCG_HIP_TRY(hipMemcpyAsync(F->hist, neg.data(), neg.size() * 8, hipMemcpyHostToDevice, c->stream));
    double *hrec = F->hist, *htrue = F->hist + F->hist_cap;
    const Matrix &A = c->A;
    const int Gs = sell_grid(A), Gt = true_blocks(A.nslices), Gv = vec_grid_for(c->info.num_cu, n);
    const int *skip = &F->st->done;
    const float *s_in = &F->st->s_in, *s_out = &F->st->s_out;

    auto iteration = [&]() {
        hipLaunchKernelGGL(kf_spmv, dim3(Gs), dim3(kSellThreads), 0, c->stream, sell16(A), F->p, F->Ap, F->st,
                           A.red_part);
        hipLaunchKernelGGL(kf_spmv_finish, dim3(1), dim3(kFinishThreads), 0, c->stream, A.red_part, Gs, F->st);
        hipLaunchKernelGGL(kf_update_xr, dim3(Gv), dim3(kVecThreads), 0, c->stream, x, F->r, F->p, F->Ap, n, F->st,
                           c->partials);
        hipLaunchKernelGGL(kf_xr_finish, dim3(1), dim3(kFinishThreads), 0, c->stream, c->partials, Gv, n, F->st,
                           hrec);
        hipLaunchKernelGGL(kf_true<false>, dim3(Gt), dim3(kTrueThreads), 0, c->stream, sell16(A), x, b, F->r, F->st,
                           A.red_part);
        hipLaunchKernelGGL(kf_true_finish<false>, dim3(1), dim3(kFinishThreads), 0, c->stream, A.red_part, Gt, n,
                           F->st, htrue);
        forward_fused(c, F->r, s_in, F->z, s_out, g, skip);
        hipLaunchKernelGGL(kf_dots<false>, dim3(Gv), dim3(kVecThreads), 0, c->stream, F->z, F->Ap, F->r, n, F->st,
                           c->partials);
        hipLaunchKernelGGL(kf_update_p<false>, dim3(Gv), dim3(kVecThreads), 0, c->stream, F->z, F->p, n, F->st,
                           c->partials, Gv);
    };

    // graphs of 1, 2, 4, ... iterations, launched in chunks that never cross a stagnation
    // check point (every `window` iterations) and are sized from the residual rate
    auto get = [&](int size, hipGraphExec_t *out) -> int {
        return cached_graph(F->graphs, c->stream, GraphKey{b, x, 200, size}, [&]() {
            for (int i = 0; i < size; ++i) iteration();
            return 0;
        }, out);
    };
    HPCCG_TRY(capture_iteration_graphs(window, max_iter, get));

    HPCCG_HIP_TRY(hipEventRecord(c->ev2, c->stream));
    // Prepare: r = b - A x, z = M(r), rz = z.r, p = z
    // x = 0 (the usual start) needs no A x: the flag is set otherwise
    hipLaunchKernelGGL(k_nonzero<>, dim3(Gv), dim3(kVecThreads), 0, c->stream, x, (int64_t)n, &F->st->x_nonzero);
    hipLaunchKernelGGL(kf_true<true>, dim3(Gt), dim3(kTrueThreads), 0, c->stream, sell16(A), x, b, F->r, F->st,
                       A.red_part);
    hipLaunchKernelGGL(kf_true_finish<true>, dim3(1), dim3(kFinishThreads), 0, c->stream, A.red_part, Gt, n, F->st,
                       htrue);
    forward_fused(c, F->r, s_in, F->z, s_out, g, skip);
    hipLaunchKernelGGL(kf_dots<true>, dim3(Gv), dim3(kVecThreads), 0, c->stream, F->z, F->Ap, F->r, n, F->st,
                       c->partials);
    hipLaunchKernelGGL(kf_update_p<true>, dim3(Gv), dim3(kVecThreads), 0, c->stream, F->z, F->p, n, F->st,
                       c->partials, Gv);
    HPCCG_HIP_TRY(hipGetLastError());

    bool fallback = false;
    FcgState &S = *F->st_h;
    HPCCG_HIP_TRY(hipMemcpyAsync(F->st_h, F->st, sizeof(FcgState), hipMemcpyDeviceToHost, c->stream));
    HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
    double check_rr = S.rr;  // r.r at the last stagnation check point
    int since = 0;           // iterations run since that point
    if (S.breakdown) fallback = !S.converged;
    ChunkPlanner plan{window, tolerance, 0};
    int m = std::min(window, max_iter), launched = 0;
    while (!fallback && !S.done && max_iter > 0) {
        const int k_before = S.k;
        HPCCG_TRY(launch_iterations(c->stream, m, get));
        launched += m;
        HPCCG_HIP_TRY(hipMemcpyAsync(F->st_h, F->st, sizeof(FcgState), hipMemcpyDeviceToHost, c->stream));
        HPCCG_HIP_TRY(hipStreamSynchronize(c->stream));
        if (S.breakdown) {
            fallback = true;
            break;
        }
        if (S.done) break;
        since += S.k - k_before;
        if (since >= window) {
            // stagnation: ||r|| shrank by less than min_factor over the last `window` iterations
            if (!(std::sqrt(S.rr) * min_factor <= std::sqrt(check_rr))) {
                fallback = true;
                break;
            }
            check_rr = S.rr;
            since = 0;
        }
        if (launched > 4 * max_iter + 64) break;
        m = plan.next(S.rr, S.k - 1, std::min(window - since, std::max(1, max_iter - (S.k - 1))));
    }
    const int fcg_done = S.k - 1 < 0 ? 0 : (S.breakdown ? S.stop_at - 1 : S.k - 1);
    int cg_n = 0;
    double tn = std::sqrt(S.rr_true);
    if (fallback) {
        // plain FP64 CG from the current x for the remaining budget
        double nr = 0.0;
        const int rem = std::max(0, max_iter - fcg_done);
        const int mode = c->dot_mode;
        c->dot_mode = HPCCG_DOT_TREE;
        const int st = cg_device_fp64(c, b, x, rem, tolerance, &cg_n, &nr, nullptr, nullptr);
        c->dot_mode = mode;
        if (st < 0) return st;
        tn = -1.0;  // the caller measures the FP64 true residual of the final x
    }
    HPCCG_HIP_TRY(hipEventRecord(c->ev3, c->stream));
    HPCCG_HIP_TRY(hipEventSynchronize(c->ev3));
    if (elapsed_ms) {
        float ms = 0.f;
        HPCCG_HIP_TRY(hipEventElapsedTime(&ms, c->ev2, c->ev3));
        *elapsed_ms = ms;
    }
    if (fcg_iters) *fcg_iters = fcg_done;
    if (cg_iters) *cg_iters = cg_n;
    if (fell_back) *fell_back = fallback ? 1 : 0;
    if (true_normr) *true_normr = tn;
    if (hist_rec) {
        std::vector<double> h(F->hist_cap);
        HPCCG_HIP_TRY(hipMemcpy(h.data(), hrec, F->hist_cap * 8, hipMemcpyDeviceToHost));
        for (int j = 0; j <= max_iter; ++j) hist_rec[j] = h[j] != -1.0 ? std::sqrt(h[j]) : -1.0;  // -1: not run
    }
    return 0;
}

}  // extern "C"
