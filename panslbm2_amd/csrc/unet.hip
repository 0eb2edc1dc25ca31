// panslbm2_amd/csrc/unet.hip -- learned warm start for the steady solves.
//
// A 3-level 3D U-Net (channels c, 2c, 4c with c = 16; two 3x3x3 convolutions + ReLU per level,
// 2x2x2 average pooling, nearest upsampling with skip concatenation [upsampled, skip], a 1x1x1
// linear head) maps the design and the boundary masks to the fields that the InitialCondition
// calls take.  It runs and trains on the GPU:
//
//   * convolutions are implicit GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16):
//     bfloat16 weights and activations, FP32 accumulation.  Activations are channels-last
//     ([voxel][channel], voxel = x + nx (y + ny z)), so the 8 consecutive K values one lane feeds
//     the matrix core (8 channels of one tap) are one 16-byte load;
//   * the input gradient of a convolution is the same kernel run on the output gradient with the
//     weights transposed and the taps mirrored;
//   * the weight gradient reduces over voxels: 32-voxel tiles of the output gradient and of the
//     shifted input are staged through LDS in [channel][voxel] order, so the matrix-core operands
//     are again 16-byte reads; partial sums per voxel range are reduced in a fixed order, so the
//     gradient is deterministic;
//   * Adam keeps FP32 master weights and repacks the bfloat16 copies after each update.
//
// The fields the solver uses are FP64 arrays; the network reads and writes them directly
// (normalised by a per-channel mean and scale, optionally relative to base fields, so that it can
// learn the change from a previous solution).  The network has its own HIP stream: its training
// kernels (matrix-core bound) overlap the solver's memory-bound steps on the context stream; the
// calls that read or write solver fields are ordered against the context stream with events.
// The CPU restatement of the same network is oracle/unet_oracle.py.  No reference counterpart:
// the reference starts every solve from rest.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/panslbm2_amd.h"
#include "context.h"

namespace {

typedef uint16_t bf16_t;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;

constexpr int kMaxFields = 16;
constexpr int kMaxCg = 64;   // widest output gradient a weight-gradient launch stages
constexpr int kMaxCx = 96;   // widest input it stages
constexpr int kTileV = 64;   // voxels per LDS tile (two matrix-core K steps of 32)

__host__ __device__ inline bf16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__device__ inline bf16x8 ld8(const bf16_t *p) { return __builtin_bit_cast(bf16x8, *(const uint4 *)p); }

__host__ __device__ inline int tap_dx(int t, int taps) { return taps == 27 ? t % 3 - 1 : 0; }
__host__ __device__ inline int tap_dy(int t, int taps) { return taps == 27 ? (t / 3) % 3 - 1 : 0; }
__host__ __device__ inline int tap_dz(int t, int taps) { return taps == 27 ? t / 9 - 1 : 0; }

// ---- convolution (forward, and input gradient with transposed weights) ----------------------------
// y[v][o] = bias[o] + sum_k w[o][k] x[v + off(tap)][ci], k = tap cin + ci.  One wave computes
// 16 MT output channels x 64 voxels (4 tiles of 16); a workgroup is 4 waves along the voxels.
// The matrix core sums A[m][k] B[k][n] over k: A = weights (row m = lane % 16), B = shifted input
// (column n = lane % 16); both operands give slot j of lane group q the same k = 8q + j, so the
// sum is the same whatever order the hardware pairs slots in.  D[m = 4q + j][n = lane % 16].
template <int MT>
__global__ void __launch_bounds__(256) k_conv(const bf16_t *__restrict__ x, int cin, int taps,
                                              const bf16_t *__restrict__ w, int kpad,
                                              const float *__restrict__ bias, bf16_t *__restrict__ y,
                                              float *__restrict__ y32, int cstore, int creal, int relu, int nx,
                                              int ny, int nz) {
    constexpr int NT = 4;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
    const int V = nx * ny * nz;
    const int v0 = (blockIdx.x * 4 + wv) * (16 * NT);
    const int co0 = blockIdx.y * 16 * MT;
    int px[NT], py[NT], pz[NT];
    bool pin[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int v = v0 + nt * 16 + r;
        pin[nt] = v < V;
        const int vv = pin[nt] ? v : 0;
        px[nt] = vv % nx;
        py[nt] = (vv / nx) % ny;
        pz[nt] = vv / (nx * ny);
    }
    f4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
    const int K = taps * cin;
    for (int kc = 0; kc < kpad; kc += 32) {
        const int k = kc + 8 * q;
        bf16x8 a[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) a[mt] = ld8(w + (size_t)(co0 + mt * 16 + r) * kpad + k);
        const int tap = k / cin, ci = k - tap * cin;
        const int dx = tap_dx(tap, taps), dy = tap_dy(tap, taps), dz = tap_dz(tap, taps);
        const bool kin = k < K;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int xx = px[nt] + dx, yy = py[nt] + dy, zz = pz[nt] + dz;
            const bool ok = kin && pin[nt] && (unsigned)xx < (unsigned)nx && (unsigned)yy < (unsigned)ny &&
                            (unsigned)zz < (unsigned)nz;
            bf16x8 b = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
            if (ok) b = ld8(x + (size_t)(xx + nx * (yy + ny * zz)) * cin + ci);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b, acc[mt][nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int v = v0 + nt * 16 + r;
        if (v >= V) continue;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = co0 + mt * 16 + 4 * q + j;
                if (co >= cstore) continue;
                float val = acc[mt][nt][j] + (bias ? bias[co] : 0.f);
                if (relu) val = fmaxf(val, 0.f);
                if (co >= creal) val = 0.f;  // padded channels stay zero
                if (y) y[(size_t)v * cstore + co] = f2bf(val);
                if (y32) y32[(size_t)v * cstore + co] = val;
            }
    }
}

// ---- weight gradient ----------------------------------------------------------------------------
// part[split][tap][o][i] = sum over the split's voxels v of g[v][o] x[v + off(tap)][i].
// Matrix core: A[m = o][k = v], B[k = v][n = i], two K steps of 32 voxels per LDS tile.
// The tap-0 workgroups also sum the staged g tiles: bpart[split][o] = sum over the split of g[v][o].
__global__ void __launch_bounds__(256) k_conv_wgrad(const bf16_t *__restrict__ g, int cg,
                                                    const bf16_t *__restrict__ x, int cx, int taps, int nx, int ny,
                                                    int nz, int chunk, float *__restrict__ part,
                                                    float *__restrict__ bpart) {
    __shared__ __align__(16) bf16_t sg[kMaxCg][kTileV + 8];
    __shared__ __align__(16) bf16_t sx[kMaxCx][kTileV + 8];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, r = lane & 15, q = lane >> 4;
    const int split = blockIdx.x, tap = blockIdx.y;
    const int V = nx * ny * nz;
    const int dx = tap_dx(tap, taps), dy = tap_dy(tap, taps), dz = tap_dz(tap, taps);
    const int vbeg = split * chunk, vend = min(V, vbeg + chunk);
    const int cxp = (cx + 15) & ~15;
    const int nmt = cg / 16, nnt = cxp / 16, tiles = nmt * nnt;
    f4 acc[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) acc[s] = f4{0.f, 0.f, 0.f, 0.f};
    const int gu = cg / 8, xu = cxp / 8;
    // bias: in tap-0 workgroups thread (o, quarter) = (tid / 4, tid % 4) sums a quarter of row o
    const bool bias = tap == 0 && tid < 4 * cg;
    const int bo = tid >> 2, bq = tid & 3;
    float bsum = 0.f;
    for (int vb = vbeg; vb < vend; vb += kTileV) {
        for (int u = tid; u < kTileV * gu; u += 256) {
            const int vl = u / gu, c8 = (u - vl * gu) * 8, v = vb + vl;
            uint4 d = {0u, 0u, 0u, 0u};
            if (v < vend) d = *(const uint4 *)(g + (size_t)v * cg + c8);
            const bf16_t *e = (const bf16_t *)&d;
#pragma unroll
            for (int j = 0; j < 8; ++j) sg[c8 + j][vl] = e[j];
        }
        for (int u = tid; u < kTileV * xu; u += 256) {
            const int vl = u / xu, c8 = (u - vl * xu) * 8, v = vb + vl;
            uint4 d = {0u, 0u, 0u, 0u};
            if (v < vend && c8 < cx) {
                const int xx = v % nx + dx, yy = (v / nx) % ny + dy, zz = v / (nx * ny) + dz;
                if ((unsigned)xx < (unsigned)nx && (unsigned)yy < (unsigned)ny && (unsigned)zz < (unsigned)nz)
                    d = *(const uint4 *)(x + (size_t)(xx + nx * (yy + ny * zz)) * cx + c8);
            }
            const bf16_t *e = (const bf16_t *)&d;
#pragma unroll
            for (int j = 0; j < 8; ++j) sx[c8 + j][vl] = e[j];
        }
        __syncthreads();
        if (bias)
#pragma unroll
            for (int vl = 0; vl < kTileV / 4; ++vl) bsum += bf2f(sg[bo][bq * (kTileV / 4) + vl]);
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const int t = wv + 4 * s;
            if (t < tiles) {
                const int mt = t / nnt, nt = t - mt * nnt;
#pragma unroll
                for (int kh = 0; kh < kTileV / 32; ++kh) {
                    const bf16x8 a = __builtin_bit_cast(bf16x8, *(const uint4 *)&sg[mt * 16 + r][32 * kh + 8 * q]);
                    const bf16x8 b = __builtin_bit_cast(bf16x8, *(const uint4 *)&sx[nt * 16 + r][32 * kh + 8 * q]);
                    acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[s], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float *out = part + ((size_t)split * taps + tap) * cg * cxp;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int t = wv + 4 * s;
        if (t < tiles) {
            const int mt = t / nnt, nt = t - mt * nnt;
#pragma unroll
            for (int j = 0; j < 4; ++j) out[(size_t)(mt * 16 + 4 * q + j) * cxp + nt * 16 + r] = acc[s][j];
        }
    }
    if (tap == 0) {  // uniform per workgroup
        __shared__ float bred[4 * kMaxCg];
        if (bias) bred[tid] = bsum;
        __syncthreads();
        if (tid < cg) bpart[(size_t)split * cg + tid] = (bred[4 * tid] + bred[4 * tid + 1]) + (bred[4 * tid + 2] + bred[4 * tid + 3]);
    }
}

// grad[(o taps + tap) ci + i] = sum over splits (in order) of part[split][tap][o][i]; the bias
// gradient follows the weights: grad[co taps ci + o] = sum over splits of bpart[split][o]
__global__ void k_wgrad_reduce(const float *__restrict__ part, const float *__restrict__ bpart, int nsplit, int taps,
                               int cg, int cxp, int co, int ci, float *__restrict__ grad) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = co * taps * ci;
    if (n >= nw + co) return;
    float s = 0.f;
    if (n < nw) {
        const int i = n % ci, t = (n / ci) % taps, o = n / (ci * taps);
        for (int p = 0; p < nsplit; ++p) s += part[(((size_t)p * taps + t) * cg + o) * cxp + i];
    } else {
        for (int p = 0; p < nsplit; ++p) s += bpart[(size_t)p * cg + (n - nw)];
    }
    grad[n] = s;
}

// ---- weight packing -----------------------------------------------------------------------------
// forward: wp[o][tap cis + i] = W[o][tap][i] (W is [co][taps][ci] FP32), zero elsewhere
__global__ void k_pack_fwd(const float *__restrict__ W, int co, int taps, int ci, int cis, int kpad, int rows,
                           bf16_t *__restrict__ wp) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= rows * kpad) return;
    const int o = n / kpad, k = n - o * kpad, t = k / cis, i = k - t * cis;
    float v = 0.f;
    if (o < co && t < taps && i < ci) v = W[((size_t)o * taps + t) * ci + i];
    wp[n] = f2bf(v);
}
// input gradient: wt[i][tap cos + o] = W[o][taps - 1 - tap][i] (mirrored taps, transposed channels)
__global__ void k_pack_bwd(const float *__restrict__ W, int co, int taps, int ci, int cos, int kpad, int rows,
                           bf16_t *__restrict__ wt) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= rows * kpad) return;
    const int i = n / kpad, k = n - i * kpad, t = k / cos, o = k - t * cos;
    float v = 0.f;
    if (i < ci && t < taps && o < co) v = W[((size_t)o * taps + (taps - 1 - t)) * ci + i];
    wt[n] = f2bf(v);
}
__global__ void k_pack_bias(const float *__restrict__ b, int co, int rows, float *__restrict__ bp) {
    const int o = threadIdx.x;
    if (o < rows) bp[o] = o < co ? b[o] : 0.f;
}

// ---- elementwise pieces ---------------------------------------------------------------------------
struct FieldPtrs {
    const double *f[kMaxFields];
    const double *base[kMaxFields];  // null: no base
    double mean[kMaxFields], scale[kMaxFields];
};
struct FieldOut {
    double *f[kMaxFields];
    const double *base[kMaxFields];
    double mean[kMaxFields], scale[kMaxFields], lo[kMaxFields], hi[kMaxFields];
};

// xs[v][c] = (f_c[v] - base_c[v] - mean_c) / scale_c for c < nf, zero up to cs; bf16 or FP32
__global__ void k_gather(FieldPtrs F, int nf, int cs, int V, bf16_t *__restrict__ xb, float *__restrict__ xf) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= V * cs) return;
    const int v = n / cs, c = n - v * cs;
    float val = 0.f;
    if (c < nf) {
        const double b = F.base[c] ? F.base[c][v] : 0.0;
        val = (float)((F.f[c][v] - b - F.mean[c]) / F.scale[c]);
    }
    if (xb) xb[n] = f2bf(val);
    if (xf) xf[n] = val;
}
// f_c[v] = base_c[v] + mean_c + scale_c y[v][c], clamped to [lo_c, hi_c]
__global__ void k_scatter(const float *__restrict__ y, int cs, int nf, int V, FieldOut F) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= V * nf) return;
    const int c = n / V, v = n - c * V;
    const double b = F.base[c] ? F.base[c][v] : 0.0;
    const double f = b + F.mean[c] + F.scale[c] * (double)y[(size_t)v * cs + c];
    F.f[c][v] = fmin(fmax(f, F.lo[c]), F.hi[c]);
}

// 2x2x2 average pooling, fine grid (nx, ny, nz)
__global__ void k_pool(const bf16_t *__restrict__ x, int C, int nx, int ny, int nz, bf16_t *__restrict__ out) {
    const int mx = nx / 2, my = ny / 2, mz = nz / 2;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= mx * my * mz * C) return;
    const int c = n % C, vc = n / C;
    const int i = vc % mx, j = (vc / mx) % my, k = vc / (mx * my);
    float s = 0.f;
    for (int e = 0; e < 2; ++e)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) s += bf2f(x[((size_t)(2 * i + a) + nx * ((2 * j + b) + ny * (2 * k + e))) * C + c]);
    out[n] = f2bf(s * 0.125f);
}
// out[v] = [coarse[parent(v)] (cu channels), skip[v] (cs channels)]
__global__ void k_upcat(const bf16_t *__restrict__ coarse, int cu, const bf16_t *__restrict__ skip, int cs, int nx,
                        int ny, int nz, bf16_t *__restrict__ out) {
    const int C = cu + cs;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nx * ny * nz * C) return;
    const int c = n % C, v = n / C;
    if (c < cu) {
        const int i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
        out[n] = coarse[((size_t)(i / 2) + (nx / 2) * ((j / 2) + (ny / 2) * (k / 2))) * cu + c];
    } else {
        out[n] = skip[(size_t)v * cs + (c - cu)];
    }
}
// g[v][c] = bf16(2 (y - t) / (V nf)) for c < nf, zero to cs; lpart[block] = sum of (y - t)^2
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ y, const float *__restrict__ t, int cs, int nf,
                                              int V, bf16_t *__restrict__ g, float *__restrict__ lpart) {
    __shared__ float red[256];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    float sq = 0.f;
    if (n < V * cs) {
        const int c = n % cs;
        float d = 0.f;
        if (c < nf) d = y[n] - t[n];
        sq = d * d;
        g[n] = f2bf(2.f * d / ((float)V * (float)nf));
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}
// one workgroup of 256: strided partial sums, then a tree (fixed order, deterministic)
__global__ void __launch_bounds__(256) k_loss_final(const float *__restrict__ lpart, int n, double inv,
                                                    double *__restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += lpart[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * inv;
}
// out[v][c] = act[v][c] > 0 ? bf16(d[v][c]) : 0   (d has stride cd)
__global__ void k_relu_bwd(const float *__restrict__ d, int cd, const bf16_t *__restrict__ act, int C, int V,
                           bf16_t *__restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= V * C) return;
    const int v = n / C, c = n - v * C;
    out[n] = bf2f(act[n]) > 0.f ? f2bf(d[(size_t)v * cd + c]) : (bf16_t)0;
}
// coarse grid: out[vc][c] = act > 0 ? bf16(sum over the 8 children of d[v][c]) : 0  (fine dims given)
__global__ void k_up_bwd(const float *__restrict__ d, int cd, const bf16_t *__restrict__ act, int C, int nx, int ny,
                         int nz, bf16_t *__restrict__ out) {
    const int mx = nx / 2, my = ny / 2, mz = nz / 2;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= mx * my * mz * C) return;
    const int c = n % C, vc = n / C;
    const int i = vc % mx, j = (vc / mx) % my, k = vc / (mx * my);
    float s = 0.f;
    for (int e = 0; e < 2; ++e)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) s += d[((size_t)(2 * i + a) + nx * ((2 * j + b) + ny * (2 * k + e))) * cd + c];
    out[n] = bf2f(act[n]) > 0.f ? f2bf(s) : (bf16_t)0;
}
// fine grid: out[v][c] = act > 0 ? bf16(skip[v][off + c] + dc[parent(v)][c] / 8) : 0
__global__ void k_pool_bwd(const float *__restrict__ dc, const float *__restrict__ skip, int cskip, int off,
                           const bf16_t *__restrict__ act, int C, int nx, int ny, int nz, bf16_t *__restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nx * ny * nz * C) return;
    const int c = n % C, v = n / C;
    const int i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
    const size_t p = (size_t)(i / 2) + (nx / 2) * ((j / 2) + (ny / 2) * (k / 2));
    const float s = skip[(size_t)v * cskip + off + c] + dc[p * C + c] * 0.125f;
    out[n] = bf2f(act[n]) > 0.f ? f2bf(s) : (bf16_t)0;
}
// Adam (Kingma & Ba 2015, Algorithm 1); bc1 = 1 - beta1^t, bc2 = 1 - beta2^t
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, size_t n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}
__global__ void k_bf_to_f64(const bf16_t *__restrict__ a, int cs, int cr, int V, double *__restrict__ b) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= V * cr) return;
    const int v = n / cr, c = n - v * cr;
    b[n] = bf2f(a[(size_t)v * cs + c]);
}
__global__ void k_f_to_f64(const float *__restrict__ a, int cs, int cr, int V, double *__restrict__ b) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= V * cr) return;
    const int v = n / cr, c = n - v * cr;
    b[n] = a[(size_t)v * cs + c];
}

inline unsigned blocks(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }
inline int r16(int c) { return (c + 15) & ~15; }
inline int r32(int c) { return (c + 31) & ~31; }

}  // namespace

// ==================================================================================================
struct lbm_unet {
    lbm_context *ctx = nullptr;
    hipStream_t s = nullptr;                          // the network's own stream
    hipEvent_t ev_in = nullptr, ev_out = nullptr;     // ordering against the context stream
    int dims[3][3] = {};  // per level: nx, ny, nz
    int V[3] = {};
    int cin = 0, cout = 0, base = 16;
    struct Conv {
        std::string name;
        int ci, co, taps, cis, level;  // real in/out channels, taps, stored input channels, grid level
        size_t wo, bo;                 // offsets of W and b in the parameter vector
        int kpad, kpadT, rowsT;        // packed K (forward), packed K and rows (input gradient)
        bf16_t *wp = nullptr, *wt = nullptr;
        float *bp = nullptr;
    };
    std::vector<Conv> conv;
    size_t np = 0;
    int adam_t = 0;
    std::vector<void *> allocs;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    // activations
    bf16_t *x0, *e1a, *e1b, *p1, *e2a, *e2b, *p2, *e3a, *e3b, *cat2, *d2a, *d2b, *cat1, *d1a, *d1b;
    float *y32, *tgt;
    // backward scratch
    bf16_t *gb0, *gb1;
    float *dx1, *dx2, *dxt, *wpart, *bpart, *lpart;
    double *dloss, *hloss;
    size_t wpart_n = 0, bpart_n = 0;
    bool have_target = false;
};

namespace {

int hipfail(hipError_t e, const char *what) {
    return lbm::set_error(LBM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}
#define UCHK(expr)                                                                                     \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return hipfail(e_, #expr);                                               \
    } while (0)

template <class T>
int alloc(lbm_unet *N, T **p, size_t n) {
    void *q = nullptr;
    UCHK(hipMalloc(&q, n * sizeof(T) + 16));
    N->allocs.push_back(q);
    *p = (T *)q;
    return LBM_OK;
}

int launch_conv(lbm_unet *N, const bf16_t *x, int cin, int taps, const bf16_t *w, int kpad, const float *bias,
                bf16_t *y, float *y32, int cstore, int creal, int relu, const int *d) {
    const int V = d[0] * d[1] * d[2];
    const int tiles = cstore > 0 ? r16(cstore) / 16 : 0;
    int MT = tiles % 4 == 0 ? 4 : (tiles % 3 == 0 ? 3 : (tiles % 2 == 0 ? 2 : 1));
    if (cin % 8 || kpad % 32 || tiles <= 0) return lbm::set_error(LBM_ERR_ARG, "conv shape: cin %d kpad %d", cin, kpad);
    dim3 grid(blocks(V, 256), tiles / MT);
    switch (MT) {
    case 4: k_conv<4><<<grid, 256, 0, N->s>>>(x, cin, taps, w, kpad, bias, y, y32, cstore, creal, relu, d[0], d[1], d[2]); break;
    case 3: k_conv<3><<<grid, 256, 0, N->s>>>(x, cin, taps, w, kpad, bias, y, y32, cstore, creal, relu, d[0], d[1], d[2]); break;
    case 2: k_conv<2><<<grid, 256, 0, N->s>>>(x, cin, taps, w, kpad, bias, y, y32, cstore, creal, relu, d[0], d[1], d[2]); break;
    default: k_conv<1><<<grid, 256, 0, N->s>>>(x, cin, taps, w, kpad, bias, y, y32, cstore, creal, relu, d[0], d[1], d[2]); break;
    }
    UCHK(hipGetLastError());
    return LBM_OK;
}

// splits of a voxel range for the weight gradient: about 1024 workgroups over all taps
void wsplit(int V, int taps, int *nsplit, int *chunk) {
    int want = (1024 + taps - 1) / taps;
    int maxs = (V + kTileV - 1) / kTileV;
    int ns = want < maxs ? want : maxs;
    if (ns < 1) ns = 1;
    int ch = ((V + ns - 1) / ns + kTileV - 1) / kTileV * kTileV;
    *chunk = ch;
    *nsplit = (V + ch - 1) / ch;
}

// gradient of conv c: g is the output gradient (bf16, r16(co) channels), x its input (cis channels)
int conv_grads(lbm_unet *N, const lbm_unet::Conv &c, const bf16_t *g, const bf16_t *x) {
    const int *d = N->dims[c.level];
    const int V = N->V[c.level], cg = r16(c.co), cxp = r16(c.cis);
    if (cg > kMaxCg || c.cis > kMaxCx) return lbm::set_error(LBM_ERR_ARG, "weight gradient: %d x %d channels", cg, c.cis);
    int nsplit, chunk;
    wsplit(V, c.taps, &nsplit, &chunk);
    if ((size_t)nsplit * c.taps * cg * cxp > N->wpart_n) return lbm::set_error(LBM_ERR_ARG, "weight partials overflow");
    if ((size_t)nsplit * cg > N->bpart_n) return lbm::set_error(LBM_ERR_ARG, "bias partials overflow");
    k_conv_wgrad<<<dim3(nsplit, c.taps), 256, 0, N->s>>>(g, cg, x, c.cis, c.taps, d[0], d[1], d[2], chunk, N->wpart,
                                                         N->bpart);
    UCHK(hipGetLastError());
    const int nw = c.co * c.taps * c.ci;  // the bias follows the weights in the parameter vector
    k_wgrad_reduce<<<blocks(nw + c.co), 256, 0, N->s>>>(N->wpart, N->bpart, nsplit, c.taps, cg, cxp, c.co, c.ci,
                                                         N->g + c.wo);
    UCHK(hipGetLastError());
    return LBM_OK;
}

int conv_fwd(lbm_unet *N, int i, const bf16_t *x, bf16_t *y, float *y32, int relu) {
    const auto &c = N->conv[i];
    return launch_conv(N, x, c.cis, c.taps, c.wp, c.kpad, c.bp, y, y32, r16(c.co), c.co, relu, N->dims[c.level]);
}
// input gradient of conv i into dx (FP32, cis channels)
int conv_dx(lbm_unet *N, int i, const bf16_t *g, float *dx) {
    const auto &c = N->conv[i];
    return launch_conv(N, g, r16(c.co), c.taps, c.wt, c.kpadT, nullptr, nullptr, dx, c.cis, c.cis, 0, N->dims[c.level]);
}

int repack(lbm_unet *N) {
    for (auto &c : N->conv) {
        const int rows = r16(c.co);
        k_pack_fwd<<<blocks((size_t)rows * c.kpad), 256, 0, N->s>>>(N->p + c.wo, c.co, c.taps, c.ci, c.cis, c.kpad,
                                                                   rows, c.wp);
        k_pack_bwd<<<blocks((size_t)c.rowsT * c.kpadT), 256, 0, N->s>>>(N->p + c.wo, c.co, c.taps, c.ci, r16(c.co),
                                                                        c.kpadT, c.rowsT, c.wt);
        k_pack_bias<<<1, 64, 0, N->s>>>(N->p + c.bo, c.co, rows, c.bp);
    }
    UCHK(hipGetLastError());
    return LBM_OK;
}

enum { E1A, E1B, E2A, E2B, E3A, E3B, D2A, D2B, D1A, D1B, HEAD };

int forward(lbm_unet *N) {
    const int c1 = N->base, c2 = 2 * c1, c3 = 4 * c1;
    int rc;
#define F(call)                                                                                        \
    if ((rc = (call)) != LBM_OK) return rc
    F(conv_fwd(N, E1A, N->x0, N->e1a, nullptr, 1));
    F(conv_fwd(N, E1B, N->e1a, N->e1b, nullptr, 1));
    const int *d1 = N->dims[0], *d2 = N->dims[1];
    k_pool<<<blocks((size_t)N->V[1] * c1), 256, 0, N->s>>>(N->e1b, c1, d1[0], d1[1], d1[2], N->p1);
    F(conv_fwd(N, E2A, N->p1, N->e2a, nullptr, 1));
    F(conv_fwd(N, E2B, N->e2a, N->e2b, nullptr, 1));
    k_pool<<<blocks((size_t)N->V[2] * c2), 256, 0, N->s>>>(N->e2b, c2, d2[0], d2[1], d2[2], N->p2);
    F(conv_fwd(N, E3A, N->p2, N->e3a, nullptr, 1));
    F(conv_fwd(N, E3B, N->e3a, N->e3b, nullptr, 1));
    k_upcat<<<blocks((size_t)N->V[1] * (c3 + c2)), 256, 0, N->s>>>(N->e3b, c3, N->e2b, c2, d2[0], d2[1], d2[2], N->cat2);
    F(conv_fwd(N, D2A, N->cat2, N->d2a, nullptr, 1));
    F(conv_fwd(N, D2B, N->d2a, N->d2b, nullptr, 1));
    k_upcat<<<blocks((size_t)N->V[0] * (c2 + c1)), 256, 0, N->s>>>(N->d2b, c2, N->e1b, c1, d1[0], d1[1], d1[2], N->cat1);
    F(conv_fwd(N, D1A, N->cat1, N->d1a, nullptr, 1));
    F(conv_fwd(N, D1B, N->d1a, N->d1b, nullptr, 1));
    F(conv_fwd(N, HEAD, N->d1b, nullptr, N->y32, 0));
    UCHK(hipGetLastError());
    return LBM_OK;
}

int backward(lbm_unet *N) {
    const int c1 = N->base, c2 = 2 * c1, c3 = 4 * c1, co = r16(N->cout);
    const int *d1 = N->dims[0], *d2 = N->dims[1];
    const int V1 = N->V[0], V2 = N->V[1], V3 = N->V[2];
    int rc;
    const int nl = (int)blocks((size_t)V1 * co);
    k_loss<<<nl, 256, 0, N->s>>>(N->y32, N->tgt, co, N->cout, V1, N->gb0, N->lpart);
    k_loss_final<<<1, 256, 0, N->s>>>(N->lpart, nl, 1.0 / ((double)V1 * N->cout), N->dloss);
    // head and level-1 decoder
    F(conv_grads(N, N->conv[HEAD], N->gb0, N->d1b));
    F(conv_dx(N, HEAD, N->gb0, N->dxt));
    k_relu_bwd<<<blocks((size_t)V1 * c1), 256, 0, N->s>>>(N->dxt, c1, N->d1b, c1, V1, N->gb1);
    F(conv_grads(N, N->conv[D1B], N->gb1, N->d1a));
    F(conv_dx(N, D1B, N->gb1, N->dxt));
    k_relu_bwd<<<blocks((size_t)V1 * c1), 256, 0, N->s>>>(N->dxt, c1, N->d1a, c1, V1, N->gb0);
    F(conv_grads(N, N->conv[D1A], N->gb0, N->cat1));
    F(conv_dx(N, D1A, N->gb0, N->dx1));  // [V1][c2 + c1]: upsampled part, then the skip of e1b
    k_up_bwd<<<blocks((size_t)V2 * c2), 256, 0, N->s>>>(N->dx1, c2 + c1, N->d2b, c2, d1[0], d1[1], d1[2], N->gb1);
    // level-2 decoder
    F(conv_grads(N, N->conv[D2B], N->gb1, N->d2a));
    F(conv_dx(N, D2B, N->gb1, N->dxt));
    k_relu_bwd<<<blocks((size_t)V2 * c2), 256, 0, N->s>>>(N->dxt, c2, N->d2a, c2, V2, N->gb0);
    F(conv_grads(N, N->conv[D2A], N->gb0, N->cat2));
    F(conv_dx(N, D2A, N->gb0, N->dx2));  // [V2][c3 + c2]
    k_up_bwd<<<blocks((size_t)V3 * c3), 256, 0, N->s>>>(N->dx2, c3 + c2, N->e3b, c3, d2[0], d2[1], d2[2], N->gb1);
    // bottleneck
    F(conv_grads(N, N->conv[E3B], N->gb1, N->e3a));
    F(conv_dx(N, E3B, N->gb1, N->dxt));
    k_relu_bwd<<<blocks((size_t)V3 * c3), 256, 0, N->s>>>(N->dxt, c3, N->e3a, c3, V3, N->gb0);
    F(conv_grads(N, N->conv[E3A], N->gb0, N->p2));
    F(conv_dx(N, E3A, N->gb0, N->dxt));  // [V3][c2]
    k_pool_bwd<<<blocks((size_t)V2 * c2), 256, 0, N->s>>>(N->dxt, N->dx2, c3 + c2, c3, N->e2b, c2, d2[0], d2[1], d2[2],
                                                         N->gb1);
    // level-2 encoder
    F(conv_grads(N, N->conv[E2B], N->gb1, N->e2a));
    F(conv_dx(N, E2B, N->gb1, N->dxt));
    k_relu_bwd<<<blocks((size_t)V2 * c2), 256, 0, N->s>>>(N->dxt, c2, N->e2a, c2, V2, N->gb0);
    F(conv_grads(N, N->conv[E2A], N->gb0, N->p1));
    F(conv_dx(N, E2A, N->gb0, N->dxt));  // [V2][c1]
    k_pool_bwd<<<blocks((size_t)V1 * c1), 256, 0, N->s>>>(N->dxt, N->dx1, c2 + c1, c2, N->e1b, c1, d1[0], d1[1], d1[2],
                                                         N->gb1);
    // level-1 encoder
    F(conv_grads(N, N->conv[E1B], N->gb1, N->e1a));
    F(conv_dx(N, E1B, N->gb1, N->dxt));
    k_relu_bwd<<<blocks((size_t)V1 * c1), 256, 0, N->s>>>(N->dxt, c1, N->e1a, c1, V1, N->gb0);
    F(conv_grads(N, N->conv[E1A], N->gb0, N->x0));
#undef F
    UCHK(hipGetLastError());
    return LBM_OK;
}

int adam(lbm_unet *N, double lr, double b1, double b2, double eps) {
    N->adam_t += 1;
    const double bc1 = 1.0 - std::pow(b1, N->adam_t), bc2 = 1.0 - std::pow(b2, N->adam_t);
    k_adam<<<blocks(N->np), 256, 0, N->s>>>(N->p, N->g, N->m, N->v, N->np, (float)lr, (float)b1, (float)b2,
                                            (float)eps, (float)bc1, (float)bc2);
    UCHK(hipGetLastError());
    return repack(N);
}

int fields_in(const double *const *f, const double *const *base, const double *mean, const double *scale, int nf,
              FieldPtrs *F) {
    if (nf < 1 || nf > kMaxFields || !f) return lbm::set_error(LBM_ERR_ARG, "1..%d fields expected", kMaxFields);
    for (int c = 0; c < nf; ++c) {
        if (!f[c]) return lbm::set_error(LBM_ERR_ARG, "field %d is null", c);
        F->f[c] = f[c];
        F->base[c] = base ? base[c] : nullptr;
        F->mean[c] = mean ? mean[c] : 0.0;
        F->scale[c] = scale ? scale[c] : 1.0;
        if (F->scale[c] == 0.0) return lbm::set_error(LBM_ERR_ARG, "field %d has scale 0", c);
    }
    return LBM_OK;
}

void destroy(lbm_unet *N) {
    if (!N) return;
    (void)hipSetDevice(lbm::context_device(N->ctx));
    if (N->s) (void)hipStreamSynchronize(N->s);
    for (void *q : N->allocs) (void)hipFree(q);
    if (N->hloss) (void)hipHostFree(N->hloss);
    if (N->ev_in) (void)hipEventDestroy(N->ev_in);
    if (N->ev_out) (void)hipEventDestroy(N->ev_out);
    if (N->s) (void)hipStreamDestroy(N->s);
    delete N;
}

// the network stream waits for what the context stream has queued so far
int after_ctx(lbm_unet *N) {
    UCHK(hipEventRecord(N->ev_in, lbm::context_stream(N->ctx)));
    UCHK(hipStreamWaitEvent(N->s, N->ev_in, 0));
    return LBM_OK;
}
// the context stream waits for what the network stream has queued so far
int before_ctx(lbm_unet *N) {
    UCHK(hipEventRecord(N->ev_out, N->s));
    UCHK(hipStreamWaitEvent(lbm::context_stream(N->ctx), N->ev_out, 0));
    return LBM_OK;
}

// xorshift64* stream for the initial weights
struct Rng {
    uint64_t s;
    double uniform() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return (double)((s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
    }
};

}  // namespace

extern "C" {

int lbm_unet_create(lbm_context *ctx, int nx, int ny, int nz, int cin, int cout, int base, unsigned long long seed,
                    lbm_unet **out) {
    if (!ctx || !out) return lbm::set_error(LBM_ERR_ARG, "null argument");
    *out = nullptr;
    if (nx < 4 || ny < 4 || nz < 4 || nx % 4 || ny % 4 || nz % 4)
        return lbm::set_error(LBM_ERR_ARG, "grid %dx%dx%d: each side a multiple of 4 (two poolings)", nx, ny, nz);
    if (cin < 1 || cin > 8) return lbm::set_error(LBM_ERR_ARG, "cin %d: 1..8 input fields", cin);
    if (cout < 1 || cout > 16) return lbm::set_error(LBM_ERR_ARG, "cout %d: 1..16 output fields", cout);
    if (base != 16) return lbm::set_error(LBM_ERR_UNSUPPORTED, "base %d: only 16 (channels 16/32/64)", base);
    if ((size_t)nx * ny * nz > (size_t)1 << 28) return lbm::set_error(LBM_ERR_ARG, "grid too large");
    UCHK(hipSetDevice(lbm::context_device(ctx)));
    lbm_unet *N = new lbm_unet;
    N->ctx = ctx;
    if (hipStreamCreateWithFlags(&N->s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&N->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&N->ev_out, hipEventDisableTiming) != hipSuccess) {
        destroy(N);
        return lbm::set_error(LBM_ERR_HIP, "stream or event creation failed");
    }
    N->cin = cin;
    N->cout = cout;
    N->base = base;
    for (int l = 0; l < 3; ++l) {
        N->dims[l][0] = nx >> l;
        N->dims[l][1] = ny >> l;
        N->dims[l][2] = nz >> l;
        N->V[l] = N->dims[l][0] * N->dims[l][1] * N->dims[l][2];
    }
    const int c1 = base, c2 = 2 * base, c3 = 4 * base;
    struct L {
        const char *name;
        int ci, co, taps, cis, level;
    } layers[] = {
        {"e1a", cin, c1, 27, 8, 0},        {"e1b", c1, c1, 27, c1, 0},      {"e2a", c1, c2, 27, c1, 1},
        {"e2b", c2, c2, 27, c2, 1},        {"e3a", c2, c3, 27, c2, 2},      {"e3b", c3, c3, 27, c3, 2},
        {"d2a", c3 + c2, c2, 27, c3 + c2, 1}, {"d2b", c2, c2, 27, c2, 1},   {"d1a", c2 + c1, c1, 27, c2 + c1, 0},
        {"d1b", c1, c1, 27, c1, 0},        {"head", c1, cout, 1, c1, 0},
    };
    size_t off = 0;
    for (const auto &l : layers) {
        lbm_unet::Conv c;
        c.name = l.name;
        c.ci = l.ci;
        c.co = l.co;
        c.taps = l.taps;
        c.cis = l.cis;
        c.level = l.level;
        c.wo = off;
        c.bo = off + (size_t)l.co * l.taps * l.ci;
        off = c.bo + l.co;
        c.kpad = r32(l.taps * l.cis);
        c.kpadT = r32(l.taps * r16(l.co));
        c.rowsT = r16(l.cis);
        N->conv.push_back(c);
    }
    N->np = off;
    const size_t V1 = N->V[0], V2 = N->V[1], V3 = N->V[2];
    int rc = LBM_OK;
#define A(p, n)                                                                                        \
    if (rc == LBM_OK) rc = alloc(N, &(p), (n))
    A(N->p, N->np);
    A(N->g, N->np);
    A(N->m, N->np);
    A(N->v, N->np);
    for (auto &c : N->conv) {
        A(c.wp, (size_t)r16(c.co) * c.kpad);
        A(c.wt, (size_t)c.rowsT * c.kpadT);
        A(c.bp, (size_t)r16(c.co));
    }
    A(N->x0, V1 * 8);
    A(N->e1a, V1 * c1);
    A(N->e1b, V1 * c1);
    A(N->p1, V2 * c1);
    A(N->e2a, V2 * c2);
    A(N->e2b, V2 * c2);
    A(N->p2, V3 * c2);
    A(N->e3a, V3 * c3);
    A(N->e3b, V3 * c3);
    A(N->cat2, V2 * (c3 + c2));
    A(N->d2a, V2 * c2);
    A(N->d2b, V2 * c2);
    A(N->cat1, V1 * (c2 + c1));
    A(N->d1a, V1 * c1);
    A(N->d1b, V1 * c1);
    A(N->y32, V1 * r16(cout));
    A(N->tgt, V1 * r16(cout));
    size_t gmax = std::max(std::max(V1 * c1, V2 * c2), V3 * c3);
    gmax = std::max(gmax, V1 * (size_t)r16(cout));
    A(N->gb0, gmax);
    A(N->gb1, gmax);
    A(N->dx1, V1 * (c2 + c1));
    A(N->dx2, V2 * (c3 + c2));
    A(N->dxt, gmax);
    size_t wpn = 0;
    for (auto &c : N->conv) {
        int ns, ch;
        wsplit(N->V[c.level], c.taps, &ns, &ch);
        wpn = std::max(wpn, (size_t)ns * c.taps * r16(c.co) * r16(c.cis));
    }
    N->wpart_n = wpn;
    A(N->wpart, wpn);
    size_t bpn = 0;
    for (auto &c : N->conv) {
        int ns, ch;
        wsplit(N->V[c.level], c.taps, &ns, &ch);
        bpn = std::max(bpn, (size_t)ns * r16(c.co));
    }
    N->bpart_n = bpn;
    A(N->bpart, N->bpart_n);
    A(N->lpart, blocks(V1 * r16(cout)));
    A(N->dloss, 1);
#undef A
    if (rc == LBM_OK && hipHostMalloc(&N->hloss, sizeof(double), hipHostMallocDefault) != hipSuccess)
        rc = lbm::set_error(LBM_ERR_HIP, "hipHostMalloc failed");
    if (rc != LBM_OK) {
        destroy(N);
        return rc;
    }
    // He-uniform initial weights (bound sqrt(6 / fan_in)), zero biases, and a zero linear head:
    // the untrained network outputs zero (the base fields, when it predicts a change)
    std::vector<float> h(N->np, 0.f);
    Rng rng{seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull};
    for (auto &c : N->conv) {
        if (c.name == "head") continue;
        const double b = std::sqrt(6.0 / ((double)c.taps * c.ci));
        for (size_t i = c.wo; i < c.bo; ++i) h[i] = (float)((2.0 * rng.uniform() - 1.0) * b);
    }
    hipError_t e = hipMemcpyAsync(N->p, h.data(), N->np * sizeof(float), hipMemcpyHostToDevice, N->s);
    if (e == hipSuccess) e = hipMemsetAsync(N->m, 0, N->np * sizeof(float), N->s);
    if (e == hipSuccess) e = hipMemsetAsync(N->v, 0, N->np * sizeof(float), N->s);
    if (e == hipSuccess) e = hipMemsetAsync(N->g, 0, N->np * sizeof(float), N->s);
    if (e == hipSuccess) e = hipMemsetAsync(N->x0, 0, V1 * 8 * sizeof(bf16_t), N->s);
    if (e == hipSuccess) e = hipMemsetAsync(N->tgt, 0, V1 * r16(cout) * sizeof(float), N->s);
    if (e == hipSuccess) e = hipStreamSynchronize(N->s);
    if (e != hipSuccess) {
        destroy(N);
        return hipfail(e, "initial weights");
    }
    if ((rc = repack(N)) != LBM_OK) {
        destroy(N);
        return rc;
    }
    *out = N;
    return LBM_OK;
}

int lbm_unet_destroy(lbm_unet *net) {
    destroy(net);
    return LBM_OK;
}

int lbm_unet_num_params(lbm_unet *net, size_t *n) {
    if (!net || !n) return lbm::set_error(LBM_ERR_ARG, "null argument");
    *n = net->np;
    return LBM_OK;
}

static int vec_io(lbm_unet *N, float *dev, double *host_out, const double *host_in, size_t n) {
    if (!N || n != N->np) return lbm::set_error(LBM_ERR_ARG, "expected %zu parameters", N ? N->np : (size_t)0);
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    std::vector<float> h(n);
    if (host_in) {
        for (size_t i = 0; i < n; ++i) h[i] = (float)host_in[i];
        UCHK(hipMemcpyAsync(dev, h.data(), n * sizeof(float), hipMemcpyHostToDevice, N->s));
        UCHK(hipStreamSynchronize(N->s));
        return repack(N);
    }
    UCHK(hipMemcpyAsync(h.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, N->s));
    UCHK(hipStreamSynchronize(N->s));
    for (size_t i = 0; i < n; ++i) host_out[i] = h[i];
    return LBM_OK;
}
int lbm_unet_get_params(lbm_unet *net, double *host, size_t n) { return vec_io(net, net ? net->p : nullptr, host, nullptr, n); }
int lbm_unet_set_params(lbm_unet *net, const double *host, size_t n) {
    if (!host) return lbm::set_error(LBM_ERR_ARG, "null argument");
    return vec_io(net, net ? net->p : nullptr, nullptr, host, n);
}
int lbm_unet_get_grads(lbm_unet *net, double *host, size_t n) { return vec_io(net, net ? net->g : nullptr, host, nullptr, n); }

int lbm_unet_reset_optimizer(lbm_unet *N) {
    if (!N) return lbm::set_error(LBM_ERR_ARG, "null argument");
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    UCHK(hipMemsetAsync(N->m, 0, N->np * sizeof(float), N->s));
    UCHK(hipMemsetAsync(N->v, 0, N->np * sizeof(float), N->s));
    N->adam_t = 0;
    return LBM_OK;
}

int lbm_unet_set_input(lbm_unet *N, const double *const *fields, const double *mean, const double *scale, int nfields) {
    if (!N) return lbm::set_error(LBM_ERR_ARG, "null argument");
    if (nfields != N->cin) return lbm::set_error(LBM_ERR_ARG, "%d input fields, network takes %d", nfields, N->cin);
    FieldPtrs F;
    int rc = fields_in(fields, nullptr, mean, scale, nfields, &F);
    if (rc != LBM_OK) return rc;
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    if ((rc = after_ctx(N)) != LBM_OK) return rc;
    k_gather<<<blocks((size_t)N->V[0] * 8), 256, 0, N->s>>>(F, nfields, 8, N->V[0], N->x0, nullptr);
    UCHK(hipGetLastError());
    return LBM_OK;
}

int lbm_unet_set_target(lbm_unet *N, const double *const *fields, const double *const *base, const double *mean,
                        const double *scale, int nfields) {
    if (!N) return lbm::set_error(LBM_ERR_ARG, "null argument");
    if (nfields != N->cout) return lbm::set_error(LBM_ERR_ARG, "%d target fields, network gives %d", nfields, N->cout);
    FieldPtrs F;
    int rc = fields_in(fields, base, mean, scale, nfields, &F);
    if (rc != LBM_OK) return rc;
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    if ((rc = after_ctx(N)) != LBM_OK) return rc;
    const int cs = r16(N->cout);
    k_gather<<<blocks((size_t)N->V[0] * cs), 256, 0, N->s>>>(F, nfields, cs, N->V[0], nullptr, N->tgt);
    UCHK(hipGetLastError());
    N->have_target = true;
    return LBM_OK;
}

int lbm_unet_forward(lbm_unet *N) {
    if (!N) return lbm::set_error(LBM_ERR_ARG, "null argument");
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    return forward(N);
}

int lbm_unet_output(lbm_unet *N, double *const *fields, const double *const *base, const double *mean,
                    const double *scale, const double *lo, const double *hi, int nfields) {
    if (!N || !fields) return lbm::set_error(LBM_ERR_ARG, "null argument");
    if (nfields != N->cout) return lbm::set_error(LBM_ERR_ARG, "%d output fields, network gives %d", nfields, N->cout);
    FieldOut F;
    for (int c = 0; c < nfields; ++c) {
        if (!fields[c]) return lbm::set_error(LBM_ERR_ARG, "field %d is null", c);
        F.f[c] = fields[c];
        F.base[c] = base ? base[c] : nullptr;
        F.mean[c] = mean ? mean[c] : 0.0;
        F.scale[c] = scale ? scale[c] : 1.0;
        F.lo[c] = lo ? lo[c] : -HUGE_VAL;
        F.hi[c] = hi ? hi[c] : HUGE_VAL;
        if (!(F.lo[c] <= F.hi[c])) return lbm::set_error(LBM_ERR_ARG, "field %d: empty clamp range", c);
    }
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    int rc = after_ctx(N);
    if (rc != LBM_OK) return rc;
    k_scatter<<<blocks((size_t)N->V[0] * nfields), 256, 0, N->s>>>(N->y32, r16(N->cout), nfields, N->V[0], F);
    UCHK(hipGetLastError());
    return before_ctx(N);
}

int lbm_unet_backward(lbm_unet *N, double *loss) {
    if (!N) return lbm::set_error(LBM_ERR_ARG, "null argument");
    if (!N->have_target) return lbm::set_error(LBM_ERR_ARG, "no target set");
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    int rc = backward(N);
    if (rc != LBM_OK) return rc;
    if (loss) {
        UCHK(hipMemcpyAsync(N->hloss, N->dloss, sizeof(double), hipMemcpyDeviceToHost, N->s));
        UCHK(hipStreamSynchronize(N->s));
        *loss = *N->hloss;
    }
    return LBM_OK;
}

int lbm_unet_adam(lbm_unet *N, double lr, double beta1, double beta2, double eps) {
    if (!N) return lbm::set_error(LBM_ERR_ARG, "null argument");
    if (!(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
        return lbm::set_error(LBM_ERR_ARG, "Adam hyper-parameters out of range");
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    return adam(N, lr, beta1, beta2, eps);
}

int lbm_unet_train(lbm_unet *N, int steps, double lr, double *loss) {
    if (!N || steps < 0) return lbm::set_error(LBM_ERR_ARG, "bad argument");
    if (!N->have_target) return lbm::set_error(LBM_ERR_ARG, "no target set");
    if (!(lr > 0)) return lbm::set_error(LBM_ERR_ARG, "lr must be positive");
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    int rc;
    for (int s = 0; s < steps; ++s) {
        if ((rc = forward(N)) != LBM_OK) return rc;
        if ((rc = backward(N)) != LBM_OK) return rc;
        if ((rc = adam(N, lr, 0.9, 0.999, 1e-8)) != LBM_OK) return rc;
    }
    if (loss) {
        // loss of the last step's forward pass (before its update)
        UCHK(hipMemcpyAsync(N->hloss, N->dloss, sizeof(double), hipMemcpyDeviceToHost, N->s));
        UCHK(hipStreamSynchronize(N->s));
        *loss = steps ? *N->hloss : NAN;
    }
    return LBM_OK;
}

int lbm_unet_synchronize(lbm_unet *N) {
    if (!N) return lbm::set_error(LBM_ERR_ARG, "null argument");
    UCHK(hipSetDevice(lbm::context_device(N->ctx)));
    UCHK(hipStreamSynchronize(N->s));
    return LBM_OK;
}

int lbm_unet_read(lbm_unet *N, const char *name, double *host, size_t n) {
    if (!N || !name || !host) return lbm::set_error(LBM_ERR_ARG, "null argument");
    const int c1 = N->base, c2 = 2 * c1, c3 = 4 * c1;
    struct T {
        const char *name;
        const void *p;
        int level, cs, cr, f32;
    } tab[] = {
        {"x0", N->x0, 0, 8, N->cin, 0},   {"e1a", N->e1a, 0, c1, c1, 0},   {"e1b", N->e1b, 0, c1, c1, 0},
        {"p1", N->p1, 1, c1, c1, 0},      {"e2a", N->e2a, 1, c2, c2, 0},   {"e2b", N->e2b, 1, c2, c2, 0},
        {"p2", N->p2, 2, c2, c2, 0},      {"e3a", N->e3a, 2, c3, c3, 0},   {"e3b", N->e3b, 2, c3, c3, 0},
        {"cat2", N->cat2, 1, c3 + c2, c3 + c2, 0}, {"d2a", N->d2a, 1, c2, c2, 0}, {"d2b", N->d2b, 1, c2, c2, 0},
        {"cat1", N->cat1, 0, c2 + c1, c2 + c1, 0}, {"d1a", N->d1a, 0, c1, c1, 0}, {"d1b", N->d1b, 0, c1, c1, 0},
        {"y", N->y32, 0, r16(N->cout), N->cout, 1}, {"target", N->tgt, 0, r16(N->cout), N->cout, 1},
    };
    for (const auto &t : tab) {
        if (strcmp(t.name, name)) continue;
        const int V = N->V[t.level];
        const size_t cnt = (size_t)V * t.cr;
        if (n != cnt) return lbm::set_error(LBM_ERR_ARG, "%s holds %zu values, %zu requested", name, cnt, n);
        UCHK(hipSetDevice(lbm::context_device(N->ctx)));
        double *d = nullptr;
        UCHK(hipMalloc(&d, cnt * sizeof(double)));
        if (t.f32)
            k_f_to_f64<<<blocks(cnt), 256, 0, N->s>>>((const float *)t.p, t.cs, t.cr, V, d);
        else
            k_bf_to_f64<<<blocks(cnt), 256, 0, N->s>>>((const bf16_t *)t.p, t.cs, t.cr, V, d);
        hipError_t e = hipMemcpyAsync(host, d, cnt * sizeof(double), hipMemcpyDeviceToHost, N->s);
        if (e == hipSuccess) e = hipStreamSynchronize(N->s);
        (void)hipFree(d);
        if (e != hipSuccess) return hipfail(e, "read");
        return LBM_OK;
    }
    return lbm::set_error(LBM_ERR_ARG, "unknown tensor %s", name);
}

}  // extern "C"
