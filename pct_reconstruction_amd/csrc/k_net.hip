// k_net.hip — learned warm start: a 4-level 3D U-Net mapping (FBP image, hull mask) to the DROP
// initial iterate x0, with its training (forward, backward, Adam) in hand-written gfx950 kernels.
//
// The reference has no learned component; its x0 is FBP, the hull, or their product
// (define_initial_iterate, DS.cu:8612-8646).  The net predicts a correction y to the FBP image and
// x0 = hull * (FBP + y), so a zero head reproduces the reference's hybrid x0.  DROP still defines the
// final image, and pct_net_apply keeps the FBP start when the net's x0 has the larger data residual.
//
// Network (channels 16-32-64-128, 3x3x3 convolutions, "same" zero padding):
//   enc1  conv(in8 -> 16)                       r0      in8 = (FBP, hull, 6 zero channels)
//   enc2  conv(avgpool(enc1) -> 32)             r0/2
//   enc3  conv(avgpool(enc2) -> 64)             r0/4
//   enc4  conv(avgpool(enc3) -> 128)            r0/8
//   dec3  conv(concat(up(enc4), enc3) -> 64)    r0/4    up = nearest x2, fused into the conv loader
//   dec2  conv(concat(up(dec3), enc2) -> 32)    r0/2
//   dec1  conv(concat(up(dec2), enc1) -> 16)    r0
//   head  1x1x1 conv(dec1 -> 1), linear
// ReLU after every 3x3x3 convolution.  Loss: mean over hull voxels of (x0 - target)^2.
//
// Numerics: BF16 weights and activations (round to nearest even), FP32 accumulation on MFMA
// (v_mfma_f32_16x16x32_bf16), FP32 master weights, gradients of weights in FP32, Adam in FP32.
// Layout: NDHWC (channels innermost) so a lane's 8 consecutive input channels are one 16-byte load.
// Convolution as implicit GEMM: M = voxels, N = output channels, K = 27 * input channels
// (k = tap * Cin + c, tap = (dz+1)*9 + (dy+1)*3 + (dx+1)), padded to a multiple of 64.
// Backward data = the same forward kernel over the masked output gradient with the taps flipped and
// the channel roles swapped.  Weight gradients: LDS-staged MFMA over 32-voxel slices, split over
// voxel chunks into FP32 partials that a second kernel sums in chunk order (deterministic).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kBlock = 256;
constexpr int kLayers = 7;  // 3x3x3 convolutions; the head is separate

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint4 ld16(const uint16_t* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Input of a 3x3x3 convolution: channels [0, ca) from tensor a (at half resolution when a_coarse:
// nearest-neighbour upsampling), channels [ca, ca + cb) from tensor b at full resolution.
struct ConvIn {
    const uint16_t* a;
    const uint16_t* b;
    int ca, cb, a_coarse;
};

// 8 consecutive input channels [c, c+8) of voxel (z, y, x) of a D x H x W grid (zeros outside)
__device__ __forceinline__ uint4 load_in8(const ConvIn& in, int D, int H, int W, int z, int y, int x, int c)
{
    uint4 r = make_uint4(0, 0, 0, 0);
    if (z < 0 || y < 0 || x < 0 || z >= D || y >= H || x >= W) return r;
    if (c < in.ca) {
        if (in.a_coarse) {
            const int h2 = H >> 1, w2 = W >> 1;
            r = ld16(in.a + ((((int64_t)(z >> 1) * h2 + (y >> 1)) * w2 + (x >> 1)) * in.ca + c));
        } else {
            r = ld16(in.a + (((int64_t)z * H + y) * W + x) * in.ca + c);
        }
    } else {
        r = ld16(in.b + (((int64_t)z * H + y) * W + x) * in.cb + (c - in.ca));
    }
    return r;
}

// Forward 3x3x3 convolution (also the backward-data pass with flipped weights).
// A wave computes 16 voxels x 16*NT output channels; a workgroup 64 voxels.
// Wt: bf16 [Cout][Kp], k = tap * Cin + c; bias may be null.
template <int NT>
__global__ void __launch_bounds__(kBlock) k_conv_fwd(ConvIn in, int D, int H, int W, int Kp,
                                                      const uint16_t* __restrict__ Wt,
                                                      const float* __restrict__ bias, int relu,
                                                      uint16_t* __restrict__ out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t nvox = (int64_t)D * H * W;
    const int64_t m0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (m0 >= nvox) return;  // uniform over the wave
    const int Cin = in.ca + in.cb, K = 27 * Cin, Cout = 16 * NT;
    const int g = lane >> 4, i = lane & 15;
    const int64_t m = m0 + i;
    const bool mv = m < nvox;
    const int x = (int)(m % W), y = (int)((m / W) % H), z = (int)(m / ((int64_t)W * H));
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < Kp; kc += 32) {
        const int k = kc + 8 * g;
        uint4 av = make_uint4(0, 0, 0, 0);
        if (mv && k < K) {
            const int tap = k / Cin, c = k - tap * Cin;
            av = load_in8(in, D, H, W, z + tap / 9 - 1, y + (tap / 3) % 3 - 1, x + tap % 3 - 1, c);
        }
        const bf16x8 a = as_bf16x8(av);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bf16x8 b = as_bf16x8(ld16(Wt + (int64_t)(16 * t + i) * Kp + k));
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
    // D[row][col]: lane holds rows 4g..4g+3 of column i
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = 16 * t + i;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t vox = m0 + 4 * g + r;
            if (vox < nvox) {
                float v = acc[t][r] + bv;
                if (relu) v = v > 0.0f ? v : 0.0f;
                out[vox * Cout + co] = f2bf(v);
            }
        }
    }
}

// k_conv_fwd with the input staged in LDS.  A workgroup takes an output tile of 16 (x) x TY (y) x 2 (z)
// voxels (TY = 4 or 2): it loads the tile's 18 x (TY + 2) x 4 input halo (all Cin channels, zeros
// outside the grid) once, and each wave computes TY / 2 rows of 16 voxels from LDS, so every input
// voxel is read from memory about once per tile instead of once per tap.  The k order, and so every
// output value, is that of k_conv_fwd.  LDS: halo voxels * Cin bf16 (dynamic).
constexpr int kTileX = 16, kTileZ = 2;
constexpr int halo_vox(int ty) { return (kTileX + 2) * (ty + 2) * (kTileZ + 2); }

template <int NT, int TY>
__global__ void __launch_bounds__(kBlock) k_conv_fwd_tiled(ConvIn in, int D, int H, int W, int Kp,
                                                            const uint16_t* __restrict__ Wt,
                                                            const float* __restrict__ bias, int relu,
                                                            uint16_t* __restrict__ out)
{
    constexpr int HX = kTileX + 2, HY = TY + 2, NV = halo_vox(TY), RPW = TY * kTileZ / 4;  // rows per wave
    extern __shared__ uint4 tile4[];
    const int Cin = in.ca + in.cb, K = 27 * Cin, Cout = 16 * NT;
    const int tx = blockIdx.x * kTileX, ty = blockIdx.y * TY, tz = blockIdx.z * kTileZ;
    const int groups = Cin / 8;
    for (int t = threadIdx.x; t < NV * groups; t += kBlock) {
        const int v = t / groups, cg = t - v * groups;
        const int hx = v % HX, hy = (v / HX) % HY, hz = v / (HX * HY);
        tile4[(v * Cin + cg * 8) / 8] = load_in8(in, D, H, W, tz + hz - 1, ty + hy - 1, tx + hx - 1, cg * 8);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, i = lane & 15;
    f32x4 acc[RPW][NT];
#pragma unroll
    for (int rr = 0; rr < RPW; rr++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[rr][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < Kp; kc += 32) {
        const int k = kc + 8 * g;
        const int tap = k / Cin, c = k - tap * Cin;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;  // offsets + 1
        bf16x8 b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) b[t] = as_bf16x8(ld16(Wt + (int64_t)(16 * t + i) * Kp + k));
#pragma unroll
        for (int rr = 0; rr < RPW; rr++) {
            const int r = RPW * wave + rr, ry = r % TY, rz = r / TY;
            uint4 av = make_uint4(0, 0, 0, 0);
            if (k < K) av = tile4[((((rz + dz) * HY + (ry + dy)) * HX + (i + dx)) * Cin + c) / 8];
            const bf16x8 a = as_bf16x8(av);
#pragma unroll
            for (int t = 0; t < NT; t++) acc[rr][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[t], acc[rr][t], 0, 0, 0);
        }
    }
    // D[row][col]: lane holds rows 4g..4g+3 (x = tx + 4g + r) of column i
#pragma unroll
    for (int rr = 0; rr < RPW; rr++) {
        const int rw = RPW * wave + rr, y = ty + rw % TY, z = tz + rw / TY;
        if (y >= H || z >= D) continue;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int co = 16 * t + i;
            const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int x = tx + 4 * g + r;
                if (x < W) {
                    float v = acc[rr][t][r] + bv;
                    if (relu) v = v > 0.0f ? v : 0.0f;
                    out[(((int64_t)z * H + y) * W + x) * Cout + co] = f2bf(v);
                }
            }
        }
    }
}

// Weight gradient partials: partial[chunk][co][k] = sum over the chunk's voxels v of
// A[v][k] * G[v][co], A the convolution input gathered as in the forward pass, G = dL/d(pre-ReLU).
// Workgroup: one 64-wide k tile (16 k per wave) x all Cout, over vchunk voxels in slices of 32.
template <int NT>
__global__ void __launch_bounds__(kBlock) k_conv_wgrad(ConvIn in, int D, int H, int W, int Kp,
                                                        const uint16_t* __restrict__ G, int vchunk,
                                                        float* __restrict__ partial)
{
    constexpr int Cout = 16 * NT;
    __shared__ __align__(16) uint16_t As[64][40];      // [k][v], rows padded to 80 B
    __shared__ __align__(16) uint16_t Bs[Cout][40];    // [co][v]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, i = lane & 15;
    const int64_t nvox = (int64_t)D * H * W;
    const int Cin = in.ca + in.cb, K = 27 * Cin;
    const int kt = blockIdx.x * 64;
    const int64_t v_begin = (int64_t)blockIdx.y * vchunk;
    int64_t v_end = v_begin + vchunk;
    if (v_end > nvox) v_end = nvox;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // staging roles: A: thread -> (voxel tid/8, k group tid%8); B: thread -> (voxel tid%32, co group tid/32)
    const int av = tid >> 3, ak = kt + 8 * (tid & 7);
    for (int64_t v0 = v_begin; v0 < v_end; v0 += 32) {
        {
            const int64_t v = v0 + av;
            uint4 a = make_uint4(0, 0, 0, 0);
            if (v < v_end && ak < K) {
                const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((int64_t)W * H));
                const int tap = ak / Cin, c = ak - tap * Cin;
                a = load_in8(in, D, H, W, z + tap / 9 - 1, y + (tap / 3) % 3 - 1, x + tap % 3 - 1, c);
            }
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&a);
            const int kl = 8 * (tid & 7);
#pragma unroll
            for (int j = 0; j < 8; j++) As[kl + j][av] = e[j];
        }
        for (int q = tid; q < 32 * (Cout / 8); q += kBlock) {
            const int bv = q & 31, cg = (q >> 5) * 8;
            const int64_t v = v0 + bv;
            uint4 b = make_uint4(0, 0, 0, 0);
            if (v < v_end) b = ld16(G + v * Cout + cg);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&b);
#pragma unroll
            for (int j = 0; j < 8; j++) Bs[cg + j][bv] = e[j];
        }
        __syncthreads();
        const bf16x8 a = as_bf16x8(*reinterpret_cast<const uint4*>(&As[16 * wave + i][8 * g]));
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bf16x8 b = as_bf16x8(*reinterpret_cast<const uint4*>(&Bs[16 * t + i][8 * g]));
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
        __syncthreads();
    }
    // D[k][co]: lane holds k rows 4g..4g+3 (within the wave's 16) of column co = 16t + i
    float* P = partial + (int64_t)blockIdx.y * Cout * Kp;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) P[(int64_t)(16 * t + i) * Kp + kt + 16 * wave + 4 * g + r] = acc[t][r];
}

// grad[co][k] (k < K) = sum over chunks of partial[chunk][co][k], in chunk order
__global__ void k_reduce_partials(int nchunks, int Cout, int K, int Kp, const float* __restrict__ partial,
                                  float* __restrict__ grad)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)Cout * K) return;
    const int co = (int)(idx / K), k = (int)(idx % K);
    float s = 0.0f;
    for (int c = 0; c < nchunks; c++) s += partial[((int64_t)c * Cout + co) * Kp + k];
    grad[idx] = s;
}

// bias gradient partials: partial[chunk][co] = sum of G[v][co] over the chunk's voxels
__global__ void k_bias_partials(int64_t nvox, int Cout, int vchunk, const uint16_t* __restrict__ G,
                                float* __restrict__ partial)
{
    const int co = threadIdx.x;
    if (co >= Cout) return;
    const int64_t v0 = (int64_t)blockIdx.x * vchunk;
    int64_t v1 = v0 + vchunk;
    if (v1 > nvox) v1 = nvox;
    float s = 0.0f;
    for (int64_t v = v0; v < v1; v++) s += bf2f(G[v * Cout + co]);
    partial[(int64_t)blockIdx.x * Cout + co] = s;
}
__global__ void k_reduce_bias(int nchunks, int Cout, const float* __restrict__ partial, float* __restrict__ grad)
{
    const int co = threadIdx.x;
    if (co >= Cout) return;
    float s = 0.0f;
    for (int c = 0; c < nchunks; c++) s += partial[(int64_t)c * Cout + co];
    grad[co] = s;
}

// master [Cout][27][Cin] fp32 -> forward bf16 [Cout][Kp] and flipped bf16 [Cin][Kpf] (k' = tap' * Cout + co,
// tap' = 26 - tap), zero padding beyond K
__global__ void k_pack_weights(const float* __restrict__ w, int Cin, int Cout, int Kp, int Kpf,
                               uint16_t* __restrict__ wt, uint16_t* __restrict__ wf)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n1 = (int64_t)Cout * Kp, n2 = (int64_t)Cin * Kpf;
    if (idx < n1) {
        const int co = (int)(idx / Kp), k = (int)(idx % Kp);
        wt[idx] = k < 27 * Cin ? f2bf(w[(int64_t)co * 27 * Cin + k]) : (uint16_t)0;
    } else if (idx < n1 + n2) {
        const int64_t j = idx - n1;
        const int c = (int)(j / Kpf), kf = (int)(j % Kpf);
        uint16_t v = 0;
        if (kf < 27 * Cout) {
            const int tapf = kf / Cout, co = kf % Cout;
            v = f2bf(w[((int64_t)co * 27 + (26 - tapf)) * Cin + c]);
        }
        wf[j] = v;
    }
}

// in8 = (FBP, hull, 0...) as bf16
__global__ void k_pack_input(int64_t nvox, const float* __restrict__ fbp, const uint8_t* __restrict__ hull,
                             uint16_t* __restrict__ in8)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    uint16_t e[8] = {f2bf(fbp[v]), f2bf(hull[v] ? 1.0f : 0.0f), 0, 0, 0, 0, 0, 0};
    *reinterpret_cast<uint4*>(in8 + v * 8) = *reinterpret_cast<const uint4*>(e);
}

// 2x2x2 average pooling; out dims D/2 x H/2 x W/2
__global__ void k_avgpool(int D, int H, int W, int C, const uint16_t* __restrict__ in, uint16_t* __restrict__ out)
{
    const int64_t n = (int64_t)(D / 2) * (H / 2) * (W / 2) * C;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int c = (int)(idx % C);
    const int64_t p = idx / C;
    const int w2 = W / 2, h2 = H / 2;
    const int x = (int)(p % w2), y = (int)((p / w2) % h2), z = (int)(p / ((int64_t)w2 * h2));
    float s = 0.0f;
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++)
                s += bf2f(in[(((int64_t)(2 * z + dz) * H + (2 * y + dy)) * W + (2 * x + dx)) * C + c]);
    out[idx] = f2bf(s * 0.125f);
}

// dPre[v][c] = act[v][c] > 0 ? skip_grad[v][c_off + c] + (pool_grad ? pool_grad[parent][c] / 8 : 0) : 0
// skip_grad is the concat gradient at full resolution with cs channels (this tensor at offset c_off).
__global__ void k_merge_grad(int D, int H, int W, int C, const uint16_t* __restrict__ act,
                             const uint16_t* __restrict__ skip_grad, int cs, int c_off,
                             const uint16_t* __restrict__ pool_grad, uint16_t* __restrict__ dpre)
{
    const int64_t n = (int64_t)D * H * W * C;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int c = (int)(idx % C);
    const int64_t v = idx / C;
    float gsum = bf2f(skip_grad[v * cs + c_off + c]);
    if (pool_grad) {
        const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((int64_t)W * H));
        const int64_t p = ((int64_t)(z >> 1) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1);
        gsum += bf2f(pool_grad[p * C + c]) * 0.125f;
    }
    dpre[idx] = bf2f(act[idx]) > 0.0f ? f2bf(gsum) : (uint16_t)0;
}

// Gradient of a coarse tensor that was upsampled into channels [0, C) of a concat at full resolution
// (cs channels): dPre[p][c] = act[p][c] > 0 ? sum over the 8 children of cat_grad[child][c] : 0
__global__ void k_up_grad(int D, int H, int W, int C, int cs, const uint16_t* __restrict__ cat_grad,
                          const uint16_t* __restrict__ act, uint16_t* __restrict__ dpre)
{
    const int64_t n = (int64_t)(D / 2) * (H / 2) * (W / 2) * C;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int c = (int)(idx % C);
    const int64_t p = idx / C;
    const int w2 = W / 2, h2 = H / 2;
    const int x = (int)(p % w2), y = (int)((p / w2) % h2), z = (int)(p / ((int64_t)w2 * h2));
    float s = 0.0f;
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++)
                s += bf2f(cat_grad[(((int64_t)(2 * z + dz) * H + (2 * y + dy)) * W + (2 * x + dx)) * cs + c]);
    dpre[idx] = bf2f(act[idx]) > 0.0f ? f2bf(s) : (uint16_t)0;
}

// head: y = w . d1 + b; x0 = hull ? FBP + y : 0; per-block partial sums of the masked squared error
// and of the hull count
__global__ void k_head_fwd(int64_t nvox, const uint16_t* __restrict__ d1, const float* __restrict__ wh,
                           const float* __restrict__ fbp, const uint8_t* __restrict__ hull,
                           const float* __restrict__ target, float* __restrict__ x0, double* __restrict__ part)
{
    __shared__ double s_err[kBlock], s_cnt[kBlock];
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, c = 0.0;
    if (v < nvox) {
        const uint4 q0 = ld16(d1 + v * 16), q1 = ld16(d1 + v * 16 + 8);
        const uint16_t* a = reinterpret_cast<const uint16_t*>(&q0);
        const uint16_t* b = reinterpret_cast<const uint16_t*>(&q1);
        float y = wh[16];
#pragma unroll
        for (int j = 0; j < 8; j++) y += wh[j] * bf2f(a[j]);
#pragma unroll
        for (int j = 0; j < 8; j++) y += wh[8 + j] * bf2f(b[j]);
        const float xv = hull[v] ? fbp[v] + y : 0.0f;
        x0[v] = xv;
        if (target && hull[v]) {
            const double d = (double)xv - (double)target[v];
            e = d * d;
            c = 1.0;
        }
    }
    s_err[threadIdx.x] = e;
    s_cnt[threadIdx.x] = c;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_err[threadIdx.x] += s_err[threadIdx.x + o];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s_err[0];
        part[2 * blockIdx.x + 1] = s_cnt[0];
    }
}

// head backward: dy = 2 (x0 - t) / n_hull on hull voxels; dPre_dec1[v][c] = d1 > 0 ? dy * w[c] : 0;
// per-block partials of dw[c] = sum dy * d1[v][c] and db = sum dy
__global__ void k_head_bwd(int64_t nvox, double inv_n, const uint16_t* __restrict__ d1, const float* __restrict__ wh,
                           const uint8_t* __restrict__ hull, const float* __restrict__ x0,
                           const float* __restrict__ target, uint16_t* __restrict__ dpre, float* __restrict__ part)
{
    __shared__ float s[17][kBlock];
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float dy = 0.0f;
    float a[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = 0.0f;
    if (v < nvox) {
        const uint4 q0 = ld16(d1 + v * 16), q1 = ld16(d1 + v * 16 + 8);
        const uint16_t* p0 = reinterpret_cast<const uint16_t*>(&q0);
        const uint16_t* p1 = reinterpret_cast<const uint16_t*>(&q1);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            a[j] = bf2f(p0[j]);
            a[8 + j] = bf2f(p1[j]);
        }
        if (hull[v]) dy = (float)(2.0 * ((double)x0[v] - (double)target[v]) * inv_n);
        uint16_t o[16];
#pragma unroll
        for (int j = 0; j < 16; j++) o[j] = a[j] > 0.0f ? f2bf(dy * wh[j]) : (uint16_t)0;
        *reinterpret_cast<uint4*>(dpre + v * 16) = *reinterpret_cast<const uint4*>(o);
        *reinterpret_cast<uint4*>(dpre + v * 16 + 8) = *reinterpret_cast<const uint4*>(o + 8);
    }
#pragma unroll
    for (int j = 0; j < 16; j++) s[j][threadIdx.x] = dy * a[j];
    s[16][threadIdx.x] = dy;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int j = 0; j < 17; j++) s[j][threadIdx.x] += s[j][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 17) part[(int64_t)blockIdx.x * 17 + threadIdx.x] = s[threadIdx.x][0];
}
__global__ void k_reduce_cols(int nrows, int ncols, const float* __restrict__ part, float* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    float s = 0.0f;
    for (int r = 0; r < nrows; r++) s += part[(int64_t)r * ncols + c];
    out[c] = s;
}

__global__ void k_adam(int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned blocks_for(int64_t n, int b = kBlock) { return pct_grid_for(n, b); }

}  // namespace

// ------------------------------------------------------------------------------------ host side
struct pct_net {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int D = 0, H = 0, W = 0;  // slices, rows, columns at full resolution
    int64_t nv[4] = {};       // voxels per level
    struct Layer {
        int cin, cout, level, Kp, Kpf;
        int64_t w_off, b_off;
        uint16_t* wt = nullptr;
        uint16_t* wf = nullptr;
    } L[kLayers];
    int64_t head_off = 0, n_params = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    int64_t adam_t = 0;
    // activations (bf16)
    uint16_t *in8 = nullptr, *e[4] = {}, *pool[3] = {}, *dd[3] = {};  // dd[0] = dec1, dd[1] = dec2, dd[2] = dec3
    float* x0 = nullptr;
    // gradients (bf16)
    uint16_t *g_dd[3] = {}, *g_cat[3] = {}, *g_e[4] = {}, *g_pool[3] = {};
    float* partial = nullptr;
    int64_t partial_n = 0;
    double* loss_part = nullptr;
    float* head_part = nullptr;
    // scratch input / target (device)
    float *fbp = nullptr, *target = nullptr;
    uint8_t* hull = nullptr;
    struct Sample {
        float* fbp;
        uint8_t* hull;
        float* target;
    };
    std::vector<Sample> samples;
};

namespace {

#define NET_HIP(net, expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            (net)->err = std::string(#expr) + ": " + hipGetErrorString(e_);       \
            return PCT_E_HIP;                                                     \
        }                                                                         \
    } while (0)

template <class T>
int net_alloc(pct_net* net, T** ptr, int64_t n)
{
    if (hipMalloc((void**)ptr, (size_t)(n > 0 ? n : 1) * sizeof(T)) != hipSuccess) {
        net->err = "hipMalloc failed (" + std::to_string(n * sizeof(T)) + " B)";
        *ptr = nullptr;
        return PCT_E_NOMEM;
    }
    return PCT_OK;
}
template <class T>
void net_free(T** ptr)
{
    if (*ptr) (void)hipFree(*ptr);
    *ptr = nullptr;
}

// deterministic normal numbers (splitmix64 + Box-Muller)
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

int pack_weights(pct_net* net)
{
    for (int l = 0; l < kLayers; l++) {
        auto& Ly = net->L[l];
        const int64_t n = (int64_t)Ly.cout * Ly.Kp + (int64_t)Ly.cin * Ly.Kpf;
        hipLaunchKernelGGL(k_pack_weights, dim3(blocks_for(n)), dim3(kBlock), 0, net->stream, net->p + Ly.w_off, Ly.cin,
                           Ly.cout, Ly.Kp, Ly.Kpf, Ly.wt, Ly.wf);
    }
    NET_HIP(net, hipGetLastError());
    return PCT_OK;
}

int launch_conv(pct_net* net, const ConvIn& in, int level, int Kp, const uint16_t* w, const float* bias, int relu,
                int cout, uint16_t* out)
{
    const int D = net->D >> level, H = net->H >> level, W = net->W >> level;
    const int64_t nvox = (int64_t)D * H * W;
    const unsigned grid = (unsigned)((nvox + 63) / 64);
    hipStream_t s = net->stream;
    // the input halo in LDS: 16 x 4 x 2 output tiles, or 16 x 2 x 2 for wider inputs
    const int cin = in.ca + in.cb;
    const int ty = (size_t)halo_vox(4) * cin * sizeof(uint16_t) <= 64 * 1024 ? 4
                   : (size_t)halo_vox(2) * cin * sizeof(uint16_t) <= 64 * 1024 ? 2 : 0;
    if (ty > 0) {
        const size_t lds = (size_t)halo_vox(ty) * cin * sizeof(uint16_t);
        const dim3 tg((W + kTileX - 1) / kTileX, (H + ty - 1) / ty, (D + kTileZ - 1) / kTileZ);
#define PCT_CONV_T(NT)                                                                                            \
    if (ty == 4)                                                                                                  \
        hipLaunchKernelGGL((k_conv_fwd_tiled<NT, 4>), tg, dim3(kBlock), lds, s, in, D, H, W, Kp, w, bias, relu, out); \
    else                                                                                                          \
        hipLaunchKernelGGL((k_conv_fwd_tiled<NT, 2>), tg, dim3(kBlock), lds, s, in, D, H, W, Kp, w, bias, relu, out);
        switch (cout) {
        case 16: PCT_CONV_T(1) break;
        case 32: PCT_CONV_T(2) break;
        case 48: PCT_CONV_T(3) break;
        case 64: PCT_CONV_T(4) break;
        case 96: PCT_CONV_T(6) break;
        case 128: PCT_CONV_T(8) break;
        case 192: PCT_CONV_T(12) break;
        default: net->err = "unsupported channel count"; return PCT_E_ARG;
        }
#undef PCT_CONV_T
        NET_HIP(net, hipGetLastError());
        return PCT_OK;
    }
    switch (cout) {
    case 16: hipLaunchKernelGGL(k_conv_fwd<1>, dim3(grid), dim3(kBlock), 0, s, in, D, H, W, Kp, w, bias, relu, out); break;
    case 32: hipLaunchKernelGGL(k_conv_fwd<2>, dim3(grid), dim3(kBlock), 0, s, in, D, H, W, Kp, w, bias, relu, out); break;
    case 48: hipLaunchKernelGGL(k_conv_fwd<3>, dim3(grid), dim3(kBlock), 0, s, in, D, H, W, Kp, w, bias, relu, out); break;
    case 64: hipLaunchKernelGGL(k_conv_fwd<4>, dim3(grid), dim3(kBlock), 0, s, in, D, H, W, Kp, w, bias, relu, out); break;
    case 96: hipLaunchKernelGGL(k_conv_fwd<6>, dim3(grid), dim3(kBlock), 0, s, in, D, H, W, Kp, w, bias, relu, out); break;
    case 128: hipLaunchKernelGGL(k_conv_fwd<8>, dim3(grid), dim3(kBlock), 0, s, in, D, H, W, Kp, w, bias, relu, out); break;
    case 192: hipLaunchKernelGGL(k_conv_fwd<12>, dim3(grid), dim3(kBlock), 0, s, in, D, H, W, Kp, w, bias, relu, out); break;
    default: net->err = "unsupported channel count"; return PCT_E_ARG;
    }
    NET_HIP(net, hipGetLastError());
    return PCT_OK;
}

constexpr int64_t kBiasChunk = 256;  // voxels per bias-gradient partial

// split of a weight-gradient launch: about 1024 workgroups of (64-wide k tile, voxel chunk)
void wgrad_shape(int64_t nvox, int Kp, int64_t* nchunks, int64_t* vchunk)
{
    const int ktiles = Kp / 64;
    int64_t nc = 1024 / ktiles;
    if (nc < 1) nc = 1;
    int64_t vc = (nvox + nc - 1) / nc;
    vc = (vc + 31) / 32 * 32;
    *vchunk = vc;
    *nchunks = (nvox + vc - 1) / vc;
}

// floats of partial sums the weight and bias gradients of layer l need
int64_t partial_need(const pct_net* net, int l)
{
    const auto& Ly = net->L[l];
    const int64_t nvox = net->nv[Ly.level];
    int64_t nc = 0, vc = 0;
    wgrad_shape(nvox, Ly.Kp, &nc, &vc);
    const int64_t a = nc * Ly.cout * Ly.Kp, b = (nvox + kBiasChunk - 1) / kBiasChunk * Ly.cout;
    return a > b ? a : b;
}

int launch_wgrad(pct_net* net, const ConvIn& in, int level, int l, const uint16_t* G)
{
    auto& Ly = net->L[l];
    const int D = net->D >> level, H = net->H >> level, W = net->W >> level;
    const int64_t nvox = (int64_t)D * H * W;
    const int ktiles = Ly.Kp / 64;
    int64_t nchunks = 0, vchunk = 0;
    wgrad_shape(nvox, Ly.Kp, &nchunks, &vchunk);
    if (partial_need(net, l) > net->partial_n) {
        net->err = "gradient scratch too small";
        return PCT_E_LIMIT;
    }
    hipStream_t s = net->stream;
    dim3 grid((unsigned)ktiles, (unsigned)nchunks);
    switch (Ly.cout) {
    case 16: hipLaunchKernelGGL(k_conv_wgrad<1>, grid, dim3(kBlock), 0, s, in, D, H, W, Ly.Kp, G, (int)vchunk, net->partial); break;
    case 32: hipLaunchKernelGGL(k_conv_wgrad<2>, grid, dim3(kBlock), 0, s, in, D, H, W, Ly.Kp, G, (int)vchunk, net->partial); break;
    case 64: hipLaunchKernelGGL(k_conv_wgrad<4>, grid, dim3(kBlock), 0, s, in, D, H, W, Ly.Kp, G, (int)vchunk, net->partial); break;
    case 128: hipLaunchKernelGGL(k_conv_wgrad<8>, grid, dim3(kBlock), 0, s, in, D, H, W, Ly.Kp, G, (int)vchunk, net->partial); break;
    default: net->err = "unsupported channel count"; return PCT_E_ARG;
    }
    const int K = 27 * Ly.cin;
    hipLaunchKernelGGL(k_reduce_partials, dim3(blocks_for((int64_t)Ly.cout * K)), dim3(kBlock), 0, s, (int)nchunks,
                       Ly.cout, K, Ly.Kp, net->partial, net->g + Ly.w_off);
    // bias
    const int64_t nb = (nvox + kBiasChunk - 1) / kBiasChunk;
    hipLaunchKernelGGL(k_bias_partials, dim3((unsigned)nb), dim3(128), 0, s, nvox, Ly.cout, (int)kBiasChunk, G,
                       net->partial);
    hipLaunchKernelGGL(k_reduce_bias, dim3(1), dim3(128), 0, s, (int)nb, Ly.cout, net->partial, net->g + Ly.b_off);
    NET_HIP(net, hipGetLastError());
    return PCT_OK;
}

ConvIn conv_input(pct_net* net, int l)
{
    ConvIn in{};
    switch (l) {
    case 0: in = ConvIn{net->in8, nullptr, 8, 0, 0}; break;
    case 1: in = ConvIn{net->pool[0], nullptr, 16, 0, 0}; break;
    case 2: in = ConvIn{net->pool[1], nullptr, 32, 0, 0}; break;
    case 3: in = ConvIn{net->pool[2], nullptr, 64, 0, 0}; break;
    case 4: in = ConvIn{net->e[3], net->e[2], 128, 64, 1}; break;   // dec3
    case 5: in = ConvIn{net->dd[2], net->e[1], 64, 32, 1}; break;   // dec2
    default: in = ConvIn{net->dd[1], net->e[0], 32, 16, 1}; break;  // dec1
    }
    return in;
}

// forward pass from the device arrays fbp/hull; x0 (and the loss when target != null)
int net_forward(pct_net* net, const float* fbp, const uint8_t* hull, const float* target, double* loss)
{
    hipStream_t s = net->stream;
    const int64_t n0 = net->nv[0];
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(n0)), dim3(kBlock), 0, s, n0, fbp, hull, net->in8);
    uint16_t* outs[kLayers] = {net->e[0], net->e[1], net->e[2], net->e[3], net->dd[2], net->dd[1], net->dd[0]};
    const int level[kLayers] = {0, 1, 2, 3, 2, 1, 0};
    for (int l = 0; l < kLayers; l++) {
        auto& Ly = net->L[l];
        PCT_TRY(launch_conv(net, conv_input(net, l), level[l], Ly.Kp, Ly.wt, net->p + Ly.b_off, 1, Ly.cout, outs[l]));
        if (l < 3) {
            const int D = net->D >> l, H = net->H >> l, W = net->W >> l;
            const int64_t n = (int64_t)(D / 2) * (H / 2) * (W / 2) * Ly.cout;
            hipLaunchKernelGGL(k_avgpool, dim3(blocks_for(n)), dim3(kBlock), 0, s, D, H, W, Ly.cout, outs[l], net->pool[l]);
        }
    }
    const unsigned nb = blocks_for(n0);
    hipLaunchKernelGGL(k_head_fwd, dim3(nb), dim3(kBlock), 0, s, n0, net->dd[0], net->p + net->head_off, fbp, hull, target,
                       net->x0, net->loss_part);
    NET_HIP(net, hipGetLastError());
    if (loss) {
        std::vector<double> part(2 * (size_t)nb);
        NET_HIP(net, hipMemcpyAsync(part.data(), net->loss_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        NET_HIP(net, hipStreamSynchronize(s));
        double e = 0, c = 0;
        for (unsigned b = 0; b < nb; b++) {
            e += part[2 * b];
            c += part[2 * b + 1];
        }
        *loss = c > 0 ? e / c : 0.0;
    }
    return PCT_OK;
}

int hull_count(pct_net* net, double* cnt)
{
    const unsigned nb = blocks_for(net->nv[0]);
    std::vector<double> part(2 * (size_t)nb);
    NET_HIP(net, hipMemcpyAsync(part.data(), net->loss_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                net->stream));
    NET_HIP(net, hipStreamSynchronize(net->stream));
    double c = 0;
    for (unsigned b = 0; b < nb; b++) c += part[2 * b + 1];
    *cnt = c;
    return PCT_OK;
}

// backward pass after net_forward with a target: fills net->g
int net_backward(pct_net* net, const uint8_t* hull, const float* target)
{
    hipStream_t s = net->stream;
    const int64_t n0 = net->nv[0];
    double cnt = 0;
    PCT_TRY(hull_count(net, &cnt));
    const double inv_n = cnt > 0 ? 1.0 / cnt : 0.0;
    const unsigned nb = blocks_for(n0);
    hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(kBlock), 0, s, n0, inv_n, net->dd[0], net->p + net->head_off, hull,
                       net->x0, target, net->g_dd[0], net->head_part);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(64), 0, s, (int)nb, 17, net->head_part, net->g + net->head_off);
    // decoder: l = 6 (dec1, level 0), 5 (dec2, level 1), 4 (dec3, level 2)
    for (int d = 0; d < 3; d++) {
        const int l = 6 - d, level = d;
        auto& Ly = net->L[l];
        const ConvIn in = conv_input(net, l);
        PCT_TRY(launch_wgrad(net, in, level, l, net->g_dd[d]));
        // concat gradient: conv of dPre with flipped weights (Cin' = Cout, Cout' = Cin)
        ConvIn gin{net->g_dd[d], nullptr, Ly.cout, 0, 0};
        PCT_TRY(launch_conv(net, gin, level, Ly.Kpf, Ly.wf, nullptr, 0, Ly.cin, net->g_cat[d]));
        const int D = net->D >> level, H = net->H >> level, W = net->W >> level;
        const int ca = in.ca;
        // coarse part -> dPre of the tensor one level down (dec2, dec3 or enc4)
        uint16_t* coarse_act = d < 2 ? net->dd[d + 1] : net->e[3];
        uint16_t* coarse_g = d < 2 ? net->g_dd[d + 1] : net->g_e[3];
        const int64_t nc = (int64_t)(D / 2) * (H / 2) * (W / 2) * ca;
        hipLaunchKernelGGL(k_up_grad, dim3(blocks_for(nc)), dim3(kBlock), 0, s, D, H, W, ca, Ly.cin, net->g_cat[d],
                           coarse_act, coarse_g);
    }
    // encoder: enc4 (l = 3) .. enc1 (l = 0); g_e[l] holds dPre of enc l+1
    for (int l = 3; l >= 0; l--) {
        auto& Ly = net->L[l];
        const int level = l;
        if (l < 3) {
            // dPre(enc l) = mask * (skip part of the decoder concat at this level + pool gradient)
            const int D = net->D >> level, H = net->H >> level, W = net->W >> level;
            const int d = level;  // decoder index consuming enc l as skip
            const int cs = net->L[6 - d].cin, c_off = net->L[6 - d].cin - Ly.cout;
            const int64_t n = (int64_t)D * H * W * Ly.cout;
            hipLaunchKernelGGL(k_merge_grad, dim3(blocks_for(n)), dim3(kBlock), 0, s, D, H, W, Ly.cout, net->e[l],
                               net->g_cat[d], cs, c_off, net->g_pool[l], net->g_e[l]);
        }
        const ConvIn in = conv_input(net, l);
        PCT_TRY(launch_wgrad(net, in, level, l, net->g_e[l]));
        if (l > 0) {
            // gradient of the pooled input of this layer
            ConvIn gin{net->g_e[l], nullptr, Ly.cout, 0, 0};
            PCT_TRY(launch_conv(net, gin, level, Ly.Kpf, Ly.wf, nullptr, 0, Ly.cin, net->g_pool[l - 1]));
        }
    }
    NET_HIP(net, hipGetLastError());
    return PCT_OK;
}

int upload_sample(pct_net* net, const float* fbp, const uint8_t* hull, const float* target)
{
    const int64_t n0 = net->nv[0];
    hipStream_t s = net->stream;
    NET_HIP(net, hipMemcpyAsync(net->fbp, fbp, n0 * sizeof(float), hipMemcpyHostToDevice, s));
    NET_HIP(net, hipMemcpyAsync(net->hull, hull, n0, hipMemcpyHostToDevice, s));
    if (target) NET_HIP(net, hipMemcpyAsync(net->target, target, n0 * sizeof(float), hipMemcpyHostToDevice, s));
    return PCT_OK;
}

}  // namespace

extern "C" {

int pct_net_create(int columns, int rows, int slices, uint64_t seed, int device, pct_net** out)
{
    if (!out || columns <= 0 || rows <= 0 || slices <= 0 || columns % 8 || rows % 8 || slices % 8) return PCT_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PCT_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return PCT_E_HIP;
    pct_net* net = new pct_net();
    net->device = device;
    net->D = slices;
    net->H = rows;
    net->W = columns;
    for (int lv = 0; lv < 4; lv++) net->nv[lv] = (int64_t)(slices >> lv) * (rows >> lv) * (columns >> lv);
    const int cin[kLayers] = {8, 16, 32, 64, 192, 96, 48};
    const int cout[kLayers] = {16, 32, 64, 128, 64, 32, 16};
    const int level[kLayers] = {0, 1, 2, 3, 2, 1, 0};
    int64_t off = 0;
    for (int l = 0; l < kLayers; l++) {
        auto& Ly = net->L[l];
        Ly.cin = cin[l];
        Ly.cout = cout[l];
        Ly.level = level[l];
        Ly.Kp = (27 * cin[l] + 63) / 64 * 64;
        Ly.Kpf = (27 * cout[l] + 63) / 64 * 64;
        Ly.w_off = off;
        off += (int64_t)cout[l] * 27 * cin[l];
        Ly.b_off = off;
        off += cout[l];
    }
    net->head_off = off;
    off += 17;
    net->n_params = off;
    auto fail = [&](int rc) {
        pct_net_destroy(net);
        return rc;
    };
    if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess) return fail(PCT_E_HIP);
    int rc = PCT_OK;
    if ((rc = net_alloc(net, &net->p, off)) || (rc = net_alloc(net, &net->g, off)) || (rc = net_alloc(net, &net->m, off)) ||
        (rc = net_alloc(net, &net->v, off)))
        return fail(rc);
    for (int l = 0; l < kLayers; l++) {
        auto& Ly = net->L[l];
        if ((rc = net_alloc(net, &Ly.wt, (int64_t)Ly.cout * Ly.Kp)) || (rc = net_alloc(net, &Ly.wf, (int64_t)Ly.cin * Ly.Kpf)))
            return fail(rc);
    }
    const int64_t* nv = net->nv;
    if ((rc = net_alloc(net, &net->in8, nv[0] * 8))) return fail(rc);
    for (int l = 0; l < 4; l++) {
        if ((rc = net_alloc(net, &net->e[l], nv[l] * cout[l]))) return fail(rc);
        if ((rc = net_alloc(net, &net->g_e[l], nv[l] * cout[l]))) return fail(rc);
    }
    for (int l = 0; l < 3; l++) {
        if ((rc = net_alloc(net, &net->pool[l], nv[l + 1] * cout[l]))) return fail(rc);
        if ((rc = net_alloc(net, &net->g_pool[l], nv[l + 1] * cout[l]))) return fail(rc);
        // decoder d at level d: dd[d] has cout[6 - d] channels, its concat cin[6 - d]
        if ((rc = net_alloc(net, &net->dd[l], nv[l] * cout[6 - l]))) return fail(rc);
        if ((rc = net_alloc(net, &net->g_dd[l], nv[l] * cout[6 - l]))) return fail(rc);
        if ((rc = net_alloc(net, &net->g_cat[l], nv[l] * cin[6 - l]))) return fail(rc);
    }
    if ((rc = net_alloc(net, &net->x0, nv[0])) || (rc = net_alloc(net, &net->fbp, nv[0])) ||
        (rc = net_alloc(net, &net->target, nv[0])) || (rc = net_alloc(net, &net->hull, nv[0])) ||
        (rc = net_alloc(net, &net->loss_part, 2 * (int64_t)blocks_for(nv[0]))) ||
        (rc = net_alloc(net, &net->head_part, 17 * (int64_t)blocks_for(nv[0]))))
        return fail(rc);
    for (int l = 0; l < kLayers; l++) {
        const int64_t need = partial_need(net, l);
        net->partial_n = need > net->partial_n ? need : net->partial_n;
    }
    if ((rc = net_alloc(net, &net->partial, net->partial_n))) return fail(rc);
    // He-normal weights, zero biases, zero head (x0 starts as hull * FBP)
    std::vector<float> h(off, 0.0f);
    Rng rng{seed * 0x2545f4914f6cdd1dull + 1};
    for (int l = 0; l < kLayers; l++) {
        auto& Ly = net->L[l];
        const double sd = std::sqrt(2.0 / (27.0 * Ly.cin));
        for (int64_t j = 0; j < (int64_t)Ly.cout * 27 * Ly.cin; j++) h[Ly.w_off + j] = (float)(sd * rng.normal());
    }
    if (hipMemcpy(net->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return fail(PCT_E_HIP);
    (void)hipMemset(net->m, 0, off * sizeof(float));
    (void)hipMemset(net->v, 0, off * sizeof(float));
    if (pack_weights(net) != PCT_OK) return fail(PCT_E_HIP);
    if (hipStreamSynchronize(net->stream) != hipSuccess) return fail(PCT_E_HIP);
    *out = net;
    return PCT_OK;
}

int pct_net_destroy(pct_net* net)
{
    if (!net) return PCT_E_ARG;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    net_free(&net->p);
    net_free(&net->g);
    net_free(&net->m);
    net_free(&net->v);
    for (auto& Ly : net->L) {
        net_free(&Ly.wt);
        net_free(&Ly.wf);
    }
    net_free(&net->in8);
    for (int l = 0; l < 4; l++) {
        net_free(&net->e[l]);
        net_free(&net->g_e[l]);
    }
    for (int l = 0; l < 3; l++) {
        net_free(&net->pool[l]);
        net_free(&net->g_pool[l]);
        net_free(&net->dd[l]);
        net_free(&net->g_dd[l]);
        net_free(&net->g_cat[l]);
    }
    net_free(&net->x0);
    net_free(&net->fbp);
    net_free(&net->target);
    net_free(&net->hull);
    net_free(&net->loss_part);
    net_free(&net->head_part);
    net_free(&net->partial);
    for (auto& sm : net->samples) {
        net_free(&sm.fbp);
        net_free(&sm.hull);
        net_free(&sm.target);
    }
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
    return PCT_OK;
}

const char* pct_net_last_error(pct_net* net) { return net ? net->err.c_str() : "null net"; }

int64_t pct_net_num_params(pct_net* net) { return net ? net->n_params : -1; }

int pct_net_get_params(pct_net* net, float* out, int64_t n)
{
    if (!net || !out || n != net->n_params) return PCT_E_ARG;
    NET_HIP(net, hipMemcpy(out, net->p, n * sizeof(float), hipMemcpyDeviceToHost));
    return PCT_OK;
}

int pct_net_set_params(pct_net* net, const float* in, int64_t n)
{
    if (!net || !in || n != net->n_params) return PCT_E_ARG;
    NET_HIP(net, hipMemcpy(net->p, in, n * sizeof(float), hipMemcpyHostToDevice));
    PCT_TRY(pack_weights(net));
    NET_HIP(net, hipStreamSynchronize(net->stream));
    return PCT_OK;
}

int pct_net_loss_grad(pct_net* net, const float* fbp, const uint8_t* hull, const float* target, double* loss,
                      float* grad)
{
    if (!net || !fbp || !hull || !target || !loss) return PCT_E_ARG;
    PCT_TRY(upload_sample(net, fbp, hull, target));
    PCT_TRY(net_forward(net, net->fbp, net->hull, net->target, loss));
    if (grad) {
        PCT_TRY(net_backward(net, net->hull, net->target));
        NET_HIP(net, hipMemcpyAsync(grad, net->g, net->n_params * sizeof(float), hipMemcpyDeviceToHost, net->stream));
        NET_HIP(net, hipStreamSynchronize(net->stream));
    }
    return PCT_OK;
}

int pct_net_predict(pct_net* net, const float* fbp, const uint8_t* hull, float* x0)
{
    if (!net || !fbp || !hull || !x0) return PCT_E_ARG;
    PCT_TRY(upload_sample(net, fbp, hull, nullptr));
    PCT_TRY(net_forward(net, net->fbp, net->hull, nullptr, nullptr));
    NET_HIP(net, hipMemcpyAsync(x0, net->x0, net->nv[0] * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    NET_HIP(net, hipStreamSynchronize(net->stream));
    return PCT_OK;
}

int pct_net_add_sample(pct_net* net, const float* fbp, const uint8_t* hull, const float* target)
{
    if (!net || !fbp || !hull || !target) return PCT_E_ARG;
    const int64_t n0 = net->nv[0];
    pct_net::Sample sm{nullptr, nullptr, nullptr};
    int rc;
    if ((rc = net_alloc(net, &sm.fbp, n0)) || (rc = net_alloc(net, &sm.hull, n0)) || (rc = net_alloc(net, &sm.target, n0))) {
        net_free(&sm.fbp);
        net_free(&sm.hull);
        net_free(&sm.target);
        return rc;
    }
    net->samples.push_back(sm);
    NET_HIP(net, hipMemcpy(sm.fbp, fbp, n0 * sizeof(float), hipMemcpyHostToDevice));
    NET_HIP(net, hipMemcpy(sm.hull, hull, n0, hipMemcpyHostToDevice));
    NET_HIP(net, hipMemcpy(sm.target, target, n0 * sizeof(float), hipMemcpyHostToDevice));
    return PCT_OK;
}

int pct_net_num_samples(pct_net* net) { return net ? (int)net->samples.size() : -1; }

int pct_net_train(pct_net* net, int steps, double lr, uint64_t seed, double* mean_loss)
{
    if (!net || steps < 0 || net->samples.empty()) return PCT_E_ARG;
    hipStream_t s = net->stream;
    const int ns = (int)net->samples.size();
    std::vector<int> order(ns);
    Rng rng{seed + 0x51ed27ull};
    double lsum = 0.0;
    int pos = ns;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < steps; it++) {
        if (pos == ns) {  // reshuffle each epoch (Fisher-Yates)
            for (int k = 0; k < ns; k++) order[k] = k;
            for (int k = ns - 1; k > 0; k--) {
                int j = (int)(rng.next() % (uint64_t)(k + 1));
                std::swap(order[k], order[j]);
            }
            pos = 0;
        }
        const auto& sm = net->samples[order[pos++]];
        double loss = 0;
        PCT_TRY(net_forward(net, sm.fbp, sm.hull, sm.target, &loss));
        PCT_TRY(net_backward(net, sm.hull, sm.target));
        net->adam_t++;
        const float c1 = 1.0f - std::pow(b1, (float)net->adam_t), c2 = 1.0f - std::pow(b2, (float)net->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(net->n_params)), dim3(kBlock), 0, s, net->n_params, (float)lr, b1, b2,
                           eps, c1, c2, net->p, net->g, net->m, net->v);
        PCT_TRY(pack_weights(net));
        lsum += loss;
    }
    NET_HIP(net, hipStreamSynchronize(s));
    if (mean_loss) *mean_loss = steps > 0 ? lsum / steps : 0.0;
    return PCT_OK;
}

int pct_net_save(pct_net* net, const char* path)
{
    if (!net || !path) return PCT_E_ARG;
    std::vector<float> h(net->n_params);
    NET_HIP(net, hipMemcpy(h.data(), net->p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    FILE* f = std::fopen(path, "wb");
    if (!f) return PCT_E_IO;
    const char magic[8] = {'P', 'C', 'T', 'N', 'E', 'T', '0', '1'};
    int64_t n = net->n_params;
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&n, sizeof(n), 1, f) == 1 &&
              std::fwrite(h.data(), sizeof(float), h.size(), f) == h.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? PCT_OK : PCT_E_IO;
}

int pct_net_load(pct_net* net, const char* path)
{
    if (!net || !path) return PCT_E_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return PCT_E_IO;
    char magic[8];
    int64_t n = 0;
    std::vector<float> h(net->n_params);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "PCTNET01", 8) == 0 &&
              std::fread(&n, sizeof(n), 1, f) == 1 && n == net->n_params &&
              std::fread(h.data(), sizeof(float), h.size(), f) == h.size();
    std::fclose(f);
    if (!ok) {
        net->err = "not a weight file of this network";
        return PCT_E_IO;
    }
    return pct_net_set_params(net, h.data(), n);
}

// x0 := net(FBP, hull) on the device of ctx; the FBP-based x0 is kept instead when the net's x0 has
// the larger DROP data residual ||Ax - b|| (needs the MLP lists).
int pct_net_apply(pct_net* net, pct_ctx* ctx, int* used_net, double* residual_fbp, double* residual_net)
{
    if (!net || !ctx) return PCT_E_ARG;
    const pct_params& P = ctx->P;
    if (P.columns != net->W || P.rows != net->H || P.slices != net->D) {
        net->err = "network grid differs from the reconstruction grid";
        return PCT_E_ARG;
    }
    if (ctx->stage < ST_LISTS || !ctx->fbp || !ctx->hull || !ctx->x0 || !ctx->x) {
        net->err = "pct_net_apply needs FBP, hull, x0 and the MLP lists";
        return PCT_E_STATE;
    }
    const int64_t n0 = net->nv[0];
    // the context's stream finishes its work before the net reads FBP and hull
    NET_HIP(net, hipStreamSynchronize(ctx->stream));
    PCT_TRY(net_forward(net, ctx->fbp, ctx->hull, nullptr, nullptr));
    NET_HIP(net, hipStreamSynchronize(net->stream));
    // copies into the context run on its stream, ordered with its residual kernels
    hipStream_t cs = ctx->stream;
    double r_fbp = 0, r_net = 0;
    NET_HIP(net, hipMemcpyAsync(ctx->x, ctx->x0, n0 * sizeof(float), hipMemcpyDeviceToDevice, cs));
    if (pct_stage_residual(ctx, &r_fbp) != PCT_OK) {
        net->err = ctx->err;
        return PCT_E_HIP;
    }
    NET_HIP(net, hipMemcpyAsync(ctx->x, net->x0, n0 * sizeof(float), hipMemcpyDeviceToDevice, cs));
    if (pct_stage_residual(ctx, &r_net) != PCT_OK) {
        net->err = ctx->err;
        return PCT_E_HIP;
    }
    const int use = r_net < r_fbp;
    if (use) {
        NET_HIP(net, hipMemcpyAsync(ctx->x0, net->x0, n0 * sizeof(float), hipMemcpyDeviceToDevice, cs));
    } else {
        NET_HIP(net, hipMemcpyAsync(ctx->x, ctx->x0, n0 * sizeof(float), hipMemcpyDeviceToDevice, cs));
    }
    NET_HIP(net, hipStreamSynchronize(cs));
    if (used_net) *used_net = use;
    if (residual_fbp) *residual_fbp = r_fbp;
    if (residual_net) *residual_net = r_net;
    return PCT_OK;
}

int pct_net_apply_iterate(pct_net* net, pct_ctx* ctx, int* used_net, double* residual_x, double* residual_net)
{
    if (!net || !ctx) return PCT_E_ARG;
    const pct_params& P = ctx->P;
    if (P.columns != net->W || P.rows != net->H || P.slices != net->D) {
        net->err = "network grid differs from the reconstruction grid";
        return PCT_E_ARG;
    }
    if (ctx->stage < ST_LISTS || !ctx->hull || !ctx->x) {
        net->err = "pct_net_apply_iterate needs the hull, an iterate and the MLP lists";
        return PCT_E_STATE;
    }
    const int64_t n0 = net->nv[0];
    NET_HIP(net, hipStreamSynchronize(ctx->stream));
    PCT_TRY(net_forward(net, ctx->x, ctx->hull, nullptr, nullptr));
    NET_HIP(net, hipStreamSynchronize(net->stream));
    hipStream_t cs = ctx->stream;
    double r_x = 0, r_net = 0;
    if (pct_stage_residual(ctx, &r_x) != PCT_OK) {
        net->err = ctx->err;
        return PCT_E_HIP;
    }
    NET_HIP(net, hipMemcpyAsync(net->fbp, ctx->x, n0 * sizeof(float), hipMemcpyDeviceToDevice, cs));  // keep x
    NET_HIP(net, hipMemcpyAsync(ctx->x, net->x0, n0 * sizeof(float), hipMemcpyDeviceToDevice, cs));
    if (pct_stage_residual(ctx, &r_net) != PCT_OK) {
        net->err = ctx->err;
        return PCT_E_HIP;
    }
    const int use = r_net < r_x;
    if (!use) NET_HIP(net, hipMemcpyAsync(ctx->x, net->fbp, n0 * sizeof(float), hipMemcpyDeviceToDevice, cs));
    NET_HIP(net, hipStreamSynchronize(cs));
    if (used_net) *used_net = use;
    if (residual_x) *residual_x = r_x;
    if (residual_net) *residual_net = r_net;
    return PCT_OK;
}

}  // extern "C"
