// k_tvs.hip — total variation superiorization between DROP iterations.
//
// NTVS_iteration (DS.cu:10583) -> iteratively_perturb_image (:10476) or, unconditioned,
// iteratively_perturb_image_unconditional (:10562), called by DROP_GPU (:9889-9915) before
// (TVS_FIRST) or after each DROP iteration k = 1, 2, ...:
//   L = draw in [L_0 + k - 1, L]                       (randi :11478; not under TVS_OLD)
//   repeat TVS_REPETITIONS times (once under TVS_OLD):
//     G = subgradient of TV(x), v = G / ||G||          (generate_perturbation_array :10262)
//     BETA = A^L; TV_x = TV(x); x_TVS = x; x = x_TVS - BETA v on hull voxels
//     loop: L++; TV_y = TV(x); stop if TV_y <= TV_x, |TV_y - TV_prev| / TV_y < TV_THRESHOLD or
//           L > 300; else BETA = A^L, TV_prev = TV_y, x = x_TVS - BETA v on hull voxels
//
// The image stays in its linear layout.  x holds x_TVS for the whole repetition: the candidate
// x_TVS - BETA v is never stored while it is being tried; k_tvs_candidate_tv sums the TV of the
// candidate computed on the fly, and k_tvs_apply writes the accepted one.
//
// Reductions: the reference sums the TV and ||G||^2 serially into a float.  Here both are FP64 sums in
// a fixed order — tiles of 1024 voxels, thread j of 256 adding voxels j, j+256, j+512, j+768 in
// turn, a halving tree over the 256 sums; the tile sums reduced the same way by one workgroup — and
// then rounded to float where the reference holds a float.  The result is reproducible run to run.
// G itself, the TV terms, v and every image update are the reference's float operations.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kTile = 1024;

__device__ __forceinline__ void tree256(double* a)
{
    for (int s = 128; s >= 1; s >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < s) a[threadIdx.x] = a[threadIdx.x] + a[threadIdx.x + s];
    }
    __syncthreads();
}

// ||(gx, gy)|| of generate_perturbation_array step 1 (:10281): squares and sum in double, sqrt in
// double, stored as float.
__device__ __forceinline__ float grad_norm(float gx, float gy)
{
    return (float)sqrt((double)gx * (double)gx + (double)gy * (double)gy);
}

// G[w] of generate_perturbation_array step 2 (:10287-10303).  The reference adds to G[w] from the
// sources w - COLUMNS, w - 1 and w, in that order (its loop visits voxels in increasing index).
__device__ __forceinline__ float subgradient(const float* __restrict__ x, int64_t w, int col, int row, int C, int R)
{
    float g = 0.0f;
    if (row >= 1 && col < C - 1) {  // source w - C: G[src + C] += gy / n
        const int64_t s = w - C;
        const float gx = x[s + 1] - x[s], gy = x[w] - x[s];
        const float n = grad_norm(gx, gy);
        if (n > 0.0f) g += gy / n;
    }
    if (col >= 1 && row < R - 1) {  // source w - 1: G[src + 1] += gx / n
        const int64_t s = w - 1;
        const float gx = x[w] - x[s], gy = x[s + C] - x[s];
        const float n = grad_norm(gx, gy);
        if (n > 0.0f) g += gx / n;
    }
    if (col < C - 1 && row < R - 1) {  // source w: G[w] -= (gx + gy) / n
        const float gx = x[w + 1] - x[w], gy = x[w + C] - x[w];
        const float n = grad_norm(gx, gy);
        if (n > 0.0f) g -= (gx + gy) / n;
    }
    return g;
}

// Column and row of voxel w (w % C, (w / C) % R), and the step to voxel w + kThreads: one
// division per lane instead of two 64-bit divisions per voxel.
struct ColRow {
    int col, row;
    __device__ ColRow(int64_t w, int C, int R) : col((int)(w % C)), row((int)((w / C) % R)) {}
    __device__ void advance(int by, int C, int R)
    {
        col += by;
        while (col >= C) {
            col -= C;
            if (++row == R) row = 0;
        }
    }
};

// One term of calculate_total_variation (:10178): sqrt(powf(dy, 2) + powf(dx, 2)) in float.
__device__ __forceinline__ float tv_term(float c, float right, float below)
{
    const float dy = below - c, dx = right - c;
    return sqrtf(dy * dy + dx * dx);
}

// Per tile: G (when GRAD) and the FP64 tile sums of G^2 and of the TV terms of x.
template <bool GRAD>
__global__ void __launch_bounds__(kThreads) k_tvs_grad(int C, int R, int64_t nv, const float* __restrict__ x,
                                                       float* __restrict__ G, double* __restrict__ part_norm,
                                                       double* __restrict__ part_tv)
{
    __shared__ double a[kThreads], b[kThreads];
    double sn = 0.0, st = 0.0;
    const int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x;
    ColRow cr(base, C, R);
    for (int i = 0; i < 4; i++, cr.advance(kThreads, C, R)) {
        const int64_t w = base + (int64_t)kThreads * i;
        if (w >= nv) break;
        const int col = cr.col, row = cr.row;
        if (GRAD) {
            const float g = subgradient(x, w, col, row, C, R);
            G[w] = g;
            sn += (double)g * (double)g;
        }
        const float t = (col < C - 1 && row < R - 1) ? tv_term(x[w], x[w + 1], x[w + C]) : 0.0f;
        st += (double)t;
    }
    a[threadIdx.x] = sn;
    b[threadIdx.x] = st;
    tree256(a);
    tree256(b);
    if (threadIdx.x == 0) {
        if (GRAD) part_norm[blockIdx.x] = a[0];
        part_tv[blockIdx.x] = b[0];
    }
}

// y = x - BETA * v on hull voxels, x elsewhere; v = G / norm (0 when norm is 0) — the float operations
// of generate_perturbation_array step 4 (:10313-10320) and apply_TVS_perturbation (:10417).
__device__ __forceinline__ float perturbed(const float* __restrict__ x, const float* __restrict__ G,
                                           const uint8_t* __restrict__ hull, int64_t w, float norm, float beta)
{
    const float xv = x[w];
    if (!hull[w]) return xv;
    const float v = norm != 0.0f ? G[w] / norm : 0.0f;
    return xv - beta * v;
}

// Tile sums of the TV terms of the candidate image x - BETA v.
__global__ void __launch_bounds__(kThreads) k_tvs_candidate_tv(int C, int R, int64_t nv, const float* __restrict__ x,
                                                               const float* __restrict__ G,
                                                               const uint8_t* __restrict__ hull, float norm,
                                                               float beta, double* __restrict__ part_tv)
{
    __shared__ double a[kThreads];
    double st = 0.0;
    const int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x;
    ColRow cr(base, C, R);
    for (int i = 0; i < 4; i++, cr.advance(kThreads, C, R)) {
        const int64_t w = base + (int64_t)kThreads * i;
        if (w >= nv) break;
        const int col = cr.col, row = cr.row;
        float t = 0.0f;
        if (col < C - 1 && row < R - 1)
            t = tv_term(perturbed(x, G, hull, w, norm, beta), perturbed(x, G, hull, w + 1, norm, beta),
                        perturbed(x, G, hull, w + C, norm, beta));
        st += (double)t;
    }
    a[threadIdx.x] = st;
    tree256(a);
    if (threadIdx.x == 0) part_tv[blockIdx.x] = a[0];
}

__global__ void k_tvs_apply(int64_t nv, float* __restrict__ x, const float* __restrict__ G,
                            const uint8_t* __restrict__ hull, float norm, float beta)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nv && hull[w]) {
        const float v = norm != 0.0f ? G[w] / norm : 0.0f;  // the operations of perturbed()
        x[w] = x[w] - beta * v;
    }
}

// Second level of the fixed-order sum: out[i] = sum of part[i * stride .. i * stride + nt) for the
// arrays selected by `count` (1 or 2).
__global__ void __launch_bounds__(kThreads) k_tvs_final(int64_t nt, int count, const double* __restrict__ part,
                                                        int64_t stride, double* __restrict__ out)
{
    __shared__ double a[kThreads];
    for (int q = 0; q < count; q++) {
        const double* p = part + q * stride;
        double s = 0.0;
        for (int64_t i = threadIdx.x; i < nt; i += kThreads) s += p[i];
        a[threadIdx.x] = s;
        tree256(a);
        if (threadIdx.x == 0) out[q] = a[0];
        __syncthreads();
    }
}

uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// L = randi(L_0 + k - 1, L) (DS.cu:10481).  The reference seeds std::default_random_engine from the
// clock; this is a seeded SplitMix64 draw, uniform over the same range.  hi < lo gives lo.
int draw_L(int seed, int k, int lo, int hi)
{
    if (hi <= lo) return lo;
    const uint64_t r = splitmix64(((uint64_t)(uint32_t)seed << 32) ^ (uint64_t)(uint32_t)k);
    return lo + (int)(r % (uint64_t)(hi - lo + 1));
}

// TVS_eta_sequence_h[L - L_0] (generate_TVS_eta_sequence :10149-10156): powf(A, L) for
// L_0 <= L < length = (int)(log(TVS_MIN_ETA) / log(A)), 0 at L = length (the calloc'd last element)
// and beyond it (the reference reads past its array there).
float eta(const pct_params& P, int L)
{
    const int len = (int)(std::log(P.tvs_min_eta) / std::log(P.tvs_a));
    if (L < P.tvs_l0 || L >= len) return 0.0f;
    return (float)std::pow((double)(float)P.tvs_a, (double)(float)L);
}

struct Tvs {
    pct_ctx* ctx;
    int C, R;
    int64_t nv, nt;
    hipStream_t s;

    int ensure()
    {
        if (ctx->tvs_G && ctx->tvs_nt == nt) return PCT_OK;
        PCT_TRY(pct_alloc(ctx, &ctx->tvs_G, (size_t)nv));
        PCT_TRY(pct_alloc(ctx, &ctx->tvs_part, (size_t)(2 * nt)));
        PCT_TRY(pct_alloc(ctx, &ctx->tvs_res, 2));
        if (!ctx->tvs_res_h) PCT_CHECK_HIP(ctx, hipHostMalloc((void**)&ctx->tvs_res_h, 2 * sizeof(double)));
        ctx->tvs_nt = nt;
        return PCT_OK;
    }
    int fetch(int count)
    {
        hipLaunchKernelGGL(k_tvs_final, dim3(1), dim3(kThreads), 0, s, nt, count, ctx->tvs_part, nt, ctx->tvs_res);
        PCT_CHECK_HIP(ctx, hipGetLastError());
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->tvs_res_h, ctx->tvs_res, count * sizeof(double), hipMemcpyDeviceToHost, s));
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
        return PCT_OK;
    }
    // G of x, ||G|| (float) and TV(x) (float).
    int grad(float* norm, float* tv)
    {
        hipLaunchKernelGGL(k_tvs_grad<true>, dim3((unsigned)nt), dim3(kThreads), 0, s, C, R, nv, ctx->x, ctx->tvs_G,
                           ctx->tvs_part, ctx->tvs_part + nt);
        PCT_TRY(fetch(2));
        *norm = sqrtf((float)ctx->tvs_res_h[0]);
        *tv = (float)ctx->tvs_res_h[1];
        return PCT_OK;
    }
    // TV(x): the FP64 sum (rounded to float where the reference keeps a float).
    int tv_of_x(double* tv)
    {
        hipLaunchKernelGGL(k_tvs_grad<false>, dim3((unsigned)nt), dim3(kThreads), 0, s, C, R, nv, ctx->x, nullptr,
                           nullptr, ctx->tvs_part);
        PCT_TRY(fetch(1));
        *tv = ctx->tvs_res_h[0];
        return PCT_OK;
    }
    int candidate_tv(float norm, float beta, float* tv)
    {
        hipLaunchKernelGGL(k_tvs_candidate_tv, dim3((unsigned)nt), dim3(kThreads), 0, s, C, R, nv, ctx->x,
                           ctx->tvs_G, ctx->hull, norm, beta, ctx->tvs_part);
        PCT_TRY(fetch(1));
        *tv = (float)ctx->tvs_res_h[0];
        return PCT_OK;
    }
    int apply(float norm, float beta)
    {
        const unsigned g = (unsigned)((nv + kThreads - 1) / kThreads);
        hipLaunchKernelGGL(k_tvs_apply, dim3(g), dim3(kThreads), 0, s, nv, ctx->x, ctx->tvs_G, ctx->hull, norm, beta);
        PCT_CHECK_HIP(ctx, hipGetLastError());
        return PCT_OK;
    }
};

Tvs make_tvs(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    Tvs t{ctx, P.columns, P.rows, (int64_t)P.num_voxels, ((int64_t)P.num_voxels + kTile - 1) / kTile, ctx->stream};
    return t;
}

}  // namespace

// One NTVS_iteration on ctx->x for DROP iteration k (1-based); appends a record to ctx->tvs_log.
int pct_tvs_iteration(pct_ctx* ctx, int k)
{
    const pct_params& P = ctx->P;
    if (!ctx->hull || !ctx->x) {
        ctx->err = "TVS needs the hull and the iterate";
        return PCT_E_STATE;
    }
    Tvs t = make_tvs(ctx);
    PCT_TRY(t.ensure());
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PCT_CHECK_HIP(ctx, hipEventCreate(&e0));
    PCT_CHECK_HIP(ctx, hipEventCreate(&e1));
    PCT_CHECK_HIP(ctx, hipEventRecord(e0, t.s));

    int L = std::max(ctx->tvs_L, P.tvs_l0);
    const int reps = P.tvs_old ? 1 : P.tvs_repetitions;
    if (!P.tvs_old) L = draw_L(P.tvs_seed, k, P.tvs_l0 + k - 1, L);
    float norm = 0.0f, tv_x = 0.0f, tv_before = 0.0f, tv_after = 0.0f;
    double tv64 = 0.0;
    if (reps == 0) {
        PCT_TRY(t.tv_of_x(&tv64));
        tv_before = tv_after = (float)tv64;
    } else if (P.tvs_conditioned) {
        for (int n = 0; n < reps; n++) {
            PCT_TRY(t.grad(&norm, &tv_x));
            if (n == 0) tv_before = tv_x;
            float beta = eta(P, L), tv_prev = 0.0f, tv_y = 0.0f;
            for (;;) {
                L++;
                PCT_TRY(t.candidate_tv(norm, beta, &tv_y));
                if (tv_y <= tv_x || (double)(std::fabs(tv_y - tv_prev) / tv_y) < P.tv_threshold || L > 300) break;
                beta = eta(P, L);
                tv_prev = tv_y;
            }
            PCT_TRY(t.apply(norm, beta));
            tv_after = tv_y;
        }
    } else {
        for (int n = 0; n < reps; n++, L++) {
            PCT_TRY(t.grad(&norm, &tv_x));
            if (n == 0) tv_before = tv_x;
            PCT_TRY(t.apply(norm, eta(P, L)));
        }
        PCT_TRY(t.tv_of_x(&tv64));
        tv_after = (float)tv64;
    }
    ctx->tvs_L = L;
    PCT_CHECK_HIP(ctx, hipEventRecord(e1, t.s));
    PCT_CHECK_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->tvs_log.push_back({(double)tv_before, (double)tv_after, L, (double)ms});
    return PCT_OK;
}

int pct_stage_total_variation(pct_ctx* ctx, double* tv)
{
    Tvs t = make_tvs(ctx);
    PCT_TRY(t.ensure());
    return t.tv_of_x(tv);
}
