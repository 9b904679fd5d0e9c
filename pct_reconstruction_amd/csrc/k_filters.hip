// k_filters.hip — optional image filters of the reference's preprocessing.
//
//   median filter of the FBP image (median_filter_2D, DS.cu:7482-7519; applied in FBP() :6150-6166):
//     for every voxel whose (2r+1) x (2r+1) in-slice window lies inside the image, the element at
//     position (2r+1)^2 / 2 of the sorted window; every other voxel is 0 (the output array is
//     zero-initialised and fully copied back).  The filtered image feeds the initial iterate only;
//     the FBP hull is thresholded from the unfiltered image (FBP_image_2_hull uses the device copy).
//   averaging filter of the hull (averaging_filter_GPU, DS.cu:7420-7443; hull_selection :7375-7386):
//     window [x-r, x+r] x [y-r, y+r] clipped to the slice, n = its voxel count; the voxel is kept if
//     the window sum > n * HULL_AVG_FILTER_THRESHOLD (double).  The filtered hull replaces the device
//     hull (MLP endpoints); the initial iterate keeps the unfiltered host hull.
//
// The median runs per 16 x 16 tile of a slice with the (16 + 2r)^2 input patch in LDS; each lane
// selects its median by rank: the window element v with #(w < v) <= mid < #(w <= v) is exactly the
// element a sort puts at position mid.  For r <= 5 each lane holds its window in registers
// (k_median2d_fixed); larger radii compare from LDS.
#include "internal.h"

namespace {

constexpr int kTile = 16;
constexpr int kMaxR = 10;

__global__ void __launch_bounds__(kTile* kTile) k_median2d(int C, int R, int r, const float* __restrict__ in,
                                                            float* __restrict__ out)
{
    __shared__ float patch[(kTile + 2 * kMaxR) * (kTile + 2 * kMaxR)];
    const int P = kTile + 2 * r;
    const int z = blockIdx.z;
    const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
    const float* slice = in + (int64_t)z * C * R;
    for (int t = threadIdx.y * kTile + threadIdx.x; t < P * P; t += kTile * kTile) {
        const int px = x0 - r + t % P, py = y0 - r + t / P;
        patch[t] = (px >= 0 && px < C && py >= 0 && py < R) ? slice[(int64_t)py * C + px] : 0.0f;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= C || y >= R) return;
    float res = 0.0f;
    if (x >= r && x < C - r && y >= r && y < R - r) {
        const int w = 2 * r + 1, n = w * w, mid = n / 2;
        const int bx = threadIdx.x, by = threadIdx.y;  // window origin in the patch
        for (int i = 0; i < n; i++) {
            const float v = patch[(by + i / w) * P + bx + i % w];
            int lt = 0, le = 0;
            for (int j = 0; j < n; j++) {
                const float u = patch[(by + j / w) * P + bx + j % w];
                lt += u < v;
                le += u <= v;
            }
            if (lt <= mid && mid < le) {
                res = v;
                break;
            }
        }
    }
    out[(int64_t)z * C * R + (int64_t)y * C + x] = res;
}

// k_median2d for a radius known at compile time (RR <= 5): each lane holds its (2RR+1)^2 window in
// registers and ranks the candidates, taken from LDS in the same order, against it.  The first
// candidate with lt <= mid < le is the one k_median2d returns, so the images are identical.
template <int RR>
__global__ void __launch_bounds__(kTile* kTile) k_median2d_fixed(int C, int R, const float* __restrict__ in,
                                                                 float* __restrict__ out)
{
    constexpr int W = 2 * RR + 1, N = W * W, MID = N / 2, P = kTile + 2 * RR;
    __shared__ float patch[P * P];
    const int z = blockIdx.z;
    const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
    const float* slice = in + (int64_t)z * C * R;
    for (int t = threadIdx.y * kTile + threadIdx.x; t < P * P; t += kTile * kTile) {
        const int px = x0 - RR + t % P, py = y0 - RR + t / P;
        patch[t] = (px >= 0 && px < C && py >= 0 && py < R) ? slice[(int64_t)py * C + px] : 0.0f;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= C || y >= R) return;
    float res = 0.0f;
    if (x >= RR && x < C - RR && y >= RR && y < R - RR) {
        const float* wp = patch + threadIdx.y * P + threadIdx.x;
        float win[N];
#pragma unroll
        for (int j = 0; j < N; j++) win[j] = wp[(j / W) * P + j % W];
        for (int i = 0; i < N; i++) {
            const float v = wp[(i / W) * P + i % W];
            int lt = 0, le = 0;
#pragma unroll
            for (int j = 0; j < N; j++) {
                lt += win[j] < v;
                le += win[j] <= v;
            }
            if (lt <= MID && MID < le) {
                res = v;
                break;
            }
        }
    }
    out[(int64_t)z * C * R + (int64_t)y * C + x] = res;
}

__global__ void k_avg_hull(int C, int R, int64_t nv, int r, double thr, const uint8_t* __restrict__ in,
                           uint8_t* __restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int x = (int)(v % C), y = (int)((v / C) % R);
    const int64_t z = v / ((int64_t)C * R);
    const int left = x - r > 0 ? x - r : 0, right = x + r < C - 1 ? x + r : C - 1;
    const int top = y - r > 0 ? y - r : 0, bottom = y + r < R - 1 ? y + r : R - 1;
    const int n = (right - left + 1) * (bottom - top + 1);
    const double sum_threshold = n * thr;
    double sum = 0.0;
    for (int c = left; c <= right; c++)
        for (int rr = top; rr <= bottom; rr++) sum += in[z * C * R + (int64_t)rr * C + c] ? 1.0 : 0.0;
    out[v] = sum > sum_threshold ? 1 : 0;
}

}  // namespace

int pct_median_filter_fbp(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    const int r = P.fbp_med_filter_radius;
    if (!(P.median_filter_fbp && r > 0 && !P.avg_filter_fbp)) {
        pct_free(&ctx->fbp_x0);
        return PCT_OK;
    }
    if (r > kMaxR) {
        ctx->err = "FBP_MED_FILTER_RADIUS above 10";
        return PCT_E_LIMIT;
    }
    PCT_TRY(pct_alloc(ctx, &ctx->fbp_x0, P.num_voxels));
    dim3 grid((P.columns + kTile - 1) / kTile, (P.rows + kTile - 1) / kTile, P.slices), block(kTile, kTile);
    switch (r) {
    case 1: hipLaunchKernelGGL(k_median2d_fixed<1>, grid, block, 0, ctx->stream, P.columns, P.rows, ctx->fbp, ctx->fbp_x0); break;
    case 2: hipLaunchKernelGGL(k_median2d_fixed<2>, grid, block, 0, ctx->stream, P.columns, P.rows, ctx->fbp, ctx->fbp_x0); break;
    case 3: hipLaunchKernelGGL(k_median2d_fixed<3>, grid, block, 0, ctx->stream, P.columns, P.rows, ctx->fbp, ctx->fbp_x0); break;
    case 4: hipLaunchKernelGGL(k_median2d_fixed<4>, grid, block, 0, ctx->stream, P.columns, P.rows, ctx->fbp, ctx->fbp_x0); break;
    case 5: hipLaunchKernelGGL(k_median2d_fixed<5>, grid, block, 0, ctx->stream, P.columns, P.rows, ctx->fbp, ctx->fbp_x0); break;
    default:
        hipLaunchKernelGGL(k_median2d, grid, block, 0, ctx->stream, P.columns, P.rows, r, ctx->fbp, ctx->fbp_x0);
    }
    PCT_CHECK_HIP(ctx, hipGetLastError());
    return PCT_OK;
}

int pct_avg_filter_hull(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    const int r = P.hull_avg_filter_radius;
    if (!(P.avg_filter_hull && r > 0)) {
        pct_free(&ctx->hull_endpoints);
        return PCT_OK;
    }
    PCT_TRY(pct_alloc(ctx, &ctx->hull_endpoints, P.num_voxels));
    const int64_t nv = P.num_voxels;
    hipLaunchKernelGGL(k_avg_hull, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, ctx->stream, P.columns, P.rows, nv,
                       r, P.hull_avg_filter_threshold, ctx->hull, ctx->hull_endpoints);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    return PCT_OK;
}
