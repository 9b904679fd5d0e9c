// k_preproc.hip — stage (a): recon-volume intersections, binning, compaction, per-bin statistics,
// statistical cuts and sinogram sums.
//
// Layout: every per-history quantity is a separate float/int array in HBM (SoA), so each kernel
// streams only the fields it needs with fully coalesced 4-byte loads.  One lane per history.
//
// Determinism: the reference accumulates the per-bin sums with float atomicAdd (DS.cu:5543-5546,
// 5713-5715, 5968-5969), whose order is unspecified.  Here histories are stably sorted by bin
// (LSD radix sort keeps history order inside a bin) and each bin's sums are formed by one lane in
// ascending history order: the sums a single-threaded run of the reference produces, bit for bit,
// on every run.
#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) { return pct_grid_for(n, block); }

// convert_mm_2_cm (DS.cu:4070-4089) and apply_tuv_shifts (:4090-4110): float *= 0.1; float += shift.
__global__ void k_units(int64_t n, float* t0, float* t1, float* t2, float* t3, float* v0, float* v1, float* v2,
                        float* v3, float* u0, float* u1, float* u2, float* u3, float* wepl, int mm, double ts,
                        double us, double vs)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mm) {
        v0[i] = (float)((double)v0[i] * 0.1); v1[i] = (float)((double)v1[i] * 0.1);
        v2[i] = (float)((double)v2[i] * 0.1); v3[i] = (float)((double)v3[i] * 0.1);
        t0[i] = (float)((double)t0[i] * 0.1); t1[i] = (float)((double)t1[i] * 0.1);
        t2[i] = (float)((double)t2[i] * 0.1); t3[i] = (float)((double)t3[i] * 0.1);
        u0[i] = (float)((double)u0[i] * 0.1); u1[i] = (float)((double)u1[i] * 0.1);
        u2[i] = (float)((double)u2[i] * 0.1); u3[i] = (float)((double)u3[i] * 0.1);
        wepl[i] = (float)((double)wepl[i] * 0.1);
    }
    if (ts != 0.0 || us != 0.0 || vs != 0.0) {
        t0[i] = (float)((double)t0[i] + ts); t1[i] = (float)((double)t1[i] + ts);
        t2[i] = (float)((double)t2[i] + ts); t3[i] = (float)((double)t3[i] + ts);
        u0[i] = (float)((double)u0[i] + us); u1[i] = (float)((double)u1[i] + us);
        u2[i] = (float)((double)u2[i] + us); u3[i] = (float)((double)u3[i] + us);
        v0[i] = (float)((double)v0[i] + vs); v1[i] = (float)((double)v1[i] + vs);
        v2[i] = (float)((double)v2[i] + vs); v3[i] = (float)((double)v3[i] + vs);
    }
}

// calculate_intercepts, DS.cu:5286-5372
__device__ __forceinline__ int calc_intercepts(double R, double u, double t, double ut_angle, double* u_int, double* t_int)
{
    const double PI_OVER_4 = atan(1.0), THREE_PI_OVER_4 = 3 * atan(1.0);
    const double FIVE_PI_OVER_4 = 5 * atan(1.0), SEVEN_PI_OVER_4 = 7 * atan(1.0), PI_OVER_2 = 2 * atan(1.0);
    int in_cone = ((ut_angle > PI_OVER_4) && (ut_angle < THREE_PI_OVER_4)) ||
                  ((ut_angle > FIVE_PI_OVER_4) && (ut_angle < SEVEN_PI_OVER_4));
    if (in_cone) {
        double u_temp = u;
        u = -t;
        t = u_temp;
        ut_angle += PI_OVER_2;
    }
    double m = tan(ut_angle);
    double b_in = t - m * u;
    double a = (double)(1 + pct_sqf(m));
    double b = 2 * m * b_in;
    double c = (double)(pct_sqf(b_in) - pct_sqf(R));
    double disc = (double)pct_sqf(b) - (4 * a * c);
    int intersected = (disc > 0);
    if (intersected) {
        double sq = sqrt(disc);
        double u1 = (sq - b) / (2 * a);
        double u2 = (sq + b) / (2 * a);
        double t1 = m * u1 + b_in;
        double t2 = m * u2 - b_in;
        double d1 = (double)(pct_sqf(u1 - u) + pct_sqf(t1 - t));
        double d2 = (double)(pct_sqf(u2 + u) + pct_sqf(t2 + t));
        *u_int = u1 * (double)(d1 <= d2) - u2 * (double)(d1 > d2);
        *t_int = t1 * (double)(d1 <= d2) - t2 * (double)(d1 > d2);
    }
    if (in_cone) {
        double u_temp = *u_int;
        *u_int = *t_int;
        *t_int = -u_temp;
    }
    return intersected;
}

struct RawPtrs {
    const float *t0, *t1, *t2, *t3, *v0, *v1, *v2, *v3, *u0, *u1, *u2, *u3, *ang;
};
struct HPtrs {
    float* h[10];
};

// recon_volume_intersections_GPU, DS.cu:4890-5021 (float angles: the _exact_angles variant :5022,
// identical arithmetic for integer gantry angles).
__global__ void k_intersections(int64_t n, RawPtrs r, HPtrs o, uint8_t* missed, double R, double H, double A2R)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t0 = r.t0[i], t1 = r.t1[i], t2 = r.t2[i], t3 = r.t3[i];
    const float v0 = r.v0[i], v1 = r.v1[i], v2 = r.v2[i], v3 = r.v3[i];
    const float u0 = r.u0[i], u1 = r.u1[i], u2 = r.u2[i], u3 = r.u3[i];
    double rot = (double)r.ang[i] * A2R;
    double cr = cos(rot), sr = sin(rot);
    double ut_entry_angle = (double)pct_atan2f(t1 - t0, u1 - u0);
    double u_entry = 0.0, t_entry = 0.0;
    int entered = calc_intercepts(R, u1, t1, ut_entry_angle, &u_entry, &t_entry);
    float xya_e = (float)(ut_entry_angle + rot);
    float xe = (float)((cr * u_entry) - (sr * t_entry));
    float ye = (float)((sr * u_entry) + (cr * t_entry));
    double ut_exit_angle = (double)pct_atan2f(t3 - t2, u3 - u2);
    double u_exit = 0.0, t_exit = 0.0;
    int exited = calc_intercepts(R, u2, t2, ut_exit_angle, &u_exit, &t_exit);
    float xya_x = (float)(ut_exit_angle + rot);
    float xx = (float)((cr * u_exit) - (sr * t_exit));
    float yx = (float)((sr * u_exit) + (cr * t_exit));
    double uv_entry_slope = (double)((v1 - v0) / (u1 - u0));
    double uv_exit_slope = (double)((v3 - v2) / (u3 - u2));
    float xza_e = pct_atan2f(v1 - v0, u1 - u0);
    float xza_x = pct_atan2f(v3 - v2, u3 - u2);
    u_entry = (cr * (double)xe) + (sr * (double)ye);
    u_exit = (cr * (double)xx) + (sr * (double)yx);
    float ze = (float)((double)v1 + uv_entry_slope * (u_entry - (double)u1));
    float zx = (float)((double)v2 - uv_exit_slope * ((double)u2 - u_exit));
    if (entered && exited) {
        if ((fabsf(ze) < H * 0.5) && (fabsf(zx) > H * 0.5)) {
            double sgn = (double)((zx >= 0) - (zx < 0));
            double frac = fabs((sgn * H * 0.5 - ze) / (double)(zx - ze));
            xx = (float)((double)xe + frac * (double)(xx - xe));
            yx = (float)((double)ye + frac * (double)(yx - ye));
            zx = (float)(sgn * H * 0.5);
        } else if (fabsf(ze) > H * 0.5) {
            entered = 0;
            exited = 0;
        }
        if ((fabsf(ze) > H * 0.5) && (fabsf(zx) < H * 0.5)) {
            double sgn_x = (double)((zx >= 0) - (zx < 0));
            double sgn_e = (double)((ze >= 0) - (ze < 0));
            double frac = fabs((sgn_x * H * 0.5 - zx) / (double)(zx - ze));
            xe = (float)((double)xx + frac * (double)(xx - xe));
            ye = (float)((double)yx + frac * (double)(yx - ye));
            ze = (float)(sgn_e * H * 0.5);
        }
        if ((fabsf(t2 - t1) > 5) || (fabsf(v2 - v1) > 5)) {
            entered = 0;
            exited = 0;
        }
    }
    o.h[0][i] = xe; o.h[1][i] = ye; o.h[2][i] = ze;
    o.h[3][i] = xx; o.h[4][i] = yx; o.h[5][i] = zx;
    o.h[6][i] = xya_e; o.h[7][i] = xza_e; o.h[8][i] = xya_x; o.h[9][i] = xza_x;
    missed[i] = (uint8_t)(!entered || !exited);
}

// binning_GPU, DS.cu:5482-5553 (bin number only; sums are formed per bin below).
// cos/sin of the angle-bin centres come from a host table so that the values equal the oracle's.
__global__ void k_binning(int64_t n, HPtrs hp, const uint8_t* missed, int32_t* bin_num, uint8_t* in_range,
                          const double* bin_cos, const double* bin_sin, double PI, double abs_deg, double t_bin_size,
                          double v_bin_size, int T, int V, int A)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (missed[i]) {
        bin_num[i] = 0;
        in_range[i] = 0;
        return;
    }
    const float xe = hp.h[0][i], ye = hp.h[1][i], ze = hp.h[2][i], xx = hp.h[3][i], yx = hp.h[4][i], zx = hp.h[5][i];
    double x_mid = (double)((xe + xx) / 2);
    double y_mid = (double)((ye + yx) / 2);
    double z_mid = (double)((ze + zx) / 2);
    double path_angle = (double)pct_atan2f(yx - ye, xx - xe);
    if (path_angle < 0) path_angle += 2 * PI;
    int angle_bin = (int)((path_angle * (180.0 / PI) / abs_deg) + 0.5) % A;
    double tt = y_mid * bin_cos[angle_bin] - x_mid * bin_sin[angle_bin];
    int t_bin = (int)((tt / t_bin_size) + T / 2);
    int v_bin = (int)((z_mid / v_bin_size) + V / 2);
    if ((t_bin >= 0) && (v_bin >= 0) && (t_bin < T) && (v_bin < V)) {
        bin_num[i] = t_bin + angle_bin * T + v_bin * T * A;
        in_range[i] = 1;
    } else {
        bin_num[i] = 0;  // quirk A1 (DS.cu:5380, 5431)
        in_range[i] = 0;
    }
}

__global__ void k_not(int64_t n, const uint8_t* in, uint8_t* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint8_t)!in[i];
}

template <class T>
__global__ void k_scatter(int64_t n, const uint8_t* keep, const int64_t* pos, const T* in, T* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && keep[i]) out[pos[i]] = in[i];
}

__global__ void k_iota(int64_t n, uint32_t* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)i;
}

__global__ void k_bin_hist(int64_t n, const int32_t* bin_num, int32_t* counts)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&counts[bin_num[i]], 1);
}

__device__ __forceinline__ double wrap_pi(double a, double PI)
{
    if (a > PI) a -= 2 * PI;
    if (a < -PI) a += 2 * PI;
    return a;
}

// The per-history inputs of the bin statistics in bin order (j = position in the sorted order): in-range
// flag, WEPL and the two wrapped angle differences, so that the three passes of k_bin_stats stream
// contiguous memory instead of gathering through the order.
__global__ void k_bin_gather(int64_t n, const uint32_t* __restrict__ perm, const uint8_t* __restrict__ in_range,
                             const float* __restrict__ wepl, const float* __restrict__ xya_e,
                             const float* __restrict__ xza_e, const float* __restrict__ xya_x,
                             const float* __restrict__ xza_x, double PI, uint8_t* __restrict__ g_ir,
                             float* __restrict__ g_w, double* __restrict__ g_ut, double* __restrict__ g_uv)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = perm[j];
    g_ir[j] = in_range[i];
    g_w[j] = wepl[i];
    g_ut[j] = wrap_pi((double)(xya_x[i] - xya_e[i]), PI);
    g_uv[j] = wrap_pi((double)(xza_x[i] - xza_e[i]), PI);
}

// One lane per bin: binning sums (DS.cu:5518-5546), calculate_means_GPU (:5644), the squared
// deviations (:5689-5717), calculate_standard_deviations_GPU (:5737) and statistical_cuts_GPU
// (:5937-5971), each over the bin's histories in ascending history order, read from the bin-ordered
// inputs of k_bin_gather.
__global__ void k_bin_stats(int NB, const int32_t* off, const uint32_t* perm, const uint8_t* __restrict__ g_ir,
                            const float* __restrict__ g_w, const double* __restrict__ g_ut,
                            const double* __restrict__ g_uv, int32_t* counts_pre, float* m_w, float* m_ut,
                            float* m_uv, float* s_w, float* s_ut, float* s_uv, int32_t* counts_post, float* sinogram,
                            uint8_t* failed, float k, int quirk_a1)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= NB) return;
    const int32_t beg = off[b], end = off[b + 1];
    int cnt = 0;
    float sw = 0.0f, su = 0.0f, sv = 0.0f;
    for (int32_t j = beg; j < end; j++) {
        if (!g_ir[j]) continue;
        cnt += 1;
        sw += g_w[j];
        su += (float)g_ut[j];
        sv += (float)g_uv[j];
    }
    if (cnt > 0) {
        sw /= cnt;
        su /= cnt;
        sv /= cnt;
    }
    counts_pre[b] = cnt;
    m_w[b] = sw;
    m_ut[b] = su;
    m_uv[b] = sv;
    float qw = 0.0f, qu = 0.0f, qv = 0.0f;
    for (int32_t j = beg; j < end; j++) {
        if (!quirk_a1 && !g_ir[j]) continue;
        double d_w = (double)(g_w[j] - sw);
        double d_ut = g_ut[j] - su;
        double d_uv = g_uv[j] - sv;
        qw += pct_sqf(d_w);
        qu += pct_sqf(d_ut);
        qv += pct_sqf(d_uv);
    }
    if (cnt > 1) {
        qw = sqrtf(qw / (cnt - 1));
        qu = sqrtf(qu / (cnt - 1));
        qv = sqrtf(qv / (cnt - 1));
    }
    s_w[b] = qw;
    s_ut[b] = qu;
    s_uv[b] = qv;
    int cpost = 0;
    float sino = 0.0f;
    for (int32_t j = beg; j < end; j++) {
        const uint32_t i = perm[j];
        if (!quirk_a1 && !g_ir[j]) {
            failed[i] = 1;
            continue;
        }
        const float w = g_w[j];
        int pass_ut = fabs(g_ut[j] - su) < (double)(k * qu);
        int pass_uv = fabs(g_uv[j] - sv) < (double)(k * qv);
        int pass_w = fabsf(sw - w) <= k * qw;
        int f = !pass_ut || !pass_uv || !pass_w;
        failed[i] = (uint8_t)f;
        if (!f) {
            cpost += 1;
            sino += w;
        }
    }
    counts_post[b] = cpost;
    sinogram[b] = sino;
}

}  // namespace

int pct_compact(pct_ctx* ctx, const uint8_t* keep, int64_t n, int64_t* n_out, bool with_first_voxel)
{
    hipStream_t s = ctx->stream;
    ctx->bin_perm_n = -1;  // the histories move
    if (n == 0) {
        *n_out = 0;
        return PCT_OK;
    }
    // exclusive scan of keep flags (as int64) -> positions
    size_t need = 0;
    hipcub::TransformInputIterator<int64_t, hipcub::CastOp<int64_t>, const uint8_t*> it(keep, hipcub::CastOp<int64_t>());
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, it, ctx->pos, n, s));
    if (need > ctx->cub_tmp_bytes) {
        PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, need));
        ctx->cub_tmp_bytes = need;
    }
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, need, it, ctx->pos, n, s));
    int64_t last_pos = 0;
    uint8_t last_keep = 0;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&last_pos, ctx->pos + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&last_keep, keep + n - 1, 1, hipMemcpyDeviceToHost, s));
    unsigned g = grid_for(n);
    for (int k = 0; k < 10; k++) {
        hipLaunchKernelGGL(k_scatter<float>, dim3(g), dim3(kBlock), 0, s, n, keep, ctx->pos, ctx->h[k], ctx->h_tmp[k]);
        std::swap(ctx->h[k], ctx->h_tmp[k]);
    }
    hipLaunchKernelGGL(k_scatter<float>, dim3(g), dim3(kBlock), 0, s, n, keep, ctx->pos, ctx->wepl, ctx->wepl_tmp);
    std::swap(ctx->wepl, ctx->wepl_tmp);
    hipLaunchKernelGGL(k_scatter<float>, dim3(g), dim3(kBlock), 0, s, n, keep, ctx->pos, ctx->gantry, ctx->gantry_tmp);
    std::swap(ctx->gantry, ctx->gantry_tmp);
    hipLaunchKernelGGL(k_scatter<int32_t>, dim3(g), dim3(kBlock), 0, s, n, keep, ctx->pos, ctx->bin_num, ctx->bin_tmp);
    std::swap(ctx->bin_num, ctx->bin_tmp);
    hipLaunchKernelGGL(k_scatter<uint8_t>, dim3(g), dim3(kBlock), 0, s, n, keep, ctx->pos, ctx->in_range, ctx->u8_tmp);
    std::swap(ctx->in_range, ctx->u8_tmp);
    if (with_first_voxel) {
        hipLaunchKernelGGL(k_scatter<uint32_t>, dim3(g), dim3(kBlock), 0, s, n, keep, ctx->pos, ctx->first_voxel,
                           ctx->u32_d);
        std::swap(ctx->first_voxel, ctx->u32_d);
    }
    PCT_CHECK_HIP(ctx, hipGetLastError());
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    *n_out = last_pos + (last_keep ? 1 : 0);
    return PCT_OK;
}

// Stable sort of the indices of the first n histories by bin number into ctx->u32_d (an LSD radix sort keeps
// history order inside a bin).  The order stays valid until the next compaction or binning.
int pct_bin_order(pct_ctx* ctx, int64_t n)
{
    if (ctx->bin_perm_n == n) return PCT_OK;
    hipStream_t s = ctx->stream;
    if (n > 0x7fffffff) {
        ctx->err = "bin order: more than 2^31 histories";
        return PCT_E_LIMIT;
    }
    int end_bit = 1;
    while ((1ll << end_bit) < (int64_t)ctx->P.num_bins) end_bit++;
    uint32_t* keys_in = ctx->u32_a;
    uint32_t* keys_out = ctx->u32_b;
    uint32_t* vals_in = ctx->u32_c;
    uint32_t* perm = ctx->u32_d;
    if (n > 0) {
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(keys_in, ctx->bin_num, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, s, n, vals_in);
        size_t need = 0;
        PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys_in, keys_out, vals_in, perm, (int)n, 0,
                                                              end_bit, s));
        if (need > ctx->cub_tmp_bytes) {
            PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, need));
            ctx->cub_tmp_bytes = need;
        }
        PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp, need, keys_in, keys_out, vals_in, perm,
                                                              (int)n, 0, end_bit, s));
        PCT_CHECK_HIP(ctx, hipGetLastError());
    }
    ctx->bin_perm_n = n;
    return PCT_OK;
}

// Sort key of a history's chord: the bin with angle bins a and a + A/2 paired (the same line traversed in
// the opposite direction, at t mirrored), so that histories on one line, either way, become neighbours.
__global__ void k_line_keys(int64_t n, const int32_t* __restrict__ bin_num, int T, int A, uint32_t* __restrict__ key)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = (uint32_t)bin_num[i];
    const uint32_t t = b % (uint32_t)T, r = b / (uint32_t)T, a = r % (uint32_t)A, v = r / (uint32_t)A;
    const uint32_t h = (uint32_t)A / 2u, half = a >= h ? 1u : 0u;
    const uint32_t tc = half ? (uint32_t)T - 1u - (t < (uint32_t)T ? t : (uint32_t)T - 1u) : t;
    key[i] = (((v * h + (a - half * h)) * (uint32_t)T + tc) << 1) | half;
}

// History indices ordered by chord line (k_line_keys) into ctx->u32_d.  Replaces a bin order held there.
int pct_line_order(pct_ctx* ctx, int64_t n)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    if (n > 0x7fffffff || (P.angular_bins & 1)) return pct_bin_order(ctx, n);
    ctx->bin_perm_n = -1;
    int end_bit = 2;
    while ((1ll << end_bit) < 2ll * P.num_bins) end_bit++;
    if (n > 0) {
        hipLaunchKernelGGL(k_line_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, n, ctx->bin_num, P.t_bins,
                           P.angular_bins, ctx->u32_a);
        hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, s, n, ctx->u32_c);
        size_t need = 0;
        PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, need, ctx->u32_a, ctx->u32_b, ctx->u32_c,
                                                              ctx->u32_d, (int)n, 0, end_bit, s));
        if (need > ctx->cub_tmp_bytes) {
            PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, need));
            ctx->cub_tmp_bytes = need;
        }
        PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp, need, ctx->u32_a, ctx->u32_b, ctx->u32_c,
                                                              ctx->u32_d, (int)n, 0, end_bit, s));
        PCT_CHECK_HIP(ctx, hipGetLastError());
    }
    return PCT_OK;
}

int pct_ensure_hist_capacity(pct_ctx* ctx, int64_t n)
{
    if (n <= ctx->cap_hist) return PCT_OK;
    for (int k = 0; k < 10; k++) {
        PCT_TRY(pct_alloc(ctx, &ctx->h[k], n));
        PCT_TRY(pct_alloc(ctx, &ctx->h_tmp[k], n));
    }
    PCT_TRY(pct_alloc(ctx, &ctx->wepl, n));
    PCT_TRY(pct_alloc(ctx, &ctx->wepl_tmp, n));
    PCT_TRY(pct_alloc(ctx, &ctx->gantry, n));
    PCT_TRY(pct_alloc(ctx, &ctx->gantry_tmp, n));
    PCT_TRY(pct_alloc(ctx, &ctx->bin_num, n));
    PCT_TRY(pct_alloc(ctx, &ctx->bin_tmp, n));
    PCT_TRY(pct_alloc(ctx, &ctx->in_range, n));
    PCT_TRY(pct_alloc(ctx, &ctx->u8_tmp, n));
    PCT_TRY(pct_alloc(ctx, &ctx->flag, n));
    PCT_TRY(pct_alloc(ctx, &ctx->pos, n));
    PCT_TRY(pct_alloc(ctx, &ctx->first_voxel, n));
    PCT_TRY(pct_alloc(ctx, &ctx->u32_a, n));
    PCT_TRY(pct_alloc(ctx, &ctx->u32_b, n));
    PCT_TRY(pct_alloc(ctx, &ctx->u32_c, n));
    PCT_TRY(pct_alloc(ctx, &ctx->u32_d, n));
    ctx->cap_hist = n;
    return PCT_OK;
}

int pct_carve_counts(pct_ctx* ctx);  // k_hull.hip

int pct_stage_intersect_bin_carve(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n_loaded;
    PCT_TRY(pct_ensure_hist_capacity(ctx, n));
    ctx->bin_perm_n = -1;
    unsigned g = grid_for(n);
    RawPtrs r{ctx->raw[0], ctx->raw[1], ctx->raw[2], ctx->raw[3], ctx->raw[4], ctx->raw[5], ctx->raw[6],
              ctx->raw[7], ctx->raw[8], ctx->raw[9], ctx->raw[10], ctx->raw[11], ctx->raw[13]};
    HPtrs o;
    for (int k = 0; k < 10; k++) o.h[k] = ctx->h[k];
    uint8_t* missed = ctx->u8_tmp;
    if (n > 0) {
        hipLaunchKernelGGL(k_intersections, dim3(g), dim3(kBlock), 0, s, n, r, o, missed, P.recon_cyl_radius,
                           P.recon_cyl_height, P.pi / 180.0);
        // WEPL travels with the histories
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->wepl, ctx->raw[12], n * sizeof(float), hipMemcpyDeviceToDevice, s));
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->gantry, ctx->raw[13], n * sizeof(float), hipMemcpyDeviceToDevice, s));
        // angle-bin centres: angle_bin * ANGULAR_BIN_SIZE * ANGLE_TO_RADIANS (DS.cu:5506)
        std::vector<double> cs(2 * P.angular_bins);
        for (int a = 0; a < P.angular_bins; a++) {
            double angle = a * P.angular_bin_size * (P.pi / 180.0);
            cs[a] = std::cos(angle);
            cs[P.angular_bins + a] = std::sin(angle);
        }
        pct_tmp<double> cs_buf;
        PCT_TRY(pct_alloc(ctx, &cs_buf.p, cs.size()));
        const double* d_cs = cs_buf.p;
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(cs_buf.p, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_binning, dim3(g), dim3(kBlock), 0, s, n, o, missed, ctx->bin_num, ctx->in_range, d_cs,
                           d_cs + P.angular_bins, P.pi, P.angular_bin_size, P.t_bin_size, P.v_bin_size, P.t_bins,
                           P.v_bins, P.angular_bins);
        hipLaunchKernelGGL(k_not, dim3(g), dim3(kBlock), 0, s, n, missed, ctx->flag);
        PCT_CHECK_HIP(ctx, hipGetLastError());
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
        pct_free(&cs_buf.p);
    }
    // keep histories that traversed the reconstruction volume (DS.cu:5429-5449)
    int64_t kept = 0;
    PCT_TRY(pct_compact(ctx, ctx->flag, n, &kept, false));
    ctx->n_hist = kept;
    ctx->n_recon_volume = kept;
    // hull-detection counts on the same histories (hull_detection, DS.cu:6607)
    PCT_TRY(pct_carve_counts(ctx));
    ctx->stage = ST_BINNED;
    return PCT_OK;
}

int pct_stage_stats_and_cuts(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n_hist;
    const int NB = P.num_bins;
    if (n > 0x7fffffff) {
        ctx->err = "statistical cuts: more than 2^31 histories";
        return PCT_E_LIMIT;
    }
    PCT_TRY(pct_alloc(ctx, &ctx->counts_pre, NB));
    PCT_TRY(pct_alloc(ctx, &ctx->counts_post, NB));
    PCT_TRY(pct_alloc(ctx, &ctx->bin_off, NB + 1));
    for (int k = 0; k < 3; k++) {
        PCT_TRY(pct_alloc(ctx, &ctx->mean[k], NB));
        PCT_TRY(pct_alloc(ctx, &ctx->stddev[k], NB));
    }
    PCT_TRY(pct_alloc(ctx, &ctx->sinogram, NB));
    unsigned g = grid_for(n);
    // history indices in bin order (kept from the hull carving when no compaction came between)
    PCT_TRY(pct_bin_order(ctx, n));
    const uint32_t* perm = ctx->u32_d;
    int32_t* hist = ctx->counts_post;  // bin occupancy, reused below
    PCT_CHECK_HIP(ctx, hipMemsetAsync(hist, 0, NB * sizeof(int32_t), s));
    if (n > 0) hipLaunchKernelGGL(k_bin_hist, dim3(g), dim3(kBlock), 0, s, n, ctx->bin_num, hist);
    size_t need = 0;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, hist, ctx->bin_off, NB + 1, s));
    if (need > ctx->cub_tmp_bytes) {
        PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, need));
        ctx->cub_tmp_bytes = need;
    }
    // NB + 1 inputs: the last read is past `hist`; give the scan an input of NB+1 by scanning NB then
    // appending the total.
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, need, hist, ctx->bin_off, NB, s));
    int32_t n32 = (int32_t)n;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->bin_off + NB, &n32, sizeof(int32_t), hipMemcpyHostToDevice, s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    uint8_t* failed = ctx->flag;
    if (n > 0) PCT_CHECK_HIP(ctx, hipMemsetAsync(failed, 0, n, s));
    pct_tmp<uint8_t> g_ir;
    pct_tmp<float> g_w;
    pct_tmp<double> g_ut, g_uv;
    PCT_TRY(pct_alloc(ctx, &g_ir.p, n > 0 ? n : 1));
    PCT_TRY(pct_alloc(ctx, &g_w.p, n > 0 ? n : 1));
    PCT_TRY(pct_alloc(ctx, &g_ut.p, n > 0 ? n : 1));
    PCT_TRY(pct_alloc(ctx, &g_uv.p, n > 0 ? n : 1));
    if (n > 0)
        hipLaunchKernelGGL(k_bin_gather, dim3(g), dim3(kBlock), 0, s, n, perm, ctx->in_range, ctx->wepl, ctx->h[6],
                           ctx->h[7], ctx->h[8], ctx->h[9], P.pi, g_ir.p, g_w.p, g_ut.p, g_uv.p);
    hipLaunchKernelGGL(k_bin_stats, dim3(grid_for(NB)), dim3(kBlock), 0, s, NB, ctx->bin_off, perm, g_ir.p, g_w.p,
                       g_ut.p, g_uv.p, ctx->counts_pre, ctx->mean[0], ctx->mean[1], ctx->mean[2], ctx->stddev[0],
                       ctx->stddev[1], ctx->stddev[2], ctx->counts_post, ctx->sinogram, failed,
                       (float)P.sigmas_2_keep, P.quirk_a1_bin0);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    // keep the histories that passed (DS.cu:5892-5911)
    if (n > 0) hipLaunchKernelGGL(k_not, dim3(g), dim3(kBlock), 0, s, n, failed, ctx->u8_tmp);
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->flag, ctx->u8_tmp, n > 0 ? n : 1, hipMemcpyDeviceToDevice, s));
    int64_t kept = 0;
    PCT_TRY(pct_compact(ctx, ctx->flag, n, &kept, false));
    ctx->n_hist = kept;
    ctx->n_post_cuts = kept;
    ctx->stage = ST_CUT;
    return PCT_OK;
}

int pct_launch_units(hipStream_t s, int64_t n, float* const raw[14], int mm, double ts, double us, double vs)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_units, dim3(grid_for(n)), dim3(kBlock), 0, s, n, raw[0], raw[1], raw[2], raw[3], raw[4],
                       raw[5], raw[6], raw[7], raw[8], raw[9], raw[10], raw[11], raw[12], mm, ts, us, vs);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
