// k_fbp.hip — stage (c): sinogram, ramp filter and voxel-driven backprojection.
//
// filter: one workgroup per (v, angle) detector strip; the strip, the cosine weights of each
// detector element and the 2T-1 filter taps are staged in LDS, so each lane's O(T) convolution
// reads only LDS.  backprojection: one lane per voxel, the per-angle cos/sin come from a host table
// (the same double values the oracle uses), the filtered sinogram (a few MB) stays in L2/MALL.
#include "internal.h"

namespace {

constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n, int block = kBlock) { return pct_grid_for(n, block); }

// construct_sinogram_GPU, DS.cu:6073-6079
__global__ void k_sino_div(int NB, const int32_t* counts, float* sino)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < NB && counts[b] > 0) sino[b] /= counts[b];
}

// filter_GPU, DS.cu:6206-6240.  Dynamic LDS: taps[2T-1] double, scale[T] double, strip[T] float.
__global__ void k_filter(const float* sino, float* out, int T, int A, int V, double t_bin_size, double v_bin_size,
                         double SR, double PI, int filter_type, double tau)
{
    extern __shared__ double lds[];
    double* taps = lds;
    double* scale = lds + (2 * T - 1);
    float* strip = (float*)(scale + T);
    const int vb = blockIdx.x, ab = blockIdx.y;
    const int64_t so = ((int64_t)vb * A + ab) * T;
    const double v = (vb - V / 2 + 0.5) * v_bin_size;
    for (int k = threadIdx.x; k < 2 * T - 1; k += blockDim.x) {
        int sep = k - (T - 1);
        double filt;
        if (filter_type == PCT_FILTER_RAM_LAK) {
            if (sep == 0) filt = 1.0 / (4.0 * (tau * tau));
            else if (sep % 2 == 0) filt = 0;
            else {
                double q = tau * PI * sep;
                filt = -1.0 / (q * q);
            }
        } else if (filter_type == PCT_FILTER_SHEPP_LOGAN) {
            filt = 1.0 / ((t_bin_size * PI * t_bin_size * PI) * (1.0 - (4 * sep * sep)));
        } else {
            filt = (sep == 0) ? 1.0 / t_bin_size : 0.0;
        }
        taps[k] = filt;
    }
    for (int k = threadIdx.x; k < T; k += blockDim.x) {
        double t = (k - T / 2 + 0.5) * t_bin_size;
        scale[k] = SR / sqrt(SR * SR + t * t + v * v);
        strip[k] = sino[so + k];
    }
    __syncthreads();
    for (int tb = threadIdx.x; tb < T; tb += blockDim.x) {
        float acc = 0.0f;
        for (int tr = 0; tr < T; tr++)
            acc += (float)(t_bin_size * (double)strip[tr] * taps[tb - tr + T - 1] * scale[tr]);
        out[so + tb] = acc;
    }
}

struct BPArgs {
    int T, V, A, C, R, S;
    int64_t NV;
    double VW, VH, VT, Rc, SR, t_bin_size, v_bin_size, delta;
    float rsp_air;
    int quirk_decode;
};

// backprojection_GPU2, DS.cu:6340-6479
__global__ void k_backproject(BPArgs a, const float* __restrict__ sf, const double* __restrict__ cs, float* image)
{
    int64_t voxel = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (voxel >= a.NV) return;
    int vx, vy, vz;
    const int64_t CR = (int64_t)a.C * a.R;
    if (a.quirk_decode) {  // voxel_2_3D_voxels_GPU, DS.cu:11124-11134 ('>' comparisons)
        int64_t rem = voxel, z = 0, y = 0;
        if (rem - CR > 0) {
            z = (rem - 1) / CR;
            rem -= z * CR;
        }
        if (rem - a.C > 0) {
            y = (rem - 1) / a.C;
            rem -= y * a.C;
        }
        vx = (int)rem;
        vy = (int)y;
        vz = (int)z;
    } else {
        vz = (int)(voxel / CR);
        vy = (int)((voxel / a.C) % a.R);
        vx = (int)(voxel % a.C);
    }
    double x = 1 * (vx - (a.C - 1) / 2.0) * a.VW;
    double y = -1 * (vy - (a.R - 1) / 2.0) * a.VH;
    double z = -1 * (vz - (a.S - 1) / 2.0) * a.VT;
    if ((x * x + y * y) > (a.Rc * a.Rc)) {
        image[voxel] = a.rsp_air;
        return;
    }
    const int T = a.T, V = a.V, A = a.A;
    float acc = 0.0f;
    for (int ab = 0; ab < A; ab++) {
        const double c = cs[ab], sn = cs[A + ab];
        double u = x * c + y * sn;
        double t = y * c - x * sn;
        double v = z;
        double dnt = (t - u * (t / (a.SR + u))) / a.t_bin_size + T / 2.0;
        int t_bin = (int)floor(dnt);
        double eta = dnt - t_bin;
        double dnv = (v - u * (v / (a.SR + u))) / a.v_bin_size + V / 2.0;
        int v_bin = (int)floor(dnv);
        double eps = dnv - v_bin;
        double scale = (double)pct_sqf(a.SR / (a.SR + u));
        int64_t b11 = ((int64_t)v_bin * A + ab) * T + t_bin;
        int64_t b12 = b11 + 1, b21 = b11 + (int64_t)A * T, b22 = b21 + 1;
        float xterm = 0.0f;
        if (t_bin > -1 && t_bin < T - 1) {
            if (v_bin > -1 && v_bin < V - 1)
                xterm = (float)((double)sf[b11] * (1 - eta) * (1 - eps) + (double)sf[b12] * eta * (1 - eps) +
                                (double)sf[b21] * (1 - eta) * eps + (double)sf[b22] * eta * eps);
            else if (v_bin == -1)
                xterm = (float)((double)sf[b21] * (1 - eta) * eps + (double)sf[b22] * eta * eps);
            else if (v_bin == V - 1)
                xterm = (float)((double)sf[b11] * (1 - eta) * (1 - eps) + (double)sf[b12] * eta * (1 - eps));
        } else if (t_bin == -1) {
            if (v_bin > -1 && v_bin < V - 1)
                xterm = (float)((double)sf[b12] * eta * (1 - eps) + (double)sf[b22] * eta * eps);
            else if (v_bin == -1)
                xterm = (float)((double)sf[b22] * eta * eps);
            else if (v_bin == V - 1)
                xterm = (float)((double)sf[b12] * eta * (1 - eps));
        } else if (t_bin == T - 1) {
            if (v_bin > -1 && v_bin < V - 1)
                xterm = (float)((double)sf[b11] * (1 - eta) * (1 - eps) + (double)sf[b21] * (1 - eta) * eps);
            else if (v_bin == -1)
                xterm = (float)((double)sf[b21] * (1 - eta) * eps);
            else if (v_bin == V - 1)
                xterm = (float)((double)sf[b11] * (1 - eta) * (1 - eps));
        }
        acc += (float)((double)xterm * a.delta * scale);
    }
    image[voxel] = acc;
}

}  // namespace

int pct_fbp_hull(pct_ctx* ctx);

// generate_preprocessing_images -> construct_sinogram (DS.cu:6039) and FBP (:6080)
int pct_stage_fbp(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    const int NB = P.num_bins;
    hipLaunchKernelGGL(k_sino_div, dim3(grid_for(NB)), dim3(kBlock), 0, s, NB, ctx->counts_post, ctx->sinogram);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    if (!P.fbp_on) {
        ctx->stage = ST_FBP;
        return PCT_OK;
    }
    PCT_TRY(pct_alloc(ctx, &ctx->sino_filtered, NB));
    PCT_TRY(pct_alloc(ctx, &ctx->fbp, P.num_voxels));
    const int T = P.t_bins;
    size_t lds = (size_t)(2 * T - 1) * sizeof(double) + (size_t)T * sizeof(double) + (size_t)T * sizeof(float);
    if (lds > 64 * 1024) {
        ctx->err = "T_BINS too large for the LDS filter (max ~2300)";
        return PCT_E_LIMIT;
    }
    int threads = T < 512 ? ((T + 63) / 64) * 64 : 512;
    hipLaunchKernelGGL(k_filter, dim3(P.v_bins, P.angular_bins), dim3(threads), lds, s, ctx->sinogram,
                       ctx->sino_filtered, T, P.angular_bins, P.v_bins, P.t_bin_size, P.v_bin_size, P.source_radius,
                       P.pi, P.fbp_filter_type, P.ram_lak_tau);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    // cos/sin of bin_angle = angle_bin * delta (DS.cu:6362-6364)
    const double delta = P.angular_bin_size * (P.pi / 180.0);
    std::vector<double> cs(2 * P.angular_bins);
    for (int a = 0; a < P.angular_bins; a++) {
        double ba = a * delta;
        cs[a] = std::cos(ba);
        cs[P.angular_bins + a] = std::sin(ba);
    }
    pct_tmp<double> cs_buf;
    PCT_TRY(pct_alloc(ctx, &cs_buf.p, cs.size()));
    const double* d_cs = cs_buf.p;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(cs_buf.p, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice, s));
    BPArgs a;
    a.T = T;
    a.V = P.v_bins;
    a.A = P.angular_bins;
    a.C = P.columns;
    a.R = P.rows;
    a.S = P.slices;
    a.NV = P.num_voxels;
    a.VW = P.voxel_width;
    a.VH = P.voxel_height;
    a.VT = P.voxel_thickness;
    a.Rc = P.recon_cyl_radius;
    a.SR = P.source_radius;
    a.t_bin_size = P.t_bin_size;
    a.v_bin_size = P.v_bin_size;
    a.delta = delta;
    a.rsp_air = (float)P.rsp_air;
    a.quirk_decode = P.quirk_voxel_decode;
    hipLaunchKernelGGL(k_backproject, dim3(grid_for(P.num_voxels)), dim3(kBlock), 0, s, a, ctx->sino_filtered, d_cs,
                       ctx->fbp);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    pct_free(&cs_buf.p);
    PCT_TRY(pct_median_filter_fbp(ctx));  // FBP(), DS.cu:6150-6166 (x0 only)
    if (P.hull_type == PCT_HULL_FBP) PCT_TRY(pct_fbp_hull(ctx));
    ctx->stage = ST_FBP;
    return PCT_OK;
}
