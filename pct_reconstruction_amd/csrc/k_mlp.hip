// k_mlp.hip — stage (d): MLP endpoints on the hull and the MLP voxel lists.
//
// Endpoints: one lane per history, two straight walks along the tracker angles to the first hull
// voxel (find_MLP_endpoints_GPU, DS.cu:7581-7773), the histories in sinogram-bin order and the hull
// as a bit mask.
//
// MLP lists: a one-history-per-lane kernel.  The Sigma_1 terms of every u-step (they depend only on
// the step number) are precomputed on the host in FP64 and staged in LDS at kernel start; per step a
// lane reads them from LDS, one scattering coefficient for its Sigma_2 from the FP64 table in HBM/L2,
// and does the 2x2 inversions and the t/v solve in FP64 exactly as find_MLP_path_GPU_tabulated
// (DS.cu:9104-9262).  Each history's voxel list is written, in 64-byte chunks staged in LDS, to its own
// segment of a CSR array whose capacity comes from the number of u-steps; the DROP kernels read
// entries [off, off + count).  With the transposed DROP plan the kernel runs group by group on a second
// stream while the plan's earlier groups are transposed (pct_drop_prepare).
#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace {

constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n, int block = kBlock) { return pct_grid_for(n, block); }

// The hull as one bit per voxel (bit v % 32 of word v / 32 set where hull[v] == 1): 1/8 of the bytes, so
// the whole mask stays in each XCD's L2 while the endpoint walks test it.
__global__ void k_hull_bits(int64_t nv, const uint8_t* __restrict__ hull, uint32_t* __restrict__ bits)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 32 >= nv) return;
    uint32_t b = 0;
    for (int k = 0; k < 32 && w * 32 + k < nv; k++) b |= (uint32_t)(hull[w * 32 + k] == 1) << k;
    bits[w] = b;
}

__device__ __forceinline__ int hull_bit(const uint32_t* __restrict__ bits, int64_t v)
{
    return (int)((bits[v >> 5] >> (v & 31)) & 1u);
}

// find_MLP_endpoints_GPU, DS.cu:7581-7773; the walk is bounded by 4*(C+R+S)+16 steps (the
// reference loops forever when the walk never leaves the image, e.g. z clamped at slice 0).
__device__ int find_endpoint(const DevGeom& g, const uint32_t* __restrict__ hull, double x_start, double y_start,
                             double z_start, double xy_angle, double xz_angle, double* xo, double* yo, double* zo,
                             int* vxo, int* vyo, int* vzo, int entering)
{
    const int C = g.C, R = g.R, S = g.S;
    double x = x_start, y = y_start, z = z_start;
    if (!entering) xy_angle += g.PI;
    const double cxy = cos(xy_angle), sxy = sin(xy_angle), sxz = sin(xz_angle);
    int x_move = (cxy >= 0) - (cxy <= 0);
    int y_move = (sxy >= 0) - (sxy <= 0);
    int z_move = (sxz >= 0) - (sxz <= 0);
    if (x_move < 0) z_move *= -1;
    int vx = pct_calc_voxel(g.X0, x, g.VW), vy = pct_calc_voxel(g.Y0, y, g.VH), vz = pct_calc_voxel(g.Z0, z, g.VT);
    double x_to_go = pct_dist(g.X0, x, PCT_X_INC, x_move, g.VW, vx);
    double y_to_go = pct_dist(g.Y0, y, PCT_Y_INC, y_move, g.VH, vy);
    double z_to_go = pct_dist(g.Z0, z, PCT_Z_INC, z_move, g.VT, vz);
    int64_t voxel = vx + ((int64_t)vy + (int64_t)vz * R) * C;
    const double txy = tan(xy_angle), txz = tan(xz_angle);
    double delta_yx = fabs(txy);
    double delta_zx = fabs(txz);
    double delta_zy = fabs(txz / txy);
    double dy_dx = txy;
    double dz_dx = txz;
    double dz_dy = txz / txy;
    double dx_dy = (double)pct_recipf(txy);
    double dx_dz = (double)pct_recipf(txz);
    double dy_dz = (double)pct_tanf(xy_angle) / txz;
    int hit = 0;
    bool outside = (vx >= C) || (vy >= R) || (vz >= S) || (vx < 0) || (vy < 0) || (vz < 0);
    if (!outside) hit = hull_bit(hull, voxel);
    bool end_walk = outside || hit;
    const int cap = 4 * (C + R + S) + 16;
    int steps = 0;
    while (!end_walk) {
        if (z_move != 0) {
            double x_ext = delta_zx * x_to_go;
            double y_ext = delta_zy * y_to_go;
            if ((z_to_go <= x_ext) && (z_to_go <= y_ext)) {
                vz -= z_move;
                z = pct_edge(g.Z0, vz, g.VT, PCT_Z_INC, z_move);
                x = pct_corr(dx_dz, z, z_start, x_start);
                y = pct_corr(dy_dz, z, z_start, y_start);
                x_to_go = pct_dist(g.X0, x, PCT_X_INC, x_move, g.VW, vx);
                y_to_go = pct_dist(g.Y0, y, PCT_Y_INC, y_move, g.VH, vy);
                z_to_go = g.VT;
            } else if (x_ext <= y_ext) {
                vx += x_move;
                x = pct_edge(g.X0, vx, g.VW, PCT_X_INC, x_move);
                y = pct_corr(dy_dx, x, x_start, y_start);
                z = pct_corr(dz_dx, x, x_start, z_start);
                x_to_go = g.VW;
                y_to_go = pct_dist(g.Y0, y, PCT_Y_INC, y_move, g.VH, vy);
                z_to_go = pct_dist(g.Z0, z, PCT_Z_INC, z_move, g.VT, vz);
            } else {
                vy -= y_move;
                y = pct_edge(g.Y0, vy, g.VH, PCT_Y_INC, y_move);
                x = pct_corr(dx_dy, y, y_start, x_start);
                z = pct_corr(dz_dy, y, y_start, z_start);
                x_to_go = pct_dist(g.X0, x, PCT_X_INC, x_move, g.VW, vx);
                y_to_go = g.VH;
                z_to_go = pct_dist(g.Z0, z, PCT_Z_INC, z_move, g.VT, vz);
            }
            if (x_to_go == 0) { x_to_go = g.VW; vx += x_move; }
            if (y_to_go == 0) { y_to_go = g.VH; vy -= y_move; }
            if (z_to_go == 0) { z_to_go = g.VT; vz -= z_move; }
            vz = vz > 0 ? vz : 0;
        } else {
            double y_ext = y_to_go / delta_yx;
            if (x_to_go <= y_ext) {
                vx += x_move;
                x = pct_edge(g.X0, vx, g.VW, PCT_X_INC, x_move);
                y = pct_corr(dy_dx, x, x_start, y_start);
                x_to_go = g.VW;
                y_to_go = pct_dist(g.Y0, y, PCT_Y_INC, y_move, g.VH, vy);
            } else {
                vy -= y_move;
                y = pct_edge(g.Y0, vy, g.VH, PCT_Z_INC, y_move);
                x = pct_corr(dx_dy, y, y_start, x_start);
                x_to_go = pct_dist(g.X0, x, PCT_X_INC, x_move, g.VW, vx);
                y_to_go = g.VH;
            }
            if (x_to_go == 0) { x_to_go = g.VW; vx += x_move; }
            if (y_to_go == 0) { y_to_go = g.VH; vy -= y_move; }
        }
        voxel = vx + (int64_t)vy * C + (int64_t)vz * C * R;
        outside = (vx >= C) || (vy >= R) || (vz >= S) || (vx < 0) || (vy < 0) || (vz < 0);
        if (!outside) hit = hull_bit(hull, voxel);
        end_walk = outside || hit;
        if (++steps > cap) {
            hit = 0;
            end_walk = true;
        }
    }
    if (hit) {
        *xo = x;
        *yo = y;
        *zo = z;
    }
    *vxo = vx;
    *vyo = vy;
    *vzo = vz;
    return hit;
}

struct H10 {
    float* h[10];
};

// collect_MLP_endpoints_GPU_nobool, DS.cu:7819-7854.  With `perm`, lane t takes history perm[t]: in
// sinogram-bin order the lanes of a wavefront walk nearly the same chords, so their walks to the hull take
// about as many steps.
__global__ void k_endpoints(int64_t n, const uint32_t* __restrict__ perm, DevGeom g, const uint32_t* __restrict__ hull,
                            H10 hp, uint32_t* first_voxel, uint8_t* keep)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = perm ? (int64_t)perm[t] : t;
    double xi = 0, yi = 0, zi = 0, xo = 0, yo = 0, zo = 0;
    int vx, vy, vz, vxi, vyi, vzi;
    int entered = find_endpoint(g, hull, hp.h[0][i], hp.h[1][i], hp.h[2][i], hp.h[6][i], hp.h[7][i], &xi, &yi, &zi,
                                &vx, &vy, &vz, 1);
    int exited = find_endpoint(g, hull, hp.h[3][i], hp.h[4][i], hp.h[5][i], hp.h[8][i], hp.h[9][i], &xo, &yo, &zo,
                               &vxi, &vyi, &vzi, 0);
    if (entered && exited) {
        first_voxel[i] = (uint32_t)(vx + g.C * vy + g.R * g.C * vz);
        hp.h[0][i] = (float)xi; hp.h[1][i] = (float)yi; hp.h[2][i] = (float)zi;
        hp.h[3][i] = (float)xo; hp.h[4][i] = (float)yo; hp.h[5][i] = (float)zo;
        keep[i] = 1;
    } else {
        first_voxel[i] = (uint32_t)g.NV;
        keep[i] = 0;
    }
}

struct MlpArgs {
    DevGeom g;
    double trig_min, trig_step, step, coeff_step, poly_step;
    int64_t n_trig, n_poly, n_coeff;
    unsigned s_step, p_step;
    int n_steps;
};

// Orientation and depth of a history (DS.cu:9120-9147): rotate by the tabulated sin/cos of the
// entry angle, adding pi to both angles when u_in > u_out.
struct MlpFrame {
    double sin_t, cos_t, u_in, u_out, t_in, t_out, xy_entry, xy_exit, u_2;
};
__device__ __forceinline__ MlpFrame mlp_frame(const MlpArgs& a, const double* __restrict__ sin_t,
                                              const double* __restrict__ cos_t, double x_in, double y_in,
                                              double x_out, double y_out, double xy_entry, double xy_exit)
{
    MlpFrame f;
    int64_t ti = (int64_t)(unsigned int)((xy_entry - a.trig_min) / a.trig_step + 0.5);
    ti = ti < a.n_trig ? ti : a.n_trig - 1;
    f.sin_t = sin_t[ti];
    f.cos_t = cos_t[ti];
    f.u_in = f.cos_t * x_in + f.sin_t * y_in;
    f.u_out = f.cos_t * x_out + f.sin_t * y_out;
    if (f.u_in > f.u_out) {
        xy_entry += a.g.PI;
        xy_exit += a.g.PI;
        ti = (int64_t)(unsigned int)((xy_entry - a.trig_min) / a.trig_step + 0.5);
        ti = ti < a.n_trig ? ti : a.n_trig - 1;
        f.sin_t = sin_t[ti];
        f.cos_t = cos_t[ti];
        f.u_in = f.cos_t * x_in + f.sin_t * y_in;
        f.u_out = f.cos_t * x_out + f.sin_t * y_out;
    }
    f.t_in = f.cos_t * y_in - f.sin_t * x_in;
    f.t_out = f.cos_t * y_out - f.sin_t * x_out;
    f.xy_entry = xy_entry;
    f.xy_exit = xy_exit;
    f.u_2 = fabs(f.u_out - f.u_in);
    return f;
}

// EffectiveChordLength_GPU, DS.cu:8873-8894
__device__ double chord_length(double PI, double VW, double VT, double STEP, double abs_angle_t, double abs_angle_v)
{
    double eff_t = abs_angle_t - ((int)(abs_angle_t / (PI / 2))) * (PI / 2);
    double eff_v = fabs(abs_angle_v);
    double sf = STEP / VW;
    float s2t = pct_sinf(2 * eff_t), ct = pct_cosf(eff_t), st = pct_sinf(eff_t);
    float cs = ct + st;
    double c2d = (1 / 3.0) * ((sf * sf * s2t - 6) / (sf * s2t - (double)(2 * cs)) + sf * sf * s2t / (double)(2 * cs));
    double mpw = VW / (double)cs;
    double hf = VT / mpw;
    sf = STEP / mpw;
    float s2v = pct_sinf(2 * eff_v), cv = pct_cosf(eff_v), sv = pct_sinf(eff_v);
    double den = hf * cv + sv;
    double c3d = (1 / 3.0) * ((sf * sf * s2v - 6 * hf) / (sf * s2v - 2 * den) + sf * sf * s2v / (2 * den));
    return VW * c2d * c3d;
}

// capacity of each history's list: 1 + the number of u-steps (loop while depth > STEP)
__global__ void k_mlp_capacity(int64_t n, MlpArgs a, H10 hp, const double* __restrict__ sin_t,
                               const double* __restrict__ cos_t, int64_t* cap)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    MlpFrame f = mlp_frame(a, sin_t, cos_t, hp.h[0][i], hp.h[1][i], hp.h[3][i], hp.h[4][i], hp.h[6][i], hp.h[8][i]);
    int64_t steps = 0;
    double depth = f.u_2 - a.step;
    while (depth > a.step && steps < 4 * a.n_steps) {
        depth -= a.step;
        steps++;
    }
    cap[i] = 1 + steps;
}

// find_MLP_path_GPU_tabulated, DS.cu:9104-9262, writing the voxel list of history i.
// A lane's list entries go out in whole 64-byte chunks: the lane collects them in its own LDS column
// (entry k at buf[k * kBlock], so the lanes of a wavefront use distinct banks) and writes each chunk that
// ends on a 16-entry boundary of the list array with four 16-byte stores, the partial chunks at either end
// of its list entry by entry.  Written one 4-byte entry per step instead, the lines of 64 interleaved lists
// leave the L2 partly written and the HBM write traffic grows about tenfold.
struct ListWriter {
    uint32_t* vox;   // the list array (16-entry chunks are 64-byte aligned: hipMalloc aligns to 256 bytes)
    uint32_t* buf;   // this lane's LDS column
    int64_t start;   // list index of buf[0]
    int n;           // buffered entries, at most 16

    __device__ __forceinline__ ListWriter(uint32_t* v, uint32_t* lds, int64_t base)
        : vox(v), buf(lds + threadIdx.x), start(base), n(0)
    {
    }
    __device__ __forceinline__ void put(uint32_t x)
    {
        buf[n * kBlock] = x;
        ++n;
        if (((start + n) & 15) == 0) flush();
    }
    __device__ __forceinline__ void flush()
    {
        if (n == 16) {  // ends on a 16-entry boundary with 16 entries: starts on one too
            uint4* d = reinterpret_cast<uint4*>(vox + start);
#pragma unroll
            for (int q = 0; q < 4; q++)
                d[q] = make_uint4(buf[(4 * q) * kBlock], buf[(4 * q + 1) * kBlock], buf[(4 * q + 2) * kBlock],
                                  buf[(4 * q + 3) * kBlock]);
        } else {
            for (int k = 0; k < n; k++) vox[start + k] = buf[k * kBlock];
        }
        start += n;
        n = 0;
    }
};

__global__ void __launch_bounds__(kBlock) k_mlp_lists(int64_t n, MlpArgs a, H10 hp, const uint32_t* __restrict__ first,
                                                      const double* __restrict__ sin_t,
                                                      const double* __restrict__ cos_t,
                                                      const double* __restrict__ scat, const double* __restrict__ p12,
                                                      const double* __restrict__ p26, const double* __restrict__ p312,
                                                      const double* __restrict__ step_terms,
                                                      const int64_t* __restrict__ off, uint32_t* __restrict__ cnt_out,
                                                      uint32_t* __restrict__ vox, double* __restrict__ chord,
                                                      int32_t* overflow)
{
    extern __shared__ double lds_terms[];  // n_steps x 6
    __shared__ uint32_t list_buf[16 * kBlock];
    for (int k = threadIdx.x; k < a.n_steps; k += blockDim.x) {
#pragma unroll
        for (int c = 0; c < 6; c++) lds_terms[k * 6 + c] = step_terms[k * 8 + c];
    }
    __syncthreads();
    const DevGeom& g = a.g;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double x_in = hp.h[0][i], y_in = hp.h[1][i], z_in = hp.h[2][i];
        const double x_out = hp.h[3][i], y_out = hp.h[4][i], z_out = hp.h[5][i];
        const double xz_entry = hp.h[7][i], xz_exit = hp.h[9][i];
        MlpFrame f = mlp_frame(a, sin_t, cos_t, x_in, y_in, x_out, y_out, hp.h[6][i], hp.h[8][i]);
        const double T2_1 = f.xy_exit - f.xy_entry, V2_1 = xz_exit - xz_entry;
        const double u_2 = f.u_2, STEP = a.step;
        double depth = u_2 - STEP;
        double u_shifted = f.u_in;
        unsigned s2 = (unsigned)(depth / a.coeff_step + 0.5);
        int64_t u2i = (int64_t)(unsigned)(u_2 / a.poly_step + 0.5);
        u2i = u2i < a.n_poly ? u2i : a.n_poly - 1;
        const double u2_312 = p312[u2i], u2_26 = p26[u2i], u2_12 = p12[u2i];
        const double u2_sq = (double)pct_sqf(u_2);
        const int64_t base = off[i], capn = off[i + 1] - off[i];
        uint32_t last = first[i];
        int64_t cnt = 1;
        ListWriter w(vox, list_buf, base);
        w.put(last);
        chord[i] = chord_length(g.PI, g.VW, g.VT, STEP, (f.xy_entry + f.xy_exit) / 2.0, (xz_entry + xz_exit) / 2.0);
        int k = 0;
        while (depth > STEP) {
            if (k >= a.n_steps || (int64_t)s2 >= a.n_coeff) {
                atomicOr(overflow, 1);
                break;
            }
            const double* row = lds_terms + k * 6;
            const double u1_12 = row[0], u1_23 = row[1], u1_34 = row[2];
            const double S1I0 = row[3], S1I1 = row[4], S1I2 = row[5];
            double st2 = u2_312 - u2_sq * u1_12 + 2 * u_2 * u1_23 - u1_34;
            double st2th2 = u2_26 - u_2 * u1_12 + u1_23;
            double sth2 = u2_12 - u1_12;
            double det2 = scat[s2] * (st2 * sth2 - (double)pct_sqf(st2th2));
            double S2I0 = sth2 / det2, S2I1 = st2th2 / det2, S2I2 = st2 / det2;
            double c24 = S2I0 * depth - S2I1;
            double ft0 = S1I0 + S2I0;
            double ft1 = c24 - S1I1;
            double ft2 = depth * S2I0 - S1I1 - S2I1;
            double ft3 = S1I2 + S2I2 + depth * (c24 - S2I1);
            double detft = ft0 * ft3 - ft1 * ft2;
            double c3 = S2I0 * f.t_out - S2I1 * T2_1;
            double sc0 = S1I0 * f.t_in + c3;
            double sc1 = depth * c3 + S2I2 * T2_1 - S2I1 * f.t_out - S1I1 * f.t_in;
            double t_1 = (ft3 * sc0 - ft1 * sc1) / detft;
            c3 = S2I0 * z_out - S2I1 * V2_1;
            sc0 = S1I0 * z_in + c3;
            sc1 = depth * c3 + S2I2 * V2_1 - S2I1 * z_out - S1I1 * z_in;
            double v_1 = (ft3 * sc0 - ft1 * sc1) / detft;
            u_shifted += STEP;
            double x_1 = f.cos_t * u_shifted - f.sin_t * t_1;
            double y_1 = f.sin_t * u_shifted + f.cos_t * t_1;
            int vx = pct_calc_voxel(g.X0, x_1, g.VW);
            int vy = pct_calc_voxel(g.Y0, y_1, g.VH);
            int vz = pct_calc_voxel(g.Z0, v_1, g.VT);
            int64_t voxel = vx + (int64_t)vy * g.C + (int64_t)vz * g.C * g.R;
            if (voxel != (int64_t)last && voxel < g.NV) {
                if (cnt >= capn) {
                    atomicOr(overflow, 2);
                    break;
                }
                w.put((uint32_t)voxel);
                cnt++;
                last = (uint32_t)voxel;
            }
            depth -= STEP;
            s2 -= a.s_step;
            k++;
        }
        w.flush();
        cnt_out[i] = (uint32_t)cnt;
    }
}

// ---------------------------------------------------------------- untabulated MLP ----
// find_MLP_path_GPU (MLP_ALGORITHM = STANDARD), DS.cu:8895-9103.  The scattering polynomial of every
// Sigma term is evaluated in the kernel: a0 + u(a1 + u(a2 + u(a3 + u(a4 + ((u c5) P5) / d5)))), where the
// a_k = (c_k P_k) / d_k are formed on the host (the A_k_OVER_m macros of DS.h:756-781 expand that way).
struct PolyFam {
    double a[5];
    double d5;
};
struct StdArgs {
    DevGeom g;
    double step, e0, x0, c5, p5;
    PolyFam f12, f23, f34, f26, f312;
    int max_steps;
};

__device__ __forceinline__ double std_poly(const StdArgs& a, const PolyFam& f, double u)
{
    double in5 = u * a.c5 * a.p5 / f.d5;
    return f.a[0] + u * (f.a[1] + u * (f.a[2] + u * (f.a[3] + u * (f.a[4] + in5))));
}

// powf(x, 3) of a double argument: x rounded to float, cubed in double (exact square, one rounding), rounded.
__device__ __forceinline__ float std_cubef(double x)
{
    double f = (double)(float)x;
    return (float)(f * f * f);
}

// Rotation into the entry frame with cos()/sin() of the entry angle (DS.cu:8938-8955).
struct StdFrame {
    double c, s, u_in, u_out, t_in, t_out, xy_entry, xy_exit;
};
__device__ __forceinline__ StdFrame std_frame(double PI, double x_in, double y_in, double x_out, double y_out,
                                              double xy_entry, double xy_exit)
{
    StdFrame f;
    f.c = cos(xy_entry);
    f.s = sin(xy_entry);
    f.u_in = f.c * x_in + f.s * y_in;
    f.u_out = f.c * x_out + f.s * y_out;
    f.t_in = f.c * y_in - f.s * x_in;
    f.t_out = f.c * y_out - f.s * x_out;
    if (f.u_in > f.u_out) {
        xy_entry += PI;
        xy_exit += PI;
        f.c = cos(xy_entry);
        f.s = sin(xy_entry);
        f.u_in = f.c * x_in + f.s * y_in;
        f.u_out = f.c * x_out + f.s * y_out;
        f.t_in = f.c * y_in - f.s * x_in;
        f.t_out = f.c * y_out - f.s * x_out;
    }
    f.xy_entry = xy_entry;
    f.xy_exit = xy_exit;
    return f;
}

// capacity: 1 + the u-steps of `while (u_1 < u_2 - MLP_U_STEP)` with u_1 += MLP_U_STEP (DS.cu:8978)
__global__ void k_mlp_capacity_std(int64_t n, StdArgs a, H10 hp, int64_t* cap)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    StdFrame f = std_frame(a.g.PI, hp.h[0][i], hp.h[1][i], hp.h[3][i], hp.h[4][i], hp.h[6][i], hp.h[8][i]);
    const double u_2 = fabs(f.u_out - f.u_in);
    double u_1 = a.step;
    int64_t steps = 0;
    while (u_1 < u_2 - a.step && steps < a.max_steps) {
        u_1 += a.step;
        steps++;
    }
    cap[i] = 1 + steps;
}

// One history per lane, grid-stride.  FP64 throughout, with the reference's float powf roundings.
__global__ void __launch_bounds__(kBlock) k_mlp_lists_std(int64_t n, StdArgs a, H10 hp,
                                                          const uint32_t* __restrict__ first,
                                                          const int64_t* __restrict__ off,
                                                          uint32_t* __restrict__ cnt_out, uint32_t* __restrict__ vox,
                                                          double* __restrict__ chord, int32_t* overflow)
{
    __shared__ uint32_t list_buf[16 * kBlock];
    const DevGeom& g = a.g;
    const double STEP = a.step, E0 = a.e0, X0 = a.x0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double x_in = hp.h[0][i], y_in = hp.h[1][i], z_in = hp.h[2][i];
        const double x_out = hp.h[3][i], y_out = hp.h[4][i], z_out = hp.h[5][i];
        const double xy_e0 = hp.h[6][i], xz_entry = hp.h[7][i], xy_x0 = hp.h[8][i], xz_exit = hp.h[9][i];
        // the chord comes from the angles before the +pi turn (DS.cu:8921 precedes :8945)
        chord[i] = chord_length(g.PI, g.VW, g.VT, STEP, (xy_e0 + xy_x0) / 2.0, (xz_entry + xz_exit) / 2.0);
        StdFrame f = std_frame(g.PI, x_in, y_in, x_out, y_out, xy_e0, xy_x0);
        const double T0_0 = f.t_in, T0_1 = 0.0, T2_0 = f.t_out, T2_1 = f.xy_exit - f.xy_entry;
        const double V0_0 = z_in, V0_1 = 0.0, V2_0 = z_out, V2_1 = xz_exit - xz_entry;
        const double u_0 = 0.0, u_2 = fabs(f.u_out - f.u_in);
        const double pre1 = u_2 * u_2 * u_2 * std_poly(a, a.f312, u_2);
        const double pre2 = u_2 * u_2 * std_poly(a, a.f26, u_2);
        const double pre3 = u_2 * std_poly(a, a.f12, u_2);
        const double u2_sq = (double)pct_sqf(u_2);
        const double R0_0 = 1.0, R0_2 = 0.0, R0_3 = 1.0, R1_0 = 1.0, R1_2 = 0.0, R1_3 = 1.0;
        const double R1T_0 = 1.0, R1T_1 = 0.0, R1T_3 = 1.0;
        const int64_t base = off[i], capn = off[i + 1] - off[i];
        uint32_t last = first[i];
        int64_t cnt = 1;
        ListWriter w(vox, list_buf, base);
        w.put(last);
        double u_1 = STEP;
        int k = 0;
        while (u_1 < u_2 - STEP) {
            if (k >= a.max_steps) {
                atomicOr(overflow, 1);
                break;
            }
            const double R0_1 = u_1, R1_1 = u_2 - u_1, R1T_2 = u_2 - u_1;
            const double k1 = (double)pct_sqf(E0 * (1 + 0.038 * log((u_1 - u_0) / X0))) / X0;
            const double st1 = k1 * ((double)std_cubef(u_1) * std_poly(a, a.f312, u_1));
            const double st1th1 = k1 * ((double)pct_sqf(u_1) * std_poly(a, a.f26, u_1));
            const double sth1 = k1 * (u_1 * std_poly(a, a.f12, u_1));
            const double det1 = st1 * sth1 - (double)pct_sqf(st1th1);
            const double S1_0 = sth1 / det1, S1_1 = -st1th1 / det1, S1_2 = -st1th1 / det1, S1_3 = st1 / det1;
            const double k2 = (double)pct_sqf(E0 * (1 + 0.038 * log((u_2 - u_1) / X0))) / X0;
            const double cm1 = u_1 * std_poly(a, a.f12, u_1);
            const double cm2 = (double)pct_sqf(u_1) * std_poly(a, a.f23, u_1);
            const double cm3 = (double)std_cubef(u_1) * std_poly(a, a.f34, u_1);
            const double st2 = k2 * (pre1 - u2_sq * cm1 + 2 * u_2 * cm2 - cm3);
            const double st2th2 = k2 * (pre2 - u_2 * cm1 + cm2);
            const double sth2 = k2 * (pre3 - cm1);
            const double det2 = st2 * sth2 - (double)pct_sqf(st2th2);
            const double S2_0 = sth2 / det2, S2_1 = -st2th2 / det2, S2_2 = -st2th2 / det2, S2_3 = st2 / det2;
            const double c13_1 = S2_0 * R1_0 + S2_1 * R1_2;
            const double c13_2 = S2_2 * R1_0 + S2_3 * R1_2;
            const double c24_1 = S2_0 * R1_1 + S2_1 * R1_3;
            const double c24_2 = S2_2 * R1_1 + S2_3 * R1_3;
            double ft0 = S1_0 + R1T_0 * c13_1 + R1T_1 * c13_2;
            double ft1 = S1_1 + R1T_0 * c24_1 + R1T_1 * c24_2;
            double ft2 = S1_2 + R1T_2 * c13_1 + R1T_3 * c13_2;
            double ft3 = S1_3 + R1T_2 * c24_1 + R1T_3 * c24_2;
            const double detft = ft0 * ft3 - ft1 * ft2;
            const double tmp = ft0;
            ft0 = ft3 / detft;
            ft1 = -ft1 / detft;
            ft2 = -ft2 / detft;
            ft3 = tmp / detft;
            double sc1 = R0_0 * T0_0 + R0_1 * T0_1;
            double sc2 = R0_2 * T0_0 + R0_3 * T0_1;
            double sc3 = S2_0 * T2_0 + S2_1 * T2_1;
            double sc4 = S2_2 * T2_0 + S2_3 * T2_1;
            double s0 = S1_0 * sc1 + S1_1 * sc2 + R1T_0 * sc3 + R1T_1 * sc4;
            double s1 = S1_2 * sc1 + S1_3 * sc2 + R1T_2 * sc3 + R1T_3 * sc4;
            const double t_1 = ft0 * s0 + ft1 * s1;
            sc1 = R0_0 * V0_0 + R0_1 * V0_1;
            sc2 = R0_2 * V0_0 + R0_3 * V0_1;
            sc3 = S2_0 * V2_0 + S2_1 * V2_1;
            sc4 = S2_2 * V2_0 + S2_3 * V2_1;
            s0 = S1_0 * sc1 + S1_1 * sc2 + R1T_0 * sc3 + R1T_1 * sc4;
            s1 = S1_2 * sc1 + S1_3 * sc2 + R1T_2 * sc3 + R1T_3 * sc4;
            const double v_1 = ft0 * s0 + ft1 * s1;
            const double x_1 = (f.c * (f.u_in + u_1)) - (f.s * t_1);
            const double y_1 = (f.s * (f.u_in + u_1)) + (f.c * t_1);
            const int vx = pct_calc_voxel(g.X0, x_1, g.VW);
            const int vy = pct_calc_voxel(g.Y0, y_1, g.VH);
            const int vz = pct_calc_voxel(g.Z0, v_1, g.VT);
            const int64_t voxel = vx + (int64_t)vy * g.C + (int64_t)vz * g.C * g.R;
            if (voxel != (int64_t)last && voxel < g.NV) {
                if (cnt >= capn) {
                    atomicOr(overflow, 2);
                    break;
                }
                w.put((uint32_t)voxel);
                cnt++;
                last = (uint32_t)voxel;
            }
            u_1 += STEP;
            k++;
        }
        w.flush();
        cnt_out[i] = (uint32_t)cnt;
    }
}

// total of the per-history list counts (one 64-bit atomic per wavefront)
__global__ void k_count_entries(int64_t n, const uint32_t* __restrict__ cnt, unsigned long long* total)
{
    unsigned long long part = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) part += cnt[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    if ((threadIdx.x & 63) == 0 && part) atomicAdd(total, part);
}

StdArgs std_args(const pct_params& P, int max_steps)
{
    StdArgs a;
    a.g = pct_geom(P);
    a.step = P.mlp_u_step;
    a.e0 = P.e_0;
    a.x0 = P.x_0_rad_length;
    const double c[6] = {7.457, 4.548, -5.777, 1.301, -9.228, 2.687};  // DS.h:749-754
    const float Pw[6] = {std::pow(10.0f, -6.0f), std::pow(10.0f, -7.0f), std::pow(10.0f, -8.0f),
                         std::pow(10.0f, -8.0f), std::pow(10.0f, -10.0f), std::pow(10.0f, -11.0f)};
    auto fam = [&](const int* d) {
        PolyFam f;
        for (int k = 0; k < 5; k++) f.a[k] = c[k] * Pw[k] / d[k];
        f.d5 = d[5];
        return f;
    };
    static const int d12[6] = {1, 2, 3, 4, 5, 6}, d23[6] = {2, 3, 4, 5, 6, 7}, d34[6] = {3, 4, 5, 6, 7, 8};
    static const int d26[6] = {2, 6, 12, 20, 30, 42}, d312[6] = {3, 12, 30, 60, 105, 168};
    a.c5 = c[5];
    a.p5 = Pw[5];
    a.f12 = fam(d12);
    a.f23 = fam(d23);
    a.f34 = fam(d34);
    a.f26 = fam(d26);
    a.f312 = fam(d312);
    a.max_steps = max_steps;
    return a;
}

}  // namespace

// reconstruction_cuts (DS.cu:8296): endpoints on the selected hull, keep histories that hit it
int pct_stage_endpoints(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n_hist;
    DevGeom g = pct_geom(P);
    H10 hp;
    for (int k = 0; k < 10; k++) hp.h[k] = ctx->h[k];
    const uint32_t* perm = nullptr;
    if (n > 0 && n <= 0x7fffffff && ctx->endpoint_order) {  // the bin-order sort takes up to 2^31 - 1
        PCT_TRY(pct_bin_order(ctx, n));
        perm = ctx->u32_d;
    }
    pct_tmp<uint32_t> bits;
    if (n > 0) {
        const int64_t NV = P.num_voxels, words = (NV + 31) / 32;
        PCT_TRY(pct_alloc(ctx, &bits.p, words));
        hipLaunchKernelGGL(k_hull_bits, dim3(grid_for(words)), dim3(kBlock), 0, s, NV,
                           ctx->hull_endpoints ? ctx->hull_endpoints : ctx->hull, bits.p);
        hipLaunchKernelGGL(k_endpoints, dim3(grid_for(n)), dim3(kBlock), 0, s, n, perm, g, bits.p, hp,
                           ctx->first_voxel, ctx->flag);
    }
    PCT_CHECK_HIP(ctx, hipGetLastError());
    int64_t kept = 0;
    PCT_TRY(pct_compact(ctx, ctx->flag, n, &kept, true));  // DS.cu:8286-8294
    ctx->n_hist = kept;
    ctx->n_recon = kept;
    ctx->stage = ST_ENDPOINTS;
    return PCT_OK;
}

static int upload_tables(pct_ctx* ctx)
{
    if (ctx->T.sin_t) return PCT_OK;
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    HostTables ht;
    pct_build_tables(P, &ht);
    int n_steps = (int)(2 * P.recon_cyl_radius / P.mlp_u_step) + 8;
    std::vector<double> terms;
    pct_build_step_terms(P, ht, n_steps, &terms);
    DevTables& T = ctx->T;
    T.n_trig = (int64_t)ht.sin_t.size();
    T.n_coeff = (int64_t)ht.scat.size();
    T.n_poly = (int64_t)ht.p12.size();
    T.n_steps = n_steps;
    PCT_TRY(pct_alloc(ctx, &T.sin_t, ht.sin_t.size()));
    PCT_TRY(pct_alloc(ctx, &T.cos_t, ht.cos_t.size()));
    PCT_TRY(pct_alloc(ctx, &T.scat, ht.scat.size()));
    PCT_TRY(pct_alloc(ctx, &T.p12, ht.p12.size()));
    PCT_TRY(pct_alloc(ctx, &T.p26, ht.p26.size()));
    PCT_TRY(pct_alloc(ctx, &T.p312, ht.p312.size()));
    PCT_TRY(pct_alloc(ctx, &T.step_terms, terms.size()));
    auto up = [&](double* d, const std::vector<double>& h) {
        return hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s);
    };
    PCT_CHECK_HIP(ctx, up(T.sin_t, ht.sin_t));
    PCT_CHECK_HIP(ctx, up(T.cos_t, ht.cos_t));
    PCT_CHECK_HIP(ctx, up(T.scat, ht.scat));
    PCT_CHECK_HIP(ctx, up(T.p12, ht.p12));
    PCT_CHECK_HIP(ctx, up(T.p26, ht.p26));
    PCT_CHECK_HIP(ctx, up(T.p312, ht.p312));
    PCT_CHECK_HIP(ctx, up(T.step_terms, terms));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    return PCT_OK;
}

int pct_stage_build_lists(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n_hist;
    PCT_TRY(upload_tables(ctx));
    if (!ctx->seg_inside && pct_segment_histories(ctx) > 0) {
        // lists are built per segment inside each DROP iteration and residual (segments.hip)
        ctx->list_entries = 0;
        ctx->plan.ready = false;
        ctx->stage = ST_LISTS;
        return PCT_OK;
    }
    MlpArgs a;
    a.g = pct_geom(P);
    a.trig_min = P.trig_table_min;
    a.trig_step = P.trig_table_step;
    a.step = P.mlp_u_step;
    a.coeff_step = P.coeff_table_step;
    a.poly_step = P.poly_table_step;
    a.n_trig = ctx->T.n_trig;
    a.n_poly = ctx->T.n_poly;
    a.n_coeff = ctx->T.n_coeff;
    a.s_step = (unsigned)(P.mlp_u_step / P.coeff_table_step + 0.5);
    a.p_step = (unsigned)(P.mlp_u_step / P.poly_table_step + 0.5);
    a.n_steps = ctx->T.n_steps;
    H10 hp;
    for (int k = 0; k < 10; k++) hp.h[k] = ctx->h[k];
    if (n > ctx->list_hist_cap || !ctx->list_off) {  // buffers persist across rebuilds
        PCT_TRY(pct_alloc(ctx, &ctx->list_off, n + 1));
        PCT_TRY(pct_alloc(ctx, &ctx->list_cnt, n));
        PCT_TRY(pct_alloc(ctx, &ctx->chord, n));
        PCT_TRY(pct_alloc(ctx, &ctx->cap_scratch, n + 1));
        ctx->list_hist_cap = n;
    }
    if (!ctx->list_overflow) PCT_TRY(pct_alloc(ctx, &ctx->list_overflow, 1));
    PCT_CHECK_HIP(ctx, hipMemsetAsync(ctx->list_overflow, 0, sizeof(int32_t), s));
    // capacities -> exclusive scan -> offsets
    int64_t* capv = ctx->cap_scratch;
    PCT_CHECK_HIP(ctx, hipMemsetAsync(capv, 0, (n + 1) * sizeof(int64_t), s));
    const bool standard = P.mlp_algorithm == PCT_MLP_STANDARD;
    const StdArgs sa = std_args(P, 4 * a.n_steps);
    if (n > 0 && standard)
        hipLaunchKernelGGL(k_mlp_capacity_std, dim3(grid_for(n)), dim3(kBlock), 0, s, n, sa, hp, capv);
    else if (n > 0)
        hipLaunchKernelGGL(k_mlp_capacity, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, hp, ctx->T.sin_t,
                           ctx->T.cos_t, capv);
    size_t need = 0;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, capv, ctx->list_off, n + 1, s));
    if (need > ctx->cub_tmp_bytes) {
        PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, need));
        ctx->cub_tmp_bytes = need;
    }
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, need, capv, ctx->list_off, n + 1, s));
    int64_t total = 0;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&total, ctx->list_off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (total > ctx->list_cap) {
        PCT_TRY(pct_alloc(ctx, &ctx->list_vox, total));
        ctx->list_cap = total;
    }
    ctx->plan.vox_b_ready = false;  // the brick copy of the lists belongs to the previous build
    const size_t lds = (size_t)a.n_steps * 6 * sizeof(double);
    if (lds > 64 * 1024) {
        ctx->err = "MLP: too many u-steps for the LDS step table";
        return PCT_E_LIMIT;
    }
    int dev_cus = 256;
    (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    // the MLP list kernel over the histories [h0, h1): per-history arrays offset by h0, list offsets absolute
    const MlpLaunch launch = [&, lds, dev_cus](int64_t h0, int64_t h1, hipStream_t st) -> int {
        const int64_t m = h1 - h0;
        if (m <= 0) return PCT_OK;
        H10 hq;
        for (int k = 0; k < 10; k++) hq.h[k] = ctx->h[k] + h0;
        const int64_t want = (m + kBlock - 1) / kBlock;
        // persistent grid (8 workgroups per CU) when the kernel has the GPU to itself; one lane per history
        // when it shares the GPU with the transposition, so that the two streams' workgroups interleave
        const int per_cu = ctx->mlp_blocks_per_cu >= 0 ? ctx->mlp_blocks_per_cu : (st == ctx->stream ? 8 : 0);
        const int64_t cap_blocks = per_cu > 0 ? (int64_t)dev_cus * per_cu : want;
        const unsigned blocks = (unsigned)(want < cap_blocks ? want : cap_blocks);
        if (standard)
            hipLaunchKernelGGL(k_mlp_lists_std, dim3(blocks), dim3(kBlock), 0, st, m, sa, hq, ctx->first_voxel + h0,
                               ctx->list_off + h0, ctx->list_cnt + h0, ctx->list_vox, ctx->chord + h0,
                               ctx->list_overflow);
        else
            hipLaunchKernelGGL(k_mlp_lists, dim3(blocks), dim3(kBlock), lds, st, m, a, hq, ctx->first_voxel + h0,
                               ctx->T.sin_t, ctx->T.cos_t, ctx->T.scat, ctx->T.p12, ctx->T.p26, ctx->T.p312,
                               ctx->T.step_terms, ctx->list_off + h0, ctx->list_cnt + h0, ctx->list_vox,
                               ctx->chord + h0, ctx->list_overflow);
        PCT_CHECK_HIP(ctx, hipGetLastError());
        return PCT_OK;
    };
    // Resident lists with the transposed DROP plan: the plan build runs the MLP kernel group by group on a
    // second stream, overlapping the transposition of the groups before (no segments, no bucket
    // transposition: those build their plan after the lists).
    const bool pipe = ctx->drop_pipeline && !ctx->seg_inside && !ctx->seg_lists_only &&
                      !(ctx->drop_flags & PCT_DROP_BUCKETS) && n > 0;
    if (pipe) {
        ctx->stage = ST_LISTS;
        ctx->plan.ready = false;
        const int rc = pct_drop_prepare(ctx, &launch);
        if (rc != PCT_OK) {
            ctx->stage = ST_ENDPOINTS;  // the lists are incomplete
            return rc;
        }
    } else if (n > 0) {
        hipEvent_t e0, e1;
        PCT_CHECK_HIP(ctx, hipEventCreate(&e0));
        PCT_CHECK_HIP(ctx, hipEventCreate(&e1));
        PCT_CHECK_HIP(ctx, hipEventRecord(e0, s));
        PCT_TRY(launch(0, n, s));
        PCT_CHECK_HIP(ctx, hipEventRecord(e1, s));
        PCT_CHECK_HIP(ctx, hipEventSynchronize(e1));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ctx->t_mlp_kernel = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    int32_t overflow = 0;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&overflow, ctx->list_overflow, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (overflow) {
        ctx->err = "MLP list capacity exceeded (code " + std::to_string(overflow) + ")";
        ctx->plan.ready = false;  // a plan built with the lists holds them
        ctx->stage = ST_ENDPOINTS;
        return PCT_E_LIMIT;
    }
    // entries actually listed (the capacity `total` counts every u-step)
    unsigned long long* d_total = (unsigned long long*)ctx->cap_scratch;  // free after the scan
    PCT_CHECK_HIP(ctx, hipMemsetAsync(d_total, 0, sizeof(unsigned long long), s));
    if (n > 0)
        hipLaunchKernelGGL(k_count_entries, dim3(pct_grid_capped(n, kBlock, 2048)), dim3(kBlock), 0, s, n,
                           ctx->list_cnt, d_total);
    unsigned long long listed = 0;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&listed, d_total, sizeof(listed), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    ctx->list_entries = (int64_t)listed;
    ctx->list_slots = total;
    ctx->stage = ST_LISTS;
    if (pipe) return PCT_OK;  // the plan was built with the lists
    // the DROP blocks' transposed lists belong to the list build
    ctx->plan.ready = false;
    if (!ctx->seg_lists_only) PCT_TRY(pct_drop_prepare(ctx));
    return PCT_OK;
}
