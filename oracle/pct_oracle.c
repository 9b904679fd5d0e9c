/*
 * pct_oracle.c — CPU restatement of the reference's reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 *
 * Reference file for all line citations: src/pCT_Reconstruction_Data_Segments_Blake.cu (DS.cu),
 * constants in include/pCT_Reconstruction_Data_Segments_Blake.h (DS.h), derived parameters in
 * include/Configurations.h (Cfg.h).  Build with -ffp-contract=off: the reference's expressions are
 * evaluated operation by operation, in the C types the reference uses (float + float stays float).
 *
 * Meaning given to CUDA float functions (parity unpinned by any reference test; see the header):
 *   powf(x, 2) = x*x rounded to float; powf(x, -1) = 1/x in float;
 *   sinf/cosf/tanf/atan2f(x) = the double function of the float argument, rounded to float.
 */
#include "pct_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef O_FP64
/* FP64 twin (pct_oracle_fp64.c): every value the reference holds or computes in float is double, the
 * CUDA float functions are their double versions, and the entry points carry the suffix _f64. */
typedef double real;
#define O(name) name##_f64
#define r_fabs fabs
#define r_sqrt sqrt
static inline real f_sq(double x) { return x * x; }
static inline real f_atan2(real y, real x) { return atan2(y, x); }
static inline real f_sin(double x) { return sin(x); }
static inline real f_cos(double x) { return cos(x); }
static inline real f_tan(double x) { return tan(x); }
static inline real f_recip(double x) { return 1.0 / x; }
#else
typedef float real;
#define O(name) name
#define r_fabs fabsf
#define r_sqrt sqrtf
static inline float f_sq(double x) { float f = (float)x; return f * f; }
static inline float f_atan2(float y, float x) { return (float)atan2((double)y, (double)x); }
static inline float f_sin(double x) { return (float)sin((double)(float)x); }
static inline float f_cos(double x) { return (float)cos((double)(float)x); }
static inline float f_tan(double x) { return (float)tan((double)(float)x); }
static inline float f_recip(double x) { float f = (float)x; return 1.0f / f; }
#endif
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ parameters (Cfg.h:1745-1789) */
void O(o_derive)(pct_params* p)
{
    p->pi = 4 * atan(1.0);                                                   /* DS.h:808 PI */
    p->t_bins = (int32_t)(p->ssd_t_size / p->t_bin_size + 0.5);              /* Cfg.h:1749 */
    p->v_bins = (int32_t)(p->ssd_v_size / p->v_bin_size + 0.5);              /* Cfg.h:1750 */
    p->angular_bins = (int32_t)(360 / p->angular_bin_size + 0.5);            /* Cfg.h:1751 */
    p->num_bins = p->angular_bins * p->t_bins * p->v_bins;                   /* Cfg.h:1752 */
    p->recon_cyl_height = p->ssd_v_size - 1.0;                               /* Cfg.h:1753 */
    p->slices = (int32_t)(p->recon_cyl_height / p->slice_thickness);         /* Cfg.h:1755 */
    p->num_voxels = p->columns * p->rows * p->slices;                        /* Cfg.h:1756 */
    p->image_width = 2 * p->recon_cyl_radius;                                /* Cfg.h:1754,1757 */
    p->image_height = 2 * p->recon_cyl_radius;
    p->image_thickness = p->recon_cyl_height;                                /* Cfg.h:1759 */
    p->voxel_width = p->image_width / p->columns;                            /* Cfg.h:1760 */
    p->voxel_height = p->image_height / p->rows;
    p->voxel_thickness = p->image_thickness / p->slices;                     /* Cfg.h:1762 */
    p->x_zero_coordinate = -p->recon_cyl_radius;                             /* Cfg.h:1763-1765 */
    p->y_zero_coordinate = p->recon_cyl_radius;
    p->z_zero_coordinate = p->recon_cyl_height / 2;
    p->ram_lak_tau = (double)(2 / sqrtf(2.0f)) * p->t_bin_size;              /* Cfg.h:1766, ROOT_TWO = sqrtf(2.0) */
    p->mlp_u_step = p->voxel_width / 2;                                      /* Cfg.h:1786, DS.h:411 */
    p->trig_table_min = p->trig_table_min_pis * p->pi;                       /* DS.h:795 */
    p->trig_table_max = p->trig_table_max_pis * p->pi;                       /* DS.h:796 */
    p->trig_table_step = p->trig_table_step_degrees * (p->pi / 180.0);       /* DS.h:798 */
    p->trig_table_elements = (int32_t)((p->trig_table_max - p->trig_table_min) / p->trig_table_step + 0.5); /* DS.h:799 */
    p->coeff_table_elements = (int32_t)(p->coeff_table_range / p->coeff_table_step + 0.5);                 /* DS.h:803 */
    p->poly_table_elements = (int32_t)(p->poly_table_range / p->poly_table_step + 0.5);                    /* DS.h:807 */
}

/* ------------------------------------------------------------------ (a) intersections */
/* calculate_intercepts, DS.cu:5286-5372 */
static int o_calculate_intercepts(const pct_params* p, double u, double t, double ut_angle, double* u_int, double* t_int)
{
    const double PI_OVER_4 = atan(1.0), THREE_PI_OVER_4 = 3 * atan(1.0);
    const double FIVE_PI_OVER_4 = 5 * atan(1.0), SEVEN_PI_OVER_4 = 7 * atan(1.0), PI_OVER_2 = 2 * atan(1.0);
    int in_cone = ((ut_angle > PI_OVER_4) && (ut_angle < THREE_PI_OVER_4)) ||
                  ((ut_angle > FIVE_PI_OVER_4) && (ut_angle < SEVEN_PI_OVER_4));       /* :5314 */
    if (in_cone) {                                                                      /* :5318-5324 */
        double u_temp = u;
        u = -t;
        t = u_temp;
        ut_angle += PI_OVER_2;
    }
    double m = tan(ut_angle);                                                           /* :5325 */
    double b_in = t - m * u;
    double a = (double)(1 + f_sq(m));                                  /* 1 + powf(m,2): float add */
    double b = 2 * m * b_in;
    double c = (double)(f_sq(b_in) - f_sq(p->recon_cyl_radius));       /* float - float */
    double disc = (double)f_sq(b) - (4 * a * c);                                        /* :5337 */
    int intersected = (disc > 0);
    if (intersected) {                                                                  /* :5352-5362 */
        double u1 = (sqrt(disc) - b) / (2 * a);
        double u2 = (sqrt(disc) + b) / (2 * a);
        double t1 = m * u1 + b_in;
        double t2 = m * u2 - b_in;
        double d1 = (double)(f_sq(u1 - u) + f_sq(t1 - t));               /* float sum of powf */
        double d2 = (double)(f_sq(u2 + u) + f_sq(t2 + t));
        *u_int = u1 * (double)(d1 <= d2) - u2 * (double)(d1 > d2);
        *t_int = t1 * (double)(d1 <= d2) - t2 * (double)(d1 > d2);
    }
    if (in_cone) {                                                                      /* :5364-5370 */
        double u_temp = *u_int;
        *u_int = *t_int;
        *t_int = -u_temp;
    }
    return intersected;
}

/* recon_volume_intersections_GPU, DS.cu:4890-5021 (identical to the _exact_angles variant :5022) */
void O(o_intersections)(const pct_params* p, int64_t n, const real* const t[4], const real* const v[4],
                     const real* const u[4], const float* angle_deg, real* const h[10], uint8_t* missed)
{
    const double A2R = p->pi / 180.0, H = p->recon_cyl_height;
    real *xe = h[0], *ye = h[1], *ze = h[2], *xx = h[3], *yx = h[4], *zx = h[5];
    real *xya_e = h[6], *xza_e = h[7], *xya_x = h[8], *xza_x = h[9];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double rot = (double)angle_deg[i] * A2R;                                         /* :4920 */
        double ut_entry_angle = (double)f_atan2(t[1][i] - t[0][i], u[1][i] - u[0][i]);  /* :4935 */
        double u_entry = 0.0, t_entry = 0.0;
        int entered = o_calculate_intercepts(p, u[1][i], t[1][i], ut_entry_angle, &u_entry, &t_entry);
        xya_e[i] = (real)(ut_entry_angle + rot);
        xe[i] = (real)((cos(rot) * u_entry) - (sin(rot) * t_entry));                   /* :4946 */
        ye[i] = (real)((sin(rot) * u_entry) + (cos(rot) * t_entry));
        double ut_exit_angle = (double)f_atan2(t[3][i] - t[2][i], u[3][i] - u[2][i]);   /* :4951 */
        double u_exit = 0.0, t_exit = 0.0;
        int exited = o_calculate_intercepts(p, u[2][i], t[2][i], ut_exit_angle, &u_exit, &t_exit);
        xya_x[i] = (real)(ut_exit_angle + rot);
        xx[i] = (real)((cos(rot) * u_exit) - (sin(rot) * t_exit));
        yx[i] = (real)((sin(rot) * u_exit) + (cos(rot) * t_exit));
        double uv_entry_slope = (double)((v[1][i] - v[0][i]) / (u[1][i] - u[0][i]));     /* :4967: float / float */
        double uv_exit_slope = (double)((v[3][i] - v[2][i]) / (u[3][i] - u[2][i]));
        xza_e[i] = f_atan2(v[1][i] - v[0][i], u[1][i] - u[0][i]);                       /* :4970 */
        xza_x[i] = f_atan2(v[3][i] - v[2][i], u[3][i] - u[2][i]);
        u_entry = (cos(rot) * (double)xe[i]) + (sin(rot) * (double)ye[i]);              /* :4982 */
        u_exit = (cos(rot) * (double)xx[i]) + (sin(rot) * (double)yx[i]);
        ze[i] = (real)((double)v[1][i] + uv_entry_slope * (u_entry - (double)u[1][i])); /* :4984 */
        zx[i] = (real)((double)v[2][i] - uv_exit_slope * ((double)u[2][i] - u_exit));
        if (entered && exited) {                                                         /* :4993 */
            if ((r_fabs(ze[i]) < H * 0.5) && (r_fabs(zx[i]) > H * 0.5)) {
                double sgn = (double)((zx[i] >= 0) - (zx[i] < 0));
                double frac = fabs((sgn * H * 0.5 - ze[i]) / (double)(zx[i] - ze[i]));
                xx[i] = (real)((double)xe[i] + frac * (double)(xx[i] - xe[i]));
                yx[i] = (real)((double)ye[i] + frac * (double)(yx[i] - ye[i]));
                zx[i] = (real)(sgn * H * 0.5);
            } else if (r_fabs(ze[i]) > H * 0.5) {
                entered = 0;
                exited = 0;
            }
            if ((r_fabs(ze[i]) > H * 0.5) && (r_fabs(zx[i]) < H * 0.5)) {                 /* :5003 */
                double sgn_x = (double)((zx[i] >= 0) - (zx[i] < 0));
                double sgn_e = (double)((ze[i] >= 0) - (ze[i] < 0));
                double frac = fabs((sgn_x * H * 0.5 - zx[i]) / (double)(zx[i] - ze[i]));
                xe[i] = (real)((double)xx[i] + frac * (double)(xx[i] - xe[i]));
                ye[i] = (real)((double)yx[i] + frac * (double)(yx[i] - ye[i]));
                ze[i] = (real)(sgn_e * H * 0.5);
            }
            if ((fabsf(t[2][i] - t[1][i]) > 5) || (fabsf(v[2][i] - v[1][i]) > 5)) {     /* :5011 */
                entered = 0;
                exited = 0;
            }
        }
        missed[i] = (uint8_t)(!entered || !exited);                                     /* :5018 */
    }
}

/* ------------------------------------------------------------------ (a) binning */
/* binning_GPU, DS.cu:5482-5553 */
void O(o_binning)(const pct_params* p, int64_t n, const real* const h[10], int32_t* bin_num, uint8_t* in_range)
{
    const double PI = p->pi, R2A = 180.0 / PI, A2R = PI / 180.0;
    const int T = p->t_bins, V = p->v_bins, A = p->angular_bins;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double x_mid = (double)((h[0][i] + h[3][i]) / 2);                                /* :5497: float math */
        double y_mid = (double)((h[1][i] + h[4][i]) / 2);
        double z_mid = (double)((h[2][i] + h[5][i]) / 2);
        double path_angle = (double)f_atan2(h[4][i] - h[1][i], h[3][i] - h[0][i]);       /* :5502 */
        if (path_angle < 0) path_angle += 2 * PI;
        int angle_bin = (int)((path_angle * R2A / p->angular_bin_size) + 0.5) % A;      /* :5505 */
        double angle = angle_bin * p->angular_bin_size * A2R;
        double tt = y_mid * cos(angle) - x_mid * sin(angle);                             /* :5508 */
        int t_bin = (int)((tt / p->t_bin_size) + T / 2);
        int v_bin = (int)((z_mid / p->v_bin_size) + V / 2);
        if ((t_bin >= 0) && (v_bin >= 0) && (t_bin < T) && (v_bin < V)) {
            bin_num[i] = t_bin + angle_bin * T + v_bin * T * A;                          /* :5516 */
            in_range[i] = 1;
        } else {
            bin_num[i] = 0;            /* quirk A1: calloc'd bin_num_h keeps 0 (DS.cu:5380) */
            in_range[i] = 0;
        }
    }
}

static inline double o_wrap(double a, double PI)
{
    if (a > PI) a -= 2 * PI;
    if (a < -PI) a += 2 * PI;
    return a;
}

void O(o_stats_and_cuts)(const pct_params* p, int64_t n, const int32_t* bin_num, const uint8_t* in_range,
                      const real* wepl, const real* const h[10], int32_t* counts_pre, real* const mean[3],
                      real* const stddev[3], int32_t* counts_post, real* sinogram, uint8_t* failed)
{
    const double PI = p->pi;
    const int64_t NB = p->num_bins;
    int32_t* cnt = (int32_t*)calloc((size_t)NB, sizeof(int32_t));
    for (int k = 0; k < 3; k++) {
        memset(mean[k], 0, (size_t)NB * sizeof(real));
        memset(stddev[k], 0, (size_t)NB * sizeof(real));
    }
    /* binning sums, DS.cu:5518-5546 (float atomics in history order) */
    for (int64_t i = 0; i < n; i++) {
        if (!in_range[i]) continue;
        int b = bin_num[i];
        double rel_ut = o_wrap((double)(h[8][i] - h[6][i]), PI);
        double rel_uv = o_wrap((double)(h[9][i] - h[7][i]), PI);
        cnt[b] += 1;
        mean[0][b] += wepl[i];
        mean[1][b] += (real)rel_ut;
        mean[2][b] += (real)rel_uv;
    }
    /* calculate_means_GPU, DS.cu:5644-5654 */
    for (int64_t b = 0; b < NB; b++) {
        if (cnt[b] > 0) {
            mean[0][b] /= cnt[b];
            mean[1][b] /= cnt[b];
            mean[2][b] /= cnt[b];
        }
        counts_pre[b] = cnt[b];
    }
    /* sum_squared_deviations_GPU, DS.cu:5689-5717 (all kept histories, incl. quirk-A1 bin 0) */
    for (int64_t i = 0; i < n; i++) {
        if (!p->quirk_a1_bin0 && !in_range[i]) continue;
        int b = bin_num[i];
        double rel_ut = o_wrap((double)(h[8][i] - h[6][i]), PI);
        double rel_uv = o_wrap((double)(h[9][i] - h[7][i]), PI);
        double d_w = (double)(wepl[i] - mean[0][b]);
        double d_ut = rel_ut - mean[1][b];
        double d_uv = rel_uv - mean[2][b];
        stddev[0][b] += f_sq(d_w);
        stddev[1][b] += f_sq(d_ut);
        stddev[2][b] += f_sq(d_uv);
    }
    /* calculate_standard_deviations_GPU, DS.cu:5737-5750 (SAMPLE_STD_DEV = true, DS.h:224) */
    for (int64_t b = 0; b < NB; b++) {
        if (cnt[b] > 1) {
            for (int k = 0; k < 3; k++) stddev[k][b] = r_sqrt(stddev[k][b] / (cnt[b] - 1));
        }
        cnt[b] = 0;
    }
    /* statistical_cuts_GPU, DS.cu:5937-5971; initialize_sinogram :6026 */
    memset(sinogram, 0, (size_t)NB * sizeof(real));
    const real k = (real)p->sigmas_2_keep;
    for (int64_t i = 0; i < n; i++) {
        if (!p->quirk_a1_bin0 && !in_range[i]) {
            failed[i] = 1;
            continue;
        }
        int b = bin_num[i];
        double rel_ut = o_wrap((double)(h[8][i] - h[6][i]), PI);
        double rel_uv = o_wrap((double)(h[9][i] - h[7][i]), PI);
        int pass_ut = fabs(rel_ut - mean[1][b]) < (double)(k * stddev[1][b]);
        int pass_uv = fabs(rel_uv - mean[2][b]) < (double)(k * stddev[2][b]);
        int pass_w = r_fabs(mean[0][b] - wepl[i]) <= k * stddev[0][b];
        failed[i] = (uint8_t)(!pass_ut || !pass_uv || !pass_w);
        if (!failed[i]) {
            cnt[b] += 1;
            sinogram[b] += wepl[i];
        }
    }
    for (int64_t b = 0; b < NB; b++) counts_post[b] = cnt[b];
    free(cnt);
}

/* construct_sinogram_GPU, DS.cu:6073-6079 */
void O(o_construct_sinogram)(const pct_params* p, const int32_t* counts_post, real* sinogram)
{
    for (int64_t b = 0; b < p->num_bins; b++)
        if (counts_post[b] > 0) sinogram[b] /= counts_post[b];
}

/* ------------------------------------------------------------------ (b) hull */
/* initialize_hull_GPU, DS.cu:6659-6669 */
static int o_inside_cylinder(const pct_params* p, int column, int row)
{
    double x = (column - p->columns / 2 + 0.5) * p->voxel_width;
    double y = (p->rows / 2 - row - 0.5) * p->voxel_height;
    return (f_sq(x) + f_sq(y)) < f_sq(p->recon_cyl_radius);
}
void O(o_init_hull_u8)(const pct_params* p, uint8_t* hull)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    for (int s = 0; s < S; s++)
        for (int r = 0; r < R; r++)
            for (int c = 0; c < C; c++) hull[c + r * C + (int64_t)s * C * R] = (uint8_t)o_inside_cylinder(p, c, r);
}
void O(o_init_hull_i32)(const pct_params* p, int32_t* counts)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    for (int s = 0; s < S; s++)
        for (int r = 0; r < R; r++)
            for (int c = 0; c < C; c++) counts[c + r * C + (int64_t)s * C * R] = o_inside_cylinder(p, c, r);
}

/* calculate_voxel_GPU :11163, edge_coordinate_GPU :11322, distance_remaining_GPU :11313,
 * corresponding_coordinate_GPU :11340 */
static inline int o_calc_voxel(double zero, double pos, double size) { return (int)(fabs(pos - zero) / size); }
static inline double o_edge(double zero, int voxel_entered, double size, int inc, int step)
{
    int on_edge = (step == inc) ? voxel_entered : voxel_entered + 1;
    return zero + (inc * on_edge * size);
}
static inline double o_dist(double zero, double pos, int inc, int step, double size, int cur)
{
    int next = cur + inc * step;
    return fabs(o_edge(zero, next, size, inc, step) - pos);
}
static inline double o_corr(double m, double x, double x0, double y0) { return m * (x - x0) + y0; }

#define O_X_INC 1
#define O_Y_INC (-1)
#define O_Z_INC (-1)

/* The straight-line walk shared by SC_GPU (:6681), MSC_GPU (:6799) and SM_GPU (:6997):
 * visits every voxel from the entry voxel up to, not including, the exit voxel.
 * Walks stop at a negative voxel coordinate or after 4*(C+R+S)+16 steps (the reference has
 * no such bound; both cases are out-of-bounds or non-terminating there). */
typedef void (*o_visit_fn)(void* ctx, int64_t voxel);
static void o_carve_walk(const pct_params* p, const real* const h[10], int64_t i, o_visit_fn visit, void* vctx)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    const double X0 = p->x_zero_coordinate, Y0 = p->y_zero_coordinate, Z0 = p->z_zero_coordinate;
    const double VW = p->voxel_width, VH = p->voxel_height, VT = p->voxel_thickness;
    const real xe = h[0][i], ye = h[1][i], ze = h[2][i], xx = h[3][i], yx = h[4][i], zx = h[5][i];
    double x = xe, y = ye, z = ze;
    int x_move = (xe <= xx) - (xe >= xx);
    int y_move = (ye <= yx) - (ye >= yx);
    int z_move = (ze <= zx) - (ze >= zx);
    int vx = o_calc_voxel(X0, xe, VW), vy = o_calc_voxel(Y0, ye, VH), vz = o_calc_voxel(Z0, ze, VT);
    int64_t voxel = vx + (int64_t)vy * C + (int64_t)vz * C * R;
    double x_to_go = o_dist(X0, x, O_X_INC, x_move, VW, vx);
    double y_to_go = o_dist(Y0, y, O_Y_INC, y_move, VH, vy);
    double z_to_go = o_dist(Z0, z, O_Z_INC, z_move, VT, vz);
    double dy_dx = (double)((yx - ye) / (xx - xe));   /* float / float */
    double dz_dx = (double)((zx - ze) / (xx - xe));
    double dz_dy = (double)((zx - ze) / (yx - ye));
    double dx_dy = (double)((xx - xe) / (yx - ye));
    double dx_dz = (double)((xx - xe) / (zx - ze));
    double dy_dz = (double)((yx - ye) / (zx - ze));
    int vxo = o_calc_voxel(X0, xx, VW), vyo = o_calc_voxel(Y0, yx, VH), vzo = o_calc_voxel(Z0, zx, VT);
    int64_t voxel_out = vxo + (int64_t)vyo * C + (int64_t)vzo * C * R;
    int end_walk = (voxel == voxel_out) || (vx >= C) || (vy >= R) || (vz >= S) || vx < 0 || vy < 0 || vz < 0;
    if (!end_walk) visit(vctx, voxel);
    const int cap = 4 * (C + R + S) + 16;
    int steps = 0;
    while (!end_walk) {
        if (z_move != 0) {                             /* take_3D_step_GPU :11392-11459 */
            double x_ext = fabs(dz_dx) * x_to_go;
            double y_ext = fabs(dz_dy) * y_to_go;
            if ((z_to_go <= x_ext) && (z_to_go <= y_ext)) {
                vz -= z_move;
                z = o_edge(Z0, vz, VT, O_Z_INC, z_move);
                x = o_corr(dx_dz, z, ze, xe);
                y = o_corr(dy_dz, z, ze, ye);
                x_to_go = o_dist(X0, x, O_X_INC, x_move, VW, vx);
                y_to_go = o_dist(Y0, y, O_Y_INC, y_move, VH, vy);
                z_to_go = VT;
            } else if (x_ext <= y_ext) {
                vx += x_move;
                x = o_edge(X0, vx, VW, O_X_INC, x_move);
                y = o_corr(dy_dx, x, xe, ye);
                z = o_corr(dz_dx, x, xe, ze);
                x_to_go = VW;
                y_to_go = o_dist(Y0, y, O_Y_INC, y_move, VH, vy);
                z_to_go = o_dist(Z0, z, O_Z_INC, z_move, VT, vz);
            } else {
                vy -= y_move;
                y = o_edge(Y0, vy, VH, O_Y_INC, y_move);
                x = o_corr(dx_dy, y, ye, xe);
                z = o_corr(dz_dy, y, ye, ze);
                x_to_go = o_dist(X0, x, O_X_INC, x_move, VW, vx);
                y_to_go = VH;
                z_to_go = o_dist(Z0, z, O_Z_INC, z_move, VT, vz);
            }
            if (x_to_go == 0) { x_to_go = VW; vx += x_move; }
            if (y_to_go == 0) { y_to_go = VH; vy -= y_move; }
            if (z_to_go == 0) { z_to_go = VT; vz -= z_move; }
        } else {                                       /* take_2D_step_GPU :11346-11391 */
            double y_ext = fabs(dx_dy) * y_to_go;
            if (x_to_go <= y_ext) {
                vx += x_move;
                x = o_edge(X0, vx, VW, O_X_INC, x_move);
                y = o_corr(dy_dx, x, xe, ye);
                x_to_go = VW;
                y_to_go = o_dist(Y0, y, O_Y_INC, y_move, VH, vy);
            } else {
                vy -= y_move;
                y = o_edge(Y0, vy, VH, O_Y_INC, y_move);
                x = o_corr(dx_dy, y, ye, xe);
                x_to_go = o_dist(X0, x, O_X_INC, x_move, VW, vx);
                y_to_go = VH;
            }
            if (x_to_go == 0) { x_to_go = VW; vx += x_move; }
            if (y_to_go == 0) { y_to_go = VH; vy -= y_move; }
        }
        vz = imax(vz, 0);
        voxel = vx + (int64_t)vy * C + (int64_t)vz * C * R;
        end_walk = (voxel == voxel_out) || (vx >= C) || (vy >= R) || (vz >= S);
        if (vx < 0 || vy < 0 || ++steps > cap) end_walk = 1;
        if (!end_walk) visit(vctx, voxel);
    }
}

static void o_visit_sc(void* c, int64_t v) { ((uint8_t*)c)[v] = 0; }
static void o_visit_count(void* c, int64_t v) { ((int32_t*)c)[v] += 1; }

void O(o_sc)(const pct_params* p, int64_t n, const real* const h[10], const real* wepl, uint8_t* sc_hull)
{
    for (int64_t i = 0; i < n; i++)     /* :6686 (missed histories were compacted away) */
        if ((wepl[i] <= p->sc_upper_threshold) && (wepl[i] >= p->sc_lower_threshold)) o_carve_walk(p, h, i, o_visit_sc, sc_hull);
}
void O(o_msc)(const pct_params* p, int64_t n, const real* const h[10], const real* wepl, int32_t* counts)
{
    for (int64_t i = 0; i < n; i++)     /* :6806 */
        if ((wepl[i] <= p->msc_upper_threshold) && (wepl[i] >= p->msc_lower_threshold)) o_carve_walk(p, h, i, o_visit_count, counts);
}
void O(o_sm)(const pct_params* p, int64_t n, const real* const h[10], const real* wepl, int32_t* counts)
{
    for (int64_t i = 0; i < n; i++)     /* :7004 */
        if (wepl[i] >= p->sm_lower_threshold) o_carve_walk(p, h, i, o_visit_count, counts);
}

static int o_rim(const pct_params* p, int column, int row)
{
    real x = (real)((column - p->columns / 2 + 0.5) * p->voxel_width);
    real y = (real)((p->rows / 2 - row - 0.5) * p->voxel_height);
    double m = p->voxel_width > p->voxel_height ? p->voxel_width : p->voxel_height;
    return (x * x + y * y) >= f_sq(p->recon_cyl_radius - m / 2);
}
static int o_max_diff(const pct_params* p, const int32_t* counts, int c, int r, int64_t s_off)
{
    const int C = p->columns, R = p->rows;
    int md = 0;
    if ((r != 0) && (r != R - 1) && (c != 0) && (c != C - 1)) {
        int32_t own = counts[s_off + c + r * C];
        for (int rr = r - 1; rr <= r + 1; rr++)
            for (int cc = c - 1; cc <= c + 1; cc++) {
                int d = own - counts[s_off + cc + rr * C];
                if (d > md) md = d;
            }
    }
    return md;
}
void O(o_msc_edge)(const pct_params* p, const int32_t* counts, int32_t* hull)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    for (int s = 0; s < S; s++) {
        int64_t so = (int64_t)s * C * R;
        for (int r = 0; r < R; r++)
            for (int c = 0; c < C; c++) {
                int64_t v = so + c + r * C;
                int md = o_max_diff(p, counts, c, r, so);
                int32_t out = (md > p->msc_diff_thresh || counts[v] > p->msc_diff_thresh) ? 0 : 1; /* :6950 */
                if (o_rim(p, c, r)) out = 0;
                hull[v] = out;
            }
    }
}
void O(o_sm_edge)(const pct_params* p, const int32_t* counts, int32_t* hull)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    for (int s = 0; s < S; s++) {
        int64_t so = (int64_t)s * C * R;
        int max_difference = 0, threshold = 0;                    /* :7118-7134 */
        for (int pix = 0; pix < C * R; pix++) {
            int d = o_max_diff(p, counts, pix % C, pix / C, so);
            if (d > max_difference) {
                max_difference = d;
                threshold = counts[so + pix];
            }
        }
        for (int r = 0; r < R; r++)
            for (int c = 0; c < C; c++) {                         /* SM_edge_detection_GPU :7160 */
                int64_t v = so + c + r * C;
                int32_t out = (counts[v] > p->sm_scale_threshold * threshold) ? 1 : 0;
                if (o_rim(p, c, r)) out = 0;
                hull[v] = out;
            }
    }
}

/* ------------------------------------------------------------------ (c) FBP */
void O(o_filter)(const pct_params* p, const real* sino, real* filtered)
{
    const int T = p->t_bins, V = p->v_bins, A = p->angular_bins;
    const double PI = p->pi, SR = p->source_radius, TB = p->t_bin_size;
#pragma omp parallel for collapse(2) schedule(static)
    for (int vb = 0; vb < V; vb++)
        for (int ab = 0; ab < A; ab++) {
            double v = (vb - V / 2 + 0.5) * p->v_bin_size;                               /* :6211 */
            int64_t strip = ((int64_t)vb * A + ab) * T;
            for (int tb = 0; tb < T; tb++) {
                real acc = 0.0f;
                for (int tr = 0; tr < T; tr++) {
                    double t = (tr - T / 2 + 0.5) * TB;
                    int sep = tb - tr;
                    double scale = SR / sqrt(SR * SR + t * t + v * v);
                    double filt;
                    if (p->fbp_filter_type == PCT_FILTER_RAM_LAK) {                      /* :6224-6231 */
                        double tau = p->ram_lak_tau;
                        if (sep == 0) filt = 1.0 / (4.0 * (tau * tau));
                        else if (sep % 2 == 0) filt = 0;
                        else { double q = tau * PI * sep; filt = -1.0 / (q * q); }
                    } else if (p->fbp_filter_type == PCT_FILTER_SHEPP_LOGAN) {           /* :6234 */
                        filt = 1.0 / ((TB * PI * TB * PI) * (1.0 - (4 * sep * sep)));
                    } else {
                        filt = (sep == 0) ? 1.0 / TB : 0.0;  /* UNFILTERED leaves `filtered` unset; identity here */
                    }
                    acc += (real)(TB * (double)sino[strip + tr] * filt * scale);      /* :6237 */
                }
                filtered[strip + tb] = acc;
            }
        }
}

/* voxel_2_3D_voxels_GPU :11124-11134 (quirk: '>' decodes exact multiples one row/slice early) */
static void o_voxel_decode(const pct_params* p, int64_t voxel, int* vx, int* vy, int* vz)
{
    const int64_t C = p->columns, R = p->rows;
    if (p->quirk_voxel_decode) {
        int64_t z = 0, y = 0;
        while (voxel - C * R > 0) { voxel -= C * R; z++; }
        while (voxel - C > 0) { voxel -= C; y++; }
        *vx = (int)voxel; *vy = (int)y; *vz = (int)z;
    } else {
        *vz = (int)(voxel / (C * R));
        *vy = (int)((voxel / C) % R);
        *vx = (int)(voxel % C);
    }
}
static inline double o_voxel_pos(int vi, double size, int num, int prog)
{
    double zero_voxel = (num - 1) / 2.0;                                                 /* :11146 */
    return prog * (vi - zero_voxel) * size;
}

/* backprojection_GPU2, DS.cu:6340-6479 */
void O(o_backproject)(const pct_params* p, const real* sf, real* image)
{
    const int T = p->t_bins, V = p->v_bins, A = p->angular_bins;
    const double PI = p->pi, SR = p->source_radius, R = p->recon_cyl_radius;
    const double delta = p->angular_bin_size * (PI / 180.0);
#pragma omp parallel for schedule(static)
    for (int64_t voxel = 0; voxel < p->num_voxels; voxel++) {
        int vx, vy, vz;
        o_voxel_decode(p, voxel, &vx, &vy, &vz);
        double x = o_voxel_pos(vx, p->voxel_width, p->columns, 1);
        double y = o_voxel_pos(vy, p->voxel_height, p->rows, -1);
        double z = o_voxel_pos(vz, p->voxel_thickness, p->slices, -1);
        if ((x * x + y * y) > (R * R)) {
            image[voxel] = (real)p->rsp_air;
            continue;
        }
        real acc = 0.0f;
        for (int ab = 0; ab < A; ab++) {
            double ba = ab * delta;
            double u = x * cos(ba) + y * sin(ba);
            double t = y * cos(ba) - x * sin(ba);
            double v = z;
            double dnt = (t - u * (t / (SR + u))) / p->t_bin_size + T / 2.0;
            int t_bin = (int)floor(dnt);
            double eta = dnt - t_bin;
            double dnv = (v - u * (v / (SR + u))) / p->v_bin_size + V / 2.0;
            int v_bin = (int)floor(dnv);
            double eps = dnv - v_bin;
            double scale = (double)f_sq(SR / (SR + u));
            int64_t bin11 = ((int64_t)v_bin * A + ab) * T + t_bin;
            int64_t bin12 = bin11 + 1, bin21 = bin11 + (int64_t)A * T, bin22 = bin21 + 1;
            real xterm = 0.0f;
            if (t_bin > -1 && t_bin < T - 1) {
                if (v_bin > -1 && v_bin < V - 1)
                    xterm = (real)((double)sf[bin11] * (1 - eta) * (1 - eps) + (double)sf[bin12] * eta * (1 - eps) +
                                    (double)sf[bin21] * (1 - eta) * eps + (double)sf[bin22] * eta * eps);
                else if (v_bin == -1)
                    xterm = (real)((double)sf[bin21] * (1 - eta) * eps + (double)sf[bin22] * eta * eps);
                else if (v_bin == V - 1)
                    xterm = (real)((double)sf[bin11] * (1 - eta) * (1 - eps) + (double)sf[bin12] * eta * (1 - eps));
            } else if (t_bin == -1) {
                if (v_bin > -1 && v_bin < V - 1)
                    xterm = (real)((double)sf[bin12] * eta * (1 - eps) + (double)sf[bin22] * eta * eps);
                else if (v_bin == -1)
                    xterm = (real)((double)sf[bin22] * eta * eps);
                else if (v_bin == V - 1)
                    xterm = (real)((double)sf[bin12] * eta * (1 - eps));
            } else if (t_bin == T - 1) {
                if (v_bin > -1 && v_bin < V - 1)
                    xterm = (real)((double)sf[bin11] * (1 - eta) * (1 - eps) + (double)sf[bin21] * (1 - eta) * eps);
                else if (v_bin == -1)
                    xterm = (real)((double)sf[bin21] * (1 - eta) * eps);
                else if (v_bin == V - 1)
                    xterm = (real)((double)sf[bin11] * (1 - eta) * (1 - eps));
            }
            acc += (real)((double)xterm * delta * scale);                               /* :6474 */
        }
        image[voxel] = acc;
    }
}

/* median_filter_2D (DS.cu:7482-7519) as FBP() applies it (:6150-6166): interior voxels get the
 * element at position (2r+1)^2/2 of their sorted in-slice window, the rest 0. */
static int o_cmp_float(const void* a, const void* b)
{
    real x = *(const real*)a, y = *(const real*)b;
    return (x > y) - (x < y);
}
void O(o_median_filter_2d)(const pct_params* p, const real* in, int radius, real* out)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    const int w = 2 * radius + 1, n = w * w;
    real* nb = (real*)malloc(sizeof(real) * n);
    memset(out, 0, sizeof(real) * (size_t)C * R * S);
    for (int s = 0; s < S; s++)
        for (int c = radius; c < C - radius; c++)
            for (int r = radius; r < R - radius; r++) {
                int k = 0;
                for (int cc = c - radius; cc <= c + radius; cc++)
                    for (int rr = r - radius; rr <= r + radius; rr++) nb[k++] = in[cc + rr * C + (int64_t)s * C * R];
                qsort(nb, n, sizeof(real), o_cmp_float);
                out[c + r * C + (int64_t)s * C * R] = nb[n / 2];
            }
    free(nb);
}

/* averaging_filter_GPU with perform_threshold (DS.cu:7420-7443), on a 0/1 hull */
void O(o_avg_filter_hull)(const pct_params* p, const uint8_t* in, int radius, double threshold, uint8_t* out)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    for (int s = 0; s < S; s++)
        for (int y = 0; y < R; y++)
            for (int x = 0; x < C; x++) {
                int left = imax(x - radius, 0), right = x + radius < C - 1 ? x + radius : C - 1;
                int top = imax(y - radius, 0), bottom = y + radius < R - 1 ? y + radius : R - 1;
                int nvox = (right - left + 1) * (bottom - top + 1);
                double sum_threshold = nvox * threshold, sum = 0.0;
                for (int c = left; c <= right; c++)
                    for (int r = top; r <= bottom; r++) sum += in[c + r * C + (int64_t)s * C * R] ? 1.0 : 0.0;
                out[x + y * C + (int64_t)s * C * R] = sum > sum_threshold;
            }
}

/* FBP_image_2_hull_GPU, DS.cu:6592-6603 */
void O(o_fbp_hull)(const pct_params* p, const real* image, uint8_t* hull)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    const double RC = p->recon_cyl_radius;
    for (int s = 0; s < S; s++)
        for (int r = 0; r < R; r++)
            for (int c = 0; c < C; c++) {
                int64_t v = (int64_t)s * C * R + r * C + c;
                double x = -RC + (c + 0.5) * p->voxel_width;
                double y = RC - (r + 0.5) * p->voxel_height;
                double d2 = (double)(f_sq(x) + f_sq(y));
                hull[v] = (uint8_t)(((double)image[v] > p->fbp_threshold) && (d2 < (double)f_sq(RC)));
            }
}

/* ------------------------------------------------------------------ (d) MLP tables */
/* A_0..A_5 (DS.h:749-754): c * pow((float)10, (float)e) with the float overload of pow. */
static void o_poly_coeffs(double* c, float* P)
{
    c[0] = 7.457; c[1] = 4.548; c[2] = -5.777; c[3] = 1.301; c[4] = -9.228; c[5] = 2.687;
    P[0] = powf(10.0f, -6.0f); P[1] = powf(10.0f, -7.0f); P[2] = powf(10.0f, -8.0f);
    P[3] = powf(10.0f, -8.0f); P[4] = powf(10.0f, -10.0f); P[5] = powf(10.0f, -11.0f);
}
/* A_k_OVER_m expands to `c_k * P_k / m`, i.e. (c_k*P_k)/m; the innermost `u * A_5_OVER_m` expands to
 * ((u*c_5)*P_5)/m (DS.cu:8476-8480 with the DS.h:756-781 macros). */
static double o_poly(const double* c, const float* P, double u, const int* d)
{
    double a0 = c[0] * P[0] / d[0], a1 = c[1] * P[1] / d[1], a2 = c[2] * P[2] / d[2];
    double a3 = c[3] * P[3] / d[3], a4 = c[4] * P[4] / d[4];
    double in5 = u * c[5] * P[5] / d[5];
    return a0 + u * (a1 + u * (a2 + u * (a3 + u * (a4 + in5))));
}
int O(o_tables_create)(const pct_params* p, o_tables* t)
{
    memset(t, 0, sizeof(*t));
    t->n_trig = (int64_t)p->trig_table_elements + 1;
    t->n_coeff = (int64_t)p->coeff_table_elements + 1;
    t->n_poly = (int64_t)p->poly_table_elements + 1;
    t->sin_t = (double*)malloc(sizeof(double) * t->n_trig);
    t->cos_t = (double*)malloc(sizeof(double) * t->n_trig);
    t->scat = (double*)malloc(sizeof(double) * t->n_coeff);
    double** tabs[5] = {&t->p12, &t->p23, &t->p34, &t->p26, &t->p312};
    for (int k = 0; k < 5; k++) *tabs[k] = (double*)malloc(sizeof(double) * t->n_poly);
    if (!t->sin_t || !t->cos_t || !t->scat || !t->p12 || !t->p23 || !t->p34 || !t->p26 || !t->p312) return -1;
    for (int64_t s = 0; s < t->n_trig; s++) {                                            /* :8421-8434 */
        double arg = p->trig_table_min + s * p->trig_table_step;
        t->sin_t[s] = sin(arg);
        t->cos_t[s] = cos(arg);
    }
    for (int64_t s = 0; s < t->n_coeff; s++) {                                           /* :8446-8455 */
        double depth = s * p->coeff_table_step;
        t->scat[s] = pow(p->e_0 * (1 + 0.038 * log(depth / p->x_0_rad_length)), 2.0) / p->x_0_rad_length;
    }
    double c[6];
    float P[6];
    o_poly_coeffs(c, P);
    static const int d12[6] = {1, 2, 3, 4, 5, 6}, d23[6] = {2, 3, 4, 5, 6, 7}, d34[6] = {3, 4, 5, 6, 7, 8};
    static const int d26[6] = {2, 6, 12, 20, 30, 42}, d312[6] = {3, 12, 30, 60, 105, 168};
    for (int64_t s = 0; s < t->n_poly; s++) {                                            /* :8466-8481 */
        double u = s * p->poly_table_step;
        /* poly_1_2 has A_0 (not A_0/1) as its constant term: c0*P0 is the same value either way. */
        t->p12[s] = u * o_poly(c, P, u, d12);
        t->p23[s] = pow(u, 2) * o_poly(c, P, u, d23);
        t->p34[s] = pow(u, 3) * o_poly(c, P, u, d34);
        t->p26[s] = pow(u, 2) * o_poly(c, P, u, d26);
        t->p312[s] = pow(u, 3) * o_poly(c, P, u, d312);
    }
    return 0;
}
void O(o_tables_destroy)(o_tables* t)
{
    free(t->sin_t); free(t->cos_t); free(t->scat);
    free(t->p12); free(t->p23); free(t->p34); free(t->p26); free(t->p312);
    memset(t, 0, sizeof(*t));
}

/* ------------------------------------------------------------------ (d) endpoints */
/* find_MLP_endpoints_GPU, DS.cu:7581-7773.  Walk from (x,y,z) along the tracker angles until a
 * hull voxel or the image boundary; bounded by 4*(C+R+S)+16 steps (reference: unbounded). */
static int o_find_endpoint(const pct_params* p, const uint8_t* hull, double x_start, double y_start, double z_start,
                           double xy_angle, double xz_angle, double* xo, double* yo, double* zo, int* vxo, int* vyo,
                           int* vzo, int entering)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    const double X0 = p->x_zero_coordinate, Y0 = p->y_zero_coordinate, Z0 = p->z_zero_coordinate;
    const double VW = p->voxel_width, VH = p->voxel_height, VT = p->voxel_thickness;
    double x = x_start, y = y_start, z = z_start;
    if (!entering) xy_angle += p->pi;
    int x_move = (cos(xy_angle) >= 0) - (cos(xy_angle) <= 0);
    int y_move = (sin(xy_angle) >= 0) - (sin(xy_angle) <= 0);
    int z_move = (sin(xz_angle) >= 0) - (sin(xz_angle) <= 0);
    if (x_move < 0) z_move *= -1;
    int vx = o_calc_voxel(X0, x, VW), vy = o_calc_voxel(Y0, y, VH), vz = o_calc_voxel(Z0, z, VT);
    double x_to_go = o_dist(X0, x, O_X_INC, x_move, VW, vx);
    double y_to_go = o_dist(Y0, y, O_Y_INC, y_move, VH, vy);
    double z_to_go = o_dist(Z0, z, O_Z_INC, z_move, VT, vz);
    int64_t voxel = vx + ((int64_t)vy + (int64_t)vz * R) * C;
    double delta_yx = fabs(tan(xy_angle));                                              /* :7649-7658 */
    double delta_zx = fabs(tan(xz_angle));
    double delta_zy = fabs(tan(xz_angle) / tan(xy_angle));
    double dy_dx = tan(xy_angle);
    double dz_dx = tan(xz_angle);
    double dz_dy = tan(xz_angle) / tan(xy_angle);
    double dx_dy = (double)f_recip(tan(xy_angle));       /* powf(tan(xy), -1.0) */
    double dx_dz = (double)f_recip(tan(xz_angle));
    double dy_dz = (double)f_tan(xy_angle) / tan(xz_angle); /* tanf(xy)/tan(xz) */
    int hit = 0;
    int outside = (vx >= C) || (vy >= R) || (vz >= S) || (vx < 0) || (vy < 0) || (vz < 0);
    if (!outside) hit = (hull[voxel] == 1);
    int end_walk = outside || hit;
    const int cap = 4 * (C + R + S) + 16;
    int steps = 0;
    while (!end_walk) {
        if (z_move != 0) {
            double x_ext = delta_zx * x_to_go;
            double y_ext = delta_zy * y_to_go;
            if ((z_to_go <= x_ext) && (z_to_go <= y_ext)) {
                vz -= z_move;
                z = o_edge(Z0, vz, VT, O_Z_INC, z_move);
                x = o_corr(dx_dz, z, z_start, x_start);
                y = o_corr(dy_dz, z, z_start, y_start);
                x_to_go = o_dist(X0, x, O_X_INC, x_move, VW, vx);
                y_to_go = o_dist(Y0, y, O_Y_INC, y_move, VH, vy);
                z_to_go = VT;
            } else if (x_ext <= y_ext) {
                vx += x_move;
                x = o_edge(X0, vx, VW, O_X_INC, x_move);
                y = o_corr(dy_dx, x, x_start, y_start);
                z = o_corr(dz_dx, x, x_start, z_start);
                x_to_go = VW;
                y_to_go = o_dist(Y0, y, O_Y_INC, y_move, VH, vy);
                z_to_go = o_dist(Z0, z, O_Z_INC, z_move, VT, vz);
            } else {
                vy -= y_move;
                y = o_edge(Y0, vy, VH, O_Y_INC, y_move);
                x = o_corr(dx_dy, y, y_start, x_start);
                z = o_corr(dz_dy, y, y_start, z_start);
                x_to_go = o_dist(X0, x, O_X_INC, x_move, VW, vx);
                y_to_go = VH;
                z_to_go = o_dist(Z0, z, O_Z_INC, z_move, VT, vz);
            }
            if (x_to_go == 0) { x_to_go = VW; vx += x_move; }
            if (y_to_go == 0) { y_to_go = VH; vy -= y_move; }
            if (z_to_go == 0) { z_to_go = VT; vz -= z_move; }
            vz = imax(vz, 0);
        } else {
            double y_ext = y_to_go / delta_yx;
            if (x_to_go <= y_ext) {
                vx += x_move;
                x = o_edge(X0, vx, VW, O_X_INC, x_move);
                y = o_corr(dy_dx, x, x_start, y_start);
                x_to_go = VW;
                y_to_go = o_dist(Y0, y, O_Y_INC, y_move, VH, vy);
            } else {
                vy -= y_move;
                y = o_edge(Y0, vy, VH, O_Z_INC, y_move);
                x = o_corr(dx_dy, y, y_start, x_start);
                x_to_go = o_dist(X0, x, O_X_INC, x_move, VW, vx);
                y_to_go = VH;
            }
            if (x_to_go == 0) { x_to_go = VW; vx += x_move; }
            if (y_to_go == 0) { y_to_go = VH; vy -= y_move; }
        }
        voxel = vx + (int64_t)vy * C + (int64_t)vz * C * R;
        outside = (vx >= C) || (vy >= R) || (vz >= S) || (vx < 0) || (vy < 0) || (vz < 0);
        if (!outside) hit = (hull[voxel] == 1);
        end_walk = outside || hit;
        if (++steps > cap) { hit = 0; end_walk = 1; }
    }
    if (hit) { *xo = x; *yo = y; *zo = z; }
    *vxo = vx; *vyo = vy; *vzo = vz;
    return hit;
}

/* collect_MLP_endpoints_GPU_nobool, DS.cu:7819-7854 */
void O(o_endpoints)(const pct_params* p, const uint8_t* hull, int64_t n, real* const h[10], uint32_t* first_voxel)
{
    const int C = p->columns, R = p->rows;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        double xi = 0, yi = 0, zi = 0, xo = 0, yo = 0, zo = 0;
        int vx, vy, vz, vxi, vyi, vzi;
        int entered = o_find_endpoint(p, hull, h[0][i], h[1][i], h[2][i], h[6][i], h[7][i], &xi, &yi, &zi, &vx, &vy, &vz, 1);
        int exited = o_find_endpoint(p, hull, h[3][i], h[4][i], h[5][i], h[8][i], h[9][i], &xo, &yo, &zo, &vxi, &vyi, &vzi, 0);
        if (entered && exited) {
            first_voxel[i] = (uint32_t)(vx + C * vy + R * C * vz);
            h[0][i] = (real)xi; h[1][i] = (real)yi; h[2][i] = (real)zi;
            h[3][i] = (real)xo; h[4][i] = (real)yo; h[5][i] = (real)zo;
        } else {
            first_voxel[i] = (uint32_t)p->num_voxels;
        }
    }
}

/* ------------------------------------------------------------------ (d) MLP */
/* EffectiveChordLength_GPU, DS.cu:8873-8894 */
double O(o_chord_length)(const pct_params* p, double abs_angle_t, double abs_angle_v)
{
    const double PI = p->pi, VW = p->voxel_width;
    double eff_t = abs_angle_t - ((int)(abs_angle_t / (PI / 2))) * (PI / 2);
    double eff_v = fabs(abs_angle_v);
    double sf = p->mlp_u_step / VW;
    real s2t = f_sin(2 * eff_t), ct = f_cos(eff_t), st = f_sin(eff_t);
    real cs = ct + st;                                        /* cosf + sinf: float add */
    double c2d = (1 / 3.0) * ((sf * sf * s2t - 6) / (sf * s2t - (double)(2 * cs)) + sf * sf * s2t / (double)(2 * cs));
    double mpw = VW / (double)cs;
    double hf = p->voxel_thickness / mpw;
    sf = p->mlp_u_step / mpw;
    real s2v = f_sin(2 * eff_v), cv = f_cos(eff_v), sv = f_sin(eff_v);
    double den = hf * cv + sv;
    double c3d = (1 / 3.0) * ((sf * sf * s2v - 6 * hf) / (sf * s2v - 2 * den) + sf * sf * s2v / (2 * den));
    return VW * c2d * c3d;
}

/* find_MLP_path_GPU_tabulated, DS.cu:9104-9262.  Writes up to cap voxels into path, returns the count
 * (or -1 past cap) and the chord.  Voxels whose linear index is >= NUM_VOXELS are not listed (the
 * reference reads and writes out of bounds there). */
static int64_t o_mlp_path(const pct_params* p, const o_tables* tb, const real* const h[10], int64_t i, uint32_t first,
                          uint32_t* path, int64_t cap, double* chord, double* rec, int64_t rec_cap, int64_t* n_rec)
{
    const double PI = p->pi, STEP = p->mlp_u_step;
    const int C = p->columns, R = p->rows;
    const int64_t NV = p->num_voxels;
    double x_in = h[0][i], y_in = h[1][i], z_in = h[2][i], x_out = h[3][i], y_out = h[4][i], z_out = h[5][i];
    double xy_entry = h[6][i], xz_entry = h[7][i], xy_exit = h[8][i], xz_exit = h[9][i];
    unsigned int ti = (unsigned int)((xy_entry - p->trig_table_min) / p->trig_table_step + 0.5);
    double sin_term = tb->sin_t[ti], cos_term = tb->cos_t[ti];
    double u_in = cos_term * x_in + sin_term * y_in;
    double u_out = cos_term * x_out + sin_term * y_out;
    if (u_in > u_out) {                                                                  /* :9128 */
        xy_entry += PI;
        xy_exit += PI;
        ti = (unsigned int)((xy_entry - p->trig_table_min) / p->trig_table_step + 0.5);
        sin_term = tb->sin_t[ti];
        cos_term = tb->cos_t[ti];
        u_in = cos_term * x_in + sin_term * y_in;
        u_out = cos_term * x_out + sin_term * y_out;
    }
    double t_in = cos_term * y_in - sin_term * x_in;
    double t_out = cos_term * y_out - sin_term * x_out;
    double T2_1 = xy_exit - xy_entry, V2_1 = xz_exit - xz_entry;
    double u_2 = fabs(u_out - u_in);
    double depth = u_2 - STEP;
    double u_shifted = u_in;
    unsigned int s_step = (unsigned int)(STEP / p->coeff_table_step + 0.5);
    unsigned int s1 = s_step;
    unsigned int s2 = (unsigned int)(depth / p->coeff_table_step + 0.5);
    unsigned int p_step = (unsigned int)(STEP / p->poly_table_step + 0.5);
    unsigned int u1i = p_step;
    unsigned int u2i = (unsigned int)(u_2 / p->poly_table_step + 0.5);
    double u2_312 = tb->p312[u2i], u2_26 = tb->p26[u2i], u2_12 = tb->p12[u2i];
    double u2_sq = (double)f_sq(u_2);
    int64_t n = 0;
    uint32_t last = first;
    path[n++] = first;
    *chord = O(o_chord_length)(p, (xy_entry + xy_exit) / 2.0, (xz_entry + xz_exit) / 2.0);
    while (depth > STEP) {
        double u1_12 = tb->p12[u1i], u1_23 = tb->p23[u1i];
        double st1 = tb->p312[u1i], st1th1 = tb->p26[u1i], sth1 = u1_12;
        double st2 = u2_312 - u2_sq * u1_12 + 2 * u_2 * u1_23 - tb->p34[u1i];
        double st2th2 = u2_26 - u_2 * u1_12 + u1_23;
        double sth2 = u2_12 - u1_12;
        double det1 = tb->scat[s1] * (st1 * sth1 - (double)f_sq(st1th1));
        double S1I0 = sth1 / det1, S1I1 = st1th1 / det1, S1I2 = st1 / det1;
        double det2 = tb->scat[s2] * (st2 * sth2 - (double)f_sq(st2th2));
        double S2I0 = sth2 / det2, S2I1 = st2th2 / det2, S2I2 = st2 / det2;
        double c24 = S2I0 * depth - S2I1;                                                /* :9201 */
        double ft0 = S1I0 + S2I0;
        double ft1 = c24 - S1I1;
        double ft2 = depth * S2I0 - S1I1 - S2I1;
        double ft3 = S1I2 + S2I2 + depth * (c24 - S2I1);
        double detft = ft0 * ft3 - ft1 * ft2;
        double c3 = S2I0 * t_out - S2I1 * T2_1;
        double sc0 = S1I0 * t_in + c3;
        double sc1 = depth * c3 + S2I2 * T2_1 - S2I1 * t_out - S1I1 * t_in;
        double t_1 = (ft3 * sc0 - ft1 * sc1) / detft;
        c3 = S2I0 * z_out - S2I1 * V2_1;
        sc0 = S1I0 * z_in + c3;
        sc1 = depth * c3 + S2I2 * V2_1 - S2I1 * z_out - S1I1 * z_in;
        double v_1 = (ft3 * sc0 - ft1 * sc1) / detft;
        u_shifted += STEP;
        double x_1 = cos_term * u_shifted - sin_term * t_1;
        double y_1 = sin_term * u_shifted + cos_term * t_1;
        int vx = o_calc_voxel(p->x_zero_coordinate, x_1, p->voxel_width);
        int vy = o_calc_voxel(p->y_zero_coordinate, y_1, p->voxel_height);
        int vz = o_calc_voxel(p->z_zero_coordinate, v_1, p->voxel_thickness);
        int64_t voxel = vx + (int64_t)vy * C + (int64_t)vz * C * R;
        if (rec && *n_rec < rec_cap) {   /* per-step record for o_mlp_steps */
            double* r = rec + 6 * *n_rec;
            r[0] = u_shifted; r[1] = t_1; r[2] = v_1; r[3] = x_1; r[4] = y_1; r[5] = (double)voxel;
        }
        if (rec) *n_rec += 1;
        if (voxel != (int64_t)last && voxel < NV) {
            if (n >= cap) return -1;
            path[n++] = (uint32_t)voxel;
            last = (uint32_t)voxel;
        }
        depth -= STEP;
        s1 += s_step;
        s2 -= s_step;
        u1i += p_step;
    }
    return n;
}

/* powf(x, 3) of a double argument: x rounded to float, cubed (the float-argument double cube, rounded). */
#ifdef O_FP64
static inline real f_cube(double x) { return x * x * x; }
#else
static inline float f_cube(double x) { double f = (double)(float)x; return (float)(f * f * f); }
#endif

/* find_MLP_path_GPU (MLP_ALGORITHM = STANDARD), DS.cu:8895-9103: the same MLP as the tabulated path with the
 * scattering moments evaluated from the polynomial at every step, the scattering coefficients from log(),
 * the rotation from cos()/sin() of the entry angle, and the explicit 2x2 matrix products
 * (R_0 = [1 u_1; 0 1], R_1 = [1 u_2-u_1; 0 1], R_1^T) kept as written.  Differences from the tabulated
 * path that the reference has: the chord is taken from the angles before the +pi turn (:8921 precedes
 * :8945), u_1 counts up by MLP_U_STEP and the loop runs while u_1 < u_2 - MLP_U_STEP (:8978), and the
 * powers of u_1 (powf(u_1, 2), powf(u_1, 3)), the two scattering coefficients and the off-diagonal
 * squares in the determinants are float (:8985-9005).  Voxels >= NUM_VOXELS are not listed, as in
 * o_mlp_path. */
static int64_t o_mlp_path_standard(const pct_params* p, const real* const h[10], int64_t i, uint32_t first,
                                   uint32_t* path, int64_t cap, double* chord, double* rec, int64_t rec_cap,
                                   int64_t* n_rec)
{
    const double PI = p->pi, STEP = p->mlp_u_step, E0 = p->e_0, X0 = p->x_0_rad_length;
    const int C = p->columns, R = p->rows;
    const int64_t NV = p->num_voxels;
    double c[6];
    float Pw[6];
    o_poly_coeffs(c, Pw);
    static const int d12[6] = {1, 2, 3, 4, 5, 6}, d23[6] = {2, 3, 4, 5, 6, 7}, d34[6] = {3, 4, 5, 6, 7, 8};
    static const int d26[6] = {2, 6, 12, 20, 30, 42}, d312[6] = {3, 12, 30, 60, 105, 168};
    double x_in = h[0][i], y_in = h[1][i], z_in = h[2][i], x_out = h[3][i], y_out = h[4][i], z_out = h[5][i];
    double xy_entry = h[6][i], xz_entry = h[7][i], xy_exit = h[8][i], xz_exit = h[9][i];
    *chord = O(o_chord_length)(p, (xy_entry + xy_exit) / 2.0, (xz_entry + xz_exit) / 2.0);         /* :8921 */
    double u_in = cos(xy_entry) * x_in + sin(xy_entry) * y_in;                                 /* :8938 */
    double u_out = cos(xy_entry) * x_out + sin(xy_entry) * y_out;
    double t_in = cos(xy_entry) * y_in - sin(xy_entry) * x_in;
    double t_out = cos(xy_entry) * y_out - sin(xy_entry) * x_out;
    if (u_in > u_out) {                                                                         /* :8945 */
        xy_entry += PI;
        xy_exit += PI;
        u_in = cos(xy_entry) * x_in + sin(xy_entry) * y_in;
        u_out = cos(xy_entry) * x_out + sin(xy_entry) * y_out;
        t_in = cos(xy_entry) * y_in - sin(xy_entry) * x_in;
        t_out = cos(xy_entry) * y_out - sin(xy_entry) * x_out;
    }
    const double T0[2] = {t_in, 0.0}, T2[2] = {t_out, xy_exit - xy_entry};
    const double V0[2] = {z_in, 0.0}, V2[2] = {z_out, xz_exit - xz_entry};
    const double u_0 = 0.0;
    double u_1 = STEP;
    const double u_2 = fabs(u_out - u_in);
    const double pre1 = u_2 * u_2 * u_2 * o_poly(c, Pw, u_2, d312);                            /* :8974 */
    const double pre2 = u_2 * u_2 * o_poly(c, Pw, u_2, d26);
    const double pre3 = u_2 * o_poly(c, Pw, u_2, d12);
    double R0[4] = {1.0, 0.0, 0.0, 1.0}, R1[4] = {1.0, 0.0, 0.0, 1.0}, R1T[4] = {1.0, 0.0, 0.0, 1.0};
    int64_t n = 0;
    uint32_t last = first;
    path[n++] = first;
    while (u_1 < u_2 - STEP) {
        R0[1] = u_1;
        R1[1] = u_2 - u_1;
        R1T[2] = u_2 - u_1;
        double k1 = (double)f_sq(E0 * (1 + 0.038 * log((u_1 - u_0) / X0))) / X0;              /* :8985 */
        double st1 = k1 * ((double)f_cube(u_1) * o_poly(c, Pw, u_1, d312));
        double st1th1 = k1 * ((double)f_sq(u_1) * o_poly(c, Pw, u_1, d26));
        double sth1 = k1 * (u_1 * o_poly(c, Pw, u_1, d12));
        double det1 = st1 * sth1 - (double)f_sq(st1th1);
        double S1[4] = {sth1 / det1, -st1th1 / det1, -st1th1 / det1, st1 / det1};
        double k2 = (double)f_sq(E0 * (1 + 0.038 * log((u_2 - u_1) / X0))) / X0;              /* :8997 */
        double cm1 = u_1 * o_poly(c, Pw, u_1, d12);
        double cm2 = (double)f_sq(u_1) * o_poly(c, Pw, u_1, d23);
        double cm3 = (double)f_cube(u_1) * o_poly(c, Pw, u_1, d34);
        double st2 = k2 * (pre1 - (double)f_sq(u_2) * cm1 + 2 * u_2 * cm2 - cm3);
        double st2th2 = k2 * (pre2 - u_2 * cm1 + cm2);
        double sth2 = k2 * (pre3 - cm1);
        double det2 = st2 * sth2 - (double)f_sq(st2th2);
        double S2[4] = {sth2 / det2, -st2th2 / det2, -st2th2 / det2, st2 / det2};
        double c13_1 = S2[0] * R1[0] + S2[1] * R1[2];                                           /* :9012 */
        double c13_2 = S2[2] * R1[0] + S2[3] * R1[2];
        double c24_1 = S2[0] * R1[1] + S2[1] * R1[3];
        double c24_2 = S2[2] * R1[1] + S2[3] * R1[3];
        double ft[4];
        ft[0] = S1[0] + R1T[0] * c13_1 + R1T[1] * c13_2;
        ft[1] = S1[1] + R1T[0] * c24_1 + R1T[1] * c24_2;
        ft[2] = S1[2] + R1T[2] * c13_1 + R1T[3] * c13_2;
        ft[3] = S1[3] + R1T[2] * c24_1 + R1T[3] * c24_2;
        double detft = ft[0] * ft[3] - ft[1] * ft[2];
        double tmp = ft[0];
        ft[0] = ft[3] / detft;
        ft[1] = -ft[1] / detft;
        ft[2] = -ft[2] / detft;
        ft[3] = tmp / detft;
        double sc1 = R0[0] * T0[0] + R0[1] * T0[1];                                             /* :9031 */
        double sc2 = R0[2] * T0[0] + R0[3] * T0[1];
        double sc3 = S2[0] * T2[0] + S2[1] * T2[1];
        double sc4 = S2[2] * T2[0] + S2[3] * T2[1];
        double s0 = S1[0] * sc1 + S1[1] * sc2 + R1T[0] * sc3 + R1T[1] * sc4;
        double s1 = S1[2] * sc1 + S1[3] * sc2 + R1T[2] * sc3 + R1T[3] * sc4;
        double t_1 = ft[0] * s0 + ft[1] * s1;
        sc1 = R0[0] * V0[0] + R0[1] * V0[1];                                                    /* :9050 */
        sc2 = R0[2] * V0[0] + R0[3] * V0[1];
        sc3 = S2[0] * V2[0] + S2[1] * V2[1];
        sc4 = S2[2] * V2[0] + S2[3] * V2[1];
        s0 = S1[0] * sc1 + S1[1] * sc2 + R1T[0] * sc3 + R1T[1] * sc4;
        s1 = S1[2] * sc1 + S1[3] * sc2 + R1T[2] * sc3 + R1T[3] * sc4;
        double v_1 = ft[0] * s0 + ft[1] * s1;
        double x_1 = (cos(xy_entry) * (u_in + u_1)) - (sin(xy_entry) * t_1);                   /* :9067 */
        double y_1 = (sin(xy_entry) * (u_in + u_1)) + (cos(xy_entry) * t_1);
        int vx = o_calc_voxel(p->x_zero_coordinate, x_1, p->voxel_width);
        int vy = o_calc_voxel(p->y_zero_coordinate, y_1, p->voxel_height);
        int vz = o_calc_voxel(p->z_zero_coordinate, v_1, p->voxel_thickness);
        int64_t voxel = vx + (int64_t)vy * C + (int64_t)vz * C * R;
        if (rec && *n_rec < rec_cap) {
            double* r = rec + 6 * *n_rec;
            r[0] = u_in + u_1; r[1] = t_1; r[2] = v_1; r[3] = x_1; r[4] = y_1; r[5] = (double)voxel;
        }
        if (rec) *n_rec += 1;
        if (voxel != (int64_t)last && voxel < NV) {
            if (n >= cap) return -1;
            path[n++] = (uint32_t)voxel;
            last = (uint32_t)voxel;
        }
        u_1 += STEP;
    }
    return n;
}

static int64_t o_mlp_any(const pct_params* p, const o_tables* tb, const real* const h[10], int64_t i, uint32_t first,
                         uint32_t* path, int64_t cap, double* chord, double* rec, int64_t rec_cap, int64_t* n_rec)
{
    if (p->mlp_algorithm == PCT_MLP_STANDARD)
        return o_mlp_path_standard(p, h, i, first, path, cap, chord, rec, rec_cap, n_rec);
    return o_mlp_path(p, tb, h, i, first, path, cap, chord, rec, rec_cap, n_rec);
}

int64_t O(o_mlp_lists)(const pct_params* p, const o_tables* t, int64_t n, const real* const h[10],
                    const uint32_t* first_voxel, int64_t* offsets, uint32_t* voxels, int64_t capacity, double* chord)
{
    /* Histories are split into contiguous chunks; each chunk's lists go to a private buffer that is
     * concatenated in history order afterwards, so the result does not depend on the thread count. */
    int nchunks = 1;
#ifdef _OPENMP
    nchunks = omp_get_max_threads() * 4;
#endif
    if (n < 4096) nchunks = 1;
    const int64_t per_path = (int64_t)(2 * p->recon_cyl_radius / p->mlp_u_step) + 8;
    int64_t* csize = (int64_t*)calloc((size_t)nchunks, sizeof(int64_t));
    uint32_t** cbuf = (uint32_t**)calloc((size_t)nchunks, sizeof(uint32_t*));
    int failed = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < nchunks; c++) {
        int64_t a = n * c / nchunks, b = n * (c + 1) / nchunks;
        int64_t cap = (b - a) * per_path + 1;
        uint32_t* buf = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
        int64_t used = 0;
        for (int64_t i = a; i < b && buf; i++) {
            double L;
            int64_t k = o_mlp_any(p, t, h, i, first_voxel[i], buf + used, cap - used, &L, NULL, 0, NULL);
            if (k < 0) { failed = 1; break; }
            used += k;
            offsets[i + 1] = k;   /* count for now */
            if (chord) chord[i] = L;
        }
        if (!buf) failed = 1;
        cbuf[c] = buf;
        csize[c] = used;
    }
    int64_t total = 0;
    for (int c = 0; c < nchunks; c++) total += csize[c];
    if (failed || total > capacity) total = -1;
    if (total >= 0) {
        offsets[0] = 0;
        for (int64_t i = 0; i < n; i++) offsets[i + 1] += offsets[i];
        int64_t pos = 0;
        for (int c = 0; c < nchunks; c++) {
            memcpy(voxels + pos, cbuf[c], sizeof(uint32_t) * (size_t)csize[c]);
            pos += csize[c];
        }
    }
    for (int c = 0; c < nchunks; c++) free(cbuf[c]);
    free(cbuf);
    free(csize);
    return total;
}

int64_t O(o_mlp_steps)(const pct_params* p, const o_tables* t, const real* hist10, uint32_t first, double* rec,
                    int64_t rec_cap, uint32_t* path, int64_t path_cap, int64_t* path_len, double* chord)
{
    real one[10];
    const real* h[10];
    for (int k = 0; k < 10; k++) { one[k] = hist10[k]; h[k] = &one[k]; }
    int64_t n_rec = 0;
    double L = 0.0;
    int64_t n = o_mlp_any(p, t, h, 0, first, path, path_cap, &L, rec, rec_cap, &n_rec);
    if (path_len) *path_len = n;
    if (chord) *chord = L;
    return n_rec;
}

/* ------------------------------------------------------------------ (e) DROP */
static void o_drop_block_range(const pct_params* p, int64_t n, int64_t start, int64_t* proc)
{
    int64_t m = n - start < p->drop_block_size ? n - start : p->drop_block_size;
    *proc = p->quirk_a3_drop_floor ? (m / p->histories_per_block) * p->histories_per_block : m;
}

static double o_row_chord(const double* chord, int64_t i) { return chord[i]; }

void O(o_drop_iteration)(const pct_params* p, int64_t n, const int64_t* offsets, const uint32_t* voxels,
                      const double* chord, const real* wepl, real* x, int accum, int nthreads)
{
    const int64_t NV = p->num_voxels;
    const double lambda = p->lambda;
    float* U32 = (float*)calloc((size_t)NV, sizeof(float));
    float* K32 = (float*)calloc((size_t)NV, sizeof(float)); /* Kahan compensation */
    double* U64 = (double*)calloc((size_t)NV, sizeof(double));
    uint32_t* S = (uint32_t*)calloc((size_t)NV, sizeof(uint32_t));
    /* per-thread sums, counts and touched-voxel lists of the parallel CPU baseline */
    float* priv_u = NULL;
    uint32_t *priv_s = NULL, *priv_t = NULL;
    if (nthreads < -1) {
        priv_u = (float*)calloc((size_t)(-nthreads) * NV, sizeof(float));
        priv_s = (uint32_t*)calloc((size_t)(-nthreads) * NV, sizeof(uint32_t));
        priv_t = (uint32_t*)malloc((size_t)(-nthreads) * NV * sizeof(uint32_t));
    }
    for (int64_t start = 0; start < n; start += p->drop_block_size) {                   /* :9549-9582 */
        int64_t proc;
        o_drop_block_range(p, n, start, &proc);
        if (nthreads >= -1 && nthreads <= 1) {
            for (int64_t i = start; i < start + proc; i++) {                             /* :9321-9380 */
                int64_t cnt = offsets[i + 1] - offsets[i];
                const uint32_t* a = voxels + offsets[i];
                double L = o_row_chord(chord, i);
                double ax = 0.0;
                for (int64_t j = 0; j < cnt; j++) ax += x[a[j]];
                ax *= L;
                double aa = (double)f_sq(L) * cnt;
                if (cnt >= p->min_mlp_length) {                                         /* :9522-9536 */
                    double val = L * ((wepl[i] - ax) / aa) * lambda;
                    for (int64_t j = 0; j < cnt; j++) {
                        uint32_t v = a[j];
                        S[v] += 1;
                        if (accum == PCT_ACCUM_FP64) U64[v] += val;
                        else if (accum == PCT_ACCUM_KAHAN32) {
                            float y = (float)val - K32[v];
                            float t = U32[v] + y;
                            K32[v] = (t - U32[v]) - y;
                            U32[v] = t;
                        } else U32[v] += (float)val;
                    }
                }
            }
        } else {
#ifdef _OPENMP
          if (nthreads > 1) {
            /* CPU baseline: the histories of a block in parallel, float atomic sums (order varies, as the
             * reference's float atomicAdd) */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
            for (int64_t i = start; i < start + proc; i++) {
                int64_t cnt = offsets[i + 1] - offsets[i];
                const uint32_t* a = voxels + offsets[i];
                double L = o_row_chord(chord, i);
                double ax = 0.0;
                for (int64_t j = 0; j < cnt; j++) ax += x[a[j]];
                ax *= L;
                double aa = (double)f_sq(L) * cnt;
                if (cnt >= p->min_mlp_length) {
                    double val = L * ((wepl[i] - ax) / aa) * lambda;
                    float fv = (float)val;
                    for (int64_t j = 0; j < cnt; j++) {
#pragma omp atomic
                        S[a[j]] += 1;
#pragma omp atomic
                        U32[a[j]] += fv;
                    }
                }
            }
          } else {
            /* the same with per-thread sums (nthreads < -1): each thread sums its histories' updates and
             * counts into its own x_update / S with a list of the voxels it touched, and adds them into the
             * block's at the block end */
            const int nth = -nthreads;
#pragma omp parallel num_threads(nth)
            {
                const int tid = omp_get_thread_num();
                float* Ut = priv_u + (size_t)tid * NV;
                uint32_t* St = priv_s + (size_t)tid * NV;
                uint32_t* Tt = priv_t + (size_t)tid * NV;
                int64_t nt = 0;
#pragma omp for schedule(dynamic, 64)
                for (int64_t i = start; i < start + proc; i++) {
                    int64_t cnt = offsets[i + 1] - offsets[i];
                    const uint32_t* a = voxels + offsets[i];
                    double L = o_row_chord(chord, i);
                    double ax = 0.0;
                    for (int64_t j = 0; j < cnt; j++) ax += x[a[j]];
                    ax *= L;
                    double aa = (double)f_sq(L) * cnt;
                    if (cnt >= p->min_mlp_length) {
                        double val = L * ((wepl[i] - ax) / aa) * lambda;
                        float fv = (float)val;
                        for (int64_t j = 0; j < cnt; j++) {
                            uint32_t v = a[j];
                            if (St[v]++ == 0) Tt[nt++] = v;
                            Ut[v] += fv;
                        }
                    }
                }
                for (int64_t k = 0; k < nt; k++) {   /* the block-end reduction of this thread's sums */
                    uint32_t v = Tt[k];
#pragma omp atomic
                    S[v] += St[v];
#pragma omp atomic
                    U32[v] += Ut[v];
                    St[v] = 0;
                    Ut[v] = 0.0f;
                }
            }
          }
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : -nthreads) schedule(static)
            for (int64_t v = 0; v < NV; v++) {
                if (S[v] > 0) {
                    x[v] += U32[v] / (float)S[v];
                    U32[v] = 0.0f;
                    S[v] = 0;
                }
            }
            continue;
#endif
        }
        for (int64_t v = 0; v < NV; v++) {                                               /* :9398-9419 */
            if (S[v] > 0) {
                if (accum == PCT_ACCUM_FP64) x[v] += (real)(U64[v] / (double)S[v]);
                else x[v] += U32[v] / (float)S[v];
                U32[v] = 0.0f;
                K32[v] = 0.0f;
                U64[v] = 0.0;
                S[v] = 0;
            }
        }
    }
    free(K32);
    free(U32);
    free(U64);
    free(S);
    free(priv_u);
    free(priv_s);
    free(priv_t);
}

void O(o_drop_iteration_fp64)(const pct_params* p, int64_t n, const int64_t* offsets, const uint32_t* voxels,
                           const double* chord, const real* wepl, double* x)
{
    const int64_t NV = p->num_voxels;
    double* U = (double*)calloc((size_t)NV, sizeof(double));
    uint32_t* S = (uint32_t*)calloc((size_t)NV, sizeof(uint32_t));
    for (int64_t start = 0; start < n; start += p->drop_block_size) {
        int64_t proc;
        o_drop_block_range(p, n, start, &proc);
        for (int64_t i = start; i < start + proc; i++) {
            int64_t cnt = offsets[i + 1] - offsets[i];
            const uint32_t* a = voxels + offsets[i];
            double L = o_row_chord(chord, i);
            double ax = 0.0;
            for (int64_t j = 0; j < cnt; j++) ax += x[a[j]];
            ax *= L;
            double aa = L * L * cnt;
            if (cnt >= p->min_mlp_length) {
                double val = L * ((wepl[i] - ax) / aa) * p->lambda;
                for (int64_t j = 0; j < cnt; j++) { S[a[j]] += 1; U[a[j]] += val; }
            }
        }
        for (int64_t v = 0; v < NV; v++)
            if (S[v] > 0) { x[v] += U[v] / S[v]; U[v] = 0.0; S[v] = 0; }
    }
    free(U);
    free(S);
}

double O(o_residual)(int64_t n, const int64_t* offsets, const uint32_t* voxels, const double* chord, const real* wepl,
                  const real* x, int min_len)
{
    double acc = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (offsets[i + 1] - offsets[i] < min_len) continue;
        double s = 0.0;
        for (int64_t j = offsets[i]; j < offsets[i + 1]; j++) s += x[voxels[j]];
        double r = (double)wepl[i] - (double)chord[i] * s;
        acc += r * r;
    }
    return sqrt(acc);
}
