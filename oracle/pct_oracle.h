/*
 * pct_oracle.h — CPU restatement of the reference's reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP implementation: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.  The product
 * library never links or calls it.
 *
 * Every function names the reference lines it restates (file src/pCT_Reconstruction_Data_Segments_Blake.cu
 * unless stated).  Arithmetic follows the reference's types: double where the reference computes in
 * double, float where it stores or calls a float function.  Float library functions of CUDA are
 * given one fixed meaning (the IEEE round-to-nearest value of the exact result, computed as the
 * double function rounded to float; powf(x, 2) = x*x in float).  Float atomic sums are accumulated
 * in ascending history order, the order a single-threaded run produces.
 */
#ifndef PCT_ORACLE_H
#define PCT_ORACLE_H
#include <stdint.h>
#include "../include/pct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/Configurations.h:1745-1789 restated. */
void o_derive(pct_params* p);

/* recon_volume_intersections_GPU (:4890-5021) with calculate_intercepts (:5286-5372).
 * in: t[4], v[4], u[4] (t_in_1, t_in_2, t_out_1, t_out_2 ...), angle [deg];
 * out: h[10] = x_entry, y_entry, z_entry, x_exit, y_exit, z_exit, xy_entry, xz_entry, xy_exit, xz_exit. */
void o_intersections(const pct_params* p, int64_t n, const float* const t[4], const float* const v[4],
                     const float* const u[4], const float* angle_deg, float* const h[10], uint8_t* missed);

/* binning_GPU (:5482-5553): bin number and whether t/v bins were in range (quirk A1 gives bin 0). */
void o_binning(const pct_params* p, int64_t n, const float* const h[10], int32_t* bin_num, uint8_t* in_range);

/* binning sums (:5543-5546), calculate_means_GPU (:5644), sum_squared_deviations_GPU (:5689),
 * calculate_standard_deviations_GPU (:5737), statistical_cuts_GPU (:5937).
 * mean[3] / stddev[3] = WEPL, rel ut angle, rel uv angle.  counts_post and sinogram are the sums
 * after the cuts (construct_sinogram_GPU is o_construct_sinogram). */
void o_stats_and_cuts(const pct_params* p, int64_t n, const int32_t* bin_num, const uint8_t* in_range,
                      const float* wepl, const float* const h[10], int32_t* counts_pre, float* const mean[3],
                      float* const stddev[3], int32_t* counts_post, float* sinogram, uint8_t* failed);

/* construct_sinogram_GPU (:6073): sinogram /= count. */
void o_construct_sinogram(const pct_params* p, const int32_t* counts_post, float* sinogram);

/* initialize_hull_GPU (:6659): 1 inside the cylinder (used for SC bool hull and MSC/SM int counts). */
void o_init_hull_u8(const pct_params* p, uint8_t* hull);
void o_init_hull_i32(const pct_params* p, int32_t* counts);
/* SC_GPU (:6681), MSC_GPU (:6799), SM_GPU (:6997) with take_2D/3D_step_GPU (:11346-11459). */
void o_sc(const pct_params* p, int64_t n, const float* const h[10], const float* wepl, uint8_t* sc_hull);
void o_msc(const pct_params* p, int64_t n, const float* const h[10], const float* wepl, int32_t* counts);
void o_sm(const pct_params* p, int64_t n, const float* const h[10], const float* wepl, int32_t* counts);
/* MSC_edge_detection_GPU (:6911) evaluated on the input counts (two-buffer semantics of :6958). */
void o_msc_edge(const pct_params* p, const int32_t* counts, int32_t* hull);
/* SM_edge_detection (:7100): carve_differences (:6616), per-slice host threshold (:7118-7134),
 * SM_edge_detection_GPU (:7160). */
void o_sm_edge(const pct_params* p, const int32_t* counts, int32_t* hull);

/* filter_GPU (:6206-6240). */
void o_filter(const pct_params* p, const float* sinogram, float* filtered);
/* backprojection_GPU2 (:6340-6479). */
void o_backproject(const pct_params* p, const float* filtered, float* image);
/* FBP_image_2_hull_GPU (:6592). */
void o_fbp_hull(const pct_params* p, const float* image, uint8_t* hull);
void o_median_filter_2d(const pct_params* p, const float* in, int radius, float* out);
void o_avg_filter_hull(const pct_params* p, const uint8_t* in, int radius, double threshold, uint8_t* out);

/* Total variation superiorization (pct_tvs_oracle.c; DS.cu:10149-10602).  mode 0 = the reference's
 * serial float sums, 1 = the library's tiled FP64 sums. */
double o_tile_sum(const double* t, int64_t n);
int o_tvs_eta_len(const pct_params* p);
float o_tvs_eta(const pct_params* p, int L);
int o_tvs_draw(int seed, int k, int lo, int hi);
float o_tv(const pct_params* p, const float* img, int mode);
float o_tvs_perturbation(const pct_params* p, const float* img, float* G, float* v, int mode);
void o_tvs_iteration(const pct_params* p, const uint8_t* hull, float* x, int k, int* L, int mode,
                     float* tv_before, float* tv_after);

/* MLP lookup tables (:8421-8502), FP64. */
typedef struct o_tables {
    int64_t n_trig, n_coeff, n_poly;
    double *sin_t, *cos_t, *scat, *p12, *p23, *p34, *p26, *p312;
} o_tables;
int o_tables_create(const pct_params* p, o_tables* t);
void o_tables_destroy(o_tables* t);

/* collect_MLP_endpoints_GPU_nobool (:7819) with find_MLP_endpoints_GPU (:7581).  Rewrites the
 * entry/exit coordinates in h for histories that hit the hull; first_voxel = NUM_VOXELS otherwise. */
void o_endpoints(const pct_params* p, const uint8_t* hull, int64_t n, float* const h[10], uint32_t* first_voxel);

/* EffectiveChordLength_GPU (:8873-8894). */
double o_chord_length(const pct_params* p, double abs_angle_t, double abs_angle_v);

/* find_MLP_path_GPU_tabulated (:9104-9262) for all histories: voxel lists and chord lengths.
 * offsets has n+1 entries; returns total entries or -1 if capacity is exceeded. */
int64_t o_mlp_lists(const pct_params* p, const o_tables* t, int64_t n, const float* const h[10],
                    const uint32_t* first_voxel, int64_t* offsets, uint32_t* voxels, int64_t capacity, double* chord);

/* The MLP loop of o_mlp_lists for one history (hist10 = the 10 geometry values of one history, in the
 * h[] order), recording every u-step: rec[6 k + 0..5] = u (the rotated depth u_shifted), t_1, v_1, x_1,
 * y_1 and the voxel index of step k.  Returns the number of steps (at most rec_cap are recorded); the
 * voxel list goes to path (*path_len entries, -1 past path_cap) and the chord to *chord. */
int64_t o_mlp_steps(const pct_params* p, const o_tables* t, const float* hist10, uint32_t first, double* rec,
                    int64_t rec_cap, uint32_t* path, int64_t path_cap, int64_t* path_len, double* chord);

/* One DROP iteration over precomputed lists (DROP_full_tx_iteration :9537, calculate_x_update_GPU
 * :9321, DROP_block_update_GPU :9522, image_update_GPU :9398) in history order.
 * accum = PCT_ACCUM_FP32: x_update in float (reference); PCT_ACCUM_FP64: x_update in double.
 * nthreads > 1: OpenMP over histories of a block with float atomic sums (CPU baseline timing; the
 * summation order then varies); nthreads < -1: |nthreads| threads with per-thread sums, counts and
 * touched-voxel lists added into the block's at the block end. */
void o_drop_iteration(const pct_params* p, int64_t n, const int64_t* offsets, const uint32_t* voxels,
                      const double* chord, const float* wepl, float* x, int accum, int nthreads);
/* The same iteration with the image and every sum in double (the FP64 oracle). */
void o_drop_iteration_fp64(const pct_params* p, int64_t n, const int64_t* offsets, const uint32_t* voxels,
                           const double* chord, const float* wepl, double* x);
/* ||A x - b|| with A_ij = chord_i for each listed voxel, over rows with at least min_len voxels. */
double o_residual(int64_t n, const int64_t* offsets, const uint32_t* voxels, const double* chord,
                  const float* wepl, const float* x, int min_len);

#ifdef __cplusplus
}
#endif
#endif
