/*
 * pct.h — C-ABI of the MI355X proton-CT reconstruction path.
 *
 * The reference (BlakeSchultze/pCT_Reconstruction) has no library boundary: its stages are
 * `void stage(int num_histories)` functions over global SoA arrays, called in a fixed order
 * from `preprocessing()` (src/pCT_Reconstruction_Data_Segments_Blake.cu:7394) and
 * `image_reconstruction()` (same file :10650).  Each entry point below replaces one of those
 * stage functions; the comment on each names the reference function (file:line, file is
 * src/pCT_Reconstruction_Data_Segments_Blake.cu unless stated) whose effect it reproduces.
 *
 * Conventions
 *   - every function returns PCT_OK (0) or a negative PCT_E* code; pct_last_error() gives text;
 *   - host pointers are caller-owned; device memory is owned by the context;
 *   - one context is used by one host thread at a time; all work runs on the context's stream;
 *   - units are cm and degrees at the boundary, radians inside (as in the reference).
 */
#ifndef PCT_H
#define PCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCT_OK 0
#define PCT_E_ARG -1      /* invalid argument / size                        */
#define PCT_E_STATE -2    /* stage called out of order                      */
#define PCT_E_HIP -3      /* HIP runtime error                              */
#define PCT_E_IO -4       /* file I/O or format error                       */
#define PCT_E_NOMEM -5    /* allocation failure                             */
#define PCT_E_LIMIT -6    /* size beyond a documented limit                 */

/* Enumerations use the integers of settings.cfg (include/pCT_Reconstruction.h:92-94). */
enum pct_hull_type { PCT_HULL_IMPORT = 0, PCT_HULL_SC = 1, PCT_HULL_MSC = 2, PCT_HULL_SM = 3, PCT_HULL_FBP = 4 };
enum pct_filter_type { PCT_FILTER_RAM_LAK = 0, PCT_FILTER_SHEPP_LOGAN = 1, PCT_FILTER_NONE = 2 };
enum pct_x0_type { PCT_X0_IMPORT = 0, PCT_X0_HULL = 1, PCT_X0_FBP = 2, PCT_X0_HYBRID = 3, PCT_X0_ZEROS = 4 };
enum pct_mlp_algorithm { PCT_MLP_STANDARD = 0, PCT_MLP_TABULATED = 1 };
/* DROP per-voxel accumulation precision (the reference uses float atomicAdd, DS.cu:9533):
 * FP32 sums in history order (reference precision), FP64 sums, or FP32 with Kahan compensation. */
enum pct_drop_accum { PCT_ACCUM_FP32 = 0, PCT_ACCUM_FP64 = 1, PCT_ACCUM_KAHAN32 = 2 };
/* DROP scheduling flags (pct_set_launch).  None changes any result bit:
 * SPATIAL_ORDER visits a block's histories sorted by sinogram bin (v, angle, t) so that a wavefront
 * gathers along neighbouring parallel paths; XCD_SLABS gives each of the 8 XCDs a contiguous slab
 * of those histories (row kernel) and of the touched voxels (column kernel); BRICKS keeps the image
 * in 4x4x2-voxel bricks (one 128-byte line each) during DROP, so consecutive voxels of a path and
 * neighbouring paths share cache lines; COLS_LDS runs the per-voxel sums with one wavefront per 64
 * voxels that stages their entries' updates in LDS with coalesced loads; COLS_SLICED stores the
 * ranks of each wavefront's 64 voxels column-major (entry m of the 64 voxels contiguous, padded to
 * the largest count), so each rank load of the per-voxel sums is one coalesced read.  COLS_LDS takes
 * precedence over COLS_SLICED.  BUCKETS (with BRICKS) transposes the lists by tiles of 4 x 4 bricks, each
 * sorted in LDS, instead of one device-wide radix sort per group of blocks; a group with a tile of more
 * than 8192 entries in one block takes the sort. */
enum pct_drop_flags {
    PCT_DROP_SPATIAL_ORDER = 1,
    PCT_DROP_XCD_SLABS = 2,
    PCT_DROP_BRICKS = 4,
    PCT_DROP_COLS_LDS = 8,
    PCT_DROP_COLS_SLICED = 16,
    PCT_DROP_BUCKETS = 32
};

/*
 * Parameters.  Field names follow the settings.cfg keys (settings.cfg:1-164) and the
 * `configurations` struct (include/Configurations.h:14-362, defaults at :60-200).
 * Fields below "derived" are filled by pct_params_derive() with the rules of
 * set_dependent_parameters (include/Configurations.h:1745-1789).
 */
typedef struct pct_params {
    /* scan geometry and binning */
    double gantry_angle_interval; /* GANTRY_ANGLE_INTERVAL [deg] */
    double angular_bin_size;      /* ANGULAR_BIN_SIZE [deg] */
    double ssd_t_size;            /* SSD_T_SIZE [cm] */
    double ssd_v_size;            /* SSD_V_SIZE [cm] */
    double t_shift, u_shift, v_shift; /* T_SHIFT/U_SHIFT/V_SHIFT [cm] */
    double t_bin_size;            /* T_BIN_SIZE [cm] */
    double v_bin_size;            /* V_BIN_SIZE [cm] */
    int32_t sigmas_2_keep;        /* SIGMAS_2_KEEP */
    double recon_cyl_radius;      /* RECON_CYL_RADIUS [cm] */
    int32_t columns, rows;        /* COLUMNS, ROWS */
    double slice_thickness;       /* SLICE_THICKNESS [cm] */
    /* hull detection */
    int32_t msc_diff_thresh;      /* MSC_DIFF_THRESH */
    double sc_lower_threshold, sc_upper_threshold;
    double msc_lower_threshold, msc_upper_threshold;
    double sm_lower_threshold, sm_upper_threshold, sm_scale_threshold;
    int32_t sc_on, msc_on, sm_on, fbp_on; /* SC_ON, MSC_ON, SM_ON, FBP_ON */
    int32_t hull_type;            /* HULL_TYPE (enum pct_hull_type) */
    int32_t fbp_filter_type;      /* FBP_FILTER_TYPE (enum pct_filter_type) */
    int32_t x_0_type;             /* X_0_TYPE (enum pct_x0_type) */
    double fbp_threshold;         /* FBP_THRESHOLD (DS.h:370) RSP threshold of the FBP hull */
    /* reconstruction */
    double lambda;                /* LAMBDA */
    int32_t iterations;           /* ITERATIONS */
    int32_t drop_block_size;      /* DROP_BLOCK_SIZE */
    int32_t histories_per_block;  /* HISTORIES_PER_BLOCK: granularity of the DROP launch (quirk A3) */
    int32_t min_mlp_length;       /* MIN_MLP_LENGTH (DS.h:409) */
    /* MLP lookup tables (DS.h:795-807) */
    double trig_table_min_pis, trig_table_max_pis; /* TRIG_TABLE_MIN/MAX in units of pi */
    double trig_table_step_degrees;                /* TRIG_TABLE_STEP_DEGREES */
    double coeff_table_range, coeff_table_step;    /* scattering-coefficient table [cm] */
    double poly_table_range, poly_table_step;      /* polynomial tables [cm] */
    /* physics (DS.h:741-745) */
    double e_0, x_0_rad_length, rsp_air, source_radius;
    /* behaviour switches; 1 = reproduce the reference */
    int32_t data_in_mm;           /* DATA_IN_MM: inputs given in mm, converted with *= 0.1 */
    int32_t quirk_a1_bin0;        /* out-of-range t/v bins keep bin 0 (DS.cu:5380,5431) */
    int32_t quirk_a3_drop_floor;  /* DROP block processes floor(n/HISTORIES_PER_BLOCK)*HPB (DS.cu:9558) */
    int32_t quirk_voxel_decode;   /* voxel_2_3D_voxels '>' decode used by FBP (DS.cu:11124-11134) */
    int32_t drop_accum;           /* enum pct_drop_accum */
    /* optional image filters, off by default (FBP() DS.cu:6080-6168, hull_selection :7375-7386) */
    int32_t median_filter_fbp;       /* MEDIAN_FILTER_FBP: median of the (2r+1)^2 in-slice window -> x0 */
    int32_t fbp_med_filter_radius;   /* FBP_MED_FILTER_RADIUS r (0..10) */
    int32_t avg_filter_fbp;          /* AVG_FILTER_FBP: takes precedence over the median filter (else-if) */
    int32_t avg_filter_hull;         /* AVG_FILTER_HULL: thresholded window average of the hull -> endpoints */
    int32_t hull_avg_filter_radius;  /* HULL_AVG_FILTER_RADIUS */
    int32_t mlp_algorithm;           /* MLP_ALGORITHM (DS.h:174): 0 = STANDARD (find_MLP_path_GPU :8895), 1 = TABULATED (:9104) */
    double hull_avg_filter_threshold; /* HULL_AVG_FILTER_THRESHOLD (DS.h:483, 0.1) */
    /* total variation superiorization between DROP iterations (DROP_GPU DS.cu:9889-9915, NTVS_iteration
     * :10583, constants DS.h:427-444).  Compiled on in the reference; off here by default (the plain method). */
    int32_t tvs_on;                  /* TVS_ON */
    int32_t tvs_first;               /* TVS_FIRST: 1 = before each DROP iteration, 0 = after */
    int32_t tvs_conditioned;         /* TVS_CONDITIONED: shrink BETA until TV does not grow (:10476) */
    int32_t tvs_old;                 /* TVS_OLD: one repetition and no random draw of L (:9890-9892) */
    int32_t tvs_repetitions;         /* TVS_REPETITIONS (DS.h:443, 5) */
    int32_t tvs_l0;                  /* L_0 (DS.h:437, 0) */
    int32_t tvs_seed;                /* seed of the draw of L (the reference seeds from the clock, :11480) */
    int32_t list_segment_histories;  /* LIST_SEGMENT_HISTORIES (no reference key; the role of MLP_IN_LOOP and the
                                      * MAX_*_HISTORIES data segments, DS.cu:9590-9742): 0 = every MLP list and the
                                      * DROP plan stay on the device; > 0 = they are built per segment of that many
                                      * reconstruction histories (rounded down to whole DROP blocks) inside each
                                      * DROP iteration and each residual, so device memory holds one segment */
    double tvs_a;                    /* A (DS.h:436, 0.5): BETA = A^L */
    double tvs_min_eta;              /* TVS_MIN_ETA (DS.h:434, 1e-40): length of the A^L table */
    double tv_threshold;             /* TV_THRESHOLD: the reference's (1/1000) is integer division, 0 */

    /* ---- derived (pct_params_derive) ---- */
    int32_t t_bins, v_bins, angular_bins, num_bins;
    int32_t slices, num_voxels;
    double recon_cyl_height, image_width, image_height, image_thickness;
    double voxel_width, voxel_height, voxel_thickness;
    double x_zero_coordinate, y_zero_coordinate, z_zero_coordinate;
    double mlp_u_step, ram_lak_tau, pi;
    double trig_table_min, trig_table_max, trig_table_step; /* radians */
    int32_t trig_table_elements, coeff_table_elements, poly_table_elements;
} pct_params;

/* sizeof(pct_params), for bindings that mirror the struct. */
size_t pct_params_size(void);
/* Defaults of the configurations constructor (include/Configurations.h:60-200). */
void pct_params_default(pct_params* p);
/* Parse a settings.cfg file ("KEY = value // comment", typing of V.cu:9275-9344) over *p. */
int pct_params_load_cfg(pct_params* p, const char* path, char* err, size_t err_len);
/* Set one key from text, same typing rules.  Unknown keys return PCT_E_ARG. */
int pct_params_set(pct_params* p, const char* key, const char* value);
/* Fill the derived fields (include/Configurations.h:1745-1789); validates sizes. */
int pct_params_derive(pct_params* p, char* err, size_t err_len);

/* ---------------------------------------------------------------- context ---- */
typedef struct pct_ctx pct_ctx;

/* Create a context on HIP device `device`; copies *p (which must be derived). */
int pct_ctx_create(const pct_params* p, int device, pct_ctx** out);
int pct_ctx_destroy(pct_ctx* ctx);
const char* pct_last_error(const pct_ctx* ctx);
/* Block until all queued work of the context has finished. */
int pct_ctx_synchronize(pct_ctx* ctx);
/* Number of histories currently held at each stage. */
enum pct_count_kind { PCT_N_LOADED = 0, PCT_N_RECON_VOLUME = 1, PCT_N_POST_CUTS = 2, PCT_N_RECONSTRUCTION = 3,
                      PCT_N_LIST_ENTRIES = 4, PCT_N_CURRENT = 5 /* histories held by the latest stage */ };
int pct_get_count(pct_ctx* ctx, int kind, int64_t* out);

/*
 * Load raw tracker histories (replaces read_data_chunk_v0, DS.cu:4361-4467, after the file read):
 * 13 SoA float arrays t0..t3, v0..v3, u0..u3, WEPL in the PCTD order (t_in_1 = t0, t_in_2 = t1,
 * t_out_1 = t2, t_out_2 = t3), plus the per-history gantry angle in degrees (the integer file
 * angle for PCTD, the float per-history angle for PCTN).  If params.data_in_mm, coordinates and
 * WEPL are scaled by 0.1 (convert_mm_2_cm, DS.cu:4070); then the t/u/v shifts are applied
 * (apply_tuv_shifts, DS.cu:4090).  May be called repeatedly to append.
 */
int pct_load_histories(pct_ctx* ctx, size_t n, const float* const t[4], const float* const v[4],
                       const float* const u[4], const float* wepl, const float* angle_deg);

/* (a)+(b) part 1: recon_volume_intersections (DS.cu:4767, kernel :4890, :5286), binning
 * (DS.cu:5373, kernel :5482) with the compaction of :5429-5449, and the SC/MSC/SM carving
 * counts of hull_detection (DS.cu:6607) for the enabled methods. */
int pct_intersect_bin_carve(pct_ctx* ctx);
/* (b) part 2: hull_detection_finish (DS.cu:7276): MSC/SM edge detection. */
int pct_hull_finish(pct_ctx* ctx);
/* (a) part 2: statistical_calculations_and_cuts (DS.cu:5972) incl. the sinogram sums. */
int pct_stats_and_cuts(pct_ctx* ctx);
/* (c): construct_sinogram (DS.cu:6039) + FBP (DS.cu:6080: filter_GPU :6206,
 * backprojection_GPU2 :6340, FBP_image_2_hull :6592 when HULL_TYPE = FBP). */
int pct_fbp(pct_ctx* ctx);
/* hull_selection (DS.cu:7351). */
int pct_select_hull(pct_ctx* ctx);
/* define_initial_iterate (DS.cu:8612); x0 is read only for X_0_TYPE = IMPORT (NUM_VOXELS floats). */
int pct_initial_iterate(pct_ctx* ctx, const float* x0_import);
/* (d) prerequisite: reconstruction_cuts (DS.cu:8296) with collect_MLP_endpoints_GPU_nobool
 * (:7819) and the compaction of :8286-8294. */
int pct_mlp_endpoints(pct_ctx* ctx);
/* (d): MLP path lists: for every reconstruction history the voxel list of
 * find_MLP_path_GPU_tabulated (DS.cu:9104-9262) and its effective chord length
 * (EffectiveChordLength_GPU, :8873).  Built once, reused by every DROP iteration. */
int pct_mlp_build_lists(pct_ctx* ctx);
/* (e): `iterations` DROP iterations (DROP_GPU, DS.cu:9848; DROP_full_tx_iteration :9537;
 * DROP_block_update_GPU :9522; image_update_GPU :9398) on the current image. */
int pct_drop_iterate(pct_ctx* ctx, int iterations);
/* Run every stage above in the reference order, with params.iterations DROP iterations. */
int pct_run_pipeline(pct_ctx* ctx);
/* Reset the iterate x to the initial iterate x0 (on the device). */
int pct_reset_iterate(pct_ctx* ctx);
/* Launch shape of the DROP kernels: lanes per history in the row kernel (4, 8, 16, 32, 64), voxels
 * per lane in the column kernel (1, 2, 4, 8) and enum pct_drop_flags (a change of flags rebuilds
 * the transposed lists at the next DROP call).  Results do not depend on any of them. */
int pct_set_launch(pct_ctx* ctx, int drop_row_lanes, int drop_col_voxels, int drop_flags);
/* Run `iterations` DROP iterations with a HIP event pair around every kernel launch and report the
 * average duration of one block's row kernel (per-history updates) and column kernel (per-voxel
 * sums) in ms, the number of blocks per iteration, and per iteration the histories processed,
 * list entries gathered, list entries summed per voxel and voxels touched.  Updates x like
 * pct_drop_iterate. */
int pct_drop_profile(pct_ctx* ctx, int iterations, double* rows_ms, double* cols_ms, int64_t* launches,
                     int64_t* histories, int64_t* gathered, int64_t* scattered, int64_t* touched);
/* Total variation superiorization record of the DROP iterations run since the initial iterate (or
 * pct_set_image / pct_reset_iterate): for each NTVS_iteration (DS.cu:10583) the total variation of x
 * before and after it (the TV_x_values pushes of :10587 and :10596) and the exponent L after it.
 * *count receives the number of records; at most `capacity` are copied (arrays may be NULL). */
int pct_get_tvs_log(pct_ctx* ctx, size_t capacity, double* tv_before, double* tv_after, int32_t* l_after,
                    int64_t* count);
/* Total variation of the current iterate (calculate_total_variation, DS.cu:10165): the float terms
 * sqrt(dy^2 + dx^2) over rows < ROWS-1, columns < COLUMNS-1 of every slice, summed in FP64 in the
 * library's fixed order (tiles of 1024 voxels). */
int pct_total_variation(pct_ctx* ctx, double* tv);

/* ---------------------------------------------------------------- outputs ---- */
enum pct_image_kind { PCT_IMG_X = 0, PCT_IMG_FBP = 1, PCT_IMG_X0 = 2, PCT_IMG_HULL = 3,
                      PCT_IMG_SC_HULL = 4, PCT_IMG_MSC_COUNTS = 5, PCT_IMG_SM_COUNTS = 6,
                      PCT_IMG_MSC_HULL = 7, PCT_IMG_SM_HULL = 8, PCT_IMG_FBP_HULL = 9,
                      PCT_IMG_FBP_MEDIAN = 10,    /* median-filtered FBP (MEDIAN_FILTER_FBP) */
                      PCT_IMG_HULL_FILTERED = 11  /* average-filtered hull (AVG_FILTER_HULL) */ };
/* Copy an image (NUM_VOXELS values, index x + y*COLUMNS + z*COLUMNS*ROWS) to host as float. */
int pct_get_image(pct_ctx* ctx, int kind, float* out);
/* Overwrite the current iterate x (NUM_VOXELS floats), e.g. a learned warm start. */
int pct_set_image(pct_ctx* ctx, const float* x);
/* Hull for HULL_TYPE = IMPORT (NUM_VOXELS bytes, 1 = inside); call before pct_select_hull. */
int pct_set_hull(pct_ctx* ctx, const uint8_t* hull);
enum pct_bin_kind { PCT_BIN_COUNTS_PRE = 0, PCT_BIN_MEAN_WEPL = 1, PCT_BIN_MEAN_UT = 2, PCT_BIN_MEAN_UV = 3,
                    PCT_BIN_STD_WEPL = 4, PCT_BIN_STD_UT = 5, PCT_BIN_STD_UV = 6, PCT_BIN_COUNTS_POST = 7,
                    PCT_BIN_SINOGRAM = 8, PCT_BIN_SINOGRAM_FILTERED = 9 };
/* Copy a per-bin array (NUM_BINS values) to host as float (counts converted exactly). */
int pct_get_bins(pct_ctx* ctx, int kind, float* out);
/* Per-history arrays of the current stage (count = pct_get_count of that stage). */
enum pct_hist_kind { PCT_H_X_ENTRY = 0, PCT_H_Y_ENTRY, PCT_H_Z_ENTRY, PCT_H_X_EXIT, PCT_H_Y_EXIT, PCT_H_Z_EXIT,
                     PCT_H_XY_ENTRY_ANGLE, PCT_H_XZ_ENTRY_ANGLE, PCT_H_XY_EXIT_ANGLE, PCT_H_XZ_EXIT_ANGLE,
                     PCT_H_WEPL, PCT_H_BIN_NUM /* int32 */, PCT_H_FIRST_VOXEL /* uint32 */,
                     PCT_H_PATH_COUNT /* uint32 */, PCT_H_CHORD /* double */,
                     PCT_H_GANTRY_ANGLE /* float, degrees: the angle loaded with the history */ };
int pct_get_histories(pct_ctx* ctx, int kind, void* out, size_t capacity);
/* MLP list of the reconstruction histories: offsets (n+1 int64) and voxel indices (uint32). */
int pct_get_lists(pct_ctx* ctx, int64_t* offsets, size_t offsets_cap, uint32_t* voxels, size_t voxels_cap);
/* Data residual ||A x - b||_2 over the reconstruction histories for the current x, where
 * row i of A is chord_i on each listed voxel (DROP's system, DS.cu:9163-9174). */
int pct_residual(pct_ctx* ctx, double* norm);
/* Stage timings (ms, HIP events) and counts of the last run as a JSON object. */
const char* pct_metrics_json(pct_ctx* ctx);

/* ---------------------------------------------------------------- DROP kernel boundary ---- */
/*
 * One DROP iteration on caller-owned device buffers (no context, no allocation).  This is the hot
 * loop of DROP_full_tx_iteration (DS.cu:9537-9589) over precomputed path lists:
 *   offsets[n+1] (int64) into voxels[] (uint32): history i owns voxels[offsets[i] .. offsets[i+1]);
 *   chord[n] (double, the per-history effective chord length), wepl[n] (float);
 *   x[num_voxels] (float, updated in place); accum[num_voxels] (double) and counts[num_voxels]
 *   (uint32) must be zero on entry and are zero on return.
 * Blocks of block_size histories, each processing floor(m/histories_per_block)*hpb histories
 * when quirk_a3 != 0.  Work is queued on hip_stream (NULL = default stream); not synchronised.
 */
int pct_drop_iteration_device(int64_t n, const int64_t* offsets, const uint32_t* voxels, const double* chord,
                              const float* wepl, float* x, double* accum, uint32_t* counts, int64_t num_voxels,
                              int block_size, int histories_per_block, int quirk_a3, int min_mlp_length,
                              double lambda, void* hip_stream);
/*
 * No counterpart in the reference.  pct_drop_iteration_device for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same iteration and result, with the per-voxel sums and counts in a library-owned device array of
 * 2^23 voxels (96 MiB, held by each device the library runs on until the process ends) instead of
 * caller workspace.  Calls no allocating routine.  Returns PCT_E_LIMIT when num_voxels > 2^23.  The
 * array is cleared at the start of each call; calls that overlap on one device share it and race.
 */
int pct_drop_iteration_device_noalloc(int64_t n, const int64_t* offsets, const uint32_t* voxels,
                                      const double* chord, const float* wepl, float* x, int64_t num_voxels,
                                      int block_size, int histories_per_block, int quirk_a3, int min_mlp_length,
                                      double lambda, void* hip_stream);

/* ---------------------------------------------------------------- I/O ---- */
/*
 * Preprocessed data of the reference's variant program (src/pCT_Reconstruction.cu, "V.cu"):
 * accumulate_and_export_preprocessed_data (V.cu:6613-6658) with export_hull (:6540), export_x_0
 * (:6551), export_histories (:6571), export_voxels_per_path (:6594) and export_avg_chord_lengths
 * (:6603).  Writes into directory `dir` (names: include/Constants.h:84-88 and the write-mode
 * extensions of include/Configurations.h:1802-1822):
 *   MLP.bin               uint32 n, then every path's voxel indices, concatenated (uint32)
 *   voxels_per_path.bin   uint32 n, then n uint32 path lengths
 *   avg_chord_lengths.bin uint32 n, then n float effective chord lengths
 *   WEPL.bin              uint32 n, then n float WEPL
 *   histories.bin         uint32 n, then 15 arrays of n: entry voxel x, y, z (int32); x, y, z entry;
 *                         x, y, z exit; xy, xz entry angle; xy, xz exit angle (float); bin number and
 *                         gantry angle (int32)
 *   hull.txt              NUM_VOXELS bytes, 1 = inside (binary despite the name, as the reference writes it)
 *   x_0.txt               NUM_VOXELS floats (binary, likewise)
 * for the n reconstruction histories in order (after pct_mlp_build_lists).  With
 * PCT_EXPORT_CHORD_F64 it also writes avg_chord_lengths_f64.bin (uint32 n, n doubles): the chords at
 * full precision, so that an import reproduces DROP bit for bit.
 */
enum pct_export_flags { PCT_EXPORT_CHORD_F64 = 1 };
int pct_export_preprocessed(pct_ctx* ctx, const char* dir, int flags);
/*
 * init_import_preprocessed_data (V.cu:6775-6861), import_hull (:6659) and import_x_0 (:6672): reads
 * the files above into a context at any stage and leaves it ready for pct_drop_iterate, with x = x0.
 * MLP.bin, voxels_per_path.bin, avg_chord_lengths.bin (or _f64), WEPL.bin and x_0.txt are
 * required; the four per-history files must agree on n (V.cu:6810).  hull.txt (needed by TVS) and
 * histories.bin are optional.  MLP.bin and WEPL.bin may carry the "%u
" text header that
 * accumulate_and_export_preprocessed_data writes (V.cu:6628-6629) in place of the uint32.  Voxel
 * indices >= NUM_VOXELS and paths longer than 4 (COLUMNS + ROWS + SLICES) + 16 are rejected.
 */
int pct_import_preprocessed(pct_ctx* ctx, const char* dir);

/* Read one PCTD/PCTN projection file (count_histories_v0 :3578, read_data_chunk_v0 :4361,
 * format SupplementalDocumentation/ProposedDataFormat_Version1-1.txt).  With arrays NULL only
 * the header is read into *n and *angle_deg.  Version 1 / PCTN files carry per-history angles
 * which are returned in angles (may be NULL). */
int pct_read_projection(const char* path, int64_t* n, float* header_angle_deg, float* const t[4], float* const v[4],
                        float* const u[4], float* wepl, float* angles);
/* Write a PCTD (version 0) or PCTN (version 1, angles != NULL) projection file. */
int pct_write_projection(const char* path, int64_t n, float angle_deg, float energy_mev, const float* const t[4],
                         const float* const v[4], const float* const u[4], const float* wepl, const float* angles);
/* Text image in the array_2_disk layout (DS.cu:10743): COLUMNS values + ' ' per line, ROWS*SLICES lines,
 * default ostream formatting (6 significant digits).  Slices are formatted on up to 16 host threads
 * and written in order; PCT_E_NOMEM if formatting runs out of memory, PCT_E_IO on a write error. */
int pct_write_image_text(const char* path, const float* image, int columns, int rows, int slices);
/* Binary image of write_PNG (DS.cu:9939): float32 value/2 in x-fastest order. */
int pct_write_image_dat(const char* path, const float* image, int columns, int rows, int slices);
/* The PNG write_PNG makes from that file with ImageMagick (DS.cu:9969-9971): 16-bit grayscale,
 * COLUMNS x ROWS*SLICES, sample = round(clamp(value/2, 0, 1) * 65535). */
int pct_write_image_png(const char* path, const float* image, int columns, int rows, int slices);
/*
 * The context's images under the reference's output names in directory `dir` (which must exist), as its
 * stages write them when their WRITE_* switches are on (DS.h:315-322):
 *   PCT_WRITE_FBP   FBP.txt                       (FBP, DS.cu:6112-6116)
 *   PCT_WRITE_HULL  hull.txt, hull.dat, hull.png  (hull_selection, DS.cu:7365-7370)
 *   PCT_WRITE_X0    x_0.txt, x_0.dat, x_0.png     (define_initial_iterate, DS.cu:8637-8641)
 *   PCT_WRITE_X     x_<k>.txt, x_<k>.dat, x_<k>.png for the current iterate after k DROP iterations since the
 *                   initial iterate (DROP_GPU, DS.cu:9915-9922)
 * Text in the array_2_disk layout, .dat and .png as write_PNG (pct_write_image_text/_dat/_png).  Images
 * whose stage has not run are skipped.
 */
enum pct_write_flags { PCT_WRITE_FBP = 1, PCT_WRITE_HULL = 2, PCT_WRITE_X0 = 4, PCT_WRITE_X = 8 };
int pct_write_reference_layout(pct_ctx* ctx, const char* dir, int flags);

/* ---------------------------------------------------------------- synthetic data ---- */
/* Phantom made of up to 16 objects; each is a cylinder (kind 0, axis z) or an ellipsoid (kind 1)
 * with centre, semi-axes (cylinder: radius a, half height c), rotation about z [deg] and an RSP
 * that is ADDED where the object lies (Shepp-Logan convention). */
typedef struct pct_object { int32_t kind; double cx, cy, cz, a, b, c, phi_deg, rsp; } pct_object;
typedef struct pct_phantom { int32_t n_objects; pct_object obj[16]; } pct_phantom;
typedef struct pct_beam {
    double energy_mev;             /* kinetic energy */
    double u_planes[4];            /* tracker u positions [cm], e.g. -20,-15,15,20 */
    double t_half_range, v_half_range; /* entry position ranges [cm] at u_planes[1] */
    int32_t continuous_angles;     /* 0: n_angles equally spaced integer angles; 1: U[0,360) */
    int32_t n_angles; double angle_interval;
    double wepl_noise_sigma;       /* [cm] */
    double outlier_fraction;       /* fraction of histories with WEPL replaced by U(0, 2*max) */
    int32_t scatter;               /* 1: Highland multiple-scattering deviation of the exit */
} pct_beam;
void pct_phantom_cylinder_c1(pct_phantom* ph);   /* water cylinder + 3 inserts */
void pct_phantom_head(pct_phantom* ph);          /* 10-ellipsoid head */
void pct_phantom_random(pct_phantom* ph, uint64_t seed, int n_ellipsoids, double radius, double half_height);
/* Generate n histories (cm, angles in degrees) sorted by projection (angle-major) as in the
 * per-angle projection files; deterministic in (seed, n) for any thread count. */
int pct_generate_histories(const pct_phantom* ph, const pct_beam* beam, uint64_t seed, int64_t n, float* const t[4],
                           float* const v[4], float* const u[4], float* wepl, float* angle_deg);
/* Rasterise the phantom's RSP on the reconstruction grid of *p (voxel centres, NUM_VOXELS floats). */
int pct_phantom_rasterize(const pct_phantom* ph, const pct_params* p, float* out);

/*
 * Learned warm start (no reference counterpart: the reference's initial iterate is FBP, the hull or
 * their product, define_initial_iterate DS.cu:8612-8646).  A 4-level 3D U-Net (16-32-64-128
 * channels, 3x3x3 convolutions, BF16 weights and activations, FP32 accumulation on MFMA) maps
 * (FBP image, hull) to x0 = hull * (FBP + correction).  Grid dimensions must be multiples of 8.
 * Parameters are one FP32 vector: per convolution W[Cout][27][Cin] then b[Cout] for enc1..enc4,
 * dec3, dec2, dec1 (Cin 8/16/32/64/192/96/48, Cout 16/32/64/128/64/32/16), then the head w[16], b.
 */
typedef struct pct_net pct_net;
int pct_net_create(int columns, int rows, int slices, uint64_t seed, int device, pct_net** out);
int pct_net_destroy(pct_net* net);
const char* pct_net_last_error(pct_net* net);
int64_t pct_net_num_params(pct_net* net);
int pct_net_get_params(pct_net* net, float* out, int64_t n);
int pct_net_set_params(pct_net* net, const float* in, int64_t n);
/* Masked MSE loss over hull voxels of x0 against `target`, and its gradient (grad may be NULL). */
int pct_net_loss_grad(pct_net* net, const float* fbp, const uint8_t* hull, const float* target, double* loss,
                      float* grad);
int pct_net_predict(pct_net* net, const float* fbp, const uint8_t* hull, float* x0);
/* Training set held in HBM: (FBP, hull, converged DROP image) triples. */
int pct_net_add_sample(pct_net* net, const float* fbp, const uint8_t* hull, const float* target);
int pct_net_num_samples(pct_net* net);
/* `steps` Adam steps (beta 0.9/0.999), one sample per step in a seeded shuffled order. */
int pct_net_train(pct_net* net, int steps, double lr, uint64_t seed, double* mean_loss);
int pct_net_save(pct_net* net, const char* path);
int pct_net_load(pct_net* net, const char* path);
/* Replace the initial iterate of `ctx` (after pct_mlp_build_lists) by the net's x0 when its DROP data
 * residual ||Ax - b|| is below that of the current x0; x is reset to the chosen x0. */
int pct_net_apply(pct_net* net, pct_ctx* ctx, int* used_net, double* residual_fbp, double* residual_net);
/* The same judgement for a net whose input is the current iterate x (e.g. after a few DROP iterations
 * from the FBP start, for a net trained on such iterates): x is replaced by net(x, hull) when that
 * lowers ||Ax - b||; x0 is left unchanged. */
int pct_net_apply_iterate(pct_net* net, pct_ctx* ctx, int* used_net, double* residual_x, double* residual_net);

#ifdef __cplusplus
}
#endif
#endif /* PCT_H */
