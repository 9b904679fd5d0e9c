// params.cpp — parameter defaults, settings.cfg parsing and derived quantities.
//
// Defaults: the `configurations` constructor (include/Configurations.h:60-200), with the
// physics and table constants of include/pCT_Reconstruction_Data_Segments_Blake.h:741-807.
// Parsing: read_key_value_pair (src/pCT_Reconstruction.cu:9275-9344): "KEY = value // comment";
// `true`/`false` -> bool, no '.' -> integer, '.' -> double, quoted -> string.  A double given for an
// integer key is truncated (static_cast<int>, V.cu:9330), so `ROWS = 200.4` means 200.
// Derivation: set_dependent_parameters (include/Configurations.h:1745-1789).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/pct.h"

extern "C" size_t pct_params_size(void) { return sizeof(pct_params); }

extern "C" void pct_params_default(pct_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->gantry_angle_interval = 4;
    p->angular_bin_size = 4.0;
    p->ssd_t_size = 35.0;
    p->ssd_v_size = 9.0;
    p->t_bin_size = 0.1;
    p->v_bin_size = 0.25;
    p->sigmas_2_keep = 3;
    p->recon_cyl_radius = 10.0;
    p->columns = 200;
    p->rows = 200;
    p->slice_thickness = 0.25;
    p->msc_diff_thresh = 50;
    p->sc_lower_threshold = -10.0;
    p->sc_upper_threshold = 0.0;
    p->msc_lower_threshold = -10.0;
    p->msc_upper_threshold = 0.0;
    p->sm_lower_threshold = 6.0;
    p->sm_upper_threshold = 21.0;
    p->sm_scale_threshold = 1.0;
    p->sc_on = 0;
    p->msc_on = 1;
    p->sm_on = 0;
    p->fbp_on = 1;
    p->hull_type = PCT_HULL_MSC;
    p->fbp_filter_type = PCT_FILTER_SHEPP_LOGAN;
    p->x_0_type = PCT_X0_HYBRID;
    p->fbp_threshold = 0.6;
    p->lambda = 0.0001;
    p->iterations = 12;
    p->drop_block_size = 3200;
    p->histories_per_block = 320;
    p->min_mlp_length = 30;
    p->trig_table_min_pis = -2.0;
    p->trig_table_max_pis = 4.0;
    p->trig_table_step_degrees = 0.001;
    // DS.h:802-806 (COEFF/POLY_TABLE_RANGE 40 cm): covers chords up to 2 * RECON_CYL_RADIUS = 40 cm.
    p->coeff_table_range = 40.0;
    p->coeff_table_step = 0.00005;
    p->poly_table_range = 40.0;
    p->poly_table_step = 0.00005;
    p->e_0 = 13.6;
    p->x_0_rad_length = 36.08;
    p->rsp_air = 0.00113;
    p->source_radius = 260.7;
    p->data_in_mm = 0;
    p->quirk_a1_bin0 = 1;
    p->quirk_a3_drop_floor = 1;
    p->quirk_voxel_decode = 1;
    p->drop_accum = PCT_ACCUM_FP64;
    // filters: Configurations.h:154-168 (off, radius 0); threshold DS.h:483
    p->median_filter_fbp = 0;
    p->fbp_med_filter_radius = 0;
    p->avg_filter_fbp = 0;
    p->avg_filter_hull = 0;
    p->hull_avg_filter_radius = 0;
    p->mlp_algorithm = PCT_MLP_TABULATED;  // DS.h:174
    p->hull_avg_filter_threshold = 0.1;
    // total variation superiorization: DS.h:427-444 except TVS_ON (off here, the plain method)
    p->tvs_on = 0;
    p->tvs_first = 1;
    p->tvs_conditioned = 1;
    p->tvs_old = 0;
    p->tvs_repetitions = 5;
    p->tvs_l0 = 0;
    p->tvs_seed = 0;
    p->list_segment_histories = 0;
    p->tvs_a = 0.5;
    p->tvs_min_eta = 1e-40;
    p->tv_threshold = 0.0;
}

namespace {

enum Kind { K_DOUBLE, K_INT, K_BOOL };
struct Key {
    const char* name;
    Kind kind;
    size_t offset;
};
#define PK_D(k, f) {k, K_DOUBLE, offsetof(pct_params, f)}
#define PK_I(k, f) {k, K_INT, offsetof(pct_params, f)}
#define PK_B(k, f) {k, K_BOOL, offsetof(pct_params, f)}
const Key kKeys[] = {
    PK_D("GANTRY_ANGLE_INTERVAL", gantry_angle_interval), PK_D("ANGULAR_BIN_SIZE", angular_bin_size),
    PK_D("SSD_T_SIZE", ssd_t_size), PK_D("SSD_V_SIZE", ssd_v_size), PK_D("T_SHIFT", t_shift),
    PK_D("U_SHIFT", u_shift), PK_D("V_SHIFT", v_shift), PK_D("T_BIN_SIZE", t_bin_size),
    PK_D("V_BIN_SIZE", v_bin_size), PK_I("SIGMAS_2_KEEP", sigmas_2_keep), PK_D("RECON_CYL_RADIUS", recon_cyl_radius),
    PK_I("COLUMNS", columns), PK_I("ROWS", rows), PK_D("SLICE_THICKNESS", slice_thickness),
    PK_I("MSC_DIFF_THRESH", msc_diff_thresh), PK_D("SC_LOWER_THRESHOLD", sc_lower_threshold),
    PK_D("SC_UPPER_THRESHOLD", sc_upper_threshold), PK_D("MSC_LOWER_THRESHOLD", msc_lower_threshold),
    PK_D("MSC_UPPER_THRESHOLD", msc_upper_threshold), PK_D("SM_LOWER_THRESHOLD", sm_lower_threshold),
    PK_D("SM_UPPER_THRESHOLD", sm_upper_threshold), PK_D("SM_SCALE_THRESHOLD", sm_scale_threshold),
    PK_B("SC_ON", sc_on), PK_B("MSC_ON", msc_on), PK_B("SM_ON", sm_on), PK_B("FBP_ON", fbp_on),
    PK_I("HULL_TYPE", hull_type), PK_I("FBP_FILTER_TYPE", fbp_filter_type), PK_I("X_0_TYPE", x_0_type),
    PK_D("FBP_THRESHOLD", fbp_threshold), PK_D("LAMBDA", lambda), PK_I("ITERATIONS", iterations),
    PK_I("DROP_BLOCK_SIZE", drop_block_size), PK_I("HISTORIES_PER_BLOCK", histories_per_block),
    PK_I("MIN_MLP_LENGTH", min_mlp_length), PK_D("TRIG_TABLE_MIN_PIS", trig_table_min_pis),
    PK_D("TRIG_TABLE_MAX_PIS", trig_table_max_pis), PK_D("TRIG_TABLE_STEP_DEGREES", trig_table_step_degrees),
    PK_D("DEPTH_TABLE_RANGE", coeff_table_range), PK_D("DEPTH_TABLE_STEP", coeff_table_step),
    PK_D("POLY_TABLE_RANGE", poly_table_range), PK_D("POLY_TABLE_STEP", poly_table_step),
    PK_D("SOURCE_RADIUS", source_radius), PK_D("RSP_AIR", rsp_air), PK_B("DATA_IN_MM", data_in_mm),
    PK_B("QUIRK_A1_BIN0", quirk_a1_bin0), PK_B("QUIRK_A3_DROP_FLOOR", quirk_a3_drop_floor),
    PK_B("QUIRK_VOXEL_DECODE", quirk_voxel_decode), PK_I("DROP_ACCUM", drop_accum),
    PK_B("MEDIAN_FILTER_FBP", median_filter_fbp), PK_I("FBP_MED_FILTER_RADIUS", fbp_med_filter_radius),
    PK_B("AVG_FILTER_FBP", avg_filter_fbp), PK_B("AVG_FILTER_HULL", avg_filter_hull),
    PK_I("HULL_AVG_FILTER_RADIUS", hull_avg_filter_radius), PK_I("MLP_ALGORITHM", mlp_algorithm),
    PK_D("HULL_AVG_FILTER_THRESHOLD", hull_avg_filter_threshold),
    PK_B("TVS_ON", tvs_on), PK_B("TVS_FIRST", tvs_first), PK_B("TVS_CONDITIONED", tvs_conditioned),
    PK_B("TVS_OLD", tvs_old), PK_I("TVS_REPETITIONS", tvs_repetitions), PK_I("L_0", tvs_l0),
    PK_I("TVS_SEED", tvs_seed), PK_D("A", tvs_a), PK_D("TVS_MIN_ETA", tvs_min_eta),
    PK_D("TV_THRESHOLD", tv_threshold), PK_I("LIST_SEGMENT_HISTORIES", list_segment_histories),
};

// Keys of settings.cfg that configure parts of the reference outside this path (I/O directories,
// filters, output switches, transfer modes).  They are accepted and ignored so that an unmodified
// settings.cfg loads.
const char* kIgnored[] = {
    "PROJECTION_DATA_DIR", "PREPROCESSING_DIR", "RECONSTRUCTION_DIR", "PATH_2_PCT_DATA_DIR", "OBJECT", "RUN_DATE",
    "RUN_NUMBER", "PROJECTION_DATA_DATE", "PREPROCESS_DATE", "USER_NAME", "DATA_TYPE", "RECONSTRUCTION_METHOD",
    "ENDPOINTS_TX_MODE", "DROP_TX_MODE", "ENDPOINTS_ALG", "NUM_SCANS", "T_BINS", "V_BINS",
    "RECON_CYL_HEIGHT", "SLICES", "IMAGE_WIDTH", "IMAGE_HEIGHT", "IMAGE_THICKNESS", "VOXEL_WIDTH", "VOXEL_HEIGHT",
    "VOXEL_THICKNESS", "ETA", "PSI_SIGN", "THREADS_PER_BLOCK", "ENDPOINTS_PER_BLOCK", "ENDPOINTS_PER_THREAD",
    "HISTORIES_PER_THREAD", "VOXELS_PER_THREAD", "MAX_GPU_HISTORIES", "MAX_CUTS_HISTORIES", "MAX_ENDPOINTS_HISTORIES",
    "HULL_RSP_THRESHOLD", "EXPORT_PREPROCESSING", "IMPORT_PREPROCESSING", "PERFORM_RECONSTRUCTION",
    "PREPROCESS_OVERWRITE_OK", "RECON_OVERWRITE_OK", "MLP_IN_LOOP", "IMPORT_DATA_ITERATIVELY",
    "IMPORT_MLP_ITERATIVELY", "TOTAL_DATA_2_GPU", "TOTAL_MLP_2_GPU", "MEDIAN_FILTER_HULL",
    "MEDIAN_FILTER_X_0", "MEDIAN_FILTER_X_K", "MEDIAN_FILTER_X",
    "AVG_FILTER_X_0", "AVG_FILTER_X_K", "AVG_FILTER_X", "HULL_MED_FILTER_RADIUS",
    "X_0_MED_FILTER_RADIUS", "X_K_MED_FILTER_RADIUS", "X_MED_FILTER_RADIUS",
    "FBP_AVG_FILTER_RADIUS", "X_0_AVG_FILTER_RADIUS", "X_K_AVG_FILTER_RADIUS", "X_AVG_FILTER_RADIUS",
    "WRITE_MSC_COUNTS", "WRITE_SM_COUNTS", "WRITE_X_FBP", "WRITE_FBP_HULL", "WRITE_AVG_FBP", "WRITE_MEDIAN_FBP",
    "WRITE_BIN_WEPLS", "WRITE_WEPL_DISTS", "WRITE_SSD_ANGLES", "STDOUT_2_DISK", "USER_INPUT_REQUESTS_OFF",
    "ADD_DATA_LOG_ENTRY", "DEBUG_TEXT_ON", "EXIT_AFTER_BINNING", "EXIT_AFTER_HULLS", "EXIT_AFTER_CUTS",
    "EXIT_AFTER_SINOGRAM", "EXIT_AFTER_FBP",
};

bool parse_value(const char* value, Kind kind, double* d, int* i)
{
    if (std::strcmp(value, "true") == 0 || std::strcmp(value, "false") == 0) {
        *i = std::strcmp(value, "true") == 0;
        *d = *i;
        return true;
    }
    if (std::strchr(value, '"')) return false;  // strings are not numeric parameters
    char* end = nullptr;
    if (std::strchr(value, '.') == nullptr) {  // integer (V.cu:9318)
        long v = std::strtol(value, &end, 10);
        if (end == value) return false;
        *i = (int)v;
        *d = (double)*i;
    } else {  // double (V.cu:9329); its integer reading truncates
        double v = std::strtod(value, &end);
        if (end == value) return false;
        *d = v;
        *i = (int)v;
    }
    (void)kind;
    return true;
}

}  // namespace

extern "C" int pct_params_set(pct_params* p, const char* key, const char* value)
{
    for (const Key& k : kKeys) {
        if (std::strcmp(k.name, key) != 0) continue;
        double d;
        int i;
        if (!parse_value(value, k.kind, &d, &i)) return PCT_E_ARG;
        char* base = reinterpret_cast<char*>(p) + k.offset;
        if (k.kind == K_DOUBLE) *reinterpret_cast<double*>(base) = d;
        else *reinterpret_cast<int32_t*>(base) = i;
        return PCT_OK;
    }
    for (const char* ign : kIgnored)
        if (std::strcmp(ign, key) == 0) return PCT_OK;
    return PCT_E_ARG;
}

extern "C" int pct_params_load_cfg(pct_params* p, const char* path, char* err, size_t err_len)
{
    FILE* f = std::fopen(path, "r");
    if (!f) {
        if (err) std::snprintf(err, err_len, "cannot open %s", path);
        return PCT_E_IO;
    }
    char line[2048];
    int lineno = 0;
    while (std::fgets(line, sizeof(line), f)) {
        lineno++;
        char* s = line;
        while (*s == ' ' || *s == '\t') s++;
        if (s[0] == '\0' || s[0] == '\n' || s[0] == '\r' || (s[0] == '/' && s[1] == '/')) continue;
        char key[256] = {0}, eq[16] = {0}, value[1024] = {0};
        if (std::sscanf(s, " %255s %15s %1023s", key, eq, value) < 3 || std::strcmp(eq, "=") != 0) {
            if (err) std::snprintf(err, err_len, "%s:%d: expected KEY = value", path, lineno);
            std::fclose(f);
            return PCT_E_IO;
        }
        if (value[0] == '"') continue;  // string values (paths, names) configure I/O only
        int rc = pct_params_set(p, key, value);
        if (rc != PCT_OK) {
            if (err) std::snprintf(err, err_len, "%s:%d: unknown key or bad value: %s = %s", path, lineno, key, value);
            std::fclose(f);
            return rc;
        }
    }
    std::fclose(f);
    return PCT_OK;
}

extern "C" int pct_params_derive(pct_params* p, char* err, size_t err_len)
{
    auto fail = [&](const char* msg) {
        if (err) std::snprintf(err, err_len, "%s", msg);
        return PCT_E_ARG;
    };
    if (p->t_bin_size <= 0 || p->v_bin_size <= 0 || p->angular_bin_size <= 0 || p->slice_thickness <= 0)
        return fail("bin sizes and slice thickness must be positive");
    if (p->columns <= 0 || p->rows <= 0 || p->recon_cyl_radius <= 0) return fail("image size must be positive");
    if (p->ssd_v_size <= 1.0) return fail("SSD_V_SIZE must exceed 1 cm (RECON_CYL_HEIGHT = SSD_V_SIZE - 1)");
    p->pi = 4 * std::atan(1.0);
    p->t_bins = (int32_t)(p->ssd_t_size / p->t_bin_size + 0.5);
    p->v_bins = (int32_t)(p->ssd_v_size / p->v_bin_size + 0.5);
    p->angular_bins = (int32_t)(360 / p->angular_bin_size + 0.5);
    double nb = (double)p->angular_bins * p->t_bins * p->v_bins;
    if (nb > 2147483647.0) return fail("NUM_BINS exceeds int32");
    p->num_bins = p->angular_bins * p->t_bins * p->v_bins;
    p->recon_cyl_height = p->ssd_v_size - 1.0;
    p->slices = (int32_t)(p->recon_cyl_height / p->slice_thickness);
    if (p->slices <= 0) return fail("SLICES must be positive");
    double nv = (double)p->columns * p->rows * p->slices;
    if (nv >= 4294967295.0 || nv > 2147483647.0) return fail("NUM_VOXELS exceeds int32");
    p->num_voxels = p->columns * p->rows * p->slices;
    p->image_width = 2 * p->recon_cyl_radius;
    p->image_height = 2 * p->recon_cyl_radius;
    p->image_thickness = p->recon_cyl_height;
    p->voxel_width = p->image_width / p->columns;
    p->voxel_height = p->image_height / p->rows;
    p->voxel_thickness = p->image_thickness / p->slices;
    p->x_zero_coordinate = -p->recon_cyl_radius;
    p->y_zero_coordinate = p->recon_cyl_radius;
    p->z_zero_coordinate = p->recon_cyl_height / 2;
    p->ram_lak_tau = (double)(2 / sqrtf(2.0f)) * p->t_bin_size;
    p->mlp_u_step = p->voxel_width / 2;
    p->trig_table_min = p->trig_table_min_pis * p->pi;
    p->trig_table_max = p->trig_table_max_pis * p->pi;
    p->trig_table_step = p->trig_table_step_degrees * (p->pi / 180.0);
    p->trig_table_elements = (int32_t)((p->trig_table_max - p->trig_table_min) / p->trig_table_step + 0.5);
    p->coeff_table_elements = (int32_t)(p->coeff_table_range / p->coeff_table_step + 0.5);
    p->poly_table_elements = (int32_t)(p->poly_table_range / p->poly_table_step + 0.5);
    if (p->coeff_table_range < 2 * p->recon_cyl_radius || p->poly_table_range < 2 * p->recon_cyl_radius)
        return fail("MLP depth tables must cover 2 * RECON_CYL_RADIUS");
    if (p->trig_table_min > -p->pi || p->trig_table_max < 4 * p->pi)
        return fail("trig table must cover [-pi, 4pi] (entry angle, +pi when u_in > u_out)");
    if (p->drop_block_size <= 0 || p->histories_per_block <= 0) return fail("DROP sizes must be positive");
    if (p->drop_accum < PCT_ACCUM_FP32 || p->drop_accum > PCT_ACCUM_KAHAN32) return fail("DROP_ACCUM must be 0, 1 or 2");
    if (p->hull_type < PCT_HULL_IMPORT || p->hull_type > PCT_HULL_FBP) return fail("HULL_TYPE out of range");
    if (p->x_0_type < PCT_X0_IMPORT || p->x_0_type > PCT_X0_ZEROS) return fail("X_0_TYPE out of range");
    if (p->fbp_filter_type < 0 || p->fbp_filter_type > 2) return fail("FBP_FILTER_TYPE out of range");
    if (p->mlp_algorithm != PCT_MLP_STANDARD && p->mlp_algorithm != PCT_MLP_TABULATED)
        return fail("MLP_ALGORITHM out of range");
    if (p->list_segment_histories < 0) return fail("LIST_SEGMENT_HISTORIES must be >= 0");
    if (p->fbp_med_filter_radius < 0 || p->fbp_med_filter_radius > 10) return fail("FBP_MED_FILTER_RADIUS must be 0..10");
    if (p->hull_avg_filter_radius < 0 || p->hull_avg_filter_radius > 64)
        return fail("HULL_AVG_FILTER_RADIUS must be 0..64");
    if (p->tvs_repetitions < 0 || p->tvs_repetitions > 10000) return fail("TVS_REPETITIONS must be 0..10000");
    if (p->tvs_l0 < 0 || p->tvs_l0 > 300) return fail("L_0 must be 0..300");
    if (!(p->tvs_a > 0 && p->tvs_a < 1)) return fail("A must lie in (0, 1)");
    if (!(p->tvs_min_eta > 0 && p->tvs_min_eta < 1)) return fail("TVS_MIN_ETA must lie in (0, 1)");
    return PCT_OK;
}
