// context.hip — the C-ABI: context lifetime, data loading, stage sequencing, outputs.
//
// Stage order is the reference's (preprocessing, DS.cu:7394-7403; image_reconstruction :10650-10685):
//   import_and_process_data   -> pct_intersect_bin_carve
//   hull_detection_finish     -> pct_hull_finish
//   statistical_calculations_and_cuts -> pct_stats_and_cuts
//   generate_preprocessing_images (construct_sinogram, FBP, hull_selection, define_initial_iterate)
//                             -> pct_fbp, pct_select_hull, pct_initial_iterate
//   reconstruction_cuts       -> pct_mlp_endpoints
//   DROP_GPU                  -> pct_mlp_build_lists + pct_drop_iterate
// Every stage runs on the context's stream; the host only synchronises to learn compaction counts.
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <cstring>

#include "internal.h"

namespace {

constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n, int block = kBlock) { return pct_grid_for(n, block); }

enum TimerSlot { TM_INTERSECT = 0, TM_HULLS, TM_CUTS, TM_FBP, TM_HULL_SEL, TM_X0, TM_ENDPOINTS, TM_LISTS, TM_DROP };
const char* kTimerName[] = {"intersect_bin_carve", "hull_finish", "stats_and_cuts", "fbp", "select_hull",
                            "initial_iterate", "mlp_endpoints", "mlp_build_lists", "drop"};

// define_initial_iterate, DS.cu:8612-8646 (HYBRID: initial_iterate_generate_hybrid :8582)
__global__ void k_x0(int64_t nv, int kind, const uint8_t* hull, const float* fbp, float* x0)
{
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    float r = 0.0f;
    if (kind == PCT_X0_HULL) r = (float)hull[v];
    else if (kind == PCT_X0_FBP) r = fbp[v];
    else if (kind == PCT_X0_HYBRID) r = hull[v] ? fbp[v] : 0.0f;
    x0[v] = r;
}

template <class T>
__global__ void k_to_float(int64_t n, const T* in, float* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// A new iterate restarts the DROP iteration count and the TVS exponent L (a global starting at 0 in
// the reference, DS.h:439).
void restart_iterations(pct_ctx* ctx)
{
    ctx->drop_k = 0;
    ctx->tvs_L = 0;
    ctx->tvs_log.clear();
}

struct StageTimer {
    pct_ctx* ctx;
    int slot;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(pct_ctx* c, int s) : ctx(c), slot(s)
    {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, ctx->stream);
    }
    void done()
    {
        (void)hipEventRecord(b, ctx->stream);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        ctx->t_stage[slot] = ms;
    }
    ~StageTimer()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

int require(pct_ctx* ctx, int stage, const char* what)
{
    if (ctx->stage < stage) {
        ctx->err = std::string(what) + ": previous stage has not run";
        return PCT_E_STATE;
    }
    return PCT_OK;
}

}  // namespace

extern "C" {

int pct_ctx_create(const pct_params* p, int device, pct_ctx** out)
{
    if (!p || !out) return PCT_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PCT_E_HIP;
    if (device < 0 || device >= ndev) return PCT_E_ARG;
    pct_params q = *p;
    char err[256];
    if (pct_params_derive(&q, err, sizeof(err)) != PCT_OK) return PCT_E_ARG;
    pct_ctx* ctx = new pct_ctx();
    ctx->P = q;
    ctx->device = device;
    // launch shapes of the DROP kernels (tuning; the arithmetic does not depend on them)
    if (const char* e = std::getenv("PCT_DROP_ROW_LANES")) {
        int g = std::atoi(e);
        if (g == 4 || g == 8 || g == 16 || g == 32 || g == 64) ctx->drop_row_lanes = g;
    }
    if (const char* e = std::getenv("PCT_DROP_COL_VOXELS")) {
        int v = std::atoi(e);
        if (v == 1 || v == 2 || v == 4 || v == 8) ctx->drop_col_voxels = v;
    }
    if (const char* e = std::getenv("PCT_DROP_FLAGS")) {
        int f = std::atoi(e);
        if (f >= 0 && f <= 63) ctx->drop_flags = f;
    }
    // hull carving: LDS hash slots of the fused walk, or 0 for one walk per method (the counts are the same)
    if (const char* e = std::getenv("PCT_CARVE_TABLE")) {
        int t = std::atoi(e);
        if (t == 0 || t == 4096 || t == 8192 || t == 16384) ctx->carve_table = t;
    }
    if (const char* e = std::getenv("PCT_CARVE_BLOCK")) {
        int t = std::atoi(e);
        if (t == 256 || t == 512) ctx->carve_block = t;
    }
    if (const char* e = std::getenv("PCT_ENDPOINT_ORDER")) ctx->endpoint_order = std::atoi(e) != 0;
    if (const char* e = std::getenv("PCT_CARVE_LINES")) ctx->carve_line_order = std::atoi(e) != 0;
    if (const char* e = std::getenv("PCT_CARVE_SPILL")) {
        int t = std::atoi(e);
        if (t >= 1 && t <= 512) ctx->carve_spill = t;
    }
    if (const char* e = std::getenv("PCT_DROP_PIPELINE")) ctx->drop_pipeline = std::atoi(e) != 0;
    if (const char* e = std::getenv("PCT_MLP_BLOCKS_PER_CU")) {
        int b = std::atoi(e);
        if (b >= -1 && b <= 64) ctx->mlp_blocks_per_cu = b;
    }
    if (const char* e = std::getenv("PCT_DROP_GROUP_ENTRIES")) {
        long long g = std::atoll(e);
        if (g >= 1 && g <= (1ll << 30)) ctx->drop_group_entries = g;
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PCT_E_HIP;
    }
    *out = ctx;
    return PCT_OK;
}

int pct_ctx_destroy(pct_ctx* ctx)
{
    if (!ctx) return PCT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& r : ctx->raw) pct_free(&r);
    for (int k = 0; k < 10; k++) {
        pct_free(&ctx->h[k]);
        pct_free(&ctx->h_tmp[k]);
    }
    pct_free(&ctx->wepl); pct_free(&ctx->wepl_tmp); pct_free(&ctx->bin_num); pct_free(&ctx->bin_tmp);
    pct_free(&ctx->gantry); pct_free(&ctx->gantry_tmp);
    pct_free(&ctx->in_range); pct_free(&ctx->u8_tmp); pct_free(&ctx->flag); pct_free(&ctx->pos);
    pct_free(&ctx->first_voxel); pct_free(&ctx->u32_a); pct_free(&ctx->u32_b); pct_free(&ctx->u32_c);
    pct_free(&ctx->u32_d); pct_free(&ctx->cub_tmp);
    pct_free(&ctx->counts_pre); pct_free(&ctx->counts_post); pct_free(&ctx->bin_off);
    for (int k = 0; k < 3; k++) {
        pct_free(&ctx->mean[k]);
        pct_free(&ctx->stddev[k]);
    }
    pct_free(&ctx->sinogram); pct_free(&ctx->sino_filtered);
    pct_free(&ctx->sc_carved); pct_free(&ctx->sc_hull); pct_free(&ctx->msc_counts); pct_free(&ctx->sm_counts);
    pct_free(&ctx->msc_hull); pct_free(&ctx->sm_hull); pct_free(&ctx->fbp_hull); pct_free(&ctx->hull);
    pct_free(&ctx->hull_import); pct_free(&ctx->fbp); pct_free(&ctx->x0); pct_free(&ctx->x);
    pct_free(&ctx->fbp_x0); pct_free(&ctx->hull_endpoints);
    pct_free(&ctx->list_off); pct_free(&ctx->list_cnt); pct_free(&ctx->list_vox); pct_free(&ctx->chord);
    pct_free(&ctx->cap_scratch);
    pct_free(&ctx->res_r2);
    pct_free(&ctx->T.sin_t); pct_free(&ctx->T.cos_t); pct_free(&ctx->T.scat); pct_free(&ctx->T.p12);
    pct_free(&ctx->T.p26); pct_free(&ctx->T.p312); pct_free(&ctx->T.step_terms);
    pct_free(&ctx->accum64); pct_free(&ctx->accum32); pct_free(&ctx->S); pct_free(&ctx->touched);
    pct_free(&ctx->n_touched); pct_free(&ctx->list_overflow);
    pct_drop_plan_free(&ctx->plan);
    pct_free(&ctx->tvs_G); pct_free(&ctx->tvs_part); pct_free(&ctx->tvs_res);
    if (ctx->tvs_res_h) (void)hipHostFree(ctx->tvs_res_h);
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
    return PCT_OK;
}

const char* pct_last_error(const pct_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pct_ctx_synchronize(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PCT_OK;
}

int pct_get_count(pct_ctx* ctx, int kind, int64_t* out)
{
    if (!ctx || !out) return PCT_E_ARG;
    switch (kind) {
    case PCT_N_LOADED: *out = ctx->n_loaded; break;
    case PCT_N_RECON_VOLUME: *out = ctx->n_recon_volume; break;
    case PCT_N_POST_CUTS: *out = ctx->n_post_cuts; break;
    case PCT_N_RECONSTRUCTION: *out = ctx->n_recon; break;
    case PCT_N_CURRENT: *out = ctx->n_hist; break;
    case PCT_N_LIST_ENTRIES: {
        if (ctx->stage < ST_LISTS || ctx->n_recon == 0) { *out = 0; break; }
        if (pct_segment_histories(ctx) > 0) {  // lists of the segments of the last DROP iteration
            *out = ctx->seg.entries;
            break;
        }
        // sum of the per-history counts
        std::vector<uint32_t> c(ctx->n_recon);
        PCT_CHECK_HIP(ctx, hipMemcpy(c.data(), ctx->list_cnt, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        int64_t s = 0;
        for (uint32_t v : c) s += v;
        *out = s;
        break;
    }
    default: return PCT_E_ARG;
    }
    return PCT_OK;
}

int pct_load_histories(pct_ctx* ctx, size_t n, const float* const t[4], const float* const v[4],
                       const float* const u[4], const float* wepl, const float* angle_deg)
{
    if (!ctx || !t || !v || !u || !wepl || !angle_deg) return PCT_E_ARG;
    for (int k = 0; k < 4; k++)
        if (!t[k] || !v[k] || !u[k]) return PCT_E_ARG;
    (void)hipSetDevice(ctx->device);
    const int64_t total = ctx->n_loaded + (int64_t)n;
    if (total > 0x7fffffffll) {
        ctx->err = "more than 2^31 histories in one context";
        return PCT_E_LIMIT;
    }
    if (total > ctx->cap_loaded) {
        int64_t cap = total > 2 * ctx->cap_loaded ? total : 2 * ctx->cap_loaded;
        for (int k = 0; k < 14; k++) {
            pct_tmp<float> nb;
            PCT_TRY(pct_alloc(ctx, &nb.p, cap));
            if (ctx->raw[k] && ctx->n_loaded)
                PCT_CHECK_HIP(ctx, hipMemcpyAsync(nb.p, ctx->raw[k], ctx->n_loaded * sizeof(float),
                                                  hipMemcpyDeviceToDevice, ctx->stream));
            PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
            pct_free(&ctx->raw[k]);
            ctx->raw[k] = nb.p;
            nb.p = nullptr;  // owned by the context from here on
        }
        ctx->cap_loaded = cap;
    }
    const float* src[14] = {t[0], t[1], t[2], t[3], v[0], v[1], v[2], v[3], u[0], u[1], u[2], u[3], wepl, angle_deg};
    for (int k = 0; k < 14; k++)
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->raw[k] + ctx->n_loaded, src[k], n * sizeof(float),
                                          hipMemcpyHostToDevice, ctx->stream));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    extern int pct_convert_units(pct_ctx * ctx, int64_t first, int64_t n);
    PCT_TRY(pct_convert_units(ctx, ctx->n_loaded, (int64_t)n));
    ctx->n_loaded = total;
    ctx->stage = ST_LOADED;
    return PCT_OK;
}

int pct_intersect_bin_carve(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_LOADED, "pct_intersect_bin_carve"));
    (void)hipSetDevice(ctx->device);
    StageTimer tm(ctx, TM_INTERSECT);
    PCT_TRY(pct_stage_intersect_bin_carve(ctx));
    tm.done();
    return PCT_OK;
}

int pct_hull_finish(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_BINNED, "pct_hull_finish"));
    StageTimer tm(ctx, TM_HULLS);
    PCT_TRY(pct_stage_hull_finish(ctx));
    tm.done();
    return PCT_OK;
}

int pct_stats_and_cuts(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_HULLS, "pct_stats_and_cuts"));
    StageTimer tm(ctx, TM_CUTS);
    PCT_TRY(pct_stage_stats_and_cuts(ctx));
    tm.done();
    return PCT_OK;
}

int pct_fbp(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_CUT, "pct_fbp"));
    StageTimer tm(ctx, TM_FBP);
    PCT_TRY(pct_stage_fbp(ctx));
    tm.done();
    return PCT_OK;
}

int pct_select_hull(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_FBP, "pct_select_hull"));
    StageTimer tm(ctx, TM_HULL_SEL);
    PCT_TRY(pct_stage_select_hull(ctx));
    tm.done();
    return PCT_OK;
}

int pct_initial_iterate(pct_ctx* ctx, const float* x0_import)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_HULL_SEL, "pct_initial_iterate"));
    const pct_params& P = ctx->P;
    const int64_t NV = P.num_voxels;
    StageTimer tm(ctx, TM_X0);
    PCT_TRY(pct_alloc(ctx, &ctx->x0, NV));
    PCT_TRY(pct_alloc(ctx, &ctx->x, NV));
    if (P.x_0_type == PCT_X0_IMPORT) {
        if (!x0_import) {
            ctx->err = "X_0_TYPE = IMPORT needs an image";
            return PCT_E_ARG;
        }
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->x0, x0_import, NV * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    } else {
        if ((P.x_0_type == PCT_X0_FBP || P.x_0_type == PCT_X0_HYBRID) && !ctx->fbp) {
            ctx->err = "X_0_TYPE = FBP/HYBRID requires FBP_ON";
            return PCT_E_STATE;
        }
        hipLaunchKernelGGL(k_x0, dim3(grid_for(NV)), dim3(kBlock), 0, ctx->stream, NV, P.x_0_type, ctx->hull,
                           ctx->fbp_x0 ? ctx->fbp_x0 : ctx->fbp, ctx->x0);
    }
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->x, ctx->x0, NV * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    PCT_CHECK_HIP(ctx, hipGetLastError());
    tm.done();
    ctx->stage = ST_X0;
    restart_iterations(ctx);
    return PCT_OK;
}

int pct_mlp_endpoints(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_X0, "pct_mlp_endpoints"));
    StageTimer tm(ctx, TM_ENDPOINTS);
    PCT_TRY(pct_stage_endpoints(ctx));
    tm.done();
    return PCT_OK;
}

int pct_mlp_build_lists(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_ENDPOINTS, "pct_mlp_build_lists"));
    StageTimer tm(ctx, TM_LISTS);
    PCT_TRY(pct_stage_build_lists(ctx));
    tm.done();
    return PCT_OK;
}

int pct_drop_iterate(pct_ctx* ctx, int iterations)
{
    if (!ctx || iterations < 0) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_LISTS, "pct_drop_iterate"));
    StageTimer tm(ctx, TM_DROP);
    PCT_TRY(pct_stage_drop(ctx, iterations));
    tm.done();
    return PCT_OK;
}

int pct_run_pipeline(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(pct_intersect_bin_carve(ctx));
    PCT_TRY(pct_hull_finish(ctx));
    PCT_TRY(pct_stats_and_cuts(ctx));
    PCT_TRY(pct_fbp(ctx));
    PCT_TRY(pct_select_hull(ctx));
    PCT_TRY(pct_initial_iterate(ctx, nullptr));
    PCT_TRY(pct_mlp_endpoints(ctx));
    PCT_TRY(pct_mlp_build_lists(ctx));
    PCT_TRY(pct_drop_iterate(ctx, ctx->P.iterations));
    return PCT_OK;
}

int pct_get_image(pct_ctx* ctx, int kind, float* out)
{
    if (!ctx || !out) return PCT_E_ARG;
    const int64_t NV = ctx->P.num_voxels;
    hipStream_t s = ctx->stream;
    pct_tmp<float> tmp;
    const void* src = nullptr;
    int type = 0;  // 0 float, 1 u8, 2 i32
    switch (kind) {
    case PCT_IMG_X: src = ctx->x; break;
    case PCT_IMG_FBP: src = ctx->fbp; break;
    case PCT_IMG_X0: src = ctx->x0; break;
    case PCT_IMG_HULL: src = ctx->hull; type = 1; break;
    case PCT_IMG_SC_HULL: src = ctx->sc_hull; type = 1; break;
    case PCT_IMG_MSC_COUNTS: src = ctx->msc_counts; type = 2; break;
    case PCT_IMG_SM_COUNTS: src = ctx->sm_counts; type = 2; break;
    case PCT_IMG_MSC_HULL: src = ctx->msc_hull; type = 2; break;
    case PCT_IMG_SM_HULL: src = ctx->sm_hull; type = 2; break;
    case PCT_IMG_FBP_HULL: src = ctx->fbp_hull; type = 1; break;
    case PCT_IMG_FBP_MEDIAN: src = ctx->fbp_x0; break;
    case PCT_IMG_HULL_FILTERED: src = ctx->hull_endpoints; type = 1; break;
    default: return PCT_E_ARG;
    }
    if (!src) {
        ctx->err = "image not available at this stage";
        return PCT_E_STATE;
    }
    if (type == 0) {
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(out, src, NV * sizeof(float), hipMemcpyDeviceToHost, s));
    } else {
        PCT_TRY(pct_alloc(ctx, &tmp.p, NV));
        if (type == 1)
            hipLaunchKernelGGL(k_to_float<uint8_t>, dim3(grid_for(NV)), dim3(kBlock), 0, s, NV, (const uint8_t*)src, tmp.p);
        else
            hipLaunchKernelGGL(k_to_float<int32_t>, dim3(grid_for(NV)), dim3(kBlock), 0, s, NV, (const int32_t*)src, tmp.p);
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(out, tmp.p, NV * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    return PCT_OK;
}

int pct_set_image(pct_ctx* ctx, const float* x)
{
    if (!ctx || !x) return PCT_E_ARG;
    if (!ctx->x) PCT_TRY(pct_alloc(ctx, &ctx->x, ctx->P.num_voxels));
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->x, x, ctx->P.num_voxels * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    restart_iterations(ctx);
    return PCT_OK;
}

int pct_get_bins(pct_ctx* ctx, int kind, float* out)
{
    if (!ctx || !out) return PCT_E_ARG;
    const int NB = ctx->P.num_bins;
    hipStream_t s = ctx->stream;
    const void* src = nullptr;
    bool is_int = false;
    switch (kind) {
    case PCT_BIN_COUNTS_PRE: src = ctx->counts_pre; is_int = true; break;
    case PCT_BIN_MEAN_WEPL: src = ctx->mean[0]; break;
    case PCT_BIN_MEAN_UT: src = ctx->mean[1]; break;
    case PCT_BIN_MEAN_UV: src = ctx->mean[2]; break;
    case PCT_BIN_STD_WEPL: src = ctx->stddev[0]; break;
    case PCT_BIN_STD_UT: src = ctx->stddev[1]; break;
    case PCT_BIN_STD_UV: src = ctx->stddev[2]; break;
    case PCT_BIN_COUNTS_POST: src = ctx->counts_post; is_int = true; break;
    case PCT_BIN_SINOGRAM: src = ctx->sinogram; break;
    case PCT_BIN_SINOGRAM_FILTERED: src = ctx->sino_filtered; break;
    default: return PCT_E_ARG;
    }
    if (!src) {
        ctx->err = "bins not available at this stage";
        return PCT_E_STATE;
    }
    if (is_int) {
        pct_tmp<float> tmp;
        PCT_TRY(pct_alloc(ctx, &tmp.p, NB));
        hipLaunchKernelGGL(k_to_float<int32_t>, dim3(grid_for(NB)), dim3(kBlock), 0, s, NB, (const int32_t*)src, tmp.p);
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(out, tmp.p, NB * sizeof(float), hipMemcpyDeviceToHost, s));
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    } else {
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(out, src, NB * sizeof(float), hipMemcpyDeviceToHost, s));
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    }
    return PCT_OK;
}

int pct_get_histories(pct_ctx* ctx, int kind, void* out, size_t capacity)
{
    if (!ctx || !out) return PCT_E_ARG;
    const int64_t n = ctx->n_hist;
    if ((size_t)n > capacity) return PCT_E_ARG;
    const void* src = nullptr;
    size_t esz = 4;
    if (kind >= PCT_H_X_ENTRY && kind <= PCT_H_XZ_EXIT_ANGLE) src = ctx->h[kind];
    else if (kind == PCT_H_WEPL) src = ctx->wepl;
    else if (kind == PCT_H_GANTRY_ANGLE) src = ctx->gantry;
    else if (kind == PCT_H_BIN_NUM) src = ctx->bin_num;
    else if (kind == PCT_H_FIRST_VOXEL) src = ctx->first_voxel;
    else if ((kind == PCT_H_PATH_COUNT || kind == PCT_H_CHORD) && pct_segment_histories(ctx) > 0) {
        ctx->err = "path counts and chords are built per segment (LIST_SEGMENT_HISTORIES) and not kept";
        return PCT_E_STATE;
    } else if (kind == PCT_H_PATH_COUNT) src = ctx->stage >= ST_LISTS ? ctx->list_cnt : nullptr;
    else if (kind == PCT_H_CHORD) {
        src = ctx->stage >= ST_LISTS ? ctx->chord : nullptr;
        esz = 8;
    } else return PCT_E_ARG;
    if (!src) {
        ctx->err = "history array not available at this stage";
        return PCT_E_STATE;
    }
    if (n > 0) PCT_CHECK_HIP(ctx, hipMemcpy(out, src, n * esz, hipMemcpyDeviceToHost));
    return PCT_OK;
}

int pct_get_lists(pct_ctx* ctx, int64_t* offsets, size_t offsets_cap, uint32_t* voxels, size_t voxels_cap)
{
    if (!ctx || !offsets) return PCT_E_ARG;
    if (ctx->stage < ST_LISTS) {
        ctx->err = "lists not built";
        return PCT_E_STATE;
    }
    if (pct_segment_histories(ctx) > 0) {
        ctx->err = "lists are built per segment (LIST_SEGMENT_HISTORIES) and not kept";
        return PCT_E_STATE;
    }
    const int64_t n = ctx->n_recon;
    if (offsets_cap < (size_t)n + 1) return PCT_E_ARG;
    std::vector<int64_t> off(n + 1);
    std::vector<uint32_t> cnt(n);
    if (n > 0) {
        PCT_CHECK_HIP(ctx, hipMemcpy(off.data(), ctx->list_off, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        PCT_CHECK_HIP(ctx, hipMemcpy(cnt.data(), ctx->list_cnt, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) total += cnt[i];
    offsets[0] = 0;
    if (!voxels) {  // sizes only
        int64_t acc = 0;
        for (int64_t i = 0; i < n; i++) {
            acc += cnt[i];
            offsets[i + 1] = acc;
        }
        return PCT_OK;
    }
    if (voxels_cap < (size_t)total) return PCT_E_ARG;
    std::vector<uint32_t> all(off.empty() ? 0 : off[n]);
    if (!all.empty()) PCT_CHECK_HIP(ctx, hipMemcpy(all.data(), ctx->list_vox, all.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    int64_t acc = 0;
    for (int64_t i = 0; i < n; i++) {
        std::memcpy(voxels + acc, all.data() + off[i], cnt[i] * sizeof(uint32_t));
        acc += cnt[i];
        offsets[i + 1] = acc;
    }
    return PCT_OK;
}

int pct_residual(pct_ctx* ctx, double* norm)
{
    if (!ctx || !norm) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_LISTS, "pct_residual"));
    return pct_stage_residual(ctx, norm);
}

const char* pct_metrics_json(pct_ctx* ctx)
{
    if (!ctx) return "{}";
    char buf[256];
    std::string s = "{\"counts\": {";
    std::snprintf(buf, sizeof(buf), "\"loaded\": %lld, \"recon_volume\": %lld, \"post_cuts\": %lld, \"reconstruction\": %lld",
                  (long long)ctx->n_loaded, (long long)ctx->n_recon_volume, (long long)ctx->n_post_cuts,
                  (long long)ctx->n_recon);
    s += buf;
    size_t mem_free = 0, mem_total = 0;
    (void)hipMemGetInfo(&mem_free, &mem_total);
    const bool segmented = ctx->stage >= ST_LISTS && pct_segment_histories(ctx) > 0;
    std::snprintf(buf, sizeof(buf),
                  ", \"list_entries\": %lld, \"list_slots\": %lld, \"list_capacity\": %lld, \"device_bytes_used\": %lld",
                  (long long)(segmented ? ctx->seg.entries : ctx->list_entries), (long long)ctx->list_slots,
                  (long long)ctx->list_cap, (long long)(mem_total - mem_free));
    s += buf;
    if (segmented) {
        std::snprintf(buf, sizeof(buf),
                      ", \"segments\": {\"histories\": %lld, \"count\": %lld, \"lists_ms\": %.3f, \"drop_ms\": %.3f}",
                      (long long)pct_segment_histories(ctx), (long long)ctx->seg.segments, ctx->seg.lists_ms,
                      ctx->seg.drop_ms);
        s += buf;
    }
    s += "}, \"stage_ms\": {";
    for (int k = 0; k <= TM_DROP; k++) {
        std::snprintf(buf, sizeof(buf), "%s\"%s\": %.4f", k ? ", " : "", kTimerName[k], ctx->t_stage[k]);
        s += buf;
    }
    std::snprintf(buf, sizeof(buf), "}, \"kernel_ms\": {\"mlp_lists\": %.4f}", ctx->t_mlp_kernel);
    s += buf;
    std::snprintf(buf, sizeof(buf), ", \"drop_plan\": {\"flags\": %d, \"groups_bucketed\": %lld, \"groups_sorted\": %lld}",
                  ctx->plan.flags, (long long)ctx->plan.groups_bucketed, (long long)ctx->plan.groups_sorted);
    s += buf;
    double tvs_ms = 0;
    for (const TvsRecord& r : ctx->tvs_log) tvs_ms += r.ms;
    std::snprintf(buf, sizeof(buf), ", \"tvs\": {\"iterations\": %zu, \"ms\": %.4f, \"L\": %d}}", ctx->tvs_log.size(),
                  tvs_ms, ctx->tvs_L);
    s += buf;
    ctx->metrics = s;
    return ctx->metrics.c_str();
}

int pct_reset_iterate(pct_ctx* ctx)
{
    if (!ctx) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_X0, "pct_reset_iterate"));
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(ctx->x, ctx->x0, ctx->P.num_voxels * sizeof(float), hipMemcpyDeviceToDevice,
                                      ctx->stream));
    restart_iterations(ctx);
    return PCT_OK;
}

int pct_get_tvs_log(pct_ctx* ctx, size_t capacity, double* tv_before, double* tv_after, int32_t* l_after,
                    int64_t* count)
{
    if (!ctx || !count) return PCT_E_ARG;
    *count = (int64_t)ctx->tvs_log.size();
    for (size_t i = 0; i < ctx->tvs_log.size() && i < capacity; i++) {
        if (tv_before) tv_before[i] = ctx->tvs_log[i].tv_before;
        if (tv_after) tv_after[i] = ctx->tvs_log[i].tv_after;
        if (l_after) l_after[i] = ctx->tvs_log[i].L;
    }
    return PCT_OK;
}

int pct_total_variation(pct_ctx* ctx, double* tv)
{
    if (!ctx || !tv) return PCT_E_ARG;
    PCT_TRY(require(ctx, ST_X0, "pct_total_variation"));
    return pct_stage_total_variation(ctx, tv);
}

int pct_drop_iteration_device(int64_t n, const int64_t* offsets, const uint32_t* voxels, const double* chord,
                              const float* wepl, float* x, double* accum, uint32_t* counts, int64_t num_voxels,
                              int block_size, int histories_per_block, int quirk_a3, int min_mlp_length, double lambda,
                              void* hip_stream)
{
    if (n < 0 || !offsets || !voxels || !chord || !wepl || !x || !accum || !counts || num_voxels <= 0 ||
        block_size <= 0 || histories_per_block <= 0)
        return PCT_E_ARG;
    std::string err;
    return pct_drop_run((hipStream_t)hip_stream, n, offsets, nullptr, voxels, chord, wepl, x, accum, nullptr, counts,
                        nullptr, nullptr, num_voxels, block_size, histories_per_block, quirk_a3, min_mlp_length, lambda,
                        PCT_ACCUM_FP64, &err);
}

}  // extern "C"

// unit conversion of freshly loaded raw data (k_preproc.hip has the kernel)
int pct_launch_units(hipStream_t s, int64_t n, float* const raw[14], int mm, double ts, double us, double vs);
int pct_convert_units(pct_ctx* ctx, int64_t first, int64_t n)
{
    const pct_params& P = ctx->P;
    if (!P.data_in_mm && P.t_shift == 0.0 && P.u_shift == 0.0 && P.v_shift == 0.0) return PCT_OK;
    float* r[14];
    for (int k = 0; k < 14; k++) r[k] = ctx->raw[k] + first;
    if (pct_launch_units(ctx->stream, n, r, P.data_in_mm, P.t_shift, P.u_shift, P.v_shift) != 0) {
        ctx->err = "unit conversion launch failed";
        return PCT_E_HIP;
    }
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PCT_OK;
}

extern "C" int pct_set_launch(pct_ctx* ctx, int drop_row_lanes, int drop_col_voxels, int drop_flags)
{
    if (!ctx) return PCT_E_ARG;
    const int g = drop_row_lanes, v = drop_col_voxels;
    if (!(g == 4 || g == 8 || g == 16 || g == 32 || g == 64) || !(v == 1 || v == 2 || v == 4 || v == 8)) return PCT_E_ARG;
    if (drop_flags & ~(PCT_DROP_SPATIAL_ORDER | PCT_DROP_XCD_SLABS | PCT_DROP_BRICKS | PCT_DROP_COLS_LDS |
                       PCT_DROP_COLS_SLICED | PCT_DROP_BUCKETS))
        return PCT_E_ARG;
    ctx->drop_row_lanes = g;
    ctx->drop_col_voxels = v;
    ctx->drop_flags = drop_flags;
    return PCT_OK;
}

// Output files of the reference's stages (array_2_disk + write_PNG) for the context's images.
extern "C" int pct_write_reference_layout(pct_ctx* ctx, const char* dir, int flags)
{
    if (!ctx || !dir || (flags & ~(PCT_WRITE_FBP | PCT_WRITE_HULL | PCT_WRITE_X0 | PCT_WRITE_X))) return PCT_E_ARG;
    const pct_params& P = ctx->P;
    std::vector<float> img((size_t)P.num_voxels);
    auto write = [&](int kind, const std::string& base, bool with_png) -> int {
        PCT_TRY(pct_get_image(ctx, kind, img.data()));
        const std::string stem = std::string(dir) + "/" + base;
        int rc = pct_write_image_text((stem + ".txt").c_str(), img.data(), P.columns, P.rows, P.slices);
        if (rc == PCT_OK && with_png) rc = pct_write_image_dat((stem + ".dat").c_str(), img.data(), P.columns, P.rows, P.slices);
        if (rc == PCT_OK && with_png) rc = pct_write_image_png((stem + ".png").c_str(), img.data(), P.columns, P.rows, P.slices);
        if (rc != PCT_OK) ctx->err = "writing " + stem + " failed";
        return rc;
    };
    if ((flags & PCT_WRITE_FBP) && ctx->fbp) PCT_TRY(write(PCT_IMG_FBP, "FBP", false));
    if ((flags & PCT_WRITE_HULL) && ctx->hull && ctx->stage >= ST_HULL_SEL) PCT_TRY(write(PCT_IMG_HULL, "hull", true));
    if ((flags & PCT_WRITE_X0) && ctx->x0 && ctx->stage >= ST_X0) PCT_TRY(write(PCT_IMG_X0, "x_0", true));
    if ((flags & PCT_WRITE_X) && ctx->x && ctx->stage >= ST_X0)
        PCT_TRY(write(PCT_IMG_X, "x_" + std::to_string(ctx->drop_k), true));
    return PCT_OK;
}

extern "C" int pct_set_hull(pct_ctx* ctx, const uint8_t* hull)
{
    if (!ctx || !hull) return PCT_E_ARG;
    // the reference's hull is bool (import_hull, V.cu:6659): any non-zero byte is inside
    std::vector<uint8_t> b((size_t)ctx->P.num_voxels);
    for (size_t i = 0; i < b.size(); i++) b[i] = hull[i] != 0;
    PCT_TRY(pct_alloc(ctx, &ctx->hull_import, ctx->P.num_voxels));
    PCT_CHECK_HIP(ctx, hipMemcpy(ctx->hull_import, b.data(), b.size(), hipMemcpyHostToDevice));
    return PCT_OK;
}
