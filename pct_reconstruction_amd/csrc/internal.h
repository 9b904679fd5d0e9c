// internal.h — context layout and launcher declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/pct.h"
#include "devmath.h"
#include "launch.h"

#define PCT_CHECK_HIP(ctx, expr)                                                        \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
            return PCT_E_HIP;                                                           \
        }                                                                               \
    } while (0)

#define PCT_TRY(expr)               \
    do {                            \
        int rc_ = (expr);           \
        if (rc_ != PCT_OK) return rc_; \
    } while (0)

enum pct_stage {
    ST_EMPTY = 0,
    ST_LOADED,
    ST_BINNED,      // intersections + binning + carving counts done
    ST_HULLS,       // edge detection done
    ST_CUT,         // statistical cuts done
    ST_FBP,         // sinogram + FBP done
    ST_HULL_SEL,    // hull selected
    ST_X0,          // initial iterate defined
    ST_ENDPOINTS,   // MLP endpoints + compaction
    ST_LISTS,       // MLP lists built
};

// Device copies of the MLP lookup tables (DS.cu:8421-8502) plus the per-step Sigma_1 terms.
struct DevTables {
    double* sin_t = nullptr;
    double* cos_t = nullptr;
    double* scat = nullptr;
    double* p12 = nullptr;
    double* p26 = nullptr;
    double* p312 = nullptr;
    double* step_terms = nullptr;  // [k][8]: u1_12, u1_23, p34(u1), S1I0, S1I1, S1I2, unused, unused
    int n_steps = 0;               // rows of step_terms
    int64_t n_trig = 0, n_coeff = 0, n_poly = 0;
};

// Per history, in visiting order: what the row kernel needs before its list (one 24-byte load).
struct RowDesc {
    int64_t off;   // first list entry
    double chord;  // effective chord length
    uint32_t cnt;  // list length
    float wepl;
};

// Per-block transposed lists for the atomic-free DROP (k_drop2.hip).  Block b covers histories
// [start, start + proc); its `entries` list entries (histories with >= MIN_MLP_LENGTH voxels) are
// grouped by voxel: t_vox[u] touched voxel, t_off[u..u+1] its range in t_hist (visiting rank of
// each history through the voxel, in ascending history order).  Blocks are transposed in groups
// (one sort per group by (block, voxel)): a block's runs start at run_base in t_vox / t_off, and
// t_off counts entries from its group's first entry, rank_base in t_hist.
// Sliced columns (PCT_DROP_COLS_SLICED): a block's runs in slices of 64 (the runs one column
// wavefront takes); slice s holds the ranks of its runs column-major, entry m of run 64 j + l at
// sell[sell_off[s] + 64 m + l], in 64 x (largest entry count of the slice) slots.
struct SliceBlock {
    int64_t slice_base;  // first slice of the block (the entry after the last block: total slices)
    int64_t run;         // absolute index of the block's first run in t_vox / t_off
    int64_t entry_base;  // the block's group's first entry in t_hist
    uint32_t n_vox, pad;
};

struct DropPlan {
    struct Block {
        int64_t start, proc, entry_base, entries;
        int64_t group, rank_base, run_base;  // run_base: absolute index of the block's first run
        uint32_t n_vox;                      // touched voxels
        int64_t slice_base;                  // first slice (sliced columns)
    };
    struct Group {
        int64_t first_block, n_blocks, entry_base, entries, hist_begin, hist_end;
    };
    std::vector<Block> blocks;
    std::vector<Group> groups;
    uint32_t* t_vox = nullptr;   // per run: voxel
    uint32_t* t_off = nullptr;   // per run (+1 per group): first entry of the run, relative to the group
    uint32_t* t_hist = nullptr;  // per entry: the visiting rank of the history (ascending history order)
    uint32_t* perm = nullptr;    // per reconstruction history: in-block index of the history visited at rank k
    uint32_t* rank = nullptr;    // per reconstruction history: its visiting rank within its block
    RowDesc* desc = nullptr;     // per reconstruction history, in visiting order
    double* r = nullptr;         // per rank of the current block: the update value of that history
    // scratch of the group transposition (kept between list builds)
    uint32_t *keys = nullptr, *keys_sorted = nullptr, *vals = nullptr, *runs = nullptr, *n_runs = nullptr;
    uint32_t *hkey = nullptr, *hkey_sorted = nullptr, *hidx = nullptr, *hidx_sorted = nullptr;
    uint32_t *lc = nullptr, *pos = nullptr;
    int32_t* run_start = nullptr;  // per block of a group: index of its first run, -1 if none
    uint32_t* sell = nullptr;      // sliced columns: ranks, column-major per slice
    uint64_t* sell_off = nullptr;  // per slice (+1): first slot in sell
    uint64_t* slice_w = nullptr;   // per slice (+1): slots (scratch of the build)
    SliceBlock* slice_blocks = nullptr;
    int64_t cap_sell = 0, cap_slices = 0, cap_slice_w = 0, cap_slice_blocks = 0;
    // brick layout (PCT_DROP_BRICKS): list entries and image in 4x4x2 bricks of 32 voxels
    uint32_t* vox_b = nullptr;     // list entries as brick indices (same CSR layout as list_vox)
    float* x_b = nullptr;          // image in brick order during DROP
    int64_t cap_vox_b = 0, cap_x_b = 0, nv_b = 0;
    bool vox_b_ready = false;      // vox_b holds the current lists (written by the MLP kernel)
    int nbx = 0, nby = 0;
    int64_t cap_entries = 0, cap_off = 0, cap_r = 0, cap_perm = 0, cap_group_e = 0, cap_group_h = 0,
            cap_group_b = 0;
    // bucket transposition (PCT_DROP_BUCKETS): per bucket counts, offsets, cursors, runs, run offsets; size
    // class lists (3 x buckets) and counters (3 + overflow flag)
    uint32_t *bk_cnt = nullptr, *bk_off = nullptr, *bk_cur = nullptr, *bk_runs = nullptr, *bk_roff = nullptr;
    uint32_t *bk_list = nullptr, *bk_misc = nullptr;
    int64_t cap_bk = 0;
    int64_t groups_bucketed = 0, groups_sorted = 0;  // groups transposed by each path in the last build
    int flags = -1;              // PCT_DROP_* flags the plan was built with
    bool desc_all = false;       // desc holds a descriptor for every reconstruction history
    bool ready = false;
};

struct TvsRecord {
    double tv_before, tv_after;  // TV of x before / after the NTVS iteration
    int L;                       // exponent L after it
    double ms;                   // device time of the iteration
};

// Totals of one segmented DROP iteration, summed over its segments.
struct SegStats {
    int64_t segments = 0, entries = 0, histories = 0, gathered = 0, scattered = 0, touched = 0;
    double lists_ms = 0, drop_ms = 0;  // host wall time in the list builds (incl. plans) and in the DROP blocks
};

struct pct_ctx {
    pct_params P;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // the MLP list kernels while the DROP plan is built (default priority)
    std::string err;
    std::string metrics;
    int stage = ST_EMPTY;

    // raw tracker data (kept so the pipeline can be re-run): t0..t3, v0..v3, u0..u3, wepl, angle
    int64_t n_loaded = 0, cap_loaded = 0;
    float* raw[14] = {};

    // per-history arrays of the current stage
    int64_t n_hist = 0, cap_hist = 0;
    int64_t n_recon_volume = 0, n_post_cuts = 0, n_recon = 0;
    float* h[10] = {};        // x/y/z entry, x/y/z exit, xy/xz entry angle, xy/xz exit angle
    float* wepl = nullptr;
    float* gantry = nullptr;  // gantry angle of the history's projection [deg]
    int32_t* bin_num = nullptr;
    uint8_t* in_range = nullptr;
    uint32_t* first_voxel = nullptr;
    // scratch of the same capacity
    float* h_tmp[10] = {};
    float* wepl_tmp = nullptr;
    float* gantry_tmp = nullptr;
    int32_t* bin_tmp = nullptr;
    uint8_t* u8_tmp = nullptr;
    uint8_t* flag = nullptr;
    int64_t* pos = nullptr;
    uint32_t* u32_a = nullptr;
    uint32_t* u32_b = nullptr;
    uint32_t* u32_c = nullptr;
    uint32_t* u32_d = nullptr;
    void* cub_tmp = nullptr;
    size_t cub_tmp_bytes = 0;
    int64_t bin_perm_n = -1;  // u32_d holds the bin order of the first bin_perm_n histories (pct_bin_order)
    int carve_spill = 512;    // LDS count at which a carve visit goes to global memory, 1..512 (PCT_CARVE_SPILL)
    int carve_table = 8192;   // LDS hash slots of the fused carve (4096, 8192, 16384); 0: one walk per method (PCT_CARVE_TABLE)
    int carve_line_order = 1;  // hull carving in chord-line order (both directions of a line together; PCT_CARVE_LINES)
    int endpoint_order = 1;   // MLP endpoints in sinogram-bin order (PCT_ENDPOINT_ORDER)
    int carve_block = 512;    // lanes per workgroup of the fused carve (256, 512; PCT_CARVE_BLOCK)

    // bins
    int32_t* counts_pre = nullptr;
    int32_t* counts_post = nullptr;
    int32_t* bin_off = nullptr;  // NB + 1
    float* mean[3] = {};
    float* stddev[3] = {};
    float* sinogram = nullptr;
    float* sino_filtered = nullptr;

    // images
    uint32_t* sc_carved = nullptr;  // bit-packed carved mask for SC
    uint8_t* sc_hull = nullptr;
    int32_t* msc_counts = nullptr;
    int32_t* sm_counts = nullptr;
    int32_t* msc_hull = nullptr;
    int32_t* sm_hull = nullptr;
    uint8_t* fbp_hull = nullptr;
    uint8_t* hull = nullptr;
    uint8_t* hull_import = nullptr;
    float* fbp = nullptr;
    float* fbp_x0 = nullptr;          // median-filtered FBP (MEDIAN_FILTER_FBP), feeds x0 only
    uint8_t* hull_endpoints = nullptr; // average-filtered hull (AVG_FILTER_HULL), feeds the MLP endpoints
    float* x0 = nullptr;
    float* x = nullptr;

    // MLP path lists (d): CSR with capacity per history; entries [off, off + cnt)
    int64_t* list_off = nullptr;   // n_recon + 1 (capacity offsets)
    uint32_t* list_cnt = nullptr;  // n_recon
    uint32_t* list_vox = nullptr;
    double* chord = nullptr;       // n_recon
    int64_t* cap_scratch = nullptr;  // n_recon + 1
    double* res_r2 = nullptr;        // residual: per history (b - a.x)^2, and their sum
    int64_t res_cap = 0;
    int64_t list_cap = 0, list_entries = 0, list_hist_cap = 0;
    int64_t list_slots = 0;  // capacity slots of the last list build (1 + u-steps per history)
    DevTables T;

    // DROP plan (transposed lists) and the scratch of the atomic DROP kernels
    DropPlan plan;
    int drop_row_lanes = 16;   // lanes per history in the row kernel (PCT_DROP_ROW_LANES: 4..64)
    int drop_col_voxels = 1;   // voxels per lane in the column kernel (PCT_DROP_COL_VOXELS: 1, 2, 4, 8)
    int drop_flags = PCT_DROP_SPATIAL_ORDER | PCT_DROP_BRICKS | PCT_DROP_COLS_SLICED;  // PCT_DROP_FLAGS
    int mlp_blocks_per_cu = -1;  // MLP list kernel grid: workgroups per CU, 0 = one lane per history, -1 = auto (PCT_MLP_BLOCKS_PER_CU)
    int drop_pipeline = 1;  // MLP lists of the next transposition groups built while a group is transposed (PCT_DROP_PIPELINE)
    int64_t drop_group_entries = 1ll << 30;  // list entries per transposition group (PCT_DROP_GROUP_ENTRIES)
    double* accum64 = nullptr;
    float* accum32 = nullptr;
    uint32_t* S = nullptr;
    uint32_t* touched = nullptr;
    uint32_t* n_touched = nullptr;  // 2 counters (ping-pong)
    int32_t* list_overflow = nullptr;

    // total variation superiorization between DROP iterations (k_tvs.hip)
    float* tvs_G = nullptr;        // subgradient of the TV, NUM_VOXELS
    double* tvs_part = nullptr;    // 2 x tiles of 1024 voxels: FP64 tile sums
    double* tvs_res = nullptr;     // 2 reduced sums
    double* tvs_res_h = nullptr;   // pinned host copy of tvs_res
    int64_t tvs_nt = 0;
    int tvs_L = 0;                 // the exponent L carried between iterations (DS.h:439)
    int drop_k = 0;                // DROP iterations run since the initial iterate
    std::vector<TvsRecord> tvs_log;

    // timing (ms) of the last run of each stage, and of the last MLP list kernel launch
    double t_stage[16] = {};
    double t_mlp_kernel = 0;

    // list segments (LIST_SEGMENT_HISTORIES): the lists and the DROP plan describe one segment at a time
    bool seg_inside = false;      // a SegmentView is active: the per-history arrays are a segment's window
    bool seg_lists_only = false;  // build a segment's lists without its DROP plan (residual)
    SegStats seg;                 // totals of the last segmented DROP iteration
};

// allocation helpers
template <class T>
int pct_alloc(pct_ctx* ctx, T** p, size_t count)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        ctx->err = std::string("hipMalloc(") + std::to_string(count * sizeof(T)) + " B): " + hipGetErrorString(e);
        *p = nullptr;
        return PCT_E_NOMEM;
    }
    return PCT_OK;
}
template <class T>
void pct_free(T** p)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
}
// A device buffer owned by one stage call: freed when the call returns, on its error paths too.
template <class T>
struct pct_tmp {
    T* p = nullptr;
    pct_tmp() = default;
    pct_tmp(const pct_tmp&) = delete;
    pct_tmp& operator=(const pct_tmp&) = delete;
    ~pct_tmp() { pct_free(&p); }
};

DevGeom pct_geom(const pct_params& P);

// host-side tables (tables.cpp)
struct HostTables {
    std::vector<double> sin_t, cos_t, scat, p12, p23, p34, p26, p312;
};
void pct_build_tables(const pct_params& P, HostTables* t);
// per-step Sigma_1 terms for steps k = 0..n-1 (u_1 = (k+1) * MLP_U_STEP)
void pct_build_step_terms(const pct_params& P, const HostTables& t, int n, std::vector<double>* out);

// stage launchers (one per translation unit)
int pct_stage_intersect_bin_carve(pct_ctx* ctx);
int pct_stage_hull_finish(pct_ctx* ctx);
int pct_stage_stats_and_cuts(pct_ctx* ctx);
int pct_stage_fbp(pct_ctx* ctx);
int pct_stage_select_hull(pct_ctx* ctx);
int pct_stage_initial_iterate(pct_ctx* ctx, const float* x0_import);
int pct_stage_endpoints(pct_ctx* ctx);
int pct_stage_build_lists(pct_ctx* ctx);
struct DropTiming;
int pct_stage_drop(pct_ctx* ctx, int iterations, DropTiming* timing = nullptr);
int pct_drop_entries(pct_ctx* ctx, int64_t* histories, int64_t* gathered, int64_t* scattered);
// Launches the MLP list kernel for the histories [h0, h1) on a stream (k_mlp.hip).
using MlpLaunch = std::function<int(int64_t h0, int64_t h1, hipStream_t s)>;
// Builds the DROP plan from the lists.  With `mlp`, the lists are not built yet: the MLP kernel runs
// group by group on ctx->stream2 while the earlier groups are transposed on ctx->stream.
int pct_drop_prepare(pct_ctx* ctx, const MlpLaunch* mlp = nullptr);
int pct_median_filter_fbp(pct_ctx* ctx);
int pct_avg_filter_hull(pct_ctx* ctx);
void pct_drop_plan_free(DropPlan* plan);
int pct_drop_enter(pct_ctx* ctx, bool enter);  // image into (true) / out of the DROP layout
int pct_drop_iterate_plan(pct_ctx* ctx, DropTiming* timing);
int pct_stage_residual(pct_ctx* ctx, double* norm);
// list segments: the segment length in histories (0 = no segments), and the segmented DROP iteration and
// residual over every segment (segments.hip)
int64_t pct_segment_histories(const pct_ctx* ctx);
int pct_stage_drop_segments(pct_ctx* ctx, int iterations, DropTiming* timing);
int pct_stage_residual_segments(pct_ctx* ctx, double* norm);
int pct_tvs_iteration(pct_ctx* ctx, int k);               // NTVS_iteration for DROP iteration k (1-based)
int pct_stage_total_variation(pct_ctx* ctx, double* tv);  // TV of ctx->x, fixed-order FP64 sum

// compaction: keep entries with flag != 0 (flag array has ctx->n_hist entries); returns new count
int pct_compact(pct_ctx* ctx, const uint8_t* keep, int64_t n, int64_t* n_out, bool with_first_voxel);
// per-history arrays (and their compaction scratch) for n histories
int pct_ensure_hist_capacity(pct_ctx* ctx, int64_t n);
int pct_bin_order(pct_ctx* ctx, int64_t n);
int pct_line_order(pct_ctx* ctx, int64_t n);
// per-kernel timing of DROP launches (HIP events around every launch on the launch stream)
struct DropTiming {
    double scatter_ms = 0, apply_ms = 0;
    int64_t scatter_launches = 0, apply_launches = 0;
    int64_t entries = 0, histories = 0;  // algorithmic work of the timed scatter launches
};
// DROP iteration on raw device buffers
int pct_drop_run(hipStream_t s, int64_t n, const int64_t* off, const uint32_t* cnt, const uint32_t* vox,
                 const double* chord, const float* wepl, float* x, double* accum64, float* accum32, uint32_t* S,
                 uint32_t* touched, uint32_t* n_touched, int64_t num_voxels, int block_size, int hpb, int quirk_a3,
                 int min_len, double lambda, int accum_mode, std::string* err, DropTiming* timing = nullptr);
