// segments.hip — MLP lists and DROP plans built one segment of histories at a time.
//
// The reference bounds device memory on large runs by moving the histories of each DROP block to the
// GPU and recomputing their MLP paths inside the block update (DROP_partial_tx_iteration and
// DROP_partial_tx_preallocated_iteration, DS.cu:9590-9742; MLP_IN_LOOP).  Here the histories stay
// resident (48 bytes each) and, with LIST_SEGMENT_HISTORIES > 0, every DROP iteration walks the
// reconstruction histories in segments of whole DROP blocks: build the segment's lists and its
// transposed DROP plan, run its blocks, move on.  Device memory holds the lists and plan of one segment.
//
// A segment is a window [a, b) of the per-history arrays with a a multiple of DROP_BLOCK_SIZE, so its
// blocks are exactly the blocks of the unsegmented run (same histories, same order, the same quirk-A3
// floor on the final partial block) and the DROP image is bit-identical to the resident run.
#include <algorithm>
#include <chrono>

#include "internal.h"

int64_t pct_segment_histories(const pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    if (P.list_segment_histories <= 0) return 0;
    const int64_t BS = P.drop_block_size;
    int64_t seg = ((int64_t)P.list_segment_histories / BS) * BS;
    if (seg < BS) seg = BS;
    return ctx->n_recon > seg ? seg : 0;
}

namespace {

// The per-history arrays of the reconstruction histories narrowed to [a, b) for the lifetime of the
// view; the context's own pointers and counts come back on destruction, on every path.
class SegmentView {
public:
    SegmentView(pct_ctx* ctx, int64_t a, int64_t b, bool lists_only) : ctx_(ctx)
    {
        for (int k = 0; k < 10; k++) h_[k] = ctx->h[k];
        wepl_ = ctx->wepl;
        bin_ = ctx->bin_num;
        first_ = ctx->first_voxel;
        gantry_ = ctx->gantry;
        n_hist_ = ctx->n_hist;
        n_recon_ = ctx->n_recon;
        for (int k = 0; k < 10; k++) ctx->h[k] = h_[k] + a;
        ctx->wepl = wepl_ + a;
        ctx->bin_num = bin_ + a;
        ctx->first_voxel = first_ + a;
        if (gantry_) ctx->gantry = gantry_ + a;
        ctx->n_hist = b - a;
        ctx->n_recon = b - a;
        ctx->seg_inside = true;
        ctx->seg_lists_only = lists_only;
    }
    ~SegmentView()
    {
        for (int k = 0; k < 10; k++) ctx_->h[k] = h_[k];
        ctx_->wepl = wepl_;
        ctx_->bin_num = bin_;
        ctx_->first_voxel = first_;
        ctx_->gantry = gantry_;
        ctx_->n_hist = n_hist_;
        ctx_->n_recon = n_recon_;
        ctx_->seg_inside = false;
        ctx_->seg_lists_only = false;
        ctx_->plan.ready = false;  // the plan described the segment's window
        ctx_->stage = ST_LISTS;
    }
    SegmentView(const SegmentView&) = delete;
    SegmentView& operator=(const SegmentView&) = delete;

private:
    pct_ctx* ctx_;
    float* h_[10];
    float *wepl_, *gantry_;
    int32_t* bin_;
    uint32_t* first_;
    int64_t n_hist_, n_recon_;
};

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// histories processed, list entries scattered and voxels touched by the blocks of the current segment's plan;
// with `all`, also the entries gathered (every history's list, read back from the device)
int plan_totals(pct_ctx* ctx, SegStats* st, bool all)
{
    for (const auto& bk : ctx->plan.blocks) {
        st->histories += bk.proc;
        st->scattered += bk.entries;
        st->touched += bk.n_vox;
    }
    if (all) {
        int64_t h = 0, g = 0, sc = 0;
        PCT_TRY(pct_drop_entries(ctx, &h, &g, &sc));
        st->gathered += g;
    }
    return PCT_OK;
}

// One DROP iteration over every segment.
int drop_iteration_segments(pct_ctx* ctx, int64_t seg, DropTiming* timing)
{
    const int64_t n = ctx->n_recon;
    SegStats st;
    for (int64_t a = 0; a < n; a += seg) {
        const int64_t b = std::min(n, a + seg);
        SegmentView view(ctx, a, b, false);
        auto t0 = std::chrono::steady_clock::now();
        PCT_TRY(pct_stage_build_lists(ctx));
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        st.lists_ms += ms_since(t0);
        st.entries += ctx->list_entries;
        PCT_TRY(plan_totals(ctx, &st, timing != nullptr));
        t0 = std::chrono::steady_clock::now();
        PCT_TRY(pct_drop_enter(ctx, true));
        PCT_TRY(pct_drop_iterate_plan(ctx, timing));
        PCT_TRY(pct_drop_enter(ctx, false));
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        st.drop_ms += ms_since(t0);
        st.segments++;
    }
    ctx->seg = st;
    return PCT_OK;
}

}  // namespace

// DROP_GPU (DS.cu:9848-9938) with segmented lists; TVS between iterations as in pct_stage_drop.
int pct_stage_drop_segments(pct_ctx* ctx, int iterations, DropTiming* timing)
{
    const int64_t seg = pct_segment_histories(ctx);
    if (seg <= 0 || ctx->seg_inside) {
        ctx->err = "segmented DROP without segments";
        return PCT_E_STATE;
    }
    for (int it = 0; it < iterations; it++) {
        const int k = ++ctx->drop_k;
        if (ctx->P.tvs_on && ctx->P.tvs_first) PCT_TRY(pct_tvs_iteration(ctx, k));
        PCT_TRY(drop_iteration_segments(ctx, seg, timing));
        if (ctx->P.tvs_on && !ctx->P.tvs_first) PCT_TRY(pct_tvs_iteration(ctx, k));
    }
    return PCT_OK;
}

// ||b - A x|| with the lists of one segment at a time (no DROP plan); the squares are summed per
// segment on the device and the segment sums in order on the host.
int pct_stage_residual_segments(pct_ctx* ctx, double* norm)
{
    const int64_t seg = pct_segment_histories(ctx);
    if (seg <= 0 || ctx->seg_inside) {
        ctx->err = "segmented residual without segments";
        return PCT_E_STATE;
    }
    const int64_t n = ctx->n_recon;
    double sum = 0.0;
    for (int64_t a = 0; a < n; a += seg) {
        const int64_t b = std::min(n, a + seg);
        SegmentView view(ctx, a, b, true);
        PCT_TRY(pct_stage_build_lists(ctx));
        double part = 0.0;
        PCT_TRY(pct_stage_residual(ctx, &part));
        sum += part * part;
    }
    *norm = std::sqrt(sum);
    return PCT_OK;
}
