// k_drop2.hip — stage (e): atomic-free DROP over per-block transposed path lists.
//
// DROP (DROP_full_tx_iteration, DS.cu:9537-9589) updates every voxel of a block by the mean of the
// updates of the block's histories through it:
//     x[v] += (sum over entries e of the block with voxel v of u_i(e)) / S[v],
//     u_i = L_i * ((b_i - L_i * sum_j x[v_ij]) / (L_i^2 n_i)) * lambda      (DS.cu:9321-9380, 9522-9536)
// The reference scatters u_i with float atomicAdd into x_update and counts S with atomicAdd
// (2 atomics per list entry).  Here the block's lists are transposed once, when the lists are built:
// for every voxel the block touches, the ascending list of its histories (an LSD radix sort by voxel
// is stable, so equal voxels keep history order).  Then each iteration runs, per block,
//   k_drop_rows   one wavefront per history: coalesced reads of its voxel list, gather of x, FP64
//                 reduction of a_i.x, u_i stored in r[i];
//   k_drop_cols   one lane per touched voxel: sum of r over its histories in history order
//                 (FP32 as the reference, FP32 with Kahan compensation, or FP64), S = the number of
//                 entries, x[v] += sum / S.
// No atomics, no scratch clearing, and the per-voxel sums are formed in the same order as a
// single-threaded run of the reference: the result is reproducible bit for bit.
#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// per history of a block: number of entries it contributes (0 for paths shorter than MIN_MLP_LENGTH)
__global__ void k_long_counts(int64_t start, int64_t proc, const uint32_t* __restrict__ cnt, int min_len,
                              uint32_t* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= proc) return;
    uint32_t c = cnt[start + i];
    out[i] = (int)c >= min_len ? c : 0u;
}

// (voxel, history-in-block) pairs in history order; one wavefront per history, coalesced.
__global__ void k_emit_pairs(int64_t start, int64_t proc, const int64_t* __restrict__ off,
                             const uint32_t* __restrict__ vox, const uint32_t* __restrict__ long_cnt,
                             const uint32_t* __restrict__ pos, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; i < proc; i += n_waves) {
        const uint32_t c = long_cnt[i];
        const int64_t base = off[start + i];
        const uint32_t p = pos[i];
        for (uint32_t j = lane; j < c; j += kWave) {
            keys[p + j] = vox[base + j];
            vals[p + j] = (uint32_t)i;
        }
    }
}

// per history: u_i (DS.cu:9321-9380 with the list of find_MLP_path_GPU_tabulated).
// G lanes per history (G | 64): lane l sums entries l, l+G, l+2G, ... with four gathers in flight,
// then the G partial sums are combined with a butterfly.  Several histories per wavefront keep
// enough independent loads in flight to cover the HBM/MALL latency.
template <int G>
__global__ void __launch_bounds__(kBlock) k_drop_rows(int64_t start, int64_t proc, const int64_t* __restrict__ off,
                                                       const uint32_t* __restrict__ cnt_arr,
                                                       const uint32_t* __restrict__ vox,
                                                       const double* __restrict__ chord,
                                                       const float* __restrict__ wepl, const float* __restrict__ x,
                                                       double lambda, double* __restrict__ r)
{
    const int lane = threadIdx.x & (G - 1);
    const int64_t n_groups = ((int64_t)gridDim.x * blockDim.x) / G;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; i < proc; i += n_groups) {
        const int64_t p = start + i;
        const uint32_t* __restrict__ lst = vox + off[p];
        const int cnt = (int)cnt_arr[p];
        double part = 0.0;
        int j = lane;
        for (; j + 3 * G < cnt; j += 4 * G) {
            const uint32_t a0 = lst[j], a1 = lst[j + G], a2 = lst[j + 2 * G], a3 = lst[j + 3 * G];
            const float x0 = x[a0], x1 = x[a1], x2 = x[a2], x3 = x[a3];
            part += (double)x0;
            part += (double)x1;
            part += (double)x2;
            part += (double)x3;
        }
        for (; j < cnt; j += G) part += (double)x[lst[j]];
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, G);
        if (lane == 0) {
            const double L = chord[p];
            const double ax = part * L;
            const double aa = (double)pct_sqf(L) * (double)cnt;
            r[i] = L * (((double)wepl[p] - ax) / aa) * lambda;
        }
    }
}

template <int ACC>
__device__ __forceinline__ float col_update(uint32_t b, uint32_t e, double first, const uint32_t* __restrict__ t_hist,
                                            const double* __restrict__ r)
{
    if (ACC == PCT_ACCUM_FP64) {
        double s = first;
        for (uint32_t k = b + 1; k < e; k++) s += r[t_hist[k]];
        return (float)(s / (double)(e - b));
    } else if (ACC == PCT_ACCUM_FP32) {
        float s = (float)first;
        for (uint32_t k = b + 1; k < e; k++) s += (float)r[t_hist[k]];
        return s / (float)(e - b);
    } else {  // Kahan-compensated FP32
        float s = 0.0f, c = 0.0f;
        float y = (float)first - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
        for (uint32_t k = b + 1; k < e; k++) {
            y = (float)r[t_hist[k]] - c;
            t = s + y;
            c = (t - s) - y;
            s = t;
        }
        return s / (float)(e - b);
    }
}

// per touched voxel: image_update_GPU (DS.cu:9398-9419) with the block's sums formed in history
// order.  Each lane takes V voxels (strided by the grid) and issues their loads together before any
// result is needed: segment bounds, first history, its update and x.
template <int ACC, int V>
__global__ void __launch_bounds__(kBlock) k_drop_cols(const uint32_t* __restrict__ t_vox,
                                                       const uint32_t* __restrict__ t_off,
                                                       const uint32_t* __restrict__ t_hist,
                                                       const uint32_t* __restrict__ n_vox,
                                                       const double* __restrict__ r, float* __restrict__ x)
{
    const uint32_t U = *n_vox;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t u0 = blockIdx.x * blockDim.x + threadIdx.x; u0 < U; u0 += stride * V) {
        uint32_t b[V], e[V], v[V], h0[V];
        float xv[V];
        double r0[V];
#pragma unroll
        for (int q = 0; q < V; q++) {
            const uint32_t u = u0 + q * stride;
            const bool ok = u < U;
            b[q] = ok ? t_off[u] : 0;
            e[q] = ok ? t_off[u + 1] : 0;
            v[q] = ok ? t_vox[u] : 0;
        }
#pragma unroll
        for (int q = 0; q < V; q++) h0[q] = e[q] > b[q] ? t_hist[b[q]] : 0;
#pragma unroll
        for (int q = 0; q < V; q++) {
            r0[q] = e[q] > b[q] ? r[h0[q]] : 0.0;
            xv[q] = e[q] > b[q] ? x[v[q]] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < V; q++)
            if (e[q] > b[q]) x[v[q]] = xv[q] + col_update<ACC>(b[q], e[q], r0[q], t_hist, r);
    }
}

inline unsigned grid_for(int64_t n, int block = kBlock)
{
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

// Build the per-block transposed lists (once per list build).
int pct_drop_prepare(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n_recon;
    DropPlan& D = ctx->plan;
    D.blocks.clear();
    // block ranges and entry counts (host pass over the list counts)
    std::vector<uint32_t> cnt(n > 0 ? n : 1);
    if (n > 0) PCT_CHECK_HIP(ctx, hipMemcpy(cnt.data(), ctx->list_cnt, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    int64_t total = 0, max_e = 0, max_proc = 0;
    for (int64_t start = 0; start < n; start += P.drop_block_size) {
        int64_t m = n - start < P.drop_block_size ? n - start : P.drop_block_size;
        int64_t proc = P.quirk_a3_drop_floor ? (m / P.histories_per_block) * P.histories_per_block : m;
        if (proc <= 0) continue;
        int64_t e = 0;
        for (int64_t i = start; i < start + proc; i++)
            if ((int)cnt[i] >= P.min_mlp_length) e += cnt[i];
        DropPlan::Block bk{start, proc, total, e};
        D.blocks.push_back(bk);
        total += e;
        max_e = e > max_e ? e : max_e;
        max_proc = proc > max_proc ? proc : max_proc;
    }
    if (max_e > 0xffffffffll) {
        ctx->err = "DROP block with more than 2^32 list entries";
        return PCT_E_LIMIT;
    }
    const int64_t nb = (int64_t)D.blocks.size();
    if (total > D.cap_entries || !D.t_hist) {
        PCT_TRY(pct_alloc(ctx, &D.t_hist, total));
        PCT_TRY(pct_alloc(ctx, &D.t_vox, total));
        D.cap_entries = total;
    }
    if (total + nb > D.cap_off || !D.t_off) {
        PCT_TRY(pct_alloc(ctx, &D.t_off, total + nb));
        D.cap_off = total + nb;
    }
    if (nb > D.cap_blocks || !D.n_vox) {
        PCT_TRY(pct_alloc(ctx, &D.n_vox, nb));
        D.cap_blocks = nb;
    }
    if (max_proc > D.cap_r || !D.r) {
        PCT_TRY(pct_alloc(ctx, &D.r, max_proc));
        D.cap_r = max_proc;
    }
    // scratch: keys/values (max_e + 1), long counts / positions (max_proc + 1)
    uint32_t *keys = nullptr, *keys_sorted = nullptr, *vals = nullptr, *runs = nullptr, *lc = nullptr, *pos = nullptr;
    PCT_TRY(pct_alloc(ctx, &keys, max_e + 1));
    PCT_TRY(pct_alloc(ctx, &keys_sorted, max_e + 1));
    PCT_TRY(pct_alloc(ctx, &vals, max_e + 1));
    PCT_TRY(pct_alloc(ctx, &runs, max_e + 1));
    PCT_TRY(pct_alloc(ctx, &lc, max_proc + 1));
    PCT_TRY(pct_alloc(ctx, &pos, max_proc + 1));
    int end_bit = 1;
    while ((1ll << end_bit) < (int64_t)P.num_voxels) end_bit++;
    size_t need = 0, t1 = 0, t2 = 0, t3 = 0;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, t1, lc, pos, (int)(max_proc + 1), s));
    PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, t2, keys, keys_sorted, vals, vals, (int)(max_e + 1),
                                                          0, end_bit, s));
    PCT_CHECK_HIP(ctx, hipcub::DeviceRunLengthEncode::Encode(nullptr, t3, keys_sorted, keys, runs, D.n_vox,
                                                             (int)(max_e + 1), s));
    need = std::max(t1, std::max(t2, t3));
    size_t t4 = 0;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, t4, runs, D.t_off, (int)(max_e + 1), s));
    need = std::max(need, t4);
    if (need > ctx->cub_tmp_bytes) {
        PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, need));
        ctx->cub_tmp_bytes = need;
    }
    for (int64_t bi = 0; bi < nb; bi++) {
        const DropPlan::Block& bk = D.blocks[bi];
        const int64_t E = bk.entries, proc = bk.proc;
        uint32_t* t_hist = D.t_hist + bk.entry_base;
        uint32_t* t_vox = D.t_vox + bk.entry_base;
        uint32_t* t_off = D.t_off + bk.entry_base + bi;
        if (E == 0) {
            PCT_CHECK_HIP(ctx, hipMemsetAsync(D.n_vox + bi, 0, sizeof(uint32_t), s));
            PCT_CHECK_HIP(ctx, hipMemsetAsync(t_off, 0, sizeof(uint32_t), s));
            continue;
        }
        hipLaunchKernelGGL(k_long_counts, dim3(grid_for(proc)), dim3(kBlock), 0, s, bk.start, proc, ctx->list_cnt,
                           P.min_mlp_length, lc);
        size_t tb = ctx->cub_tmp_bytes;
        PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, tb, lc, pos, (int)proc, s));
        int64_t blocks = (proc * kWave + kBlock - 1) / kBlock;
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(k_emit_pairs, dim3((unsigned)blocks), dim3(kBlock), 0, s, bk.start, proc, ctx->list_off,
                           ctx->list_vox, lc, pos, keys, vals);
        tb = ctx->cub_tmp_bytes;
        PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp, tb, keys, keys_sorted, vals, t_hist, (int)E,
                                                              0, end_bit, s));
        PCT_CHECK_HIP(ctx, hipMemsetAsync(runs, 0, (E + 1) * sizeof(uint32_t), s));
        tb = ctx->cub_tmp_bytes;
        PCT_CHECK_HIP(ctx, hipcub::DeviceRunLengthEncode::Encode(ctx->cub_tmp, tb, keys_sorted, t_vox, runs,
                                                                 D.n_vox + bi, (int)E, s));
        tb = ctx->cub_tmp_bytes;
        PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, tb, runs, t_off, (int)(E + 1), s));
        PCT_CHECK_HIP(ctx, hipGetLastError());
    }
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    pct_free(&keys);
    pct_free(&keys_sorted);
    pct_free(&vals);
    pct_free(&runs);
    pct_free(&lc);
    pct_free(&pos);
    D.ready = true;
    return PCT_OK;
}

// One DROP iteration with the prepared plan; optional per-launch timing.
int pct_drop_iterate_plan(pct_ctx* ctx, DropTiming* timing)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    DropPlan& D = ctx->plan;
    std::vector<hipEvent_t> ev;
    auto mark = [&]() {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        ev.push_back(e);
    };
    const int G = ctx->drop_row_lanes, V = ctx->drop_col_voxels;
    for (size_t bi = 0; bi < D.blocks.size(); bi++) {
        const DropPlan::Block& bk = D.blocks[bi];
        int64_t blocks = (bk.proc * G + kBlock - 1) / kBlock;
        if (blocks > 16384) blocks = 16384;
        if (timing) mark();
        dim3 rg((unsigned)blocks), rb(kBlock);
#define PCT_ROWS(GG)                                                                                             \
    hipLaunchKernelGGL(k_drop_rows<GG>, rg, rb, 0, s, bk.start, bk.proc, ctx->list_off, ctx->list_cnt, ctx->list_vox, \
                       ctx->chord, ctx->wepl, ctx->x, P.lambda, D.r)
        switch (G) {
        case 4: PCT_ROWS(4); break;
        case 8: PCT_ROWS(8); break;
        case 16: PCT_ROWS(16); break;
        case 32: PCT_ROWS(32); break;
        default: PCT_ROWS(64); break;
        }
#undef PCT_ROWS
        if (timing) mark();
        if (bk.entries > 0) {
            int64_t cb = (bk.entries + (int64_t)kBlock * V - 1) / ((int64_t)kBlock * V);
            if (cb > 4096) cb = 4096;
            const uint32_t* t_vox = D.t_vox + bk.entry_base;
            const uint32_t* t_off = D.t_off + bk.entry_base + bi;
            const uint32_t* t_hist = D.t_hist + bk.entry_base;
            dim3 cg((unsigned)cb), cbk(kBlock);
#define PCT_COLS(A, VV) \
    hipLaunchKernelGGL((k_drop_cols<A, VV>), cg, cbk, 0, s, t_vox, t_off, t_hist, D.n_vox + bi, D.r, ctx->x)
#define PCT_COLS_V(A)                  \
    switch (V) {                       \
    case 1: PCT_COLS(A, 1); break;     \
    case 2: PCT_COLS(A, 2); break;     \
    case 8: PCT_COLS(A, 8); break;     \
    default: PCT_COLS(A, 4); break;    \
    }
            if (P.drop_accum == PCT_ACCUM_FP64) {
                PCT_COLS_V(PCT_ACCUM_FP64)
            } else if (P.drop_accum == PCT_ACCUM_FP32) {
                PCT_COLS_V(PCT_ACCUM_FP32)
            } else {
                PCT_COLS_V(PCT_ACCUM_KAHAN32)
            }
#undef PCT_COLS_V
#undef PCT_COLS
        }
        if (timing) {
            mark();
            timing->histories += bk.proc;
        }
    }
    PCT_CHECK_HIP(ctx, hipGetLastError());
    if (timing) {
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
        for (size_t j = 0; j + 2 < ev.size(); j += 3) {
            float a = 0, b = 0;
            (void)hipEventElapsedTime(&a, ev[j], ev[j + 1]);
            (void)hipEventElapsedTime(&b, ev[j + 1], ev[j + 2]);
            timing->scatter_ms += a;
            timing->apply_ms += b;
            timing->scatter_launches++;
            timing->apply_launches++;
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    return PCT_OK;
}
