// k_drop2.hip — stage (e): atomic-free DROP over per-block transposed path lists.
//
// DROP (DROP_full_tx_iteration, DS.cu:9537-9589) updates every voxel of a block by the mean of the
// updates of the block's histories through it:
//     x[v] += (sum over entries e of the block with voxel v of u_i(e)) / S[v],
//     u_i = L_i * ((b_i - L_i * sum_j x[v_ij]) / (L_i^2 n_i)) * lambda      (DS.cu:9321-9380, 9522-9536)
// The reference scatters u_i with float atomicAdd into x_update and counts S with atomicAdd
// (2 atomics per list entry).  Here the block's lists are transposed once, when the lists are built:
// for every voxel the block touches, the list of its histories in ascending history order (an LSD
// radix sort by voxel is stable, so equal voxels keep history order).  Then each iteration runs,
// per block,
//   k_drop_rows   G lanes per history: coalesced reads of its voxel list, gather of x, FP64
//                 reduction of a_i.x, u_i stored in r[rank];
//   k_drop_cols   one lane per touched voxel: sum of r over its histories in history order
//                 (FP32 as the reference, FP32 with Kahan compensation, or FP64), S = the number of
//                 entries, x[v] += sum / S.
// No atomics, no scratch clearing, and the per-voxel sums are formed in the same order as a
// single-threaded run of the reference: the result is reproducible bit for bit.
//
// Locality.  Both kernels are limited by the rate of scattered 4-8 byte accesses (x gathers in the
// rows, r gathers in the columns), which grows with the number of distinct cache lines a wavefront
// instruction touches.  The histories of a block are therefore visited in sinogram-bin order
// (v bin, angle bin, t bin — DS.cu:5644-5660 numbering): neighbouring ranks are neighbouring
// parallel paths, so the histories sharing a wavefront gather from the same lines of x, and the
// histories through neighbouring voxels have neighbouring ranks in r.  The rank only decides where
// u_i is stored; the column sums still run in history order.  Workgroups are dispatched round-robin
// over the 8 XCDs; remapping the workgroup index gives each XCD one contiguous slab of ranks (rows)
// or of touched voxels (columns), so each XCD's L2 holds a slab of x instead of all of it.
#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr unsigned kXcds = 8;  // MI355X: 8 XCDs of 32 CUs; workgroup w runs on XCD w % 8

// Logical workgroup index.  With per > 0 the XCD (blockIdx % 8) takes the logical indices
// [xcd * per, (xcd + 1) * per): a contiguous slab per XCD.  per == 0: identity.
__device__ __forceinline__ uint32_t slab_index(uint32_t per)
{
    return per ? (blockIdx.x % kXcds) * per + blockIdx.x / kXcds : blockIdx.x;
}

// per history of a block: sort key (its sinogram bin) and in-block index
__global__ void k_hist_keys(int64_t start, int64_t proc, const int32_t* __restrict__ bin_num,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ idx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= proc) return;
    keys[i] = (uint32_t)bin_num[start + i];
    idx[i] = (uint32_t)i;
}

__global__ void k_iota32(int64_t n, uint32_t* __restrict__ a, uint32_t* __restrict__ b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i] = (uint32_t)i;
    b[i] = (uint32_t)i;
}

// rank[perm[k]] = k
__global__ void k_rank(int64_t proc, const uint32_t* __restrict__ perm, uint32_t* __restrict__ rank)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= proc) return;
    rank[perm[k]] = (uint32_t)k;
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

// (voxel, rank) pairs in history order; one wavefront per history, coalesced.
__global__ void k_emit_pairs(int64_t start, int64_t proc, const int64_t* __restrict__ off,
                             const uint32_t* __restrict__ vox, const uint32_t* __restrict__ long_cnt,
                             const uint32_t* __restrict__ pos, const uint32_t* __restrict__ rank,
                             uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; i < proc; i += n_waves) {
        const uint32_t c = long_cnt[i];
        const int64_t base = off[start + i];
        const uint32_t p = pos[i];
        const uint32_t rk = rank[i];
        for (uint32_t j = lane; j < c; j += kWave) {
            keys[p + j] = vox[base + j];
            vals[p + j] = rk;
        }
    }
}

// per rank k of a block: the row descriptor of the history visited at k
__global__ void k_row_desc(int64_t start, int64_t proc, const uint32_t* __restrict__ perm,
                           const int64_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                           const double* __restrict__ chord, const float* __restrict__ wepl,
                           RowDesc* __restrict__ desc)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= proc) return;
    const int64_t p = start + perm[start + k];
    RowDesc d;
    d.off = off[p];
    d.chord = chord[p];
    d.cnt = cnt[p];
    d.wepl = wepl[p];
    desc[start + k] = d;
}

// per history: u_i (DS.cu:9321-9380 with the list of find_MLP_path_GPU_tabulated).
// G lanes per history (G | 64), histories taken in rank order: lane l sums entries l, l+G, l+2G, ...
// with kUnroll list loads and then kUnroll gathers in flight, then the G partial sums are combined
// with a butterfly.
constexpr int kUnroll = 8;

template <int G>
__global__ void __launch_bounds__(kBlock) k_drop_rows(int64_t start, int64_t proc, uint32_t per,
                                                       const RowDesc* __restrict__ desc,
                                                       const uint32_t* __restrict__ vox,
                                                       const float* __restrict__ x, double lambda,
                                                       double* __restrict__ r)
{
    const int lane = threadIdx.x & (G - 1);
    const int64_t k = (int64_t)slab_index(per) * (kBlock / G) + threadIdx.x / G;
    if (k >= proc) return;  // the G lanes of a history leave together
    const RowDesc d = desc[start + k];
    const uint32_t* __restrict__ lst = vox + d.off;
    const int cnt = (int)d.cnt;
    double part = 0.0;
    for (int j = lane; j < cnt; j += kUnroll * G) {
        uint32_t a[kUnroll];
        float xv[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; q++) a[q] = j + q * G < cnt ? lst[j + q * G] : 0u;
#pragma unroll
        for (int q = 0; q < kUnroll; q++) xv[q] = j + q * G < cnt ? x[a[q]] : 0.0f;
#pragma unroll
        for (int q = 0; q < kUnroll; q++)
            if (j + q * G < cnt) part += (double)xv[q];
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, G);
    if (lane == 0) {
        const double L = d.chord;
        const double ax = part * L;
        const double aa = (double)pct_sqf(L) * (double)cnt;
        r[k] = L * (((double)d.wepl - ax) / aa) * lambda;
    }
}

// Sum of the n = e - b updates of one voxel in history order, divided by n (the voxel's S).
// The first min(n, K) updates arrive prefetched in rv; the rest are loaded here.
template <int ACC, int K>
__device__ __forceinline__ float col_update(uint32_t b, uint32_t e, const double* rv,
                                            const uint32_t* __restrict__ t_rank, const double* __restrict__ r)
{
    const uint32_t n = e - b;
    if (ACC == PCT_ACCUM_FP64) {
        double s = rv[0];
#pragma unroll
        for (int m = 1; m < K; m++)
            if ((uint32_t)m < n) s += rv[m];
        for (uint32_t k = b + K; k < e; k++) s += r[t_rank[k]];
        return (float)(s / (double)n);
    } else if (ACC == PCT_ACCUM_FP32) {
        float s = (float)rv[0];
#pragma unroll
        for (int m = 1; m < K; m++)
            if ((uint32_t)m < n) s += (float)rv[m];
        for (uint32_t k = b + K; k < e; k++) s += (float)r[t_rank[k]];
        return s / (float)n;
    } else {  // Kahan-compensated FP32
        float s = 0.0f, c = 0.0f, y, t;
#pragma unroll
        for (int m = 0; m < K; m++)
            if ((uint32_t)m < n) {
                y = (float)rv[m] - c;
                t = s + y;
                c = (t - s) - y;
                s = t;
            }
        for (uint32_t k = b + K; k < e; k++) {
            y = (float)r[t_rank[k]] - c;
            t = s + y;
            c = (t - s) - y;
            s = t;
        }
        return s / (float)n;
    }
}

// per touched voxel: image_update_GPU (DS.cu:9398-9419) with the block's sums formed in history
// order.  A workgroup takes kBlock * V consecutive touched voxels; each lane issues the loads of
// its V voxels (segment bounds, then up to K ranks each, then their updates and x) before any
// result is needed.
template <int ACC, int V>
__global__ void __launch_bounds__(kBlock) k_drop_cols(uint32_t U, uint32_t per, const uint32_t* __restrict__ t_vox,
                                                       const uint32_t* __restrict__ t_off,
                                                       const uint32_t* __restrict__ t_rank,
                                                       const double* __restrict__ r, float* __restrict__ x)
{
    constexpr int K = V <= 2 ? 8 : 4;
    const uint32_t u0 = slab_index(per) * (uint32_t)(kBlock * V) + threadIdx.x;
    uint32_t b[V], e[V], v[V], rk[V][K];
    float xv[V];
    double rv[V][K];
#pragma unroll
    for (int q = 0; q < V; q++) {
        const uint32_t u = u0 + q * kBlock;
        const bool ok = u < U;
        b[q] = ok ? t_off[u] : 0;
        e[q] = ok ? t_off[u + 1] : 0;
        v[q] = ok ? t_vox[u] : 0;
    }
#pragma unroll
    for (int q = 0; q < V; q++)
#pragma unroll
        for (int m = 0; m < K; m++) rk[q][m] = b[q] + m < e[q] ? t_rank[b[q] + m] : 0u;
#pragma unroll
    for (int q = 0; q < V; q++) {
        xv[q] = e[q] > b[q] ? x[v[q]] : 0.0f;
#pragma unroll
        for (int m = 0; m < K; m++) rv[q][m] = b[q] + m < e[q] ? r[rk[q][m]] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < V; q++)
        if (e[q] > b[q]) x[v[q]] = xv[q] + col_update<ACC, K>(b[q], e[q], rv[q], t_rank, r);
}

inline unsigned grid_for(int64_t n, int block = kBlock)
{
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

// grid for n_logical workgroups; with XCD slabs the grid is rounded up to a multiple of 8 and
// `per` logical workgroups go to each XCD (the extra ones find no work and return).
inline unsigned slab_grid(int64_t n_logical, bool xcd, uint32_t* per)
{
    if (n_logical < 1) n_logical = 1;
    if (!xcd) {
        *per = 0;
        return (unsigned)n_logical;
    }
    const int64_t p = (n_logical + kXcds - 1) / kXcds;
    *per = (uint32_t)p;
    return (unsigned)(p * kXcds);
}

}  // namespace

// Build the per-block visiting order and transposed lists (once per list build).
int pct_drop_prepare(pct_ctx* ctx)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n_recon;
    DropPlan& D = ctx->plan;
    D.blocks.clear();
    D.ready = false;
    const bool spatial = (ctx->drop_flags & PCT_DROP_SPATIAL_ORDER) != 0;
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
    if (max_e > 0xffffffffll || max_proc > 0xffffffffll) {
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
    if (n > D.cap_perm || !D.perm || !D.desc) {
        PCT_TRY(pct_alloc(ctx, &D.perm, n));
        PCT_TRY(pct_alloc(ctx, &D.desc, n));
        D.cap_perm = n;
    }
    // scratch: pair keys/values (max_e + 1); per-history counts, positions, sort keys, ranks (max_proc + 1)
    uint32_t *keys = nullptr, *keys_sorted = nullptr, *vals = nullptr, *runs = nullptr, *lc = nullptr, *pos = nullptr;
    uint32_t *hkey = nullptr, *hkey_sorted = nullptr, *hidx = nullptr, *rank = nullptr;
    PCT_TRY(pct_alloc(ctx, &keys, max_e + 1));
    PCT_TRY(pct_alloc(ctx, &keys_sorted, max_e + 1));
    PCT_TRY(pct_alloc(ctx, &vals, max_e + 1));
    PCT_TRY(pct_alloc(ctx, &runs, max_e + 1));
    PCT_TRY(pct_alloc(ctx, &lc, max_proc + 1));
    PCT_TRY(pct_alloc(ctx, &pos, max_proc + 1));
    PCT_TRY(pct_alloc(ctx, &hkey, max_proc + 1));
    PCT_TRY(pct_alloc(ctx, &hkey_sorted, max_proc + 1));
    PCT_TRY(pct_alloc(ctx, &hidx, max_proc + 1));
    PCT_TRY(pct_alloc(ctx, &rank, max_proc + 1));
    int end_bit = 1;
    while ((1ll << end_bit) < (int64_t)P.num_voxels) end_bit++;
    int bin_bits = 1;
    while ((1ll << bin_bits) < (int64_t)P.num_bins) bin_bits++;
    size_t need = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, t1, lc, pos, (int)(max_proc + 1), s));
    PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, t2, keys, keys_sorted, vals, vals, (int)(max_e + 1),
                                                          0, end_bit, s));
    PCT_CHECK_HIP(ctx, hipcub::DeviceRunLengthEncode::Encode(nullptr, t3, keys_sorted, keys, runs, D.n_vox,
                                                             (int)(max_e + 1), s));
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, t4, runs, D.t_off, (int)(max_e + 1), s));
    PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, t5, hkey, hkey_sorted, hidx, hidx,
                                                          (int)(max_proc + 1), 0, 32, s));
    need = std::max(std::max(t1, t2), std::max(std::max(t3, t4), t5));
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
        uint32_t* perm = D.perm + bk.start;
        size_t tb = ctx->cub_tmp_bytes;
        // visiting order: perm[k] = in-block index of the history of rank k; rank = inverse
        if (spatial) {
            hipLaunchKernelGGL(k_hist_keys, dim3(grid_for(proc)), dim3(kBlock), 0, s, bk.start, proc, ctx->bin_num,
                               hkey, hidx);
            PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp, tb, hkey, hkey_sorted, hidx, perm,
                                                                  (int)proc, 0, bin_bits, s));
            hipLaunchKernelGGL(k_rank, dim3(grid_for(proc)), dim3(kBlock), 0, s, proc, perm, rank);
        } else {
            hipLaunchKernelGGL(k_iota32, dim3(grid_for(proc)), dim3(kBlock), 0, s, proc, perm, rank);
        }
        hipLaunchKernelGGL(k_row_desc, dim3(grid_for(proc)), dim3(kBlock), 0, s, bk.start, proc, D.perm,
                           ctx->list_off, ctx->list_cnt, ctx->chord, ctx->wepl, D.desc);
        if (E == 0) {
            PCT_CHECK_HIP(ctx, hipMemsetAsync(D.n_vox + bi, 0, sizeof(uint32_t), s));
            PCT_CHECK_HIP(ctx, hipMemsetAsync(t_off, 0, sizeof(uint32_t), s));
            continue;
        }
        hipLaunchKernelGGL(k_long_counts, dim3(grid_for(proc)), dim3(kBlock), 0, s, bk.start, proc, ctx->list_cnt,
                           P.min_mlp_length, lc);
        tb = ctx->cub_tmp_bytes;
        PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, tb, lc, pos, (int)proc, s));
        int64_t blocks = (proc * kWave + kBlock - 1) / kBlock;
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(k_emit_pairs, dim3((unsigned)blocks), dim3(kBlock), 0, s, bk.start, proc, ctx->list_off,
                           ctx->list_vox, lc, pos, rank, keys, vals);
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
    D.n_vox_host.assign(nb > 0 ? nb : 1, 0);
    if (nb > 0)
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(D.n_vox_host.data(), D.n_vox, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    pct_free(&keys);
    pct_free(&keys_sorted);
    pct_free(&vals);
    pct_free(&runs);
    pct_free(&lc);
    pct_free(&pos);
    pct_free(&hkey);
    pct_free(&hkey_sorted);
    pct_free(&hidx);
    pct_free(&rank);
    D.flags = ctx->drop_flags;
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
    const bool xcd = (D.flags & PCT_DROP_XCD_SLABS) != 0;
    for (size_t bi = 0; bi < D.blocks.size(); bi++) {
        const DropPlan::Block& bk = D.blocks[bi];
        const int64_t hpw = kBlock / G;  // histories per workgroup
        uint32_t rper = 0;
        const unsigned rgrid = slab_grid((bk.proc + hpw - 1) / hpw, xcd, &rper);
        if (timing) mark();
        dim3 rg(rgrid), rb(kBlock);
#define PCT_ROWS(GG)                                                                                                  \
    hipLaunchKernelGGL(k_drop_rows<GG>, rg, rb, 0, s, bk.start, bk.proc, rper, D.desc, ctx->list_vox, ctx->x, \
                       P.lambda, D.r)
        switch (G) {
        case 4: PCT_ROWS(4); break;
        case 8: PCT_ROWS(8); break;
        case 16: PCT_ROWS(16); break;
        case 32: PCT_ROWS(32); break;
        default: PCT_ROWS(64); break;
        }
#undef PCT_ROWS
        if (timing) mark();
        const uint32_t U = D.n_vox_host[bi];
        if (bk.entries > 0 && U > 0) {
            uint32_t cper = 0;
            const int64_t vpw = (int64_t)kBlock * V;  // voxels per workgroup
            const unsigned cgrid = slab_grid((U + vpw - 1) / vpw, xcd, &cper);
            const uint32_t* t_vox = D.t_vox + bk.entry_base;
            const uint32_t* t_off = D.t_off + bk.entry_base + bi;
            const uint32_t* t_rank = D.t_hist + bk.entry_base;
            dim3 cg(cgrid), cbk(kBlock);
#define PCT_COLS(A, VV) \
    hipLaunchKernelGGL((k_drop_cols<A, VV>), cg, cbk, 0, s, U, cper, t_vox, t_off, t_rank, D.r, ctx->x)
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
