// k_drop2.hip — stage (e): atomic-free DROP over per-block transposed path lists.
//
// DROP (DROP_full_tx_iteration, DS.cu:9537-9589) updates every voxel of a block by the mean of the
// updates of the block's histories through it:
//     x[v] += (sum over entries e of the block with voxel v of u_i(e)) / S[v],
//     u_i = L_i * ((b_i - L_i * sum_j x[v_ij]) / (L_i^2 n_i)) * lambda      (DS.cu:9321-9380, 9522-9536)
// The reference scatters u_i with float atomicAdd into x_update and counts S with atomicAdd
// (2 atomics per list entry).  Here the blocks' lists are transposed once, when the lists are built:
// for every voxel a block touches, the list of its histories in ascending history order (an LSD
// radix sort by (block, voxel) is stable, so equal keys keep history order; blocks are sorted in
// groups of up to 2^30 entries).  Then each iteration runs, per block,
//   k_drop_rows   G lanes per history: coalesced reads of its voxel list, gather of x, FP64
//                 reduction of a_i.x (regrouped over the lanes only where no partial sum can round,
//                 else in list order), u_i stored in r[rank];
//   k_drop_cols   one lane per touched voxel: sum of r over its histories in history order
//                 (FP32 as the reference, FP32 with Kahan compensation, or FP64), S = the number of
//                 entries, x[v] += sum / S.
// No atomics, no scratch clearing, and the sums equal those of a single-threaded run of the reference:
// the result is reproducible bit for bit.  By default the
// column kernel reads the ranks from a sliced copy of the transposed lists (column-major per 64
// touched voxels, the voxels of one wavefront), so that each rank load is one coalesced read.
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
#include <hip/hip_ext.h>
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

// The lists' entries as brick indices: the cnt[h] entries of every history h (16 lanes per history;
// capacity slots past cnt are never read).  Grid-stride: lists can hold more entries than one
// dispatch has work-items.
__global__ void k_bricks_list(int64_t n, const int64_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                              const uint32_t* __restrict__ in, uint32_t* __restrict__ out, BrickGeom g, uint32_t nv)
{
    const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 4;
    const uint32_t l = threadIdx.x & 15;
    for (int64_t h = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; h < n; h += stride) {
        const int64_t o = off[h];
        const uint32_t c = cnt[h];
        for (uint32_t j = l; j < c; j += 16) {
            const uint32_t v = in[o + j];
            out[o + j] = v < nv ? brick_of(v, g) : 0u;
        }
    }
}

// image to (to_bricks = 1) or from brick order
__global__ void k_x_bricks(int64_t nv, float* __restrict__ x, float* __restrict__ xb, BrickGeom g, int to_bricks)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const uint32_t b = brick_of((uint32_t)v, g);
    if (to_bricks) xb[b] = x[v];
    else x[v] = xb[b];
}

// entries of every raw block: sum of the counts of its processed histories with at least
// min_len voxels (one workgroup per block)
__global__ void __launch_bounds__(256) k_block_entries(int64_t BS, const int32_t* __restrict__ proc_raw,
                                                        const uint32_t* __restrict__ cnt, int min_len,
                                                        int64_t* __restrict__ out)
{
    __shared__ int64_t part[256];
    const int64_t rb = blockIdx.x, start = rb * BS, proc = proc_raw[rb];
    int64_t s = 0;
    for (int64_t i = threadIdx.x; i < proc; i += blockDim.x) {
        const uint32_t c = cnt[start + i];
        if ((int)c >= min_len) s += c;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[rb] = part[0];
}

// per raw block: the list capacity of its processed histories (an upper bound of its entries)
__global__ void k_block_caps(int64_t n_raw, int64_t BS, const int32_t* __restrict__ proc_raw,
                             const int64_t* __restrict__ off, int64_t* __restrict__ out)
{
    const int64_t rb = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (rb >= n_raw) return;
    const int64_t start = rb * BS;
    out[rb] = off[start + proc_raw[rb]] - off[start];
}

// ---- transposition of a group of blocks (once per list build) ----
// History i belongs to raw block i / BS (BS = DROP_BLOCK_SIZE) at in-block index i % BS and is
// processed when that index is below the block's proc (quirk A3 may leave a tail unprocessed).

// sort key of a history: (block in group, sinogram bin) with spatial order, (block in group, 0)
// without; unprocessed histories take the largest bin so they sort to the end of their block.
__global__ void k_group_hist_keys(int64_t h0, int64_t n, int64_t BS, int64_t rb0, const int32_t* __restrict__ proc_raw,
                                  const int32_t* __restrict__ bin_num, int spatial, int bin_bits,
                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ idx)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = h0 + t, rb = i / BS;
    const uint32_t marker = (1u << bin_bits) - 1u;
    uint32_t b = marker;
    if (i - rb * BS < proc_raw[rb]) {
        b = 0u;
        if (spatial) {
            const uint32_t bn = (uint32_t)bin_num[i];
            b = bn < marker ? bn : marker - 1u;
        }
    }
    keys[t] = ((uint32_t)(rb - rb0) << bin_bits) | b;
    idx[t] = (uint32_t)t;
}

// sorted position s -> visiting rank within the block; perm, rank and the row descriptor
__global__ void k_group_rank(int64_t h0, int64_t n, int64_t BS, const uint32_t* __restrict__ sorted_idx,
                             const int64_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                             const double* __restrict__ chord, const float* __restrict__ wepl,
                             uint32_t* __restrict__ perm, uint32_t* __restrict__ rank, RowDesc* __restrict__ desc)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t i = h0 + sorted_idx[s];
    const int64_t start = (i / BS) * BS;
    const int64_t k = h0 + s - start;  // the block's histories occupy sorted positions [start - h0, ...)
    perm[start + k] = (uint32_t)(i - start);
    rank[i] = (uint32_t)k;
    RowDesc d;
    d.off = off[i];
    d.chord = chord[i];
    d.cnt = cnt[i];
    d.wepl = wepl[i];
    desc[start + k] = d;
}

// entries a history contributes to the column sums (0 if unprocessed or shorter than MIN_MLP_LENGTH)
__global__ void k_group_long_counts(int64_t h0, int64_t n, int64_t BS, const int32_t* __restrict__ proc_raw,
                                    const uint32_t* __restrict__ cnt, int min_len, uint32_t* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = h0 + t, rb = i / BS;
    const uint32_t c = cnt[i];
    out[t] = (i - rb * BS < proc_raw[rb] && (int)c >= min_len) ? c : 0u;
}

// ((block in group, voxel), rank) pairs in history order; 16 lanes per history, coalesced
__global__ void k_group_emit(int64_t h0, int64_t n, int64_t BS, int64_t rb0, int vbits,
                             const int64_t* __restrict__ off, const uint32_t* __restrict__ vox,
                             const uint32_t* __restrict__ long_cnt, const uint32_t* __restrict__ pos,
                             const uint32_t* __restrict__ rank, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    constexpr int L = 16;  // lanes per history: four histories' loads in flight per wavefront
    const int lane = threadIdx.x & (L - 1);
    const int64_t n_groups = ((int64_t)gridDim.x * blockDim.x) / L;
    for (int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / L; t < n; t += n_groups) {
        const uint32_t c = long_cnt[t];
        if (c == 0) continue;
        const int64_t i = h0 + t;
        const uint32_t tag = (uint32_t)(i / BS - rb0) << vbits;
        const int64_t base = off[i];
        const uint32_t p = pos[t], rk = rank[i];
        for (uint32_t j = lane; j < c; j += L) {
            keys[p + j] = tag | vox[base + j];
            vals[p + j] = rk;
        }
    }
}

// runs of the sorted keys: voxel of each run, and the first run of every block of the group
__global__ void k_group_runs(const uint32_t* __restrict__ run_keys, const uint32_t* __restrict__ n_runs, int vbits,
                             uint32_t n_blocks, uint32_t* __restrict__ t_vox, int32_t* __restrict__ run_start)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= *n_runs) return;
    const uint32_t key = run_keys[u];
    const uint32_t lb = key >> vbits;
    t_vox[u] = key & ((1u << vbits) - 1u);
    // lb < n_blocks for every key k_group_emit writes; the host checks the runs it gets
    if (lb < n_blocks && (u == 0 || (run_keys[u - 1] >> vbits) != lb)) run_start[lb] = (int32_t)u;
}

// ---- bucket transposition (once per list build; PCT_DROP_BUCKETS) ----
// The same transposed lists as the sort path, built without a device-wide sort.  A bucket is a tile of
// 4 x 4 bricks (16 x 16 x 2 voxels, 512 voxel slots) of one block.  (1) Every history's entries are cut
// into runs of one tile and counted per bucket (one atomic per run); (2) after a scan the runs are
// scattered into their buckets (unordered between histories); (3) one workgroup per bucket sorts it in
// LDS by (voxel slot in the tile, history index in the block) with 8-bit LSD passes, writes the ranks in
// that order and the bucket's runs of equal voxels; (4) the runs are compacted in bucket order.  Within a
// voxel the entries end in history order, as in the sort path, so every column sum is formed in the same
// order; only the order of the runs differs (tile-major instead of brick-major), which no sum depends on.
constexpr int kTileBits = 9;  // 16 bricks x 32 voxels

struct TileGeom {
    uint32_t nbx, nby, ntbx, ntby;  // bricks and 4 x 4-brick tiles per row / column
    uint32_t nt;                     // tiles per block
};
__device__ __forceinline__ uint32_t tile_of(uint32_t v, TileGeom t)
{
    const uint32_t b = v >> 5, bx = b % t.nbx, r = b / t.nbx, by = r % t.nby, bz = r / t.nby;
    return (bz * t.ntby + (by >> 2)) * t.ntbx + (bx >> 2);
}
__device__ __forceinline__ uint32_t slot_of(uint32_t v, TileGeom t)
{
    const uint32_t b = v >> 5, bx = b % t.nbx, by = (b / t.nbx) % t.nby;
    return (((by & 3u) << 2) | (bx & 3u)) << 5 | (v & 31u);
}
__device__ __forceinline__ uint32_t voxel_of(uint32_t tile, uint32_t slot, TileGeom t)
{
    const uint32_t tbx = tile % t.ntbx, r = tile / t.ntbx, tby = r % t.ntby, bz = r / t.ntby;
    const uint32_t by = (tby << 2) | (slot >> 7), bx = (tbx << 2) | ((slot >> 5) & 3u);
    return ((bz * t.nby + by) * t.nbx + bx) << 5 | (slot & 31u);
}

// Lane l of the 16 lanes of a history takes the contiguous entries [l c / 16, (l + 1) c / 16) and cuts
// them into runs of one bucket.  COUNT: one atomic add of the run length per run.  Otherwise: the run's
// place in its bucket from one atomic on the bucket's cursor, and its (slot << hb | history index, rank)
// pairs written there.
template <bool COUNT>
__global__ void k_tr_runs(int64_t h0, int64_t n, int64_t BS, int64_t rb0, TileGeom tg, int hb,
                          const int64_t* __restrict__ off, const uint32_t* __restrict__ vox,
                          const uint32_t* __restrict__ long_cnt, const uint32_t* __restrict__ rank,
                          uint32_t* __restrict__ bucket, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    constexpr int L = 16;
    const int lane = threadIdx.x & (L - 1);
    const int64_t n_groups = ((int64_t)gridDim.x * blockDim.x) / L;
    for (int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / L; t < n; t += n_groups) {
        const uint32_t c = long_cnt[t];
        if (c == 0) continue;
        const int64_t i = h0 + t, rb = i / BS;
        const uint32_t j0 = (uint32_t)(((uint64_t)c * lane) / L), j1 = (uint32_t)(((uint64_t)c * (lane + 1)) / L);
        if (j0 >= j1) continue;
        const uint32_t* lst = vox + off[i];
        const uint32_t gbase = (uint32_t)(rb - rb0) * tg.nt;
        const uint32_t hi = (uint32_t)(i - rb * BS), rk = COUNT ? 0u : rank[i];
        uint32_t j = j0;
        while (j < j1) {
            const uint32_t g = gbase + tile_of(lst[j], tg);
            uint32_t e = j + 1;
            while (e < j1 && gbase + tile_of(lst[e], tg) == g) e++;
            if (COUNT) {
                atomicAdd(&bucket[g], e - j);
            } else {
                const uint32_t p = atomicAdd(&bucket[g], e - j);
                for (uint32_t q = j; q < e; q++) {
                    keys[p + q - j] = (slot_of(lst[q], tg) << hb) | hi;
                    vals[p + q - j] = rk;
                }
            }
            j = e;
        }
    }
}

// Buckets by size: class 0 up to 1024 entries, 1 up to 4096, 2 up to 8192; larger buckets raise *over.
__global__ void k_tr_classify(uint32_t nbk, const uint32_t* __restrict__ cnt, uint32_t* __restrict__ lists,
                              uint32_t* __restrict__ n_in, uint32_t* __restrict__ over)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nbk) return;
    const uint32_t c = cnt[g];
    if (c == 0) return;
    const int k = c <= 1024u ? 0 : c <= 4096u ? 1 : c <= 8192u ? 2 : 3;
    if (k == 3) {
        atomicOr(over, 1u);
        return;
    }
    lists[(size_t)k * nbk + atomicAdd(&n_in[k], 1u)] = g;
}

// One workgroup of W waves per bucket of at most CAP entries: LSD radix sort of the (key, rank) pairs in LDS
// with 8-bit digits over `bits` key bits (stable: wave w owns a contiguous quarter of the items and ranks
// its lanes in order), then the ranks in sorted order to t_hist and the bucket's runs of equal slots.
template <int W, int CAP>
__global__ void __launch_bounds__(W * 64) k_tr_sort(const uint32_t* __restrict__ list, int bits, int hb,
                                                   uint32_t nt, TileGeom tg,
                                                   const uint32_t* __restrict__ boff, const uint32_t* __restrict__ bcnt,
                                                   const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                   uint32_t* __restrict__ t_hist, uint32_t* __restrict__ r_vox,
                                                   uint32_t* __restrict__ r_off, uint32_t* __restrict__ b_runs)
{
    constexpr int T = W * 64;
    __shared__ uint32_t ka[CAP], va[CAP], kb[CAP], vb[CAP];
    __shared__ uint32_t cur[W][256];
    __shared__ uint32_t tot[256];
    __shared__ uint32_t wsum[W];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t g = list[blockIdx.x];
    const uint32_t n = bcnt[g], base = boff[g];
    for (uint32_t j = threadIdx.x; j < n; j += T) {
        ka[j] = keys[base + j];
        va[j] = vals[base + j];
    }
    const uint32_t s0 = (uint32_t)(((uint64_t)n * w) / W), s1 = (uint32_t)(((uint64_t)n * (w + 1)) / W);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t *src_k = ka, *src_v = va, *dst_k = kb, *dst_v = vb;
    for (int shift = 0; shift < bits; shift += 8) {
        for (int d = threadIdx.x; d < W * 256; d += T) cur[d / 256][d % 256] = 0u;
        __syncthreads();
        for (uint32_t j = s0 + lane; j < s1; j += 64) atomicAdd(&cur[w][(src_k[j] >> shift) & 255u], 1u);
        __syncthreads();
        // cursor of (wave w, digit d) = items of smaller digits + items of digit d in waves before w
        for (int d = threadIdx.x; d < 256; d += T) {
            uint32_t s = 0;
            for (int q = 0; q < W; q++) s += cur[q][d];
            tot[d] = s;
        }
        __syncthreads();
        if (w == 0) {
            uint32_t a[4], run = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                a[q] = tot[lane * 4 + q];
                run += a[q];
            }
            uint32_t incl = run;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            uint32_t ex = incl - run;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t d = lane * 4 + q;
                uint32_t c = ex;
                for (int p = 0; p < W; p++) {
                    const uint32_t h = cur[p][d];
                    cur[p][d] = c;
                    c += h;
                }
                ex += a[q];
            }
        }
        __syncthreads();
        for (uint32_t j0 = s0; j0 < s1; j0 += 64) {
            const uint32_t j = j0 + lane;
            const bool ok = j < s1;
            const uint32_t k = ok ? src_k[j] : 0u;
            const uint32_t d = (k >> shift) & 255u;
            uint64_t m = __ballot(ok);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const uint64_t bb = __ballot(ok && ((d >> b) & 1u));
                m &= ((d >> b) & 1u) ? bb : ~bb;
            }
            if (ok) {
                const uint32_t p = cur[w][d] + (uint32_t)__popcll(m & lt);
                dst_k[p] = k;
                dst_v[p] = src_v[j];
            }
            __builtin_amdgcn_wave_barrier();
            if (ok && (m & lt) == 0ull) cur[w][d] += (uint32_t)__popcll(m);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        uint32_t* tk = src_k; src_k = dst_k; dst_k = tk;
        uint32_t* tv = src_v; src_v = dst_v; dst_v = tv;
    }
    // ranks in sorted order; runs of equal slots (a run starts at item 0 and wherever the slot changes)
    const uint32_t tile = g % nt;
    uint32_t runs_before = 0;
    for (uint32_t r0 = 0; r0 < n; r0 += T) {
        const uint32_t j = r0 + threadIdx.x;
        const bool ok = j < n;
        bool head = false;
        uint32_t slot = 0;
        if (ok) {
            t_hist[base + j] = src_v[j];
            slot = src_k[j] >> hb;
            head = (j == 0) || ((src_k[j - 1] >> hb) != slot);
        }
        const uint64_t hm = __ballot(head);
        if (lane == 0) wsum[w] = (uint32_t)__popcll(hm);
        __syncthreads();
        uint32_t before = runs_before;
        for (int q = 0; q < w; q++) before += wsum[q];
        uint32_t all = runs_before;
        for (int q = 0; q < W; q++) all += wsum[q];
        if (head) {
            const uint32_t u = before + (uint32_t)__popcll(hm & lt);
            r_vox[base + u] = voxel_of(tile, slot, tg);
            r_off[base + u] = base + j;
        }
        runs_before = all;
        __syncthreads();
    }
    if (threadIdx.x == 0) b_runs[g] = runs_before;
}

// Runs of every bucket, in bucket order, into the group's t_vox / t_off (one wavefront per bucket), and
// the total entries as the closing offset.
__global__ void k_tr_compact(uint32_t nbk, const uint32_t* __restrict__ boff, const uint32_t* __restrict__ b_runs,
                             const uint32_t* __restrict__ b_roff, const uint32_t* __restrict__ r_vox,
                             const uint32_t* __restrict__ r_off, uint32_t* __restrict__ t_vox,
                             uint32_t* __restrict__ t_off)
{
    const uint32_t g = (uint32_t)(((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (g > nbk) return;
    if (g == nbk) {
        if (lane == 0) t_off[b_roff[nbk]] = boff[nbk];
        return;
    }
    const uint32_t nr = b_runs[g], src = boff[g], dst = b_roff[g];
    for (uint32_t k = lane; k < nr; k += 64) {
        t_vox[dst + k] = r_vox[src + k];
        t_off[dst + k] = r_off[src + k];
    }
}

// first run of every block of the group (-1 when it has none)
__global__ void k_tr_run_start(uint32_t n_blocks, uint32_t nt, const uint32_t* __restrict__ b_roff,
                               int32_t* __restrict__ run_start)
{
    const uint32_t lb = blockIdx.x * blockDim.x + threadIdx.x;
    if (lb >= n_blocks) return;
    const uint32_t a = b_roff[lb * nt], b = b_roff[(lb + 1) * nt];
    run_start[lb] = b > a ? (int32_t)a : -1;
}

// ---- sliced columns (once per list build, after the transposition) ----
// Slice s of all blocks' slices belongs to the block k with sb[k].slice_base <= s < sb[k + 1].slice_base
// (sb holds nb + 1 entries, the last one the total).
__device__ __forceinline__ int64_t slice_block(const SliceBlock* __restrict__ sb, int64_t nb, int64_t s)
{
    int64_t lo = 0, hi = nb;  // sb[lo].slice_base <= s < sb[hi].slice_base
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (sb[mid].slice_base <= s) lo = mid;
        else hi = mid;
    }
    return lo;
}

// per slice (one wavefront; workgroup g takes slices [g chunk, (g + 1) chunk)): 64 x the largest
// entry count of its runs
__global__ void __launch_bounds__(kWave) k_slice_width(int64_t NS, int64_t chunk, int64_t nb,
                                                        const SliceBlock* __restrict__ sb,
                                                        const uint32_t* __restrict__ t_off, uint64_t* __restrict__ w)
{
    const int64_t s0 = (int64_t)blockIdx.x * chunk, s1 = s0 + chunk < NS ? s0 + chunk : NS;
    int64_t k = s0 < s1 ? slice_block(sb, nb, s0) : 0;
    for (int64_t s = s0; s < s1; s++) {
        while (sb[k + 1].slice_base <= s) k++;
        const int64_t j = (s - sb[k].slice_base) * kWave + threadIdx.x;
        uint32_t c = 0;
        if (j < (int64_t)sb[k].n_vox) {
            const int64_t u = sb[k].run + j;
            c = t_off[u + 1] - t_off[u];
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const uint32_t y = (uint32_t)__shfl_xor((int)c, o, kWave);
            c = y > c ? y : c;
        }
        if (threadIdx.x == 0) w[s] = (uint64_t)c * kWave;
    }
}

// per slice: lane l copies the ranks of the slice's run l to slots sell_off[s] + l + 64 m.  The
// slice's entries are one contiguous range of t_hist; up to kFillStage of them are first loaded
// into LDS with coalesced reads, so that both the reads and the column-major writes coalesce.
constexpr int kFillStage = 1024;

__global__ void __launch_bounds__(kWave) k_slice_fill(int64_t NS, int64_t chunk, int64_t nb,
                                                       const SliceBlock* __restrict__ sb,
                                                       const uint32_t* __restrict__ t_off,
                                                       const uint32_t* __restrict__ t_hist,
                                                       const uint64_t* __restrict__ sell_off, uint32_t* __restrict__ sell)
{
    __shared__ uint32_t stage[kFillStage];
    const int lane = threadIdx.x;
    const int64_t s0 = (int64_t)blockIdx.x * chunk, s1 = s0 + chunk < NS ? s0 + chunk : NS;
    int64_t k = s0 < s1 ? slice_block(sb, nb, s0) : 0;
    for (int64_t s = s0; s < s1; s++) {
        while (sb[k + 1].slice_base <= s) k++;
        const int64_t j0 = (s - sb[k].slice_base) * kWave;
        const int64_t left = (int64_t)sb[k].n_vox - j0;
        const int n_runs = left < kWave ? (int)left : kWave;  // >= 1
        const int64_t u0 = sb[k].run + j0;
        const uint32_t E0 = t_off[u0], E1 = t_off[u0 + n_runs];
        const uint32_t b = lane < n_runs ? t_off[u0 + lane] : E1, c = lane < n_runs ? t_off[u0 + lane + 1] - b : 0u;
        const uint32_t* __restrict__ src = t_hist + sb[k].entry_base;
        uint32_t* __restrict__ dst = sell + sell_off[s] + lane;
        if (E1 - E0 <= (uint32_t)kFillStage) {
            for (uint32_t i = lane; i < E1 - E0; i += kWave) stage[i] = src[E0 + i];
            __syncthreads();
            for (uint32_t m = 0; m < c; m++) dst[(int64_t)m * kWave] = stage[b - E0 + m];
            __syncthreads();
        } else {
            for (uint32_t m = 0; m < c; m++) dst[(int64_t)m * kWave] = src[b + m];
        }
    }
}

constexpr int kUnroll = 8;

// Bounds of a nonzero float term of a row sum: |f| < 2^top, and f is a multiple of 2^lsb (its lowest set
// significand bit).  Zero terms leave the bounds alone.
__device__ __forceinline__ void term_bounds(float f, int& top, int& lsb)
{
    const uint32_t b = __float_as_uint(f) & 0x7fffffffu;
    if (!b) return;
    const int ef = (int)(b >> 23);
    uint32_t m = b & 0x7fffffu;
    int t, l;
    if (ef == 0) {  // subnormal: m * 2^-149
        t = -149 + 32 - __clz(m);
        l = -149 + __ffs(m) - 1;
    } else {        // (2^23 + m) * 2^(ef - 150); Inf and NaN (ef = 255) get top 129 and force the ordered sum
        m |= 0x800000u;
        t = ef - 126;
        l = ef - 150 + __ffs(m) - 1;
    }
    top = t > top ? t : top;
    lsb = l < lsb ? l : lsb;
}

// a_i.x / L_i of one history: the sum of x over its list with the row kernel's lanes and loads
// (G lanes, kUnroll list loads and then kUnroll gathers in flight per lane, lane l summing entries
// l, l + G, l + 2G, ... in order, then a butterfly over the G partial sums).  The reference sums the
// list in order (DS.cu:9245), and the regrouped sum equals the ordered one whenever no partial sum rounds:
// with every term a multiple of 2^lsb and below 2^top in magnitude, any partial sum of at most cnt terms
// is a multiple of 2^lsb below 2^(top + ceil(log2 cnt)), exact in double when top + ceil(log2 cnt) - lsb
// <= 53.  The G lanes of a failing row gather its list again in chunks and all add the chunk in list order
// from the lanes' registers.  The sum is valid in lane 0.  (Tracking the bounds this way costs about 1 us
// per DROP block; comparing raw bit patterns instead measured 7 us.)
template <int G>
__device__ inline double row_sum(const uint32_t* __restrict__ lst, int cnt, const float* __restrict__ x, int lane)
{
    double part = 0.0;
    int top = -1000, lsb = 1000;
    for (int j = lane; j < cnt; j += kUnroll * G) {
        uint32_t a[kUnroll];
        float xv[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; q++) a[q] = j + q * G < cnt ? lst[j + q * G] : 0u;
#pragma unroll
        for (int q = 0; q < kUnroll; q++) xv[q] = j + q * G < cnt ? x[a[q]] : 0.0f;
#pragma unroll
        for (int q = 0; q < kUnroll; q++)
            if (j + q * G < cnt) {
                part += (double)xv[q];
                term_bounds(xv[q], top, lsb);
            }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        part += __shfl_xor(part, o, G);
        const int t = __shfl_xor(top, o, G), l = __shfl_xor(lsb, o, G);
        top = t > top ? t : top;
        lsb = l < lsb ? l : lsb;
    }
    const int lg = cnt > 1 ? 32 - __clz((uint32_t)cnt - 1u) : 0;  // ceil(log2 cnt)
    if (top + lg - lsb > 53) {  // the same in the G lanes after the butterfly; all-zero rows (top -1000) never fail
        // the ordered sum: chunks of kUnroll * G entries gathered as above (entry j + q G + l in lane l), then
        // added in list order by every lane from the lanes' registers
        double s = 0.0;
        for (int j = 0; j < cnt; j += kUnroll * G) {
            float xv[kUnroll];
#pragma unroll
            for (int q = 0; q < kUnroll; q++) {
                const int e = j + q * G + lane;
                xv[q] = e < cnt ? x[lst[e]] : 0.0f;
            }
#pragma unroll
            for (int q = 0; q < kUnroll; q++)
                for (int l = 0; l < G; l++) {
                    const float v = __shfl(xv[q], l, G);
                    if (j + q * G + l < cnt) s += (double)v;
                }
        }
        part = s;
    }
    return part;
}

// per history: u_i (DS.cu:9321-9380 with the list of find_MLP_path_GPU_tabulated).
// G lanes per history (G | 64), histories taken in rank order.
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
    const int cnt = (int)d.cnt;
    const double part = row_sum<G>(vox + d.off, cnt, x, lane);
    if (lane == 0) {
        const double L = d.chord;
        const double ax = part * L;
        const double aa = (double)pct_sqf(L) * (double)cnt;
        r[k] = L * (((double)d.wepl - ax) / aa) * lambda;
    }
}

// per history: (b_i - L_i sum_j x[v_ij])^2 for the data residual (0 below MIN_MLP_LENGTH), G lanes
// per history as in the row kernel.  Grid-stride over histories: any history count fits one launch.
template <int G>
__global__ void __launch_bounds__(kBlock) k_residual(int64_t n, const int64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ cnt_arr,
                                                      const uint32_t* __restrict__ vox, const double* __restrict__ chord,
                                                      const float* __restrict__ wepl, const float* __restrict__ x,
                                                      int min_len, double* __restrict__ r2)
{
    const int lane = threadIdx.x & (G - 1);
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / G);
    for (int64_t p = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; p < n; p += stride) {
        const int cnt = (int)cnt_arr[p];   // the G lanes of a history loop together
        const double part = row_sum<G>(vox + off[p], cnt, x, lane);
        if (lane == 0) {
            const double r = (double)wepl[p] - chord[p] * part;
            r2[p] = cnt >= min_len ? r * r : 0.0;
        }
    }
}

// k_residual over the DROP plan's row descriptors: every history once, in the blocks' visiting order
// (neighbouring parallel paths share a wavefront, as in k_drop_rows).  Grid-stride over histories.
template <int G>
__global__ void __launch_bounds__(kBlock) k_residual_desc(int64_t n, const RowDesc* __restrict__ desc,
                                                           const uint32_t* __restrict__ vox, const float* __restrict__ x,
                                                           int min_len, double* __restrict__ r2)
{
    const int lane = threadIdx.x & (G - 1);
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / G);
    for (int64_t p = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; p < n; p += stride) {
        const RowDesc d = desc[p];
        const int cnt = (int)d.cnt;
        const double part = row_sum<G>(vox + d.off, cnt, x, lane);
        if (lane == 0) {
            const double r = (double)d.wepl - d.chord * part;
            r2[p] = cnt >= min_len ? r * r : 0.0;
        }
    }
}

// Running per-voxel sum in history order: FP64 (s = u_1 + u_2 + ...), FP32 (the reference's float
// sums), or FP32 with Kahan compensation (from s = c = 0).
template <int ACC>
struct ColAcc {
    double s64 = 0.0;
    float s32 = 0.0f, c32 = 0.0f;
    bool first = true;
    __device__ __forceinline__ void add(double rv)
    {
        if (ACC == PCT_ACCUM_FP64) {
            s64 = first ? rv : s64 + rv;
        } else if (ACC == PCT_ACCUM_FP32) {
            s32 = first ? (float)rv : s32 + (float)rv;
        } else {
            const float y = (float)rv - c32;
            const float t = s32 + y;
            c32 = (t - s32) - y;
            s32 = t;
        }
        first = false;
    }
    __device__ __forceinline__ float mean(uint32_t n) const
    {
        return ACC == PCT_ACCUM_FP64 ? (float)(s64 / (double)n) : s32 / (float)n;
    }
};

// Sum of the n = e - b updates of one voxel in history order, divided by n (the voxel's S).
// The first min(n, K) updates arrive prefetched in rv; the rest are loaded K at a time (K rank
// loads, then K gathers) so that a long column costs two memory round trips per K entries.
template <int ACC, int K>
__device__ __forceinline__ float col_update(uint32_t b, uint32_t e, const double* rv,
                                            const uint32_t* __restrict__ t_rank, const double* __restrict__ r)
{
    ColAcc<ACC> a;
#pragma unroll
    for (int m = 0; m < K; m++)
        if (b + m < e) a.add(rv[m]);
    for (uint32_t c0 = b + K; c0 < e; c0 += K) {
        uint32_t rk[K];
        double rw[K];
#pragma unroll
        for (int m = 0; m < K; m++) rk[m] = c0 + m < e ? t_rank[c0 + m] : 0u;
#pragma unroll
        for (int m = 0; m < K; m++) rw[m] = c0 + m < e ? r[rk[m]] : 0.0;
#pragma unroll
        for (int m = 0; m < K; m++)
            if (c0 + m < e) a.add(rw[m]);
    }
    return a.mean(e - b);
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

// col_update over a sliced column: entry m's rank at sp[64 m].
template <int ACC, int K>
__device__ __forceinline__ float col_update_sliced(uint32_t n, const double* rv, const uint32_t* __restrict__ sp,
                                                   const double* __restrict__ r)
{
    ColAcc<ACC> a;
#pragma unroll
    for (int m = 0; m < K; m++)
        if ((uint32_t)m < n) a.add(rv[m]);
    for (uint32_t c0 = K; c0 < n; c0 += K) {
        uint32_t rk[K];
        double rw[K];
#pragma unroll
        for (int m = 0; m < K; m++) rk[m] = c0 + m < n ? sp[(int64_t)(c0 + m) * kWave] : 0u;
#pragma unroll
        for (int m = 0; m < K; m++) rw[m] = c0 + m < n ? r[rk[m]] : 0.0;
#pragma unroll
        for (int m = 0; m < K; m++)
            if (c0 + m < n) a.add(rw[m]);
    }
    return a.mean(n);
}

// k_drop_cols over sliced columns (PCT_DROP_COLS_SLICED): the same sums in the same order, with
// each rank load one coalesced 256-byte read of the wavefront's slice.  sell_off points at the
// block's first slice; lane l of the wavefront that takes runs [64 j, 64 j + 64) reads slice j.
template <int ACC, int V>
__global__ void __launch_bounds__(kBlock) k_drop_cols_sliced(uint32_t U, uint32_t per,
                                                              const uint32_t* __restrict__ t_vox,
                                                              const uint32_t* __restrict__ t_off,
                                                              const uint64_t* __restrict__ sell_off,
                                                              const uint32_t* __restrict__ sell,
                                                              const double* __restrict__ r, float* __restrict__ x)
{
    constexpr int K = V <= 2 ? 8 : 4;
    const uint32_t u0 = slab_index(per) * (uint32_t)(kBlock * V) + threadIdx.x;
    uint32_t n[V], v[V], rk[V][K];
    const uint32_t* sp[V];
    float xv[V];
    double rv[V][K];
#pragma unroll
    for (int q = 0; q < V; q++) {
        const uint32_t u = u0 + q * kBlock;
        const bool ok = u < U;
        const uint32_t b = ok ? t_off[u] : 0u, e = ok ? t_off[u + 1] : 0u;
        n[q] = e - b;
        v[q] = ok ? t_vox[u] : 0u;
        sp[q] = sell + (ok ? sell_off[u / kWave] : 0ull) + (u % kWave);
    }
#pragma unroll
    for (int q = 0; q < V; q++)
#pragma unroll
        for (int m = 0; m < K; m++) rk[q][m] = (uint32_t)m < n[q] ? sp[q][m * kWave] : 0u;
#pragma unroll
    for (int q = 0; q < V; q++) {
        xv[q] = n[q] > 0 ? x[v[q]] : 0.0f;
#pragma unroll
        for (int m = 0; m < K; m++) rv[q][m] = (uint32_t)m < n[q] ? r[rk[q][m]] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < V; q++)
        if (n[q] > 0) x[v[q]] = xv[q] + col_update_sliced<ACC, K>(n[q], rv[q], sp[q], r);
}

// Column update with the entries staged per wavefront (PCT_DROP_COLS_LDS).  A one-wave workgroup
// takes 64 consecutive touched voxels; their entries are one contiguous range of t_rank.  All 64
// lanes load that range coalesced and gather the updates r into LDS (kStage entries per round);
// then each lane sums its own voxel's segment from LDS in history order, exactly as col_update.
constexpr int kStage = 1024;

template <int ACC>
__global__ void __launch_bounds__(kWave) k_drop_cols_lds(uint32_t U, const uint32_t* __restrict__ t_vox,
                                                          const uint32_t* __restrict__ t_off,
                                                          const uint32_t* __restrict__ t_rank,
                                                          const double* __restrict__ r, float* __restrict__ x)
{
    __shared__ double stage[kStage];
    const int lane = threadIdx.x;
    const uint32_t u0 = blockIdx.x * kWave;
    const uint32_t u = u0 + lane;
    const bool ok = u < U;
    const uint32_t b = ok ? t_off[u] : 0u, e = ok ? t_off[u + 1] : 0u;
    const uint32_t v = ok ? t_vox[u] : 0u;
    const float xv = ok && e > b ? x[v] : 0.0f;
    const uint32_t last = u0 + kWave < U ? u0 + kWave : U;
    const uint32_t E0 = t_off[u0], E1 = t_off[last];
    ColAcc<ACC> a;
    for (uint32_t R0 = E0; R0 < E1; R0 += kStage) {
        const uint32_t R1 = R0 + kStage < E1 ? R0 + kStage : E1;
        for (uint32_t j = R0 + lane; j < R1; j += kWave) stage[j - R0] = r[t_rank[j]];
        __syncthreads();
        const uint32_t j0 = b > R0 ? b : R0, j1 = e < R1 ? e : R1;
        for (uint32_t j = j0; j < j1; j++) a.add(stage[j - R0]);
        __syncthreads();
    }
    if (ok && e > b) x[v] = xv + a.mean(e - b);
}

inline unsigned grid_for(int64_t n, int block = kBlock) { return pct_grid_for(n, block); }

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

namespace {

int bits_for(int64_t n)  // smallest b with 2^b >= n
{
    int b = 0;
    while ((1ll << b) < n) b++;
    return b;
}

template <class T>
int grow(pct_ctx* ctx, T** p, int64_t* cap, int64_t need)
{
    if (need <= *cap && *p) return PCT_OK;
    PCT_TRY(pct_alloc(ctx, p, need));
    *cap = need;
    return PCT_OK;
}

// Sliced columns of every block (PCT_DROP_COLS_SLICED), from the transposed runs: slice widths,
// their prefix sum (slots), then the ranks copied column-major.  Sets each block's slice_base.
int build_slices(pct_ctx* ctx, DropPlan& D)
{
    hipStream_t s = ctx->stream;
    const int64_t nb = (int64_t)D.blocks.size();
    std::vector<SliceBlock> sb(nb + 1);
    int64_t NS = 0;
    for (int64_t k = 0; k < nb; k++) {
        DropPlan::Block& bk = D.blocks[k];
        bk.slice_base = NS;
        sb[k] = SliceBlock{NS, bk.run_base, bk.rank_base, bk.n_vox, 0u};
        NS += ((int64_t)bk.n_vox + kWave - 1) / kWave;
    }
    sb[nb] = SliceBlock{NS, 0, 0, 0u, 0u};
    PCT_TRY(grow(ctx, &D.slice_blocks, &D.cap_slice_blocks, nb + 1));
    PCT_TRY(grow(ctx, &D.sell_off, &D.cap_slices, NS + 1));
    PCT_TRY(grow(ctx, &D.slice_w, &D.cap_slice_w, NS + 1));
    uint64_t* w = D.slice_w;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(D.slice_blocks, sb.data(), sb.size() * sizeof(SliceBlock), hipMemcpyHostToDevice, s));
    PCT_CHECK_HIP(ctx, hipMemsetAsync(w + NS, 0, sizeof(uint64_t), s));
    // contiguous chunks of slices per workgroup: one block search per workgroup
    const int64_t chunk = std::max<int64_t>(1, (NS + 65535) / 65536);
    const unsigned grid = (unsigned)std::max<int64_t>(1, (NS + chunk - 1) / chunk);
    if (NS > 0)
        hipLaunchKernelGGL(k_slice_width, dim3(grid), dim3(kWave), 0, s, NS, chunk, nb, D.slice_blocks, D.t_off, w);
    size_t tb = 0;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, w, D.sell_off, (int)(NS + 1), s));
    if (tb > ctx->cub_tmp_bytes) {
        PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, tb));
        ctx->cub_tmp_bytes = tb;
    }
    tb = ctx->cub_tmp_bytes;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, tb, w, D.sell_off, (int)(NS + 1), s));
    uint64_t total = 0;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&total, D.sell_off + NS, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    PCT_TRY(grow(ctx, &D.sell, &D.cap_sell, (int64_t)total + 1));
    if (NS > 0)
        hipLaunchKernelGGL(k_slice_fill, dim3(grid), dim3(kWave), 0, s, NS, chunk, nb, D.slice_blocks, D.t_off,
                           D.t_hist, D.sell_off, D.sell);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    return PCT_OK;
}

}  // namespace

void pct_drop_plan_free(DropPlan* D)
{
    pct_free(&D->t_vox);
    pct_free(&D->t_off);
    pct_free(&D->t_hist);
    pct_free(&D->perm);
    pct_free(&D->rank);
    pct_free(&D->desc);
    pct_free(&D->r);
    pct_free(&D->keys);
    pct_free(&D->keys_sorted);
    pct_free(&D->vals);
    pct_free(&D->runs);
    pct_free(&D->n_runs);
    pct_free(&D->hkey);
    pct_free(&D->hkey_sorted);
    pct_free(&D->hidx);
    pct_free(&D->hidx_sorted);
    pct_free(&D->lc);
    pct_free(&D->pos);
    pct_free(&D->run_start);
    pct_free(&D->sell);
    pct_free(&D->sell_off);
    pct_free(&D->slice_w);
    pct_free(&D->slice_blocks);
    D->cap_sell = D->cap_slices = D->cap_slice_w = D->cap_slice_blocks = 0;
    pct_free(&D->bk_cnt);
    pct_free(&D->bk_off);
    pct_free(&D->bk_cur);
    pct_free(&D->bk_runs);
    pct_free(&D->bk_roff);
    pct_free(&D->bk_list);
    pct_free(&D->bk_misc);
    D->cap_bk = 0;
    pct_free(&D->vox_b);
    pct_free(&D->x_b);
    D->cap_vox_b = D->cap_x_b = 0;
    D->vox_b_ready = false;
    D->cap_entries = D->cap_off = D->cap_r = D->cap_perm = D->cap_group_e = D->cap_group_h = D->cap_group_b = 0;
    D->ready = false;
}

namespace {

// The bucket transposition of one group (see k_tr_runs).  *done is false when a bucket holds more than 8192
// entries; the caller then transposes the group with the sort.
int transpose_group_buckets(pct_ctx* ctx, DropPlan& D, const DropPlan::Group& g, int64_t gi, int64_t run_off,
                            int64_t BS, const TileGeom& tg, const uint32_t* lvox, uint32_t* n_runs_out, bool* done)
{
    hipStream_t s = ctx->stream;
    const int64_t h0 = g.hist_begin, H = g.hist_end - g.hist_begin, E = g.entries, rb0 = h0 / BS;
    const uint32_t nbk = (uint32_t)(g.n_blocks * (int64_t)tg.nt);
    const int hb = bits_for(BS);
    *done = false;
    PCT_CHECK_HIP(ctx, hipMemsetAsync(D.bk_cnt, 0, (size_t)(nbk + 1) * sizeof(uint32_t), s));
    PCT_CHECK_HIP(ctx, hipMemsetAsync(D.bk_misc, 0, 4 * sizeof(uint32_t), s));
    int64_t wg = (H * 16 + kBlock - 1) / kBlock;
    if (wg > 65536) wg = 65536;
    hipLaunchKernelGGL(k_tr_runs<true>, dim3((unsigned)wg), dim3(kBlock), 0, s, h0, H, BS, rb0, tg, hb, ctx->list_off,
                       lvox, D.lc, D.rank, D.bk_cnt, (uint32_t*)nullptr, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_tr_classify, dim3(grid_for(nbk)), dim3(kBlock), 0, s, nbk, D.bk_cnt, D.bk_list, D.bk_misc,
                       D.bk_misc + 3);
    uint32_t misc[4] = {0, 0, 0, 0};
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(misc, D.bk_misc, sizeof(misc), hipMemcpyDeviceToHost, s));
    size_t tb = ctx->cub_tmp_bytes;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, tb, D.bk_cnt, D.bk_off, (int)(nbk + 1), s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (misc[3]) return PCT_OK;  // a bucket beyond 8192 entries: the sort path
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(D.bk_cur, D.bk_off, (size_t)nbk * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_tr_runs<false>, dim3((unsigned)wg), dim3(kBlock), 0, s, h0, H, BS, rb0, tg, hb, ctx->list_off,
                       lvox, D.lc, D.rank, D.bk_cur, D.keys, D.vals);
    const int bits = kTileBits + hb;
    uint32_t* t_hist = D.t_hist + g.entry_base;
    PCT_CHECK_HIP(ctx, hipMemsetAsync(D.bk_runs, 0, (size_t)(nbk + 1) * sizeof(uint32_t), s));
    if (misc[0])
        hipLaunchKernelGGL((k_tr_sort<1, 1024>), dim3(misc[0]), dim3(64), 0, s, D.bk_list, bits, hb, tg.nt, tg,
                           D.bk_off, D.bk_cnt, D.keys, D.vals, t_hist, D.keys_sorted, D.runs, D.bk_runs);
    if (misc[1])
        hipLaunchKernelGGL((k_tr_sort<4, 4096>), dim3(misc[1]), dim3(256), 0, s, D.bk_list + nbk, bits, hb, tg.nt, tg,
                           D.bk_off, D.bk_cnt, D.keys, D.vals, t_hist, D.keys_sorted, D.runs, D.bk_runs);
    if (misc[2])
        hipLaunchKernelGGL((k_tr_sort<8, 8192>), dim3(misc[2]), dim3(512), 0, s, D.bk_list + 2 * (size_t)nbk, bits, hb,
                           tg.nt, tg, D.bk_off, D.bk_cnt, D.keys, D.vals, t_hist, D.keys_sorted, D.runs, D.bk_runs);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    tb = ctx->cub_tmp_bytes;
    PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, tb, D.bk_runs, D.bk_roff, (int)(nbk + 1), s));
    const int64_t cw = ((int64_t)nbk + 1) * 64;
    hipLaunchKernelGGL(k_tr_compact, dim3((unsigned)((cw + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nbk, D.bk_off,
                       D.bk_runs, D.bk_roff, D.keys_sorted, D.runs, D.t_vox + run_off + gi, D.t_off + run_off + gi);
    hipLaunchKernelGGL(k_tr_run_start, dim3(grid_for(g.n_blocks)), dim3(kBlock), 0, s, (uint32_t)g.n_blocks, tg.nt,
                       D.bk_roff, D.run_start);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    uint32_t nr = 0, total = 0;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&nr, D.bk_roff + nbk, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&total, D.bk_off + nbk, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if ((int64_t)total != E || nr < 1 || (int64_t)nr > E) {
        ctx->err = "DROP bucket transposition: entry or run count out of range in a group";
        return PCT_E_LIMIT;
    }
    *n_runs_out = nr;
    *done = true;
    return PCT_OK;
}

}  // namespace

// Build the visiting order and the transposed lists of every block (once per list build).  Blocks
// are processed in groups of up to 2^30 list entries: per group one radix sort of the histories by
// (block, sinogram bin) gives the visiting order, and one radix sort of the (block, voxel) keys of
// all its list entries, a run-length encoding and a scan give every block's touched voxels with
// their histories in history order (LSD radix sorts are stable).
int pct_drop_prepare(pct_ctx* ctx, const MlpLaunch* mlp)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n_recon;
    const bool pipe = mlp != nullptr;  // the lists are built group by group below
    const int64_t BS = P.drop_block_size;
    DropPlan& D = ctx->plan;
    D.blocks.clear();
    D.groups.clear();
    D.ready = false;
    const bool spatial = (ctx->drop_flags & PCT_DROP_SPATIAL_ORDER) != 0;
    const bool bricks = (ctx->drop_flags & PCT_DROP_BRICKS) != 0;
    int64_t nv_index = P.num_voxels;  // index space of the voxel keys
    if (bricks) {
        D.nbx = (P.columns + 3) / 4;
        D.nby = (P.rows + 3) / 4;
        D.nv_b = (int64_t)D.nbx * D.nby * ((P.slices + 1) / 2) * 32;
        nv_index = D.nv_b;
    }
    const int vbits = bits_for(nv_index);
    const int bin_bits = bits_for((int64_t)P.num_bins + 1);
    if (vbits > 31 || bin_bits > 31) {
        ctx->err = "DROP transposition: voxel or bin index wider than 31 bits";
        return PCT_E_LIMIT;
    }
    const int block_bits = 32 - (vbits > bin_bits ? vbits : bin_bits);
    const int64_t max_group_blocks = 1ll << block_bits;
    const int64_t kGroupEntries = ctx->drop_group_entries, kGroupHistories = 1ll << 30;

    // block ranges (DS.cu:9558 floors the processed count) and their entry counts (on the device)
    const int64_t n_raw = (n + BS - 1) / BS;
    std::vector<int32_t> proc_raw(n_raw > 0 ? n_raw : 1, 0);
    for (int64_t rb = 0; rb < n_raw; rb++) {
        const int64_t m = n - rb * BS < BS ? n - rb * BS : BS;
        const int64_t proc = P.quirk_a3_drop_floor ? (m / P.histories_per_block) * P.histories_per_block : m;
        proc_raw[rb] = proc > 0 ? (int32_t)proc : 0;
    }
    pct_tmp<int32_t> proc_buf;
    pct_tmp<int64_t> e_buf;
    PCT_TRY(pct_alloc(ctx, &proc_buf.p, n_raw + 1));
    PCT_TRY(pct_alloc(ctx, &e_buf.p, n_raw + 1));
    const int32_t* proc_dev = proc_buf.p;
    int64_t* e_dev = e_buf.p;
    std::vector<int64_t> e_raw(n_raw > 0 ? n_raw : 1, 0);
    if (n_raw > 0) {
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(proc_buf.p, proc_raw.data(), n_raw * sizeof(int32_t), hipMemcpyHostToDevice, s));
        if (pipe)  // list capacities bound the entries until the lists exist: groups and allocations use them
            hipLaunchKernelGGL(k_block_caps, dim3(grid_for(n_raw)), dim3(kBlock), 0, s, n_raw, BS, proc_dev,
                               ctx->list_off, e_dev);
        else
            hipLaunchKernelGGL(k_block_entries, dim3((unsigned)n_raw), dim3(256), 0, s, BS, proc_dev, ctx->list_cnt,
                               P.min_mlp_length, e_dev);
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(e_raw.data(), e_dev, n_raw * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    }
    int64_t total = 0, max_proc = 0;
    for (int64_t rb = 0; rb < n_raw; rb++) {
        const int64_t start = rb * BS;
        const int64_t m = n - start < BS ? n - start : BS;
        const int64_t proc = proc_raw[rb];
        if (proc <= 0) continue;
        const int64_t e = e_raw[rb];
        if (e >= (1ll << 30)) {  // a group holds at least one whole block
            ctx->err = "DROP block with 2^30 or more list entries";
            return PCT_E_LIMIT;
        }
        DropPlan::Block bk{};
        bk.start = start;
        bk.proc = proc;
        bk.entry_base = total;
        bk.entries = e;
        // group: start a new one when this block would overflow the current one
        bool fresh = D.groups.empty();
        if (!fresh) {
            const DropPlan::Group& g = D.groups.back();
            const int64_t span = rb - g.hist_begin / BS + 1;  // raw blocks the group would span
            fresh = g.entries + e > kGroupEntries || span > max_group_blocks || start + m - g.hist_begin > kGroupHistories;
        }
        if (fresh) D.groups.push_back(DropPlan::Group{(int64_t)D.blocks.size(), 0, total, 0, start, start});
        DropPlan::Group& g = D.groups.back();
        bk.group = (int64_t)D.groups.size() - 1;
        g.n_blocks = rb - g.hist_begin / BS + 1;
        g.entries += e;
        g.hist_end = start + m;
        D.blocks.push_back(bk);
        total += e;
        max_proc = proc > max_proc ? proc : max_proc;
    }
    const int64_t nb = (int64_t)D.blocks.size(), ng = (int64_t)D.groups.size();
    int64_t max_ge = 0, max_gh = 0, max_gb = 0;
    for (const auto& g : D.groups) {
        max_ge = g.entries > max_ge ? g.entries : max_ge;
        max_gh = g.hist_end - g.hist_begin > max_gh ? g.hist_end - g.hist_begin : max_gh;
        max_gb = g.n_blocks > max_gb ? g.n_blocks : max_gb;
    }
    // persistent arrays and scratch
    PCT_TRY(grow(ctx, &D.t_hist, &D.cap_entries, total + 1));
    if (D.cap_off < total + ng + 1 || !D.t_vox || !D.t_off) {
        PCT_TRY(pct_alloc(ctx, &D.t_vox, total + ng + 1));
        PCT_TRY(pct_alloc(ctx, &D.t_off, total + ng + 1));
        D.cap_off = total + ng + 1;
    }
    PCT_TRY(grow(ctx, &D.r, &D.cap_r, max_proc + 1));
    if (D.cap_perm < n + 1 || !D.perm || !D.rank || !D.desc) {
        PCT_TRY(pct_alloc(ctx, &D.perm, n + 1));
        PCT_TRY(pct_alloc(ctx, &D.rank, n + 1));
        PCT_TRY(pct_alloc(ctx, &D.desc, n + 1));
        D.cap_perm = n + 1;
    }
    if (D.cap_group_e < max_ge + 1 || !D.keys) {
        PCT_TRY(pct_alloc(ctx, &D.keys, max_ge + 1));
        PCT_TRY(pct_alloc(ctx, &D.keys_sorted, max_ge + 1));
        PCT_TRY(pct_alloc(ctx, &D.vals, max_ge + 1));
        PCT_TRY(pct_alloc(ctx, &D.runs, max_ge + 1));
        D.cap_group_e = max_ge + 1;
    }
    if (D.cap_group_h < max_gh + 1 || !D.hkey) {
        PCT_TRY(pct_alloc(ctx, &D.hkey, max_gh + 1));
        PCT_TRY(pct_alloc(ctx, &D.hkey_sorted, max_gh + 1));
        PCT_TRY(pct_alloc(ctx, &D.hidx, max_gh + 1));
        PCT_TRY(pct_alloc(ctx, &D.hidx_sorted, max_gh + 1));
        PCT_TRY(pct_alloc(ctx, &D.lc, max_gh + 1));
        PCT_TRY(pct_alloc(ctx, &D.pos, max_gh + 1));
        D.cap_group_h = max_gh + 1;
    }
    if (D.cap_group_b < max_gb + 1 || !D.run_start || !D.n_runs) {
        PCT_TRY(pct_alloc(ctx, &D.run_start, max_gb + 1));
        PCT_TRY(pct_alloc(ctx, &D.n_runs, 2));
        D.cap_group_b = max_gb + 1;
    }
    const uint32_t* lvox = ctx->list_vox;
    if (bricks) {
        if (!D.vox_b_ready && (D.cap_vox_b < ctx->list_cap + 1 || !D.vox_b)) {
            PCT_TRY(pct_alloc(ctx, &D.vox_b, ctx->list_cap + 1));
            D.cap_vox_b = ctx->list_cap + 1;
        }
        if (D.cap_x_b < D.nv_b || !D.x_b) {
            PCT_TRY(pct_alloc(ctx, &D.x_b, D.nv_b));
            D.cap_x_b = D.nv_b;
            PCT_CHECK_HIP(ctx, hipMemsetAsync(D.x_b, 0, D.nv_b * sizeof(float), s));  // padding stays 0
        }
        const BrickGeom bg{(uint32_t)P.columns, (uint32_t)P.rows, (uint32_t)D.nbx, (uint32_t)D.nby};
        if (!pipe && !D.vox_b_ready && ctx->list_cap > 0)  // lists built without the brick index (per group with pipe)
            hipLaunchKernelGGL(k_bricks_list, dim3(pct_grid_capped(ctx->n_recon * 16, kBlock, 1u << 20)), dim3(kBlock), 0,
                               s, ctx->n_recon, ctx->list_off, ctx->list_cnt, ctx->list_vox, D.vox_b, bg,
                               (uint32_t)P.num_voxels);
        D.vox_b_ready = !pipe;
        lvox = D.vox_b;
    }
    // bucket transposition: tiles of 4 x 4 bricks per block
    const bool buckets = bricks && (ctx->drop_flags & PCT_DROP_BUCKETS) != 0;
    TileGeom tg{};
    if (buckets) {
        tg.nbx = (uint32_t)D.nbx;
        tg.nby = (uint32_t)D.nby;
        tg.ntbx = (tg.nbx + 3) / 4;
        tg.ntby = (tg.nby + 3) / 4;
        tg.nt = tg.ntbx * tg.ntby * (uint32_t)((P.slices + 1) / 2);
        const int64_t nbk = max_gb * (int64_t)tg.nt + 1;
        if (nbk >= (1ll << 31)) {
            ctx->err = "DROP bucket transposition: too many buckets in a group";
            return PCT_E_LIMIT;
        }
        if (D.cap_bk < nbk || !D.bk_cnt) {
            PCT_TRY(pct_alloc(ctx, &D.bk_cnt, nbk));
            PCT_TRY(pct_alloc(ctx, &D.bk_off, nbk));
            PCT_TRY(pct_alloc(ctx, &D.bk_cur, nbk));
            PCT_TRY(pct_alloc(ctx, &D.bk_runs, nbk));
            PCT_TRY(pct_alloc(ctx, &D.bk_roff, nbk));
            PCT_TRY(pct_alloc(ctx, &D.bk_list, 3 * nbk));
            PCT_TRY(pct_alloc(ctx, &D.bk_misc, 4));
            D.cap_bk = nbk;
        }
    }
    D.groups_bucketed = D.groups_sorted = 0;
    // library scratch size for the largest group
    {
        size_t need = 0, t = 0;
        if (buckets) {
            PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, t, D.bk_cnt, D.bk_off, (int)D.cap_bk, s));
            need = std::max(need, t);
        }
        const int E1 = (int)(max_ge + 1), H1 = (int)(max_gh + 1);
        PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, t, D.hkey, D.hkey_sorted, D.hidx, D.hidx_sorted,
                                                              H1, 0, 32, s));
        need = std::max(need, t);
        PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, t, D.lc, D.pos, H1, s));
        need = std::max(need, t);
        PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, t, D.keys, D.keys_sorted, D.vals, D.t_hist, E1, 0,
                                                              32, s));
        need = std::max(need, t);
        PCT_CHECK_HIP(ctx, hipcub::DeviceRunLengthEncode::Encode(nullptr, t, D.keys_sorted, D.keys, D.runs, D.n_runs, E1,
                                                                 s));
        need = std::max(need, t);
        PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, t, D.runs, D.t_off, E1, s));
        need = std::max(need, t);
        if (need > ctx->cub_tmp_bytes) {
            PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, need));
            ctx->cub_tmp_bytes = need;
        }
    }
    std::vector<std::vector<int32_t>> run_start_h(ng);
    std::vector<uint32_t> n_runs_h(ng, 0);
    int64_t run_off = 0;
    // With `mlp`: the MLP kernel for each group's histories (and the histories of no group) is queued on
    // stream2 with an event after each group; a group is transposed on `s` once its event has fired, while
    // stream2 builds the lists of the next groups.  The events and stream2 are drained on every return.
    struct MlpQueue {
        pct_ctx* ctx;
        std::vector<hipEvent_t> ev;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        ~MlpQueue()
        {
            if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
            if (t0) (void)hipEventDestroy(t0);
            if (t1) (void)hipEventDestroy(t1);
        }
    } q{ctx, {}};
    if (pipe) {
        if (!ctx->stream2) PCT_CHECK_HIP(ctx, hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
        hipStream_t s2 = ctx->stream2;
        hipEvent_t ready = nullptr;  // the capacity pass, the scan and the buffers on `s` come first
        PCT_CHECK_HIP(ctx, hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        q.ev.push_back(ready);
        PCT_CHECK_HIP(ctx, hipEventRecord(ready, s));
        PCT_CHECK_HIP(ctx, hipStreamWaitEvent(s2, ready, 0));
        PCT_CHECK_HIP(ctx, hipEventCreate(&q.t0));
        PCT_CHECK_HIP(ctx, hipEventCreate(&q.t1));
        PCT_CHECK_HIP(ctx, hipEventRecord(q.t0, s2));
        int64_t prev = 0;
        for (int64_t gi = 0; gi <= ng; gi++) {
            const int64_t h1 = gi < ng ? D.groups[gi].hist_end : n;
            if (h1 > prev) PCT_TRY((*mlp)(prev, h1, s2));
            prev = h1 > prev ? h1 : prev;
            hipEvent_t e = nullptr;
            PCT_CHECK_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
            q.ev.push_back(e);
            PCT_CHECK_HIP(ctx, hipEventRecord(e, s2));
        }
        PCT_CHECK_HIP(ctx, hipEventRecord(q.t1, s2));
    }
    int64_t entry_total = 0;  // entries of the groups transposed so far
    std::vector<int64_t> group_run_off(ng, 0);
    for (int64_t gi = 0; gi < ng; gi++) {
        DropPlan::Group& gw = D.groups[gi];
        if (pipe) {
            PCT_CHECK_HIP(ctx, hipStreamWaitEvent(s, q.ev[gi + 1], 0));
            const int64_t gh0 = gw.hist_begin, gH = gw.hist_end - gw.hist_begin, grb0 = gh0 / BS;
            if (bricks && gH > 0) {
                const BrickGeom bg{(uint32_t)P.columns, (uint32_t)P.rows, (uint32_t)D.nbx, (uint32_t)D.nby};
                hipLaunchKernelGGL(k_bricks_list, dim3(pct_grid_capped(gH * 16, kBlock, 1u << 20)), dim3(kBlock), 0, s,
                                   gH, ctx->list_off + gh0, ctx->list_cnt + gh0, ctx->list_vox, D.vox_b, bg,
                                   (uint32_t)P.num_voxels);
            }
            // the group's entries, now that its lists exist
            hipLaunchKernelGGL(k_block_entries, dim3((unsigned)gw.n_blocks), dim3(256), 0, s, BS, proc_dev + grb0,
                               ctx->list_cnt + grb0 * BS, P.min_mlp_length, e_dev + grb0);
            PCT_CHECK_HIP(ctx, hipMemcpyAsync(e_raw.data() + grb0, e_dev + grb0, gw.n_blocks * sizeof(int64_t),
                                              hipMemcpyDeviceToHost, s));
            PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
            gw.entry_base = entry_total;
            gw.entries = 0;
            for (int64_t b = gw.first_block; b < nb && D.blocks[b].group == gi; b++) {
                DropPlan::Block& bk = D.blocks[b];
                const int64_t e = e_raw[bk.start / BS];
                if (e > bk.entries) {  // the capacity bounds the entries
                    ctx->err = "DROP transposition: more entries than list capacity in a block";
                    return PCT_E_LIMIT;
                }
                bk.entries = e;
                bk.entry_base = entry_total + gw.entries;
                gw.entries += e;
            }
        }
        entry_total += gw.entries;
        const DropPlan::Group& g = gw;
        const int64_t h0 = g.hist_begin, H = g.hist_end - g.hist_begin, E = g.entries;
        const int64_t rb0 = h0 / BS;
        const int sort_bits = bits_for(g.n_blocks) + bin_bits;
        size_t tb = ctx->cub_tmp_bytes;
        hipLaunchKernelGGL(k_group_hist_keys, dim3(grid_for(H)), dim3(kBlock), 0, s, h0, H, BS, rb0, proc_dev,
                           ctx->bin_num, spatial ? 1 : 0, bin_bits, D.hkey, D.hidx);
        PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp, tb, D.hkey, D.hkey_sorted, D.hidx,
                                                              D.hidx_sorted, (int)H, 0, sort_bits, s));
        hipLaunchKernelGGL(k_group_rank, dim3(grid_for(H)), dim3(kBlock), 0, s, h0, H, BS, D.hidx_sorted, ctx->list_off,
                           ctx->list_cnt, ctx->chord, ctx->wepl, D.perm, D.rank, D.desc);
        group_run_off[gi] = run_off;
        PCT_CHECK_HIP(ctx, hipMemsetAsync(D.run_start, 0xff, (g.n_blocks + 1) * sizeof(int32_t), s));
        bool bucketed = false;
        if (E > 0) {
            hipLaunchKernelGGL(k_group_long_counts, dim3(grid_for(H)), dim3(kBlock), 0, s, h0, H, BS, proc_dev,
                               ctx->list_cnt, P.min_mlp_length, D.lc);
            if (buckets)
                PCT_TRY(transpose_group_buckets(ctx, D, g, gi, run_off, BS, tg, lvox, &n_runs_h[gi], &bucketed));
        }
        if (bucketed) {
            D.groups_bucketed++;
        } else if (E > 0) {
            D.groups_sorted++;
            tb = ctx->cub_tmp_bytes;
            PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, tb, D.lc, D.pos, (int)H, s));
            int64_t wg = (H * 16 + kBlock - 1) / kBlock;  // 16 lanes per history
            if (wg > 65536) wg = 65536;
            hipLaunchKernelGGL(k_group_emit, dim3((unsigned)wg), dim3(kBlock), 0, s, h0, H, BS, rb0, vbits, ctx->list_off,
                               lvox, D.lc, D.pos, D.rank, D.keys, D.vals);
            tb = ctx->cub_tmp_bytes;
            PCT_CHECK_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp, tb, D.keys, D.keys_sorted, D.vals,
                                                                  D.t_hist + g.entry_base, (int)E, 0,
                                                                  bits_for(g.n_blocks) + vbits, s));
            tb = ctx->cub_tmp_bytes;
            PCT_CHECK_HIP(ctx, hipcub::DeviceRunLengthEncode::Encode(ctx->cub_tmp, tb, D.keys_sorted, D.keys, D.runs,
                                                                     D.n_runs, (int)E, s));
            // the run count sizes the scan and the runs kernel
            PCT_CHECK_HIP(ctx, hipMemcpyAsync(&n_runs_h[gi], D.n_runs, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
            const int64_t nr = (int64_t)n_runs_h[gi];
            if (nr < 1 || nr > E) {
                ctx->err = "DROP transposition: run count out of range in a group";
                return PCT_E_LIMIT;
            }
            // t_off[u] = first entry of run u; t_off[nr] = E (the zero appended to the counts)
            PCT_CHECK_HIP(ctx, hipMemsetAsync(D.runs + nr, 0, sizeof(uint32_t), s));
            tb = ctx->cub_tmp_bytes;
            PCT_CHECK_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, tb, D.runs, D.t_off + run_off + gi,
                                                                (int)(nr + 1), s));
            hipLaunchKernelGGL(k_group_runs, dim3(grid_for(nr)), dim3(kBlock), 0, s, D.keys, D.n_runs, vbits,
                               (uint32_t)g.n_blocks, D.t_vox + run_off + gi, D.run_start);
            PCT_CHECK_HIP(ctx, hipGetLastError());
        }
        run_start_h[gi].assign(g.n_blocks + 1, -1);
        PCT_CHECK_HIP(ctx, hipMemcpyAsync(run_start_h[gi].data(), D.run_start, g.n_blocks * sizeof(int32_t),
                                          hipMemcpyDeviceToHost, s));
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));  // run_start and n_runs are reused by the next group
        if ((int64_t)n_runs_h[gi] > E) {
            ctx->err = "DROP transposition: more runs than entries in a group";
            return PCT_E_LIMIT;
        }
        run_off += (int64_t)n_runs_h[gi];
    }
    if (pipe) {
        PCT_CHECK_HIP(ctx, hipStreamWaitEvent(s, q.ev[ng + 1], 0));  // the histories of no group
        if (bricks) {  // their lists too hold brick indices (the residual reads every history's list)
            const BrickGeom bg{(uint32_t)P.columns, (uint32_t)P.rows, (uint32_t)D.nbx, (uint32_t)D.nby};
            int64_t prev = 0;
            for (int64_t gi = 0; gi <= ng; gi++) {
                const int64_t a = prev, b = gi < ng ? D.groups[gi].hist_begin : n;
                if (b > a)
                    hipLaunchKernelGGL(k_bricks_list, dim3(pct_grid_capped((b - a) * 16, kBlock, 1u << 20)), dim3(kBlock),
                                       0, s, b - a, ctx->list_off + a, ctx->list_cnt + a, ctx->list_vox, D.vox_b, bg,
                                       (uint32_t)P.num_voxels);
                if (gi < ng) prev = D.groups[gi].hist_end > prev ? D.groups[gi].hist_end : prev;
            }
            PCT_CHECK_HIP(ctx, hipGetLastError());
        }
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream2));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, q.t0, q.t1);
        ctx->t_mlp_kernel = ms;  // first MLP launch to the last one's end, sharing the GPU with the transposition
        if (bricks) D.vox_b_ready = true;
    }
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    pct_free(&proc_buf.p);
    pct_free(&e_buf.p);
    // per block: its runs [run_start, next run_start) within the group
    for (auto& bk : D.blocks) {
        const DropPlan::Group& g = D.groups[bk.group];
        const int64_t lb = bk.start / BS - g.hist_begin / BS;
        const std::vector<int32_t>& rs = run_start_h[bk.group];
        bk.rank_base = g.entry_base;
        bk.n_vox = 0;
        bk.run_base = group_run_off[bk.group] + bk.group;
        if (rs[lb] < 0) continue;
        int64_t end = n_runs_h[bk.group];
        for (int64_t j = lb + 1; j < g.n_blocks; j++)
            if (rs[j] >= 0) {
                end = rs[j];
                break;
            }
        if (rs[lb] >= end || end - rs[lb] > std::min<int64_t>(bk.entries, nv_index)) {
            ctx->err = "DROP transposition: inconsistent runs for a block";
            return PCT_E_LIMIT;
        }
        bk.run_base += rs[lb];
        bk.n_vox = (uint32_t)(end - rs[lb]);
    }
    if ((ctx->drop_flags & PCT_DROP_COLS_SLICED) && !(ctx->drop_flags & PCT_DROP_COLS_LDS))
        PCT_TRY(build_slices(ctx, D));
    // the groups' descriptors cover every history when the groups tile [0, n)
    D.desc_all = !D.groups.empty() && D.groups.front().hist_begin == 0 && D.groups.back().hist_end == n;
    for (size_t gi = 1; gi < D.groups.size(); gi++)
        D.desc_all = D.desc_all && D.groups[gi].hist_begin == D.groups[gi - 1].hist_end;
    D.flags = ctx->drop_flags;
    D.ready = true;
    (void)nb;
    return PCT_OK;
}

// Launch `kernel`.  With `ev`, through hipExtLaunchKernelGGL with events recorded at the kernel's own
// start and end (the duration a kernel trace reports), appended to *ev as a pair.
template <typename K, typename... Args>
void drop_launch(K kernel, dim3 g, dim3 b, hipStream_t s, std::vector<hipEvent_t>* ev, Args... args)
{
    if (ev) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        hipExtLaunchKernelGGL(kernel, g, b, 0, s, e0, e1, 0, args...);
        ev->push_back(e0);
        ev->push_back(e1);
    } else {
        hipLaunchKernelGGL(kernel, g, b, 0, s, args...);
    }
}

// Image into / out of brick order around a DROP call (no-op without PCT_DROP_BRICKS).
int pct_drop_enter(pct_ctx* ctx, bool enter)
{
    const DropPlan& D = ctx->plan;
    if (!(D.flags & PCT_DROP_BRICKS)) return PCT_OK;
    const pct_params& P = ctx->P;
    const BrickGeom bg{(uint32_t)P.columns, (uint32_t)P.rows, (uint32_t)D.nbx, (uint32_t)D.nby};
    hipLaunchKernelGGL(k_x_bricks, dim3(grid_for(P.num_voxels)), dim3(kBlock), 0, ctx->stream, (int64_t)P.num_voxels,
                       ctx->x, D.x_b, bg, enter ? 1 : 0);
    PCT_CHECK_HIP(ctx, hipGetLastError());
    return PCT_OK;
}

// One DROP iteration with the prepared plan; optional per-kernel timing.
int pct_drop_iterate_plan(pct_ctx* ctx, DropTiming* timing)
{
    const pct_params& P = ctx->P;
    hipStream_t s = ctx->stream;
    DropPlan& D = ctx->plan;
    std::vector<hipEvent_t> ev_rows, ev_cols;
    std::vector<hipEvent_t>* er = timing ? &ev_rows : nullptr;
    std::vector<hipEvent_t>* ec = timing ? &ev_cols : nullptr;
    const int G = ctx->drop_row_lanes, V = ctx->drop_col_voxels;
    const bool xcd = (D.flags & PCT_DROP_XCD_SLABS) != 0;
    const bool bricks = (D.flags & PCT_DROP_BRICKS) != 0;
    const uint32_t* lv = bricks ? D.vox_b : ctx->list_vox;
    float* xx = bricks ? D.x_b : ctx->x;
    for (size_t bi = 0; bi < D.blocks.size(); bi++) {
        const DropPlan::Block& bk = D.blocks[bi];
        const int64_t hpw = kBlock / G;  // histories per workgroup
        uint32_t rper = 0;
        const dim3 rg(slab_grid((bk.proc + hpw - 1) / hpw, xcd, &rper)), blk(kBlock);
        switch (G) {
        case 4: drop_launch(k_drop_rows<4>, rg, blk, s, er, bk.start, bk.proc, rper, D.desc, lv, xx, P.lambda, D.r); break;
        case 8: drop_launch(k_drop_rows<8>, rg, blk, s, er, bk.start, bk.proc, rper, D.desc, lv, xx, P.lambda, D.r); break;
        case 16: drop_launch(k_drop_rows<16>, rg, blk, s, er, bk.start, bk.proc, rper, D.desc, lv, xx, P.lambda, D.r); break;
        case 32: drop_launch(k_drop_rows<32>, rg, blk, s, er, bk.start, bk.proc, rper, D.desc, lv, xx, P.lambda, D.r); break;
        default: drop_launch(k_drop_rows<64>, rg, blk, s, er, bk.start, bk.proc, rper, D.desc, lv, xx, P.lambda, D.r); break;
        }
        const uint32_t U = bk.n_vox;
        if (bk.entries > 0 && U > 0) {
            uint32_t cper = 0;
            const int64_t vpw = (int64_t)kBlock * V;  // voxels per workgroup
            const dim3 cg(slab_grid((U + vpw - 1) / vpw, xcd, &cper));
            const uint32_t* t_vox = D.t_vox + bk.run_base;
            const uint32_t* t_off = D.t_off + bk.run_base;
            const uint32_t* t_rank = D.t_hist + bk.rank_base;
#define PCT_COLS(A, VV) drop_launch(k_drop_cols<A, VV>, cg, blk, s, ec, U, cper, t_vox, t_off, t_rank, D.r, xx)
#define PCT_COLS_V(A)                  \
    switch (V) {                       \
    case 1: PCT_COLS(A, 1); break;     \
    case 2: PCT_COLS(A, 2); break;     \
    case 8: PCT_COLS(A, 8); break;     \
    default: PCT_COLS(A, 4); break;    \
    }
            const dim3 lg((U + kWave - 1) / kWave), lb(kWave);
            const uint64_t* sell_off = D.sell_off + bk.slice_base;
#define PCT_COLS_S(A, VV)                                                                                       \
    drop_launch(k_drop_cols_sliced<A, VV>, cg, blk, s, ec, U, cper, t_vox, t_off, sell_off, (const uint32_t*)D.sell, \
                (const double*)D.r, xx)
#define PCT_COLS_SV(A)                   \
    switch (V) {                         \
    case 1: PCT_COLS_S(A, 1); break;     \
    case 2: PCT_COLS_S(A, 2); break;     \
    case 8: PCT_COLS_S(A, 8); break;     \
    default: PCT_COLS_S(A, 4); break;    \
    }
            if (D.flags & PCT_DROP_COLS_LDS) {
                if (P.drop_accum == PCT_ACCUM_FP64)
                    drop_launch(k_drop_cols_lds<PCT_ACCUM_FP64>, lg, lb, s, ec, U, t_vox, t_off, t_rank, D.r, xx);
                else if (P.drop_accum == PCT_ACCUM_FP32)
                    drop_launch(k_drop_cols_lds<PCT_ACCUM_FP32>, lg, lb, s, ec, U, t_vox, t_off, t_rank, D.r, xx);
                else
                    drop_launch(k_drop_cols_lds<PCT_ACCUM_KAHAN32>, lg, lb, s, ec, U, t_vox, t_off, t_rank, D.r, xx);
            } else if (D.flags & PCT_DROP_COLS_SLICED) {
                if (P.drop_accum == PCT_ACCUM_FP64) {
                    PCT_COLS_SV(PCT_ACCUM_FP64)
                } else if (P.drop_accum == PCT_ACCUM_FP32) {
                    PCT_COLS_SV(PCT_ACCUM_FP32)
                } else {
                    PCT_COLS_SV(PCT_ACCUM_KAHAN32)
                }
            } else if (P.drop_accum == PCT_ACCUM_FP64) {
                PCT_COLS_V(PCT_ACCUM_FP64)
            } else if (P.drop_accum == PCT_ACCUM_FP32) {
                PCT_COLS_V(PCT_ACCUM_FP32)
            } else {
                PCT_COLS_V(PCT_ACCUM_KAHAN32)
            }
#undef PCT_COLS_SV
#undef PCT_COLS_S
#undef PCT_COLS_V
#undef PCT_COLS
        }
        if (timing) timing->histories += bk.proc;
    }
    PCT_CHECK_HIP(ctx, hipGetLastError());
    if (timing) {
        PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
        for (size_t j = 0; j + 1 < ev_rows.size(); j += 2) {
            float a = 0;
            (void)hipEventElapsedTime(&a, ev_rows[j], ev_rows[j + 1]);
            timing->scatter_ms += a;
            timing->scatter_launches++;
        }
        for (size_t j = 0; j + 1 < ev_cols.size(); j += 2) {
            float a = 0;
            (void)hipEventElapsedTime(&a, ev_cols[j], ev_cols[j + 1]);
            timing->apply_ms += a;
            timing->apply_launches++;
        }
        for (hipEvent_t e : ev_rows) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_cols) (void)hipEventDestroy(e);
    }
    return PCT_OK;
}

// ||b - A x|| over the reconstruction histories: the criterion that keeps or rejects the learned start.
// With the brick layout prepared, the lists' brick indices gather from a brick-ordered copy of x.
int pct_stage_residual(pct_ctx* ctx, double* norm)
{
    if (!ctx->seg_inside && pct_segment_histories(ctx) > 0) return pct_stage_residual_segments(ctx, norm);
    const int64_t n = ctx->n_recon;
    hipStream_t s = ctx->stream;
    DropPlan& D = ctx->plan;
    if (ctx->res_cap < n + 1 || !ctx->res_r2) {
        PCT_TRY(pct_alloc(ctx, &ctx->res_r2, n + 1));
        ctx->res_cap = n + 1;
    }
    double* out = ctx->res_r2 + n;
    PCT_CHECK_HIP(ctx, hipMemsetAsync(out, 0, sizeof(double), s));
    if (n > 0) {
        const uint32_t* vox = ctx->list_vox;
        const float* x = ctx->x;
        if (D.ready && (D.flags & PCT_DROP_BRICKS) && D.vox_b_ready && D.x_b) {
            const pct_params& P = ctx->P;
            const BrickGeom bg{(uint32_t)P.columns, (uint32_t)P.rows, (uint32_t)D.nbx, (uint32_t)D.nby};
            hipLaunchKernelGGL(k_x_bricks, dim3(grid_for(P.num_voxels)), dim3(kBlock), 0, s, (int64_t)P.num_voxels,
                               ctx->x, D.x_b, bg, 1);
            vox = D.vox_b;
            x = D.x_b;
        }
        constexpr int G = 16;
        const unsigned rg = pct_grid_capped(n * G, kBlock, 1u << 20);  // grid-stride kernels
        if (vox == D.vox_b && D.desc_all)  // the plan's descriptors cover every history
            hipLaunchKernelGGL(k_residual_desc<G>, dim3(rg), dim3(kBlock), 0, s, n, D.desc, vox, x,
                               ctx->P.min_mlp_length, ctx->res_r2);
        else
            hipLaunchKernelGGL(k_residual<G>, dim3(rg), dim3(kBlock), 0, s, n, ctx->list_off,
                               ctx->list_cnt, vox, ctx->chord, ctx->wepl, x, ctx->P.min_mlp_length, ctx->res_r2);
        PCT_CHECK_HIP(ctx, hipGetLastError());
        size_t need = 0;
        PCT_CHECK_HIP(ctx, hipcub::DeviceReduce::Sum(nullptr, need, ctx->res_r2, out, n, s));
        if (need > ctx->cub_tmp_bytes) {
            PCT_TRY(pct_alloc(ctx, (uint8_t**)&ctx->cub_tmp, need));
            ctx->cub_tmp_bytes = need;
        }
        need = ctx->cub_tmp_bytes;
        PCT_CHECK_HIP(ctx, hipcub::DeviceReduce::Sum(ctx->cub_tmp, need, ctx->res_r2, out, n, s));
    }
    double h = 0.0;
    PCT_CHECK_HIP(ctx, hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, s));
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));
    *norm = std::sqrt(h);
    return PCT_OK;
}
