// k_drop.hip — stage (e): DROP block-iterative updates over precomputed MLP lists.
//
// Per DROP block (DROP_full_tx_iteration, DS.cu:9537-9589):
//   k_drop_scatter  one wavefront per history: 64 lanes read 64 consecutive list entries
//                   (coalesced), gather x, reduce a_i.x in FP64 across the wavefront, then every
//                   lane adds the history's update to its voxels' sums (calculate_x_update_GPU :9321
//                   + DROP_block_update_GPU :9522).  The per-voxel numerator is accumulated in FP64
//                   (PCT_ACCUM_FP64: 53-bit sums, order effects ~1e-16 relative) or in FP32 as the
//                   reference does; the denominator S is an exact u32 count.  The first lane to
//                   raise S[v] from 0 appends v to a touched-voxel list (wavefront-aggregated
//                   append: one atomic per wavefront and step).
//   k_drop_apply    x[v] += sum[v]/S[v] for the touched voxels only, and clears sum/S
//                   (image_update_GPU :9398 does this densely over every voxel).
// The touched list makes the apply cost proportional to the block's footprint instead of NUM_VOXELS.
#include <mutex>

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

template <bool FP64, bool TOUCHED>
__global__ void __launch_bounds__(kBlock) k_drop_scatter(int64_t start, int64_t proc, const int64_t* __restrict__ off,
                                                          const uint32_t* __restrict__ cnt_arr,
                                                          const uint32_t* __restrict__ vox,
                                                          const double* __restrict__ chord,
                                                          const float* __restrict__ wepl, const float* __restrict__ x,
                                                          double* __restrict__ acc64, float* __restrict__ acc32,
                                                          uint32_t* __restrict__ S, uint32_t* __restrict__ touched,
                                                          uint32_t* __restrict__ n_touched, int slot, int min_len,
                                                          double lambda)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    if (TOUCHED && wave == 0 && lane == 0) n_touched[slot ^ 1] = 0;  // consumed by the previous block's apply
    for (int64_t p = start + wave; p < start + proc; p += n_waves) {
        const int64_t base = off[p];
        const int64_t cnt = cnt_arr ? (int64_t)cnt_arr[p] : off[p + 1] - base;
        const double L = chord[p];
        double part = 0.0;
        for (int64_t j = lane; j < cnt; j += kWave) part += (double)x[vox[base + j]];
        double ax = wave_sum(part) * L;
        double aa = (double)pct_sqf(L) * (double)cnt;
        if (cnt < min_len) continue;  // uniform across the wavefront
        const double val = L * (((double)wepl[p] - ax) / aa) * lambda;
        for (int64_t j0 = 0; j0 < cnt; j0 += kWave) {
            const int64_t j = j0 + lane;
            bool fresh = false;
            uint32_t v = 0;
            if (j < cnt) {
                v = vox[base + j];
                uint32_t old = atomicAdd(&S[v], 1u);
                fresh = (old == 0);
                if (FP64) atomicAdd(&acc64[v], val);
                else atomicAdd(&acc32[v], (float)val);
            }
            if (TOUCHED) {
                uint64_t m = __ballot(fresh);
                if (m) {
                    int leader = __ffsll((unsigned long long)m) - 1;
                    uint32_t b = 0;
                    if (lane == leader) b = atomicAdd(&n_touched[slot], (uint32_t)__popcll(m));
                    b = __shfl(b, leader, kWave);
                    if (fresh) touched[b + __popcll(m & ((1ull << lane) - 1ull))] = v;
                }
            }
        }
    }
}

template <bool FP64>
__global__ void k_drop_apply_touched(const uint32_t* __restrict__ touched, const uint32_t* __restrict__ n_touched,
                                     int slot, float* __restrict__ x, double* __restrict__ acc64,
                                     float* __restrict__ acc32, uint32_t* __restrict__ S)
{
    const uint32_t n = n_touched[slot];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t v = touched[t];
        const uint32_t s = S[v];
        if (FP64) {
            x[v] += (float)(acc64[v] / (double)s);
            acc64[v] = 0.0;
        } else {
            x[v] += acc32[v] / (float)s;
            acc32[v] = 0.0f;
        }
        S[v] = 0;
    }
}

template <bool FP64>
__global__ void k_drop_apply_dense(int64_t nv, float* __restrict__ x, double* __restrict__ acc64,
                                   float* __restrict__ acc32, uint32_t* __restrict__ S)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t s = S[v];
        if (s > 0) {
            if (FP64) {
                x[v] += (float)(acc64[v] / (double)s);
                acc64[v] = 0.0;
            } else {
                x[v] += acc32[v] / (float)s;
                acc32[v] = 0.0f;
            }
            S[v] = 0;
        }
    }
}

}  // namespace

int pct_drop_run(hipStream_t s, int64_t n, const int64_t* off, const uint32_t* cnt, const uint32_t* vox,
                 const double* chord, const float* wepl, float* x, double* acc64, float* acc32, uint32_t* S,
                 uint32_t* touched, uint32_t* n_touched, int64_t num_voxels, int block_size, int hpb, int quirk_a3,
                 int min_len, double lambda, int accum_mode, std::string* err, DropTiming* timing)
{
    const bool fp64 = accum_mode == PCT_ACCUM_FP64;
    const bool use_touched = touched != nullptr && n_touched != nullptr;
    if (use_touched) {
        hipError_t e = hipMemsetAsync(n_touched, 0, 2 * sizeof(uint32_t), s);
        if (e != hipSuccess) {
            *err = std::string("hipMemsetAsync: ") + hipGetErrorString(e);
            return PCT_E_HIP;
        }
    }
    // timing: one event pair per launch, read back after the stream drains
    std::vector<hipEvent_t> ev;
    auto mark = [&]() {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        ev.push_back(e);
    };
    int k = 0;
    for (int64_t start = 0; start < n; start += block_size, k++) {
        int64_t m = n - start < block_size ? n - start : block_size;
        int64_t proc = quirk_a3 ? (m / hpb) * hpb : m;  // DS.cu:9558: integer division floors
        if (proc <= 0) continue;
        int slot = k & 1;
        int64_t waves = proc;
        int64_t blocks = (waves * kWave + kBlock - 1) / kBlock;
        if (blocks > 8192) blocks = 8192;
        dim3 grid((unsigned)blocks), blk(kBlock);
        if (timing) {
            mark();
            timing->histories += proc;
        }
        if (fp64) {
            if (use_touched)
                hipLaunchKernelGGL((k_drop_scatter<true, true>), grid, blk, 0, s, start, proc, off, cnt, vox, chord,
                                   wepl, x, acc64, acc32, S, touched, n_touched, slot, min_len, lambda);
            else
                hipLaunchKernelGGL((k_drop_scatter<true, false>), grid, blk, 0, s, start, proc, off, cnt, vox, chord,
                                   wepl, x, acc64, acc32, S, touched, n_touched, slot, min_len, lambda);
        } else {
            if (use_touched)
                hipLaunchKernelGGL((k_drop_scatter<false, true>), grid, blk, 0, s, start, proc, off, cnt, vox, chord,
                                   wepl, x, acc64, acc32, S, touched, n_touched, slot, min_len, lambda);
            else
                hipLaunchKernelGGL((k_drop_scatter<false, false>), grid, blk, 0, s, start, proc, off, cnt, vox, chord,
                                   wepl, x, acc64, acc32, S, touched, n_touched, slot, min_len, lambda);
        }
        if (timing) mark();
        if (use_touched) {
            int64_t ab = (proc * 64 + kBlock - 1) / kBlock;  // a block touches at most sum(cnt) voxels
            if (ab > 2048) ab = 2048;
            if (fp64)
                hipLaunchKernelGGL(k_drop_apply_touched<true>, dim3((unsigned)ab), dim3(kBlock), 0, s, touched, n_touched,
                                   slot, x, acc64, acc32, S);
            else
                hipLaunchKernelGGL(k_drop_apply_touched<false>, dim3((unsigned)ab), dim3(kBlock), 0, s, touched,
                                   n_touched, slot, x, acc64, acc32, S);
        } else {
            int64_t ab = (num_voxels + kBlock - 1) / kBlock;
            if (ab > 4096) ab = 4096;
            if (fp64)
                hipLaunchKernelGGL(k_drop_apply_dense<true>, dim3((unsigned)ab), dim3(kBlock), 0, s, num_voxels, x,
                                   acc64, acc32, S);
            else
                hipLaunchKernelGGL(k_drop_apply_dense<false>, dim3((unsigned)ab), dim3(kBlock), 0, s, num_voxels, x,
                                   acc64, acc32, S);
        }
        if (timing) mark();
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        *err = std::string("DROP launch: ") + hipGetErrorString(e);
        return PCT_E_HIP;
    }
    if (timing) {
        (void)hipStreamSynchronize(s);
        for (size_t j = 0; j + 2 < ev.size(); j += 3) {
            float a = 0, b = 0;
            (void)hipEventElapsedTime(&a, ev[j], ev[j + 1]);
            (void)hipEventElapsedTime(&b, ev[j + 1], ev[j + 2]);
            timing->scatter_ms += a;
            timing->apply_ms += b;
            timing->scatter_launches++;
            timing->apply_launches++;
        }
        for (hipEvent_t x_ : ev) (void)hipEventDestroy(x_);
    }
    return PCT_OK;
}

int pct_stage_drop(pct_ctx* ctx, int iterations, DropTiming* timing)
{
    if (!ctx->seg_inside && pct_segment_histories(ctx) > 0) return pct_stage_drop_segments(ctx, iterations, timing);
    if (!ctx->plan.ready || ctx->plan.flags != ctx->drop_flags) PCT_TRY(pct_drop_prepare(ctx));
    if (!ctx->P.tvs_on) {
        PCT_TRY(pct_drop_enter(ctx, true));
        for (int it = 0; it < iterations; it++) PCT_TRY(pct_drop_iterate_plan(ctx, timing));
        PCT_TRY(pct_drop_enter(ctx, false));
        ctx->drop_k += iterations;
        return PCT_OK;
    }
    // DROP_GPU (DS.cu:9885-9915): NTVS_iteration before (TVS_FIRST) or after each DROP iteration; TVS
    // works on the linear image, DROP on its own layout.
    for (int it = 0; it < iterations; it++) {
        const int k = ++ctx->drop_k;
        if (ctx->P.tvs_first) PCT_TRY(pct_tvs_iteration(ctx, k));
        PCT_TRY(pct_drop_enter(ctx, true));
        PCT_TRY(pct_drop_iterate_plan(ctx, timing));
        PCT_TRY(pct_drop_enter(ctx, false));
        if (!ctx->P.tvs_first) PCT_TRY(pct_tvs_iteration(ctx, k));
    }
    return PCT_OK;
}

// Algorithmic work of one DROP iteration: histories processed by the blocks, list entries they
// gather (all of them) and list entries they scatter (histories with >= MIN_MLP_LENGTH voxels).
int pct_drop_entries(pct_ctx* ctx, int64_t* histories, int64_t* gathered, int64_t* scattered)
{
    const pct_params& P = ctx->P;
    const int64_t n = ctx->n_recon;
    std::vector<uint32_t> cnt(n > 0 ? n : 1);
    if (n > 0) PCT_CHECK_HIP(ctx, hipMemcpy(cnt.data(), ctx->list_cnt, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    int64_t g = 0, sc = 0, h = 0;
    for (int64_t start = 0; start < n; start += P.drop_block_size) {
        int64_t m = n - start < P.drop_block_size ? n - start : P.drop_block_size;
        int64_t proc = P.quirk_a3_drop_floor ? (m / P.histories_per_block) * P.histories_per_block : m;
        for (int64_t i = start; i < start + proc; i++) {
            g += cnt[i];
            if ((int64_t)cnt[i] >= P.min_mlp_length) sc += cnt[i];
        }
        h += proc;
    }
    *histories = h;
    *gathered = g;
    *scattered = sc;
    return PCT_OK;
}

extern "C" int pct_drop_profile(pct_ctx* ctx, int iterations, double* rows_ms, double* cols_ms,
                                int64_t* launches, int64_t* histories, int64_t* gathered, int64_t* scattered,
                                int64_t* touched)
{
    if (!ctx || iterations <= 0 || !rows_ms || !cols_ms || !launches || !histories || !gathered || !scattered ||
        !touched)
        return PCT_E_ARG;
    if (ctx->stage < ST_LISTS) {
        ctx->err = "pct_drop_profile: lists not built";
        return PCT_E_STATE;
    }
    DropTiming t;
    PCT_TRY(pct_stage_drop(ctx, iterations, &t));
    if (pct_segment_histories(ctx) > 0) {  // totals of the last iteration's segments
        *histories = ctx->seg.histories;
        *gathered = ctx->seg.gathered;
        *scattered = ctx->seg.scattered;
        *touched = ctx->seg.touched;
    } else {
        PCT_TRY(pct_drop_entries(ctx, histories, gathered, scattered));
        int64_t u = 0;
        for (const auto& bk : ctx->plan.blocks) u += bk.n_vox;
        *touched = u;
    }
    *rows_ms = t.scatter_ms / (t.scatter_launches ? t.scatter_launches : 1);
    *cols_ms = t.apply_ms / (t.apply_launches ? t.apply_launches : 1);
    *launches = t.scatter_launches / iterations;
    return PCT_OK;
}


// ---------------------------------------------------------------- allocation-free DROP iteration ----
// pct_drop_iteration_device without the workspace arguments: the per-voxel FP64 sums and u32 counts
// live in `scratch`, a file-scope device array of kNoallocMaxVoxels * 12 bytes (96 MiB) that every
// device running this library holds until the process ends.  Sums occupy the first 8 * kNoallocMaxVoxels
// bytes, counts the next 4 * kNoallocMaxVoxels.  Its device address is looked up once per device
// (hipGetSymbolAddress) and only that address is passed to kernels.  Calls that overlap on one device
// share the array and race; calls on one stream, or synchronised between streams, do not.
namespace {

constexpr int64_t kNoallocMaxVoxels = int64_t(1) << 23;
constexpr int kNoallocMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoallocMaxVoxels * 12];

std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kNoallocMaxDevices];

// Device address of the current device's scratch array, looked up on the first call on that device.
int noalloc_scratch(uint8_t** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return PCT_E_HIP;
    if (dev < 0 || dev >= kNoallocMaxDevices) return PCT_E_LIMIT;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PCT_E_HIP;
        g_scratch_ptr[dev] = (uint8_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return PCT_OK;
}

// Zero the sums and counts of the first nv voxels (a fill kernel rather than a memset node, so that
// the call can be captured in a graph).
__global__ void k_noalloc_clear(int64_t nv, double* __restrict__ acc, uint32_t* __restrict__ cnt)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride) {
        acc[v] = 0.0;
        cnt[v] = 0u;
    }
}

}  // namespace

extern "C" int pct_drop_iteration_device_noalloc(int64_t n, const int64_t* offsets, const uint32_t* voxels,
                                                 const double* chord, const float* wepl, float* x,
                                                 int64_t num_voxels, int block_size, int histories_per_block,
                                                 int quirk_a3, int min_mlp_length, double lambda, void* hip_stream)
{
    if (n < 0 || !offsets || !voxels || !chord || !wepl || !x || num_voxels <= 0 || block_size <= 0 ||
        histories_per_block <= 0)
        return PCT_E_ARG;
    if (num_voxels > kNoallocMaxVoxels) return PCT_E_LIMIT;
    uint8_t* base = nullptr;
    const int rc = noalloc_scratch(&base);
    if (rc != PCT_OK) return rc;
    double* acc = (double*)base;
    uint32_t* cnt = (uint32_t*)(base + 8 * kNoallocMaxVoxels);
    hipStream_t s = (hipStream_t)hip_stream;
    hipLaunchKernelGGL(k_noalloc_clear, dim3(pct_grid_capped(num_voxels, kBlock, 4096)), dim3(kBlock), 0, s,
                       num_voxels, acc, cnt);
    if (hipGetLastError() != hipSuccess) return PCT_E_HIP;
    std::string err;
    return pct_drop_run(s, n, offsets, nullptr, voxels, chord, wepl, x, acc, nullptr, cnt, nullptr, nullptr,
                        num_voxels, block_size, histories_per_block, quirk_a3, min_mlp_length, lambda,
                        PCT_ACCUM_FP64, &err);
}
