// gather_rate.hip — gather throughput of one MI355X against the number of wave lanes that share a
// 128-byte line, for a 25 MB table (the DROP image: larger than the 4 MB L2 of an XCD, inside the
// 256 MB Infinity Cache) and a 256 KB table (cache resident).  It supports the DROP roofline
// discussion: the DROP kernels gather 4-8 byte values whose lines are shared by few lanes.
//
// Built by pct_reconstruction_amd/build.py (next to this file); run pct_reconstruction_amd/microbench/gather_rate.
// Every lane reads inside its table: line < number of lines, offset < values per line.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <vector>

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16; return x;
}
// L lanes per line; each iteration every lane gathers one float (or double) from its line
template <typename T>
__global__ void k_gather(const T* __restrict__ x, uint32_t nlines, int L, int iters, T* sink) {
    const uint32_t lane = threadIdx.x & 63, wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    constexpr uint32_t per_line = 128 / sizeof(T);
    T acc = 0;
    for (int i = 0; i < iters; i += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t grp = lane / L;
            const uint32_t line = hash32(wave * 7919u + (uint32_t)(i + u) * 104729u + grp * 15485863u) % nlines;
            const uint32_t off = (lane % L) * (per_line / L);
            v[u] = x[(size_t)line * per_line + off];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc += v[u];
    }
    if (acc == (T)-1.2345) sink[0] = acc;
}
#define CHECK(expr)                                                                  \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s: %s\n", #expr, hipGetErrorString(e_));         \
            return 1;                                                                \
        }                                                                            \
    } while (0)

int main()
{
    const size_t nbytes = 25ull << 20;  // 25 MB, the size of the DROP image
    float *xf, *sf;
    double *xd, *sd;
    CHECK(hipMalloc(&xf, nbytes));
    CHECK(hipMalloc(&xd, nbytes));
    CHECK(hipMalloc(&sf, 64));
    CHECK(hipMalloc(&sd, 64));
    CHECK(hipMemset(xf, 0, nbytes));
    CHECK(hipMemset(xd, 0, nbytes));
    const uint32_t nlines = nbytes / 128;
    const int iters = 256, blocks = 256 * 16, threads = 256;  // 16 workgroups of 4 waves per CU
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    std::printf("| table | value | lanes per line | ms | G lane gathers/s | G instructions/s | G lines/s |\n");
    std::printf("|---|---|---|---|---|---|---|\n");
    for (int small = 0; small < 2; small++) {
        const uint32_t nl = small ? 2048u : nlines;  // 256 KB (cache resident) or 25 MB
        for (int L : {1, 2, 4, 8, 16, 32, 64}) {
            for (int dbl = 0; dbl < 2; dbl++) {
                if (dbl && L > 16) continue;
                for (int rep = 0; rep < 2; rep++) {  // the first run warms the caches
                    CHECK(hipEventRecord(a));
                    if (dbl)
                        hipLaunchKernelGGL(k_gather<double>, dim3(blocks), dim3(threads), 0, 0, xd, nl, L, iters, sd);
                    else
                        hipLaunchKernelGGL(k_gather<float>, dim3(blocks), dim3(threads), 0, 0, xf, nl, L, iters, sf);
                    CHECK(hipGetLastError());
                    CHECK(hipEventRecord(b));
                    CHECK(hipEventSynchronize(b));
                    float ms = 0;
                    CHECK(hipEventElapsedTime(&ms, a, b));
                    if (rep) {
                        const double g = (double)blocks * threads * iters;  // lane gathers
                        const double ins = g / 64, lines = ins * (64 / L);
                        std::printf("| %s | %s | %d | %.3f | %.1f | %.2f | %.1f |\n", small ? "256 KB" : "25 MB",
                                    dbl ? "f64" : "f32", L, ms, g / ms / 1e6, ins / ms / 1e6, lines / ms / 1e6);
                    }
                }
            }
        }
    }
    CHECK(hipFree(xf));
    CHECK(hipFree(xd));
    CHECK(hipFree(sf));
    CHECK(hipFree(sd));
    return 0;
}
