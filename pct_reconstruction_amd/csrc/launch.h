// launch.h — workgroup counts of one-dimensional dispatches (host code only, no HIP types).
//
// One dispatch holds fewer than 2^32 work-items: the grid size times the workgroup size must fit in
// 32 bits.  Kernels that take one work-item per element are sized with pct_grid_for, which returns 0
// when the elements do not fit one dispatch; a launch with a zero-sized grid fails, and the caller's
// error check reports it instead of a dispatch that silently covers a wrapped count.  Kernels over
// larger ranges are grid-stride loops sized with pct_grid_capped.
#pragma once
#include <cstdint>

constexpr int64_t kPctMaxDispatchItems = 0xffffffffll;

inline unsigned pct_grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g * (int64_t)block > kPctMaxDispatchItems) return 0u;
    return (unsigned)g;
}

// ceil(n / block) workgroups, at most cap (and never more than one dispatch holds), at least 1.
inline unsigned pct_grid_capped(int64_t n, int block, unsigned cap)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (int64_t)cap) g = cap;
    const int64_t most = kPctMaxDispatchItems / block;
    if (g > most) g = most;
    return (unsigned)g;
}
