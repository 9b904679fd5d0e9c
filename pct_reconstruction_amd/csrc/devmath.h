// devmath.h — device-side arithmetic helpers shared by the gfx950 kernels.
//
// The reference evaluates its expressions in the C types of the operands (float + float is a float
// add, powf of a double argument rounds the argument to float).  These helpers give every CUDA float
// library call used on the path one fixed meaning so that the kernels are reproducible:
//   powf(x, 2)       -> x*x in float (correctly rounded square)
//   powf(x, -1)      -> 1/x in float
//   sinf/cosf/tanf/atan2f -> the double-precision function of the float argument, rounded to float.
// All code is compiled with -ffp-contract=off so that a*b+c stays two rounded operations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PCT_HD __host__ __device__ __forceinline__

PCT_HD float pct_sqf(double x)
{
    float f = (float)x;
    return f * f;
}
PCT_HD float pct_atan2f(float y, float x) { return (float)atan2((double)y, (double)x); }
PCT_HD float pct_sinf(double x) { return (float)sin((double)(float)x); }
PCT_HD float pct_cosf(double x) { return (float)cos((double)(float)x); }
PCT_HD float pct_tanf(double x) { return (float)tan((double)(float)x); }
PCT_HD float pct_recipf(double x)
{
    float f = (float)x;
    return 1.0f / f;
}

// Voxel-walk geometry (calculate_voxel_GPU DS.cu:11163, edge_coordinate_GPU :11322,
// distance_remaining_GPU :11313, corresponding_coordinate_GPU :11340).
PCT_HD int pct_calc_voxel(double zero, double pos, double size) { return (int)(fabs(pos - zero) / size); }
PCT_HD double pct_edge(double zero, int voxel_entered, double size, int inc, int step)
{
    int on_edge = (step == inc) ? voxel_entered : voxel_entered + 1;
    return zero + (inc * on_edge * size);
}
PCT_HD double pct_dist(double zero, double pos, int inc, int step, double size, int cur)
{
    int next = cur + inc * step;
    return fabs(pct_edge(zero, next, size, inc, step) - pos);
}
PCT_HD double pct_corr(double m, double x, double x0, double y0) { return m * (x - x0) + y0; }

// Coordinate conventions (DS.h:916-921): voxel numbers increase with +x, -y, -z.
#define PCT_X_INC 1
#define PCT_Y_INC (-1)
#define PCT_Z_INC (-1)

// Geometry subset of pct_params passed by value to kernels.
struct DevGeom {
    int C, R, S;
    int64_t NV;
    double X0, Y0, Z0, VW, VH, VT;
    double Rc, H, PI;
};

// Brick index used by DROP: voxel (x, y, z) of the C x R x S grid -> brick (z/2, y/4, x/4) of a grid
// of nbx x nby brick columns, times 32, plus (z%2, y%4, x%4) within the brick (32 floats = 128 B).
struct BrickGeom {
    uint32_t C, R, nbx, nby;
};
PCT_HD uint32_t brick_of(uint32_t v, BrickGeom g)
{
    const uint32_t x = v % g.C, t = v / g.C, y = t % g.R, z = t / g.R;
    const uint32_t b = ((z >> 1) * g.nby + (y >> 2)) * g.nbx + (x >> 2);
    return b * 32u + (z & 1u) * 16u + (y & 3u) * 4u + (x & 3u);
}
