/*
 * pct_oracle_fp64.c — the FP64 twin of the oracle: pct_oracle.c compiled with O_FP64.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Every value the reference holds or computes in float (the
 * per-history geometry, bin statistics, sinogram, FBP image, iterate and the float rounding points of the
 * MLP and intersection arithmetic) is double, and the CUDA float functions are their double versions.
 * Stage order, algorithms and integer decisions are those of pct_oracle.c; entry points carry the suffix
 * _f64 and take double arrays where the float build takes float arrays.  Against it each image of the
 * reference-precision pipeline states its arithmetic error.
 */
#define O_FP64 1
#include "pct_oracle.c"
