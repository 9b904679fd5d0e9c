/*
 * pct_tvs_oracle.c — CPU restatement of the reference's total variation superiorization (TVS).
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/ and __graft_entry__.smoke(), never by the
 * product library.
 *
 * Reference: src/pCT_Reconstruction_Data_Segments_Blake.cu (DS.cu) :10149-10602 and its call sites in
 * DROP_GPU :9889-9915; constants in include/pCT_Reconstruction_Data_Segments_Blake.h (DS.h) :427-444.
 * Built with -ffp-contract=off; every expression is evaluated in the C types the reference uses.
 * pow(float, int) has its C++ <cmath> meaning: both arguments promoted to double.
 *
 * The two reductions (total variation, squared norm of the subgradient) come in two orders:
 *   O_TVS_SUM_SERIAL_FLOAT (0): the reference's serial loops into a float accumulator
 *       (calculate_total_variation :10165-10190, generate_perturbation_array step 3 :10306-10309);
 *   O_TVS_SUM_TILED_FP64 (1): the library's order.  FP64 sum over tiles of 1024 voxels: thread j of
 *       256 adds voxels j, j+256, j+512, j+768 of the tile in turn (starting from 0.0), then a halving
 *       tree a[j] += a[j+s] for s = 128, 64, ..., 1; the tile sums are reduced the same way (thread j
 *       adds tiles j, j+256, ... in turn, then the same tree).
 * Mode 1 is what the GPU computes, bit for bit; mode 0 agrees with it to float rounding on small
 * images (tests/test_tvs.py).
 *
 * The draw of L (randi :11478, a std::default_random_engine seeded from the system clock) is replaced
 * by a seeded SplitMix64 draw (Steele, Lea and Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014) uniform over the same range; any value in the range is a possible
 * reference outcome, so the draw itself is parity unpinned.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pct_oracle.h"

enum { O_TVS_SUM_SERIAL_FLOAT = 0, O_TVS_SUM_TILED_FP64 = 1 };

static void o_tree256(double* a)
{
    for (int s = 128; s >= 1; s >>= 1)
        for (int j = 0; j < s; j++) a[j] = a[j] + a[j + s];
}

/* The library's FP64 order (mode 1 above). */
double o_tile_sum(const double* t, int64_t n)
{
    const int64_t nt = (n + 1023) / 1024;
    double* part = (double*)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(double));
    double a[256];
    for (int64_t tile = 0; tile < nt; tile++) {
        for (int j = 0; j < 256; j++) {
            double s = 0.0;
            for (int i = 0; i < 4; i++) {
                int64_t idx = tile * 1024 + j + 256 * i;
                if (idx < n) s += t[idx];
            }
            a[j] = s;
        }
        o_tree256(a);
        part[tile] = a[0];
    }
    for (int j = 0; j < 256; j++) {
        double s = 0.0;
        for (int64_t i = j; i < nt; i += 256) s += part[i];
        a[j] = s;
    }
    o_tree256(a);
    free(part);
    return a[0];
}

/* Length of the A^L table: TVS_ETA_SEQUENCE_LENGTH = log(TVS_MIN_ETA)/log(A), truncated to int (:10151). */
int o_tvs_eta_len(const pct_params* p) { return (int)(log(p->tvs_min_eta) / log(p->tvs_a)); }

/* TVS_eta_sequence_h[L - L_0] (generate_TVS_eta_sequence :10149-10156): powf(A, L) for L_0 <= L < length,
 * 0 at L = length (the calloc'd last element).  Beyond it the reference reads past the array; those
 * entries are 0 here. */
float o_tvs_eta(const pct_params* p, int L)
{
    const int len = o_tvs_eta_len(p);
    if (L < p->tvs_l0 || L >= len) return 0.0f;
    return (float)pow((double)(float)p->tvs_a, (double)(float)L);
}

static uint64_t o_splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* L = randi(L_0 + k - 1, L) (:10481): uniform over [lo, hi]; hi < lo (undefined in the reference) gives lo. */
int o_tvs_draw(int seed, int k, int lo, int hi)
{
    if (hi <= lo) return lo;
    uint64_t r = o_splitmix64(((uint64_t)(uint32_t)seed << 32) ^ (uint64_t)(uint32_t)k);
    return lo + (int)(r % (uint64_t)(hi - lo + 1));
}

/* One term of calculate_total_variation (:10178): sqrt(powf(dy, 2) + powf(dx, 2)) in float; 0 on the
 * last row and column of every slice, which the loops skip. */
static inline float o_tv_term(const pct_params* p, const float* img, int64_t v)
{
    const int C = p->columns, R = p->rows;
    const int col = (int)(v % C), row = (int)((v / C) % R);
    if (col >= C - 1 || row >= R - 1) return 0.0f;
    float dy = img[v + C] - img[v];
    float dx = img[v + 1] - img[v];
    return sqrtf(dy * dy + dx * dx);
}

float o_tv(const pct_params* p, const float* img, int mode)
{
    const int64_t NV = p->num_voxels;
    if (mode == O_TVS_SUM_SERIAL_FLOAT) {
        float tv = 0.0f;
        for (int64_t v = 0; v < NV; v++) tv += o_tv_term(p, img, v);
        return tv;
    }
    double* t = (double*)malloc((size_t)NV * sizeof(double));
    for (int64_t v = 0; v < NV; v++) t[v] = (double)o_tv_term(p, img, v);
    double s = o_tile_sum(t, NV);
    free(t);
    return (float)s;
}

/* generate_perturbation_array (:10262-10326): the subgradient G of the TV, accumulated in the
 * reference's loop order, its norm, and v = G / norm (0 when the norm is 0).  Returns the norm. */
float o_tvs_perturbation(const pct_params* p, const float* img, float* G, float* v, int mode)
{
    const int C = p->columns, R = p->rows, S = p->slices;
    const int64_t NV = p->num_voxels;
    float* gx = (float*)calloc((size_t)NV, sizeof(float));
    float* gy = (float*)calloc((size_t)NV, sizeof(float));
    float* gn = (float*)calloc((size_t)NV, sizeof(float));
    memset(G, 0, (size_t)NV * sizeof(float));
    for (int s = 0; s < S; s++)
        for (int r = 0; r < R - 1; r++)
            for (int c = 0; c < C - 1; c++) {
                int64_t w = c + (int64_t)r * C + (int64_t)s * R * C;
                gx[w] = img[w + 1] - img[w];
                gy[w] = img[w + C] - img[w];
                gn[w] = (float)sqrt(pow(gx[w], 2) + pow(gy[w], 2));
            }
    for (int s = 0; s < S; s++)
        for (int r = 0; r < R - 1; r++)
            for (int c = 0; c < C - 1; c++) {
                int64_t w = c + (int64_t)r * C + (int64_t)s * R * C;
                if (gn[w] > 0.0) {
                    G[w] -= (gx[w] + gy[w]) / gn[w];
                    G[w + C] += gy[w] / gn[w];
                    G[w + 1] += gx[w] / gn[w];
                }
            }
    float norm;
    if (mode == O_TVS_SUM_SERIAL_FLOAT) {
        norm = 0.0f;
        for (int64_t w = 0; w < NV; w++) norm += pow(G[w], 2);
        norm = sqrtf(norm);
    } else {
        double* t = (double*)malloc((size_t)NV * sizeof(double));
        for (int64_t w = 0; w < NV; w++) t[w] = (double)G[w] * (double)G[w];
        norm = sqrtf((float)o_tile_sum(t, NV));
        free(t);
    }
    for (int64_t w = 0; w < NV; w++) v[w] = norm != 0 ? G[w] / norm : 0.0f;
    free(gx);
    free(gy);
    free(gn);
    return norm;
}

/* apply_TVS_perturbation (:10417): out = in - BETA * v on hull voxels. */
static void o_tvs_apply(int64_t NV, const uint8_t* hull, const float* in, float* out, const float* v, float beta)
{
    for (int64_t w = 0; w < NV; w++)
        if (hull[w]) out[w] = in[w] - beta * v[w];
}

/*
 * NTVS_iteration (:10583) for DROP iteration k (1-based): iteratively_perturb_image (:10476) when
 * TVS_CONDITIONED, else iteratively_perturb_image_unconditional (:10562).  *L is the exponent carried
 * between iterations (the reference's global L, DS.h:439, 0 at start); a value below L_0 is raised to
 * L_0 (the reference would index before its table).  TVS_OLD: one repetition, no draw (:9890-9892).
 * tv_before / tv_after: the TV_x_values pushes of :10587 and :10596.
 */
void o_tvs_iteration(const pct_params* p, const uint8_t* hull, float* x, int k, int* L_io, int mode,
                     float* tv_before, float* tv_after)
{
    const int64_t NV = p->num_voxels;
    float* G = (float*)malloc((size_t)NV * sizeof(float));
    float* v = (float*)malloc((size_t)NV * sizeof(float));
    float* xt = (float*)malloc((size_t)NV * sizeof(float));
    int L = *L_io < p->tvs_l0 ? p->tvs_l0 : *L_io;
    const int reps = p->tvs_old ? 1 : p->tvs_repetitions;
    *tv_before = o_tv(p, x, mode);
    if (!p->tvs_old) L = o_tvs_draw(p->tvs_seed, k, p->tvs_l0 + k - 1, L);
    if (p->tvs_conditioned) {
        for (int n = 0; n < reps; n++) {
            o_tvs_perturbation(p, x, G, v, mode);
            float beta = o_tvs_eta(p, L);
            float tv_x = o_tv(p, x, mode);
            float tv_prev = 0.0f;
            memcpy(xt, x, (size_t)NV * sizeof(float));
            o_tvs_apply(NV, hull, xt, x, v, beta);
            for (;;) {
                L++;
                float tv_y = o_tv(p, x, mode);
                if (tv_y <= tv_x || (double)(fabsf(tv_y - tv_prev) / tv_y) < p->tv_threshold || L > 300) break;
                beta = o_tvs_eta(p, L);
                tv_prev = tv_y;
                o_tvs_apply(NV, hull, xt, x, v, beta);
            }
        }
    } else {
        for (int n = 0; n < reps; n++, L++) {
            o_tvs_perturbation(p, x, G, v, mode);
            o_tvs_apply(NV, hull, x, x, v, o_tvs_eta(p, L));
        }
    }
    *tv_after = o_tv(p, x, mode);
    *L_io = L;
    free(G);
    free(v);
    free(xt);
}
