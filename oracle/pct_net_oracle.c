/*
 * pct_net_oracle.c — FP64 CPU reference of the learned warm-start network (forward, loss, backward).
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/ to check the GPU network's forward pass and
 * gradients; never by the product library.  The network has no counterpart in the reference; this
 * restates the network defined in the library's header (pct_net_*): a 4-level 3D U-Net, channels
 * 16-32-64-128, 3x3x3 "same" convolutions + ReLU, 2x2x2 average pooling, nearest x2 upsampling with
 * channel concatenation (upsampled first, skip second), a linear 1x1x1 head, x0 = hull * (FBP + y),
 * loss = mean over hull voxels of (x0 - target)^2.
 *
 * bf16 != 0 rounds to bfloat16 (nearest even) where the GPU stores bfloat16: the packed input, the
 * weights, every stored activation and every stored activation gradient; all arithmetic is double.
 * bf16 == 0 is the plain FP64 network (pinned against central finite differences in the tests).
 *
 * Layout: NDHWC (x fastest, channels innermost).  Parameter vector: per convolution W[Cout][27][Cin]
 * then b[Cout] (enc1, enc2, enc3, enc4, dec3, dec2, dec1), then the head w[16], b.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double rbf(double x, int on)
{
    if (!on) return x;
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

enum { NL = 7 };
static const int CIN[NL] = {8, 16, 32, 64, 192, 96, 48};
static const int COUT[NL] = {16, 32, 64, 128, 64, 32, 16};
static const int LEVEL[NL] = {0, 1, 2, 3, 2, 1, 0};

typedef struct {
    int D, H, W;
} dims;

static dims lv(dims d, int l) { dims r = {d.D >> l, d.H >> l, d.W >> l}; return r; }
static int64_t nvox(dims d) { return (int64_t)d.D * d.H * d.W; }

int64_t on_num_params(void)
{
    int64_t n = 0;
    for (int l = 0; l < NL; l++) n += (int64_t)COUT[l] * 27 * CIN[l] + COUT[l];
    return n + 17;
}

/* input channel c of voxel (z,y,x): channels [0, ca) of a (coarse if ac), then b */
static double in_at(const double* a, int ca, int ac, const double* b, int cb, dims d, int z, int y, int x, int c)
{
    if (z < 0 || y < 0 || x < 0 || z >= d.D || y >= d.H || x >= d.W) return 0.0;
    if (c < ca) {
        if (ac) return a[(((int64_t)(z >> 1) * (d.H >> 1) + (y >> 1)) * (d.W >> 1) + (x >> 1)) * ca + c];
        return a[(((int64_t)z * d.H + y) * d.W + x) * ca + c];
    }
    return b[(((int64_t)z * d.H + y) * d.W + x) * cb + (c - ca)];
}

/* out[v][co] = sum_{tap, c} in[v + off(tap)][c] * w[co][tap][c] (+ bias, ReLU) */
static void conv_fwd(const double* a, int ca, int ac, const double* b, int cb, dims d, const double* w,
                     const double* bias, int relu, int cout, int bf, double* out)
{
    const int cin = ca + cb;
    /* every output is one sequential sum: the z-slabs are independent */
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < d.D; z++)
        for (int y = 0; y < d.H; y++)
            for (int x = 0; x < d.W; x++) {
                const int64_t v = ((int64_t)z * d.H + y) * d.W + x;
                for (int co = 0; co < cout; co++) {
                    double s = 0.0;
                    for (int tap = 0; tap < 27; tap++) {
                        const int zz = z + tap / 9 - 1, yy = y + (tap / 3) % 3 - 1, xx = x + tap % 3 - 1;
                        if (zz < 0 || yy < 0 || xx < 0 || zz >= d.D || yy >= d.H || xx >= d.W) continue;
                        const double* wr = w + ((int64_t)co * 27 + tap) * cin;
                        for (int c = 0; c < cin; c++) s += in_at(a, ca, ac, b, cb, d, zz, yy, xx, c) * wr[c];
                    }
                    if (bias) s += bias[co];
                    if (relu && s < 0) s = 0;
                    out[v * cout + co] = rbf(s, bf);
                }
            }
}

/* gradients of conv_fwd given dpre = dL/d(pre-activation): dw, db accumulate; dinput (concat layout,
 * full resolution, cin channels) is written (before any bf16 rounding of the caller) */
static void conv_bwd(const double* a, int ca, int ac, const double* b, int cb, dims d, const double* w, int cout,
                     const double* dpre, double* dw, double* db, double* din, int bf)
{
    const int cin = ca + cb;
    const int64_t n = nvox(d);
    if (din) memset(din, 0, sizeof(double) * n * cin);
    for (int z = 0; z < d.D; z++)
        for (int y = 0; y < d.H; y++)
            for (int x = 0; x < d.W; x++) {
                const int64_t v = ((int64_t)z * d.H + y) * d.W + x;
                for (int co = 0; co < cout; co++) {
                    const double g = dpre[v * cout + co];
                    if (g == 0.0) continue;
                    db[co] += g;
                    for (int tap = 0; tap < 27; tap++) {
                        const int zz = z + tap / 9 - 1, yy = y + (tap / 3) % 3 - 1, xx = x + tap % 3 - 1;
                        if (zz < 0 || yy < 0 || xx < 0 || zz >= d.D || yy >= d.H || xx >= d.W) continue;
                        const int64_t u = ((int64_t)zz * d.H + yy) * d.W + xx;
                        const double* wr = w + ((int64_t)co * 27 + tap) * cin;
                        double* dwr = dw + ((int64_t)co * 27 + tap) * cin;
                        for (int c = 0; c < cin; c++) {
                            dwr[c] += g * in_at(a, ca, ac, b, cb, d, zz, yy, xx, c);
                            if (din) din[u * cin + c] += g * wr[c];
                        }
                    }
                }
            }
    if (din && bf)
        for (int64_t j = 0; j < n * cin; j++) din[j] = rbf(din[j], bf);
}

static void avgpool(const double* in, dims d, int C, int bf, double* out)
{
    dims h = lv(d, 1);
    for (int z = 0; z < h.D; z++)
        for (int y = 0; y < h.H; y++)
            for (int x = 0; x < h.W; x++)
                for (int c = 0; c < C; c++) {
                    double s = 0;
                    for (int k = 0; k < 8; k++)
                        s += in[(((int64_t)(2 * z + (k >> 2)) * d.H + (2 * y + ((k >> 1) & 1))) * d.W + (2 * x + (k & 1))) * C + c];
                    out[(((int64_t)z * h.H + y) * h.W + x) * C + c] = rbf(s * 0.125, bf);
                }
}

/* Forward (and backward when grad != NULL).  Returns the loss; x0 (nvox) may be NULL. */
double on_loss_grad(int D, int H, int W, const float* params, const float* fbp, const uint8_t* hull,
                    const float* target, int bf, double* grad, double* x0)
{
    dims d0 = {D, H, W};
    const int64_t np = on_num_params();
    double* p = (double*)malloc(sizeof(double) * np);
    for (int64_t j = 0; j < np; j++) p[j] = (double)params[j];
    int64_t woff[NL], boff[NL], off = 0;
    for (int l = 0; l < NL; l++) {
        woff[l] = off;
        off += (int64_t)COUT[l] * 27 * CIN[l];
        boff[l] = off;
        off += COUT[l];
    }
    const int64_t hoff = off;
    /* bf16 weights (biases and head stay FP32 on the GPU) */
    for (int l = 0; l < NL; l++)
        for (int64_t j = 0; j < (int64_t)COUT[l] * 27 * CIN[l]; j++) p[woff[l] + j] = rbf(p[woff[l] + j], bf);
    const int64_t n0 = nvox(d0);
    int64_t nv[4];
    for (int l = 0; l < 4; l++) nv[l] = nvox(lv(d0, l));
    double* in8 = (double*)calloc(n0 * 8, sizeof(double));
    for (int64_t v = 0; v < n0; v++) {
        in8[v * 8] = rbf(fbp[v], bf);
        in8[v * 8 + 1] = hull[v] ? 1.0 : 0.0;
    }
    double *e[4], *pool[3], *dd[3];
    for (int l = 0; l < 4; l++) e[l] = (double*)calloc(nv[l] * COUT[l], sizeof(double));
    for (int l = 0; l < 3; l++) {
        pool[l] = (double*)calloc(nv[l + 1] * COUT[l], sizeof(double));
        dd[l] = (double*)calloc(nv[l] * COUT[6 - l], sizeof(double));
    }
    /* encoder */
    conv_fwd(in8, 8, 0, NULL, 0, d0, p + woff[0], p + boff[0], 1, 16, bf, e[0]);
    avgpool(e[0], d0, 16, bf, pool[0]);
    conv_fwd(pool[0], 16, 0, NULL, 0, lv(d0, 1), p + woff[1], p + boff[1], 1, 32, bf, e[1]);
    avgpool(e[1], lv(d0, 1), 32, bf, pool[1]);
    conv_fwd(pool[1], 32, 0, NULL, 0, lv(d0, 2), p + woff[2], p + boff[2], 1, 64, bf, e[2]);
    avgpool(e[2], lv(d0, 2), 64, bf, pool[2]);
    conv_fwd(pool[2], 64, 0, NULL, 0, lv(d0, 3), p + woff[3], p + boff[3], 1, 128, bf, e[3]);
    /* decoder */
    conv_fwd(e[3], 128, 1, e[2], 64, lv(d0, 2), p + woff[4], p + boff[4], 1, 64, bf, dd[2]);
    conv_fwd(dd[2], 64, 1, e[1], 32, lv(d0, 1), p + woff[5], p + boff[5], 1, 32, bf, dd[1]);
    conv_fwd(dd[1], 32, 1, e[0], 16, d0, p + woff[6], p + boff[6], 1, 16, bf, dd[0]);
    /* head */
    double* xo = (double*)malloc(sizeof(double) * n0);
    double err = 0, cnt = 0;
    for (int64_t v = 0; v < n0; v++) {
        double y = p[hoff + 16];
        for (int c = 0; c < 16; c++) y += p[hoff + c] * dd[0][v * 16 + c];
        xo[v] = hull[v] ? (double)fbp[v] + y : 0.0;
        if (hull[v]) {
            const double df = xo[v] - (double)target[v];
            err += df * df;
            cnt += 1;
        }
        if (x0) x0[v] = xo[v];
    }
    const double loss = cnt > 0 ? err / cnt : 0.0;
    if (grad) {
        memset(grad, 0, sizeof(double) * np);
        /* head backward -> dPre of dec1 */
        double* gd[3];
        for (int l = 0; l < 3; l++) gd[l] = (double*)calloc(nv[l] * COUT[6 - l], sizeof(double));
        for (int64_t v = 0; v < n0; v++) {
            const double dy = hull[v] && cnt > 0 ? 2.0 * (xo[v] - (double)target[v]) / cnt : 0.0;
            for (int c = 0; c < 16; c++) {
                grad[hoff + c] += dy * dd[0][v * 16 + c];
                gd[0][v * 16 + c] = dd[0][v * 16 + c] > 0 ? rbf(dy * p[hoff + c], bf) : 0.0;
            }
            grad[hoff + 16] += dy;
        }
        double* gcat[3];
        double* ge[4];
        double* gpool[3];
        for (int l = 0; l < 4; l++) ge[l] = (double*)calloc(nv[l] * COUT[l], sizeof(double));
        for (int l = 0; l < 3; l++) {
            gcat[l] = (double*)calloc(nv[l] * CIN[6 - l], sizeof(double));
            gpool[l] = (double*)calloc(nv[l + 1] * COUT[l], sizeof(double));
        }
        /* decoder d = 0 (dec1), 1 (dec2), 2 (dec3) */
        for (int dl = 0; dl < 3; dl++) {
            const int l = 6 - dl;
            dims dd_ = lv(d0, dl);
            const double* a = dl == 0 ? dd[1] : (dl == 1 ? dd[2] : e[3]);
            const int ca = CIN[l] - COUT[dl];  /* skip = enc at this level (COUT[dl]) */
            conv_bwd(a, ca, 1, e[dl], COUT[dl], dd_, p + woff[l], COUT[l], gd[dl], grad + woff[l], grad + boff[l],
                     gcat[dl], bf);
            /* coarse part: sum over children, masked by the coarse activation */
            dims h = lv(dd_, 1);
            double* cg = dl < 2 ? gd[dl + 1] : ge[3];
            for (int z = 0; z < h.D; z++)
                for (int y = 0; y < h.H; y++)
                    for (int x = 0; x < h.W; x++)
                        for (int c = 0; c < ca; c++) {
                            double s = 0;
                            for (int k = 0; k < 8; k++)
                                s += gcat[dl][(((int64_t)(2 * z + (k >> 2)) * dd_.H + (2 * y + ((k >> 1) & 1))) * dd_.W +
                                               (2 * x + (k & 1))) * CIN[l] + c];
                            const int64_t pidx = (((int64_t)z * h.H + y) * h.W + x) * ca + c;
                            cg[pidx] = a[pidx] > 0 ? rbf(s, bf) : 0.0;
                        }
        }
        /* encoder: enc4 (l = 3) .. enc1 (l = 0); ge[l] = dPre of enc l */
        for (int l = 3; l >= 0; l--) {
            dims dl = lv(d0, l);
            if (l < 3) {
                const int cs = CIN[6 - l], c_off = CIN[6 - l] - COUT[l];
                for (int z = 0; z < dl.D; z++)
                    for (int y = 0; y < dl.H; y++)
                        for (int x = 0; x < dl.W; x++) {
                            const int64_t v = ((int64_t)z * dl.H + y) * dl.W + x;
                            const int64_t pp = ((int64_t)(z >> 1) * (dl.H >> 1) + (y >> 1)) * (dl.W >> 1) + (x >> 1);
                            for (int c = 0; c < COUT[l]; c++) {
                                const double g = gcat[l][v * cs + c_off + c] + gpool[l][pp * COUT[l] + c] * 0.125;
                                ge[l][v * COUT[l] + c] = e[l][v * COUT[l] + c] > 0 ? rbf(g, bf) : 0.0;
                            }
                        }
            }
            const double* inp = l == 0 ? in8 : pool[l - 1];
            conv_bwd(inp, CIN[l], 0, NULL, 0, dl, p + woff[l], COUT[l], ge[l], grad + woff[l], grad + boff[l],
                     l > 0 ? gpool[l - 1] : NULL, bf);
        }
        for (int l = 0; l < 3; l++) {
            free(gd[l]);
            free(gcat[l]);
            free(gpool[l]);
        }
        for (int l = 0; l < 4; l++) free(ge[l]);
    }
    free(xo);
    for (int l = 0; l < 4; l++) free(e[l]);
    for (int l = 0; l < 3; l++) {
        free(pool[l]);
        free(dd[l]);
    }
    free(in8);
    free(p);
    return loss;
}
