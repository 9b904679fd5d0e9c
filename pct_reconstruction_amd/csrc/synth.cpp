// synth.cpp — seeded synthetic phantoms and proton histories.
//
// The reference ships no data (its inputs live on a cluster, README.md:19).  This generator makes
// histories of the shape the reference reads (four tracker planes, t/v per plane, WEPL, gantry
// angle), from analytic phantoms whose WEPL line integrals are exact:
//   * cone beam from a source at u = -SOURCE_RADIUS; entry positions uniform at the second plane;
//   * straight chord through the phantom; WEPL = sum of RSP x chord length per object;
//   * exit deviation by multiple Coulomb scattering: Highland angle
//     theta0 = 13.6/(beta c p) sqrt(L/X0) (1 + 0.038 ln(L/X0)) with L the water-equivalent length,
//     displacement sigma = theta0 L / sqrt(3), correlation sqrt(3)/2, independently in t and v;
//   * WEPL noise N(0, sigma) and a fraction of outliers.
// Randomness is counter based (splitmix64 of seed, history, draw), so the data do not depend on the
// number of threads.  Histories are ordered by projection angle as the per-angle files are read.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/pct.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
struct Rng {
    uint64_t base;
    uint32_t k = 0;
    Rng(uint64_t seed, uint64_t i) : base(mix64(seed * 0x632BE59BD9B4E019ull ^ mix64(i))) {}
    double uniform() { return (double)(mix64(base + 0x9E3779B97F4A7C15ull * (++k)) >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// parameter interval [s0, s1] of the line p + s d inside object o; false if none
bool object_interval(const pct_object& o, const double p[3], const double d[3], double* s0, double* s1)
{
    const double c = std::cos(o.phi_deg * M_PI / 180.0), sn = std::sin(o.phi_deg * M_PI / 180.0);
    // rotate into the object frame (rotation by -phi about z)
    double px = p[0] - o.cx, py = p[1] - o.cy, pz = p[2] - o.cz;
    double qx = c * px + sn * py, qy = -sn * px + c * py, qz = pz;
    double ex = c * d[0] + sn * d[1], ey = -sn * d[0] + c * d[1], ez = d[2];
    if (o.kind == 1) {  // ellipsoid
        double ax = qx / o.a, ay = qy / o.b, az = qz / o.c, bx = ex / o.a, by = ey / o.b, bz = ez / o.c;
        double A = bx * bx + by * by + bz * bz, B = 2 * (ax * bx + ay * by + az * bz), C = ax * ax + ay * ay + az * az - 1;
        double disc = B * B - 4 * A * C;
        if (A <= 0 || disc <= 0) return false;
        double r = std::sqrt(disc);
        *s0 = (-B - r) / (2 * A);
        *s1 = (-B + r) / (2 * A);
        return true;
    }
    // cylinder with axis z: radius a, half height c
    double lo = -1e300, hi = 1e300;
    double A = ex * ex + ey * ey, B = 2 * (qx * ex + qy * ey), C = qx * qx + qy * qy - o.a * o.a;
    if (A > 0) {
        double disc = B * B - 4 * A * C;
        if (disc <= 0) return false;
        double r = std::sqrt(disc);
        lo = (-B - r) / (2 * A);
        hi = (-B + r) / (2 * A);
    } else if (C > 0) {
        return false;
    }
    if (ez != 0) {
        double z0 = (-o.c - qz) / ez, z1 = (o.c - qz) / ez;
        if (z0 > z1) std::swap(z0, z1);
        lo = std::max(lo, z0);
        hi = std::min(hi, z1);
    } else if (std::fabs(qz) > o.c) {
        return false;
    }
    if (hi <= lo) return false;
    *s0 = lo;
    *s1 = hi;
    return true;
}

// WEPL (sum of RSP x length) along p + s d, s in [smin, smax]; |d| = 1
double line_wepl(const pct_phantom& ph, const double p[3], const double d[3])
{
    if (ph.n_objects <= 0) return 0.0;
    bool painter = ph.obj[0].rsp < 0;  // negative first RSP marks painter's order (see pct_phantom_random)
    if (!painter) {
        double w = 0.0;
        for (int k = 0; k < ph.n_objects; k++) {
            double s0, s1;
            if (object_interval(ph.obj[k], p, d, &s0, &s1)) w += ph.obj[k].rsp * (s1 - s0);
        }
        return w;
    }
    // painter's order: the last object containing a point sets its RSP (first object's RSP is -rsp)
    double iv[16][2];
    bool has[16];
    std::vector<double> br;
    for (int k = 0; k < ph.n_objects; k++) {
        has[k] = object_interval(ph.obj[k], p, d, &iv[k][0], &iv[k][1]);
        if (has[k]) {
            br.push_back(iv[k][0]);
            br.push_back(iv[k][1]);
        }
    }
    std::sort(br.begin(), br.end());
    double w = 0.0;
    for (size_t j = 0; j + 1 < br.size(); j++) {
        double a = br[j], b = br[j + 1];
        if (b <= a) continue;
        double mid = 0.5 * (a + b);
        for (int k = ph.n_objects - 1; k >= 0; k--) {
            if (has[k] && mid > iv[k][0] && mid < iv[k][1]) {
                w += std::fabs(ph.obj[k].rsp) * (b - a);
                break;
            }
        }
    }
    return w;
}

double point_rsp(const pct_phantom& ph, double x, double y, double z)
{
    bool painter = ph.n_objects > 0 && ph.obj[0].rsp < 0;
    double r = 0.0;
    for (int k = 0; k < ph.n_objects; k++) {
        const pct_object& o = ph.obj[k];
        double c = std::cos(o.phi_deg * M_PI / 180.0), sn = std::sin(o.phi_deg * M_PI / 180.0);
        double px = x - o.cx, py = y - o.cy, pz = z - o.cz;
        double qx = c * px + sn * py, qy = -sn * px + c * py;
        bool in;
        if (o.kind == 1) in = (qx / o.a) * (qx / o.a) + (qy / o.b) * (qy / o.b) + (pz / o.c) * (pz / o.c) < 1.0;
        else in = qx * qx + qy * qy < o.a * o.a && std::fabs(pz) < o.c;
        if (!in) continue;
        if (painter) r = std::fabs(o.rsp);
        else r += o.rsp;
    }
    return r;
}

template <class F>
void parallel_for(int64_t n, F f)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    if (nt > 32) nt = 32;
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            int64_t a = n * t / nt, b = n * (t + 1) / nt;
            for (int64_t i = a; i < b; i++) f(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" void pct_phantom_cylinder_c1(pct_phantom* ph)
{
    std::memset(ph, 0, sizeof(*ph));
    // water cylinder R = 8 cm, H = 8 cm; three R = 1 cm inserts at r = 4 cm, phi = 0/120/240 deg
    // with RSP 0.3 / 1.1 / 1.7 (added as RSP - 1.0 over the water)
    ph->n_objects = 4;
    ph->obj[0] = pct_object{0, 0, 0, 0, 8.0, 8.0, 4.0, 0, 1.0};
    const double rsp[3] = {0.3, 1.1, 1.7};
    for (int k = 0; k < 3; k++) {
        double phi = k * 120.0 * M_PI / 180.0;
        ph->obj[1 + k] = pct_object{0, 4.0 * std::cos(phi), 4.0 * std::sin(phi), 0, 1.0, 1.0, 4.0, 0, rsp[k] - 1.0};
    }
}

extern "C" void pct_phantom_head(pct_phantom* ph)
{
    std::memset(ph, 0, sizeof(*ph));
    // Kak & Slaney 3-D head ellipsoids, scaled by 9 cm in x/y and 5 cm in z, with additive RSPs
    // giving 0.0 (cavities) to 1.8 (dense inserts).
    struct E { double cx, cy, cz, a, b, c, phi, rsp; };
    const E e[10] = {
        {0, 0, 0, 0.69, 0.92, 0.90, 0, 1.60},           // skull
        {0, 0, 0, 0.6624, 0.874, 0.88, 0, -0.56},       // brain -> 1.04
        {-0.22, 0, -0.25, 0.41, 0.16, 0.21, 108, -1.04},  // cavity -> 0.0
        {0.22, 0, -0.25, 0.31, 0.11, 0.22, 72, -1.04},    // cavity -> 0.0
        {0, 0.35, -0.25, 0.21, 0.25, 0.50, 0, 0.02},
        {0, 0.1, -0.25, 0.046, 0.046, 0.046, 0, 0.74},  // -> 1.8
        {-0.08, -0.65, -0.25, 0.046, 0.023, 0.02, 0, 0.30},
        {0.06, -0.65, -0.25, 0.046, 0.023, 0.02, 90, 0.30},
        {0.06, -0.105, 0.625, 0.056, 0.04, 0.10, 90, 0.20},
        {0, 0.1, 0.625, 0.056, 0.056, 0.10, 0, -0.30},
    };
    ph->n_objects = 10;
    for (int k = 0; k < 10; k++)
        ph->obj[k] = pct_object{1, e[k].cx * 9.0, e[k].cy * 9.0, e[k].cz * 5.0, e[k].a * 9.0, e[k].b * 9.0,
                                e[k].c * 5.0, e[k].phi, e[k].rsp};
}

extern "C" void pct_phantom_random(pct_phantom* ph, uint64_t seed, int n_ellipsoids, double radius, double half_height)
{
    std::memset(ph, 0, sizeof(*ph));
    if (n_ellipsoids < 1) n_ellipsoids = 1;
    if (n_ellipsoids > 16) n_ellipsoids = 16;
    Rng r(seed, 0xA11CE);
    // painter's order: object 0 is the body (its RSP stored negated to mark the mode); later
    // ellipsoids lie inside it and override the RSP where they are.
    double A = radius * (0.55 + 0.25 * r.uniform()), B = radius * (0.55 + 0.25 * r.uniform());
    double Cz = half_height * (0.7 + 0.25 * r.uniform());
    ph->obj[0] = pct_object{1, 0, 0, 0, A, B, Cz, 360 * r.uniform(), -(0.8 + 0.4 * r.uniform())};
    for (int k = 1; k < n_ellipsoids; k++) {
        double rr = 0.6 * r.uniform(), th = 2 * M_PI * r.uniform();
        double cx = rr * A * std::cos(th), cy = rr * B * std::sin(th), cz = (r.uniform() - 0.5) * Cz;
        double a = radius * (0.05 + 0.2 * r.uniform()), b = radius * (0.05 + 0.2 * r.uniform());
        double c = half_height * (0.1 + 0.3 * r.uniform());
        ph->obj[k] = pct_object{1, cx, cy, cz, a, b, c, 360 * r.uniform(), 1.8 * r.uniform()};
    }
    ph->n_objects = n_ellipsoids;
}

extern "C" int pct_generate_histories(const pct_phantom* ph, const pct_beam* bm, uint64_t seed, int64_t n,
                                      float* const t[4], float* const v[4], float* const u[4], float* wepl,
                                      float* angle_deg)
{
    if (!ph || !bm || n < 0 || !t || !v || !u || !wepl || !angle_deg) return PCT_E_ARG;
    const double SR = 260.7, m = 938.272, T = bm->energy_mev;
    const double E = T + m, pc = std::sqrt(E * E - m * m), bcp = pc * pc / E;
    const double X0 = 36.08;
    parallel_for(n, [&](int64_t i) {
        Rng r(seed, (uint64_t)i);
        double ang;
        if (bm->continuous_angles) ang = 360.0 * ((double)i + r.uniform()) / (double)n;  // sorted, uniform
        else ang = (double)((i * bm->n_angles) / (n > 0 ? n : 1)) * bm->angle_interval;
        double te = (2 * r.uniform() - 1) * bm->t_half_range, ve = (2 * r.uniform() - 1) * bm->v_half_range;
        const double* up = bm->u_planes;
        // ray from the source (u = -SR, 0, 0) through (up[1], te, ve)
        double du = up[1] + SR, dt = te, dv = ve;
        double norm = std::sqrt(du * du + dt * dt + dv * dv);
        auto at = [&](double uu, double* tt, double* vv) {
            *tt = dt * (uu + SR) / du;
            *vv = dv * (uu + SR) / du;
        };
        double t0, v0, t1, v1;
        at(up[0], &t0, &v0);
        at(up[1], &t1, &v1);
        // object frame: x = cos(a) u - sin(a) t, y = sin(a) u + cos(a) t, z = v
        double ca = std::cos(ang * M_PI / 180.0), sa = std::sin(ang * M_PI / 180.0);
        double pu = up[1], pt = t1, pv = v1;
        double p[3] = {ca * pu - sa * pt, sa * pu + ca * pt, pv};
        double d[3] = {(ca * du - sa * dt) / norm, (sa * du + ca * dt) / norm, dv / norm};
        double w_true = line_wepl(*ph, p, d);
        // exit: straight line to up[2], up[3] plus Highland scattering
        double t2, v2, t3, v3;
        at(up[2], &t2, &v2);
        at(up[3], &t3, &v3);
        if (bm->scatter && w_true > 1e-6) {
            double L = w_true;
            double th0 = 13.6 / bcp * std::sqrt(L / X0) * (1 + 0.038 * std::log(L / X0));
            double sy = th0 * L / std::sqrt(3.0), rho = std::sqrt(3.0) / 2;
            for (int axis = 0; axis < 2; axis++) {
                double z1 = r.normal(), z2 = r.normal();
                double dth = th0 * z1, dy = sy * (rho * z1 + std::sqrt(1 - rho * rho) * z2);
                double slope = (axis == 0 ? dt : dv) / du;
                double a2 = std::atan(slope) + dth;
                double& q2 = axis == 0 ? t2 : v2;
                double& q3 = axis == 0 ? t3 : v3;
                q2 += dy;
                q3 = q2 + (up[3] - up[2]) * std::tan(a2);
            }
        }
        double w = w_true + bm->wepl_noise_sigma * r.normal();
        if (r.uniform() < bm->outlier_fraction) w = w_true + (r.uniform() * 20.0 - 10.0);
        t[0][i] = (float)t0; t[1][i] = (float)t1; t[2][i] = (float)t2; t[3][i] = (float)t3;
        v[0][i] = (float)v0; v[1][i] = (float)v1; v[2][i] = (float)v2; v[3][i] = (float)v3;
        u[0][i] = (float)up[0]; u[1][i] = (float)up[1]; u[2][i] = (float)up[2]; u[3][i] = (float)up[3];
        wepl[i] = (float)w;
        angle_deg[i] = (float)ang;
    });
    return PCT_OK;
}

extern "C" int pct_phantom_rasterize(const pct_phantom* ph, const pct_params* p, float* out)
{
    if (!ph || !p || !out || p->num_voxels <= 0) return PCT_E_ARG;
    const int C = p->columns, R = p->rows, S = p->slices;
    parallel_for((int64_t)S * R, [&](int64_t sr) {
        int s = (int)(sr / R), r = (int)(sr % R);
        double y = p->recon_cyl_radius - (r + 0.5) * p->voxel_height;
        double z = p->recon_cyl_height / 2 - (s + 0.5) * p->voxel_thickness;
        for (int c = 0; c < C; c++) {
            double x = -p->recon_cyl_radius + (c + 0.5) * p->voxel_width;
            out[c + (int64_t)r * C + (int64_t)s * C * R] = (float)point_rsp(*ph, x, y, z);
        }
    });
    return PCT_OK;
}
