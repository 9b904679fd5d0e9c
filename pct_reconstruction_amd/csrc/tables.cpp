// tables.cpp — host construction of the MLP lookup tables (FP64).
//
// generate_trig_tables (src/pCT_Reconstruction_Data_Segments_Blake.cu:8421-8445),
// generate_scattering_coefficient_table (:8446-8464), generate_polynomial_tables (:8466-8502).
// The polynomial coefficients A_0..A_5 (include/pCT_Reconstruction_Data_Segments_Blake.h:749-754)
// are `c * pow((float)10, (float)e)`: C++ overload resolution picks the float pow, so 10^e is the
// float nearest to it.  The A_k_OVER_m macros (DS.h:756-781) expand unparenthesised, which makes
// the innermost `u * A_5_OVER_m` evaluate as ((u * c5) * P5) / m; that order is kept.
#include <cmath>

#include "internal.h"

namespace {

struct PolyCoeffs {
    double c[6] = {7.457, 4.548, -5.777, 1.301, -9.228, 2.687};
    float P[6];
    PolyCoeffs()
    {
        P[0] = std::pow(10.0f, -6.0f);
        P[1] = std::pow(10.0f, -7.0f);
        P[2] = std::pow(10.0f, -8.0f);
        P[3] = std::pow(10.0f, -8.0f);
        P[4] = std::pow(10.0f, -10.0f);
        P[5] = std::pow(10.0f, -11.0f);
    }
    // a0 + u*(a1 + u*(a2 + u*(a3 + u*(a4 + ((u*c5)*P5)/d5))))  with a_k = (c_k*P_k)/d_k
    double eval(double u, const int* d) const
    {
        double a0 = c[0] * P[0] / d[0], a1 = c[1] * P[1] / d[1], a2 = c[2] * P[2] / d[2];
        double a3 = c[3] * P[3] / d[3], a4 = c[4] * P[4] / d[4];
        double in5 = u * c[5] * P[5] / d[5];
        return a0 + u * (a1 + u * (a2 + u * (a3 + u * (a4 + in5))));
    }
};

}  // namespace

void pct_build_tables(const pct_params& P, HostTables* t)
{
    const int64_t n_trig = (int64_t)P.trig_table_elements + 1;
    const int64_t n_coeff = (int64_t)P.coeff_table_elements + 1;
    const int64_t n_poly = (int64_t)P.poly_table_elements + 1;
    t->sin_t.resize(n_trig);
    t->cos_t.resize(n_trig);
    for (int64_t s = 0; s < n_trig; s++) {
        double arg = P.trig_table_min + s * P.trig_table_step;
        t->sin_t[s] = std::sin(arg);
        t->cos_t[s] = std::cos(arg);
    }
    t->scat.resize(n_coeff);
    for (int64_t s = 0; s < n_coeff; s++) {
        double depth = s * P.coeff_table_step;
        t->scat[s] = std::pow(P.e_0 * (1 + 0.038 * std::log(depth / P.x_0_rad_length)), 2.0) / P.x_0_rad_length;
    }
    static const int d12[6] = {1, 2, 3, 4, 5, 6}, d23[6] = {2, 3, 4, 5, 6, 7}, d34[6] = {3, 4, 5, 6, 7, 8};
    static const int d26[6] = {2, 6, 12, 20, 30, 42}, d312[6] = {3, 12, 30, 60, 105, 168};
    PolyCoeffs pc;
    t->p12.resize(n_poly);
    t->p23.resize(n_poly);
    t->p34.resize(n_poly);
    t->p26.resize(n_poly);
    t->p312.resize(n_poly);
    for (int64_t s = 0; s < n_poly; s++) {
        double u = s * P.poly_table_step;
        t->p12[s] = u * pc.eval(u, d12);
        t->p23[s] = std::pow(u, 2) * pc.eval(u, d23);
        t->p34[s] = std::pow(u, 3) * pc.eval(u, d34);
        t->p26[s] = std::pow(u, 2) * pc.eval(u, d26);
        t->p312[s] = std::pow(u, 3) * pc.eval(u, d312);
    }
}

// The Sigma_1 terms of find_MLP_path_GPU_tabulated (DS.cu:9181-9192) depend only on the step
// number k: u_1 = (k+1)*MLP_U_STEP is read at table index (k+1)*poly_step, and the scattering
// coefficient at (k+1)*sigma_step.  They are evaluated here once, with the kernel's operation
// order, so the device loop only reads them.
void pct_build_step_terms(const pct_params& P, const HostTables& t, int n, std::vector<double>* out)
{
    const unsigned s_step = (unsigned)(P.mlp_u_step / P.coeff_table_step + 0.5);
    const unsigned p_step = (unsigned)(P.mlp_u_step / P.poly_table_step + 0.5);
    out->assign((size_t)n * 8, 0.0);
    for (int k = 0; k < n; k++) {
        size_t u1i = (size_t)p_step * (k + 1), s1 = (size_t)s_step * (k + 1);
        if (u1i >= t.p12.size() || s1 >= t.scat.size()) break;
        double u1_12 = t.p12[u1i], u1_23 = t.p23[u1i];
        double st1 = t.p312[u1i], st1th1 = t.p26[u1i], sth1 = u1_12;
        float f = (float)st1th1;
        double det1 = t.scat[s1] * (st1 * sth1 - (double)(f * f));
        double* row = out->data() + (size_t)k * 8;
        row[0] = u1_12;
        row[1] = u1_23;
        row[2] = t.p34[u1i];
        row[3] = sth1 / det1;
        row[4] = st1th1 / det1;
        row[5] = st1 / det1;
    }
}
