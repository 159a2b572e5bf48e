// CPU oracle for the ATHAD_COND scoped time step. TEST INFRASTRUCTURE ONLY: the HIP product
// never links or calls this file. See athad_oracle.h for the scope and the update orders.
//
// Every function cites the reference file:line it restates. Arithmetic follows the
// reference's expression order (the parity bar is 1e-10 relative after 100 steps, and the
// reference's own -O0/-O2 sensitivity is ~1e-14), but the code is written from scratch for a
// flat [i][j][k] layout.
#include "athad_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------------------
// Constants: MixtureAtm.h:40-57, SaturationH2O.h:50-54
// ---------------------------------------------------------------------------------------
constexpr double kTcrit   = 647.096;
constexpr double kPcrit   = 220640.0;
constexpr double kTtriple = 273.16;
constexpr double kPtriple = 6.11657;
constexpr double kNoSat   = -1.0;
constexpr double kRstar   = 8.314462618;
constexpr double kMH2O    = 0.018015;
constexpr double kMCO2    = 0.044010;
constexpr double kMN2     = 0.028014;
const double kRH2O = kRstar / kMH2O;
const double kRCO2 = kRstar / kMCO2;

inline bool finite_bits(double v) {  // lib/Utils.h:121-125
    uint64_t b; std::memcpy(&b, &v, 8);
    return (b & 0x7FF0000000000000ULL) != 0x7FF0000000000000ULL;
}
inline double dmax(double a, double b) { return (a < b) ? b : a; }   // std::max semantics
inline double dmin(double a, double b) { return (b < a) ? b : a; }   // std::min semantics
inline double dclamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// ---------------------------------------------------------------------------------------
// Saturation thermodynamics: SaturationH2O.h:68-301
// ---------------------------------------------------------------------------------------
double p_sat(double T) {                      // SaturationH2O.h:68-93 (Wagner-Pruss)
    if (T >= kTcrit) return kNoSat;
    if (T <= 0.0) return 0.0;
    const double th = 1.0 - T / kTcrit;
    const double sq = std::sqrt(th);
    const double t15 = th * sq;
    const double t3 = th * th * th;
    const double t35 = t3 * sq;
    const double t4 = t3 * th;
    const double t75 = t4 * t3 * sq;
    const double s = -7.85951783 * th + 1.84408259 * t15 + -11.7866497 * t3
                   + 22.6807411 * t35 + -15.9618719 * t4 + 1.80122502 * t75;
    return kPcrit * std::exp((kTcrit / T) * s);
}
double p_subl(double T) {                     // SaturationH2O.h:105-120 (IAPWS 2011)
    if (T <= 0.0) return 0.0;
    const double th = T / kTtriple;
    const double s = -0.212144006e2 * std::pow(th, 0.333333333e-2)
                   + 0.273203819e2 * std::pow(th, 0.120666667e1)
                   + -0.610598130e1 * std::pow(th, 0.170333333e1);
    return kPtriple * std::exp(s / th);
}
double p_auto(double T) {                     // SaturationH2O.h:123-127
    if (T >= kTcrit) return kNoSat;
    return (T >= kTtriple) ? p_sat(T) : p_subl(T);
}
double L_vap(double T) {                      // SaturationH2O.h:138-147 (Watson)
    if (T >= kTcrit) return 0.0;
    return 2.501e6 * std::pow((kTcrit - T) / (kTcrit - 273.15), 0.38);
}
double L_sub(double T) { return L_vap(T) + 3.337e5; }             // :150-154
double L_auto(double T) { return (T >= kTtriple) ? L_vap(T) : L_sub(T); }  // :207-210
double q_sat(double E, double p, double Mo) {  // :168-178 exact conversion
    if (E == kNoSat) return 1.0;
    if (!(p > 0.0)) return 0.0;
    if (E >= p) return 1.0;
    const double x = E / p;
    const double num = x * kMH2O;
    const double den = num + (1.0 - x) * Mo;
    return (den > 0.0) ? num / den : 1.0;
}
double sat_A(double E, double p, double Mo) {  // :230-239
    if (E == kNoSat || !(p > 0.0) || E >= p) return 0.0;
    const double x = E / p;
    const double den = x * kMH2O + (1.0 - x) * Mo;
    if (!(den > 0.0)) return 0.0;
    return x * kMH2O * Mo / (den * den);
}
double dqdT_from(double E, double L, double T, double p, double Mo) {  // :251-255
    if (!(T > 0.0)) return 0.0;
    return sat_A(E, p, Mo) * L / (kRH2O * T * T);
}
double moist_lapse(double T, double p, double Mo, double cp, double Rm, double g) {  // :280-293
    if (!(cp > 0.0) || !(T > 0.0) || !(Rm > 0.0)) return 0.0;
    const double A = sat_A(p_auto(T), p, Mo);
    if (!(A > 0.0)) return g / cp;
    const double L = L_auto(T);
    const double num = 1.0 + L * A / (Rm * T);
    const double den = cp + L * L * A / (kRH2O * T * T);
    return (den > 0.0) ? g * num / den : g / cp;
}
double q_sat_at(double T, double p, double Mo) { return q_sat(p_auto(T), p, Mo); }  // :298-301

// ---------------------------------------------------------------------------------------
// Mixture properties: MixtureAtm.h:82-217
// ---------------------------------------------------------------------------------------
struct Comp { double M_mean, R_mix, q_H2O, q_CO2, q_bg, M_bg, R_bg; };
Comp resolve(double xh, double xc, double xn) {  // MixtureAtm.h:82-109 (other species zero)
    Comp C;
    const double mh = xh * kMH2O, mc = xc * kMCO2;
    const double mb = xn * kMN2 + 0.0 * 0.016043 + 0.0 * 0.017031 + 0.0 * 0.002016
                    + 0.0 * 0.028010 + 0.0 * 0.064066;
    C.M_mean = mh + mc + mb;
    C.R_mix = kRstar / C.M_mean;
    C.q_H2O = mh / C.M_mean;
    C.q_CO2 = mc / C.M_mean;
    C.q_bg = mb / C.M_mean;
    const double xb = xn + 0.0 + 0.0 + 0.0 + 0.0 + 0.0;
    C.M_bg = (xb > 0.0) ? (mb / xb) : kMN2;
    C.R_bg = kRstar / C.M_bg;
    return C;
}
inline double shomate(double T, double A, double B, double Cc, double D, double E) {  // :123-128
    const double t = dmin(dmax(T, 298.0), 2000.0) * 1.0e-3;
    const double t2 = t * t;
    return A + B * t + Cc * t2 + D * t2 * t + E / t2;
}
inline double cp_h2o(double T) { return shomate(T, 30.09200, 6.832514, 6.793435, -2.534480, 0.082139) / kMH2O; }
inline double cp_co2(double T) { return shomate(T, 24.99735, 55.18696, -33.69137, 7.948387, -0.136638) / kMCO2; }
inline double cp_bg(double T, double Mb) { return shomate(T, 28.98641, 1.853978, -9.647459, 16.63537, 0.000117) / Mb; }
inline void split(double c, double co2, double& qv, double& qc, double& qb) {  // :152-162
    qv = dmin(dmax(c, 0.0), 1.0);
    qc = dmin(dmax(co2, 0.0), 1.0);
    const double s = qv + qc;
    if (s > 1.0) { qv /= s; qc /= s; }
    qb = dmax(0.0, 1.0 - qv - qc);
}
inline double R_of(double c, double co2, double Rb) {  // :165-170
    double qv, qc, qb; split(c, co2, qv, qc, qb);
    return qv * kRH2O + qc * kRCO2 + qb * Rb;
}
inline double cp_of(double c, double co2, double T, double Mb) {  // :173-178
    double qv, qc, qb; split(c, co2, qv, qc, qb);
    return qv * cp_h2o(T) + qc * cp_co2(T) + qb * cp_bg(T, Mb);
}
inline double M_nonwater(double c, double co2, double Mb) {  // :209-217
    double qv, qc, qb; split(c, co2, qv, qc, qb);
    const double s = qc + qb;
    if (!(s > 0.0)) return Mb;
    const double inv = (qc / s) / kMCO2 + (qb / s) / Mb;
    return (inv > 0.0) ? 1.0 / inv : Mb;
}

// ---------------------------------------------------------------------------------------
// Parameters: cli/config_cond.xml / atmosphere/cAtmosphereDefaults.cpp.inc
// ---------------------------------------------------------------------------------------
struct Params {
    double omega = 0.000317, g = 9.8066, r_Earth = 6370.001, t_0 = 273.15;
    double t_surf_equator = 513.15, t_surf_pole = 503.15, t_skin = 254.0, r_air = 40.8;
    double dt_visc = 4e-05, L_atm = 6287.5, zeta = 3.0;
    double sc_WaterVapour = 0.61, sc_CO2 = 0.96, pr_turb = 0.9, abl_height = 10000.0;
    double x_H2O = 0.5578, x_CO2 = 0.4109, x_N2 = 0.0313;
    double u_0 = 8.0, cloud_cap = 0.05, t_max_phys = 2000.0, t_00 = 236.15, t_000 = 253.15;
    double c_0 = 0.3464, c_h2o_dry_top = 0.00423, co2_0 = 0.6233, co2_scale = 1.0;
    // radiation (cAtmosphereDefaults.cpp.inc:31-43,72-73; omega_rad: cAtmosphereModel.cpp:801)
    double sigma = 5.67028e-08, kappa_H2O = 0.01, kappa_CO2 = 0.001, kappa_bg = 1e-06;
    double geothermal_flux = 150.0, rad_equator_short = 298.0, rad_pole_short = 0.0;
    double albedo_surface = 0.08, albedo_cloud = 0.5, t_skin_relax = 0.25, omega_rad = 0.05;
    Comp comp;
};

// Turbulence constants: TurbulenceAtm.h:26-35,158-161
constexpr double kCnue = 0.028;
constexpr double kNueAir = 1.5e-5;
constexpr double kDisMin = 1.0e-10;

}  // namespace

// ---------------------------------------------------------------------------------------
// State
// ---------------------------------------------------------------------------------------
struct orc_state {
    int im = 0, jm = 0, km = 0;
    size_t n = 0;
    double dr = 0, dthe = 0, dphi = 0;
    std::vector<double> rad, the, phi;
    std::vector<float> hgt;            // layer heights, float like cAtmosphereModel.h:387
    double shell_len = 0, metric_r0 = 0, re_turb = 0;
    Params P;
    std::vector<std::vector<double>> f;
    std::vector<std::vector<double>> f2;
    std::vector<double> t_ref;
    // radiation outputs: upward long-wave flux and emissivity per cell, albedo per column,
    // incoming short wave per latitude
    std::vector<double> radiation, epsilon, albedo, sw;
    // precipitation fluxes, their stored copies and the microphysics sources of the two-category
    // ice scheme (ORC_MP_* order); allocated, zero, at first use
    std::vector<std::vector<double>> mp;
    bool ice_on = false;               // the ice scheme runs inside the moist block
    bool conv_on = false;              // and moist convection after it (cAtmosphereModel.cpp:1247-1262)
    // moist-convection arrays (ORC_MC_* order) and the evaporation diagnostic [mm/d] per
    // column that its surface-flux perturbation reads; allocated, zero, at first use
    std::vector<std::vector<double>> mc;
    std::vector<double> evap;
    std::vector<double> cape;          // computeCAPE's im values (MoistConvection.h:1259-1299)
    int iter_n = 0, total_iter = 0;
    // the reference's O(jm + km) distance-to-land scans of the SST closure over the land mask
    // h (all water here) instead of their closed form min(i, j, k); for timing the reference's
    // own cost only, the results are the same
    bool sst_scan = false;
    std::vector<double> h;
    // the turbulence closure, TurbulenceAtm.h:24 order: 0 k-epsilon (Chien 1982), 1 k-omega
    // (Wilcox 1988), 2 k-omega SST (Menter 1994, the configuration's choice)
    int turb_model = 2;

    inline size_t ix(int i, int j, int k) const { return ((size_t)i * jm + j) * km + k; }
    inline double* F(int id) { return f[id].data(); }
    double metric_radius(double rm) const { return (metric_r0 > 0.0) ? (metric_r0 + (rm - rad[0])) : rm; }
};

namespace {

using S = orc_state;

// Grid coordinates and layer heights: cAtmosphereModel.cpp:44-60,327-331,382-399;
// cAtmosphereModel.h:441-469; lib/Array_1D.cpp:49-55 (coordinates by accumulation).
void init_grid(S& s, int im, int jm, int km) {
    s.im = im; s.jm = jm; s.km = km;
    s.n = (size_t)im * jm * km;
    const double pi180 = 180.0 / M_PI;
    const double the_degree = 180.0 / (double)(jm - 1);
    const double phi_degree = 360.0 / (double)(km - 1);
    s.dr = 1.0 / (double)(im - 1);
    s.dthe = the_degree / pi180;
    s.dphi = phi_degree / pi180;
    s.rad.assign(im, 0.0); s.the.assign(jm, 0.0); s.phi.assign(km, 0.0);
    s.rad[0] = 1.0;
    for (int l = 1; l < im; l++) s.rad[l] = s.rad[l - 1] + s.dr;
    for (int l = 1; l < jm; l++) s.the[l] = s.the[l - 1] + s.dthe;
    for (int l = 1; l < km; l++) s.phi[l] = s.phi[l - 1] + s.dphi;
    const float h = (float)s.P.L_atm;
    s.hgt.assign(im, 0.0f);
    for (int i = 0; i < im; i++)
        s.hgt[i] = (float)((std::exp(s.P.zeta * (s.rad[i] - s.rad[0])) - 1) * h);
    const double span = s.rad[im - 1] - s.rad[0];
    s.shell_len = (span > 0.0) ? (std::exp(s.P.zeta * span) - 1.0) * s.P.L_atm / span : s.P.L_atm;
    s.metric_r0 = (s.shell_len > 0.0) ? (s.P.r_Earth * 1.0e3 / s.shell_len) : 0.0;
    s.re_turb = 0.4 * 0.1 / 1.5e-5;   // TurbulenceAtm.h:35,46
    s.f.assign(ORC_NFIELD, std::vector<double>(s.n, 0.0));
    s.f2.assign(ORC2_NFIELD, std::vector<double>((size_t)jm * km, 0.0));
    s.t_ref.assign(im, 1.0);
}

// ---------------------------------------------------------------------------------------
// AtomUtils::damp_wiggles, lib/Utils.cpp:429-532 (strength 1, one pass, i_surface = 0)
// ---------------------------------------------------------------------------------------
void damp_wiggles(S& s, double* x) {
    const int im = s.im, jm = s.jm, km = s.km;
    std::vector<double> tmp(x, x + s.n);
    const double cf = 0.25;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) {
                const int km1 = (k - 1 + km) % km, kp1 = (k + 1) % km;
                const double c0 = tmp[s.ix(i, j, k)];
                x[s.ix(i, j, k)] = c0 + cf * (tmp[s.ix(i, j, km1)] - 2.0 * c0 + tmp[s.ix(i, j, kp1)]);
            }
    std::copy(x, x + s.n, tmp.begin());
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++) {
            const int jm1 = std::max(j - 1, 0), jp1 = std::min(j + 1, jm - 1);
            for (int k = 0; k < km; k++) {
                const double c0 = tmp[s.ix(i, j, k)];
                x[s.ix(i, j, k)] = c0 + cf * (tmp[s.ix(i, jm1, k)] - 2.0 * c0 + tmp[s.ix(i, jp1, k)]);
            }
        }
    std::copy(x, x + s.n, tmp.begin());
    #pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++)
            for (int i = 0; i < im; i++) {
                const int im1 = std::max(i - 1, 0), ip1 = std::min(i + 1, im - 1);
                const double c0 = tmp[s.ix(i, j, k)];
                x[s.ix(i, j, k)] = c0 + cf * (tmp[s.ix(im1, j, k)] - 2.0 * c0 + tmp[s.ix(ip1, j, k)]);
            }
}

// ---------------------------------------------------------------------------------------
// ThermoAtm::densities, ThermoAtm.h:1436-1614 (anelastic base state omitted: it feeds only
// the ATM_ANELASTIC / ATM_TRACER_DIFF_FLUX paths, both off by default)
// ---------------------------------------------------------------------------------------
void densities(S& s, bool set_water) {
    const Params& P = s.P;
    const double R_mix = P.comp.R_mix, M_bg = P.comp.M_bg, R_bg = P.comp.R_bg;
    const double p_sl_factor = 1e-2 * P.r_air * R_mix;
    double *t = s.F(ORC_T), *c = s.F(ORC_C), *co2 = s.F(ORC_CO2), *cl = s.F(ORC_CLOUD),
           *ice = s.F(ORC_ICE), *ps = s.F(ORC_P_STAT), *rh = s.F(ORC_R_HUMID), *rd = s.F(ORC_R_DRY);
    double* tti = s.f2[ORC2_T_TOP_INIT].data();
    #pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < s.jm; j++)
        for (int k = 0; k < s.km; k++) {
            double T_prev = dmax(180.0, t[s.ix(0, j, k)] * P.t_0);
            double p_prev = p_sl_factor * T_prev;
            double q_sat_min = 1.0;
            for (int i = 0; i < s.im; i++) {
                const size_t id = s.ix(i, j, k);
                const double qv = c[id], qc = co2[id];
                const double R_loc = R_of(qv, qc, R_bg);
                double Ti, pi;
                if (i == 0) { Ti = T_prev; pi = p_prev; }
                else {
                    const double dz = (double)s.hgt[i] - (double)s.hgt[i - 1];
                    const size_t ip = s.ix(i - 1, j, k);
                    const double qvp = c[ip], qcp = co2[ip];
                    const double cp_loc = cp_of(qvp, qcp, T_prev, M_bg);
                    const double M_nw = M_nonwater(qvp, qcp, M_bg);
                    const double q_s = q_sat_at(T_prev, p_prev, M_nw);
                    const double gamma = (qvp >= 0.99 * q_s)
                        ? moist_lapse(T_prev, p_prev, M_nw, cp_loc, R_loc, P.g) : P.g / cp_loc;
                    const double T_ad = T_prev - gamma * dz;
                    Ti = dmax(P.t_skin, T_ad);
                    const double T_mean = 0.5 * (T_prev + Ti);
                    pi = p_prev * std::exp(-P.g * dz / (R_loc * T_mean));
                }
                t[id] = Ti / P.t_0;
                ps[id] = pi;
                if (set_water && i > 0) {
                    const double M_nw_i = M_nonwater(qv, qc, M_bg);
                    const double q_s_i = q_sat_at(Ti, pi, M_nw_i);
                    if (q_s_i < q_sat_min) q_sat_min = q_s_i;
                    c[id] = dmax(q_sat_min, P.c_h2o_dry_top);
                }
                const double wf = dmax(0.5, 1.0 - cl[id] - ice[id]);
                rh[id] = 1e2 * pi / (R_loc * Ti * wf);
                const double R_dry = R_of(0.0, qc, R_bg);
                rd[id] = 1e2 * pi / (R_dry * Ti);
                T_prev = Ti; p_prev = pi;
            }
            tti[(size_t)j * s.km + k] = t[s.ix(s.im - 1, j, k)];
        }
}

// ThermoAtm::waterVapourEvaporation, sea branch: ThermoAtm.h:115-258
void evaporation(S& s) {
    const Params& P = s.P;
    const int n_spread = 3;
    const double r = std::exp(-1.0);
    const double w_norm = (1.0 - r) / (1.0 - std::pow(r, n_spread + 1));
    double *t = s.F(ORC_T), *c = s.F(ORC_C), *co2 = s.F(ORC_CO2), *ps = s.F(ORC_P_STAT);
    const double *u = s.F(ORC_U), *v = s.F(ORC_V), *w = s.F(ORC_W);
    // the Evaporation diagnostic [mm/d] moist convection reads, held once its arrays exist
    double* ev = s.evap.empty() ? nullptr : s.evap.data();
    #pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < s.jm; j++)
        for (int k = 0; k < s.km; k++) {
            const size_t i0 = s.ix(0, j, k);
            if (t[i0] * P.t_0 >= kTcrit) {
                if (ev) ev[(size_t)j * s.km + k] = 0.0;   // ThermoAtm.h:179-185
                continue;
            }
            const double T_s = dmax(180.0, t[i0] * P.t_0);
            const double p_s = ps[i0];
            const double M_nw = M_nonwater(c[i0], co2[i0], P.comp.M_bg);
            const double E_s = p_auto(T_s);
            const double q_sea = q_sat(E_s, p_s, M_nw);
            if (ev) {
                // Dalton's law on the saturation deficit (ThermoAtm.h:217-240) with
                // AtomUtils::C_Dalton (lib/Utils.cpp:157-172, coeff_Dalton 0.7)
                const double vel = std::sqrt((u[i0] * u[i0] + v[i0] * v[i0] + w[i0] * w[i0]) / 3.0) * P.u_0;
                const double cD = 0.7 * std::sqrt(0.053 * 0.053 * vel / 10.0);
                const double ca = c[i0];
                const double e_air = (q_sea > 0.0 && q_sea < 1.0) ? ca * p_s * M_nw / (ca * M_nw + (1.0 - ca) * kMH2O)
                                                                   : 0.0;
                ev[(size_t)j * s.km + k] = cD * 24.0 * dmax(0.0, E_s - e_air);
            }
            const double c_fix = c[i0];
            c[i0] = c_fix + (q_sea - c_fix) * w_norm;
            for (int i = 1; i <= n_spread && i < s.im; i++) {
                const size_t id = s.ix(i, j, k);
                const double weight = std::pow(r, i) * w_norm;
                const double T_i = dmax(180.0, t[id] * P.t_0);
                const double M_nw_i = M_nonwater(c[id], co2[id], P.comp.M_bg);
                const double q_s_i = q_sat_at(T_i, ps[id], M_nw_i);
                c[id] = dmin(c[id] + (q_sea - c_fix) * weight, q_s_i);
            }
        }
}

// ---------------------------------------------------------------------------------------
// SaturationAdjustment: SaturationAdjustment.h:56-287 (adjustSaturation), :289-301
// (applyTopography, identity without land), :303-451 (clampAndFade) and
// IceSchemeCommon.h:121-190 (evaporateWhereImpossible; the precipitation fluxes and sources
// are zero in the scoped step and exist once the ice scheme has run).
// ---------------------------------------------------------------------------------------
struct SatConsts { double t_0, t_00, M_bg, T_max, cloud_cap; };

void satadj_cell(const SatConsts& K, double& t, double& c, double& cloud, double& ice,
                 double co2, double p_local, double& s_c_c, double dt_dim) {
    const double inv_t_0 = 1.0 / K.t_0;
    const double t_range_inv = 1.0 / (K.t_0 - K.t_00);
    s_c_c = 0.0;
    const double qv_old = dmax(0.0, c), qc_old = dmax(0.0, cloud), qi_old = dmax(0.0, ice);
    double T = t * K.t_0;
    if (T >= kTcrit) return;
    const double M_other = M_nonwater(c, co2, K.M_bg);
    const double qs0 = q_sat(p_auto(T), p_local, M_other);
    const double alpha = 1.0 / (1.0 + std::exp(-(T - K.t_00) / 5.0));
    if (!((qv_old > qs0 && alpha > 0.01) || (qv_old < qs0 && (qc_old > 1e-12 || qi_old > 1e-12))))
        return;
    double qvb = qv_old, qcb = qc_old, qib = qi_old, qv_hyp = qvb;
    const double T_orig = t * K.t_0;
    for (int it = 1; it <= 20; it++) {
        const double CND = dmax(0.0, dmin(1.0, (T - K.t_00) * t_range_inv));
        const double DEP = 1.0 - CND;
        double dqv = qv_hyp - qvb;
        if (dqv > 0) { const double me = qcb + qib; if (dqv > me) dqv = me; }
        const double dcnd = dqv * CND, ddep = dqv * DEP;
        qvb += dqv;
        qcb = dmax(0.0, qcb - dcnd);
        qib = dmax(0.0, qib - ddep);
        const double cp_loc = cp_of(qvb, co2, T, K.M_bg);
        T -= (L_vap(T) * dcnd + L_sub(T) * ddep) / cp_loc;
        const double Es = p_sat(T), Ei = p_subl(T);
        const double qs = q_sat(Es, p_local, M_other), qi = q_sat(Ei, p_local, M_other);
        const double qsum = qcb + qib;
        const double target = (qsum > 1e-12) ? (qcb * qs + qib * qi) / qsum : ((T >= K.t_0) ? qs : qi);
        const double cp_g = cp_of(qvb, co2, T, K.M_bg);
        const double Lc = L_vap(T), Ld = L_sub(T);
        const double gain = CND * (Lc / cp_g) * dqdT_from(Es, Lc, T, p_local, M_other)
                          + DEP * (Ld / cp_g) * dqdT_from(Ei, Ld, T, p_local, M_other);
        const double om = 1.0 / (1.0 + gain);
        qv_hyp = qvb + om * (target - qvb);
        if (std::fabs(qvb / qv_hyp - 1.0) <= 1.0e-6) break;
    }
    qcb = dmax(0.0, qcb);
    qib = dmax(0.0, qib);
    if (T > kTcrit) T = kTcrit;
    if (!std::isnan(T) && !std::isnan(qvb)) {
        s_c_c = alpha * (qcb - qc_old) / dt_dim;
        c = qv_old + alpha * (qvb - qv_old);
        cloud = qc_old + alpha * (qcb - qc_old);
        ice = qi_old + alpha * (qib - qi_old);
        t = (T_orig + alpha * (T - T_orig)) * inv_t_0;
    }
    if (T < K.t_00) { cloud = 0.0; ice = 0.0; }
}

// mp: the ice-scheme arrays where they exist (else null) with the cell's r_humid and index.
void clamp_fade_cell(const SatConsts& K, double& t, double& c, double& cloud, double& ice,
                     double& gr, double co2, double p_local, double& s_c_c,
                     std::vector<std::vector<double>>* mp = nullptr, double r_humid = 0.0, size_t id = 0) {
    const double inv_t_0 = 1.0 / K.t_0;
    if (c < 0.0) c = 0.0;
    if (cloud < 0.0) cloud = 0.0;
    if (ice < 0.0) ice = 0.0;
    const double c_max = dmax(0.0, 1.0 - co2);
    if (c > c_max) c = c_max;
    double Td = t * K.t_0;
    if (Td > K.T_max) { Td = K.T_max; t = K.T_max * inv_t_0; }
    {
        const double M_o = M_nonwater(c, co2, K.M_bg);
        const double q_s = q_sat_at(Td, p_local, M_o);
        if (Td >= kTcrit || q_s >= 1.0) {
            // IceSchemeCommon.h:121-190
            const double q_cond = dmax(0.0, cloud) + dmax(0.0, ice) + dmax(0.0, gr);
            if (q_cond > 0.0) {
                const double cm = dmax(0.0, 1.0 - co2);
                const double c_new = dmin(cm, c + q_cond);
                const double q_used = dmax(0.0, c_new - c);
                c = c_new; cloud = 0.0; ice = 0.0; gr = 0.0;
                if (Td < kTcrit) {
                    const double cp_loc = cp_of(c, co2, Td, K.M_bg);
                    if (cp_loc > 0.0) {
                        double T_new = Td - L_vap(Td) * q_used / cp_loc;
                        if (T_new < 1.0) T_new = 1.0;
                        t = T_new / K.t_0;
                    }
                }
            }
            if (mp) {   // falling precipitation evaporates, the sources clear (:175-189)
                std::vector<std::vector<double>>& A = *mp;
                const double rho = dmax(1.0e-6, r_humid);
                const double q_p = dmax(0.0, A[ORC_MP_P_RAIN][id]) / (1.6 * rho)
                                 + dmax(0.0, A[ORC_MP_P_SNOW][id]) / (0.96 * rho) + 0.0 / (2.0 * rho);
                if (q_p > 0.0) {
                    const double cm = dmax(0.0, 1.0 - co2);
                    c = dmin(cm, c + q_p);
                }
                for (int q : {ORC_MP_P_RAIN, ORC_MP_P_SNOW, ORC_MP_PRECIP, ORC_MP_S_V, ORC_MP_S_C, ORC_MP_S_I,
                              ORC_MP_S_R, ORC_MP_S_S})
                    A[q][id] = 0.0;
            }
            s_c_c = 0.0;
            return;
        }
    }
    const double M_other = M_nonwater(c, co2, K.M_bg);
    const double qs = q_sat(p_auto(Td), p_local, M_other);
    if (c > qs) {
        const double excess = c - qs;
        c = qs;
        if (Td >= K.t_00) {
            cloud += excess;
            Td += L_vap(Td) * excess / cp_of(c, co2, Td, K.M_bg);
        } else {
            ice += excess;
            Td += L_sub(Td) * excess / cp_of(c, co2, Td, K.M_bg);
        }
        if (Td > K.T_max) Td = K.T_max;
        t = Td * inv_t_0;
    }
    if (cloud > K.cloud_cap) cloud = K.cloud_cap;
    if (ice > K.cloud_cap) ice = K.cloud_cap;
    const double al = 1.0 / (1.0 + std::exp(-(Td - K.t_00) / 5.0));
    c *= al; cloud *= al; ice *= al;
}

void satadj(S& s) {
    const Params& P = s.P;
    SatConsts K{P.t_0, P.t_00, P.comp.M_bg, P.t_max_phys, P.cloud_cap};
    double *t = s.F(ORC_T), *c = s.F(ORC_C), *cl = s.F(ORC_CLOUD), *ice = s.F(ORC_ICE),
           *gr = s.F(ORC_GR), *co2 = s.F(ORC_CO2), *ps = s.F(ORC_P_STAT), *scc = s.F(ORC_S_C_C);
    // SaturationAdjustment.h:56-60: step[i] is a float difference of float heights.
    std::vector<double> step(s.im, 0.0);
    for (int i = 0; i + 1 < s.im; i++) step[i] = (double)(float)(s.hgt[i + 1] - s.hgt[i]);
    for (int j = 0; j < s.jm; j++)
        for (int k = 0; k < s.km; k++) scc[s.ix(0, j, k)] = 0.0;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < s.im - 1; i++)
        for (int j = 0; j < s.jm; j++) {
            const double dt_dim = step[i] / 1.6;
            for (int k = 0; k < s.km; k++) {
                const size_t id = s.ix(i, j, k);
                satadj_cell(K, t[id], c[id], cl[id], ice[id], co2[id], ps[id], scc[id], dt_dim);
            }
        }
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < s.im; i++)
        for (int j = 0; j < s.jm; j++)
            for (int k = 0; k < s.km; k++) {
                const size_t id = s.ix(i, j, k);
                clamp_fade_cell(K, t[id], c[id], cl[id], ice[id], gr[id], co2[id], ps[id], scc[id],
                                s.mp.empty() ? nullptr : &s.mp, s.F(ORC_R_HUMID)[id], id);
            }
}

// cAtmosphereModel.cpp:1278-1399: cap_S (only S_c_c is nonzero in the scoped step), the
// microphysics clamp and the water-vapour ceiling.
void moist_clamp(S& s) {
    const double cap = s.P.cloud_cap;
    double *c = s.F(ORC_C), *cl = s.F(ORC_CLOUD), *ice = s.F(ORC_ICE), *co2 = s.F(ORC_CO2), *scc = s.F(ORC_S_C_C);
    #pragma omp parallel for schedule(static)
    for (size_t id = 0; id < s.n; id++) {
        double v = scc[id];
        if (!finite_bits(v)) scc[id] = 0.0;
        else if (v > 1.0e-3) scc[id] = 1.0e-3;
        else if (v < -1.0e-3) scc[id] = -1.0e-3;
        double cv = c[id], cld = cl[id], ic = ice[id];
        if (!finite_bits(cv)) cv = 0.0;
        if (!finite_bits(cld)) cld = 0.0;
        if (!finite_bits(ic)) ic = 0.0;
        if (cv < 0.0) cv = 0.0;
        if (cld < 0.0) cld = 0.0;
        if (ic < 0.0) ic = 0.0;
        const double cv_max = dmax(0.0, 1.0 - co2[id]);
        if (cv > cv_max) cv = cv_max;
        const double sum = cld + ic;
        if (sum > cap) { const double sc = cap / sum; cld *= sc; ic *= sc; }
        c[id] = cv; cl[id] = cld; ice[id] = ic;
        // cap_S of the ice scheme's sources, where it has run
        if (!s.mp.empty())
            for (int q : {ORC_MP_S_V, ORC_MP_S_C, ORC_MP_S_I, ORC_MP_S_R, ORC_MP_S_S}) {
                const double x = s.mp[q][id];
                s.mp[q][id] = !finite_bits(x) ? 0.0 : (x > 1.0e-3) ? 1.0e-3 : (x < -1.0e-3) ? -1.0e-3 : x;
            }
        // :1369-1390 (ceiling, every moist iteration)
        if (c[id] < 0.0) c[id] = 0.0;
        const double cc = dmax(0.0, 1.0 - co2[id]);
        if (c[id] > cc) c[id] = cc;
    }
}

// ---------------------------------------------------------------------------------------
// PressureSolverAtm::run, PressureSolverAtm.h:27-594 (land-free branches only)
// ---------------------------------------------------------------------------------------
void poisson(S& s) {
    const int im = s.im, jm = s.jm, km = s.km;
    double *au = s.F(ORC_AUX_U), *av = s.F(ORC_AUX_V), *aw = s.F(ORC_AUX_W), *p = s.F(ORC_P_DYN);
    std::vector<double> sint(jm);
    for (int j = 0; j < jm; j++) { sint[j] = std::sin(s.the[j]); if (sint[j] < 0.55) sint[j] = 0.55; }
    // :106-121 aux radial walls (int c43/c13 -> plain copy)
    for (int j = 1; j < jm - 1; j++)
        for (int k = 1; k < km - 1; k++) {
            au[s.ix(0, j, k)] = au[s.ix(1, j, k)]; au[s.ix(im - 1, j, k)] = au[s.ix(im - 2, j, k)];
            av[s.ix(0, j, k)] = av[s.ix(1, j, k)]; av[s.ix(im - 1, j, k)] = av[s.ix(im - 2, j, k)];
            aw[s.ix(0, j, k)] = aw[s.ix(1, j, k)]; aw[s.ix(im - 1, j, k)] = aw[s.ix(im - 2, j, k)];
        }
    const double inv_2dr = 1.0 / (2.0 * s.dr), inv_2dthe = 1.0 / (2.0 * s.dthe), inv_2dphi = 1.0 / (2.0 * s.dphi);
    const double inv_dr2 = 1.0 / (s.dr * s.dr), inv_dthe2 = 1.0 / (s.dthe * s.dthe), inv_dphi2 = 1.0 / (s.dphi * s.dphi);
    const size_t si = (size_t)jm * km, sj = km;
    for (int colour = 0; colour < 2; colour++) {
        #pragma omp parallel for collapse(2) schedule(static)
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++) {
                const double rm = s.rad[i];
                const double exp_rm = 1.0 / (rm + 1.0);
                const double exp_2_rm = exp_rm * exp_rm;
                const double sinthe = sint[j];
                const double sinthe2 = sinthe * sinthe;
                const double rmet = s.metric_radius(rm);
                const double rmet2 = rmet * rmet;
                const double inv_rm = 1.0 / rmet;
                const double inv_rm2 = 1.0 / rmet2;
                const double inv_rmsinthe = 1.0 / (rmet * sinthe);
                const double inv_rm2sinthe2 = inv_rm2 / sinthe2;
                const double m_the = inv_rm2, m_phi = inv_rm2sinthe2;
                const double denom = 2.0 * exp_2_rm * inv_dr2 + 2.0 * m_the * inv_dthe2 + 2.0 * m_phi * inv_dphi2;
                const double inv_denom = 1.0 / denom;
                const double num1 = exp_2_rm * inv_dr2, num2 = m_the * inv_dthe2, num3 = m_phi * inv_dphi2;
                for (int k = 1; k < km - 1; k++) {
                    if (((i + j + k) & 1) != colour) continue;
                    const size_t id = s.ix(i, j, k);
                    const double du_dr = (au[id + si] - au[id - si]) * inv_2dr;
                    const double dv_dthe = (av[id + sj] - av[id - sj]) * inv_2dthe;
                    const double dw_dphi = (aw[id + 1] - aw[id - 1]) * inv_2dphi;
                    double div_src = du_dr * exp_rm + dv_dthe * inv_rm + dw_dphi * inv_rmsinthe;
                    const double src_max = denom * 2.0;
                    if (!finite_bits(div_src)) div_src = 0.0;
                    else if (div_src > src_max) div_src = src_max;
                    else if (div_src < -src_max) div_src = -src_max;
                    p[id] = ((p[id + si] + p[id - si]) * num1 + (p[id + sj] + p[id - sj]) * num2
                           + (p[id + 1] + p[id - 1]) * num3 - div_src) * inv_denom;
                }
            }
    }
    // :504-552 boundaries
    for (int k = 0; k < km; k++)
        for (int j = 0; j < jm; j++) {
            p[s.ix(0, j, k)] = p[s.ix(1, j, k)];
            p[s.ix(im - 1, j, k)] = p[s.ix(im - 2, j, k)];
        }
    for (int k = 0; k < km; k++)
        for (int i = 0; i < im; i++) {
            p[s.ix(i, 0, k)] = p[s.ix(i, 1, k)];
            p[s.ix(i, jm - 1, k)] = p[s.ix(i, jm - 2, k)];
        }
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++) {
            const double a = p[s.ix(i, j, 1)], b = p[s.ix(i, j, km - 2)];
            p[s.ix(i, j, 0)] = p[s.ix(i, j, km - 1)] = (a + b) / 2.0;
        }
    // :583-594 ceiling
    const double ceil_ = (s.total_iter > 300) ? 3.0 : 10.0;
    for (size_t id = 0; id < s.n; id++) {
        const double v = p[id];
        if (!finite_bits(v)) p[id] = 0.0;
        else if (v > ceil_) p[id] = ceil_;
        else if (v < -ceil_) p[id] = -ceil_;
    }
}

// PressureSolverAtm::project_initial_velocity (PressureSolverAtm.h:686-767), called once in
// RunTimeSlice (cAtmosphereModel.cpp:665): the velocity as the Poisson source, n sweeps of the
// solver's run(), the gradient correction v -= grad p in the RHS metric, then p_dyn and aux
// cleared.
void project_initial_velocity(S& s, int n_sweeps) {
    const int im = s.im, jm = s.jm, km = s.km;
    double *u = s.F(ORC_U), *v = s.F(ORC_V), *w = s.F(ORC_W), *p = s.F(ORC_P_DYN);
    double *au = s.F(ORC_AUX_U), *av = s.F(ORC_AUX_V), *aw = s.F(ORC_AUX_W);
    for (size_t q = 0; q < s.n; q++) { au[q] = u[q]; av[q] = v[q]; aw[q] = w[q]; p[q] = 0.0; }   // :694-706
    for (int it = 0; it < n_sweeps; it++) poisson(s);                                         // :711-713
    const double inv_2dr = 1.0 / (2.0 * s.dr), inv_2dthe = 1.0 / (2.0 * s.dthe), inv_2dphi = 1.0 / (2.0 * s.dphi);
    std::vector<double> sint(jm);
    for (int j = 0; j < jm; j++) { sint[j] = std::sin(s.the[j]); if (sint[j] < 0.55) sint[j] = 0.55; }
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < im - 1; i++)                                                          // :726-749
        for (int j = 1; j < jm - 1; j++) {
            const double rm = s.rad[i];
            const double exp_rm = 1.0 / (rm + 1.0);
            const double rmet = s.metric_radius(rm);
            const double inv_rm = 1.0 / rmet;
            const double inv_rmsinthe = 1.0 / (rmet * sint[j]);
            for (int k = 1; k < km - 1; k++) {
                const size_t id = s.ix(i, j, k);
                const double dpdr = (p[s.ix(i + 1, j, k)] - p[s.ix(i - 1, j, k)]) * inv_2dr;
                const double dpdthe = (p[s.ix(i, j + 1, k)] - p[s.ix(i, j - 1, k)]) * inv_2dthe;
                const double dpdphi = (p[s.ix(i, j, k + 1)] - p[s.ix(i, j, k - 1)]) * inv_2dphi;
                u[id] -= dpdr * exp_rm;
                v[id] -= dpdthe * inv_rm;
                w[id] -= dpdphi * inv_rmsinthe;
            }
        }
    for (size_t q = 0; q < s.n; q++) { p[q] = 0.0; au[q] = 0.0; av[q] = 0.0; aw[q] = 0.0; }   // :752-763
}

// ---------------------------------------------------------------------------------------
// TurbulenceAtm::run: compute_vel_star (:88-102), compute_sources (:323-477, serial in
// (j,k,i) order = the OMP_NUM_THREADS=1 sweep), compute_k_omega_SST (:656-773),
// zero_land_cells (no-op), clamp_nue (:271-297)
// ---------------------------------------------------------------------------------------
void vel_star(S& s) {
    const int im = s.im, jm = s.jm, km = s.km;
    const double *v = s.F(ORC_V), *w = s.F(ORC_W);
    double* vst = s.f2[ORC2_VEL_STAR].data();
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++) {
            const int i_surf = std::min(0 + 1, im - 1);
            const double vs = v[s.ix(i_surf, j, k)], ws = w[s.ix(i_surf, j, k)];
            const double U = s.P.u_0 * std::sqrt(vs * vs + ws * ws);
            vst[(size_t)j * km + k] = dmax(0.05, std::sqrt(0.002) * U);
        }
}

void turbulence(S& s) {
    const Params& P = s.P;
    const int im = s.im, jm = s.jm, km = s.km;
    double *u = s.F(ORC_U), *v = s.F(ORC_V), *w = s.F(ORC_W), *tke = s.F(ORC_TKE), *dis = s.F(ORC_DIS),
           *nue = s.F(ORC_NUE), *prod = s.F(ORC_PROD), *tks = s.F(ORC_TKE_SRC), *dss = s.F(ORC_DIS_SRC);
    const double* vst = s.f2[ORC2_VEL_STAR].data();
    vel_star(s);
    const double nue_max = 1000.0 / (P.u_0 * P.L_atm);
    const double y_mount = (double)s.hgt[0];
    const double inv_2dr = 1.0 / (2.0 * s.dr), inv_2dthe = 1.0 / (2.0 * s.dthe), inv_2dphi = 1.0 / (2.0 * s.dphi);
    const size_t si = (size_t)jm * km, sj = km;
    // SST constants :664-673
    const double a1 = 0.31, bet_star = 0.09, bet1 = 0.0333, bet2 = 0.0368, gam1 = 0.413, gam2 = 0.2;
    const double sig_w2 = 1.168;
    const double nue_air_nd = kNueAir / (P.u_0 * P.L_atm);
    for (int j = 1; j < jm - 1; j++)
        for (int k = 1; k < km - 1; k++) {
            const double vs = vst[(size_t)j * km + k];
            const double k_bg_nd = 0.01 * vs * vs / (std::sqrt(kCnue) * P.u_0 * P.u_0);
            for (int i = 1; i < im - 1; i++) {
                const size_t id = s.ix(i, j, k);
                const double rm = s.rad[i];
                const double exp_rm = 1.0 / (rm + 1.0);
                double sinthe = std::sin(s.the[j]);
                if (sinthe == 0.0) sinthe = 1.0e-5;
                const double rmsinthe = rm * sinthe;
                const double rmet = s.metric_radius(rm);
                const double inv_rm2dthe = inv_2dthe / rmet;
                const double inv_rmsinthe2dphi = inv_2dphi / (rmet * sinthe);
                tke[id] = dmax(0.0, tke[id]);
                dis[id] = dmax(kDisMin, dis[id]);
                const double dudr = (u[id + si] - u[id - si]) * inv_2dr * exp_rm;
                const double dvdr = (v[id + si] - v[id - si]) * inv_2dr * exp_rm;
                const double dwdr = (w[id + si] - w[id - si]) * inv_2dr * exp_rm;
                const double dudthe = (u[id + sj] - u[id - sj]) * inv_rm2dthe;
                const double dvdthe = (v[id + sj] - v[id - sj]) * inv_rm2dthe;
                const double dwdthe = (w[id + sj] - w[id - sj]) * inv_rm2dthe;
                const double dudphi = (u[id + 1] - u[id - 1]) * inv_rmsinthe2dphi;
                const double dvdphi = (v[id + 1] - v[id - 1]) * inv_rmsinthe2dphi;
                const double dwdphi = (w[id + 1] - w[id - 1]) * inv_rmsinthe2dphi;
                const double dtkedr = (tke[id + si] - tke[id - si]) * inv_2dr * exp_rm;
                const double ddisdr = (dis[id + si] - dis[id - si]) * inv_2dr * exp_rm;
                const double tk = tke[id], ds = dis[id];
                double cnue;
                {
                    const double dis_here = dmax(ds, kDisMin);
                    const double tke_here = dmax(tk, 0.0);
                    cnue = (s.turb_model == 0) ? kCnue * tke_here * tke_here / dis_here : tke_here / dis_here;
                    cnue = dmin(cnue, nue_max);
                }
                const double der = 0.66667 * (dudr + dvdthe + dwdphi);
                prod[id] = dmax(0.0,
                      (cnue * (2.0 * dudr - der) - 0.66667 * tk) * dudr
                    + (cnue * (dudthe + dvdr)) * dudthe
                    + (cnue * (dudphi + dwdr)) * dudphi
                    + (cnue * (2.0 * dvdthe - der) - 0.66667 * tk) * dvdthe
                    + (cnue * (dvdr + dudthe)) * dvdr
                    + (cnue * (dvdphi + dwdthe)) * dvdphi
                    + (cnue * (2.0 * dwdphi - der) - 0.66667 * tk) * dwdphi
                    + (cnue * (dwdr + dudphi)) * dwdr
                    + (cnue * (dwdthe + dvdphi)) * dwdthe);
                const double W12 = dudthe - dvdr, W13 = dudphi - dwdr, W23 = dvdphi - dwdthe;
                const double Omega = std::sqrt(W12 * W12 + W13 * W13 + W23 * W23);
                if (s.turb_model == 0) {
                    // compute_k_epsilon, TurbulenceAtm.h:481-530 (no land: y_mount = hgt[0])
                    const double C_eps_1 = 1.35, C_eps_2 = 1.80;
                    const double y_phys = dmax((double)s.hgt[i] - y_mount, P.L_atm);
                    const double y_star = y_phys / P.L_atm;
                    nue[id] = kCnue * tk * tk / dmax(ds, kDisMin);
                    const double d_plus = y_phys * vs / kNueAir;
                    const double f_nue = 1.0 - std::exp(-0.0115 * d_plus);
                    const double Re_T = tk * tk * P.u_0 * P.L_atm / dmax(ds * kNueAir, kDisMin * kNueAir);
                    const double f_2 = 1.0 - 0.4 / 1.8 * std::exp(-Re_T * Re_T / 36.0);
                    nue[id] *= f_nue;
                    const double y_star2 = y_star * y_star;
                    const double L_k = -2.0 * tk / y_star2;
                    const double L_w = -2.0 * ds / y_star2 * std::exp(-d_plus / 2.0);
                    const double P_k = prod[id];
                    const double Y_k = ds;
                    const double tke_safe = dmax(tk, kDisMin);
                    const double P_w = C_eps_1 * ds / tke_safe * P_k;
                    const double Y_w = C_eps_2 * f_2 * ds * ds / tke_safe;
                    tks[id] = P_k - Y_k + L_k / s.re_turb;
                    dss[id] = P_w - Y_w + L_w / s.re_turb;
                } else if (s.turb_model == 1) {
                    // compute_k_omega, TurbulenceAtm.h:533-651 (no land: the Neumann variants
                    // are the plain central differences; grid rm in the theta/phi gradients)
                    const double bet_star = 0.09, gam = 0.52, C_lim = 0.875, bet_0 = 0.0708;
                    const double dtkedthe_n = (tke[id + sj] - tke[id - sj]) / (rm * 2.0 * s.dthe);
                    const double ddisdthe_n = (dis[id + sj] - dis[id - sj]) / (rm * 2.0 * s.dthe);
                    const double dtkedphi_n = (tke[id + 1] - tke[id - 1]) / (rmsinthe * 2.0 * s.dphi);
                    const double ddisdphi_n = (dis[id + 1] - dis[id - 1]) / (rmsinthe * 2.0 * s.dphi);
                    const double dudr_k = (u[id + si] - u[id - si]) / (2.0 * s.dr) / (rm + 1.0);
                    const double dvdthe_k = (v[id + sj] - v[id - sj]) / (rm * 2.0 * s.dthe);
                    const double dwdphi_k = (w[id + 1] - w[id - 1]) / (rmsinthe * 2.0 * s.dphi);
                    const double dudthe_u = (u[id + sj] - u[id - sj]) / (rm * 2.0 * s.dthe);
                    const double dudphi_u = (u[id + 1] - u[id - 1]) / (rmsinthe * 2.0 * s.dphi);
                    const double dvdr_u = (v[id + si] - v[id - si]) / (2.0 * s.dr) / (rm + 1.0);
                    const double dwdr_u = (w[id + si] - w[id - si]) / (2.0 * s.dr) / (rm + 1.0);
                    const double dvdphi_u = (v[id + 1] - v[id - 1]) / (rmsinthe * 2.0 * s.dphi);
                    const double dwdthe_u = (w[id + sj] - w[id - sj]) / (rm * 2.0 * s.dthe);
                    const double S11 = dudr_k, S22 = dvdthe_k, S33 = dwdphi_k;
                    const double S12 = 0.5 * (dudthe_u + dvdr_u);
                    const double S13 = 0.5 * (dudphi_u + dwdr_u);
                    const double S23 = 0.5 * (dvdphi_u + dwdthe_u);
                    const double S_mag = std::sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33
                                                          + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
                    const double w_hat = dmax(ds, C_lim * S_mag / std::sqrt(bet_star));
                    const double cross = dtkedr * ddisdr + dtkedthe_n * ddisdthe_n + dtkedphi_n * ddisdphi_n;
                    const double sig_d = (cross <= 0.0) ? 0.0 : 0.125;
                    const double chi_w = std::fabs(Omega * Omega * S_mag / std::pow(bet_star * dmax(ds, 1.0e-20), 3));
                    const double f_bet = (1.0 + 85.0 * chi_w) / (1.0 + 100.0 * chi_w);
                    const double bet_wc = bet_0 * f_bet;
                    nue[id] = tk / dmax(w_hat, 1.0e-20);
                    const double P_k = dmin(prod[id], 20.0 * bet_star * tk * ds);
                    const double Y_k = bet_star * tk * ds;
                    const double P_w = gam * ds / dmax(tk, 1.0e-20) * P_k;
                    const double Y_w = bet_wc * ds * ds;
                    const double D_w = sig_d / dmax(ds, 1.0e-20) * cross;
                    tks[id] = P_k - Y_k;
                    dss[id] = P_w - Y_w + D_w;
                } else {
                    // compute_k_omega_SST, TurbulenceAtm.h:656-773
                    const double y_phys = dmax((double)s.hgt[i] - y_mount, 1.0e-6);
                    const double y_star = y_phys / P.L_atm;
                    const double dtkedthe_n = (tke[id + sj] - tke[id - sj]) / (rm * 2.0 * s.dthe);
                    const double ddisdthe_n = (dis[id + sj] - dis[id - sj]) / (rm * 2.0 * s.dthe);
                    const double dtkedphi_n = (tke[id + 1] - tke[id - 1]) / (rmsinthe * 2.0 * s.dphi);
                    const double ddisdphi_n = (dis[id + 1] - dis[id - 1]) / (rmsinthe * 2.0 * s.dphi);
                    const double cross = dtkedr * ddisdr + dtkedthe_n * ddisdthe_n + dtkedphi_n * ddisdphi_n;
                    const double CD_kw = dmax(2.0 * sig_w2 / dmax(ds, 1.0e-20) * cross, 1.0e-20);
                    const double dis_safe = dmax(ds, 1.0e-20);
                    const double arg1 = dmin(
                        dmax(std::sqrt(tk) / (bet_star * dis_safe * y_star),
                             500.0 * nue_air_nd / (y_star * y_star * dis_safe)),
                        4.0 * sig_w2 * tk / (CD_kw * y_star * y_star));
                    const double arg2 = dmax(2.0 * std::sqrt(tk) / (bet_star * dis_safe * y_star),
                                             500.0 * nue_air_nd / (y_star * y_star * dis_safe));
                    const double F1 = std::tanh(std::pow(arg1, 4));
                    const double F2 = std::tanh(std::pow(arg2, 2));
                    nue[id] = a1 * tk / dmax(a1 * dis_safe, Omega * F2);
                    const double tke_safe = dmax(tk, 1.0e-20);
                    const double P_k = dmin(prod[id], 20.0 * bet_star * tke_safe * dis_safe);
                    const double Y_k = bet_star * tke_safe * dis_safe;
                    const double P_w = (F1 * gam1 + (1.0 - F1) * gam2) * P_k * dis_safe / tke_safe;
                    const double Y_w = (F1 * bet1 + (1.0 - F1) * bet2) * dis_safe * dis_safe;
                    const double D_w = 2.0 * (1.0 - F1) * sig_w2 / dis_safe * cross;
                    tks[id] = P_k - Y_k;
                    dss[id] = P_w - Y_w + D_w;
                }
                tke[id] = dmax(k_bg_nd, tke[id]);
                dis[id] = dmax(kDisMin, dis[id]);
                nue[id] = dclamp(nue[id], 0.0, nue_max);
                const double dsm = 20.0 * dis[id] * dis[id];
                dss[id] = dclamp(dss[id], -dsm, dsm);
                const double tsm = 20.0 * tke[id] * dis[id];
                tks[id] = dclamp(tks[id], -tsm, tsm);
            }
        }
    // clamp_nue :271-297
    int i_abl = 0;
    while (i_abl < im - 1 && (double)s.hgt[i_abl] <= P.abl_height) i_abl++;
    const double taper_dz = dmax(1.0, (double)s.hgt[std::min(i_abl, im - 1)] - (double)s.hgt[std::max(i_abl - 1, 0)]);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < im; i++) {
        const double z_agl = (double)s.hgt[i] - (double)s.hgt[0];
        const double fct = dclamp((P.abl_height + taper_dz - z_agl) / taper_dz, 0.0, 1.0);
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) {
                const size_t id = s.ix(i, j, k);
                nue[id] = dclamp(nue[id], 0.0, nue_max) * fct;
            }
    }
}

// ---------------------------------------------------------------------------------------
// Boundary conditions, every iteration: BC_Atm::bcRadius (:556-745),
// TurbulenceAtm::apply_wall_bc (:109-147), bcTheta (:750-798), bcPhi (:803-917),
// bcScalarSurfSur (:1026-1130, identity without land), bcSolidGround (:23-551: passes 1-3
// are identities without land; passes 4-5 are the clamps)
// ---------------------------------------------------------------------------------------
// TurbulenceAtm::apply_wall_bc (:109-147): the log-layer tke, zero eddy viscosity and the
// closure's wall value of dis at i = 0 (k-epsilon: 0; k-omega: 6 nu / (0.075 y1^2); SST:
// 60 nu / (0.0333 y1^2), each made dimensionless by L_atm / u_0)
void wall_bc(S& s) {
    const Params& P = s.P;
    const double* vst = s.f2[ORC2_VEL_STAR].data();
    const double y1 = (double)s.hgt[1];
    const double y1s = dmax(y1, 1.0e-6);
    const double nd_omega = P.L_atm / P.u_0;
    for (int j = 0; j < s.jm; j++)
        for (int k = 0; k < s.km; k++) {
            const double vs = vst[(size_t)j * s.km + k];
            s.f[ORC_TKE][s.ix(0, j, k)] = vs * vs / (std::sqrt(kCnue) * P.u_0 * P.u_0);
            s.f[ORC_NUE][s.ix(0, j, k)] = 0.0;
            double d0 = 0.0;
            if (s.turb_model == 1) d0 = 6.0 * kNueAir / (0.075 * y1s * y1s) * nd_omega;
            else if (s.turb_model == 2) d0 = 60.0 * kNueAir / (0.0333 * y1s * y1s) * nd_omega;
            s.f[ORC_DIS][s.ix(0, j, k)] = d0;
        }
}

// TurbulenceAtm::init (:39-59): compute_vel_star, init_fields (:193-255, the boundary-layer
// profile k = u_tau^2 / sqrt(C_mu) (1 - z/h_abl)^2 above a 1 % background floor, with
// omega = C_mu^-1/4 sqrt(k) / (kappa z) or epsilon = C_mu^3/4 k^3/2 / (kappa z) above theirs),
// apply_wall_bc, then compute_sources and clamp_nue as in run (zero_land_cells is an identity
// without land). m.re_turb is set as the state's own.
void turbulence_init(S& s) {
    const Params& P = s.P;
    const int im = s.im, jm = s.jm, km = s.km;
    vel_star(s);
    const double* vst = s.f2[ORC2_VEL_STAR].data();
    double *tke = s.F(ORC_TKE), *dis = s.F(ORC_DIS), *tkn = s.F(ORC_TKEN), *dsn = s.F(ORC_DISN);
    const double kKarman = 0.42;
    const double inv_u02 = 1.0 / (P.u_0 * P.u_0);
    const double nd_omega = P.L_atm / P.u_0;
    const double nd_eps = P.L_atm / (P.u_0 * P.u_0 * P.u_0);
    for (int j = 1; j < jm - 1; j++)
        for (int k = 1; k < km - 1; k++) {
            const double vs = vst[(size_t)j * km + k];
            const double k_bg = 0.01 * vs * vs / std::sqrt(kCnue);
            const double om_bg = std::pow(kCnue, -0.25) * std::sqrt(k_bg) / (kKarman * P.L_atm);
            const double eps_bg = std::pow(kCnue, 0.75) * std::pow(k_bg, 1.5) / (kKarman * P.L_atm);
            const double y_mount = (double)s.hgt[0];
            for (int i = 1; i < im - 1; i++) {
                const double z_i = (double)s.hgt[i] - y_mount;
                const double z_safe = dmax(z_i, 1.0e-6);
                const double z_frac = dmin(z_i, P.abl_height) / P.abl_height;
                const double k_abl = vs * vs / std::sqrt(kCnue) * std::pow(1.0 - z_frac, 2);
                const double k_phys = dmax(k_abl, k_bg);
                double dis_nd;
                if (s.turb_model == 0) {
                    const double eps_phys = std::pow(kCnue, 0.75) * std::pow(k_phys, 1.5) / (kKarman * z_safe);
                    dis_nd = dmax(eps_phys, eps_bg) * nd_eps;
                } else {
                    const double om_phys = std::pow(kCnue, -0.25) * std::sqrt(k_phys) / (kKarman * z_safe);
                    dis_nd = dmax(om_phys, om_bg) * nd_omega;
                }
                const size_t id = s.ix(i, j, k);
                tke[id] = dmax(0.0, k_phys * inv_u02);
                dis[id] = dmax(kDisMin, dis_nd);
                tkn[id] = tke[id];
                dsn[id] = dis[id];
            }
        }
    wall_bc(s);
    turbulence(s);
}

void bc_all(S& s) {
    const Params& P = s.P;
    const int im = s.im, jm = s.jm, km = s.km, iml = im - 1;
    double* t = s.F(ORC_T);
    const double* tti = s.f2[ORC2_T_TOP_INIT].data();
    const int hydro[3] = {ORC_P_STAT, ORC_R_HUMID, ORC_R_DRY};
    const int cubic_vn[10] = {ORC_CLOUD, ORC_ICE, ORC_GR, ORC_CO2, ORC_TKE, ORC_DIS, ORC_NUE,
                              ORC_PROD, ORC_TKE_SRC, ORC_DIS_SRC};
    // bcRadius
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++) {
            auto X = [&](int fld, int i) -> double& { return s.f[fld][s.ix(i, j, k)]; };
            X(ORC_T, iml) = tti[(size_t)j * km + k];
            for (int h : hydro) {
                X(h, 0) = X(h, 3) - 3.0 * X(h, 2) + 3.0 * X(h, 1);
                const double x1 = X(h, iml - 1), x2 = X(h, iml - 2);
                X(h, iml) = (x1 > 0.0 && x2 > 0.0) ? (x1 * x1 / x2) : x1;
            }
            for (int fld : {ORC_U, ORC_V, ORC_W}) { X(fld, 0) = X(fld, 1); X(fld, iml) = X(fld, iml - 1); }
            X(ORC_V, iml) = 0.0; X(ORC_W, iml) = 0.0; X(ORC_VN, iml) = 0.0; X(ORC_WN, iml) = 0.0;
            for (int fld : {ORC_V, ORC_W, ORC_VN, ORC_WN}) X(fld, iml - 1) *= (1.0 / 3.0);
            for (int fld : {ORC_V, ORC_W, ORC_VN, ORC_WN}) X(fld, iml - 2) *= (2.0 / 3.0);
            X(ORC_U, iml) = 0.0; X(ORC_UN, iml) = 0.0; X(ORC_U, 0) = 0.0; X(ORC_UN, 0) = 0.0;
            for (int fld : cubic_vn) {
                X(fld, 0) = X(fld, 3) - 3.0 * X(fld, 2) + 3.0 * X(fld, 1);
                X(fld, iml) = X(fld, iml - 1);
            }
        }
    wall_bc(s);
    // bcTheta
    const int theta_f[18] = {ORC_T, ORC_P_STAT, ORC_R_HUMID, ORC_R_DRY, ORC_U, ORC_V, ORC_W, ORC_C,
                             ORC_CLOUD, ORC_ICE, ORC_GR, ORC_CO2, ORC_TKE, ORC_DIS, ORC_NUE, ORC_PROD,
                             ORC_TKE_SRC, ORC_DIS_SRC};
    for (int fld : theta_f) {
        double* x = s.F(fld);
        for (int k = 0; k < km; k++)
            for (int i = 0; i < im; i++) {
                x[s.ix(i, 0, k)] = x[s.ix(i, 1, k)];
                x[s.ix(i, jm - 1, k)] = x[s.ix(i, jm - 2, k)];
            }
    }
    // bcPhi
    const int avg_f[12] = {ORC_T, ORC_P_STAT, ORC_R_HUMID, ORC_R_DRY, ORC_U, ORC_V, ORC_W, ORC_C,
                           ORC_CLOUD, ORC_ICE, ORC_GR, ORC_CO2};
    const int ext_f[6] = {ORC_TKE, ORC_DIS, ORC_NUE, ORC_PROD, ORC_TKE_SRC, ORC_DIS_SRC};
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++) {
            for (int fld : avg_f) {
                double* x = s.F(fld) + s.ix(i, j, 0);
                const double v0 = x[1], ve = x[km - 2];
                x[0] = x[km - 1] = (v0 + ve) * 0.5;
            }
            for (int fld : ext_f) {
                double* x = s.F(fld) + s.ix(i, j, 0);
                x[0] = x[1];
                x[km - 1] = x[km - 2];
            }
        }
    {
        const int seam_f[6] = {ORC_U, ORC_UN, ORC_V, ORC_VN, ORC_W, ORC_WN};
        for (int i = 0; i < im; i++)
            for (int j = 0; j < jm; j++)
                for (int pass = 0; pass < 2; pass++)
                    for (int fld : seam_f) {
                        double* x = s.F(fld) + s.ix(i, j, 0);
                        const double a = x[km - 2], b = x[0], c1 = x[1];
                        const double w_a = x[km - 3], e_a = b, w_b = a, e_b = c1, w_c = b, e_c = x[2];
                        x[km - 2] = a + 0.25 * (w_a - 2.0 * a + e_a);
                        x[0] = x[km - 1] = b + 0.25 * (w_b - 2.0 * b + e_b);
                        x[1] = c1 + 0.25 * (w_c - 2.0 * c1 + e_c);
                    }
    }
    // bcSolidGround passes 4-5
    const double tke_max_nd = 1000.0 / (P.u_0 * P.u_0);
    const double nue_max = 1000.0 / (P.u_0 * P.L_atm);
    const double vmax = 100.0 / P.u_0;
    auto sc = [](double v, double lo, double hi) {
        if (!finite_bits(v)) return lo;
        return (v < lo) ? lo : (v > hi) ? hi : v;
    };
    auto scs = [](double v, double m) {
        if (!finite_bits(v)) return 0.0;
        return (v < -m) ? -m : (v > m) ? m : v;
    };
    #pragma omp parallel for schedule(static)
    for (size_t id = 0; id < s.n; id++) {
        for (int fld : {ORC_C, ORC_CLOUD, ORC_ICE, ORC_TKE, ORC_DIS})
            if (s.f[fld][id] < 0.0) s.f[fld][id] = 0.0;
        s.f[ORC_TKE][id] = sc(s.f[ORC_TKE][id], 0.0, tke_max_nd);
        s.f[ORC_TKEN][id] = sc(s.f[ORC_TKEN][id], 0.0, tke_max_nd);
        s.f[ORC_DIS][id] = sc(s.f[ORC_DIS][id], 1.0e-10, 1.0e6);
        s.f[ORC_DISN][id] = sc(s.f[ORC_DISN][id], 1.0e-10, 1.0e6);
        s.f[ORC_NUE][id] = sc(s.f[ORC_NUE][id], 0.0, nue_max);
        s.f[ORC_PROD][id] = sc(s.f[ORC_PROD][id], 0.0, 1.0e4);
        s.f[ORC_TKE_SRC][id] = sc(s.f[ORC_TKE_SRC][id], -1.0e4, 1.0e4);
        s.f[ORC_DIS_SRC][id] = sc(s.f[ORC_DIS_SRC][id], -1.0e4, 1.0e4);
        for (int fld : {ORC_U, ORC_UN, ORC_V, ORC_VN, ORC_W, ORC_WN}) s.f[fld][id] = scs(s.f[fld][id], vmax);
    }
    // Microphysics sources S_c_c, S_v, S_c, S_i, S_r, S_s where the ice scheme has run:
    // bcRadius pattern A (cubic at both ends, BC_Atm.h:567-644), bcTheta copy (:761-790), bcPhi
    // seam average (:805-842). Their boundary faces are read by no operator of the step.
    if (!s.mp.empty()) {
        double* srcs[6] = {s.F(ORC_S_C_C), s.mp[ORC_MP_S_V].data(), s.mp[ORC_MP_S_C].data(),
                           s.mp[ORC_MP_S_I].data(), s.mp[ORC_MP_S_R].data(), s.mp[ORC_MP_S_S].data()};
        const double c43 = 1.0, c13 = 0.0;
        for (double* x : srcs) {
            for (int j = 0; j < jm; j++)
                for (int k = 0; k < km; k++) {
                    x[s.ix(0, j, k)] = x[s.ix(3, j, k)] - 3.0 * x[s.ix(2, j, k)] + 3.0 * x[s.ix(1, j, k)];
                    x[s.ix(iml, j, k)] = x[s.ix(iml - 3, j, k)] - 3.0 * x[s.ix(iml - 2, j, k)] + 3.0 * x[s.ix(iml - 1, j, k)];
                }
            for (int k = 0; k < km; k++)
                for (int i = 0; i < im; i++) {
                    x[s.ix(i, 0, k)] = x[s.ix(i, 1, k)];
                    x[s.ix(i, jm - 1, k)] = x[s.ix(i, jm - 2, k)];
                }
            for (int i = 0; i < im; i++)
                for (int j = 0; j < jm; j++) {
                    const double v0 = c43 * x[s.ix(i, j, 1)] - c13 * x[s.ix(i, j, 2)];
                    const double ve = c43 * x[s.ix(i, j, km - 2)] - c13 * x[s.ix(i, j, km - 3)];
                    x[s.ix(i, j, 0)] = x[s.ix(i, j, km - 1)] = (v0 + ve) * 0.5;
                }
        }
    }
    // Moist-convection arrays, where they exist. bcRadius: pattern A (cubic at both ends) for
    // q_v_u, q_c_u, u_u, v_u, w_u, s, s_u, s_d, M_u, M_d (BC_Atm.h:567-574); pattern B for P_conv
    // (c43 x[1] - c13 x[2] at i=0 with the int constants 1 and 0, copy at the lid, :587-590,
    // :679-683); pattern C (cubic at i=0, copy at the lid) for c_u, e_d, e_l, e_p, g_p (:602-606,
    // :733-737), then q_v_u, q_c_u at i=0 floored at zero (:741-742).
    // bcTheta's copy (:769-775) and bcPhi's seam average (:810-818) for all of them except M_u,
    // M_d and P_conv, which neither list names.
    if (!s.mc.empty()) {
        const int cubic_both[10] = {ORC_MC_Q_V_U, ORC_MC_Q_C_U, ORC_MC_U_U, ORC_MC_V_U, ORC_MC_W_U,
                                    ORC_MC_S, ORC_MC_S_U, ORC_MC_S_D, ORC_MC_M_U, ORC_MC_M_D};
        const int cubic_copy[5] = {ORC_MC_C_U, ORC_MC_E_D_EVAP, ORC_MC_E_L, ORC_MC_E_P, ORC_MC_G_P};
        const int sides[13] = {ORC_MC_Q_V_U, ORC_MC_Q_C_U, ORC_MC_U_U, ORC_MC_V_U, ORC_MC_W_U, ORC_MC_S,
                               ORC_MC_S_U, ORC_MC_S_D, ORC_MC_C_U, ORC_MC_E_D_EVAP, ORC_MC_E_L, ORC_MC_E_P,
                               ORC_MC_G_P};
        const double c43 = 1.0, c13 = 0.0;
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) {
                auto X = [&](int q, int i) -> double& { return s.mc[q][s.ix(i, j, k)]; };
                for (int q : cubic_both) {
                    X(q, 0) = X(q, 3) - 3.0 * X(q, 2) + 3.0 * X(q, 1);
                    X(q, iml) = X(q, iml - 3) - 3.0 * X(q, iml - 2) + 3.0 * X(q, iml - 1);
                }
                X(ORC_MC_P_CONV, 0) = c43 * X(ORC_MC_P_CONV, 1) - c13 * X(ORC_MC_P_CONV, 2);
                X(ORC_MC_P_CONV, iml) = X(ORC_MC_P_CONV, iml - 1);
                for (int q : cubic_copy) {
                    X(q, 0) = X(q, 3) - 3.0 * X(q, 2) + 3.0 * X(q, 1);
                    X(q, iml) = X(q, iml - 1);
                }
                // updraft moisture at the surface non-negative (:741-742)
                if (X(ORC_MC_Q_V_U, 0) < 0.0) X(ORC_MC_Q_V_U, 0) = 0.0;
                if (X(ORC_MC_Q_C_U, 0) < 0.0) X(ORC_MC_Q_C_U, 0) = 0.0;
            }
        for (int q : sides) {
            double* x = s.mc[q].data();
            for (int k = 0; k < km; k++)
                for (int i = 0; i < im; i++) {
                    x[s.ix(i, 0, k)] = x[s.ix(i, 1, k)];
                    x[s.ix(i, jm - 1, k)] = x[s.ix(i, jm - 2, k)];
                }
            for (int i = 0; i < im; i++)
                for (int j = 0; j < jm; j++) {
                    const double v0 = c43 * x[s.ix(i, j, 1)] - c13 * x[s.ix(i, j, 2)];
                    const double ve = c43 * x[s.ix(i, j, km - 2)] - c13 * x[s.ix(i, j, km - 3)];
                    x[s.ix(i, j, 0)] = x[s.ix(i, j, km - 1)] = (v0 + ve) * 0.5;
                }
        }
    }
    (void)t;
}

// ---------------------------------------------------------------------------------------
// RK4 + RHS: RungeKutta_Atm_Turb.cpp:24-264, RHS_Atm_Turb.cpp:31-1080 (land-free central
// stencils, traditional Coriolis, SST with d_close = min(i,j,k), dead diagnostic blocks
// dropped: CoriolisForce/BuoyancyForce/PressureGradientForce/Q_Latent/Q_Sensible at
// :448-500 and the vbud/wbud captures).
// ---------------------------------------------------------------------------------------
void level_mean(S& s) {   // RungeKutta_Atm_Turb.cpp:24-41
    const double* t = s.F(ORC_T);
    for (int i = 0; i < s.im; i++) {
        double sum = 0.0, wsum = 0.0;
        for (int j = 0; j < s.jm; j++) {
            const double w = std::sin(s.the[j]);
            for (int k = 0; k < s.km; k++) { sum += w * t[s.ix(i, j, k)]; wsum += w; }
        }
        s.t_ref[i] = (wsum > 0.0) ? sum / wsum : 1.0;
    }
}

struct Geo {
    double exp_rm, exp_2_rm, sinthe, sinthe2, costhe, inv_rm, inv_rm2, inv_rmsinthe, inv_rm2sinthe,
           inv_rm2sinthe2, costhe_inv_rm2sinthe, inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
};

struct RhsConsts {
    double coeff_energy_p, force_nd, re_turb, pr_turb, diff_prec_re_inv, diff_co2_re_inv, nue_max, nue_air_nd,
           buoy_coeff, hs_scale, u_0, L_atm;
    int turb;   // the closure, orc_state::turb_model
};

enum { FT = 0, FU, FV, FW, FC, FCL, FI, FG, FCO, FK, FD, NF };

static inline double minmod(double a, double b) {
    return (a * b <= 0.0) ? 0.0 : (std::fabs(a) < std::fabs(b) ? a : b);
}

// One RHS evaluation at (i,j,k). cen[] are the cell's current stage values, nb[f] points at
// the array the ±1 neighbours are read from. Outputs rhs[11], aux[3], prod/sources.
void rhs_cell(const S& s, const Geo& g, const RhsConsts& K, int i, int j, int k,
              const double* cen, const double* const* nb, const double* pd, double nue_c,
              double p_stat_c, double p_surf, double t_eq_c, double t_ref_i, double buoy_ramp_dt,
              double surf_drag, double* rhs, double* aux, double& prod, double& tsrc, double& dsrc) {
    const size_t id = s.ix(i, j, k);
    const size_t si = (size_t)s.jm * s.km, sj = s.km;
    double d1r[NF], d2r[NF], d1t[NF], d2t[NF], d1p[NF], d2p[NF];
    for (int f = 0; f < NF; f++) {
        const double* x = nb[f];
        const double c0 = cen[f];
        const double xp = x[id + si], xm = x[id - si];
        d1r[f] = (xp - xm) * g.inv_2dr;
        d2r[f] = (xp - 2.0 * c0 + xm) * g.inv_dr2;
        const double yp = x[id + sj], ym = x[id - sj];
        d1t[f] = (yp - ym) * g.inv_2dthe;
        d2t[f] = (yp - 2.0 * c0 + ym) * g.inv_dthe2;
        const double zp = x[id + 1], zm = x[id - 1];
        d1p[f] = (zp - zm) * g.inv_2dphi;
        d2p[f] = (zp - 2.0 * c0 + zm) * g.inv_dphi2;
    }
    const double dpdr = (pd[id + si] - pd[id - si]) * g.inv_2dr;
    const double dpdthe = (pd[id + sj] - pd[id - sj]) * g.inv_2dthe;
    const double dpdphi = (pd[id + 1] - pd[id - 1]) * g.inv_2dphi;
    const double u = cen[FU], v = cen[FV], w = cen[FW];
    const double tke = cen[FK], dis = cen[FD];

    const double coriolis_the = 2.0 * g.costhe * w;
    const double coriolis_phi = -2.0 * (g.costhe * v);

    // prod refresh :528-561
    {
        const double dudr_s = d1r[FU] * g.exp_rm, dvdr_s = d1r[FV] * g.exp_rm, dwdr_s = d1r[FW] * g.exp_rm;
        const double dudthe_s = d1t[FU] * g.inv_rm, dvdthe_s = d1t[FV] * g.inv_rm, dwdthe_s = d1t[FW] * g.inv_rm;
        const double dudphi_s = d1p[FU] * g.inv_rmsinthe, dvdphi_s = d1p[FV] * g.inv_rmsinthe,
                     dwdphi_s = d1p[FW] * g.inv_rmsinthe;
        const double dis_here = dmax(dis, 1.0e-10);
        const double tke_here = dmax(tke, 0.0);
        double cnue = (K.turb == 0) ? kCnue * tke_here * tke_here / dis_here : tke_here / dis_here;
        cnue = dmin(cnue, K.nue_max);
        const double der = 0.66667 * (dudr_s + dvdthe_s + dwdphi_s);
        prod = dmax(0.0,
              (cnue * (2.0 * dudr_s - der) - 0.66667 * tke) * dudr_s
            + (cnue * (dudthe_s + dvdr_s)) * dudthe_s
            + (cnue * (dudphi_s + dwdr_s)) * dudphi_s
            + (cnue * (2.0 * dvdthe_s - der) - 0.66667 * tke) * dvdthe_s
            + (cnue * (dvdr_s + dudthe_s)) * dvdr_s
            + (cnue * (dvdphi_s + dwdthe_s)) * dvdphi_s
            + (cnue * (2.0 * dwdphi_s - der) - 0.66667 * tke) * dwdphi_s
            + (cnue * (dwdr_s + dudphi_s)) * dwdr_s
            + (cnue * (dwdthe_s + dvdphi_s)) * dwdthe_s);
    }
    double diffusion_t_re, diffusion_vel_re, diffusion_tke_re, diffusion_dis_re;
    if (K.turb == 0) {
        // k-epsilon :563-599: the production-destruction balance, no wall terms
        const double inv_re = 1.0 / K.re_turb;
        diffusion_t_re = (inv_re + nue_c) / K.pr_turb;
        diffusion_vel_re = inv_re + nue_c;
        diffusion_tke_re = inv_re + nue_c / 1.0;
        diffusion_dis_re = inv_re + nue_c / 1.3;
        const double tke_s = dmax(tke, 1.0e-10), dis_s = dmax(dis, 1.0e-10);
        const double P_k = prod;
        const double Y_k = dis_s;
        const double P_w = 1.35 * (dis_s / tke_s) * P_k;
        const double Y_w = 1.80 * dis_s * dis_s / tke_s;
        const double dsm = 20.0 * dis_s * dis_s, tsm = 20.0 * tke_s * dis_s;
        tsrc = dclamp(P_k - Y_k, -tsm, tsm);
        dsrc = dclamp(P_w - Y_w, -dsm, dsm);
    } else if (K.turb == 1) {
        // k-omega :601-631, with the cross-diffusion limiter sigma_d
        const double inv_re = 1.0 / K.re_turb;
        diffusion_t_re = (inv_re + nue_c) / K.pr_turb;
        diffusion_vel_re = inv_re + nue_c;
        diffusion_tke_re = inv_re + 0.6 * nue_c;
        diffusion_dis_re = inv_re + 0.5 * nue_c;
        const double bet_star = 0.09, gam = 0.52, bet_0 = 0.0708;
        const double tke_s = dmax(tke, 1.0e-10), dis_s = dmax(dis, 1.0e-10);
        const double P_k = dmin(prod, 20.0 * bet_star * tke_s * dis_s);
        const double Y_k = bet_star * tke_s * dis_s;
        const double P_w = gam * (dis_s / tke_s) * P_k;
        const double Y_w = bet_0 * dis_s * dis_s;
        const double cross = d1r[FK] * d1r[FD] + d1t[FK] * g.inv_rm * d1t[FD] * g.inv_rm
                           + d1p[FK] * g.inv_rmsinthe * d1p[FD] * g.inv_rmsinthe;
        const double sig_d = (cross <= 0.0) ? 0.0 : 0.125;
        const double D_w = sig_d * cross / dis_s;
        const double dsm = 20.0 * dis_s * dis_s, tsm = 20.0 * tke_s * dis_s;
        tsrc = dclamp(P_k - Y_k, -tsm, tsm);
        dsrc = dclamp(P_w - Y_w + D_w, -dsm, dsm);
    } else {
        // SST :633-721 with d_close = min(i, j, k) (the land scans evaluate to this without land)
        const double sig_k1 = 1.176, sig_k2 = 1.0, sig_w1 = 2.0, sig_w2 = 1.168, bet_star = 0.09;
        double d_close = (double)i;
        double dj_s = (double)j, dj_n = (double)j, dk_e = (double)k, dk_w = (double)k;
        if (s.sst_scan) {
            // RHS_Atm_Turb.cpp:640-671: scans of the land mask along j and k from the cell (the
            // same land test as AtomUtils::is_land); without land they leave the mount indices
            // at 0, so every distance is j or k
            auto land = [&](int ii, int jj, int kk) {
                return std::fabs(s.h[s.ix(ii, jj, kk)] - 1.0) < 2.220446049250313e-16;
            };
            const int jm_ = s.jm, km_ = s.km;
            double j_mount = 0.0, k_mount = 0.0;   // each carried from its first scan to its second
            for (int jj = 0; jj < jm_ - 1; jj++) if (land(i, jj, k) || jj == jm_ - 1) j_mount = jj;
            dj_s = std::fabs((double)j - j_mount);
            for (int jj = jm_ - 1; jj > 0; jj--) if (land(i, jj, k) || jj == 0) j_mount = jj;
            dj_n = std::fabs((double)j - j_mount);
            for (int kk = 0; kk < km_ - 1; kk++) if (land(i, j, kk) || kk == km_ - 1) k_mount = kk;
            dk_e = std::fabs((double)k - k_mount);
            for (int kk = km_ - 1; kk > 0; kk--) if (land(i, j, kk) || kk == 0) k_mount = kk;
            dk_w = std::fabs((double)k - k_mount);
        }
        if (dj_s <= d_close) d_close = dj_s;
        if (dj_n <= d_close) d_close = dj_n;
        if (dk_e <= d_close) d_close = dk_e;
        if (dk_w <= d_close) d_close = dk_w;
        const double cross = d1r[FK] * d1r[FD] + d1t[FK] * g.inv_rm * d1t[FD] * g.inv_rm
                           + d1p[FK] * g.inv_rmsinthe * d1p[FD] * g.inv_rmsinthe;
        const double CD_kw = dmax((2.0 * sig_w2) / dis * cross, 1.0e-20);
        const double arg1 = dmin(dmax(std::sqrt(tke) / (bet_star * dis * d_close),
                                      (500.0 * K.nue_air_nd) / (d_close * d_close * dis)),
                                 (4.0 * sig_w2 * tke) / (CD_kw * d_close * d_close));
        const double F1 = std::tanh(std::pow(arg1, 4));
        const double inv_re = 1.0 / K.re_turb;
        diffusion_t_re = (inv_re + nue_c) / K.pr_turb;
        diffusion_vel_re = inv_re + nue_c;
        diffusion_tke_re = inv_re + nue_c / (F1 * sig_k1 + (1.0 - F1) * sig_k2);
        diffusion_dis_re = inv_re + nue_c / (F1 * sig_w1 + (1.0 - F1) * sig_w2);
        const double bet1 = 0.0333, bet2 = 0.0368, gam1 = 0.413, gam2 = 0.2;
        const double tke_s = dmax(tke, 1.0e-10), dis_s = dmax(dis, 1.0e-10);
        const double P_k = dmin(prod, 20.0 * bet_star * tke_s * dis_s);
        const double Y_k = bet_star * tke_s * dis_s;
        const double P_w = (F1 * gam1 + (1.0 - F1) * gam2) * P_k * dis_s / tke_s;
        const double Y_w = (F1 * bet1 + (1.0 - F1) * bet2) * dis_s * dis_s;
        const double D_w = 2.0 * (1.0 - F1) * sig_w2 / dis_s * cross;
        const double dsm = 20.0 * dis_s * dis_s, tsm = 20.0 * tke_s * dis_s;
        tsrc = dclamp(P_k - Y_k, -tsm, tsm);
        dsrc = dclamp(P_w - Y_w + D_w, -dsm, dsm);
    }
    const double u_exp = u * g.exp_rm, v_invrm = v * g.inv_rm, w_invrs = w * g.inv_rmsinthe;
    const double pressure_t = K.coeff_energy_p * (u_exp * dpdr + v_invrm * dpdthe + w_invrs * dpdphi);
    const double inv_dr = 2.0 * g.inv_2dr, inv_dthe = 2.0 * g.inv_2dthe, inv_dphi = 2.0 * g.inv_2dphi;
    double tr[NF];
    for (int f = 0; f < NF; f++) {
        double gr_, gt_, gp_;
        if (f <= FI) {   // t u v w c cloud ice: minmod-limited advective gradients
            const double* x = nb[f];
            const double c0 = cen[f];
            gr_ = minmod(c0 - x[id - si], x[id + si] - c0) * inv_dr;
            gt_ = minmod(c0 - x[id - sj], x[id + sj] - c0) * inv_dthe;
            gp_ = minmod(c0 - x[id - 1], x[id + 1] - c0) * inv_dphi;
        } else { gr_ = d1r[f]; gt_ = d1t[f]; gp_ = d1p[f]; }
        tr[f] = u_exp * gr_ + v_invrm * gt_ + w_invrs * gp_;
    }
    const double two_over_rm_exp = 2.0 * g.inv_rm * g.exp_rm;
    const double cos_rm2sin = g.costhe_inv_rm2sinthe;
    const double v_metric = (1.0 + g.costhe * g.costhe / g.sinthe2) * g.inv_rm2;
    auto lap = [&](int f) {
        return d2r[f] * g.exp_2_rm + d1r[f] * two_over_rm_exp + d2t[f] * g.inv_rm2 + d1t[f] * cos_rm2sin
             + d2p[f] * g.inv_rm2sinthe2;
    };
    const double diffusion_t = lap(FT) * diffusion_t_re;
    const double diffusion_u = (d2r[FU] * g.exp_2_rm + 2.0 * u * g.inv_rm2 + d2t[FU] * g.inv_rm2
        + 4.0 * d1r[FU] * g.inv_rm * g.exp_rm + d1t[FU] * cos_rm2sin + d2p[FU] * g.inv_rm2sinthe2) * diffusion_vel_re;
    const double diffusion_v = (d2r[FV] * g.exp_2_rm + d1r[FV] * two_over_rm_exp + d2t[FV] * g.inv_rm2
        + d1t[FV] * cos_rm2sin - v_metric * v + d2p[FV] * g.inv_rm2sinthe2 + 2.0 * d1t[FU] * g.inv_rm2
        - d1p[FW] * 2.0 * g.costhe * g.inv_rm2sinthe2) * diffusion_vel_re;
    const double diffusion_w = (d2r[FW] * g.exp_2_rm + d1r[FW] * two_over_rm_exp + d2t[FW] * g.inv_rm2
        + d1t[FW] * cos_rm2sin - v_metric * w + d2p[FW] * g.inv_rm2sinthe2 + 2.0 * d1p[FU] * g.inv_rm2sinthe
        + d1p[FV] * 2.0 * g.costhe * g.inv_rm2sinthe2) * diffusion_vel_re;
    const double diffusion_c = lap(FC) * K.diff_prec_re_inv;
    const double diffusion_cl = lap(FCL) * K.diff_prec_re_inv;
    const double diffusion_i = lap(FI) * K.diff_prec_re_inv;
    const double diffusion_g = lap(FG) * K.diff_prec_re_inv;
    const double diffusion_co = lap(FCO) * K.diff_co2_re_inv;
    const double diffusion_k = lap(FK) * diffusion_tke_re;
    const double diffusion_d = lap(FD) * diffusion_dis_re;

    const double dpdr_exp = dpdr * g.exp_rm, dpdthe_invrm = dpdthe * g.inv_rm, dpdphi_invrs = dpdphi * g.inv_rmsinthe;
    double rt = pressure_t - tr[FT] + diffusion_t;
    // Moist-convection tendencies (RHS_Atm_Turb.cpp:370-372, 964, 1023, 1042, 1061): zero until the
    // scheme has run, so they are added only where its arrays exist.
    double mc_t = 0.0, mc_v = 0.0, mc_w = 0.0, mc_q = 0.0;
    const bool have_mc = !s.mc.empty();
    if (have_mc) {
        mc_t = (s.P.L_atm / (s.P.u_0 * s.P.t_0)) * s.mc[ORC_MC_MC_T][id];
        mc_v = (s.P.L_atm / (s.P.u_0 * s.P.u_0)) * s.mc[ORC_MC_MC_V][id];
        mc_w = (s.P.L_atm / (s.P.u_0 * s.P.u_0)) * s.mc[ORC_MC_MC_W][id];
        mc_q = (s.P.L_atm / (s.P.u_0 * s.P.c_0)) * s.mc[ORC_MC_MC_Q][id];
        rt = rt + mc_t;
    }
    // Microphysics sources (RHS_Atm_Turb.cpp:963-966, 1058-1070); zero until the ice scheme
    // has run, so they are added only where its arrays exist.
    double src_c = 0.0, src_cl = 0.0, src_i = 0.0;
    if (!s.mp.empty()) {
        const double rh = s.f[ORC_R_HUMID][id];
        const double Sv = s.mp[ORC_MP_S_V][id], Sc = s.mp[ORC_MP_S_C][id], Si = s.mp[ORC_MP_S_I][id];
        const double Sr = s.mp[ORC_MP_S_R][id], Ss = s.mp[ORC_MP_S_S][id];
        const double coeff_energy = 1.0 / (s.P.c_0 * 1349.0 * s.P.t_0);   // cp_l = 1349
        rt = rt + coeff_energy * (Sc + Sr) * 2520000.0 * rh + coeff_energy * (Si + Ss + 0.0) * 2830000.0 * rh;
        src_c = 1.0 * Sv * rh;
        src_cl = 1.0 * Sc * rh;
        src_i = 1.0 * Si * rh;
    }
    {   // Held-Suarez :976-992
        const double k_a = 1.0 / (4.0 * 86400.0), k_s = 1.0 / (0.4 * 86400.0), sigma_b = 0.7;
        const double sigma = (p_surf > 0.0) ? p_stat_c / p_surf : 0.0;
        double bl = (sigma - sigma_b) / (1.0 - sigma_b);
        if (bl < 0.0) bl = 0.0;
        const double cos4lat = g.sinthe2 * g.sinthe2;
        const double k_T = k_a + (k_s - k_a) * bl * cos4lat;
        rt -= (k_T * K.L_atm / K.u_0) * (cen[FT] - t_eq_c);
    }
    rhs[FT] = rt;
    rhs[FU] = -dpdr_exp - tr[FU] + diffusion_u + buoy_ramp_dt * (cen[FT] - t_ref_i);
    if (have_mc) {
        rhs[FV] = -dpdthe_invrm - tr[FV] + diffusion_v + K.force_nd * coriolis_the + mc_v - surf_drag * v;
        rhs[FW] = -dpdphi_invrs - tr[FW] + diffusion_w + K.force_nd * coriolis_phi + mc_w - surf_drag * w;
        rhs[FC] = -tr[FC] + diffusion_c + src_c + mc_q;
    } else {
        rhs[FV] = -dpdthe_invrm - tr[FV] + diffusion_v + K.force_nd * coriolis_the - surf_drag * v;
        rhs[FW] = -dpdphi_invrs - tr[FW] + diffusion_w + K.force_nd * coriolis_phi - surf_drag * w;
        rhs[FC] = -tr[FC] + diffusion_c + src_c;
    }
    rhs[FCL] = -tr[FCL] + diffusion_cl + src_cl;
    rhs[FI] = -tr[FI] + diffusion_i + src_i;
    rhs[FG] = -tr[FG] + diffusion_g;
    rhs[FCO] = -tr[FCO] + diffusion_co;
    rhs[FK] = -tr[FK] + diffusion_k + tsrc;
    rhs[FD] = -tr[FD] + diffusion_d + dsrc;
    aux[0] = rhs[FU] + dpdr_exp;
    aux[1] = rhs[FV] + dpdthe_invrm;
    aux[2] = rhs[FW] + dpdphi_invrs;
    (void)nue_c;
}

void rk4(S& s, int order) {
    const Params& P = s.P;
    const int im = s.im, jm = s.jm, km = s.km;
    level_mean(s);
    const double dt = P.dt_visc;
    const double buoy_ramp = std::min(1.0, (double)s.total_iter / 300.0);
    const double half_dt = 0.5 * dt, dt_sixth = dt / 6.0;
    const double tke_max_nd = 1000.0 / (P.u_0 * P.u_0);
    const int fid[NF] = {ORC_T, ORC_U, ORC_V, ORC_W, ORC_C, ORC_CLOUD, ORC_ICE, ORC_GR, ORC_CO2, ORC_TKE, ORC_DIS};
    const int nid[NF] = {ORC_TN, ORC_UN, ORC_VN, ORC_WN, ORC_CN, ORC_CLOUDN, ORC_ICEN, ORC_GRN, ORC_CO2N, ORC_TKEN, ORC_DISN};
    RhsConsts K;
    K.coeff_energy_p = P.u_0 * P.u_0 / (1349.0 * P.t_0);   // cp_l = 1349
    K.force_nd = P.omega * s.shell_len / P.u_0;
    K.re_turb = s.re_turb;
    K.pr_turb = P.pr_turb;
    K.diff_prec_re_inv = 1.0 / (P.sc_WaterVapour * s.re_turb);
    K.diff_co2_re_inv = 1.0 / (P.sc_CO2 * s.re_turb);
    K.nue_max = 1000.0 / (P.u_0 * P.L_atm);
    K.nue_air_nd = 1.5e-5 / (P.u_0 * P.L_atm);
    K.u_0 = P.u_0; K.L_atm = P.L_atm;
    K.turb = s.turb_model;
    const double buoy_ramp_dt = buoy_ramp * 1.0 * P.g * dt / P.u_0;
    std::vector<double> sint(jm), cost(jm);
    for (int j = 0; j < jm; j++) {
        sint[j] = std::sin(s.the[j]);
        if (sint[j] < 0.55) sint[j] = 0.55;
        cost[j] = std::cos(s.the[j]);
    }
    // Neighbour source: the live arrays (lex) or a snapshot taken now (jacobi).
    std::vector<std::vector<double>> snap;
    const double* nb[NF];
    if (order == ORC_ORDER_JACOBI || order == ORC_ORDER_JACOBI_REVERSE) {
        snap.resize(NF);
        for (int f = 0; f < NF; f++) { snap[f] = s.f[fid[f]]; nb[f] = snap[f].data(); }
    } else {
        for (int f = 0; f < NF; f++) nb[f] = s.f[fid[f]].data();
    }
    double* X[NF]; const double* XN[NF];
    for (int f = 0; f < NF; f++) { X[f] = s.F(fid[f]); XN[f] = s.F(nid[f]); }
    const double *pd = s.F(ORC_P_DYN), *nue = s.F(ORC_NUE), *ps = s.F(ORC_P_STAT), *teq = s.F(ORC_T_EQ);
    double *prod = s.F(ORC_PROD), *tks = s.F(ORC_TKE_SRC), *dss = s.F(ORC_DIS_SRC),
           *au = s.F(ORC_AUX_U), *av = s.F(ORC_AUX_V), *aw = s.F(ORC_AUX_W);

    // Per-(i, j) metric and per-i drag, shared by every order (the same values the reference
    // forms at the top of each column).
    std::vector<Geo> geo((size_t)im * jm);
    std::vector<double> drag(im);
    for (int i = 0; i < im; i++) {
        // Rayleigh drag :1014-1019 (i_topography = 0)
        double drag_profile = 1.0 - (double)(i - 0) / 5.0;
        if (drag_profile < 0.0) drag_profile = 0.0;
        if (drag_profile > 1.0) drag_profile = 1.0;
        drag[i] = ((1.0 / 86400.0) * P.L_atm / P.u_0 * dt) * drag_profile;
        for (int j = 0; j < jm; j++) {
            Geo& g = geo[(size_t)i * jm + j];
            const double rm = s.rad[i];
            g.exp_rm = 1.0 / (rm + 1.0);
            g.exp_2_rm = g.exp_rm * g.exp_rm;
            g.sinthe = sint[j];
            g.sinthe2 = g.sinthe * g.sinthe;
            g.costhe = cost[j];
            const double rmet = s.metric_radius(rm);
            const double rmet2 = rmet * rmet;
            g.inv_rm = 1.0 / rmet;
            g.inv_rm2 = 1.0 / rmet2;
            g.inv_rmsinthe = 1.0 / (rmet * g.sinthe);
            g.inv_rm2sinthe = g.inv_rm2 / g.sinthe;
            g.inv_rm2sinthe2 = g.inv_rm2 / g.sinthe2;
            g.costhe_inv_rm2sinthe = g.costhe * g.inv_rm2sinthe;
            g.inv_2dr = 1.0 / (2.0 * s.dr); g.inv_2dthe = 1.0 / (2.0 * s.dthe); g.inv_2dphi = 1.0 / (2.0 * s.dphi);
            g.inv_dr2 = 1.0 / (s.dr * s.dr); g.inv_dthe2 = 1.0 / (s.dthe * s.dthe); g.inv_dphi2 = 1.0 / (s.dphi * s.dphi);
        }
    }
    // In place (lex orders): every stage's value is written back before the next RHS, as the
    // reference's in-place sweep leaves it (RungeKutta_Atm_Turb.cpp:167-263).
    const bool in_place = (order == ORC_ORDER_LEX || order == ORC_ORDER_LEX_PLANES);

    auto do_cell = [&](int i, int j, int k) {
        const Geo& g = geo[(size_t)i * jm + j];
        const double surf_drag = drag[i];
        const size_t id = s.ix(i, j, k);
        double base[NF], cen[NF], kk[4][NF], rhs[NF], aux[3];
        for (int f = 0; f < NF; f++) { base[f] = XN[f][id]; cen[f] = X[f][id]; }
        double pr = 0, ts = 0, ds = 0;
        const double p_surf = ps[s.ix(0, j, k)];
        for (int st = 0; st < 4; st++) {
            rhs_cell(s, g, K, i, j, k, cen, nb, pd, nue[id], ps[id], p_surf, teq[id], s.t_ref[i],
                     buoy_ramp_dt, surf_drag, rhs, aux, pr, ts, ds);
            for (int f = 0; f < NF; f++) kk[st][f] = rhs[f];
            const double h = (st < 2) ? half_dt : dt;
            for (int f = 0; f < NF; f++) {
                double val;
                if (st < 3) val = base[f] + kk[st][f] * h;
                else val = base[f] + (kk[0][f] + 2.0 * kk[1][f] + 2.0 * kk[2][f] + kk[3][f]) * dt_sixth;
                if (f == FC || f == FCL || f == FI || f == FG) val = dmax(0.0, val);
                else if (f == FK) {
                    if (!finite_bits(val)) val = 0.0;
                    else val = (val < 0.0) ? 0.0 : (val > tke_max_nd) ? tke_max_nd : val;
                } else if (f == FD) val = dmax(1.0e-10, val);
                cen[f] = val;
            }
            if (in_place)
                for (int f = 0; f < NF; f++) X[f][id] = cen[f];
        }
        for (int f = 0; f < NF; f++) X[f][id] = cen[f];
        prod[id] = pr; tks[id] = ts; dss[id] = ds;
        au[id] = aux[0]; av[id] = aux[1]; aw[id] = aux[2];
    };
    const int ni = im - 2, nj = jm - 2;
    if (order == ORC_ORDER_LEX) {
        // RungeKutta_Atm_Turb.cpp:99-264 at OMP_NUM_THREADS=1: i, then j, then k ascending.
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++)
                for (int k = 1; k < km - 1; k++) do_cell(i, j, k);
    } else if (order == ORC_ORDER_LEX_PLANES) {
        // The same sweep by hyperplanes i + j + k = d. A cell reads its six neighbours only,
        // which lie on planes d - 1 (already final in the sequential sweep) and d + 1 (not yet
        // reached), and writes only itself, so the cells of one plane are independent and the
        // planes in increasing d reproduce the sequential sweep bit for bit.
        const int dmax_ = (im - 2) + (jm - 2) + (km - 2);
        #pragma omp parallel
        for (int d = 3; d <= dmax_; d++) {
            #pragma omp for schedule(static)
            for (int q = 0; q < ni * nj; q++) {
                const int i = 1 + q / nj, j = 1 + q % nj, k = d - i - j;
                if (k >= 1 && k <= km - 2) do_cell(i, j, k);
            }
        }
    } else if (order == ORC_ORDER_JACOBI_REVERSE) {
        // Jacobi order swept backwards in i, j and k: identical results unless some read sees
        // a value written in this sweep.
        #pragma omp parallel for collapse(2) schedule(static)
        for (int i = im - 2; i >= 1; i--)
            for (int j = jm - 2; j >= 1; j--)
                for (int k = km - 2; k >= 1; k--) do_cell(i, j, k);
    } else {
        #pragma omp parallel for collapse(2) schedule(static)
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++)
                for (int k = 1; k < km - 1; k++) do_cell(i, j, k);
    }
}

// AtomUtils::polar_zonal_filter, lib/Utils.cpp:542-608 (start 30 deg, 12 passes max, gain 3)
void polar_filter(S& s, double* x) {
    const int im = s.im, jm = s.jm, km = s.km;
    const double deg2rad = M_PI / 180.0;
    const double sin_ref = std::cos(30.0 * deg2rad);
    std::vector<int> passes(jm, 0);
    int gmax = 0;
    for (int j = 0; j < jm; j++) {
        const double sj = dmax(std::sin(s.the[j]), 1.0e-6);
        if (sj >= sin_ref) continue;
        const double ratio = sin_ref / sj;
        int p = (int)std::lround(3.0 * (ratio * ratio - 1.0));
        if (p < 0) p = 0;
        if (p > 12) p = 12;
        passes[j] = p;
        if (p > gmax) gmax = p;
    }
    if (gmax == 0) return;
    std::vector<double> row(km);
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++) {
            double* r = x + s.ix(i, j, 0);
            for (int pass = 0; pass < passes[j]; pass++) {
                std::copy(r, r + km, row.begin());
                for (int k = 0; k < km; k++) {
                    const int km1 = (k - 1 + km) % km, kp1 = (k + 1) % km;
                    r[k] = row[k] + 0.25 * (row[km1] - 2.0 * row[k] + row[kp1]);
                }
            }
        }
}

// AtomUtils::radial_shapiro_filter (:831-870) and radial_shapiro_filter_ho (:884-933)
void radial_filter(S& s, double* x, bool high_order) {
    const int im = s.im;
    const size_t si = (size_t)s.jm * s.km;
    std::vector<double> col(im);
    for (size_t jk = 0; jk < si; jk++) {
        for (int pass = 0; pass < 2; pass++) {
            for (int i = 0; i < im; i++) col[i] = x[i * si + jk];
            for (int i = 1; i < im - 1; i++) {
                const double c0 = col[i];
                if (high_order && i - 2 >= 0 && i + 2 <= im - 1) {
                    const double d4 = col[i + 2] - 4.0 * col[i + 1] + 6.0 * c0 - 4.0 * col[i - 1] + col[i - 2];
                    x[i * si + jk] = c0 - (1.0 / 16.0) * d4;
                } else {
                    x[i * si + jk] = c0 + 0.25 * (col[i - 1] - 2.0 * c0 + col[i + 1]);
                }
            }
        }
    }
}

// UtilsAtm::storeIntermediateData3D, UtilsAtm.h:25-49
void store(S& s) {
    const int src[11] = {ORC_U, ORC_V, ORC_W, ORC_T, ORC_C, ORC_CLOUD, ORC_ICE, ORC_GR, ORC_CO2, ORC_TKE, ORC_DIS};
    const int dst[11] = {ORC_UN, ORC_VN, ORC_WN, ORC_TN, ORC_CN, ORC_CLOUDN, ORC_ICEN, ORC_GRN, ORC_CO2N, ORC_TKEN, ORC_DISN};
    for (int f = 0; f < 11; f++) s.f[dst[f]] = s.f[src[f]];
    // the precipitation fluxes too (UtilsAtm.h:41-42), where the ice scheme has run
    if (!s.mp.empty()) {
        s.mp[ORC_MP_P_RAINN] = s.mp[ORC_MP_P_RAIN];
        s.mp[ORC_MP_P_SNOWN] = s.mp[ORC_MP_P_SNOW];
    }
}

void ice_scheme(S& s);
void moist_convection(S& s);

void moist_block(S& s) {   // cAtmosphereModel.cpp:1185-1434 (scoped rows)
    satadj(s);
    damp_wiggles(s, s.F(ORC_T));
    damp_wiggles(s, s.F(ORC_ICE));
    damp_wiggles(s, s.F(ORC_C));
    damp_wiggles(s, s.F(ORC_CLOUD));
    if (s.ice_on) {
        // :1230-1244 with CategoryIceScheme = 2, then precipitationSum (:1263, UtilsAtm.h:52-67)
        // with P_graupel and the moist convection's P_conv zero
        ice_scheme(s);
        if (s.conv_on) {
            // :1247-1262 MoistConvection::run, then the damp_wiggles of eleven of its arrays
            moist_convection(s);
            for (int q : {ORC_MC_P_CONV, ORC_MC_E_U, ORC_MC_M_U, ORC_MC_Q_V_U, ORC_MC_Q_C_U, ORC_MC_E_D, ORC_MC_M_D,
                          ORC_MC_Q_V_D, ORC_MC_MC_T, ORC_MC_MC_V, ORC_MC_MC_W})
                damp_wiggles(s, s.mc[q].data());
        }
        const double *pr = s.mp[ORC_MP_P_RAIN].data(), *ps = s.mp[ORC_MP_P_SNOW].data();
        const double* pc = s.mc.empty() ? nullptr : s.mc[ORC_MC_P_CONV].data();
        double* pt = s.mp[ORC_MP_PRECIP].data();
        for (size_t q = 0; q < s.n; q++) pt[q] = pr[q] + ps[q] + 0.0 + (pc ? pc[q] : 0.0);
    }
    moist_clamp(s);
    densities(s, false);
    evaporation(s);
    turbulence(s);
    for (int fld : {ORC_TKE, ORC_DIS, ORC_NUE, ORC_PROD, ORC_TKE_SRC, ORC_DIS_SRC}) damp_wiggles(s, s.F(fld));
}

void step(S& s, int order) {   // cAtmosphereModel.cpp:1103-1647 (scoped rows)
    s.iter_n++;
    s.total_iter++;
    if (s.iter_n % 4 == 0) poisson(s);
    if (s.iter_n % 2 == 0) moist_block(s);
    bc_all(s);
    rk4(s, order);
    for (int fld : {ORC_U, ORC_V, ORC_W}) polar_filter(s, s.F(fld));
    radial_filter(s, s.F(ORC_U), false);
    radial_filter(s, s.F(ORC_V), true);
    radial_filter(s, s.F(ORC_W), true);
    store(s);
}

// ---------------------------------------------------------------------------------------
// Initial state (documented recipe; follows RunTimeSlice's thermodynamic init chain,
// cAtmosphereModel.cpp:527-568, with the velocity / turbulence fields of the benchmark)
// ---------------------------------------------------------------------------------------
uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void apply_member(Params& P, const orc_member* m) {
    if (m) {
        P.t_surf_equator = m->t_surf_equator; P.t_surf_pole = m->t_surf_pole; P.r_air = m->r_air;
        P.x_H2O = m->x_H2O; P.x_CO2 = m->x_CO2; P.x_N2 = m->x_N2;
        P.c_0 = m->c_0; P.co2_0 = m->co2_0; P.c_h2o_dry_top = m->c_h2o_dry_top;
    }
    P.comp = resolve(P.x_H2O, P.x_CO2, P.x_N2);
}

void init_state(S& s, uint64_t seed, int stop_after_thermo) {
    const Params& P = s.P;
    const int im = s.im, jm = s.jm, km = s.km;
    double *t = s.F(ORC_T), *ps = s.F(ORC_P_STAT), *rh = s.F(ORC_R_HUMID), *rd = s.F(ORC_R_DRY),
           *c = s.F(ORC_C), *co2 = s.F(ORC_CO2);
    // resetArrays: t = 1 (UtilsAtm.h:435-439)
    std::fill(t, t + s.n, 1.0);
    // initTemperatureData, InitValues_Atm.cpp:493-632 (no topography; c = co2 = cloud = ice = 0)
    const double d_j_half = 0.5 * (jm - 1);
    for (int k = 0; k < km; k++)
        for (int j = 0; j < jm; j++) {
            const double ratio = (double)j / d_j_half;
            const double T_s = (P.t_surf_pole - P.t_surf_equator) * (ratio * ratio - 2.0 * ratio) + P.t_surf_pole;
            t[s.ix(0, j, k)] = T_s * (1.0 / P.t_0);
        }
    for (int k = 0; k < km; k++)
        for (int j = 0; j < jm; j++) {
            const double tu = t[s.ix(0, j, k)] * P.t_0;
            ps[s.ix(0, j, k)] = 1e-2 * (P.r_air * P.comp.R_mix * tu);
            t[s.ix(0, j, k)] = tu;
            double T_prev = dmax(100.0, t[s.ix(0, j, k)]);
            double p_prev = ps[s.ix(0, j, k)];
            for (int i = 0; i < im; i++) {
                const size_t id = s.ix(i, j, k);
                const double R_loc = R_of(c[id], co2[id], P.comp.R_bg);
                double Tc, pv;
                if (i == 0) { Tc = T_prev; pv = p_prev; }
                else {
                    const double dz = (double)(s.hgt[i] - s.hgt[i - 1]);
                    const double cp_loc = cp_of(c[id], co2[id], T_prev, P.comp.M_bg);
                    const double T_ad = T_prev - (P.g / cp_loc) * dz;
                    Tc = dmax(P.t_skin, T_ad);
                    const double Tm = 0.5 * (T_prev + Tc);
                    pv = p_prev * std::exp(-P.g * dz / (R_loc * Tm));
                }
                t[id] = Tc; ps[id] = pv;
                const double wf = dmax(0.5, 1.0 - (0.0 + 0.0));
                const double R_d = R_of(0.0, co2[id], P.comp.R_bg);
                rd[id] = pv * 100.0 / (R_d * Tc);
                rh[id] = pv * 100.0 / (R_loc * Tc * wf);
                T_prev = Tc; p_prev = pv;
            }
        }
    const double inv_t0 = 1.0 / P.t_0;
    for (size_t id = 0; id < s.n; id++) t[id] *= inv_t0;
    // RunTimeSlice :528 damp_wiggles(t)
    damp_wiggles(s, t);
    // :533-536 lid snapshot
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++) s.f2[ORC2_T_TOP_INIT][(size_t)j * km + k] = t[s.ix(im - 1, j, k)];
    // initWaterWapour, InitValues_Atm.cpp:843-905
    {
        const double M_nw = M_nonwater(P.c_0, P.co2_0, P.comp.M_bg);
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) {
                double q_sat_min = 1.0;
                for (int i = 0; i < im; i++) {
                    const size_t id = s.ix(i, j, k);
                    const double q_s = q_sat_at(t[id] * P.t_0, ps[id], M_nw);
                    if (q_s < q_sat_min) q_sat_min = q_s;
                    c[id] = dmax(q_sat_min, P.c_h2o_dry_top);
                }
            }
    }
    // :565-568 two passes of co2Atmosphere (ThermoAtm.h:873-913) + densities(true)
    for (int pass = 0; pass < 2; pass++) {
        const double f_CO2 = (P.comp.q_CO2 + P.comp.q_bg > 0.0) ? P.comp.q_CO2 / (P.comp.q_CO2 + P.comp.q_bg)
                                                               : P.co2_0 * P.co2_scale;
        for (size_t id = 0; id < s.n; id++) co2[id] = (1.0 - c[id]) * f_CO2 * P.co2_scale;
        densities(s, true);
    }
    if (stop_after_thermo) return;
    // condensate pinned to zero; velocities zero plus a seeded, N-S mirrored perturbation of
    // +-1e-3 m/s (u, w symmetric, v antisymmetric and zero on the equator row)
    const double amp = 1.0e-3 / P.u_0;
    const int jeq = (jm - 1) / 2;
    const int vel[3] = {ORC_U, ORC_V, ORC_W};
    for (int f = 0; f < 3; f++) {
        double* x = s.F(vel[f]);
        for (int i = 0; i < im; i++)
            for (int j = 0; j < jm; j++)
                for (int k = 0; k < km; k++) {
                    const int jj = (j <= jeq) ? j : (jm - 1 - j);
                    const uint64_t ctr = ((((uint64_t)f * im + i) * jm + jj) * km + k);
                    const uint64_t r = splitmix64(seed * 0xD1B54A32D192ED03ULL + ctr);
                    double val = (2.0 * ((double)(r >> 11) * (1.0 / 9007199254740992.0)) - 1.0) * amp;
                    if (f == 1) {
                        if (j == jeq) val = 0.0;
                        else if (j > jeq) val = -val;
                    }
                    x[s.ix(i, j, k)] = val;
                }
    }
    // turbulence: k = 1e-4 m2/s2, omega = 1 1/s everywhere (k* = k/u0^2, w* = w L/u0)
    std::fill(s.F(ORC_TKE), s.F(ORC_TKE) + s.n, 1.0e-4 / (P.u_0 * P.u_0));
    std::fill(s.F(ORC_DIS), s.F(ORC_DIS) + s.n, 1.0 * P.L_atm / P.u_0);
    // nue, prod and sources from one TurbulenceAtm::run and the damp_wiggles of its six
    // fields (RunTimeSlice, cAtmosphereModel.cpp:612-622)
    turbulence(s);
    for (int fld : {ORC_TKE, ORC_DIS, ORC_NUE, ORC_PROD, ORC_TKE_SRC, ORC_DIS_SRC}) damp_wiggles(s, s.F(fld));
    bc_all(s);
    // RunTimeSlice :665 projection of the initial velocity (200 sweeps)
    project_initial_velocity(s, 200);
    // run_3D_loop :750-753 Held-Suarez target; RunTimeSlice :667 store
    s.f[ORC_T_EQ] = s.f[ORC_T];
    store(s);
    s.iter_n = 0;
    s.total_iter = 0;
}

// ---------------------------------------------------------------------------------------
// Radiation in run_3D_loop's mode 2 (cAtmosphereModel.cpp:799-801, 863-879, 1655-1661):
// MultiLayerRadiation::run (MultiLayerRadiation.h:59-384) on the current t, every t blended
// by omega_rad toward that radiative result, then one step of the t_skin fixed point
// (updateSkinTemperature + planetaryShortWave, cAtmosphereModel.cpp:241-311).
// ---------------------------------------------------------------------------------------
// Short wave at the top per latitude: a parabola from pole to equator, mirrored (MLR.h:98-104).
void short_wave(S& s) {
    const Params& P = s.P;
    const int jmax = s.jm - 1, jhalf = jmax / 2;
    s.sw.assign(s.jm, P.rad_pole_short);
    const double span = P.rad_pole_short - P.rad_equator_short;
    for (int j = jhalf; j >= 0; j--) {
        const double x = (double)j / (double)jhalf;
        s.sw[j] = span * (x * x - 2.0 * x) + P.rad_pole_short;   // AtomUtils::parabola
    }
    for (int j = jmax; j > jhalf; j--) s.sw[j] = s.sw[jmax - j];
}

// Grey two-stream sweeps over one column (MLR.h:291-303, 323-333): up from the black surface,
// dn from a dark sky; level i's emission eps_i sigma T_i^4.
void sweeps(int im, const double* eps, const double* T, double sigma, double* up, double* dn) {
    up[0] = sigma * std::pow(T[0], 4.0);
    for (int i = 1; i < im; i++) up[i] = up[i - 1] * (1.0 - eps[i]) + eps[i] * sigma * std::pow(T[i], 4.0);
    dn[im] = 0.0;
    for (int i = im - 1; i >= 1; i--) dn[i] = dn[i + 1] * (1.0 - eps[i]) + eps[i] * sigma * std::pow(T[i], 4.0);
}

void radiation(S& s) {
    const Params& P = s.P;
    const int im = s.im, jm = s.jm, km = s.km;
    short_wave(s);
    s.radiation.assign(s.n, 0.0);
    s.epsilon.assign(s.n, 0.0);
    s.albedo.assign((size_t)jm * km, 0.0);
    double *t = s.F(ORC_T), *c = s.F(ORC_C), *co2 = s.F(ORC_CO2), *p = s.F(ORC_P_STAT);
    double *cloud = s.F(ORC_CLOUD), *ice = s.F(ORC_ICE);
    const double inv_g = 1.0 / P.g;
    const int top = im - 1;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < jm; j++) {
        std::vector<double> T(im), eps(im), up(im), dn(im + 1);
        for (int k = 0; k < km; k++) {
            for (int i = 0; i < im; i++) T[i] = t[s.ix(i, j, k)] * P.t_0;
            // layer optical depth: gas column mass with pressure broadening plus the cloud
            // liquid and ice paths (MLR.h:156-207)
            double lwp = 0.0, iwp = 0.0;
            for (int i = 0; i < im; i++) {
                const size_t q = s.ix(i, j, k);
                double dp = (i < top) ? (p[q] - p[s.ix(i + 1, j, k)]) * 100.0 : p[q] * 100.0;
                if (dp < 0.0) dp = 0.0;
                const double qv = dmax(0.0, c[q]), qc = dmax(0.0, co2[q]);
                const double qb = dmax(0.0, 1.0 - qv - qc);
                const double mass = dp * inv_g;
                const double broad = p[q] * 100.0 / 1.0e5;
                const double tau_gas = (P.kappa_H2O * qv + P.kappa_CO2 * qc + P.kappa_bg * qb) * mass * broad;
                const float h0 = s.hgt[(i < top) ? i : i - 1], h1 = s.hgt[(i < top) ? i + 1 : i];
                const double dz = (double)(h1 - h0);   // float heights, float difference
                const double Ti = t[q] * P.t_0;
                const double rho = (Ti > 0.0) ? (p[q] * 100.0) / (R_of(qv, qc, P.comp.R_bg) * Ti) : 0.0;
                const double lw = dmax(0.0, cloud[q]) * rho * dz * 1000.0;
                const double iw = dmax(0.0, ice[q]) * rho * dz * 1000.0;
                lwp += lw;
                iwp += iw;
                const double tau = tau_gas + 0.12 * lw + 0.055 * iw;
                eps[i] = (tau > 700.0) ? 1.0 : (1.0 - std::exp(-tau));
                s.epsilon[q] = eps[i];
            }
            // clear-sky surface albedo raised toward the cloud albedo by the capped
            // condensate path (MLR.h:90-96, 231-238)
            double alb = P.albedo_surface;
            if (P.albedo_cloud > alb) {
                const double tau = (lwp + 0.5 * iwp) / 100.0;
                alb = alb + (P.albedo_cloud - alb) * (tau / (tau + 2.0));
            }
            s.albedo[(size_t)j * km + k] = alb;
            // four Lambda iterations of layer radiative equilibrium sigma T^4 = (up + dn) / 2
            // (MLR.h:287-320), then the sweeps once more on the final profile
            for (int it = 0; it < 4; it++) {
                sweeps(im, eps.data(), T.data(), P.sigma, up.data(), dn.data());
                for (int i = 1; i <= top; i++) {
                    const double emit = 0.5 * (up[i - 1] + dn[i + 1]);
                    if (emit > 0.0 && finite_bits(emit)) T[i] = std::pow(emit / P.sigma, 0.25);
                }
            }
            sweeps(im, eps.data(), T.data(), P.sigma, up.data(), dn.data());
            // surface energy balance, sigma T^4 linearised about the current surface
            // temperature, with the bulk flux to the lowest layer (MLR.h:338-371)
            const double Ts0 = T[0], c_H = 15.0;
            const double d4 = 4.0 * P.sigma * Ts0 * Ts0 * Ts0;
            const double Ts = ((1.0 - alb) * s.sw[j] + P.geothermal_flux + dn[1] - P.sigma * std::pow(Ts0, 4.0) +
                               d4 * Ts0 + c_H * T[1]) / (d4 + c_H);
            T[0] = Ts;
            s.radiation[s.ix(0, j, k)] = P.sigma * std::pow(Ts, 4.0);
            for (int i = 1; i <= top; i++) s.radiation[s.ix(i, j, k)] = up[i];
            // apply_radiative_heating: t <- (1 - omega_rad) t + omega_rad T_rad / t_0
            for (int i = 0; i < im; i++) {
                const size_t q = s.ix(i, j, k);
                t[q] = t[q] * (1.0 - P.omega_rad) + (T[i] / P.t_0) * P.omega_rad;
            }
        }
    }
    // t_skin relaxes toward the emission temperature of the absorbed short wave plus the
    // geothermal flux, with the cos(latitude)-weighted planetary means of the model's albedo
    if (P.t_skin_relax <= 0.0) return;
    double wsum = 0.0, sw_abs = 0.0, amax = 0.0;
    for (int j = 0; j < jm; j++) {
        const double w = std::cos((j / (double)(jm - 1) - 0.5) * M_PI);
        double a = 0.0;
        for (int k = 0; k < km; k++) a += s.albedo[(size_t)j * km + k];
        a /= (double)km;
        amax = dmax(amax, a);
        wsum += w;
        sw_abs += w * (1.0 - a) * s.sw[j];
    }
    if (!(wsum > 0.0) || !(amax > 0.0)) return;
    const double in = sw_abs / wsum + P.geothermal_flux;
    if (!(in > 0.0)) return;
    const double target = std::pow(in / P.sigma, 0.25);
    s.P.t_skin += dmin(1.0, P.t_skin_relax) * (target - P.t_skin);
}

// ---------------------------------------------------------------------------------------
// Two-category ice scheme: TwoCatIceScheme::run, TwoCatIceScheme.h:89-111 (initArrays
// :135-155, computeColumns :160-517, applyBoundaryConditions :522-525 with
// IceSchemeCommon::extrapolateBC IceSchemeCommon.h:28-55, applyTopography a no-op without
// land), then the two damp_wiggles of run_3D_loop (cAtmosphereModel.cpp:1243-1244).
// ---------------------------------------------------------------------------------------
void ensure_mp(S& s) {
    if ((int)s.mp.size() == ORC_NMP) return;
    s.mp.assign(ORC_NMP, std::vector<double>(s.n, 0.0));
}

// IceSchemeCommon::evaporateWhereImpossible (IceSchemeCommon.h:121-190) with the scheme's
// precipitation fluxes; graupel and its flux are zero in the two-category scheme.
void ice_evaporate(S& s, size_t id, double t_u) {
    const Params& P = s.P;
    double* t = s.F(ORC_T); double* c = s.F(ORC_C); double* cl = s.F(ORC_CLOUD);
    double* ice = s.F(ORC_ICE); double* gr = s.F(ORC_GR); const double* co2 = s.F(ORC_CO2);
    const double* rh = s.F(ORC_R_HUMID);
    const double q_cond = dmax(0.0, cl[id]) + dmax(0.0, ice[id]) + dmax(0.0, gr[id]);
    if (q_cond > 0.0) {
        const double c_max = dmax(0.0, 1.0 - co2[id]);
        const double c_new = dmin(c_max, c[id] + q_cond);
        const double q_used = dmax(0.0, c_new - c[id]);
        c[id] = c_new; cl[id] = 0.0; ice[id] = 0.0; gr[id] = 0.0;
        if (t_u < kTcrit) {
            const double cp_loc = cp_of(c[id], co2[id], t_u, P.comp.M_bg);
            if (cp_loc > 0.0) {
                double T_new = t_u - L_vap(t_u) * q_used / cp_loc;
                if (T_new < 1.0) T_new = 1.0;
                t[id] = T_new / P.t_0;
            }
        }
    }
    double* Pr = s.mp[ORC_MP_P_RAIN].data(); double* Ps = s.mp[ORC_MP_P_SNOW].data();
    {
        const double rho = dmax(1.0e-6, rh[id]);
        const double q_p = dmax(0.0, Pr[id]) / (1.6 * rho) + dmax(0.0, Ps[id]) / (0.96 * rho) + 0.0 / (2.0 * rho);
        if (q_p > 0.0) {
            const double c_max = dmax(0.0, 1.0 - co2[id]);
            c[id] = dmin(c_max, c[id] + q_p);
        }
    }
    Pr[id] = 0.0; Ps[id] = 0.0; s.mp[ORC_MP_PRECIP][id] = 0.0;
    for (int q : {ORC_MP_S_V, ORC_MP_S_C, ORC_MP_S_I, ORC_MP_S_R, ORC_MP_S_S}) s.mp[q][id] = 0.0;
    s.F(ORC_S_C_C)[id] = 0.0;
}

// One (j,k) column of computeColumns (TwoCatIceScheme.h:164-516). The rate variables are
// column-private and carried from level to level and from pass to pass exactly where the
// reference leaves them unassigned: N_i and m_i outside the mixed-phase band, S_nuc where
// ice is zero and no nucleation condition holds, S_i_dep where c equals q_Ice.
void ice_column(S& s, int j, int k) {
    const Params& P = s.P;
    constexpr double N_i_0 = 1.0e2, m_i_0 = 1.0e-12, m_i_max = 1.0e-9, m_s_0 = 3.0e-9;
    constexpr double c_i_dep = 1.3e-5, c_c_au = 1.0e-3, q_c_crit = 5.0e-4, c_i_au = 1.0e-3;
    constexpr double c_ac = 0.24, c_rim = 18.6, c_rim_snow = 18.6 / 5.0, c_agg = 10.3;
    constexpr double c_i_cri = 0.24, c_r_cri = 3.2e-5, a_ev = 1.0e-3, b_ev = 5.9;
    constexpr double c_s_dep = 1.8e-2, b_s_dep = 12.3, c_s_melt = 8.43e-5, b_s_melt = 12.05;
    constexpr double c_r_frz = 3.75e-2;
    constexpr double t_nuc = 267.15, t_d = 248.15, t_hn = 236.15, t_r_frz = 271.15;
    constexpr int iter_prec_end = 3, i_check = 1;
    constexpr double P_max = 50.0 / 86400.0;
    const double t_0 = P.t_0;
    const double* t = s.F(ORC_T); const double* c = s.F(ORC_C); const double* cl = s.F(ORC_CLOUD);
    const double* ice = s.F(ORC_ICE); const double* co2 = s.F(ORC_CO2);
    const double* ps = s.F(ORC_P_STAT); const double* rh = s.F(ORC_R_HUMID);
    const double* scc = s.F(ORC_S_C_C);
    double* Pr = s.mp[ORC_MP_P_RAIN].data(); double* Ps = s.mp[ORC_MP_P_SNOW].data();
    double* Pt = s.mp[ORC_MP_PRECIP].data();
    double* Sv = s.mp[ORC_MP_S_V].data(); double* Sc = s.mp[ORC_MP_S_C].data();
    double* Si = s.mp[ORC_MP_S_I].data(); double* Sr = s.mp[ORC_MP_S_R].data();
    double* Ss = s.mp[ORC_MP_S_S].data();
    const int im = s.im;

    const double Rain_check = Pr[s.ix(i_check, j, k)];
    double m_i = m_i_max, N_i = 0.0, S_nuc = 0.0, S_i_dep = 0.0;
    for (int iter_prec = 1; iter_prec <= iter_prec_end; iter_prec++) {
        Pr[s.ix(im - 1, j, k)] = 0.0;
        Ps[s.ix(im - 1, j, k)] = 0.0;
        for (int i = im - 2; i >= 0; i--) {
            const size_t id = s.ix(i, j, k);
            {   // IceSchemeCommon::canCondense (IceSchemeCommon.h:104-107)
                const double t_guard = t[id] * t_0;
                bool can = false;
                if (!(t_guard >= kTcrit))
                    can = q_sat(p_sat(t_guard), ps[id], M_nonwater(c[id], co2[id], P.comp.M_bg)) < 1.0;
                if (!can) { ice_evaporate(s, id, t_guard); continue; }
            }
            const double Rain = Pr[id], Snow = Ps[id];
            double R49 = 0.0, R79 = 0.0, R139 = 0.0;
            if (Rain > 1e-12) {
                const double rb = std::pow(Rain, 1.0 / 9.0);
                const double r2 = rb * rb;
                const double r4 = r2 * r2;
                R49 = r4; R79 = Rain / r2; R139 = Rain * r4;
            }
            const double t_u = t[id] * t_0;
            const double Mo = M_nonwater(c[id], co2[id], P.comp.M_bg);
            const double q_sw = q_sat(p_sat(t_u), ps[id], Mo);
            const double q_Ice = q_sat(p_subl(t_u), ps[id], Mo);
            const double step = (double)(float)(s.hgt[i + 1] - s.hgt[i]);
            const double mass_layer = rh[id] * step;
            const double dt_snow_dim = step / 0.96;
            const double cv = c[id], cld = cl[id], ic = ice[id];

            if (!(t_u > t_0) && !(t_u <= t_hn)) {
                N_i = N_i_0 * std::exp(0.2 * (t_0 - t_u));
                m_i = dmin((rh[id] * ic) / N_i, m_i_max);
                if (m_i > m_i_max) m_i = m_i_max;
                if (m_i < m_i_0) m_i = m_i_0;
            }
            if (ic == 0.0) {
                if (((t_u < t_d) && (cv >= q_Ice)) || (((t_d <= t_u) && (t_u <= t_nuc)) && (cv >= q_sw)))
                    S_nuc = m_i_0 / (rh[id] * dt_snow_dim) * N_i;
            } else S_nuc = 0.0;
            double S_c_frz = (t_u < t_hn) ? cld / dt_snow_dim : 0.0;
            if (cv > q_Ice) S_i_dep = c_i_dep * N_i * std::pow(m_i, 1.0 / 3.0) * (cv - q_Ice);
            if (cv < q_Ice) S_i_dep = dmax((-ic / dt_snow_dim), ((cv - q_Ice) / dt_snow_dim));
            double S_c_au = (cld > q_c_crit) ? c_c_au * (cld - q_c_crit) : 0.0;
            double S_i_au = (ic > 0.0) ? dmax(c_i_au * ic, 0.0) : 0.0;
            double S_d_au = (S_i_dep > 0.0) ? S_i_dep / (1.5 * (std::pow((m_s_0 / m_i), 2.0 / 3.0) - 1.0)) : 0.0;
            double S_ac = ((t_u >= t_0) && (cld > q_c_crit)) ? c_ac * (cld - q_c_crit) * R79 : 0.0;
            double S_rim = (t_u < t_0) ? c_rim_snow * cld * Snow : 0.0;
            double S_shed = (t_u >= t_0) ? c_rim * cld * Snow : 0.0;
            double S_s_melt = 0.0, S_i_melt = 0.0;
            if (t_u > t_0) {
                if (Snow > 1e-12) {
                    S_s_melt = c_s_melt * (1.0 + b_s_melt * std::pow(Snow, 5.0 / 26.0)) * (t_u - t_0)
                             * std::pow(Snow, 1.0 / 3.0);
                    S_s_melt = dmin(S_s_melt, Snow / mass_layer);
                }
                S_i_melt = ic / dt_snow_dim;
            }
            double S_agg = 0.0, S_i_cri = 0.0, S_r_cri = 0.0;
            if (t_u <= t_0) {
                S_agg = c_agg * ic * Snow;
                S_i_cri = c_i_cri * ic * R79;
                S_r_cri = c_r_cri * ic / m_i * R139;
            }
            double S_ev = 0.0;
            if (t_u > t_0 && Rain > 1e-12 && cv < q_sw) {
                S_ev = a_ev * (1.0 + b_ev * std::pow(Rain, 1.0 / 6.0)) * (q_sw - cv) * R49;
                S_ev = dmin(S_ev, Rain / mass_layer);
            }
            double S_s_dep = 0.0;
            if (t_u < t_0 && Snow > 1e-12) {
                S_s_dep = c_s_dep * (1.0 + b_s_dep * std::pow(Snow, 5.0 / 26.0)) * (cv - q_Ice)
                        * std::pow(Snow, 8.0 / 13.0);
                if (S_s_dep < 0.0) S_s_dep = dmax(S_s_dep, -Snow / mass_layer);
            }
            double S_r_frz = 0.0;
            if (t_u <= t_0) {
                const double t_frz = dmax((t_r_frz - t_u), 0.0);
                S_r_frz = c_r_frz * std::pow(t_frz, 3.0 / 2.0) * std::pow(Rain, 3.0 / 2.0);
            }
            const double S_cloud_total = S_c_au + S_ac + S_rim + S_c_frz + S_shed;
            const double max_cloud_loss = cld / dt_snow_dim;
            if (S_cloud_total > max_cloud_loss && S_cloud_total > 0) {
                const double f = max_cloud_loss / S_cloud_total;
                S_c_au *= f; S_ac *= f; S_rim *= f; S_c_frz *= f; S_shed *= f;
            }
            const double S_ice_total = S_i_au + S_d_au + S_agg + S_i_cri + S_i_melt;
            const double max_ice_loss = ic / dt_snow_dim;
            if (S_ice_total > max_ice_loss && S_ice_total > 0) {
                const double f = max_ice_loss / S_ice_total;
                S_i_au *= f; S_d_au *= f; S_agg *= f; S_i_cri *= f; S_i_melt *= f;
            }
            Sv[id] = -scc[id] + S_ev - S_i_dep - S_s_dep - S_nuc;
            Sc[id] = scc[id] - S_c_au - S_ac - S_c_frz + S_i_melt - S_rim - S_shed;
            Si[id] = S_nuc + S_c_frz + S_i_dep - S_i_melt - S_i_au - S_d_au - S_agg - S_i_cri;
            Sr[id] = S_c_au + S_ac - S_ev + S_shed - S_r_cri - S_r_frz + S_s_melt;
            Ss[id] = S_i_au + S_d_au + S_agg + S_rim + S_s_dep + S_i_cri + S_r_cri + S_r_frz - S_s_melt;
            const double dP_rain = (S_c_au + S_ac + S_shed + S_s_melt - S_ev - S_r_frz) * mass_layer;
            Pr[id] = dmin(P_max, dmax(0.0, Pr[s.ix(i + 1, j, k)] + dP_rain));
            const double dP_snow = (S_i_au + S_d_au + S_rim + S_agg + S_r_frz - S_s_melt) * mass_layer;
            Ps[id] = (t_u < t_0 && t_u >= P.t_000) ? dmin(P_max, dmax(0.0, Ps[s.ix(i + 1, j, k)] + dP_snow)) : 0.0;
            Pt[id] = dmin(P_max, Pr[id] + Ps[id]);
        }
        if (std::fabs(Pr[s.ix(i_check, j, k)] - Rain_check) * 8.64e4 <= 1.0e-3) break;
    }
}

// IceSchemeCommon::extrapolateBC (IceSchemeCommon.h:28-55) with c43 = 1, c13 = 0 (ints,
// cAtmosphereModel.h:158) and the phi-seam average.
void extrapolate_bc(S& s, double* P) {
    const double c43 = 1.0, c13 = 0.0;
    const int im = s.im, jm = s.jm, km = s.km;
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++)
            P[s.ix(im - 1, j, k)] = c43 * P[s.ix(im - 2, j, k)] - c13 * P[s.ix(im - 3, j, k)];
    for (int k = 0; k < km; k++)
        for (int i = 0; i < im; i++) {
            P[s.ix(i, 0, k)] = c43 * P[s.ix(i, 1, k)] - c13 * P[s.ix(i, 2, k)];
            P[s.ix(i, jm - 1, k)] = c43 * P[s.ix(i, jm - 2, k)] - c13 * P[s.ix(i, jm - 3, k)];
        }
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++) {
            double& a = P[s.ix(i, j, 0)];
            double& b = P[s.ix(i, j, km - 1)];
            a = c43 * P[s.ix(i, j, 1)] - c13 * P[s.ix(i, j, 2)];
            b = c43 * P[s.ix(i, j, km - 2)] - c13 * P[s.ix(i, j, km - 3)];
            a = b = (a + b) * 0.5;
        }
}

void ice_scheme(S& s) {
    ensure_mp(s);
    double* c = s.F(ORC_C); double* cl = s.F(ORC_CLOUD); double* ice = s.F(ORC_ICE);
    double* Pr = s.mp[ORC_MP_P_RAIN].data(); double* Ps = s.mp[ORC_MP_P_SNOW].data();
    double* Prn = s.mp[ORC_MP_P_RAINN].data(); double* Psn = s.mp[ORC_MP_P_SNOWN].data();
    for (size_t q = 0; q < s.n; q++) {   // initArrays, TwoCatIceScheme.h:139-155
        c[q] = dmax(0.0, c[q]); cl[q] = dmax(0.0, cl[q]); ice[q] = dmax(0.0, ice[q]);
        Prn[q] = dmax(0.0, Prn[q]); Psn[q] = dmax(0.0, Psn[q]);
        Pr[q] = Prn[q]; Ps[q] = Psn[q];
    }
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k < s.km - 1; k++)
        for (int j = 1; j < s.jm - 1; j++) ice_column(s, j, k);
    extrapolate_bc(s, Pr);
    extrapolate_bc(s, Ps);
    damp_wiggles(s, Pr);
    damp_wiggles(s, Ps);
}

// The stored copies of the precipitation fluxes, as storeIntermediateData3D keeps them
// (UtilsAtm.h:41-42).
void ice_store(S& s) {
    ensure_mp(s);
    s.mp[ORC_MP_P_RAINN] = s.mp[ORC_MP_P_RAIN];
    s.mp[ORC_MP_P_SNOWN] = s.mp[ORC_MP_P_SNOW];
}

// ---------------------------------------------------------------------------------------
// Moist convection: MoistConvection::run (MoistConvection.h:106-140) with the model's
// defaults iter_prec = 3, convection_perturbation = 1, convection_mode = 1 and no land.
// Every phase reads only its own column's convection arrays and the fields c, cloud, t, u, v,
// w, p_stat, r_humid, co2 (c also at the four horizontal neighbours, which no phase writes),
// so the phases of one column run back to back here; the result is the reference's
// phase-by-phase sweep over all columns. computeCAPE (:1259-1299), a 1-D diagnostic that
// feeds nothing, runs after all columns, in its own column order (k outer, j inner).
// Q_sensible_2D is zero throughout the reference (nothing fills it).
// ---------------------------------------------------------------------------------------
namespace conv {
constexpr double a_ev = 1.0e-3, b_ev = 5.9, t_00 = 236.15, t_00_m2 = 236.15 - 2.0;
constexpr double b_u = 0.33, alf_1 = 0.05, alf_2 = 0.05, bet_p = 2.0e-3;
const double eps_u = 0.2 / 100.0, del_u = 0.2 / 100.0;           // :30-32 (R_cloud 100 m)
constexpr double eps_d = 1.0e-4, del_d = 1.0e-4, gam_d = -0.3;
constexpr double a_u = 3.0e-2, a_d = 3.0e-2, vel = 1.1;
constexpr double f_beg = 1000.0 / 1013.25, f_end = 970.0 / 1013.25, f_cb = 900.0 / 1013.25;
constexpr double f_diff = 200.0 / 1013.25, f_mid = 700.0 / 1013.25;
constexpr double t_add_u = 0.2, q_v_u_add = 1.0e-4, coeff_recurr = 0.1, scale = 0.98;
constexpr double q_c_u_max = 1.0e-2, w_conv_max = 50.0, M_max = 3.0, c_mb = 0.003;
constexpr double alpha_sfp = 0.2;                                  // :404
constexpr double MCt_max = 0.01, MCq_max = 2.0e-5, MCv_max = 0.01;   // :1323-1339
constexpr double cp_l = 1349.0, s_0 = 274515.75, p_0 = 60000.0, lv = 2.52e6, ls = 2.83e6;

inline double cap(double v, double mag) {                          // :216-222, :1341-1346
    if (!finite_bits(v)) return 0.0;
    return (v < -mag) ? -mag : (v > mag) ? mag : v;
}
}  // namespace conv

void ensure_mc(S& s) {
    if (s.mc.size() == ORC_NMC) return;
    s.mc.assign(ORC_NMC, std::vector<double>(s.n, 0.0));
    s.evap.assign((size_t)s.jm * s.km, 0.0);
}

void moist_convection(S& s) {
    using namespace conv;
    ensure_mc(s);
    const Params& P = s.P;
    const int im = s.im, jm = s.jm, km = s.km;
    const size_t pl = (size_t)jm * km;
    const double t0 = P.t_0;
    const double *T = s.F(ORC_T), *U = s.F(ORC_U), *V = s.F(ORC_V), *W = s.F(ORC_W);
    const double *C = s.F(ORC_C), *CL = s.F(ORC_CLOUD), *CO2 = s.F(ORC_CO2);
    const double *PS = s.F(ORC_P_STAT), *RH = s.F(ORC_R_HUMID);
    auto A = [&](int q) { return s.mc[q].data(); };
    double *Sx = A(ORC_MC_S), *Mu = A(ORC_MC_M_U), *Md = A(ORC_MC_M_D), *Su = A(ORC_MC_S_U);
    double *Qvu = A(ORC_MC_Q_V_U), *Qcu = A(ORC_MC_Q_C_U), *Cu = A(ORC_MC_C_U), *El = A(ORC_MC_E_L);
    double *Edv = A(ORC_MC_E_D_EVAP), *Ep = A(ORC_MC_E_P), *Gp = A(ORC_MC_G_P), *Sd = A(ORC_MC_S_D);
    double *Qvd = A(ORC_MC_Q_V_D), *Pc = A(ORC_MC_P_CONV), *Eu = A(ORC_MC_E_U), *Du = A(ORC_MC_D_U);
    double *Uu = A(ORC_MC_U_U), *Vu = A(ORC_MC_V_U), *Wu = A(ORC_MC_W_U), *Ed = A(ORC_MC_E_D);
    double *Dd = A(ORC_MC_D_D), *Ud = A(ORC_MC_U_D), *Vd = A(ORC_MC_V_D), *Wd = A(ORC_MC_W_D);
    double *MCt = A(ORC_MC_MC_T), *MCq = A(ORC_MC_MC_Q), *MCv = A(ORC_MC_MC_V), *MCw = A(ORC_MC_MC_W);
    double *Dbeg = A(ORC_MC_DEEP_BEG), *Dend = A(ORC_MC_DEEP_END), *CB = A(ORC_MC_CLOUD_BASE);
    double *LFS = A(ORC_MC_LFS);
    // layer heights (float in the reference, cAtmosphereModel.h:241-246) and level steps (:300-309)
    std::vector<double> h(im), step(im, 0.0);
    for (int i = 0; i < im; i++) h[i] = (double)s.hgt[i];
    for (int i = 0; i + 1 < im; i++) step[i] = h[i + 1] - h[i];
    if (im >= 2) step[im - 1] = step[im - 2];
    const double Mb = P.comp.M_bg, Rb = P.comp.R_bg;
    auto qs_env = [&](double E, double p, double qv) { return q_sat(E, p, M_nonwater(qv, P.co2_0, Mb)); };  // :181-184
    auto cc_factor = [&](double TK, double p) {                                                          // :191-203
        auto Es = [&](double Tx) { return (Tx >= t0) ? p_sat(Tx) : p_subl(Tx); };
        const double a = qs_env(Es(TK), p, P.c_0), b = qs_env(Es(288.15), p, P.c_0);
        return (b > 0.0) ? a / b : 1.0;
    };
    const double inv_a_u = 1.0 / (a_u * P.u_0), inv_a_d = 1.0 / (a_d * P.u_0);
    const double uu_cap = w_conv_max / P.u_0;
    std::vector<double> qs(im);
    // each column's cloud base and level of free sinking after the passes, for computeCAPE
    std::vector<int> base_of(pl, 0), lfs_of(pl, 0);

    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++) {
            auto id = [&](int i) { return (size_t)i * pl + (size_t)j * km + k; };
            const bool inner = j >= 1 && j <= jm - 2 && k >= 1 && k <= km - 2;
            const double p_sfc = PS[id(0)];
            // entrainment gradients of c (:238-297, no land): centred differences
            auto grad = [&](int i, double& dr, double& dth, double& dph, double& cc) {
                // about the equator row (jm - 1) / 2: row 90 of the reference's 181-row grid
                // (MoistConvection.h:239), kept at the equator on every grid
                const double costhe = std::cos(s.the[(jm - 1) / 2] - s.the[j]);
                const double rth = (P.r_Earth * 1.0e3 + h[i]) * costhe;
                const double Ue = rth * 2.0 * M_PI;
                const int ip = std::min(i + 1, im - 1), imn = std::max(i - 1, 0);
                dr = (C[id(ip)] - C[id(imn)]) * (0.5 / step[i]);
                dth = (C[id(i) + km] - C[id(i) - km]) * (0.5 / (Ue / (jm - 1)));
                dph = (C[id(i) + 1] - C[id(i) - 1]) * (1.0 / (Ue / (km - 1)));
                cc = dmax(C[id(i)], 1e-4);
            };
            // precompute (:336-376): s from t, the per-call arrays cleared
            for (int i = 0; i < im; i++) {
                const size_t q = id(i);
                Sx[q] = cp_l * T[q] * t0 / s_0;
                Mu[q] = Md[q] = Su[q] = Qvu[q] = Qcu[q] = Cu[q] = El[q] = Edv[q] = Ep[q] = Gp[q] = 0.0;
                Sd[q] = Qvd[q] = Pc[q] = 0.0;
            }
            int base = 0, lfs = 0, beg = -1, end = -1;
            double MdLFS = 0.0, cape_c = 0.0, dT = 0.0, dq = 0.0;
            // findCloudBaseLFS (:627-777)
            for (int i = 0; i < im; i++)
                if (PS[id(i)] <= f_beg * p_sfc) { beg = i; Dbeg[id(i)] = h[i]; break; }
            for (int i = im - 1; i >= 0; i--)
                if (PS[id(i)] >= f_end * p_sfc) { end = i; Dend[id(i)] = h[i]; break; }
            if (beg >= 0 && end >= 0) {
                for (int i = beg; i < im; i++)
                    qs[i] = qs_env(p_auto(T[id(i)] * t0 + t_add_u), PS[id(i)], C[id(i)]);
                for (int i = beg; i < im; i++) {
                    const size_t q = id(i);
                    const double t_u = T[q] * t0;
                    const double w = 1.0 / (1.0 + std::exp(-(t_u - t_00) / 2.0));
                    Qvu[q] = C[q] + q_v_u_add;
                    if (Qvu[q] >= scale * qs[i] && t_u >= t_00) Qcu[q] = Qvu[q] - scale * qs[i];
                    if (Qvu[q] <= 0.0) Qvu[q] = 0.0;
                    if (Qcu[q] <= 0.0) Qcu[q] = 0.0;
                    if (Qvu[q] >= scale * qs[i] && w > 0.01 && PS[q] <= f_cb * p_sfc) {
                        base = i;
                        CB[q] = h[i] * w;
                        break;
                    }
                }
                if (base > 0) {   // level of neutral buoyancy of the theta-e parcel from cloud base
                    const size_t qb = id(base);
                    const double Tb = T[qb] * t0 + t_add_u;
                    const double kap = R_of(C[qb], CO2[qb], Rb) / cp_of(C[qb], CO2[qb], Tb, Mb);
                    const double Lb = (Tb >= t0) ? lv : ls;
                    const double thp = Tb * std::pow(p_0 / PS[qb], kap) * std::exp(Lb * qs[base] / (cp_l * Tb));
                    bool buoy = false;
                    int top = base;
                    double cape = 0.0;
                    for (int i = base + 1; i < im; i++) {
                        const double Te = T[id(i)] * t0;
                        const double Le = (Te >= t0) ? lv : ls;
                        const double the = Te * std::pow(p_0 / PS[id(i)], kap) * std::exp(Le * qs[i] / (cp_l * Te));
                        if (thp - the > 0.0) {
                            buoy = true;
                            top = i;
                            cape += P.g * (thp - the) / the * (h[i] - h[i - 1]);
                        } else if (buoy) {
                            break;
                        }
                    }
                    const int lnb = std::min(std::max(top, base + 1), im - 2);
                    lfs = lnb;
                    cape_c = dmax(0.0, cape);
                    LFS[id(lnb)] = h[lnb];
                }
            }
            const bool deep = PS[id(base)] - PS[id(lfs)] >= f_diff * p_sfc;
            // surfaceFluxPerturbations, switch 1 (:402-485)
            if (inner && base >= 0 && lfs > base) {
                const double Ts = T[id(0)] * t0;
                const double rho = dmax(RH[id(0)], 1e-6);
                const double qcap = q_v_u_add * cc_factor(Ts, p_sfc);
                if (PS[id(base)] < f_mid * p_sfc) {
                    dT = t_add_u;
                    dq = qcap;
                } else if (!deep) {
                    const double Hs = 0.0;
                    const double E = dmax(0.0, s.evap[(size_t)j * km + k]) / 86400.0;
                    const double z = h[base];
                    if (!(E <= 0.0 && Hs <= 0.0) && z > 0.0) {
                        const double B = (P.g / Ts) * (Hs / (rho * cp_l) + 0.61 * Ts * E / rho);
                        const double ws = std::cbrt(dmax(0.0, B * z));
                        if (ws > 0.0) {
                            const double ir = 1.0 / (rho * ws);
                            dT = dmin(alpha_sfp * Hs * ir / cp_l, t_add_u);
                            dq = dmin(alpha_sfp * E * ir, qcap);
                        }
                    }
                }
            }
            auto entrain = [&](int i, bool with_m) {   // E_u at level i (:534-543, :841-851)
                const size_t q = id(i);
                double e = 0.0;
                if (PS[id(base)] - PS[id(lfs)] >= f_diff * p_sfc) {
                    double dr, dth, dph, cc;
                    grad(i, dr, dth, dph, cc);
                    e = -RH[q] / cc * (U[q] * dr + V[q] * dth + W[q] * dph);
                    if (with_m) e += eps_u * Mu[q];
                }
                if (T[q] * t0 <= t_00_m2) e = 0.0;
                return e;
            };
            auto seed_M = [&](int i) { return cap(RH[id(i)] * c_mb * std::sqrt(2.0 * cape_c), M_max); };
            // initUpdraft (:490-622)
            if (inner && beg != -1 && end != -1) {
                Mu[id(beg)] = seed_M(beg);
                for (int i = beg + 1; i <= end; i++) {
                    const size_t q = id(i);
                    const double t_u = T[q] * t0, t_a = t_u + dT;
                    const double qsa = qs_env(p_auto(t_a), PS[q], C[q]);
                    Qvu[q] = C[q] + dq;
                    if (Qvu[q] >= scale * qsa && t_u >= t_00) Qcu[q] = Qvu[q] - scale * qsa;
                    if (Qvu[q] <= 0.0) Qvu[q] = 0.0;
                    if (Qcu[q] <= 0.0) Qcu[q] = 0.0;
                    Eu[q] = entrain(i, false);
                    Du[q] = 0.0;
                    Mu[q] = cap(Mu[id(i - 1)] + (Eu[q] - Du[q]) * step[i], M_max);
                    Sx[q] = cp_l * t_u / s_0;
                    Su[q] = cp_l * t_a / s_0;
                    Uu[q] = vel * U[q]; Vu[q] = vel * V[q]; Wu[q] = vel * W[q];
                }
                MdLFS = cap(gam_d * Mu[id(beg)], M_max);
                Md[id(lfs)] = MdLFS;
                for (int i = end + 1; i <= lfs - 1; i++) {
                    const size_t q = id(i);
                    const double t_u = T[q] * t0, t_a = t_u + dT;
                    const double qsa = qs_env(p_auto(t_a), PS[q], C[q]);
                    if (i == end + 1) Mu[id(i - 1)] = Mu[id(beg)];
                    Sx[q] = cp_l * t_u / s_0;
                    Su[q] = cp_l * t_u / s_0;
                    Qvu[q] = C[q] + dq;
                    if (Qvu[q] >= scale * qsa && t_a >= t_00) Qcu[q] = Qvu[q] - scale * qsa;
                    Eu[q] = entrain(i, false);
                    Du[q] = 0.0;
                    Mu[q] = cap(Mu[id(i - 1)] + (Eu[q] - Du[q]) * step[i], M_max);
                    Uu[q] = dmax(-uu_cap, dmin(U[q] + Mu[q] * inv_a_u / dmax(RH[q], 1e-6), uu_cap));
                    Vu[q] = V[q]; Wu[q] = W[q];
                }
            }
            // CloudBaseUpdate (:1462-1482) with simpsonCxEu (:1426-1455)
            auto cloud_base_update = [&]() {
                if (!inner || beg < 0 || lfs <= beg) return;
                int hit = -1;
                double cum = 0.0;
                int i = beg;
                auto f = [&](int l) { return -(C[id(l)] * Eu[id(l)]); };
                for (; i + 2 <= lfs && hit < 0; i += 2) {
                    const double h1 = step[i], h2 = step[i + 1];
                    const double f0 = f(i), f1 = f(i + 1), f2 = f(i + 2);
                    const double pc = (h1 + h2) / 6.0 * ((2.0 - h2 / h1) * f0 + (h1 + h2) * (h1 + h2) / (h1 * h2) * f1
                                                         + (2.0 - h1 / h2) * f2);
                    if (cum + pc > 0.0) { hit = (cum + 0.5 * h1 * (f0 + f1) > 0.0) ? i + 1 : i + 2; break; }
                    cum += pc;
                }
                if (hit < 0 && i + 1 <= lfs && cum + 0.5 * step[i] * (f(i) + f(i + 1)) > 0.0) hit = i + 1;
                if (hit >= 0 && hit < im && PS[id(hit)] <= f_cb * p_sfc && (base <= 0 || hit <= base)) {
                    base = hit;
                    CB[id(hit)] = h[hit];
                }
            };
            cloud_base_update();
            for (int ip = 1; ip <= 3; ip++) {
                // updraftEntrainment (:780-908)
                if (inner && base >= 0 && lfs > base) {
                    const bool shallow = PS[id(base)] - PS[id(lfs)] < f_diff * p_sfc;   // convection_mode 1
                    const double dic = shallow ? 1.0e9 : 500.0;
                    const double hb = h[base];
                    for (int i = 0; i < base; i++) Mu[id(i)] = 0.0;
                    Mu[id(base)] = seed_M(base);
                    Su[id(base)] = cp_l * (T[id(base)] * t0 + dT) / s_0;
                    Qvu[id(base)] = dmax(C[id(base)] + dq, 0.0);
                    MdLFS = cap(gam_d * Mu[id(base)], M_max);
                    for (int i = base + 1; i <= lfs; i++) {
                        const size_t q = id(i);
                        const double t_u = T[q] * t0;
                        Eu[q] = entrain(i, true);
                        double d = eps_u * Mu[q];
                        if (ip > 1 && i == lfs) d = del_u * Mu[q] + b_u * Mu[q] / step[i];
                        if (ip > 1 && i == lfs - 1) d = del_u * Mu[q] + (1.0 - b_u) * Mu[q] / step[i];
                        if (t_u <= t_00_m2) d = 0.0;
                        Du[q] = d;
                        Mu[q] = cap(Mu[id(i - 1)] + (Eu[q] - Du[q]) * step[i], M_max);
                        Cu[q] = 0.0;
                        if (ip == 1) {
                            El[q] = 0.0;
                        } else {
                            El[q] = Du[q] * Qcu[q] / dmax(RH[q], 0.01);
                            if (El[q] <= 0.0) El[q] = 0.0;
                        }
                        const double Kp = (h[i] > hb + dic) ? bet_p : 0.0;
                        Gp[q] = (t_u >= t0 && ip != 1) ? Kp * Qcu[q] : 0.0;
                    }
                    Mu[id(lfs)] = 0.0;
                }
                cloud_base_update();
                // updraftRecurrence (:1032-1163)
                if (inner && base >= 0 && lfs > base) {
                    for (int i = base + 1; i <= lfs - 1; i++) {
                        const size_t q = id(i), qm = id(i - 1);
                        const double t_u = T[qm] * t0;
                        const double sp = step[i - 1];
                        Mu[q] = cap(Mu[qm] + (Eu[qm] - Du[qm]) * sp, M_max);
                        if (t_u <= t_00_m2) Mu[q] = 0.0;
                        const double Mp = Mu[qm];
                        const double n_qv = Mp * Qvu[qm] + sp * (Eu[qm] * C[qm] - Du[qm] * Qvu[qm]);
                        const double n_qc = Mp * Qcu[qm] + sp * (-Du[qm] * CL[qm] - RH[qm] * Gp[qm]);
                        const double n_v = Mp * Vu[qm] + sp * Eu[qm] * V[qm];
                        const double n_w = Mp * Wu[qm] + sp * Eu[qm] * W[qm];
                        const double n_s = (Mp * Su[qm] + sp * (Eu[qm] * Sx[qm] - Du[qm] * Su[qm])) / s_0;
                        if (std::fabs(Mu[q]) > coeff_recurr) {
                            const double r = 1.0 / Mu[q];
                            Qvu[q] = n_qv * r; Qcu[q] = n_qc * r; Vu[q] = n_v * r; Wu[q] = n_w * r; Su[q] = n_s * r;
                        } else {
                            Su[q] = Sx[q]; Qvu[q] = C[q]; Qcu[q] = 0.0; Vu[q] = V[q]; Wu[q] = W[q];
                        }
                        Cu[q] = 0.0;
                        if (std::fabs(Mu[q]) > coeff_recurr) {   // parcel saturation adjustment, two passes
                            double Tu = Su[q] * s_0 / cp_l;
                            const double pu = PS[q];
                            double tot = 0.0;
                            for (int it = 0; it < 2; ++it) {
                                const double ex = Qvu[q] - qs_env(p_auto(Tu), pu, C[q]);
                                if (ex <= 0.0) break;
                                const double L = (Tu >= t0) ? lv : ls;
                                const double G = (L / cp_l) * dqdT_from(p_auto(Tu), L_auto(Tu), Tu, pu,
                                                                        M_nonwater(C[q], CO2[q], Mb));
                                const double dc = ex / (1.0 + G);
                                Qvu[q] -= dc;
                                Qcu[q] += dc;
                                Tu += (L / cp_l) * dc;
                                tot += dc;
                            }
                            Su[q] = cp_l * Tu / s_0;
                            Cu[q] = tot * Mu[q] / (dmax(RH[q], 0.01) * step[i]);
                        }
                        Uu[q] = dmax(-uu_cap, dmin(U[q] + Mu[q] * inv_a_u / dmax(RH[q], 1e-6), uu_cap));
                        if (Qvu[q] <= 0.0) Qvu[q] = 0.0;
                        if (Qcu[q] <= 0.0) Qcu[q] = 0.0;
                        if (Qcu[q] > q_c_u_max) Qcu[q] = q_c_u_max;
                        if (Su[q] <= 0.0) Su[q] = 0.0;
                        if (t_u <= t_00) {
                            Eu[q] = Du[q] = Mu[q] = Qvu[q] = Qcu[q] = Uu[q] = Vu[q] = Wu[q] = 0.0;
                            Su[q] = 1.0;
                            Gp[q] = Cu[q] = El[q] = 0.0;
                        }
                    }
                }
                // downdraftEntrainment (:913-1027)
                if (lfs > 0 && lfs > base) {
                    const size_t ql = id(lfs);
                    const double t_l = T[ql] * t0, p_l = PS[ql];
                    Md[ql] = cap(MdLFS, M_max);
                    Ed[ql] = eps_d * std::fabs(Md[ql]);
                    Dd[ql] = del_d * std::fabs(Md[ql]);
                    Qvd[ql] = 0.5 * (CL[ql] + scale * qs_env(p_auto(t_l), p_l, C[ql]));
                    const double r_l = 1e2 * p_l / (R_of(C[ql], CO2[ql], Rb) * t_l);
                    Ud[ql] = U[ql] + Md[ql] * inv_a_d / dmax(r_l, 1e-6);
                    Vd[ql] = V[ql]; Wd[ql] = W[ql]; Sd[ql] = Sx[ql];
                    Pc[ql] = 0.0;
                    double sig = 0.0;
                    const double wcb = std::fabs(Wu[id(base)]) * P.u_0;
                    if (wcb > 0.0) sig = std::fabs(Mu[id(base)]) / (dmax(RH[id(base)], 1e-6) * wcb);
                    sig = dmin(sig, 1.0);
                    for (int i = lfs - 1; i >= 0; i--) {
                        const size_t q = id(i);
                        const double t_u = T[q] * t0;
                        Sd[q] = cp_l * t_u / s_0;
                        const double qsd = qs_env(p_auto(t_u), PS[q], C[q]);
                        Qvd[q] = 0.5 * (CL[q] + scale * qsd);
                        Ed[q] = eps_d * std::fabs(Md[q]);
                        Dd[q] = del_d * std::fabs(Md[q]);
                        Md[q] = cap(Md[id(i + 1)] - (Ed[q] - Dd[q]) * step[i], M_max);
                        Ud[q] = U[q] + Md[q] * inv_a_d / dmax(RH[q], 1e-6);
                        Vd[q] = V[q]; Wd[q] = W[q];
                        if (i <= base) {
                            if (ip == 1) {
                                Ep[q] = 0.0;
                            } else {
                                const double pr = Pc[q];
                                if (pr > 0.0 && sig > 0.0)
                                    Ep[q] = sig * alf_1 * (qsd - Qvd[q]) * std::sqrt(1e1 / alf_2 * pr / sig);
                                if (Ep[q] <= 0.0) Ep[q] = 0.0;
                            }
                        }
                        if (ip == 1) {
                            Edv[q] = 0.0;
                        } else {
                            const double pr = Pc[q];
                            Edv[q] = (pr > 0.0) ? a_ev * (1.0 + b_ev * std::pow(pr, 1.0 / 6.0)) * (qsd - Qvd[q])
                                                      * std::pow(pr, 4.0 / 9.0) : 0.0;
                            if (Edv[q] <= 0.0) Edv[q] = 0.0;
                        }
                    }
                }
                // downdraftRecurrence (:1168-1254), every column
                for (int i = lfs - 1; i >= 0; i--) {
                    const size_t q = id(i), qp = id(i + 1);
                    const double t_u = T[q] * t0;
                    const double sp = step[i + 1], rp = RH[qp];
                    Md[q] = cap(Md[qp] - (Ed[q] - Dd[q]) * sp, M_max);
                    if (t_u <= t_00_m2) Md[q] = 0.0;
                    const double L = (t_u >= t0) ? lv : ls;
                    const double Mp = Md[qp];
                    const double n_qv = Mp * Qvd[qp] - sp * (Ed[qp] * C[qp] - Dd[qp] * Qvd[qp] + rp * Edv[qp]);
                    const double n_v = Mp * Vd[qp] - sp * (Ed[qp] * V[qp] - Dd[qp] * Vd[qp]);
                    const double n_w = Mp * Wd[qp] - sp * (Ed[qp] * W[qp] - Dd[qp] * Wd[qp]);
                    const double n_s = (Mp * Sd[qp] - sp * (Ed[qp] * Sx[qp] - Dd[qp] * Sd[qp] - L * rp * Edv[qp])) / s_0;
                    if (std::fabs(Md[q]) > coeff_recurr) {
                        const double r = 1.0 / Md[q];
                        Qvd[q] = n_qv * r; Vd[q] = n_v * r; Wd[q] = n_w * r; Sd[q] = n_s * r;
                    } else {
                        Qvd[q] = C[q]; Vd[q] = V[q]; Wd[q] = W[q]; Sd[q] = Sx[q];
                    }
                    Ud[q] = U[q] + Md[q] * inv_a_d / dmax(RH[q], 1e-6);
                    if (Qvd[q] <= 0.0) Qvd[q] = 0.0;
                    if (Sd[q] <= 0.0) Sd[q] = 0.0;
                    if (t_u <= t_00) {
                        Ed[q] = Dd[q] = Md[q] = Qvd[q] = Ud[q] = Vd[q] = Wd[q] = 0.0;
                        Sd[q] = 1.0;
                        Edv[q] = Ep[q] = 0.0;
                    }
                    if (i > 0) Pc[q] = dmax(0.0, Pc[qp] + step[i] * RH[q] * (Gp[q] - Edv[q] - Ep[q]));
                }
                Pc[id(0)] = Pc[id(1)];
            }
            base_of[(size_t)j * km + k] = base;
            lfs_of[(size_t)j * km + k] = lfs;
            // rhsForcing (:1304-1417)
            for (int i = 0; i < im; i++) { Mu[id(i)] = cap(Mu[id(i)], M_max); Md[id(i)] = cap(Md[id(i)], M_max); }
            for (int i = 0; i < im - 1; i++) {
                const size_t q = id(i), qp = id(i + 1);
                const double L = (T[q] * t0 >= t0) ? lv : ls;
                const double isr = 1.0 / (step[i] * dmax(RH[q], 0.01));
                auto flux = [&](size_t x, const double* up, const double* dn, const double* env) {
                    return Mu[x] * (up[x] - env[x]) + Md[x] * (dn[x] - env[x]);
                };
                const double src = Cu[q] - Edv[q] - El[q] - Ep[q];
                MCt[q] = cap(-(flux(qp, Su, Sd, Sx) - flux(q, Su, Sd, Sx)) * isr * t0 + (L / cp_l) * src * t0, MCt_max);
                MCq[q] = cap(-(flux(qp, Qvu, Qvd, C) - flux(q, Qvu, Qvd, C)) * isr - src, MCq_max);
                MCv[q] = cap(-(flux(qp, Vu, Vd, V) - flux(q, Vu, Vd, V)) * isr * P.u_0, MCv_max);
                MCw[q] = cap(-(flux(qp, Wu, Wd, W) - flux(q, Wu, Wd, W)) * isr * P.u_0, MCv_max);
            }
        }
    // computeCAPE (:1259-1299): one array of im values, cleared, then every column in the order
    // k outer, j inner writes its cloud-base value and the integral up to its level of free
    // sinking, so each level ends holding the last column that reached it
    // alf = (1 - eps) / eps with eps = 1 / (R_WaterVapour / R_Air) (:99-101; the defaults'
    // R_Air 296.8 and R_WaterVapour 461.5, cAtmosphereDefaults.cpp.inc:93-94)
    const double rwra = 461.5 / 296.8, eps_thermo = 1.0 / rwra, alf = (1.0 - eps_thermo) / eps_thermo;
    s.cape.assign(im, 0.0);
    for (int k = 0; k < km; k++)
        for (int j = 0; j < jm; j++) {
            auto id = [&](int i) { return (size_t)i * pl + (size_t)j * km + k; };
            const int ib = base_of[(size_t)j * km + k], il = lfs_of[(size_t)j * km + k];
            auto term = [&](int i) {
                const double t_u = T[id(i)] * t0, t_u_add = t_u + t_add_u;
                const double t_vir = t_u_add * (1.0 + alf * Qvu[id(i)] - Qcu[id(i)]);
                const double t_vir_env = dmax(t_u * (1.0 + alf * C[id(i)] - CL[id(i)]), 1.0);
                const double exp_rm = 1.0 / (s.rad[i] + 1.0);
                return P.g * step[i] / exp_rm * (t_vir - t_vir_env) / t_vir_env;
            };
            s.cape[ib] = term(ib);
            for (int i = ib; i < il && i < im - 1; i++) s.cape[i + 1] = s.cape[i] + term(i);
        }
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------------------
extern "C" {

orc_state* orc_create_blank(int im, int jm, int km, const orc_member* m) {
    if (im < 6 || jm < 6 || km < 6) return nullptr;
    orc_state* s = new orc_state();
    apply_member(s->P, m);
    init_grid(*s, im, jm, km);
    return s;
}
orc_state* orc_create(int im, int jm, int km, uint64_t seed, const orc_member* m) {
    orc_state* s = orc_create_blank(im, jm, km, m);
    if (!s) return nullptr;
    init_state(*s, seed, 0);
    return s;
}
orc_state* orc_create_thermo(int im, int jm, int km, const orc_member* m) {
    orc_state* s = orc_create_blank(im, jm, km, m);
    if (!s) return nullptr;
    init_state(*s, 0, 1);
    return s;
}
void orc_destroy(orc_state* s) { delete s; }
double* orc_field_ptr(orc_state* s, int f) { return (f >= 0 && f < ORC_NFIELD) ? s->F(f) : nullptr; }
double* orc_field2d_ptr(orc_state* s, int f) { return (f >= 0 && f < ORC2_NFIELD) ? s->f2[f].data() : nullptr; }
int orc_total_iter(const orc_state* s) { return s->total_iter; }
void orc_set_iter(orc_state* s, int iter_n, int total) { s->iter_n = iter_n; s->total_iter = total; }
double orc_level_mean(const orc_state* s, int i) { return s->t_ref[i]; }
void orc_radiation(orc_state* s) { radiation(*s); }
double* orc_rad_ptr(orc_state* s, int which) {
    std::vector<double>* v[4] = {&s->radiation, &s->epsilon, &s->albedo, &s->sw};
    return (which >= 0 && which < 4 && !v[which]->empty()) ? v[which]->data() : nullptr;
}
double orc_scalar(const orc_state* s, const char* name) {
    const std::string n(name);
    if (n == "t_skin") return s->P.t_skin;
    if (n == "dr") return s->dr;
    if (n == "dthe") return s->dthe;
    if (n == "dphi") return s->dphi;
    if (n == "metric_r0") return s->metric_r0;
    if (n == "shell_len") return s->shell_len;
    if (n == "re_turb") return s->re_turb;
    if (n == "R_mix") return s->P.comp.R_mix;
    if (n == "M_bg") return s->P.comp.M_bg;
    if (n == "R_bg") return s->P.comp.R_bg;
    if (n == "q_CO2") return s->P.comp.q_CO2;
    if (n == "q_bg") return s->P.comp.q_bg;
    if (n == "r_air") return s->P.r_air;
    if (n == "c_0") return s->P.c_0;
    if (n == "co2_0") return s->P.co2_0;
    if (n == "t_surf_equator") return s->P.t_surf_equator;
    if (n == "t_surf_pole") return s->P.t_surf_pole;
    if (n.rfind("hgt", 0) == 0) return (double)s->hgt[std::stoi(n.substr(3))];
    if (n.rfind("rad", 0) == 0) return s->rad[std::stoi(n.substr(3))];
    if (n.rfind("the", 0) == 0) return s->the[std::stoi(n.substr(3))];
    return NAN;
}
void orc_ice_scheme(orc_state* s) { ice_scheme(*s); }
void orc_set_ice_scheme(orc_state* s, int on) {
    s->ice_on = on != 0;
    if (s->ice_on) ensure_mp(*s);
}
void orc_ice_store(orc_state* s) { ice_store(*s); }
void orc_set_moist_convection(orc_state* s, int on) {
    s->conv_on = on != 0;
    if (on) orc_set_ice_scheme(s, 1);
}
void orc_moist_convection(orc_state* s) { moist_convection(*s); }
int orc_cape(const orc_state* s, double* out) {
    if (s->cape.size() != (size_t)s->im) return -1;
    for (int i = 0; i < s->im; i++) out[i] = s->cape[i];
    return 0;
}
double* orc_mc_ptr(orc_state* s, int which) {
    if (which < 0 || which >= ORC_NMC) return nullptr;
    ensure_mc(*s);
    return s->mc[which].data();
}
double* orc_evap_ptr(orc_state* s) {
    ensure_mc(*s);
    return s->evap.data();
}
double* orc_mp_ptr(orc_state* s, int which) {
    if (which < 0 || which >= ORC_NMP) return nullptr;
    ensure_mp(*s);
    return s->mp[which].data();
}
int orc_set_turb_model(orc_state* s, int model) {
    if (model < 0 || model > 2) return -1;
    s->turb_model = model;
    return 0;
}
void orc_turbulence_init(orc_state* s) { turbulence_init(*s); }
void orc_set_sst_scan(orc_state* s, int on) {
    s->sst_scan = on != 0;
    if (s->sst_scan && s->h.size() != s->n) s->h.assign(s->n, 0.0);
}
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
void orc_step(orc_state* s, int order) { step(*s, order); }
void orc_run(orc_state* s, int n, int order) { for (int q = 0; q < n; q++) step(*s, order); }
void orc_op(orc_state* s, int op, int order) {
    switch (op) {
        case ORC_OP_POISSON: poisson(*s); break;
        case ORC_OP_SATADJ: satadj(*s); break;
        case ORC_OP_DAMP_MOIST:
            for (int f : {ORC_T, ORC_ICE, ORC_C, ORC_CLOUD}) damp_wiggles(*s, s->F(f));
            break;
        case ORC_OP_MOIST_CLAMP: moist_clamp(*s); break;
        case ORC_OP_DENSITIES: densities(*s, false); break;
        case ORC_OP_EVAPORATION: evaporation(*s); break;
        case ORC_OP_TURBULENCE: turbulence(*s); break;
        case ORC_OP_DAMP_TURB:
            for (int f : {ORC_TKE, ORC_DIS, ORC_NUE, ORC_PROD, ORC_TKE_SRC, ORC_DIS_SRC}) damp_wiggles(*s, s->F(f));
            break;
        case ORC_OP_BC: bc_all(*s); break;
        case ORC_OP_RK4: rk4(*s, order); break;
        case ORC_OP_POLAR: for (int f : {ORC_U, ORC_V, ORC_W}) polar_filter(*s, s->F(f)); break;
        case ORC_OP_RADIAL:
            radial_filter(*s, s->F(ORC_U), false);
            radial_filter(*s, s->F(ORC_V), true);
            radial_filter(*s, s->F(ORC_W), true);
            break;
        case ORC_OP_STORE: store(*s); break;
        case ORC_OP_LEVEL_MEAN: level_mean(*s); break;
        case ORC_OP_PROJECT: project_initial_velocity(*s, 200); break;
        default: break;
    }
}

double orc_p_sat(double T) { return p_sat(T); }
double orc_p_subl(double T) { return p_subl(T); }
double orc_p_auto(double T) { return p_auto(T); }
double orc_latent_heat(double T) { return L_vap(T); }
double orc_q_sat(double E, double p, double Mo) { return q_sat(E, p, Mo); }
double orc_dqsat_dT(double T, double p, double Mo) { return dqdT_from(p_auto(T), L_auto(T), T, p, Mo); }
double orc_moist_lapse(double T, double p, double Mo, double cp, double Rm, double g) {
    return moist_lapse(T, p, Mo, cp, Rm, g);
}
double orc_cp_of(double c, double co2, double T, double Mb) { return cp_of(c, co2, T, Mb); }
double orc_R_of(double c, double co2, double Rb) { return R_of(c, co2, Rb); }
double orc_M_nonwater(double c, double co2, double Mb) { return M_nonwater(c, co2, Mb); }
void orc_resolve(double xh, double xc, double xn, double* out) {
    const Comp C = resolve(xh, xc, xn);
    out[0] = C.M_mean; out[1] = C.R_mix; out[2] = C.q_H2O; out[3] = C.q_CO2; out[4] = C.q_bg;
    out[5] = C.M_bg; out[6] = C.R_bg;
}
void orc_satadj_cells(int n, double* t, double* c, double* cloud, double* ice, double* gr,
                      const double* co2, const double* ps, double* scc, double dt_dim, int interior) {
    Params P; P.comp = resolve(P.x_H2O, P.x_CO2, P.x_N2);
    SatConsts K{P.t_0, P.t_00, P.comp.M_bg, P.t_max_phys, P.cloud_cap};
    for (int q = 0; q < n; q++) {
        if (interior) satadj_cell(K, t[q], c[q], cloud[q], ice[q], co2[q], ps[q], scc[q], dt_dim);
        clamp_fade_cell(K, t[q], c[q], cloud[q], ice[q], gr[q], co2[q], ps[q], scc[q]);
    }
}

}  // extern "C"
