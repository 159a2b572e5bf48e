// Device-side definitions shared by the ATHAD_COND HIP kernels (gfx950).
//
// One integration ("member") is described by a MemberDesc that lives in device memory:
// grid sizes, physical constants, field pointers and 1-D tables. Every kernel takes an
// array of them and uses blockIdx.z as the member index, so a single grid, an ensemble of
// uneven grids and the serving batch all run through the same launches.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace athad {

// Field ids (same numbering as include/athad_hip.h).
enum Field {
    F_T = 0, F_U, F_V, F_W, F_C, F_CLOUD, F_ICE, F_GR, F_CO2, F_TKE, F_DIS,
    F_TN, F_UN, F_VN, F_WN, F_CN, F_CLOUDN, F_ICEN, F_GRN, F_CO2N, F_TKEN, F_DISN,
    F_P_DYN, F_P_STAT, F_R_HUMID, F_R_DRY, F_T_EQ, F_NUE, F_PROD, F_TKE_SRC,
    F_DIS_SRC, F_AUX_U, F_AUX_V, F_AUX_W, F_S_C_C,
    NFIELD
};
constexpr int NPROG = 11;   // t u v w c cloud ice gr co2 tke dis
constexpr int NTMP = 12;    // scratch fields for the filter passes
// Per-(i, j) metric coefficients of the RK4 stencils (cAtmosphereModel.h:170-178, 282-284;
// RHS_Atm_Turb.cpp:86-130, 392-395, 976-992, 1014-1019), formed once on the host: they depend
// on the radius and the latitude only. Entry (i, j) holds NGEO doubles in this order.
enum GeoCoeff {
    G_AR = 0, G_AT, G_AP, G_CEXP, G_CIRM, G_CIRS, G_WR, G_WT, G_WP, G_TWO_OVER, G_CRS, G_LC2,
    G_CU, G_CVW, G_CIRM2, G_CIRM2S, G_C2COS, G_DRAG, G_COR, G_S4, NGEO
};
// Arrays of the two-category ice scheme (same numbering as include/athad_hip.h): the rain and
// snow fluxes, their stored copies, the capped total and the microphysics sources.
// Moist-convection arrays (MoistConvection.h), in the reference harness's order.
enum McField {
    MC_S = 0, MC_M_U, MC_M_D, MC_S_U, MC_Q_V_U, MC_Q_C_U, MC_C_U, MC_E_L, MC_E_D_EVAP, MC_E_P, MC_G_P,
    MC_S_D, MC_Q_V_D, MC_P_CONV, MC_E_U, MC_D_U, MC_U_U, MC_V_U, MC_W_U, MC_E_D, MC_D_D, MC_U_D, MC_V_D,
    MC_W_D, MC_MC_T, MC_MC_Q, MC_MC_V, MC_MC_W, MC_DEEP_BEG, MC_DEEP_END, MC_CLOUD_BASE, MC_LFS, NMC
};

enum MpField { MP_P_RAIN = 0, MP_P_SNOW, MP_P_RAINN, MP_P_SNOWN, MP_PRECIP, MP_S_V, MP_S_C, MP_S_I, MP_S_R, MP_S_S, NMP };

// Per-member constants, all resolved on the host in double.
struct Consts {
    int im, jm, km;
    int turb;           // turbulence closure, TurbulenceAtm.h:24 order (kTurbKEpsilon ...)
    long long n;                                   // im*jm*km
    double dr, dthe, dphi;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
    double metric_r0, rad0, shell_len;
    // configuration
    double omega, g, t_0, t_skin, r_air, dt, L_atm, u_0, t_00, cloud_cap, t_max_phys;
    double c_h2o_dry_top, pr_turb, abl_height, sc_wv, sc_co2;
    // composition
    double R_mix, M_bg, R_bg, p_sl_factor;
    // derived coefficients
    double force_nd, re_turb, coeff_energy_p, diff_prec_re_inv, diff_co2_re_inv;
    double nue_max, nue_air_nd, tke_max_nd, vel_max_nd, om_wall, sqrt_cnue;
    double hs_k_a, hs_k_s, drag_coeff;
    double inv_re_turb, inv_pr_turb, dt_sixth;   // 1/re_turb, 1/pr_turb, dt/6 (kernel-wide)
    // moist convection's tendency coefficients (RHS_Atm_Turb.cpp:370-372) and the Dalton
    // evaporation coefficient of the Evaporation diagnostic it reads (cAtmosphereDefaults.cpp.inc:60)
    double coeff_mc_t, coeff_mc_vel, coeff_mc_q, coeff_dalton;
    double evap_r, evap_w_norm;
    // radiation (MultiLayerRadiation, mode 2)
    double sigma, kappa_h2o, kappa_co2, kappa_bg, geothermal, sw_equator, sw_pole;
    double albedo_surface, albedo_cloud, omega_rad;
    // initial-state recipe
    double t_surf_eq, t_surf_pole, init_M_nw, init_f_co2, co2_scale;
    double t_000;       // lower bound of the snow band of the ice scheme
    double c_0, co2_0, r_earth;   // reference water and CO2 mass fractions, planet radius [km]
    double coeff_energy, lv, ls;   // latent heating of the microphysics sources (RHS_Atm_Turb.cpp:367, 965-966)
    unsigned long long seed;
    // per-member state knobs
    int order;          // 0 lex, 1 jacobi
    int sym;            // 1: N-S symmetric variant of the Jacobi form
};

struct MemberDesc {
    Consts K;
    void* f[NFIELD];         // storage precision S, except F_P_STAT which is always double
    void* alt[NPROG];        // Jacobi output buffers (swapped with f[0..10] every step)
    void* tmp[NTMP];         // filter scratch
    double* vel_star;        // jm*km
    double* t_top;           // jm*km
    double* t_ref;           // im
    const double* rad;       // im       grid radial coordinate
    const double* sint;      // jm       sin(theta) floored at 0.55 (RK4 / Poisson metric)
    const double* cost;      // jm       cos(theta)
    const double* sinr;      // jm       raw sin(theta) (level mean weight, turbulence)
    const double* hgt;       // im       layer heights (float-rounded, stored as double)
    const double* nue_f;     // im       ABL taper factor of clamp_nue
    const double* step_dim;  // im       SaturationAdjustment dt_dim = float(h[i+1]-h[i])/1.6
    const int* passes;       // jm       polar filter pass count
    const double* geo;       // im*jm*NGEO  per-(i, j) metric coefficients of the RK4 stencils
    int* counters;           // [0] iter_n, [1] total_iter_count (device resident)
    double* scratch;         // reductions
    double* radiation;       // n        upward long-wave flux (allocated at the first radiation call)
    double* epsilon;         // n        layer emissivity
    double* albedo;          // jm*km    surface short-wave albedo
    void* mp[NMP];           // n        ice-scheme arrays, storage precision S (allocated at first use)
    double* mc[NMC];         // n        moist-convection arrays, FP64 (allocated at first use)
    double* evap;            // jm*km    evaporation diagnostic [mm/d] the convection trigger reads
    const double* conv_cos;  // jm       cos(theta[90] - theta[j]), the convection's zonal circle factor
    double* conv_qs[2];      // n        per-cell saturation of the lifted (+0.2 K) and the resting air
    double* conv_cape;       // im       computeCAPE's diagnostic (MoistConvection.h:1259-1299)
    int* conv_cape_rng;      // 2 jm km  per column, the first and last level its CAPE profile reached
    double* conv_tmp[2];     // n        FP64 damping scratch of the convection arrays (with conv_qs,
                             //          which the next call recomputes): allocated with them
};

#define ATHAD_DEV __device__ __forceinline__
#define ATHAD_HD __host__ __device__ __forceinline__

ATHAD_DEV size_t ix3(const Consts& K, int i, int j, int k) { return ((size_t)i * K.jm + j) * K.km + k; }

template <typename T> ATHAD_HD T tmax(T a, T b) { return (a < b) ? b : a; }       // std::max
template <typename T> ATHAD_HD T tmin(T a, T b) { return (b < a) ? b : a; }       // std::min
template <typename T> ATHAD_HD T tclamp(T v, T lo, T hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

ATHAD_DEV bool finite_bits(double v) {
    return (__double_as_longlong(v) & 0x7FF0000000000000LL) != 0x7FF0000000000000LL;
}
ATHAD_DEV bool finite_bits(float v) { return (__float_as_uint(v) & 0x7F800000u) != 0x7F800000u; }

// Workgroup index remap so that consecutive ranges of the domain land on the same XCD
// (workgroups are dealt round-robin over the 8 XCDs, each with its own L2). The launch
// uses a block count that is a multiple of 8; returns -1 for the padding blocks.
ATHAD_DEV long long xcd_block(unsigned b, unsigned nb_launch, long long nb_real) {
    const unsigned per = nb_launch >> 3;
    const long long r = (long long)(b & 7u) * per + (b >> 3);
    return (r < nb_real) ? r : -1;
}

// ------------------------------------------------------------------------------------
// Water saturation and mixture thermodynamics (always FP64).
// Restated from SaturationH2O.h:68-301 and MixtureAtm.h:82-217.
// ------------------------------------------------------------------------------------
constexpr double kTcrit = 647.096, kPcrit = 220640.0, kTtriple = 273.16, kPtriple = 6.11657;
constexpr double kNoSat = -1.0;
// the turbulence closures in TurbulenceAtm.h:24 order
constexpr int kTurbKEpsilon = 0, kTurbKOmega = 1, kTurbKOmegaSST = 2;
constexpr double kRstar = 8.314462618, kMH2O = 0.018015, kMCO2 = 0.044010;
#define kRH2O (kRstar / kMH2O)
#define kRCO2 (kRstar / kMCO2)

ATHAD_HD double th_p_sat(double T) {
    if (T >= kTcrit) return kNoSat;
    if (T <= 0.0) return 0.0;
    const double th = 1.0 - T / kTcrit;
    const double sq = sqrt(th);
    const double t3 = th * th * th;
    const double t4 = t3 * th;
    const double s = -7.85951783 * th + 1.84408259 * (th * sq) + -11.7866497 * t3
                   + 22.6807411 * (t3 * sq) + -15.9618719 * t4 + 1.80122502 * (t4 * t3 * sq);
    return kPcrit * exp((kTcrit / T) * s);
}
ATHAD_HD double th_p_subl(double T) {
    if (T <= 0.0) return 0.0;
    const double th = T / kTtriple;
    const double s = -0.212144006e2 * pow(th, 0.333333333e-2) + 0.273203819e2 * pow(th, 0.120666667e1)
                   + -0.610598130e1 * pow(th, 0.170333333e1);
    return kPtriple * exp(s / th);
}
ATHAD_HD double th_p_auto(double T) {
    if (T >= kTcrit) return kNoSat;
    return (T >= kTtriple) ? th_p_sat(T) : th_p_subl(T);
}
ATHAD_HD double th_L_vap(double T) {
    if (T >= kTcrit) return 0.0;
    return 2.501e6 * pow((kTcrit - T) / (kTcrit - 273.15), 0.38);
}
ATHAD_HD double th_L_sub(double T) { return th_L_vap(T) + 3.337e5; }
ATHAD_HD double th_L_auto(double T) { return (T >= kTtriple) ? th_L_vap(T) : th_L_sub(T); }
ATHAD_HD double th_q_sat(double E, double p, double Mo) {
    if (E == kNoSat) return 1.0;
    if (!(p > 0.0)) return 0.0;
    if (E >= p) return 1.0;
    const double x = E / p;
    const double num = x * kMH2O;
    const double den = num + (1.0 - x) * Mo;
    return (den > 0.0) ? num / den : 1.0;
}
ATHAD_HD double th_sat_A(double E, double p, double Mo) {
    if (E == kNoSat || !(p > 0.0) || E >= p) return 0.0;
    const double x = E / p;
    const double den = x * kMH2O + (1.0 - x) * Mo;
    if (!(den > 0.0)) return 0.0;
    return x * kMH2O * Mo / (den * den);
}
ATHAD_HD double th_dqdT(double E, double L, double T, double p, double Mo) {
    if (!(T > 0.0)) return 0.0;
    return th_sat_A(E, p, Mo) * L / (kRH2O * T * T);
}
// moist-adiabatic lapse rate with the saturation pressure E = p_auto(T) supplied
ATHAD_HD double th_moist_lapse_E(double T, double p, double Mo, double cp, double Rm, double g, double E) {
    if (!(cp > 0.0) || !(T > 0.0) || !(Rm > 0.0)) return 0.0;
    const double A = th_sat_A(E, p, Mo);
    if (!(A > 0.0)) return g / cp;
    const double L = th_L_auto(T);
    const double num = 1.0 + L * A / (Rm * T);
    const double den = cp + L * L * A / (kRH2O * T * T);
    return (den > 0.0) ? g * num / den : g / cp;
}
ATHAD_HD double th_moist_lapse(double T, double p, double Mo, double cp, double Rm, double g) {
    return th_moist_lapse_E(T, p, Mo, cp, Rm, g, th_p_auto(T));
}
ATHAD_HD double th_q_sat_at(double T, double p, double Mo) { return th_q_sat(th_p_auto(T), p, Mo); }

ATHAD_HD double th_shomate(double T, double A, double B, double C, double D, double E) {
    const double t = tmin(tmax(T, 298.0), 2000.0) * 1.0e-3;
    const double t2 = t * t;
    return A + B * t + C * t2 + D * t2 * t + E / t2;
}
ATHAD_HD void th_split(double c, double co2, double& qv, double& qc, double& qb) {
    qv = tmin(tmax(c, 0.0), 1.0);
    qc = tmin(tmax(co2, 0.0), 1.0);
    const double s = qv + qc;
    if (s > 1.0) { qv /= s; qc /= s; }
    qb = tmax(0.0, 1.0 - qv - qc);
}
ATHAD_HD double th_R_of(double c, double co2, double Rb) {
    double qv, qc, qb; th_split(c, co2, qv, qc, qb);
    return qv * kRH2O + qc * kRCO2 + qb * Rb;
}
ATHAD_HD double th_cp_of(double c, double co2, double T, double Mb) {
    double qv, qc, qb; th_split(c, co2, qv, qc, qb);
    return qv * (th_shomate(T, 30.09200, 6.832514, 6.793435, -2.534480, 0.082139) / kMH2O)
         + qc * (th_shomate(T, 24.99735, 55.18696, -33.69137, 7.948387, -0.136638) / kMCO2)
         + qb * (th_shomate(T, 28.98641, 1.853978, -9.647459, 16.63537, 0.000117) / Mb);
}
ATHAD_HD double th_M_nonwater(double c, double co2, double Mb) {
    double qv, qc, qb; th_split(c, co2, qv, qc, qb);
    const double s = qc + qb;
    if (!(s > 0.0)) return Mb;
    const double inv = (qc / s) / kMCO2 + (qb / s) / Mb;
    return (inv > 0.0) ? 1.0 / inv : Mb;
}

}  // namespace athad
