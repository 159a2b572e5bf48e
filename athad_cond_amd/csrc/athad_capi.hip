// C-ABI of the ATHAD_COND HIP library: context lifetime, state transfer, the per-operator
// entry points and the scoped iteration (include/athad_hip.h).
//
// Host-side grid and constant setup restates cAtmosphereModel's (cAtmosphereModel.cpp:44-60,
// 327-399; cAtmosphereModel.h:441-469) so device tables match the reference bit for bit.
#include "../../include/athad_hip.h"
#include "athad_device.h"
#include "athad_kernels.h"

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <string>
#include <vector>

using namespace athad;

namespace {

std::string g_create_error;

// ---- configuration values (cli/config_cond.xml) ----------------------------------------
struct Config {
    double omega = 0.000317, g = 9.8066, r_Earth = 6370.001, t_0 = 273.15, t_skin = 254.0;
    double dt_visc = 4e-05, L_atm = 6287.5, zeta = 3.0, sc_WaterVapour = 0.61, sc_CO2 = 0.96;
    double pr_turb = 0.9, abl_height = 10000.0, u_0 = 8.0, cloud_cap = 0.05, t_max_phys = 2000.0;
    double t_00 = 236.15, t_000 = 253.15, cp_l = 1349.0, co2_scale = 1.0, lv = 2520000.0, ls = 2830000.0;
    // radiation: cAtmosphereDefaults.cpp.inc:31-43,72-73 and run_3D_loop's omega_rad
    double sigma = 5.67028e-08, kappa_H2O = 0.01, kappa_CO2 = 0.001, kappa_bg = 1e-06;
    double geothermal_flux = 150.0, rad_equator_short = 298.0, rad_pole_short = 0.0;
    double albedo_surface = 0.08, albedo_cloud = 0.5, t_skin_relax = 0.25, omega_rad = 0.05;
};

athad_member_params default_member() {
    athad_member_params p;
    p.t_surf_equator = 513.15; p.t_surf_pole = 503.15; p.r_air = 40.8;
    p.x_H2O = 0.5578; p.x_CO2 = 0.4109; p.x_N2 = 0.0313;
    p.c_0 = 0.3464; p.co2_0 = 0.6233; p.c_h2o_dry_top = 0.00423;
    return p;
}

// AtmMixture::resolve and M_nonwater (MixtureAtm.h:82-109, :152-217), host copies.
struct Comp { double M_mean, R_mix, q_H2O, q_CO2, q_bg, M_bg, R_bg; };
constexpr double hRstar = 8.314462618, hMH2O = 0.018015, hMCO2 = 0.044010, hMN2 = 0.028014;
Comp resolve(double xh, double xc, double xn) {
    Comp C;
    const double mh = xh * hMH2O, mc = xc * hMCO2, mb = xn * hMN2;
    C.M_mean = mh + mc + mb;
    C.R_mix = hRstar / C.M_mean;
    C.q_H2O = mh / C.M_mean;
    C.q_CO2 = mc / C.M_mean;
    C.q_bg = mb / C.M_mean;
    C.M_bg = (xn > 0.0) ? (mb / xn) : hMN2;
    C.R_bg = hRstar / C.M_bg;
    return C;
}
double M_nonwater_h(double c, double co2, double Mb) {
    double qv = std::min(std::max(c, 0.0), 1.0), qc = std::min(std::max(co2, 0.0), 1.0);
    const double s = qv + qc;
    if (s > 1.0) { qv /= s; qc /= s; }
    const double qb = std::max(0.0, 1.0 - qv - qc);
    const double t = qc + qb;
    if (!(t > 0.0)) return Mb;
    const double inv = (qc / t) / hMCO2 + (qb / t) / Mb;
    return (inv > 0.0) ? 1.0 / inv : Mb;
}

struct MemberHost {
    athad_grid g{};
    athad_member_params p{};
    Consts K{};
    size_t n = 0;
    void* field[NFIELD] = {};
    void* alt[NPROG] = {};
    void* tmp[NTMP] = {};
    double *vel_star = nullptr, *t_top = nullptr, *t_ref = nullptr, *tables = nullptr, *scratch = nullptr;
    int* counters = nullptr;
    int* passes = nullptr;
    double *radiation = nullptr, *epsilon = nullptr, *albedo = nullptr;   // radiation outputs, on first use
    void* mp[NMP] = {};            // ice-scheme arrays, on first use
    double* mc[NMC] = {};          // moist-convection arrays, on first use
    double *evap = nullptr, *conv_cos = nullptr;
    double* conv_qs[2] = {};
    double* conv_tmp[2] = {};
    double* conv_cape = nullptr;   // im doubles, then 2 jm km ints (each column's CAPE level range)
    int iter_n = 0, total = 0;     // host mirror of the device counters
    double wall_y1 = 1.0e-6;       // first layer height [m], floored as apply_wall_bc floors it
    // table offsets inside 'tables'
    size_t o_rad = 0, o_sint = 0, o_cost = 0, o_sinr = 0, o_hgt = 0, o_nuef = 0, o_step = 0, o_geo = 0;
};

}  // namespace

struct athad_ctx {
    int nmem = 0, precision = ATHAD_FP64, order = ATHAD_ORDER_JACOBI;
    int turb_model = kTurbKOmegaSST;   // the turbulence closure of every member
    std::vector<MemberHost> mem;
    MemberDesc* d_md[2] = {nullptr, nullptr};
    std::vector<MemberDesc> h_md[2];
    int parity = 0;
    hipStream_t stream = nullptr;
    // second stream: the turbulence update runs beside the moist thermodynamics (disjoint
    // fields), forked and joined with events (also inside a captured graph)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // the filters of an iteration may run on the second stream beside the next iteration's
    // Poisson sweep and moist thermodynamics; s2_pending: not yet joined into `stream`
    hipEvent_t ev_ffork = nullptr, ev_fjoin = nullptr;
    bool s2_pending = false;
    Batch batch;
    std::string err;
    // profiling
    bool prof_on[KID_COUNT] = {};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev[KID_COUNT];
    double prof_ms[KID_COUNT] = {};
    int prof_n[KID_COUNT] = {};
    bool capturing = false;
    // 4-iteration graph
    hipGraphExec_t graph = nullptr;
    int graph_phase = -1, graph_parity = -1;
    std::vector<double> staging;   // host staging for conversions
    // fused damp_wiggles tile: shape index and levels per block (autotuned at creation)
    int tile_shape = 0, tile_ich = 1 << 20;
    // the two-category ice scheme runs inside the moist block (CategoryIceScheme = 2)
    bool ice_on = false;
    // the fluxes may differ from their stored copies (changed outside a step's moist block)
    bool mp_dirty = false;
    // radiation (mode 2) and the skin-temperature step after every iteration's store; the
    // device descriptors then hold a newer t_skin than the host mirror (tskin_dev)
    bool rad_on = false, tskin_dev = false;
    // the on-demand arrays of the ice scheme (mp) and of moist convection (mc) are complete on
    // every member and published in the device descriptors
    bool mp_ready = false, mc_ready = false;
    // moist convection inside the moist block, after the ice scheme (cAtmosphereModel.cpp:1247-1262)
    bool conv_on = false;
};

namespace {

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                 \
            return ATHAD_ERR_DEVICE;                                                        \
        }                                                                                   \
    } while (0)

size_t elem_size(const athad_ctx* c) { return c->precision == ATHAD_MIXED ? 4 : 8; }
size_t field_bytes(const athad_ctx* c, const MemberHost& m, int f) {
    return m.n * ((f == F_P_STAT) ? 8 : elem_size(c));
}

void fill_consts(MemberHost& m, const Config& cfg, int order, int turb) {
    Consts& K = m.K;
    const int im = m.g.im, jm = m.g.jm, km = m.g.km;
    K.im = im; K.jm = jm; K.km = km;
    K.turb = turb;
    K.n = (long long)im * jm * km;
    const double pi180 = 180.0 / M_PI;
    K.dr = 1.0 / (double)(im - 1);
    K.dthe = (180.0 / (double)(jm - 1)) / pi180;
    K.dphi = (360.0 / (double)(km - 1)) / pi180;
    K.inv_2dr = 1.0 / (2.0 * K.dr); K.inv_2dthe = 1.0 / (2.0 * K.dthe); K.inv_2dphi = 1.0 / (2.0 * K.dphi);
    K.inv_dr2 = 1.0 / (K.dr * K.dr); K.inv_dthe2 = 1.0 / (K.dthe * K.dthe); K.inv_dphi2 = 1.0 / (K.dphi * K.dphi);
    K.rad0 = 1.0;
    double rlast = 1.0;
    for (int l = 1; l < im; l++) rlast = rlast + K.dr;
    const double span = rlast - 1.0;
    K.shell_len = (span > 0.0) ? (std::exp(cfg.zeta * span) - 1.0) * cfg.L_atm / span : cfg.L_atm;
    K.metric_r0 = (K.shell_len > 0.0) ? (cfg.r_Earth * 1.0e3 / K.shell_len) : 0.0;
    K.omega = cfg.omega; K.g = cfg.g; K.t_0 = cfg.t_0; K.t_skin = cfg.t_skin; K.r_air = m.p.r_air;
    K.dt = cfg.dt_visc; K.L_atm = cfg.L_atm; K.u_0 = cfg.u_0; K.t_00 = cfg.t_00; K.cloud_cap = cfg.cloud_cap;
    K.t_max_phys = cfg.t_max_phys; K.c_h2o_dry_top = m.p.c_h2o_dry_top; K.pr_turb = cfg.pr_turb;
    K.abl_height = cfg.abl_height; K.sc_wv = cfg.sc_WaterVapour; K.sc_co2 = cfg.sc_CO2;
    const Comp C = resolve(m.p.x_H2O, m.p.x_CO2, m.p.x_N2);
    K.R_mix = C.R_mix; K.M_bg = C.M_bg; K.R_bg = C.R_bg;
    K.p_sl_factor = 1e-2 * m.p.r_air * C.R_mix;
    K.force_nd = cfg.omega * K.shell_len / cfg.u_0;
    K.re_turb = 0.4 * 0.1 / 1.5e-5;
    K.coeff_energy_p = cfg.u_0 * cfg.u_0 / (cfg.cp_l * cfg.t_0);
    K.diff_prec_re_inv = 1.0 / (cfg.sc_WaterVapour * K.re_turb);
    K.diff_co2_re_inv = 1.0 / (cfg.sc_CO2 * K.re_turb);
    K.inv_re_turb = 1.0 / K.re_turb;
    K.inv_pr_turb = 1.0 / K.pr_turb;
    K.dt_sixth = K.dt / 6.0;
    K.nue_max = 1000.0 / (cfg.u_0 * cfg.L_atm);
    K.nue_air_nd = 1.5e-5 / (cfg.u_0 * cfg.L_atm);
    K.tke_max_nd = 1000.0 / (cfg.u_0 * cfg.u_0);
    K.vel_max_nd = 100.0 / cfg.u_0;
    K.sqrt_cnue = std::sqrt(0.028);
    K.hs_k_a = 1.0 / (4.0 * 86400.0);
    K.hs_k_s = 1.0 / (0.4 * 86400.0);
    K.drag_coeff = (1.0 / 86400.0) * cfg.L_atm / cfg.u_0 * cfg.dt_visc;
    K.evap_r = std::exp(-1.0);
    K.evap_w_norm = (1.0 - K.evap_r) / (1.0 - std::pow(K.evap_r, 3 + 1));
    // RHS_Atm_Turb.cpp:370-372 (c_0 of the member), cAtmosphereDefaults.cpp.inc:60
    K.coeff_mc_t = cfg.L_atm / (cfg.u_0 * cfg.t_0);
    K.coeff_mc_vel = cfg.L_atm / (cfg.u_0 * cfg.u_0);
    K.coeff_mc_q = cfg.L_atm / (cfg.u_0 * m.p.c_0);
    K.coeff_dalton = 0.7;
    K.t_surf_eq = m.p.t_surf_equator; K.t_surf_pole = m.p.t_surf_pole;
    K.init_M_nw = M_nonwater_h(m.p.c_0, m.p.co2_0, C.M_bg);
    K.init_f_co2 = (C.q_CO2 + C.q_bg > 0.0) ? C.q_CO2 / (C.q_CO2 + C.q_bg) : m.p.co2_0 * cfg.co2_scale;
    K.co2_scale = cfg.co2_scale;
    K.t_000 = cfg.t_000;
    K.coeff_energy = 1.0 / (m.p.c_0 * cfg.cp_l * cfg.t_0);
    K.lv = cfg.lv; K.ls = cfg.ls;
    K.c_0 = m.p.c_0; K.co2_0 = m.p.co2_0; K.r_earth = cfg.r_Earth;
    K.sigma = cfg.sigma; K.kappa_h2o = cfg.kappa_H2O; K.kappa_co2 = cfg.kappa_CO2; K.kappa_bg = cfg.kappa_bg;
    K.geothermal = cfg.geothermal_flux; K.sw_equator = cfg.rad_equator_short; K.sw_pole = cfg.rad_pole_short;
    K.albedo_surface = cfg.albedo_surface; K.albedo_cloud = cfg.albedo_cloud; K.omega_rad = cfg.omega_rad;
    K.order = (order == ATHAD_ORDER_SYMMETRIC) ? ATHAD_ORDER_JACOBI : order;
    K.sym = (order == ATHAD_ORDER_SYMMETRIC) ? 1 : 0;
}

// The closure's value of dis at the wall (TurbulenceAtm.h:128-143), made dimensionless by
// L_atm / u_0: k-epsilon 0; k-omega 6 nu / (0.075 y1^2); SST 60 nu / (0.0333 y1^2), y1 the first
// layer height floored at 1e-6 m.
double wall_dis(int turb, double y1s, const Config& cfg) {
    const double nd_omega = cfg.L_atm / cfg.u_0;
    if (turb == kTurbKOmega) return 6.0 * 1.5e-5 / (0.075 * y1s * y1s) * nd_omega;
    if (turb == kTurbKOmegaSST) return 60.0 * 1.5e-5 / (0.0333 * y1s * y1s) * nd_omega;
    return 0.0;
}

// Reduction scratch per member: level means, field statistics, the checkpoint diagnostics
// (kernels' diag_offset region) and the streamfunction's three im x jm arrays or one plane of a
// field, with room to spare.
int64_t scratch_doubles(int im, int jm, int km) {
    const int64_t plane = std::max<int64_t>({3LL * im * jm, (int64_t)jm * km, (int64_t)im * km});
    return std::max<int64_t>(8192, (int64_t)sf_offset(im) + plane + 1024);
}

// 1-D tables: rad, sin/cos, layer heights, ABL taper, SatAdj dt_dim, polar passes.
void build_tables(MemberHost& m, const Config& cfg, std::vector<double>& tab, std::vector<int>& passes) {
    const int im = m.g.im, jm = m.g.jm;
    const Consts& K = m.K;
    std::vector<double> rad(im), the(jm);
    rad[0] = 1.0;
    for (int l = 1; l < im; l++) rad[l] = rad[l - 1] + K.dr;
    the[0] = 0.0;
    for (int l = 1; l < jm; l++) the[l] = the[l - 1] + K.dthe;
    // the symmetric form mirrors every latitude table about the equator row jm/2, where
    // cos(theta) is exactly zero
    const int jeq = (jm - 1) / 2;
    auto src_row = [&](int j) { return (K.sym && j > jeq) ? jm - 1 - j : j; };
    const float h = (float)cfg.L_atm;
    std::vector<double> hgt(im);
    for (int i = 0; i < im; i++) hgt[i] = (double)(float)((std::exp(cfg.zeta * (rad[i] - rad[0])) - 1) * h);
    m.o_rad = 0; m.o_sint = im; m.o_cost = m.o_sint + jm; m.o_sinr = m.o_cost + jm; m.o_hgt = m.o_sinr + jm;
    m.o_nuef = m.o_hgt + im; m.o_step = m.o_nuef + im;
    tab.assign(m.o_step + im, 0.0);
    for (int i = 0; i < im; i++) tab[m.o_rad + i] = rad[i];
    for (int j = 0; j < jm; j++) {
        const int jr = src_row(j);
        double s = std::sin(the[jr]);
        tab[m.o_sinr + j] = s;
        tab[m.o_sint + j] = (s < 0.55) ? 0.55 : s;
        double cth = std::cos(the[jr]);
        if (K.sym) cth = (j == jeq) ? 0.0 : (j > jeq ? -cth : cth);
        tab[m.o_cost + j] = cth;
    }
    for (int i = 0; i < im; i++) tab[m.o_hgt + i] = hgt[i];
    // clamp_nue taper (TurbulenceAtm.h:271-297)
    int i_abl = 0;
    while (i_abl < im - 1 && hgt[i_abl] <= cfg.abl_height) i_abl++;
    const double taper_dz = std::max(1.0, hgt[std::min(i_abl, im - 1)] - hgt[std::max(i_abl - 1, 0)]);
    for (int i = 0; i < im; i++) {
        const double z_agl = hgt[i] - hgt[0];
        double f = (cfg.abl_height + taper_dz - z_agl) / taper_dz;
        f = (f < 0.0) ? 0.0 : (1.0 < f) ? 1.0 : f;
        tab[m.o_nuef + i] = f;
    }
    // SaturationAdjustment.h:56-60,99
    for (int i = 0; i + 1 < im; i++) tab[m.o_step + i] = (double)(float)((float)hgt[i + 1] - (float)hgt[i]) / 1.6;
    tab[m.o_step + im - 1] = 1.0;
    // RK4 metric coefficients per (i, j) (GeoCoeff order), with the arithmetic the cell kernel
    // used to repeat per cell: RHS_Atm_Turb.cpp:86-130 (metric radius, sin theta floored at
    // 0.55), the Laplacian and advection weights, Coriolis :392-395 with force_nd, the Rayleigh
    // drag profile :1014-1019 and sin^4 theta of the Held-Suarez rate :976-992.
    m.o_geo = (tab.size() + 1) & ~(size_t)1;   // even: 16-byte aligned entries
    tab.resize(m.o_geo + (size_t)im * jm * NGEO, 0.0);
    for (int i = 0; i < im; i++) {
        const double rm = rad[i];
        const double exp_rm = 1.0 / (rm + 1.0);
        const double rmet = (K.metric_r0 > 0.0) ? (K.metric_r0 + (rm - K.rad0)) : rm;
        const double inv_rm = 1.0 / rmet;
        const double inv_rm2 = 1.0 / (rmet * rmet);
        double dp = 1.0 - (double)i / 5.0;
        if (dp < 0.0) dp = 0.0;
        if (dp > 1.0) dp = 1.0;
        for (int j = 0; j < jm; j++) {
            const double sinthe = tab[m.o_sint + j], costhe = tab[m.o_cost + j];
            const double sinthe2 = sinthe * sinthe;
            const double inv_rms = 1.0 / (rmet * sinthe);
            const double inv_rm2sinthe = inv_rm2 / sinthe;
            const double inv_rm2s2 = inv_rm2 / sinthe2;
            const double v_metric = (1.0 + costhe * costhe / sinthe2) * inv_rm2;
            const double wr = K.inv_dr2 * (exp_rm * exp_rm), wt = K.inv_dthe2 * inv_rm2, wp = K.inv_dphi2 * inv_rm2s2;
            const double Lc2 = 2.0 * (wr + wt + wp);
            double* q = &tab[m.o_geo + ((size_t)i * jm + j) * NGEO];
            q[G_AR] = exp_rm * 2.0 * K.inv_2dr;
            q[G_AT] = inv_rm * 2.0 * K.inv_2dthe;
            q[G_AP] = inv_rms * 2.0 * K.inv_2dphi;
            q[G_CEXP] = exp_rm; q[G_CIRM] = inv_rm; q[G_CIRS] = inv_rms;
            q[G_WR] = wr; q[G_WT] = wt; q[G_WP] = wp;
            q[G_TWO_OVER] = 2.0 * inv_rm * exp_rm;
            q[G_CRS] = costhe * inv_rm2sinthe;
            q[G_LC2] = Lc2;
            q[G_CU] = 2.0 * inv_rm2 - Lc2;
            q[G_CVW] = -v_metric - Lc2;
            q[G_CIRM2] = inv_rm2; q[G_CIRM2S] = inv_rm2sinthe; q[G_C2COS] = 2.0 * costhe * inv_rm2s2;
            q[G_DRAG] = K.drag_coeff * dp;
            q[G_COR] = K.force_nd * 2.0 * costhe;
            q[G_S4] = sinthe2 * sinthe2;
        }
    }
    // polar_zonal_filter pass counts (lib/Utils.cpp:557-571)
    passes.assign(jm, 0);
    const double sin_ref = std::cos(30.0 * (M_PI / 180.0));
    for (int j = 0; j < jm; j++) {
        const double sj = std::max(std::sin(the[src_row(j)]), 1.0e-6);
        if (sj >= sin_ref) continue;
        const double ratio = sin_ref / sj;
        int p = (int)std::lround(3.0 * (ratio * ratio - 1.0));
        passes[j] = std::min(std::max(p, 0), 12);
    }
    m.wall_y1 = std::max(hgt[1], 1.0e-6);
    m.K.om_wall = wall_dis(m.K.turb, m.wall_y1, cfg);
}

void make_desc(const athad_ctx* c, const MemberHost& m, int parity, MemberDesc& d) {
    std::memset(&d, 0, sizeof(d));
    d.K = m.K;
    for (int f = 0; f < NFIELD; f++) d.f[f] = m.field[f];
    for (int f = 0; f < NPROG; f++) d.alt[f] = m.alt[f];
    if (parity) for (int f = 0; f < NPROG; f++) { d.f[f] = m.alt[f]; d.alt[f] = m.field[f]; }
    for (int q = 0; q < NTMP; q++) d.tmp[q] = m.tmp[q];
    d.vel_star = m.vel_star; d.t_top = m.t_top; d.t_ref = m.t_ref;
    d.rad = m.tables + m.o_rad; d.sint = m.tables + m.o_sint; d.cost = m.tables + m.o_cost;
    d.sinr = m.tables + m.o_sinr; d.hgt = m.tables + m.o_hgt; d.nue_f = m.tables + m.o_nuef;
    d.step_dim = m.tables + m.o_step; d.geo = m.tables + m.o_geo; d.passes = m.passes; d.counters = m.counters; d.scratch = m.scratch;
    d.radiation = m.radiation; d.epsilon = m.epsilon; d.albedo = m.albedo;
    for (int q = 0; q < NMP; q++) d.mp[q] = m.mp[q];
    for (int q = 0; q < NMC; q++) d.mc[q] = m.mc[q];
    d.evap = m.evap; d.conv_cos = m.conv_cos;
    d.conv_qs[0] = m.conv_qs[0]; d.conv_qs[1] = m.conv_qs[1];
    d.conv_tmp[0] = m.conv_tmp[0]; d.conv_tmp[1] = m.conv_tmp[1];
    d.conv_cape = m.conv_cape;
    d.conv_cape_rng = m.conv_cape ? (int*)(m.conv_cape + m.g.im) : nullptr;
    (void)c;
}

// Brings the host mirror of t_skin up to the descriptors' value where the step's radiation has
// moved it on the device.
athad_status pull_tskin(athad_ctx* c) {
    if (!c->tskin_dev) return ATHAD_OK;
    for (int q = 0; q < c->nmem; q++) {
        const char* src = (const char*)(c->d_md[0] + q) + offsetof(MemberDesc, K) + offsetof(Consts, t_skin);
        HIPCHK(c, hipMemcpyAsync(&c->mem[q].K.t_skin, src, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->tskin_dev = false;
    return ATHAD_OK;
}

athad_status upload_descs(athad_ctx* c) {
    athad_status s0 = pull_tskin(c);
    if (s0) return s0;
    for (int p = 0; p < 2; p++) {
        c->h_md[p].resize(c->nmem);
        for (int q = 0; q < c->nmem; q++) make_desc(c, c->mem[q], p, c->h_md[p][q]);
        HIPCHK(c, hipMemcpyAsync(c->d_md[p], c->h_md[p].data(), sizeof(MemberDesc) * c->nmem,
                                 hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ATHAD_OK;
}

void free_ctx(athad_ctx* c) {
    if (!c) return;
    if (c->graph) hipGraphExecDestroy(c->graph);
    for (auto& m : c->mem) {
        for (auto p : m.field) if (p) hipFree(p);
        for (auto p : m.alt) if (p) hipFree(p);
        for (auto p : m.tmp) if (p) hipFree(p);
        for (auto p : m.mp) if (p) hipFree(p);
        for (auto p : m.mc) if (p) hipFree(p);
        if (m.evap) hipFree(m.evap);
        if (m.conv_cos) hipFree(m.conv_cos);
        for (auto p : m.conv_qs) if (p) hipFree(p);
        for (auto p : m.conv_tmp) if (p) hipFree(p);
        if (m.conv_cape) hipFree(m.conv_cape);
        for (void* p : {(void*)m.vel_star, (void*)m.t_top, (void*)m.t_ref, (void*)m.tables, (void*)m.scratch,
                        (void*)m.counters, (void*)m.passes, (void*)m.radiation, (void*)m.epsilon, (void*)m.albedo})
            if (p) hipFree(p);
    }
    for (int p = 0; p < 2; p++) if (c->d_md[p]) hipFree(c->d_md[p]);
    for (int k = 0; k < KID_COUNT; k++)
        for (auto& e : c->prof_ev[k]) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->ev_ffork) hipEventDestroy(c->ev_ffork);
    if (c->ev_fjoin) hipEventDestroy(c->ev_fjoin);
    if (c->stream2) hipStreamDestroy(c->stream2);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

Batch& cur(athad_ctx* c) {
    c->batch.d_md = c->d_md[c->parity];
    return c->batch;
}

// Kernel timing bracket.
struct Prof {
    athad_ctx* c; int kid; hipStream_t s; hipEvent_t e1 = nullptr, e2 = nullptr;
    Prof(athad_ctx* cc, int k, hipStream_t st = nullptr) : c(cc), kid(k), s(st ? st : cc->stream) {
        if (!c->capturing && c->prof_on[kid]) {
            hipEventCreate(&e1); hipEventCreate(&e2);
            hipEventRecord(e1, s);
        }
    }
    ~Prof() {
        if (e1) { hipEventRecord(e2, s); c->prof_ev[kid].push_back({e1, e2}); }
    }
};

athad_status check_member(athad_ctx* c, int member, int field, int nf) {
    if (!c) return ATHAD_ERR_ARG;
    if (member < 0 || member >= c->nmem) { c->err = "member index out of range"; return ATHAD_ERR_ARG; }
    if (field < 0 || field >= nf) { c->err = "field id out of range"; return ATHAD_ERR_ARG; }
    return ATHAD_OK;
}

// Device buffer currently holding a field of a member (Jacobi parity aware).
void* field_ptr(athad_ctx* c, int member, int f) {
    MemberHost& m = c->mem[member];
    if (f < NPROG && c->parity) return m.alt[f];
    return m.field[f];
}

athad_status after_launch(athad_ctx* c) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { c->err = std::string("kernel launch: ") + hipGetErrorString(e); return ATHAD_ERR_DEVICE; }
    return ATHAD_OK;
}

// ---- the operators (all members, no cadence gate unless given) -----------------------
const int kMoistSrc[4] = {F_T, F_ICE, F_C, F_CLOUD};

// damp_wiggles of t, ice, c, cloud plus the microphysics clamp (the standalone operator):
// the fields are staged in scratch 6..9 and damped from there back into the fields.
void op_damp_moist(athad_ctx* c, int gate) {
    Prof p(c, KID_DAMP);
    Batch& b = cur(c);
    const int tmp[4] = {NFIELD + 6, NFIELD + 7, NFIELD + 8, NFIELD + 9};
    launch_copy_job(b, kMoistSrc, tmp, 4);
    launch_damp_tile(b, c->tile_shape, c->tile_ich, tmp, kMoistSrc, 4, 1, gate);
}

// The moist damping inside the moist block: ice, c and cloud at every level, t at the lowest
// level only. The densities step that follows recomputes t at every level above from the
// damped surface value, so t damped above the surface would be discarded.
void op_damp_moist_in_block(athad_ctx* c, int gate) {
    Prof p(c, KID_DAMP);
    Batch& b = cur(c);
    const int src[3] = {NFIELD + 7, NFIELD + 8, NFIELD + 9};
    const int dst[3] = {F_ICE, F_C, F_CLOUD};
    launch_damp_tile(b, c->tile_shape, c->tile_ich, src, dst, 3, 1, gate);
    launch_damp_level0(b, NFIELD + 6, F_T, gate);
}

// TurbulenceAtm::run and the damping of its six fields, on stream st (the context's own
// stream when null).
void op_turbulence(athad_ctx* c, int gate, hipStream_t st = nullptr) {
    Batch b = cur(c);
    if (st) b.stream = st;
    {
        Prof p(c, KID_TURB, st);
        launch_vel_star(b, gate);
        launch_turb_sources(b, gate);
    }
    Prof p(c, KID_DAMP, st);
    const int t0[6] = {NFIELD + 0, NFIELD + 1, NFIELD + 2, NFIELD + 3, NFIELD + 4, NFIELD + 5};
    const int dst[6] = {F_TKE, F_DIS, F_NUE, F_PROD, F_TKE_SRC, F_DIS_SRC};
    launch_damp_tile(b, c->tile_shape, c->tile_ich, t0, dst, 6, 0, gate);
}

// The moist block of run_3D_loop (cAtmosphereModel.cpp:1185-1433). The turbulence update
// reads u, v, w, tke, dis and writes only the six turbulence fields and its scratch 0..5;
// SatAdj, the moist damping and densities read and write the thermodynamic fields and
// scratch 6..9. The two sets are disjoint, so their order is immaterial and the turbulence
// runs on the second stream beside the others.
void op_damp_conv(athad_ctx* c, int gate);

void op_moist_block(athad_ctx* c, int gate, bool poisson_beside = false, bool store_fluxes = false) {
    auto fork_s2 = [&]() {
        hipEventRecord(c->ev_fork, c->stream);
        hipStreamWaitEvent(c->stream2, c->ev_fork, 0);
        if (poisson_beside) {
            // PressureSolverAtm reads aux_u/v/w and p_dyn and writes p_dyn: none of the moist
            // block's fields; after the fork it also sees this iteration's counters
            Batch b2 = cur(c);
            b2.stream = c->stream2;
            Prof p(c, KID_POISSON, c->stream2);
            launch_poisson(b2, 0);
        }
        op_turbulence(c, gate, c->stream2);
    };
    if (c->ice_on) {
        // SatAdj (S_c_c left uncapped for the ice scheme), the damping of t ice c cloud at every
        // level (the scheme reads t throughout), the ice scheme and its flux damping, then
        // precipitationSum, cap_S and the microphysics clamp (cAtmosphereModel.cpp:1218-1399).
        // The second stream's work (turbulence, the Poisson sweep) starts beside the ice-scheme
        // columns, one wave per SIMD, rather than beside SatAdj, which fills the device.
        {
            Prof p(c, KID_SATADJ);
            launch_satadj(cur(c), gate, 1 | 2);
        }
        {
            Prof p(c, KID_DAMP);
            const int src[4] = {NFIELD + 6, NFIELD + 7, NFIELD + 8, NFIELD + 9};
            launch_damp_tile(cur(c), c->tile_shape, c->tile_ich, src, kMoistSrc, 4, 0, gate);
        }
        fork_s2();
        {
            Prof p(c, KID_ICE);
            launch_ice_scheme(cur(c), gate);
        }
        if (c->conv_on) {
            Prof p(c, KID_CONV);
            launch_moist_convection(cur(c), gate);
            op_damp_conv(c, gate);
        }
        {
            Prof p(c, KID_ICE);
            launch_moist_clamp_ice(cur(c), gate, store_fluxes ? 1 : 0);
        }
    } else {
        // the second stream's work starts after SatAdj, beside the damping and the
        // one-wave-per-SIMD densities (0.921 -> 0.904 ms per iteration against a fork before
        // SatAdj; a fork after the damping, 0.917, leaves the join waiting)
        {
            Prof p(c, KID_SATADJ);
            launch_satadj(cur(c), gate, 1);
        }
        fork_s2();
        op_damp_moist_in_block(c, gate);
    }
    {
        Prof p(c, KID_DENSITIES);
        launch_densities(cur(c), 0, 1, gate);
    }
    hipEventRecord(c->ev_join, c->stream2);
    hipStreamWaitEvent(c->stream, c->ev_join, 0);
    c->s2_pending = false;
}

// Makes the context's stream wait for work left on the second stream.
void join_s2(athad_ctx* c) {
    if (!c->s2_pending) return;
    hipEventRecord(c->ev_fjoin, c->stream2);
    hipStreamWaitEvent(c->stream, c->ev_fjoin, 0);
    c->s2_pending = false;
}

// Chooses the damp tile shape and levels per block for this context's grids by timing
// every candidate on scratch buffers (results do not depend on the choice). Cached per
// grid shape, member count and precision.
athad_status tune_damp(athad_ctx* c) {
    struct Key { int prec, n, im, jm, km; bool operator<(const Key& o) const {
        return std::tie(prec, n, im, jm, km) < std::tie(o.prec, o.n, o.im, o.jm, o.km); } };
    static std::mutex mu;
    static std::map<Key, std::pair<int, int>> cache;
    Batch& b = cur(c);
    const Key key{c->precision, c->nmem, b.max_im, b.max_jm, b.max_km};
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) { c->tile_shape = it->second.first; c->tile_ich = it->second.second; return ATHAD_OK; }
    }
    for (auto& m : c->mem)
        for (int q = 0; q < NTMP; q++) HIPCHK(c, hipMemsetAsync(m.tmp[q], 0, m.n * elem_size(c), c->stream));
    const int src[6] = {NFIELD + 0, NFIELD + 1, NFIELD + 2, NFIELD + 3, NFIELD + 4, NFIELD + 5};
    const int dst[6] = {NFIELD + 6, NFIELD + 7, NFIELD + 8, NFIELD + 9, NFIELD + 10, NFIELD + 11};
    const int ichs[3] = {1 << 20, 16, 8};
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    float best = 1e30f;
    int bs = 0, bi = ichs[0];
    for (int shape = 0; shape < damp_tile_shapes(); shape++) {
        for (int ich : ichs) {
            if (ich < (1 << 20) && ich >= b.max_im) continue;
            launch_damp_tile(b, shape, ich, src, dst, 6, 0, 0);   // warm
            HIPCHK(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < 3; r++) launch_damp_tile(b, shape, ich, src, dst, 6, 0, 0);
            HIPCHK(c, hipEventRecord(e1, c->stream));
            HIPCHK(c, hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; bs = shape; bi = ich; }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c->tile_shape = bs;
    c->tile_ich = bi;
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = {bs, bi};
    return ATHAD_OK;
}

void op_bc(athad_ctx* c, bool solid) {
    {
        Prof p(c, KID_BC);
        launch_bc(cur(c));
        if (c->mem[0].mp[0]) launch_bc_sources(cur(c));
        if (c->mem[0].mc[0]) launch_bc_conv(cur(c));
    }
    if (solid) {
        Prof p(c, KID_SOLID);
        launch_solid_clamp(cur(c));
    }
}

// PressureSolverAtm::project_initial_velocity(n_sweeps) (PressureSolverAtm.h:686-767): the
// velocity as the Poisson source, n sweeps of the solver (the same launches as the step's
// sweep), the gradient correction, p_dyn and aux cleared.
void op_project(athad_ctx* c, int n_sweeps) {
    Prof p(c, KID_POISSON);
    launch_project(cur(c), 0);
    for (int q = 0; q < n_sweeps; q++) launch_poisson(cur(c), 0);
    launch_project(cur(c), 1);
    launch_project(cur(c), 2);
}

// The source terms RK4 adds: the ice scheme's once its arrays exist (bit 0), moist convection's
// tendencies once its arrays exist (bit 1); both are zero in the reference until those operators run.
int rk4_sources(const athad_ctx* c) {
    return (c->mem[0].mp[0] ? 1 : 0) | (c->mem[0].mc[0] ? 2 : 0);
}

// The damp_wiggles of eleven convection arrays after MoistConvection::run
// (cAtmosphereModel.cpp:1249-1262): FP64 arrays in both storage precisions, damped by the fused
// tile kernel in FP64 into four FP64 temporaries and copied back, four arrays at a time.
void op_damp_conv(athad_ctx* c, int gate) {
    Batch b = cur(c);
    b.fp32 = false;
    constexpr int M0 = NFIELD + NTMP + NMP, T0 = M0 + NMC;
    const int groups[3][4] = {{MC_P_CONV, MC_E_U, MC_M_U, MC_Q_V_U},
                              {MC_Q_C_U, MC_E_D, MC_M_D, MC_Q_V_D},
                              {MC_MC_T, MC_MC_V, MC_MC_W, -1}};
    for (auto& g : groups) {
        const int nf = (g[3] < 0) ? 3 : 4;
        int src[4], dst[4];
        for (int q = 0; q < nf; q++) { src[q] = M0 + g[q]; dst[q] = T0 + q; }
        launch_damp_tile(b, c->tile_shape, c->tile_ich, src, dst, nf, 0, gate);
        launch_copy_job(b, dst, src, nf, gate);
    }
}

int max_plane(const athad_ctx* c) {
    int d = 0;
    for (auto& m : c->mem) d = std::max(d, (m.g.im - 2) + (m.g.jm - 2) + (m.g.km - 2));
    return d;
}

// RK4 in the context's order. fused: the Jacobi kernel applies the solid-ground clamps on
// read and fuses the store of the unfiltered fields (only valid inside a full step).
void op_rk4(athad_ctx* c, bool fused) {
    {
        Prof p(c, KID_LEVEL_MEAN);
        launch_level_mean(cur(c));
    }
    if (c->order == ATHAD_ORDER_LEX) {
        Prof p(c, KID_RK4_PLANE);
        const int dmax = max_plane(c);
        const int src = rk4_sources(c);
        for (int d = 3; d <= dmax; d++) launch_rk4_lex(cur(c), d, src);
    } else {
        {
            Prof p(c, KID_RK4);
            launch_rk4_jacobi(cur(c), fused, rk4_sources(c));
        }
        c->parity ^= 1;
    }
}

// polar_zonal_filter and the radial Shapiro filters of u, v, w (+ their store), on stream st
// (the context's own when null).
void op_filters(athad_ctx* c, int store, hipStream_t st = nullptr) {
    Batch b = cur(c);
    if (st) b.stream = st;
    {
        Prof p(c, KID_POLAR, st);
        launch_polar(b);
    }
    Prof p(c, KID_RADIAL, st);
    launch_radial(b, store);
}

// One scoped iteration for all members. When every member is at the same phase the
// cadence is decided on the host; otherwise the gated kernels decide per member.
void op_step(athad_ctx* c) {
    launch_advance(cur(c));
    bool same = true;
    for (auto& m : c->mem) same = same && ((m.iter_n + 1) % 4 == (c->mem[0].iter_n + 1) % 4);
    const int it = c->mem[0].iter_n + 1;
    // With the ice scheme in every member's moist block in step, the fluxes change only there,
    // so their store rides on the clamp after the scheme and other iterations need none (the
    // store would copy equal values); otherwise every iteration stores them.
    const bool mp = c->mem[0].mp[0] != nullptr;
    const bool fused_store = mp && c->ice_on && same && !c->mp_dirty;
    if (same) {
        // every 4th iteration is also a moist iteration: the sweep runs beside it
        if (it % 2 == 0) op_moist_block(c, 0, it % 4 == 0, fused_store);
    } else {
        { Prof p(c, KID_POISSON); launch_poisson(cur(c), 4); }
        op_moist_block(c, 2);
    }
    const bool lex = (c->order == ATHAD_ORDER_LEX);
    join_s2(c);   // the previous iteration's filters, if no moist block joined them already
    op_bc(c, lex);
    op_rk4(c, !lex);
    if (!lex) {
        // The filters read and write only u, v, w and un, vn, wn. The next iteration's
        // Poisson sweep (aux, p_dyn) and moist thermodynamics (t, c, cloud, ice, gr, co2,
        // p_stat, r_humid, r_dry) touch none of them, and its turbulence update, which reads
        // u, v, w, is queued behind them on the same stream; the boundaries join first.
        hipEventRecord(c->ev_ffork, c->stream);
        hipStreamWaitEvent(c->stream2, c->ev_ffork, 0);
        op_filters(c, 1, c->stream2);
        c->s2_pending = true;
    } else {
        op_filters(c, 1);
    }
    if (lex) {
        Prof p(c, KID_STORE);
        launch_store(cur(c), 0x7FF & ~0xE);   // all but u, v, w (stored by the radial filter)
    }
    if (mp && !fused_store) {   // P_rainn, P_snown (UtilsAtm.h:41-42)
        launch_ice_store(cur(c));
        c->mp_dirty = false;
    }
    if (c->rad_on) {
        // radiation_mode 2 after the store (cAtmosphereModel.cpp:1655-1661): touches t and the
        // radiation outputs, none of what the filters on the second stream read or write
        launch_radiation(cur(c));
        launch_skin_step(cur(c), c->d_md[0], c->d_md[1], Config().t_skin_relax);
        c->tskin_dev = true;
    }
    for (auto& m : c->mem) { m.iter_n++; m.total++; }
}

}  // namespace

// =======================================================================================
// C-ABI
// =======================================================================================
extern "C" {

const char* athad_version(void) { return "athad_hip 0.1 (gfx950)"; }

int64_t athad_footprint_bytes(const athad_grid* g, int precision) {
    if (!g) return -1;
    const int64_t n = (int64_t)g->im * g->jm * g->km;
    const int64_t es = (precision == ATHAD_MIXED) ? 4 : 8;
    // fields, Jacobi buffers and scratch; p_stat in FP64; vel_star and t_top; reductions; the 1-D
    // tables and the per-(i, j) metric table
    const int64_t tables = 4LL * g->im + 4LL * g->jm + 2 + (int64_t)g->im * g->jm * NGEO;
    return n * (es * (NFIELD - 1 + NPROG + NTMP) + 8) + (int64_t)g->jm * g->km * 16 +
           scratch_doubles(g->im, g->jm, g->km) * 8 + tables * 8;
}

athad_status athad_create(int n_members, const athad_grid* grids, const athad_member_params* params,
                          int precision, int order, athad_ctx** out) {
    g_create_error.clear();
    if (!out || !grids || n_members <= 0) { g_create_error = "bad arguments"; return ATHAD_ERR_ARG; }
    *out = nullptr;
    if (precision != ATHAD_FP64 && precision != ATHAD_MIXED) { g_create_error = "bad precision"; return ATHAD_ERR_ARG; }
    if (order != ATHAD_ORDER_LEX && order != ATHAD_ORDER_JACOBI && order != ATHAD_ORDER_SYMMETRIC) {
        g_create_error = "bad order";
        return ATHAD_ERR_ARG;
    }
    for (int q = 0; q < n_members; q++) {
        const athad_grid& g = grids[q];
        if (g.im < 6 || g.jm < 6 || g.km < 6 || g.im > 1024 || g.jm > 8192 || g.km > 16384 || (g.jm % 2) == 0 ||
            (long long)g.im * g.jm * g.km >= (1LL << 28)) {
            // 2^28 cells: the RK4 kernel addresses a field, and its stencil margin, by 32-bit byte offsets
            g_create_error = "unsupported grid (need im,jm,km >= 6, jm odd, fewer than 2^28 cells)";
            return ATHAD_ERR_SHAPE;
        }
    }
    athad_ctx* c = new (std::nothrow) athad_ctx();
    if (!c) return ATHAD_ERR_ALLOC;
    c->nmem = n_members;
    c->precision = precision;
    c->order = order;
    c->mem.resize(n_members);
    Config cfg;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_ffork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fjoin, hipEventDisableTiming);
    if (e != hipSuccess) { g_create_error = hipGetErrorString(e); free_ctx(c); return ATHAD_ERR_DEVICE; }
    for (int q = 0; q < n_members; q++) {
        MemberHost& m = c->mem[q];
        m.g = grids[q];
        m.p = params ? params[q] : default_member();
        m.n = (size_t)m.g.im * m.g.jm * m.g.km;
        fill_consts(m, cfg, order, kTurbKOmegaSST);
        std::vector<double> tab;
        std::vector<int> passes;
        build_tables(m, cfg, tab, passes);
        auto alloc = [&](void** p, size_t bytes) {
            if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; return false; }
            return true;
        };
        bool ok = true;
        for (int f = 0; f < NFIELD && ok; f++) ok = alloc(&m.field[f], field_bytes(c, m, f));
        for (int f = 0; f < NPROG && ok; f++) ok = alloc(&m.alt[f], m.n * elem_size(c));
        for (int f = 0; f < NTMP && ok; f++) ok = alloc(&m.tmp[f], m.n * elem_size(c));
        const size_t n2 = (size_t)m.g.jm * m.g.km;
        ok = ok && alloc((void**)&m.vel_star, n2 * 8) && alloc((void**)&m.t_top, n2 * 8) &&
             alloc((void**)&m.t_ref, m.g.im * 8) && alloc((void**)&m.tables, tab.size() * 8) &&
             alloc((void**)&m.scratch, (size_t)scratch_doubles(m.g.im, m.g.jm, m.g.km) * 8) && alloc((void**)&m.counters, 64) &&
             alloc((void**)&m.passes, passes.size() * 4);
        if (!ok) { g_create_error = "device allocation failed"; free_ctx(c); return ATHAD_ERR_ALLOC; }
        for (int f = 0; f < NFIELD; f++) hipMemsetAsync(m.field[f], 0, field_bytes(c, m, f), c->stream);
        for (int f = 0; f < NPROG; f++) hipMemsetAsync(m.alt[f], 0, m.n * elem_size(c), c->stream);
        hipMemsetAsync(m.vel_star, 0, n2 * 8, c->stream);
        hipMemsetAsync(m.t_top, 0, n2 * 8, c->stream);
        hipMemsetAsync(m.counters, 0, 64, c->stream);
        hipMemcpyAsync(m.tables, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, c->stream);
        hipMemcpyAsync(m.passes, passes.data(), passes.size() * 4, hipMemcpyHostToDevice, c->stream);
        hipStreamSynchronize(c->stream);
        c->batch.max_n = std::max(c->batch.max_n, (long long)m.n);
        c->batch.max_im = std::max(c->batch.max_im, m.g.im);
        c->batch.max_jm = std::max(c->batch.max_jm, m.g.jm);
        c->batch.max_km = std::max(c->batch.max_km, m.g.km);
    }
    for (int p = 0; p < 2; p++) {
        if (hipMalloc((void**)&c->d_md[p], sizeof(MemberDesc) * n_members) != hipSuccess) {
            g_create_error = "device allocation failed"; free_ctx(c); return ATHAD_ERR_ALLOC;
        }
    }
    c->batch.nmem = n_members;
    c->batch.stream = c->stream;
    c->batch.fp32 = (precision == ATHAD_MIXED);
    if (upload_descs(c) != ATHAD_OK) { g_create_error = c->err; free_ctx(c); return ATHAD_ERR_DEVICE; }
    if (tune_damp(c) != ATHAD_OK) { g_create_error = c->err; free_ctx(c); return ATHAD_ERR_DEVICE; }
    *out = c;
    return ATHAD_OK;
}

athad_status athad_destroy(athad_ctx* ctx) {
    if (!ctx) return ATHAD_ERR_ARG;
    hipStreamSynchronize(ctx->stream);
    free_ctx(ctx);
    return ATHAD_OK;
}

const char* athad_last_error(const athad_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

double athad_dew_point(double e_hPa) {
    if (!(e_hPa > 0.0)) return 0.0;
    if (e_hPa >= kPcrit) return kTcrit;
    double lo = 150.0, hi = kTcrit;
    if (th_p_auto(lo) >= e_hPa) return lo;
    for (int it = 0; it < 60; it++) {
        const double mid = 0.5 * (lo + hi);
        if (th_p_auto(mid) < e_hPa) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

athad_status athad_member_from_surface(double p_surf_bar, double t_sea, double x_h2o_dry, double x_co2_dry,
                                       athad_member_params* out) {
    if (!out || !(p_surf_bar > 0.0) || !(t_sea > 0.0) || !(x_h2o_dry >= 0.0) || !(x_co2_dry >= 0.0) ||
        x_h2o_dry + x_co2_dry > 1.0)
        return ATHAD_ERR_ARG;
    const double e_sea = th_p_sat(t_sea);                     // [hPa]
    const double p_hpa = p_surf_bar * 1000.0;
    if (!(e_sea > 0.0) || e_sea >= p_hpa) return ATHAD_ERR_ARG;   // sea would boil
    const double x_n2_dry = 1.0 - x_h2o_dry - x_co2_dry;
    const double x_h2o = e_sea / p_hpa;                        // saturated over the sea
    const double rest = 1.0 - x_h2o, dry_non_h2o = x_co2_dry + x_n2_dry;
    const double x_co2 = (dry_non_h2o > 0.0) ? rest * x_co2_dry / dry_non_h2o : 0.0;
    const double x_n2 = (dry_non_h2o > 0.0) ? rest * x_n2_dry / dry_non_h2o : 0.0;
    const Comp S = resolve(x_h2o, x_co2, x_n2);
    const Comp D = resolve(x_h2o_dry, x_co2_dry, x_n2_dry);
    out->t_surf_equator = t_sea;
    out->t_surf_pole = t_sea - 10.0;
    out->r_air = p_surf_bar * 1.0e5 / (S.R_mix * t_sea);
    out->x_H2O = x_h2o; out->x_CO2 = x_co2; out->x_N2 = x_n2;
    out->c_0 = S.q_H2O;
    out->co2_0 = S.q_CO2;
    out->c_h2o_dry_top = D.q_H2O;
    return ATHAD_OK;
}

athad_status athad_query(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) return ATHAD_OK;
    if (e == hipErrorNotReady) return ATHAD_NOT_READY;
    c->err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
    return ATHAD_ERR_DEVICE;
}

athad_status athad_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return ATHAD_ERR_DEVICE;
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return ATHAD_OK;
}

athad_status athad_field_stats(athad_ctx* c, int member, int field, double* sum, double* sumsq) {
    athad_status s = check_member(c, member, field, NFIELD);
    if (s) return s;
    launch_field_stats(cur(c), member, field);
    double r[2] = {0.0, 0.0};
    const MemberHost& m = c->mem[member];
    HIPCHK(c, hipMemcpyAsync(r, m.scratch + field_stats_offset(m.g.im), sizeof(r), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (sum) *sum = r[0];
    if (sumsq) *sumsq = r[1];
    return ATHAD_OK;
}

athad_status athad_field_minmax(athad_ctx* c, int member, int field, double* vmin, int* pos_min, double* vmax,
                                int* pos_max) {
    athad_status s = check_member(c, member, field, NFIELD);
    if (s) return s;
    join_s2(c);
    launch_minmax(cur(c), member, field);
    double r[4];
    const MemberHost& m = c->mem[member];
    HIPCHK(c, hipMemcpyAsync(r, m.scratch + diag_offset(m.g.im) + 4 * 256, sizeof(r), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    auto ijk = [&](double q, int* pos) {
        if (!pos) return;
        const long long iq = (q < 0.0) ? 0 : (long long)q;   // none finite: (0, 0, 0) as the reference
        pos[0] = (int)(iq / ((long long)m.g.jm * m.g.km));
        pos[1] = (int)((iq / m.g.km) % m.g.jm);
        pos[2] = (int)(iq % m.g.km);
    };
    if (vmin) *vmin = r[0];
    if (vmax) *vmax = r[2];
    ijk(r[1], pos_min);
    ijk(r[3], pos_max);
    return ATHAD_OK;
}

athad_status athad_residuum(athad_ctx* c, int member, double* res_max, int* pos_max, double* avg_abs,
                            int64_t* n_finite, int64_t* n_nonfinite, int* pos_first_nonfinite) {
    athad_status s = check_member(c, member, 0, 1);
    if (s) return s;
    const MemberHost& m = c->mem[member];
    if (m.g.im < 32) { c->err = "findResiduumAtm scans levels 10..30 and needs im >= 32"; return ATHAD_ERR_SHAPE; }
    join_s2(c);
    launch_residuum(cur(c), member);
    double r[6];
    HIPCHK(c, hipMemcpyAsync(r, m.scratch + diag_offset(m.g.im) + 6 * 256, sizeof(r), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    auto ijk = [&](double q, int* pos) {
        if (!pos) return;
        if (q < 0.0) { pos[0] = pos[1] = pos[2] = -1; return; }
        const long long iq = (long long)q;
        pos[0] = (int)(iq / ((long long)m.g.jm * m.g.km));
        pos[1] = (int)((iq / m.g.km) % m.g.jm);
        pos[2] = (int)(iq % m.g.km);
    };
    if (res_max) *res_max = r[0];
    if (pos_max) {   // no residual above 0: (0, 0, 0), the reference's starting position
        if (r[1] < 0.0) pos_max[0] = pos_max[1] = pos_max[2] = 0;
        else ijk(r[1], pos_max);
    }
    if (avg_abs) *avg_abs = (r[3] > 0.0) ? r[2] / r[3] : 0.0;
    if (n_finite) *n_finite = (int64_t)r[3];
    if (n_nonfinite) *n_nonfinite = (int64_t)r[4];
    ijk(r[5], pos_first_nonfinite);
    return ATHAD_OK;
}

athad_status athad_meridional_streamfunction(athad_ctx* c, int member, double* vbar, double* rho_vbar, double* psi) {
    athad_status s = check_member(c, member, 0, 1);
    if (s) return s;
    const MemberHost& m = c->mem[member];
    const Config cfg;
    join_s2(c);
    launch_streamfunction(cur(c), member, m.g.im, m.g.jm, 2.0 * M_PI * (cfg.r_Earth * 1000.0));
    const size_t nij = (size_t)m.g.im * m.g.jm;
    double* outs[3] = {vbar, rho_vbar, psi};
    for (int a = 0; a < 3; a++)
        if (outs[a])
            HIPCHK(c, hipMemcpyAsync(outs[a], m.scratch + sf_offset(m.g.im) + a * nij, nij * 8,
                                     hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ATHAD_OK;
}

athad_status athad_get_slice(athad_ctx* c, int member, int field, int axis, int index, double* dst) {
    athad_status s = check_member(c, member, field, NFIELD);
    if (s) return s;
    if (!dst) return ATHAD_ERR_ARG;
    const MemberHost& m = c->mem[member];
    const int extent[3] = {m.g.im, m.g.jm, m.g.km};
    if (axis < 0 || axis > 2) { c->err = "slice axis must be 0 (i), 1 (j) or 2 (k)"; return ATHAD_ERR_ARG; }
    if (index < 0 || index >= extent[axis]) { c->err = "slice index outside the grid"; return ATHAD_ERR_ARG; }
    const long long n = (long long)m.g.im * m.g.jm * m.g.km / extent[axis];
    join_s2(c);
    launch_slice(cur(c), member, field, axis, index, n);
    HIPCHK(c, hipMemcpyAsync(dst, m.scratch + sf_offset(m.g.im), (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ATHAD_OK;
}

// Radiation outputs are allocated at the first radiation call, so contexts that never call it
// hold no extra memory.
athad_status ensure_radiation_buffers(athad_ctx* c) {
    bool fresh = false;
    for (auto& m : c->mem) {
        if (m.radiation && m.epsilon && m.albedo) continue;
        const size_t bytes[3] = {m.n * 8, m.n * 8, (size_t)m.g.jm * m.g.km * 8};
        double** ptr[3] = {&m.radiation, &m.epsilon, &m.albedo};
        for (int a = 0; a < 3; a++)
            if (!*ptr[a] && hipMalloc((void**)ptr[a], bytes[a]) != hipSuccess) {
                *ptr[a] = nullptr;
                c->err = "out of device memory for the radiation outputs";
                return ATHAD_ERR_ALLOC;
            }
        fresh = true;
    }
    return fresh ? upload_descs(c) : ATHAD_OK;
}

athad_status athad_radiation(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    for (auto& m : c->mem)
        if (m.g.im > radiation_max_levels()) {
            c->err = "radiation supports up to " + std::to_string(radiation_max_levels()) + " levels";
            return ATHAD_ERR_SHAPE;
        }
    athad_status s = ensure_radiation_buffers(c);
    if (s) return s;
    if ((s = pull_tskin(c))) return s;
    join_s2(c);
    launch_radiation(cur(c));
    // updateSkinTemperature (cAtmosphereModel.cpp:276-311) with planetaryShortWave (:241-270)
    // on the zonal albedo sums; host arithmetic on jm numbers, as the reference's.
    const Config cfg;
    bool changed = false;
    for (auto& m : c->mem) {
        const int jm = m.g.jm, km = m.g.km, jmax = jm - 1, jhalf = jmax / 2;
        std::vector<double> rows(jm);
        HIPCHK(c, hipMemcpyAsync(rows.data(), m.scratch + sf_offset(m.g.im), jm * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (cfg.t_skin_relax <= 0.0) continue;
        double wsum = 0.0, absorbed = 0.0, amax = 0.0;
        for (int j = 0; j < jm; j++) {
            const double x = (double)((j <= jhalf) ? j : jmax - j) / (double)jhalf;
            const double sw = (cfg.rad_pole_short - cfg.rad_equator_short) * (x * x - 2.0 * x) + cfg.rad_pole_short;
            const double w = std::cos((j / (double)(jm - 1) - 0.5) * M_PI);
            const double a = rows[j] / (double)km;
            amax = std::max(amax, a);
            wsum += w;
            absorbed += w * (1.0 - a) * sw;
        }
        if (!(wsum > 0.0) || !(amax > 0.0)) continue;
        const double in = absorbed / wsum + cfg.geothermal_flux;
        if (!(in > 0.0)) continue;
        const double target = std::pow(in / cfg.sigma, 0.25);
        m.K.t_skin += std::min(1.0, cfg.t_skin_relax) * (target - m.K.t_skin);
        changed = true;
    }
    return changed ? upload_descs(c) : ATHAD_OK;
}

athad_status athad_set_radiation(athad_ctx* c, int on) {
    if (!c) return ATHAD_ERR_ARG;
    if (on) {
        for (auto& m : c->mem)
            if (m.g.im > radiation_max_levels()) {
                c->err = "radiation supports up to " + std::to_string(radiation_max_levels()) + " levels";
                return ATHAD_ERR_SHAPE;
            }
        athad_status s = ensure_radiation_buffers(c);
        if (s) return s;
    }
    if (c->rad_on != (on != 0) && c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    c->rad_on = on != 0;
    return ATHAD_OK;
}

athad_status athad_get_radiation(athad_ctx* c, int member, double* radiation, double* epsilon, double* albedo,
                                 double* t_skin) {
    athad_status s = check_member(c, member, 0, 1);
    if (s) return s;
    if ((s = pull_tskin(c))) return s;
    const MemberHost& m = c->mem[member];
    if (t_skin) *t_skin = m.K.t_skin;
    if (!m.radiation && (radiation || epsilon || albedo)) { c->err = "radiation has not run on this context"; return ATHAD_ERR_STATE; }
    join_s2(c);
    if (radiation) HIPCHK(c, hipMemcpyAsync(radiation, m.radiation, m.n * 8, hipMemcpyDeviceToHost, c->stream));
    if (epsilon) HIPCHK(c, hipMemcpyAsync(epsilon, m.epsilon, m.n * 8, hipMemcpyDeviceToHost, c->stream));
    if (albedo) HIPCHK(c, hipMemcpyAsync(albedo, m.albedo, (size_t)m.g.jm * m.g.km * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ATHAD_OK;
}

// The fluxes changed outside a step: the next iteration stores them in full, and a captured
// graph that stores them only after the scheme is dropped.
static void mark_mp_dirty(athad_ctx* c) {
    c->mp_dirty = true;
    if (c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
}

// The ice-scheme arrays are allocated, zero, at the first call that needs them, for every
// member at once; a 4-iteration graph captured before then did not store the fluxes, so it is
// dropped and recaptured.

// Publishes the members' pointers in the device descriptors after on-demand arrays were freed
// again: kernels test those pointers, so none may keep a freed one.
static void republish(athad_ctx* c) {
    const std::string e = c->err;
    upload_descs(c);
    c->err = e;
}

static athad_status ensure_mp_buffers(athad_ctx* c) {
    if (c->mp_ready) return ATHAD_OK;
    if (c->capturing) { c->err = "the ice-scheme arrays cannot be allocated during capture"; return ATHAD_ERR_STATE; }
    // every failure frees every member's arrays, so a non-null mp[0] always means complete
    auto fail = [&](athad_status st, const char* msg) {
        for (auto& m2 : c->mem)
            for (auto& p : m2.mp) if (p) { hipFree(p); p = nullptr; }
        c->err = msg;
        republish(c);
        return st;
    };
    for (auto& m : c->mem)
        for (int q = 0; q < NMP; q++) {
            if (hipMalloc(&m.mp[q], m.n * elem_size(c)) != hipSuccess) {
                m.mp[q] = nullptr;
                return fail(ATHAD_ERR_ALLOC, "out of device memory for the ice-scheme arrays");
            }
            if (hipMemset(m.mp[q], 0, m.n * elem_size(c)) != hipSuccess)
                return fail(ATHAD_ERR_DEVICE, "clearing the ice-scheme arrays failed");
        }
    if (c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    if (upload_descs(c) != ATHAD_OK) return fail(ATHAD_ERR_DEVICE, "publishing the ice-scheme arrays failed");
    c->mp_ready = true;
    return ATHAD_OK;
}

athad_status athad_ice_scheme(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    athad_status s = ensure_mp_buffers(c);
    if (s) return s;
    join_s2(c);
    { Prof p(c, KID_ICE); launch_ice_scheme(cur(c), 0); }
    mark_mp_dirty(c);
    return after_launch(c);
}

athad_status athad_set_ice_scheme(athad_ctx* c, int on) {
    if (!c) return ATHAD_ERR_ARG;
    if (on) {
        athad_status s = ensure_mp_buffers(c);
        if (s) return s;
    }
    if ((c->ice_on != (on != 0) || (!on && c->conv_on)) && c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    c->ice_on = on != 0;
    if (!on) c->conv_on = false;   // convection runs after the ice scheme in the moist block
    return ATHAD_OK;
}

athad_status athad_set_mp_field(athad_ctx* ctx, int member, int which, const double* src) {
    athad_status s = check_member(ctx, member, which, NMP);
    if (s) return s;
    if (!src) return ATHAD_ERR_ARG;
    if ((s = ensure_mp_buffers(ctx))) return s;
    join_s2(ctx);
    mark_mp_dirty(ctx);
    MemberHost& m = ctx->mem[member];
    if (ctx->precision == ATHAD_FP64) {
        HIPCHK(ctx, hipMemcpyAsync(m.mp[which], src, m.n * 8, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> tmp(m.n);
        for (size_t q = 0; q < m.n; q++) tmp[q] = (float)src[q];
        HIPCHK(ctx, hipMemcpyAsync(m.mp[which], tmp.data(), m.n * 4, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return ATHAD_OK;
}

athad_status athad_get_mp_field(athad_ctx* ctx, int member, int which, double* dst) {
    athad_status s = check_member(ctx, member, which, NMP);
    if (s) return s;
    if (!dst) return ATHAD_ERR_ARG;
    join_s2(ctx);
    MemberHost& m = ctx->mem[member];
    if (!m.mp[which]) {   // never allocated: the arrays are zero, as resetArrays leaves them
        std::fill(dst, dst + m.n, 0.0);
        return ATHAD_OK;
    }
    if (ctx->precision == ATHAD_FP64) {
        HIPCHK(ctx, hipMemcpyAsync(dst, m.mp[which], m.n * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> tmp(m.n);
        HIPCHK(ctx, hipMemcpyAsync(tmp.data(), m.mp[which], m.n * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t q = 0; q < m.n; q++) dst[q] = (double)tmp[q];
    }
    return ATHAD_OK;
}

// ---- moist convection (MoistConvection.h:106-140) ---------------------------------------
// The 32 convection arrays, the evaporation input and the latitude factor table are allocated,
// zero, at the first call that needs them, for every member at once.
static athad_status ensure_mc_buffers(athad_ctx* c) {
    if (c->mc_ready) return ATHAD_OK;
    if (c->capturing) { c->err = "the convection arrays cannot be allocated during capture"; return ATHAD_ERR_STATE; }
    // every failure frees every member's arrays, so a non-null mc[0] always means complete
    auto fail = [&](athad_status st, const char* msg) {
        for (auto& m2 : c->mem) {
            for (auto& p : m2.mc) if (p) { hipFree(p); p = nullptr; }
            if (m2.evap) { hipFree(m2.evap); m2.evap = nullptr; }
            if (m2.conv_cos) { hipFree(m2.conv_cos); m2.conv_cos = nullptr; }
            for (auto& p : m2.conv_qs) if (p) { hipFree(p); p = nullptr; }
            for (auto& p : m2.conv_tmp) if (p) { hipFree(p); p = nullptr; }
            if (m2.conv_cape) { hipFree(m2.conv_cape); m2.conv_cape = nullptr; }
        }
        c->err = msg;
        republish(c);
        return st;
    };
    const char* oom = "out of device memory for the convection arrays";
    for (auto& m : c->mem) {
        const int jm = m.g.jm, km = m.g.km;
        for (int q = 0; q < NMC; q++) {
            if (hipMalloc(&m.mc[q], m.n * 8) != hipSuccess) { m.mc[q] = nullptr; return fail(ATHAD_ERR_ALLOC, oom); }
            if (hipMemset(m.mc[q], 0, m.n * 8) != hipSuccess)
                return fail(ATHAD_ERR_DEVICE, "clearing the convection arrays failed");
        }
        if (hipMalloc(&m.evap, (size_t)jm * km * 8) != hipSuccess) { m.evap = nullptr; return fail(ATHAD_ERR_ALLOC, oom); }
        if (hipMemset(m.evap, 0, (size_t)jm * km * 8) != hipSuccess)
            return fail(ATHAD_ERR_DEVICE, "clearing the evaporation input failed");
        if (hipMalloc(&m.conv_cos, (size_t)jm * 8) != hipSuccess) { m.conv_cos = nullptr; return fail(ATHAD_ERR_ALLOC, oom); }
        for (auto& p : m.conv_qs)
            if (hipMalloc(&p, m.n * 8) != hipSuccess) { p = nullptr; return fail(ATHAD_ERR_ALLOC, oom); }
        for (auto& p : m.conv_tmp)
            if (hipMalloc(&p, m.n * 8) != hipSuccess) { p = nullptr; return fail(ATHAD_ERR_ALLOC, oom); }
        const size_t cape_bytes = (size_t)m.g.im * 8 + (size_t)jm * km * 8;
        if (hipMalloc(&m.conv_cape, cape_bytes) != hipSuccess) {
            m.conv_cape = nullptr;
            return fail(ATHAD_ERR_ALLOC, oom);
        }
        if (hipMemset(m.conv_cape, 0, cape_bytes) != hipSuccess)
            return fail(ATHAD_ERR_DEVICE, "clearing the CAPE diagnostic failed");
        // theta by accumulation as the grid builds it; the reference takes the circle factor
        // about row 90 (MoistConvection.h:239), which is the equator on its 181-row grid. Here
        // it is taken about the equator row (jm - 1) / 2 of every grid: the same table at
        // jm = 181, and on coarser grids a factor that stays positive instead of crossing zero
        // at row 45 and turning negative beyond it.
        std::vector<double> the(jm), cs(jm);
        the[0] = 0.0;
        for (int l = 1; l < jm; l++) the[l] = the[l - 1] + m.K.dthe;
        const int jeq = (jm - 1) / 2;
        for (int j = 0; j < jm; j++) cs[j] = std::cos(the[jeq] - the[j]);
        if (hipMemcpy(m.conv_cos, cs.data(), (size_t)jm * 8, hipMemcpyHostToDevice) != hipSuccess)
            return fail(ATHAD_ERR_DEVICE, "uploading the convection circle factors failed");
    }
    if (c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    if (upload_descs(c) != ATHAD_OK) return fail(ATHAD_ERR_DEVICE, "publishing the convection arrays failed");
    c->mc_ready = true;
    return ATHAD_OK;
}

static athad_status check_conv_grid(athad_ctx* c) {
    for (auto& m : c->mem)
        if (m.g.jm < 3 || m.g.km < 3 || m.g.im < 3) {
            c->err = "moist convection needs jm >= 3, km >= 3, im >= 3";
            return ATHAD_ERR_ARG;
        }
    return ATHAD_OK;
}

athad_status athad_moist_convection(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    athad_status s = check_conv_grid(c);
    if (s || (s = ensure_mc_buffers(c))) return s;
    join_s2(c);
    { Prof p(c, KID_CONV); launch_moist_convection(cur(c)); }
    return after_launch(c);
}

athad_status athad_set_moist_convection(athad_ctx* c, int on) {
    if (!c) return ATHAD_ERR_ARG;
    if (on) {
        athad_status s = check_conv_grid(c);
        if (s || (s = ensure_mp_buffers(c)) || (s = ensure_mc_buffers(c))) return s;
    }
    const bool ice = on ? true : c->ice_on;
    if ((c->conv_on != (on != 0) || c->ice_on != ice) && c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    c->conv_on = on != 0;
    c->ice_on = ice;
    return ATHAD_OK;
}

athad_status athad_set_conv_field(athad_ctx* ctx, int member, int which, const double* src) {
    athad_status s = check_member(ctx, member, which, NMC + 1);
    if (s) return s;
    if (!src) return ATHAD_ERR_ARG;
    if ((s = check_conv_grid(ctx)) || (s = ensure_mc_buffers(ctx))) return s;
    join_s2(ctx);
    MemberHost& m = ctx->mem[member];
    double* dst = (which == NMC) ? m.evap : m.mc[which];
    const size_t cnt = (which == NMC) ? (size_t)m.g.jm * m.g.km : m.n;
    HIPCHK(ctx, hipMemcpyAsync(dst, src, cnt * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ATHAD_OK;
}

athad_status athad_get_conv_field(athad_ctx* ctx, int member, int which, double* dst) {
    athad_status s = check_member(ctx, member, which, NMC + 1);
    if (s) return s;
    if (!dst) return ATHAD_ERR_ARG;
    join_s2(ctx);
    MemberHost& m = ctx->mem[member];
    const size_t cnt = (which == NMC) ? (size_t)m.g.jm * m.g.km : m.n;
    const double* src = (which == NMC) ? m.evap : m.mc[which];
    if (!src) {   // never allocated: zero, as resetArrays leaves them
        std::fill(dst, dst + cnt, 0.0);
        return ATHAD_OK;
    }
    HIPCHK(ctx, hipMemcpyAsync(dst, src, cnt * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ATHAD_OK;
}

athad_status athad_get_conv_cape(athad_ctx* ctx, int member, double* dst) {
    if (!ctx || !dst || member < 0 || member >= ctx->nmem) return ATHAD_ERR_ARG;
    join_s2(ctx);
    MemberHost& m = ctx->mem[member];
    if (!m.conv_cape) {   // convection never ran: the reference's CAPE is empty, reported as zeros
        std::fill(dst, dst + m.g.im, 0.0);
        return ATHAD_OK;
    }
    HIPCHK(ctx, hipMemcpyAsync(dst, m.conv_cape, (size_t)m.g.im * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ATHAD_OK;
}

athad_status athad_set_member_params(athad_ctx* c, int member, const athad_member_params* p) {
    if (!c || !p || member < 0 || member >= c->nmem) return ATHAD_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    MemberHost& m = c->mem[member];
    m.p = *p;
    const Config cfg;
    c->tskin_dev = false;   // the member's constants, t_skin included, are reset
    fill_consts(m, cfg, c->order, c->turb_model);
    std::vector<double> tab;
    std::vector<int> passes;
    build_tables(m, cfg, tab, passes);
    HIPCHK(c, hipMemcpy(m.tables, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m.passes, passes.data(), passes.size() * 4, hipMemcpyHostToDevice));
    return upload_descs(c);
}

athad_status athad_damp_tile(athad_ctx* c, int* tj, int* tk, int* levels) {
    if (!c) return ATHAD_ERR_ARG;
    int a = 0, b = 0;
    damp_tile_shape(c->tile_shape, &a, &b);
    if (tj) *tj = a;
    if (tk) *tk = b;
    if (levels) *levels = std::min(c->tile_ich, c->batch.max_im);
    return ATHAD_OK;
}

athad_status athad_set_damp_tile(athad_ctx* c, int shape, int levels) {
    if (!c || shape < 0 || shape >= damp_tile_shapes() || levels < 1) return ATHAD_ERR_ARG;
    c->tile_shape = shape;
    c->tile_ich = levels;
    if (c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    return ATHAD_OK;
}

athad_status athad_sync(athad_ctx* ctx) {
    if (!ctx) return ATHAD_ERR_ARG;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ATHAD_OK;
}

athad_status athad_set_iter(athad_ctx* ctx, int member, int iter_n, int total) {
    athad_status s = check_member(ctx, member, 0, 1);
    if (s) return s;
    ctx->mem[member].iter_n = iter_n;
    ctx->mem[member].total = total;
    int h[2] = {iter_n, total};
    HIPCHK(ctx, hipMemcpyAsync(ctx->mem[member].counters, h, 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ATHAD_OK;
}

athad_status athad_get_iter(athad_ctx* ctx, int member, int* iter_n, int* total) {
    athad_status s = check_member(ctx, member, 0, 1);
    if (s) return s;
    int h[2];
    HIPCHK(ctx, hipMemcpyAsync(h, ctx->mem[member].counters, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (iter_n) *iter_n = h[0];
    if (total) *total = h[1];
    return ATHAD_OK;
}

athad_status athad_set_field(athad_ctx* ctx, int member, int field, const double* src) {
    athad_status s = check_member(ctx, member, field, NFIELD);
    if (s) return s;
    if (!src) return ATHAD_ERR_ARG;
    MemberHost& m = ctx->mem[member];
    void* dst = field_ptr(ctx, member, field);
    if (field == F_P_STAT || ctx->precision == ATHAD_FP64) {
        HIPCHK(ctx, hipMemcpyAsync(dst, src, m.n * 8, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> tmp(m.n);
        for (size_t q = 0; q < m.n; q++) tmp[q] = (float)src[q];
        HIPCHK(ctx, hipMemcpyAsync(dst, tmp.data(), m.n * 4, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return ATHAD_OK;
}

athad_status athad_get_field(athad_ctx* ctx, int member, int field, double* dst) {
    athad_status s = check_member(ctx, member, field, NFIELD);
    if (s) return s;
    if (!dst) return ATHAD_ERR_ARG;
    MemberHost& m = ctx->mem[member];
    const void* src = field_ptr(ctx, member, field);
    if (field == F_P_STAT || ctx->precision == ATHAD_FP64) {
        HIPCHK(ctx, hipMemcpyAsync(dst, src, m.n * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> tmp(m.n);
        HIPCHK(ctx, hipMemcpyAsync(tmp.data(), src, m.n * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t q = 0; q < m.n; q++) dst[q] = (double)tmp[q];
    }
    return ATHAD_OK;
}

athad_status athad_set_field2d(athad_ctx* ctx, int member, int field, const double* src) {
    athad_status s = check_member(ctx, member, field, ATHAD_NFIELD2D);
    if (s) return s;
    if (!src) return ATHAD_ERR_ARG;
    MemberHost& m = ctx->mem[member];
    double* dst = (field == ATHAD_VEL_STAR) ? m.vel_star : m.t_top;
    HIPCHK(ctx, hipMemcpyAsync(dst, src, (size_t)m.g.jm * m.g.km * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ATHAD_OK;
}

athad_status athad_get_field2d(athad_ctx* ctx, int member, int field, double* dst) {
    athad_status s = check_member(ctx, member, field, ATHAD_NFIELD2D);
    if (s) return s;
    if (!dst) return ATHAD_ERR_ARG;
    MemberHost& m = ctx->mem[member];
    const double* src = (field == ATHAD_VEL_STAR) ? m.vel_star : m.t_top;
    HIPCHK(ctx, hipMemcpyAsync(dst, src, (size_t)m.g.jm * m.g.km * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ATHAD_OK;
}

athad_status athad_upload_rows(athad_ctx* ctx, int member, int field, double* const* rows) {
    athad_status s = check_member(ctx, member, field, NFIELD);
    if (s) return s;
    if (!rows) return ATHAD_ERR_ARG;
    MemberHost& m = ctx->mem[member];
    std::vector<double> flat(m.n);
    const int nrow = m.g.im * m.g.jm;
    for (int r = 0; r < nrow; r++) {
        if (!rows[r]) { ctx->err = "null row pointer"; return ATHAD_ERR_ARG; }
        std::memcpy(&flat[(size_t)r * m.g.km], rows[r], (size_t)m.g.km * 8);
    }
    return athad_set_field(ctx, member, field, flat.data());
}

athad_status athad_download_rows(athad_ctx* ctx, int member, int field, double* const* rows) {
    athad_status s = check_member(ctx, member, field, NFIELD);
    if (s) return s;
    if (!rows) return ATHAD_ERR_ARG;
    MemberHost& m = ctx->mem[member];
    std::vector<double> flat(m.n);
    s = athad_get_field(ctx, member, field, flat.data());
    if (s) return s;
    const int nrow = m.g.im * m.g.jm;
    for (int r = 0; r < nrow; r++) {
        if (!rows[r]) { ctx->err = "null row pointer"; return ATHAD_ERR_ARG; }
        std::memcpy(rows[r], &flat[(size_t)r * m.g.km], (size_t)m.g.km * 8);
    }
    return ATHAD_OK;
}

// ---- restart files (FileIO_Atm.cpp:293-356) --------------------------------------------
namespace {
constexpr int32_t kRestartMagic = 0x41544D31;   // "ATM1"
constexpr int kRestartArrays = 29;
// restart_arrays() (FileIO_Atm.cpp:297-304) as field ids; -1 for P_rain, P_snow, P_rainn,
// P_snown, which the scoped step keeps at zero and the context does not hold
constexpr int kRestartF[kRestartArrays] = {
    F_T, F_U, F_V, F_W, F_C, F_CLOUD, F_ICE, F_GR, F_CO2,
    F_TN, F_UN, F_VN, F_WN, F_CN, F_CLOUDN, F_ICEN, F_GRN, F_CO2N,
    F_P_DYN, F_P_STAT, F_TKE, F_DIS, F_TKEN, F_DISN, F_NUE,
    -1 - MP_P_RAIN, -1 - MP_P_SNOW, -1 - MP_P_RAINN, -1 - MP_P_SNOWN};   // negative: ice-scheme array
int64_t restart_bytes(const athad_grid& g) {
    return 5 * 4 + (int64_t)kRestartArrays * g.im * g.jm * g.km * 8;
}
}  // namespace

int64_t athad_restart_bytes(const athad_grid* g) {
    if (!g || g->im < 1 || g->jm < 1 || g->km < 1) return -1;
    return restart_bytes(*g);
}

athad_status athad_save_state_mem(athad_ctx* ctx, int member, void* buf, int64_t size) {
    athad_status s = check_member(ctx, member, 0, 1);
    if (s) return s;
    MemberHost& m = ctx->mem[member];
    if (!buf || size < restart_bytes(m.g)) { ctx->err = "restart: buffer too small"; return ATHAD_ERR_ARG; }
    join_s2(ctx);
    int iter_n = 0, total = 0;
    if ((s = athad_get_iter(ctx, member, &iter_n, &total))) return s;
    const int32_t hdr[5] = {kRestartMagic, m.g.im, m.g.jm, m.g.km, total};
    unsigned char* out = (unsigned char*)buf;
    std::memcpy(out, hdr, sizeof(hdr));
    // the arrays start 20 bytes in, so they are not 8-byte aligned: each goes through an
    // aligned staging array and a byte copy
    std::vector<double> arr(m.n);
    for (int a = 0; a < kRestartArrays; a++) {
        // the precipitation arrays read zero where the ice scheme has never run
        if (kRestartF[a] < 0) s = athad_get_mp_field(ctx, member, -1 - kRestartF[a], arr.data());
        else s = athad_get_field(ctx, member, kRestartF[a], arr.data());
        if (s) return s;
        std::memcpy(out + sizeof(hdr) + (size_t)a * m.n * 8, arr.data(), m.n * 8);
    }
    return ATHAD_OK;
}

athad_status athad_load_state_mem(athad_ctx* ctx, int member, const void* buf, int64_t size) {
    athad_status s = check_member(ctx, member, 0, 1);
    if (s) return s;
    MemberHost& m = ctx->mem[member];
    int32_t hdr[5];
    if (!buf || size < (int64_t)sizeof(hdr)) { ctx->err = "restart: no header"; return ATHAD_ERR_ARG; }
    std::memcpy(hdr, buf, sizeof(hdr));
    if (hdr[0] != kRestartMagic || hdr[1] != m.g.im || hdr[2] != m.g.jm || hdr[3] != m.g.km) {
        ctx->err = "restart: bad header or grid mismatch";
        return ATHAD_ERR_SHAPE;
    }
    if (size < restart_bytes(m.g)) { ctx->err = "restart: truncated"; return ATHAD_ERR_ARG; }
    join_s2(ctx);
    // the arrays start 20 bytes in (not 8-byte aligned): byte-copied into an aligned array
    const unsigned char* in = (const unsigned char*)buf + sizeof(hdr);
    std::vector<double> arr(m.n);
    for (int a = 0; a < kRestartArrays; a++) {
        std::memcpy(arr.data(), in + (size_t)a * m.n * 8, m.n * 8);
        if (kRestartF[a] < 0) {
            // precipitation arrays: held once the ice scheme has run; a file holding zeros
            // (as the scoped step leaves them) allocates nothing
            if (!m.mp[0] && std::all_of(arr.begin(), arr.end(), [](double v) { return v == 0.0; })) continue;
            s = athad_set_mp_field(ctx, member, -1 - kRestartF[a], arr.data());
        } else s = athad_set_field(ctx, member, kRestartF[a], arr.data());
        if (s) return s;
    }
    return athad_set_iter(ctx, member, m.iter_n, hdr[4]);
}

athad_status athad_save_state(athad_ctx* ctx, int member, const char* path) {
    athad_status s = check_member(ctx, member, 0, 1);
    if (s) return s;
    if (!path) return ATHAD_ERR_ARG;
    std::vector<unsigned char> buf((size_t)restart_bytes(ctx->mem[member].g));
    if ((s = athad_save_state_mem(ctx, member, buf.data(), (int64_t)buf.size()))) return s;
    FILE* f = std::fopen(path, "wb");
    if (!f) { ctx->err = std::string("restart: cannot open ") + path; return ATHAD_ERR_ARG; }
    const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
    const int closed = std::fclose(f);
    if (w != buf.size() || closed != 0) { ctx->err = std::string("restart: write failed ") + path; return ATHAD_ERR_ARG; }
    return ATHAD_OK;
}

athad_status athad_load_state(athad_ctx* ctx, int member, const char* path) {
    athad_status s = check_member(ctx, member, 0, 1);
    if (s) return s;
    if (!path) return ATHAD_ERR_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) { ctx->err = std::string("restart: no file ") + path; return ATHAD_ERR_ARG; }
    std::vector<unsigned char> buf((size_t)restart_bytes(ctx->mem[member].g));
    const size_t r = std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return athad_load_state_mem(ctx, member, buf.data(), (int64_t)r);
}

athad_status athad_init_state(athad_ctx* c, uint64_t seed) {
    if (!c) return ATHAD_ERR_ARG;
    // member seeds; a new integration starts from the configured t_skin
    c->tskin_dev = false;   // superseded by the reset below
    for (int q = 0; q < c->nmem; q++) {
        c->mem[q].K.seed = seed + (uint64_t)q;
        c->mem[q].K.t_skin = Config().t_skin;
    }
    c->parity = 0;
    athad_status s = upload_descs(c);
    if (s) return s;
    Batch& b = cur(c);
    // resetArrays: all fields zero except t (1) — the init kernels below overwrite every
    // field they read; the rest are cleared here.
    for (auto& m : c->mem) {
        for (int f = 0; f < NFIELD; f++) HIPCHK(c, hipMemsetAsync(m.field[f], 0, field_bytes(c, m, f), c->stream));
        HIPCHK(c, hipMemsetAsync(m.counters, 0, 64, c->stream));
        m.iter_n = 0; m.total = 0;
    }
    launch_init_column(b);
    {   // damp_wiggles(t) (cAtmosphereModel.cpp:528)
        const int src[1] = {F_T}, t0[1] = {NFIELD + 0}, t1[1] = {NFIELD + 1};
        launch_damp(b, 0, src, t0, 1, 0, 0);
        launch_damp(b, 1, t0, t1, 1, 0, 0);
        launch_damp(b, 2, t1, src, 1, 0, 0);
    }
    launch_copy_lid(b);
    launch_init_water(b);
    for (int pass = 0; pass < 2; pass++) {
        launch_init_co2(b);
        launch_densities(b, 1, 0, 0);
    }
    launch_init_rest(b);
    op_turbulence(c, 0);
    op_bc(c, true);
    op_project(c, 200);   // RunTimeSlice :665
    launch_copy_field(b, F_T_EQ, F_T);
    launch_store(b, 0x7FF);
    if (after_launch(c)) return ATHAD_ERR_DEVICE;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ATHAD_OK;
}

athad_status athad_project_initial_velocity(athad_ctx* c, int n_sweeps) {
    if (!c || n_sweeps < 0) return ATHAD_ERR_ARG;
    join_s2(c);
    op_project(c, n_sweeps);
    return after_launch(c);
}

athad_status athad_poisson_sweep(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    { Prof p(c, KID_POISSON); launch_poisson(cur(c), 0); }
    return after_launch(c);
}
athad_status athad_satadj(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    { Prof p(c, KID_SATADJ); launch_satadj(cur(c), 0, 0); }
    return after_launch(c);
}
athad_status athad_damp_moist(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    op_damp_moist(c, 0);
    return after_launch(c);
}
athad_status athad_densities(athad_ctx* c, int set_water, int with_evap) {
    if (!c) return ATHAD_ERR_ARG;
    { Prof p(c, KID_DENSITIES); launch_densities(cur(c), set_water ? 1 : 0, with_evap ? 1 : 0, 0); }
    return after_launch(c);
}
athad_status athad_turbulence(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    op_turbulence(c, 0);
    return after_launch(c);
}
athad_status athad_set_turbulence_model(athad_ctx* c, int model) {
    if (!c) return ATHAD_ERR_ARG;
    if (model != ATHAD_TURB_K_EPSILON && model != ATHAD_TURB_K_OMEGA && model != ATHAD_TURB_K_OMEGA_SST) {
        c->err = "turbulence model must be ATHAD_TURB_K_EPSILON, ATHAD_TURB_K_OMEGA or ATHAD_TURB_K_OMEGA_SST";
        return ATHAD_ERR_ARG;
    }
    join_s2(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const Config cfg;
    if (c->turb_model != model && c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    c->turb_model = model;
    c->batch.turb = model;
    for (auto& m : c->mem) {
        m.K.turb = model;
        m.K.om_wall = wall_dis(model, m.wall_y1, cfg);
    }
    return upload_descs(c);
}
athad_status athad_turbulence_init(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    join_s2(c);
    Batch& b = cur(c);
    {
        Prof p(c, KID_TURB);
        launch_vel_star(b, 0);
        launch_turb_init_fields(b);
        launch_wall_bc(b);
        launch_turb_sources(b, 0);
    }
    const int src[6] = {NFIELD + 0, NFIELD + 1, NFIELD + 2, NFIELD + 3, NFIELD + 4, NFIELD + 5};
    const int dst[6] = {F_TKE, F_DIS, F_NUE, F_PROD, F_TKE_SRC, F_DIS_SRC};
    launch_copy_job(b, src, dst, 6, 0);
    return after_launch(c);
}
athad_status athad_boundary(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    op_bc(c, true);
    return after_launch(c);
}
athad_status athad_rk4(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    op_rk4(c, false);
    return after_launch(c);
}
athad_status athad_filters(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    op_filters(c, 0);
    return after_launch(c);
}
athad_status athad_store(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    { Prof p(c, KID_STORE); launch_store(cur(c), 0x7FF); }
    if (c->mem[0].mp[0]) {
        launch_ice_store(cur(c));
        c->mp_dirty = false;
    }
    return after_launch(c);
}

athad_status athad_step(athad_ctx* c) {
    if (!c) return ATHAD_ERR_ARG;
    op_step(c);
    join_s2(c);
    return after_launch(c);
}

athad_status athad_run(athad_ctx* c, int n_steps, int use_graph) {
    if (!c || n_steps < 0) return ATHAD_ERR_ARG;
    if (!use_graph) {
        for (int q = 0; q < n_steps; q++) op_step(c);
        join_s2(c);
        return after_launch(c);
    }
    int done = 0;
    // align to the graph's phase with plain steps
    while (done < n_steps) {
        const int phase = c->mem[0].iter_n % 4;
        if (c->graph && phase == c->graph_phase && c->parity == c->graph_parity && n_steps - done >= 4) {
            join_s2(c);
            HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
            if (c->rad_on) c->tskin_dev = true;
            for (auto& m : c->mem) { m.iter_n += 4; m.total += 4; }
            done += 4;
            continue;
        }
        if (!c->graph && n_steps - done >= 4) {
            // capture four iterations
            const int ph = c->mem[0].iter_n % 4, par = c->parity;
            std::vector<MemberHost> saved = c->mem;
            hipGraph_t g;
            join_s2(c);
            c->capturing = true;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (int q = 0; q < 4; q++) op_step(c);
            join_s2(c);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            c->capturing = false;
            if (e != hipSuccess) { c->err = std::string("graph capture: ") + hipGetErrorString(e); return ATHAD_ERR_DEVICE; }
            e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) { c->err = std::string("graph instantiate: ") + hipGetErrorString(e); return ATHAD_ERR_DEVICE; }
            // capture recorded but did not execute the launches: restore the host mirror
            for (int q = 0; q < c->nmem; q++) { c->mem[q].iter_n = saved[q].iter_n; c->mem[q].total = saved[q].total; }
            c->parity = par;
            c->graph_phase = ph;
            c->graph_parity = par;
            continue;
        }
        op_step(c);
        done++;
    }
    join_s2(c);
    return after_launch(c);
}

athad_status athad_profile(athad_ctx* c, int kid, int enable) {
    if (!c || kid < 0 || kid >= KID_COUNT) return ATHAD_ERR_ARG;
    c->prof_on[kid] = enable != 0;
    if (enable) {
        for (auto& e : c->prof_ev[kid]) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        c->prof_ev[kid].clear();
    }
    return ATHAD_OK;
}

athad_status athad_profile_result(athad_ctx* c, int kid, double* total_ms, int* launches) {
    if (!c || kid < 0 || kid >= KID_COUNT) return ATHAD_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (auto& e : c->prof_ev[kid]) {
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = (int)c->prof_ev[kid].size();
    return ATHAD_OK;
}

}  // extern "C"

// ---- min/max of a caller's field without allocation ------------------------------------
// No counterpart in the reference. For callers that cannot allocate device memory (code under
// hipGraph capture, tight serving loops): the reductions' scratch, and the one descriptor the
// shared kernels read, live in a fixed file-scope device array. Its device address is looked up
// once per device with hipGetSymbolAddress and used wherever host code needs the array.
namespace {
constexpr int kNoAllocMaxIm = 200;              // radial levels: bounds the reduction layout
constexpr size_t kNoAllocDesc = 4096;           // room for one MemberDesc
constexpr size_t kNoAllocDoubles = 16384;       // reduction partials and results
constexpr int kNoAllocMaxDevices = 64;
static_assert(sizeof(MemberDesc) <= kNoAllocDesc, "descriptor slot too small");
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocDesc + kNoAllocDoubles * sizeof(double)];

extern "C" athad_status athad_field_minmax_noalloc(const void* field, int precision, int im, int jm, int km,
                                                   void* stream, double* vmin, int* pos_min, double* vmax,
                                                   int* pos_max) {
    if (!field || im < 1 || jm < 1 || km < 1 || (precision != ATHAD_FP64 && precision != ATHAD_MIXED)) {
        g_create_error = "field_minmax_noalloc: bad arguments";
        return ATHAD_ERR_ARG;
    }
    if (im > kNoAllocMaxIm || (size_t)(diag_offset(im) + 4 * 256 + 4) > kNoAllocDoubles) {
        g_create_error = "field_minmax_noalloc: more than 200 radial levels exceeds its fixed scratch";
        return ATHAD_ERR_SHAPE;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) {
        g_create_error = "field_minmax_noalloc: no current device";
        return ATHAD_ERR_DEVICE;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (!g_noalloc_ptr[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                g_create_error = std::string("field_minmax_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e);
                return ATHAD_ERR_DEVICE;
            }
            g_noalloc_ptr[dev] = p;
        }
        base = g_noalloc_ptr[dev];
    }
    // the descriptor the min/max kernels read: the caller's field in the slot of t, and the
    // reduction scratch after the descriptor slot
    MemberDesc d;
    std::memset(&d, 0, sizeof(d));
    d.K.im = im; d.K.jm = jm; d.K.km = km;
    d.K.n = (long long)im * jm * km;
    d.f[F_T] = const_cast<void*>(field);
    d.scratch = (double*)((unsigned char*)base + kNoAllocDesc);
    hipStream_t st = (hipStream_t)stream;
    auto fail = [&](hipError_t e, const char* what) {
        g_create_error = std::string("field_minmax_noalloc: ") + what + ": " + hipGetErrorString(e);
        (void)hipStreamSynchronize(st);   // the descriptor copy reads `d` on this stack frame
        return ATHAD_ERR_DEVICE;
    };
    hipError_t e = hipMemcpyAsync(base, &d, sizeof(d), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return fail(e, "descriptor upload");
    Batch b;
    b.d_md = (MemberDesc*)base;
    b.nmem = 1;
    b.max_n = d.K.n;
    b.max_im = im; b.max_jm = jm; b.max_km = km;
    b.stream = st;
    b.fp32 = (precision == ATHAD_MIXED);
    launch_minmax(b, 0, F_T);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "launch");
    double r[4];
    e = hipMemcpyAsync(r, d.scratch + diag_offset(im) + 4 * 256, sizeof(r), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return fail(e, "result copy");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "synchronize");
    auto ijk = [&](double q, int* pos) {
        if (!pos) return;
        const long long iq = (q < 0.0) ? 0 : (long long)q;   // none finite: (0, 0, 0) as the reference
        pos[0] = (int)(iq / ((long long)jm * km));
        pos[1] = (int)((iq / km) % jm);
        pos[2] = (int)(iq % km);
    };
    if (vmin) *vmin = r[0];
    if (vmax) *vmax = r[2];
    ijk(r[1], pos_min);
    ijk(r[3], pos_max);
    return ATHAD_OK;
}
