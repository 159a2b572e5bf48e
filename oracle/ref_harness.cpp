// Drives the reference's own operators (compiled from /root/reference sources by
// oracle/Makefile.ref) on a state written by the oracle, and writes the result back, so the
// oracle restatement can be checked against the reference. TEST INFRASTRUCTURE ONLY.
//
// Access to cAtmosphereModel's private members goes through the friend declaration the
// reference already carries for its offline test driver (cAtmosphereModel.h:67).
//
// usage: ref_harness <in.bin> <out.bin> <mode> [n]
//   mode = steps            n scoped iterations (the IN rows of run_3D_loop)
//   mode = op:<name>        one operator: poisson satadj damp_moist moist_clamp densities
//                           evaporation turbulence damp_turb bc rk4 polar radial store
//                           project (project_initial_velocity(200), as RunTimeSlice calls it)
//   mode = init             the reference's thermodynamic init chain on a blank state
//   mode = radiation        run_3D_loop's mode-2 radiation once (MultiLayerRadiation, the
//                           omega_rad blend of t, updateSkinTemperature)
//   mode = steps_rad        n iterations of the scoped rows, each followed by that radiation
//   mode = ice              TwoCatIceScheme::run once, then run_3D_loop's damp_wiggles of P_rain
//                           and P_snow (cAtmosphereModel.cpp:1230-1244); reads and writes the
//                           ice-scheme arrays in <in.bin>.mp / <out.bin>.mp: P_rain, P_snow,
//                           P_rainn, P_snown, Precipitation, S_v, S_c, S_i, S_r, S_s
//   mode = steps_ice        n scoped iterations with the ice scheme in its place in the moist
//                           block (after the damping of t ice c cloud), its flux damping and
//                           precipitationSum (moist convection not run: P_conv stays zero);
//                           the .mp files as for mode ice
//   mode = steps_conv       n scoped iterations with the ice scheme and moist convection in the
//                           moist block (cAtmosphereModel.cpp:1230-1263); reads and writes the
//                           .mp, .mc and .evap files (ice-scheme arrays, convection arrays, the
//                           Evaporation diagnostic)
//   mode = conv             n calls of MoistConvection::run (its parameters as the model's
//                           defaults set them: iter_prec 3, convection_perturbation 1,
//                           convection_mode 1); reads <in.bin>.evap (the Evaporation diagnostic,
//                           jm*km doubles) and, if present, <in.bin>.mc, and writes <out.bin>.mc:
//                           the 32 convection arrays in mc_fields order (and <out.bin>.cape,
//                           computeCAPE's im values)
//   mode = turb:<model>     the turbulence closure <model> (k_epsilon, k_omega or k_omega_SST)
//                           selected as cAtmosphereModel.cpp:484-489 does, TurbulenceAtm::init,
//                           the state written to <out.bin>.init, then n scoped iterations
// The radiation modes also write <out.bin>.rad: radiation and epsilon (im*jm*km doubles each),
// albedo (jm*km), the top short wave (jm) and t_skin.
// State file: int32 im,jm,km,iter_n,total_iter; then 35 fields of im*jm*km doubles
// (athad_oracle.h order); then vel_star and t_top_init (jm*km doubles each).
#include "cAtmosphereModel.h"
#include "PressureSolverAtm.h"
#include "SaturationAdjustment.h"
#include "ThermoAtm.h"
#include "UtilsAtm.h"
#include "BC_Atm.h"
#include "TurbulenceAtm.h"
#include "MultiLayerRadiation.h"
#include "TwoCatIceScheme.h"
#include "MoistConvection.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

class RadiationSelfTest {
public:
    static std::vector<Array*> fields(cAtmosphereModel& m) {
        return {&m.t, &m.u, &m.v, &m.w, &m.c, &m.cloud, &m.ice, &m.gr, &m.co2, &m.tke, &m.dis,
                &m.tn, &m.un, &m.vn, &m.wn, &m.cn, &m.cloudn, &m.icen, &m.grn, &m.co2n, &m.tken, &m.disn,
                &m.p_dyn, &m.p_stat, &m.r_humid, &m.r_dry, &m.t_eq, &m.nue, &m.prod, &m.tke_source,
                &m.dis_source, &m.aux_u, &m.aux_v, &m.aux_w, &m.S_c_c};
    }

    // the closure selected as cAtmosphereModel.cpp:484-489 does, then TurbulenceAtm::init
    static void turb_init(cAtmosphereModel& m, const std::string& model) {
        m.turb_model = model;
        m.use_turbulence_model = true;
        m.use_k_epsilon_turbulence_model = (model == "k_epsilon");
        m.use_k_omega_turbulence_model = (model == "k_omega");
        m.use_k_omega_SST_turbulence_model = (model == "k_omega_SST");
        TurbulenceAtm(m).init();
    }

    static void setup(cAtmosphereModel& m) {
        m.initComposition();
        m.m_current_time = m.m_time_list.insert(0.0f).first;   // time slice Ma = 0
        UtilsAtm(m).resetArrays();
        m.dt = m.dt_visc;
        m.turb_model = "k_omega_SST";
        m.use_turbulence_model = true;
        m.use_k_epsilon_turbulence_model = false;
        m.use_k_omega_turbulence_model = false;
        m.use_k_omega_SST_turbulence_model = true;
        m.initGridCoordinates();
        m.initMetricRadius();
        m.init_layer_heights();
        m.re_turb = 0.4 * 0.1 / 1.5e-5;
        m.diffusion_ramp = 1.0;
        m.inviscid_phase = false;
        m.vbudget_capture = false;
        m.wbudget_capture = false;
        m.t_top_init.assign(m.jm, std::vector<double>(m.km, 0.0));
    }

    static bool load(cAtmosphereModel& m, const char* fn) {
        std::ifstream f(fn, std::ios::binary);
        int32_t hdr[5];
        f.read((char*)hdr, sizeof(hdr));
        if (!f || hdr[0] != m.im || hdr[1] != m.jm || hdr[2] != m.km) {
            std::fprintf(stderr, "state %s does not match the compiled grid %dx%dx%d\n", fn, m.im, m.jm, m.km);
            return false;
        }
        m.iter_n = hdr[3];
        m.total_iter_count = hdr[4];
        std::vector<double> buf((size_t)m.im * m.jm * m.km);
        for (Array* a : fields(m)) {
            f.read((char*)buf.data(), buf.size() * 8);
            size_t q = 0;
            for (int i = 0; i < m.im; i++)
                for (int j = 0; j < m.jm; j++)
                    for (int k = 0; k < m.km; k++) a->x[i][j][k] = buf[q++];
        }
        std::vector<double> b2((size_t)m.jm * m.km);
        f.read((char*)b2.data(), b2.size() * 8);
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) m.vel_star.y[j][k] = b2[(size_t)j * m.km + k];
        f.read((char*)b2.data(), b2.size() * 8);
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) m.t_top_init[j][k] = b2[(size_t)j * m.km + k];
        return (bool)f;
    }

    static void save(cAtmosphereModel& m, const char* fn) {
        std::ofstream f(fn, std::ios::binary);
        int32_t hdr[5] = {m.im, m.jm, m.km, m.iter_n, m.total_iter_count};
        f.write((const char*)hdr, sizeof(hdr));
        std::vector<double> buf((size_t)m.im * m.jm * m.km);
        for (Array* a : fields(m)) {
            size_t q = 0;
            for (int i = 0; i < m.im; i++)
                for (int j = 0; j < m.jm; j++)
                    for (int k = 0; k < m.km; k++) buf[q++] = a->x[i][j][k];
            f.write((const char*)buf.data(), buf.size() * 8);
        }
        std::vector<double> b2((size_t)m.jm * m.km);
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) b2[(size_t)j * m.km + k] = m.vel_star.y[j][k];
        f.write((const char*)b2.data(), b2.size() * 8);
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) b2[(size_t)j * m.km + k] = m.t_top_init[j][k];
        f.write((const char*)b2.data(), b2.size() * 8);
    }

    // The loop glue of run_3D_loop that is written inline there (cap_S, the microphysics
    // clamp and the water-vapour ceiling, cAtmosphereModel.cpp:1278-1399), repeated here
    // because a lambda inside run_3D_loop cannot be called from outside.
    static void moist_clamp(cAtmosphereModel& m) {
        constexpr double S_max = 1.0e-3;
        Array* caps[] = {&m.S_v, &m.S_c, &m.S_i, &m.S_r, &m.S_s, &m.S_g, &m.S_c_c};
        for (Array* S : caps)
            for (int i = 0; i < m.im; i++)
                for (int j = 0; j < m.jm; j++)
                    for (int k = 0; k < m.km; k++) {
                        double v = S->x[i][j][k];
                        if (!AtomUtils::is_finite_safe(v)) S->x[i][j][k] = 0.0;
                        else if (v > S_max) S->x[i][j][k] = S_max;
                        else if (v < -S_max) S->x[i][j][k] = -S_max;
                    }
        for (int i = 0; i < m.im; i++)
            for (int j = 0; j < m.jm; j++)
                for (int k = 0; k < m.km; k++) {
                    double cv = m.c.x[i][j][k], cld = m.cloud.x[i][j][k], ic = m.ice.x[i][j][k];
                    if (!AtomUtils::is_finite_safe(cv)) cv = 0.0;
                    if (!AtomUtils::is_finite_safe(cld)) cld = 0.0;
                    if (!AtomUtils::is_finite_safe(ic)) ic = 0.0;
                    if (cv < 0.0) cv = 0.0;
                    if (cld < 0.0) cld = 0.0;
                    if (ic < 0.0) ic = 0.0;
                    const double cv_max = std::max(0.0, 1.0 - m.co2.x[i][j][k]);
                    if (cv > cv_max) cv = cv_max;
                    const double sum = cld + ic;
                    if (sum > m.cloud_cap) { const double sc = m.cloud_cap / sum; cld *= sc; ic *= sc; }
                    m.c.x[i][j][k] = cv; m.cloud.x[i][j][k] = cld; m.ice.x[i][j][k] = ic;
                }
        for (int i = 0; i < m.im; i++)
            for (int j = 0; j < m.jm; j++)
                for (int k = 0; k < m.km; k++) {
                    if (m.c.x[i][j][k] < 0.0) m.c.x[i][j][k] = 0.0;
                    const double cc = std::max(0.0, 1.0 - m.co2.x[i][j][k]);
                    if (m.c.x[i][j][k] > cc) m.c.x[i][j][k] = cc;
                }
    }

    static void damp(cAtmosphereModel& m, Array& a) { AtomUtils::damp_wiggles(a, &m.i_topography, true, true, true); }

    static void op(cAtmosphereModel& m, const std::string& name) {
        m.buoyancy_ramp = std::min(1.0, (double)m.total_iter_count / 300.0);
        if (name == "poisson") PressureSolverAtm(m).run();
        else if (name == "satadj") SaturationAdjustment(m).run();
        else if (name == "project") PressureSolverAtm(m).project_initial_velocity(200);
        else if (name == "damp_moist") { damp(m, m.t); damp(m, m.ice); damp(m, m.c); damp(m, m.cloud); }
        else if (name == "moist_clamp") moist_clamp(m);
        else if (name == "densities") ThermoAtm(m).densities();
        else if (name == "evaporation") ThermoAtm(m).waterVapourEvaporation();
        else if (name == "turbulence") TurbulenceAtm(m).run();
        else if (name == "damp_turb") {
            damp(m, m.tke); damp(m, m.dis); damp(m, m.nue); damp(m, m.prod); damp(m, m.tke_source); damp(m, m.dis_source);
        } else if (name == "bc") {
            BC_Atm(m).bcRadius();
            TurbulenceAtm(m).apply_wall_bc();
            BC_Atm(m).bcTheta();
            BC_Atm(m).bcPhi();
            BC_Atm(m).bcScalarSurfSur();
            BC_Atm(m).bcSolidGround();
        } else if (name == "rk4") m.solveRungeKutta_Atmosphere_Turb();
        else if (name == "polar") {
            AtomUtils::polar_zonal_filter(m.u, m.the.z, &m.i_topography, 30.0, 12, 3.0);
            AtomUtils::polar_zonal_filter(m.v, m.the.z, &m.i_topography, 30.0, 12, 3.0);
            AtomUtils::polar_zonal_filter(m.w, m.the.z, &m.i_topography, 30.0, 12, 3.0);
        } else if (name == "radial") {
            AtomUtils::orographic_shapiro_filter(m.u, m.i_topography, 2, 3, 3);
            AtomUtils::orographic_shapiro_filter(m.v, m.i_topography, 2, 3, 3);
            AtomUtils::orographic_shapiro_filter(m.w, m.i_topography, 2, 3, 3);
            AtomUtils::radial_shapiro_filter(m.u, m.i_topography, 2, 1.0);
            AtomUtils::radial_shapiro_filter_ho(m.v, m.i_topography, 2, 1.0);
            AtomUtils::radial_shapiro_filter_ho(m.w, m.i_topography, 2, 1.0);
            AtomUtils::orographic_radial_shapiro_filter(m.t, m.i_topography, 2, 10, 2);
            AtomUtils::orographic_radial_shapiro_filter(m.c, m.i_topography, 2, 10, 2);
            AtomUtils::orographic_radial_shapiro_filter(m.cloud, m.i_topography, 2, 10, 2);
            AtomUtils::orographic_radial_shapiro_filter(m.ice, m.i_topography, 2, 10, 2);
            AtomUtils::steep_massif_field_smoothing(m.u, m.i_topography, 6, 3, 30, 1);
            AtomUtils::steep_massif_field_smoothing(m.v, m.i_topography, 6, 3, 30, 1);
            AtomUtils::steep_massif_field_smoothing(m.w, m.i_topography, 6, 3, 30, 1);
            AtomUtils::steep_massif_field_smoothing(m.p_dyn, m.i_topography, 6, 3, 30, 1);
            AtomUtils::coastal_velocity_sponge(m.u, m.i_topography, 3.0, 1.0, 12);
            AtomUtils::coastal_velocity_sponge(m.v, m.i_topography, 3.0, 1.0, 12);
            AtomUtils::coastal_velocity_sponge(m.w, m.i_topography, 3.0, 1.0, 12);
        } else if (name == "store") UtilsAtm(m).storeIntermediateData3D(1.0);
        else std::fprintf(stderr, "unknown op %s\n", name.c_str());
    }

    // One iteration of run_3D_loop restricted to the scoped rows (cAtmosphereModel.cpp:1103-1647).
    static void step(cAtmosphereModel& m, bool ice_scheme = false, bool convection = false) {
        m.iter_n++;
        m.total_iter_count++;
        m.diffusion_ramp = 1.0;
        m.dt = m.dt_visc;
        if (m.iter_n % 4 == 0) op(m, "poisson");
        if (m.iter_n % 2 == 0) {
            op(m, "satadj");
            op(m, "damp_moist");
            if (ice_scheme) {
                ice(m);
                if (convection) {
                    // cAtmosphereModel.cpp:1247-1262
                    MoistConvection(m).run(m.iter_n);
                    for (Array* a : {&m.P_conv, &m.E_u, &m.M_u, &m.q_v_u, &m.q_c_u, &m.E_d, &m.M_d, &m.q_v_d,
                                     &m.MC_t, &m.MC_v, &m.MC_w})
                        damp(m, *a);
                }
                UtilsAtm(m).precipitationSum();
            }
            op(m, "moist_clamp");
            op(m, "densities");
            op(m, "evaporation");
            op(m, "turbulence");
            op(m, "damp_turb");
        }
        op(m, "bc");
        op(m, "rk4");
        op(m, "polar");
        op(m, "radial");
        op(m, "store");
    }

    // apply_radiative_heating + updateSkinTemperature, the radiation_mode == 2 branch of
    // run_3D_loop (cAtmosphereModel.cpp:863-879, 1655-1661), repeated here because the blend is
    // a lambda inside run_3D_loop. omega_rad is its constexpr 0.05 (:801).
    static void radiation(cAtmosphereModel& m) {
        constexpr double omega_rad = 0.05;
        std::vector<double> tb((size_t)m.im * m.jm * m.km);
        size_t q = 0;
        for (int i = 0; i < m.im; i++)
            for (int j = 0; j < m.jm; j++)
                for (int k = 0; k < m.km; k++) tb[q++] = m.t.x[i][j][k];
        MultiLayerRadiation(m).run();
        q = 0;
        for (int i = 0; i < m.im; i++)
            for (int j = 0; j < m.jm; j++)
                for (int k = 0; k < m.km; k++) {
                    const double re = m.t.x[i][j][k];
                    m.t.x[i][j][k] = tb[q++] * (1.0 - omega_rad) + re * omega_rad;
                }
        m.updateSkinTemperature(false);
    }

    static void save_rad(cAtmosphereModel& m, const std::string& fn) {
        std::ofstream f(fn, std::ios::binary);
        std::vector<double> buf((size_t)m.im * m.jm * m.km);
        for (Array* a : {&m.radiation, &m.epsilon}) {
            size_t q = 0;
            for (int i = 0; i < m.im; i++)
                for (int j = 0; j < m.jm; j++)
                    for (int k = 0; k < m.km; k++) buf[q++] = a->x[i][j][k];
            f.write((const char*)buf.data(), buf.size() * 8);
        }
        std::vector<double> b2((size_t)m.jm * m.km);
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) b2[(size_t)j * m.km + k] = m.albedo.y[j][k];
        f.write((const char*)b2.data(), b2.size() * 8);
        f.write((const char*)m.short_wave_radiation.data(), m.jm * 8);
        f.write((const char*)&m.t_skin, 8);
    }

    static std::vector<Array*> mp_fields(cAtmosphereModel& m) {
        return {&m.P_rain, &m.P_snow, &m.P_rainn, &m.P_snown, &m.Precipitation,
                &m.S_v, &m.S_c, &m.S_i, &m.S_r, &m.S_s};
    }

    static bool load_mp(cAtmosphereModel& m, const std::string& fn) {
        std::ifstream f(fn, std::ios::binary);
        std::vector<double> buf((size_t)m.im * m.jm * m.km);
        for (Array* a : mp_fields(m)) {
            f.read((char*)buf.data(), buf.size() * 8);
            size_t q = 0;
            for (int i = 0; i < m.im; i++)
                for (int j = 0; j < m.jm; j++)
                    for (int k = 0; k < m.km; k++) a->x[i][j][k] = buf[q++];
        }
        return (bool)f;
    }

    static void save_mp(cAtmosphereModel& m, const std::string& fn) {
        std::ofstream f(fn, std::ios::binary);
        std::vector<double> buf((size_t)m.im * m.jm * m.km);
        for (Array* a : mp_fields(m)) {
            size_t q = 0;
            for (int i = 0; i < m.im; i++)
                for (int j = 0; j < m.jm; j++)
                    for (int k = 0; k < m.km; k++) buf[q++] = a->x[i][j][k];
            f.write((const char*)buf.data(), buf.size() * 8);
        }
    }

    static std::vector<Array*> mc_fields(cAtmosphereModel& m) {
        return {&m.s, &m.M_u, &m.M_d, &m.s_u, &m.q_v_u, &m.q_c_u, &m.c_u, &m.e_l, &m.e_d, &m.e_p,
                &m.g_p, &m.s_d, &m.q_v_d, &m.P_conv, &m.E_u, &m.D_u, &m.u_u, &m.v_u, &m.w_u,
                &m.E_d, &m.D_d, &m.u_d, &m.v_d, &m.w_d, &m.MC_t, &m.MC_q, &m.MC_v, &m.MC_w,
                &m.Deep_beg, &m.Deep_end, &m.CloudBase, &m.LevelFreeSinking};
    }

    static void io_fields(cAtmosphereModel& m, const std::vector<Array*>& arrs, const std::string& fn, bool rd) {
        std::vector<double> buf((size_t)m.im * m.jm * m.km);
        std::fstream f(fn, rd ? std::ios::in | std::ios::binary : std::ios::out | std::ios::binary);
        if (!f) return;
        for (Array* a : arrs) {
            if (rd) f.read((char*)buf.data(), buf.size() * 8);
            size_t q = 0;
            for (int i = 0; i < m.im; i++)
                for (int j = 0; j < m.jm; j++)
                    for (int k = 0; k < m.km; k++) {
                        if (rd) a->x[i][j][k] = buf[q++]; else buf[q++] = a->x[i][j][k];
                    }
            if (!rd) f.write((const char*)buf.data(), buf.size() * 8);
        }
    }

    static bool load_evap(cAtmosphereModel& m, const std::string& fn) {
        std::ifstream fe(fn, std::ios::binary);
        std::vector<double> b2((size_t)m.jm * m.km);
        fe.read((char*)b2.data(), b2.size() * 8);
        if (!fe) return false;
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) m.Evaporation.y[j][k] = b2[(size_t)j * m.km + k];
        return true;
    }
    static void save_evap(cAtmosphereModel& m, const std::string& fn) {
        std::ofstream fo(fn, std::ios::binary);
        std::vector<double> b2((size_t)m.jm * m.km);
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) b2[(size_t)j * m.km + k] = m.Evaporation.y[j][k];
        fo.write((const char*)b2.data(), b2.size() * 8);
    }

    static bool conv(cAtmosphereModel& m, const std::string& in, const std::string& out, int n) {
        std::ifstream fe(in + ".evap", std::ios::binary);
        std::vector<double> b2((size_t)m.jm * m.km);
        fe.read((char*)b2.data(), b2.size() * 8);
        if (!fe) return false;
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) m.Evaporation.y[j][k] = b2[(size_t)j * m.km + k];
        io_fields(m, mc_fields(m), in + ".mc", true);
        std::fprintf(stderr, "conv: iter_prec %d perturbation %d mode %d p_0 %.17g s_0 %.17g\n",
                     m.iter_prec, m.convection_perturbation, m.convection_mode, m.p_0, m.s_0);
        for (int q = 0; q < n; q++) MoistConvection(m).run(m.iter_n);
        io_fields(m, mc_fields(m), out + ".mc", false);
        std::ofstream fc(out + ".cape", std::ios::binary);   // computeCAPE's im values
        fc.write((const char*)m.CAPE.data(), m.CAPE.size() * 8);
        return true;
    }

    static void ice(cAtmosphereModel& m) {
        TwoCatIceScheme(m).run();
        damp(m, m.P_rain);
        damp(m, m.P_snow);
    }

    // RunTimeSlice's thermodynamic init chain, cAtmosphereModel.cpp:527-568 (without
    // initCloudIce; condensate stays zero as resetArrays left it).
    static void init_chain(cAtmosphereModel& m) {
        m.initTemperatureData(0);
        AtomUtils::damp_wiggles(m.t, &m.i_topography, true, true, true);
        for (int j = 0; j < m.jm; j++)
            for (int k = 0; k < m.km; k++) m.t_top_init[j][k] = m.t.x[m.im - 1][j][k];
        m.initWaterWapour();
        for (int pass = 0; pass < 2; pass++) {
            ThermoAtm(m).co2Atmosphere();
            ThermoAtm(m).densities(true);
        }
    }
};

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s in.bin out.bin mode [n]\n", argv[0]);
        return 2;
    }
    std::ofstream devnull("/dev/null");
    std::streambuf* old = std::cout.rdbuf(devnull.rdbuf());
    cAtmosphereModel* m = new cAtmosphereModel();
    RadiationSelfTest::setup(*m);
    const std::string mode = argv[3];
    if (mode != "init" && !RadiationSelfTest::load(*m, argv[1])) return 1;
    if (mode == "steps") {
        const int n = (argc > 4) ? std::atoi(argv[4]) : 1;
        for (int q = 0; q < n; q++) RadiationSelfTest::step(*m);
    } else if (mode == "radiation") {
        RadiationSelfTest::radiation(*m);
        RadiationSelfTest::save_rad(*m, std::string(argv[2]) + ".rad");
    } else if (mode == "steps_rad") {
        const int n = (argc > 4) ? std::atoi(argv[4]) : 1;
        for (int q = 0; q < n; q++) {
            RadiationSelfTest::step(*m);
            RadiationSelfTest::radiation(*m);
        }
        RadiationSelfTest::save_rad(*m, std::string(argv[2]) + ".rad");
    } else if (mode == "steps_conv") {
        // n iterations with the ice scheme and moist convection in the moist block; the ice
        // arrays travel in .mp, the convection arrays in .mc, the Evaporation diagnostic in .evap
        if (!RadiationSelfTest::load_mp(*m, std::string(argv[1]) + ".mp")) return 1;
        if (!RadiationSelfTest::load_evap(*m, std::string(argv[1]) + ".evap")) return 1;
        RadiationSelfTest::io_fields(*m, RadiationSelfTest::mc_fields(*m), std::string(argv[1]) + ".mc", true);
        const int n = (argc > 4) ? std::atoi(argv[4]) : 1;
        for (int q = 0; q < n; q++) RadiationSelfTest::step(*m, true, true);
        RadiationSelfTest::save_mp(*m, std::string(argv[2]) + ".mp");
        RadiationSelfTest::io_fields(*m, RadiationSelfTest::mc_fields(*m), std::string(argv[2]) + ".mc", false);
        RadiationSelfTest::save_evap(*m, std::string(argv[2]) + ".evap");
    } else if (mode == "steps_ice") {
        if (!RadiationSelfTest::load_mp(*m, std::string(argv[1]) + ".mp")) return 1;
        const int n = (argc > 4) ? std::atoi(argv[4]) : 1;
        for (int q = 0; q < n; q++) RadiationSelfTest::step(*m, true);
        RadiationSelfTest::save_mp(*m, std::string(argv[2]) + ".mp");
    } else if (mode == "ice") {
        if (!RadiationSelfTest::load_mp(*m, std::string(argv[1]) + ".mp")) return 1;
        RadiationSelfTest::ice(*m);
        RadiationSelfTest::save_mp(*m, std::string(argv[2]) + ".mp");
    } else if (mode == "conv") {
        const int n = (argc > 4) ? std::atoi(argv[4]) : 1;
        if (!RadiationSelfTest::conv(*m, argv[1], argv[2], n)) return 1;
    } else if (mode.rfind("op:", 0) == 0) {
        RadiationSelfTest::op(*m, mode.substr(3));
    } else if (mode == "init") {
        RadiationSelfTest::init_chain(*m);
    } else if (mode.rfind("turb:", 0) == 0) {
        const std::string model = mode.substr(5);
        if (model != "k_epsilon" && model != "k_omega" && model != "k_omega_SST") {
            std::fprintf(stderr, "unknown turbulence model %s\n", model.c_str());
            return 2;
        }
        RadiationSelfTest::turb_init(*m, model);
        RadiationSelfTest::save(*m, (std::string(argv[2]) + ".init").c_str());
        const int n = (argc > 4) ? std::atoi(argv[4]) : 1;
        for (int q = 0; q < n; q++) RadiationSelfTest::step(*m);
    }
    RadiationSelfTest::save(*m, argv[2]);
    std::cout.rdbuf(old);
    return 0;
}
