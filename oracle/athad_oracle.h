// CPU oracle for the ATHAD_COND time step.
//
// TEST INFRASTRUCTURE ONLY. This is a clean-room restatement of the reference's scoped
// time step (RK4 of RHS_Atm_Turb, PressureSolverAtm red-black sweep, TurbulenceAtm k-omega
// SST, SaturationAdjustment + SaturationH2O/AtmMixture, ThermoAtm::densities and
// waterVapourEvaporation, BC_Atm, the damp_wiggles / polar / radial filters and the
// storeIntermediateData3D copy). It exists to check the HIP product and to serve as the
// CPU baseline; nothing in the product library links, loads or calls it.
//
// Grid-generic (any im x jm x km with a uniform 180/(jm-1) and 360/(km-1) degree spacing),
// contiguous [i][j][k] storage with k fastest. Two RK4 update orders:
//   ORDER_LEX     the reference's in-place sweep at OMP_NUM_THREADS=1
//                 (RungeKutta_Atm_Turb.cpp:99-264 with one thread)
//   ORDER_JACOBI  every cell integrates its four stages against the start-of-RK4
//                 neighbourhood (order-free; the GPU throughput form)
//   ORDER_LEX_PLANES  ORDER_LEX computed hyperplane by hyperplane (i+j+k = d) with the cells
//                 of a plane in parallel: bit-identical to ORDER_LEX, fast enough for the
//                 full grid
//   ORDER_JACOBI_REVERSE  ORDER_JACOBI swept with i, j, k descending: identical results
//                 unless a read sees a value written in the same sweep (a check of the
//                 Jacobi restatement)
#pragma once
#include <cstdint>

#ifdef __cplusplus
extern "C" {
#endif

// Field ids. Shared numbering with include/athad_hip.h.
enum orc_field {
    ORC_T = 0, ORC_U, ORC_V, ORC_W, ORC_C, ORC_CLOUD, ORC_ICE, ORC_GR, ORC_CO2, ORC_TKE, ORC_DIS,
    ORC_TN, ORC_UN, ORC_VN, ORC_WN, ORC_CN, ORC_CLOUDN, ORC_ICEN, ORC_GRN, ORC_CO2N, ORC_TKEN, ORC_DISN,
    ORC_P_DYN, ORC_P_STAT, ORC_R_HUMID, ORC_R_DRY, ORC_T_EQ, ORC_NUE, ORC_PROD, ORC_TKE_SRC,
    ORC_DIS_SRC, ORC_AUX_U, ORC_AUX_V, ORC_AUX_W, ORC_S_C_C,
    ORC_NFIELD
};
// 2-D (jm x km) fields
enum orc_field2d { ORC2_VEL_STAR = 0, ORC2_T_TOP_INIT = 1, ORC2_NFIELD = 2 };

enum orc_order { ORC_ORDER_LEX = 0, ORC_ORDER_JACOBI = 1, ORC_ORDER_LEX_PLANES = 2, ORC_ORDER_JACOBI_REVERSE = 3 };

// OpenMP threads of the oracle's parallel loops (n <= 0: the OpenMP default); returns the
// number now in use.
int orc_set_threads(int n);

// Operators that can be called one at a time (unit parity tests).
enum orc_op {
    ORC_OP_POISSON = 0, ORC_OP_SATADJ, ORC_OP_DAMP_MOIST, ORC_OP_MOIST_CLAMP, ORC_OP_DENSITIES,
    ORC_OP_EVAPORATION, ORC_OP_TURBULENCE, ORC_OP_DAMP_TURB, ORC_OP_BC, ORC_OP_RK4,
    ORC_OP_POLAR, ORC_OP_RADIAL, ORC_OP_STORE, ORC_OP_LEVEL_MEAN,
    ORC_OP_PROJECT,   // PressureSolverAtm::project_initial_velocity(200)
    ORC_NOP
};

// Member parameters that vary across the ensemble (surface composition and state).
typedef struct orc_member {
    double t_surf_equator, t_surf_pole;  // [K]
    double r_air;                        // [kg/m3] reference density at the sea surface
    double x_H2O, x_CO2, x_N2;           // surface mole fractions
    double c_0, co2_0, c_h2o_dry_top;    // [kg/kg]
} orc_member;

typedef struct orc_state orc_state;

// Creates the nominal initial state: RunTimeSlice's thermodynamic init chain
// (cAtmosphereModel.cpp:527-568), then u,v,w = +-1e-3 m/s seeded and mirrored N-S, k = 1e-4 m2/s2,
// omega = 1 1/s, one turbulence update and its damp_wiggles (:612-622), boundaries, the projection
// of the initial velocity (:665, 200 Poisson sweeps and the gradient correction), t_eq = t, store.
// member == NULL selects config_cond.xml's values.
orc_state* orc_create(int im, int jm, int km, uint64_t seed, const orc_member* member);
void       orc_destroy(orc_state* s);
// Blank state of the right size with the grid/parameters set but all fields zero (for
// loading a state from elsewhere).
orc_state* orc_create_blank(int im, int jm, int km, const orc_member* member);
// The state right after the thermodynamic init chain (surface parabola, dry column,
// damp_wiggles(t), initWaterWapour, two co2Atmosphere + densities(true) passes).
orc_state* orc_create_thermo(int im, int jm, int km, const orc_member* member);

double* orc_field_ptr(orc_state* s, int field);
double* orc_field2d_ptr(orc_state* s, int field);
int     orc_total_iter(const orc_state* s);
void    orc_set_iter(orc_state* s, int iter_n, int total_iter_count);
double  orc_scalar(const orc_state* s, const char* name);
double  orc_level_mean(const orc_state* s, int i);

// One scoped time step (iter_n and total_iter_count advance by one).
void orc_step(orc_state* s, int order);
// n steps.
void orc_run(orc_state* s, int n, int order);
// on != 0: the SST closure finds its wall distance by the reference's distance-to-land scans
// over an all-water mask (RHS_Atm_Turb.cpp:640-671, O(jm + km) per cell and stage) instead of
// their closed form min(i, j, k). Same results; for timing the reference's own cost.
void orc_set_sst_scan(orc_state* s, int on);
// The turbulence closure (TurbulenceAtm.h:24 order): 0 k-epsilon, 1 k-omega, 2 k-omega SST
// (the default). Returns -1 for any other value.
int orc_set_turb_model(orc_state* s, int model);
// TurbulenceAtm::init (TurbulenceAtm.h:39-59) with the state's closure.
void orc_turbulence_init(orc_state* s);
// One operator, with the cadence flags of the step it would run in. For ORC_OP_RK4 the
// ramps are taken from the current total_iter_count.
void orc_op(orc_state* s, int op, int order);

// Radiation of run_3D_loop's mode 2, called after a step: MultiLayerRadiation on the current t
// (MultiLayerRadiation.h:59-384), t blended 5 % toward it (cAtmosphereModel.cpp:863-879), one
// t_skin relaxation (cAtmosphereModel.cpp:276-311; later densities use the new t_skin).
void orc_radiation(orc_state* s);
// Its outputs: 0 radiation (upward long wave per cell), 1 epsilon (per cell), 2 albedo
// (jm x km), 3 top short wave (jm); NULL before the first call. orc_scalar(s, "t_skin").
double* orc_rad_ptr(orc_state* s, int which);

// Two-category ice scheme (CategoryIceScheme = 2): TwoCatIceScheme::run
// (TwoCatIceScheme.h:89-111) on the current state, then run_3D_loop's damp_wiggles of the
// two precipitation fluxes (cAtmosphereModel.cpp:1243-1244). Reads t c cloud ice co2 p_stat
// r_humid S_c_c and the stored fluxes P_rainn, P_snown; writes the fluxes, Precipitation and
// S_v S_c S_i S_r S_s, and (where no condensed phase can exist) t c cloud ice gr S_c_c.
enum orc_mp {
    ORC_MP_P_RAIN = 0, ORC_MP_P_SNOW, ORC_MP_P_RAINN, ORC_MP_P_SNOWN, ORC_MP_PRECIP,
    ORC_MP_S_V, ORC_MP_S_C, ORC_MP_S_I, ORC_MP_S_R, ORC_MP_S_S,
    ORC_NMP
};
void    orc_ice_scheme(orc_state* s);
// Puts the ice scheme in its place inside the moist block of every later step (after the
// damping of t ice c cloud, cAtmosphereModel.cpp:1230-1244, with precipitationSum :1263 and
// the cap of its sources :1278-1305). Wherever its arrays exist the RHS adds their sources
// (RHS_Atm_Turb.cpp:963-966, 1058-1070).
void    orc_set_ice_scheme(orc_state* s, int on);
// P_rainn <- P_rain, P_snown <- P_snow (storeIntermediateData3D, UtilsAtm.h:41-42).
void    orc_ice_store(orc_state* s);
// The arrays above (im*jm*km doubles, [i][j][k]); allocated, zero, at first use.
double* orc_mp_ptr(orc_state* s, int which);

// Moist convection (MoistConvection.h:106-140): the convection arrays, in the order the
// reference harness writes them, and one call of the scheme on the state's fields.
enum orc_mc {
    ORC_MC_S = 0, ORC_MC_M_U, ORC_MC_M_D, ORC_MC_S_U, ORC_MC_Q_V_U, ORC_MC_Q_C_U, ORC_MC_C_U,
    ORC_MC_E_L, ORC_MC_E_D_EVAP, ORC_MC_E_P, ORC_MC_G_P, ORC_MC_S_D, ORC_MC_Q_V_D, ORC_MC_P_CONV,
    ORC_MC_E_U, ORC_MC_D_U, ORC_MC_U_U, ORC_MC_V_U, ORC_MC_W_U, ORC_MC_E_D, ORC_MC_D_D, ORC_MC_U_D,
    ORC_MC_V_D, ORC_MC_W_D, ORC_MC_MC_T, ORC_MC_MC_Q, ORC_MC_MC_V, ORC_MC_MC_W, ORC_MC_DEEP_BEG,
    ORC_MC_DEEP_END, ORC_MC_CLOUD_BASE, ORC_MC_LFS,
    ORC_NMC
};
void    orc_moist_convection(orc_state* s);
int     orc_cape(const orc_state* s, double* out);   // computeCAPE's im values after the last call; -1 before any
// Puts moist convection in its place inside the moist block of every later step, after the ice
// scheme (which on != 0 also switches on): MoistConvection::run, the damp_wiggles of P_conv, E_u,
// M_u, q_v_u, q_c_u, E_d, M_d, q_v_d, MC_t, MC_v, MC_w, and P_conv in precipitationSum
// (cAtmosphereModel.cpp:1247-1263, UtilsAtm.h:52-64). Wherever the convection arrays exist the
// RHS adds MC_t, MC_v, MC_w, MC_q (RHS_Atm_Turb.cpp:370-372, 964, 1023, 1042, 1061).
void    orc_set_moist_convection(orc_state* s, int on);
double* orc_mc_ptr(orc_state* s, int which);      // im*jm*km doubles
double* orc_evap_ptr(orc_state* s);               // jm*km doubles, the Evaporation diagnostic [mm/d]

// Individual thermodynamic functions (known-answer tests).
double orc_p_sat(double T);
double orc_p_subl(double T);
double orc_p_auto(double T);
double orc_latent_heat(double T);
double orc_q_sat(double E, double p, double M_other);
double orc_dqsat_dT(double T, double p, double M_other);
double orc_moist_lapse(double T, double p, double M_other, double cp, double R_mix, double g);
double orc_cp_of(double c, double co2, double T, double M_bg);
double orc_R_of(double c, double co2, double R_bg);
double orc_M_nonwater(double c, double co2, double M_bg);
// Resolved composition: out[0..6] = M_mean, R_mix, q_H2O, q_CO2, q_bg, M_bg, R_bg.
void   orc_resolve(double x_H2O, double x_CO2, double x_N2, double* out);
// Per-cell saturation adjustment on arrays of n cells (the adjustSaturation + clampAndFade
// pair at one level index i_level of the grid's layer heights).
void   orc_satadj_cells(int n, double* t, double* c, double* cloud, double* ice, double* gr,
                        const double* co2, const double* p_stat, double* s_c_c,
                        double dt_dim, int interior);

#ifdef __cplusplus
}
#endif
