/*
 * athad_hip.h — C-ABI of the MI355X (gfx950) implementation of the ATHAD_COND time step.
 *
 * The reference has no C interface for this path: cAtmosphereModel::run_3D_loop
 * (atmosphere/cAtmosphereModel.cpp:744-1728) calls five operators synchronously, each a
 * friend class built on the model, plus the boundary/filter glue between them. This header
 * exports one entry point per call site (cited on each declaration) and one for the whole
 * scoped iteration, over device-resident state owned by an opaque context. Host arrays stay
 * owned by the caller; plain C types only; every function returns an athad_status and never
 * throws or aborts across the boundary.
 *
 * State layout on the device: one contiguous [i][j][k] array per field (k fastest), FP64 in
 * the drop-in form, FP32 (p_stat FP64) in the mixed form. A context holds one or more
 * members (independent integrations, possibly of different grid sizes) that share a HIP
 * stream; every operator advances all of them.
 */
#ifndef ATHAD_HIP_H
#define ATHAD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int athad_status;
#define ATHAD_OK            0
#define ATHAD_ERR_ARG       1   /* bad argument (null pointer, unknown field, bad member) */
#define ATHAD_ERR_DEVICE    2   /* HIP runtime error (message in athad_last_error) */
#define ATHAD_ERR_ALLOC     3   /* device allocation failed */
#define ATHAD_ERR_SHAPE     4   /* grid shape not supported or mismatched */
#define ATHAD_ERR_STATE     5   /* call not valid in the context's current state */
#define ATHAD_NOT_READY     6   /* athad_query: work still in flight (not an error) */

/* Storage precision */
#define ATHAD_FP64   0          /* drop-in form: FP64 storage and arithmetic */
#define ATHAD_MIXED  1          /* FP32 storage/stencils, FP64 thermodynamics and p_stat */

/* RK4 update order. RungeKutta_Atm_Turb.cpp:99-264 sweeps in place, so a cell's neighbours are
 * read at whatever time level the sweep has left them; the two orders below are the
 * reproducible choices. */
#define ATHAD_ORDER_LEX     0   /* the reference's sweep at OMP_NUM_THREADS=1 (parity form) */
#define ATHAD_ORDER_JACOBI  1   /* start-of-step neighbourhood (throughput form) */
#define ATHAD_ORDER_SYMMETRIC 2 /* Jacobi order with north-south mirror symmetry kept to the last
                                   bit: mirrored latitude tables (cos theta = 0 at the equator),
                                   latitude neighbours summed before weighting, SST wall distance
                                   from the nearer pole, turbulence floors on both latitude
                                   neighbours. Not in the reference, whose sweep order and wall
                                   distance break the mirror at rounding level and near the poles. */

/* Turbulence closures, in the order of TurbulenceAtm's Model enum (TurbulenceAtm.h:24); the
 * reference selects one from the turb_model string (cAtmosphereModel.cpp:484-489). */
#define ATHAD_TURB_K_EPSILON    0   /* Chien 1982 low-Reynolds k-epsilon */
#define ATHAD_TURB_K_OMEGA      1   /* Wilcox 1988 k-omega */
#define ATHAD_TURB_K_OMEGA_SST  2   /* Menter 1994 k-omega SST, the configuration's choice (default) */

/* Field ids: the cAtmosphereModel Array members of the scoped step
 * (atmosphere/cAtmosphereModel.h:535-662). */
enum athad_field {
    ATHAD_T = 0, ATHAD_U, ATHAD_V, ATHAD_W, ATHAD_C, ATHAD_CLOUD, ATHAD_ICE, ATHAD_GR, ATHAD_CO2,
    ATHAD_TKE, ATHAD_DIS,
    ATHAD_TN, ATHAD_UN, ATHAD_VN, ATHAD_WN, ATHAD_CN, ATHAD_CLOUDN, ATHAD_ICEN, ATHAD_GRN,
    ATHAD_CO2N, ATHAD_TKEN, ATHAD_DISN,
    ATHAD_P_DYN, ATHAD_P_STAT, ATHAD_R_HUMID, ATHAD_R_DRY, ATHAD_T_EQ, ATHAD_NUE, ATHAD_PROD,
    ATHAD_TKE_SOURCE, ATHAD_DIS_SOURCE, ATHAD_AUX_U, ATHAD_AUX_V, ATHAD_AUX_W, ATHAD_S_C_C,
    ATHAD_NFIELD
};
/* 2-D (jm x km) fields: Array_2D vel_star (cAtmosphereModel.h:533) and the lid snapshot
 * t_top_init (cAtmosphereModel.h:400). */
enum athad_field2d { ATHAD_VEL_STAR = 0, ATHAD_T_TOP_INIT = 1, ATHAD_NFIELD2D = 2 };

typedef struct athad_grid {
    int im, jm, km;             /* radial levels, latitudes (pole to pole), longitudes (seam duplicated) */
} athad_grid;

/* Per-member surface state and composition; the remaining parameters are the values of
 * cli/config_cond.xml. A NULL pointer where one is accepted selects that file's values. */
typedef struct athad_member_params {
    double t_surf_equator, t_surf_pole;   /* [K] */
    double r_air;                         /* [kg/m3] reference density at the sea surface */
    double x_H2O, x_CO2, x_N2;            /* surface mole fractions */
    double c_0, co2_0, c_h2o_dry_top;     /* [kg/kg] */
} athad_member_params;

typedef struct athad_ctx athad_ctx;

/* --- lifetime ------------------------------------------------------------------------ */

/* Replaces constructing cAtmosphereModel + LoadConfig (cAtmosphereModel.cpp:74-127) for the
 * device state of n_members integrations. grids[n_members]; params may be NULL. */
athad_status athad_create(int n_members, const athad_grid* grids, const athad_member_params* params,
                          int precision, int order, athad_ctx** out);
athad_status athad_destroy(athad_ctx* ctx);
/* Last error message of this context (or of the last failed athad_create when ctx is NULL). */
const char* athad_last_error(const athad_ctx* ctx);
/* Boiling temperature [K] of water at vapour pressure e [hPa]: bisection on the saturation
 * curve, as SaturationH2O::dewPoint (atmosphere/SaturationH2O.h:189-202). Host only. */
double athad_dew_point(double e_hPa);
/* Member parameters of an integration from its surface state: sea saturated at t_sea under
 * p_surf, CO2 : N2 at the dry ratio, r_air from the ideal gas at the surface, dry-top
 * vapour mass fraction from the dry composition; the pole is 10 K colder than the equator.
 * As test/cond_column_selftest.cpp:86-145 derives the config values. Host only; returns
 * ATHAD_ERR_ARG if the sea would boil. */
athad_status athad_member_from_surface(double p_surf_bar, double t_sea, double x_h2o_dry, double x_co2_dry,
                                       athad_member_params* out);
/* Device bytes held by a context of this shape (admission control): the fields, the Jacobi
 * buffers, scratch, the tables and reductions. Not counted, because they are allocated only
 * once their operator is first used: the ice-scheme arrays (10 fields in the storage
 * precision), the radiation outputs (2 FP64 fields and jm*km) and the moist-convection arrays
 * (36 FP64 fields, jm*km and jm doubles, and the CAPE profile: im doubles and 2 jm*km ints). */
int64_t athad_footprint_bytes(const athad_grid* grid, int precision);

/* --- state transfer ---------------------------------------------------------------- */

/* Nominal initial state for every member (init recipe of RunTimeSlice's thermodynamic
 * chain, cAtmosphereModel.cpp:527-568, plus the benchmark's velocity/turbulence fields, the
 * boundaries and the projection of the initial velocity, :665, with 200 sweeps);
 * member m uses seed + m. */
athad_status athad_init_state(athad_ctx* ctx, uint64_t seed);
/* Contiguous FP64 host arrays [im][jm][km] (jm x km for 2-D fields). */
athad_status athad_set_field(athad_ctx* ctx, int member, int field, const double* src);
athad_status athad_get_field(athad_ctx* ctx, int member, int field, double* dst);
athad_status athad_set_field2d(athad_ctx* ctx, int member, int field, const double* src);
athad_status athad_get_field2d(athad_ctx* ctx, int member, int field, double* dst);
/* Row-pointer form matching the reference Array (double*** x, one km-row per (i,j),
 * lib/Array.cpp:62-71): rows[i*jm + j] points at row (i,j). */
athad_status athad_upload_rows(athad_ctx* ctx, int member, int field, double* const* rows);
athad_status athad_download_rows(athad_ctx* ctx, int member, int field, double* const* rows);
/* Sum and sum of squares of one field of one member, reduced on the device in a fixed order
 * (equal fields give equal bits). No counterpart in the reference; for result digests that
 * avoid downloading whole fields. */
athad_status athad_field_stats(athad_ctx* ctx, int member, int field, double* sum, double* sumsq);
/* Replaces a member's surface state and composition (the constants and tables derived from
 * them) so a context can be reused for a new integration of the same grid; follow with
 * athad_init_state. No counterpart in the reference (one model per process there). */
athad_status athad_set_member_params(athad_ctx* ctx, int member, const athad_member_params* params);
/* iteration counters iter_n / total_iter_count (cAtmosphereModel.h:311-316) */
athad_status athad_set_iter(athad_ctx* ctx, int member, int iter_n, int total_iter_count);
athad_status athad_get_iter(athad_ctx* ctx, int member, int* iter_n, int* total_iter_count);

/* --- the five call sites of run_3D_loop and their glue ------------------------------ */

/* PressureSolverAtm(*this).project_initial_velocity(n_sweeps)  (cAtmosphereModel.cpp:665;
 * PressureSolverAtm.h:686-767): the velocity as the Poisson source, n_sweeps sweeps of the
 * solver, the gradient correction v -= grad p_dyn of the interior, p_dyn and aux_u/v/w cleared.
 * athad_init_state runs it with 200 sweeps, as RunTimeSlice does. */
athad_status athad_project_initial_velocity(athad_ctx* ctx, int n_sweeps);
/* PressureSolverAtm(*this).run()  (cAtmosphereModel.cpp:1158; PressureSolverAtm.h:27-669):
 * one red-black sweep of p_dyn with its boundaries and ceiling. */
athad_status athad_poisson_sweep(athad_ctx* ctx);
/* SaturationAdjustment(*this).run()  (cAtmosphereModel.cpp:1218; SaturationAdjustment.h:25-45),
 * including the S_c_c cap of cAtmosphereModel.cpp:1278-1295. */
athad_status athad_satadj(athad_ctx* ctx);
/* damp_wiggles of t, ice, c, cloud and the microphysics clamp / vapour ceiling
 * (cAtmosphereModel.cpp:1225-1228, 1307-1399). */
athad_status athad_damp_moist(athad_ctx* ctx);
/* ThermoAtm(*this).densities(set_water)  (cAtmosphereModel.cpp:1418; ThermoAtm.h:1436-1614);
 * with_evaporation also applies waterVapourEvaporation (:1421; ThermoAtm.h:115-258). */
athad_status athad_densities(athad_ctx* ctx, int set_water_profile, int with_evaporation);
/* TurbulenceAtm(*this).run() and the damp_wiggles of its six fields
 * (cAtmosphereModel.cpp:1426-1433; TurbulenceAtm.h:62-79). */
athad_status athad_turbulence(athad_ctx* ctx);
/* The turbulence closure of every member (ATHAD_TURB_*), used from then on by the turbulence
 * update (compute_sources with compute_k_epsilon, compute_k_omega or compute_k_omega_SST,
 * TurbulenceAtm.h:323-773), the RK4 right-hand side (RHS_Atm_Turb.cpp:528-721) and the wall
 * value of dis (apply_wall_bc, TurbulenceAtm.h:109-147). Replaces setting turb_model before
 * TurbulenceAtm is constructed (cAtmosphereModel.cpp:484-489). The laminar mode is not
 * offered: the reference never sets re_turb without TurbulenceAtm::init. */
athad_status athad_set_turbulence_model(athad_ctx* ctx, int model);
/* TurbulenceAtm(*this).init() with the current closure (cAtmosphereModel.cpp:613;
 * TurbulenceAtm.h:39-59): the boundary-layer profile of tke and dis with their stored copies
 * (init_fields, :193-255), the wall values, compute_sources and clamp_nue. athad_init_state
 * seeds tke and dis with the benchmark's uniform k = 1e-4 m2/s2, omega = 1/s instead. */
athad_status athad_turbulence_init(athad_ctx* ctx);
/* bcRadius, apply_wall_bc, bcTheta, bcPhi, bcScalarSurfSur, bcSolidGround
 * (cAtmosphereModel.cpp:1440-1446). */
athad_status athad_boundary(athad_ctx* ctx);
/* solveRungeKutta_Atmosphere_Turb()  (cAtmosphereModel.cpp:1509; RungeKutta_Atm_Turb.cpp:44-273)
 * in the context's update order. */
athad_status athad_rk4(athad_ctx* ctx);
/* polar_zonal_filter(u,v,w) and the radial Shapiro filters (cAtmosphereModel.cpp:1524-1575). */
athad_status athad_filters(athad_ctx* ctx);
/* UtilsAtm(*this).storeIntermediateData3D(1.0)  (cAtmosphereModel.cpp:1647). */
athad_status athad_store(athad_ctx* ctx);

/* One scoped iteration of run_3D_loop (cAtmosphereModel.cpp:1103-1647 minus the rows
 * outside the path: ice scheme, moist convection, radiation, diagnostics, output), with
 * the cadences of the reference (Poisson every 4th, moist block every 2nd iteration).
 * Advances iter_n and total_iter_count. */
athad_status athad_step(athad_ctx* ctx);
/* n_steps iterations; use_graph replays a captured 4-iteration hipGraph. */
athad_status athad_run(athad_ctx* ctx, int n_steps, int use_graph);
/* Non-blocking: ATHAD_OK when all work of the context has finished, ATHAD_NOT_READY while
 * some is in flight. For host loops that keep several contexts busy at once (serving). */
athad_status athad_query(athad_ctx* ctx);
/* Free and total device memory of the current device (admission control). */
athad_status athad_device_memory(int64_t* free_bytes, int64_t* total_bytes);
/* The tile of the fused damp_wiggles kernel (tj latitudes x tk longitudes, `levels` radial
 * levels per block) chosen for this context's grids by timing every candidate at creation.
 * No counterpart in the reference (tiling is a device choice; results do not depend on it). */
athad_status athad_damp_tile(athad_ctx* ctx, int* tj, int* tk, int* levels);
/* Forces tile candidate `shape` (0 .. n-1, n fixed by the library; ATHAD_ERR_ARG beyond) and
 * `levels` per block; for tests that check every candidate gives the same result. */
athad_status athad_set_damp_tile(athad_ctx* ctx, int shape, int levels);
/* Waits for the context's stream. */
athad_status athad_sync(athad_ctx* ctx);

/* --- restart files ------------------------------------------------------------------- */

/* The reference's checkpoint format (cAtmosphereModel::save_state / load_state,
 * atmosphere/FileIO_Atm.cpp:293-356): an int32 header {0x41544D31 "ATM1", im, jm, km,
 * total_iter_count}, then the 29 arrays of restart_arrays() (FileIO_Atm.cpp:297-304), each
 * im*jm*km FP64 values in [i][j][k] order: t u v w c cloud ice gr co2, tn un vn wn cn cloudn
 * icen grn co2n, p_dyn p_stat, tke dis tken disn nue, P_rain P_snow P_rainn P_snown. The four
 * precipitation arrays are zero in the scoped step and not held by the context: they are
 * written as zeros and skipped on load. Host byte order (little-endian on the targets). */
/* Size in bytes of a restart of grid g. */
int64_t athad_restart_bytes(const athad_grid* g);
/* Writes member's restart into buf (size >= athad_restart_bytes). Replaces save_state. */
athad_status athad_save_state_mem(athad_ctx* ctx, int member, void* buf, int64_t size);
/* Loads a restart from buf: ATHAD_ERR_SHAPE on a bad header or a grid that differs from the
 * member's, ATHAD_ERR_ARG when buf is shorter than the arrays. Sets the member's total
 * iteration count from the header and leaves its cadence counter (athad_set_iter) as it was,
 * as load_state leaves run_3D_loop's iteration to the caller. Fields outside the restart
 * keep their values. Replaces load_state. */
athad_status athad_load_state_mem(athad_ctx* ctx, int member, const void* buf, int64_t size);
/* The same through a file (FileIO_Atm.cpp names it atm_restart_<Ma>Ma_<iter>.bin). */
athad_status athad_save_state(athad_ctx* ctx, int member, const char* path);
athad_status athad_load_state(athad_ctx* ctx, int member, const char* path);

/* --- checkpoint diagnostics on the device --------------------------------------------- */

/* searchMinMax_3D (atmosphere/MinMax_Atm.cpp:51-108) of one field without downloading it: the
 * minimum and maximum of its finite values and the (i, j, k) of their first occurrence in
 * [i][j][k] order (the reference's strict comparisons keep the first of equal extremes); 0 at
 * (0, 0, 0) when no value is finite. pos_* point to 3 ints; any output may be NULL. */
athad_status athad_field_minmax(athad_ctx* ctx, int member, int field, double* vmin, int* pos_min,
                                double* vmax, int* pos_max);
/* findResiduumAtm (atmosphere/UtilsAtm.h:70-218): over levels 10..30 and the interior
 * latitudes and longitudes, the residual sqrt((du/dr^2 + dv/dtheta^2 + dw/dphi^2) / 3) of the
 * centred velocity derivatives; its maximum and position (ties: smallest (i, j, k), as the
 * reference's tie-break; 0 at (0, 0, 0) when no residual exceeds 0), the average of the finite residuals and their count, and the count
 * and first (i, j, k) of the non-finite ones ({-1,-1,-1} when none). The reference sums in its
 * thread order, this in a fixed order: the average agrees to rounding. ATHAD_ERR_SHAPE for
 * im < 32 (the reference's scan would leave the grid). The caller keeps the previous maximum
 * for the "declining" test, as the reference keeps residuum_old. */
athad_status athad_residuum(athad_ctx* ctx, int member, double* res_max, int* pos_max, double* avg_abs,
                            int64_t* n_finite, int64_t* n_nonfinite, int* pos_first_nonfinite);
/* write_meridional_streamfunction (atmosphere/MinMax_Atm.cpp:173-227) without downloading v and
 * r_humid: the zonal means vbar = <v> u_0 and rho_vbar = <r_humid v> u_0 [i][j] (non-finite v
 * skipped; non-finite or non-positive r_humid replaced by r_air), and the meridional mass
 * streamfunction psi [i][j] in kg/s, integrated down from the lid (psi = 0 at i = im-1) with
 * 2 pi a sin(theta) and the float layer-height differences. Summation and integration orders
 * are the reference's, so the arrays are bit for bit its own. Each output points to im * jm
 * doubles or is NULL. The CSV and the log line stay the caller's. */
athad_status athad_meridional_streamfunction(athad_ctx* ctx, int member, double* vbar, double* rho_vbar,
                                             double* psi);
/* One plane of one field, widened to double, for the output writers that read a plane only:
 * axis 0 fixes the level i = index ([j][k], jm * km values: paraview_vtk_radial,
 * atmosphere/Paraview_Atm.cpp:407-627, and the surface row of AtmosphereDataTransfer,
 * atmosphere/FileIO_Atm.cpp:64-120), axis 2 the meridian k = index ([i][j], im * jm:
 * paraview_vtk_zonal, :628-827), axis 1 the latitude j = index ([i][k], im * km:
 * paraview_vtk_longal, :828-). The writers' scaling, formatting and finiteness checks stay in
 * the caller. ATHAD_ERR_ARG for an axis or index outside the grid. */
athad_status athad_get_slice(athad_ctx* ctx, int member, int field, int axis, int index, double* dst);

/* --- radiation (run_3D_loop's radiation_mode 2) --------------------------------------- */

/* The radiation branch of run_3D_loop (atmosphere/cAtmosphereModel.cpp:1655-1661, called after
 * the store): apply_radiative_heating (:863-879) runs MultiLayerRadiation
 * (atmosphere/MultiLayerRadiation.h:59-384) on the current t of every member and blends t by
 * omega_rad = 0.05 toward its result; then updateSkinTemperature (:276-311) moves each
 * member's t_skin one relaxation step toward the emission temperature of its absorbed short
 * wave plus the geothermal flux, which the next densities() uses. The outputs (upward long-wave
 * flux, emissivity, albedo) are kept on the device; the first call allocates them. Grids up to
 * 80 levels (ATHAD_ERR_SHAPE above). athad_init_state resets t_skin to its configured value. */
athad_status athad_radiation(athad_ctx* ctx);
/* radiation_mode 2 inside the step (on != 0): every later athad_step / athad_run iteration ends,
 * after its store, with the radiation branch of run_3D_loop (:1655-1661) as athad_radiation
 * runs it, the skin-temperature step included, on the device, so captured graphs carry it and no
 * host round trip is needed. athad_get_radiation reports the t_skin reached. on = 0 stops it.
 * ATHAD_ERR_SHAPE above 80 levels. */
athad_status athad_set_radiation(athad_ctx* ctx, int on);
/* The radiation outputs of one member: radiation and epsilon (im*jm*km doubles, [i][j][k]),
 * albedo (jm*km) and the current t_skin in K; any may be NULL. ATHAD_ERR_STATE for the arrays
 * before the first athad_radiation. */
athad_status athad_get_radiation(athad_ctx* ctx, int member, double* radiation, double* epsilon, double* albedo,
                                 double* t_skin);

/* --- two-category ice scheme (CategoryIceScheme = 2) -------------------------------- */

/* Arrays of the ice scheme: Array members P_rain, P_snow, P_rainn, P_snown, Precipitation,
 * S_v, S_c, S_i, S_r, S_s of cAtmosphereModel (atmosphere/cAtmosphereModel.h). Held in the
 * context's storage precision from the first call that needs them, zero until then. */
enum athad_mp_field {
    ATHAD_P_RAIN = 0, ATHAD_P_SNOW, ATHAD_P_RAINN, ATHAD_P_SNOWN, ATHAD_PRECIPITATION,
    ATHAD_S_V, ATHAD_S_C, ATHAD_S_I, ATHAD_S_R, ATHAD_S_S,
    ATHAD_NMP
};
/* The ice-scheme call site of run_3D_loop (atmosphere/cAtmosphereModel.cpp:1230-1244 with
 * CategoryIceScheme = 2): TwoCatIceScheme(*this).run() (atmosphere/TwoCatIceScheme.h:89-111:
 * initArrays, computeColumns, extrapolateBC of the fluxes; applyTopography is a no-op without
 * land) and the damp_wiggles of P_rain and P_snow, on every member. Reads t c cloud ice gr co2
 * p_stat r_humid S_c_c and the stored fluxes P_rainn, P_snown; writes the fluxes,
 * Precipitation and the five sources, and where no condensed phase can exist
 * (IceSchemeCommon.h:104-190) t c cloud ice gr S_c_c. athad_store and every iteration's store
 * then copy P_rain, P_snow to P_rainn, P_snown (UtilsAtm.h:41-42). */
athad_status athad_ice_scheme(athad_ctx* ctx);
/* CategoryIceScheme = 2 inside the step (on != 0): every later moist iteration of athad_step /
 * athad_run runs the ice-scheme call site in its place in the moist block, after the damping of
 * t ice c cloud, then precipitationSum (:1263; graupel and the moist convection's P_conv are
 * zero: athad_step does not call athad_moist_convection) and cap_S of the sources (:1278-1305).
 * Independently of this switch, once the ice-scheme arrays exist the RK4 tendencies of t, c,
 * cloud, ice carry their sources (atmosphere/RHS_Atm_Turb.cpp:963-966, 1058-1067), the
 * boundaries extrapolate them (atmosphere/BC_Atm.h:567-842), and saturation adjustment
 * evaporates falling precipitation where no condensed phase can exist. on = 0 restores the
 * scoped moist block. */
athad_status athad_set_ice_scheme(athad_ctx* ctx, int on);
/* Contiguous FP64 host arrays [im][jm][km] of one ice-scheme array (athad_mp_field). */
athad_status athad_set_mp_field(athad_ctx* ctx, int member, int which, const double* src);
athad_status athad_get_mp_field(athad_ctx* ctx, int member, int which, double* dst);

/* --- moist convection (atmosphere/MoistConvection.h) ------------------------------------
 * The 32 convection arrays, in the order of athad_conv_field (FP64 in every precision), and
 * the per-column evaporation diagnostic [mm/d] (jm*km) the surface-flux trigger reads. */
typedef enum athad_conv_field {
    ATHAD_CONV_S = 0, ATHAD_CONV_M_U, ATHAD_CONV_M_D, ATHAD_CONV_S_U, ATHAD_CONV_Q_V_U, ATHAD_CONV_Q_C_U,
    ATHAD_CONV_C_U, ATHAD_CONV_E_L, ATHAD_CONV_E_D_EVAP, ATHAD_CONV_E_P, ATHAD_CONV_G_P, ATHAD_CONV_S_D,
    ATHAD_CONV_Q_V_D, ATHAD_CONV_P_CONV, ATHAD_CONV_E_U, ATHAD_CONV_D_U, ATHAD_CONV_U_U, ATHAD_CONV_V_U,
    ATHAD_CONV_W_U, ATHAD_CONV_E_D, ATHAD_CONV_D_D, ATHAD_CONV_U_D, ATHAD_CONV_V_D, ATHAD_CONV_W_D,
    ATHAD_CONV_MC_T, ATHAD_CONV_MC_Q, ATHAD_CONV_MC_V, ATHAD_CONV_MC_W, ATHAD_CONV_DEEP_BEG,
    ATHAD_CONV_DEEP_END, ATHAD_CONV_CLOUD_BASE, ATHAD_CONV_LEVEL_FREE_SINKING,
    ATHAD_CONV_EVAPORATION = 32   /* input, jm*km */
} athad_conv_field;
/* Replaces MoistConvection(*this).run(iter_n) (atmosphere/cAtmosphereModel.cpp:1247;
 * MoistConvection.h:106-140) with the model's defaults iter_prec 3, convection_perturbation 1,
 * convection_mode 1, every member, on the current fields: cloud base and level of free sinking
 * per column, the updraft and downdraft mass fluxes and their recurrences, convective
 * precipitation P_conv, and the tendencies MC_t, MC_q, MC_v, MC_w of rhsForcing. The arrays the
 * reference does not clear between calls (E_u, D_u, E_d, D_d, the updraft / downdraft winds,
 * the MC_* tendencies, the level markers) carry over as there. computeCAPE's 1-D diagnostic is
 * not produced. Arrays are allocated, zero, at the first call (the 32 arrays, two per-cell
 * saturation fractions and two damping temporaries: 36 * im*jm*km * 8 bytes per member, plus
 * jm*km and jm doubles). The zonal circle factor cos(theta_eq - theta_j) of the entrainment gradients is taken
 * about the equator row (jm - 1) / 2 (MoistConvection.h:239 takes row 90, the equator of its
 * 181-row grid). Grids need jm, km, im >= 3 (ATHAD_ERR_ARG otherwise). */
athad_status athad_moist_convection(athad_ctx* ctx);
/* The configured moist block (on != 0): every later moist iteration of athad_step / athad_run
 * runs, after the ice scheme (which this switches on too) and its flux damping, the moist
 * convection call site of run_3D_loop (atmosphere/cAtmosphereModel.cpp:1247-1262):
 * MoistConvection(*this).run(iter_n), then damp_wiggles of P_conv, E_u, M_u, q_v_u, q_c_u, E_d,
 * M_d, q_v_d, MC_t, MC_v, MC_w, and precipitationSum with P_conv (UtilsAtm.h:52-64). Once the
 * convection arrays exist, RK4 adds coeff_MC_t MC_t, coeff_MC_vel MC_v and MC_w, coeff_MC_q MC_q
 * to the tendencies of t, v, w, c (atmosphere/RHS_Atm_Turb.cpp:370-372, 964, 1023, 1042, 1061),
 * and the evaporation step writes the Evaporation diagnostic the scheme's surface-flux trigger
 * reads (Dalton's law, atmosphere/ThermoAtm.h:217-240). on = 0 leaves the ice scheme as it is;
 * athad_set_ice_scheme(ctx, 0) stops both. Allocates the convection arrays (see
 * athad_moist_convection). */
athad_status athad_set_moist_convection(athad_ctx* ctx, int on);
/* Contiguous FP64 host arrays [im][jm][km] of one convection array, or [jm][km] for
 * ATHAD_CONV_EVAPORATION. */
athad_status athad_set_conv_field(athad_ctx* ctx, int member, int which, const double* src);
athad_status athad_get_conv_field(athad_ctx* ctx, int member, int which, double* dst);
/* computeCAPE's diagnostic after the last convection call (MoistConvection.h:1259-1299,
 * cAtmosphereModel::CAPE): im doubles, each level holding the last column (k outer, j inner)
 * whose cloud-base-to-free-sinking integral reached it, zero where none did. */
athad_status athad_get_conv_cape(athad_ctx* ctx, int member, double* dst);

/* --- entry point that allocates nothing ------------------------------------------------ */

/* No counterpart in the reference. searchMinMax_3D (atmosphere/MinMax_Atm.cpp:51-108), as
 * athad_field_minmax computes it, of a field held in the caller's device memory (im*jm*km
 * values, [i][j][k], double for ATHAD_FP64, float for ATHAD_MIXED), for callers that cannot
 * allocate device memory: code under hipGraph capture set-up, tight serving loops, harnesses
 * that forbid hipMalloc. It calls no allocating routine: the reduction's partial results and
 * the descriptor the kernels read live in a fixed file-scope device array of the library
 * (4 KiB + 128 KiB per device, held until the process ends), whose address is looked up once
 * per device. Limit: im <= 200 (ATHAD_ERR_SHAPE beyond). stream: a hipStream_t, NULL for the
 * default stream; the call waits for it before it returns. Calls that overlap on one device
 * share that array and race: issue them one after the other. Errors are reported by
 * athad_last_error(NULL). Outputs as athad_field_minmax; any may be NULL. */
athad_status athad_field_minmax_noalloc(const void* field, int precision, int im, int jm, int km, void* stream,
                                        double* vmin, int* pos_min, double* vmax, int* pos_max);

/* --- measurement --------------------------------------------------------------------- */

/* Per-kernel timing with HIP events on the context's stream: kernel ids as in
 * athad_cond_amd/csrc/athad_kernels.h (0 = RK4). Collected while enabled. */
athad_status athad_profile(athad_ctx* ctx, int kernel_id, int enable);
athad_status athad_profile_result(athad_ctx* ctx, int kernel_id, double* total_ms, int* launches);

/* Library version string. */
const char* athad_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ATHAD_HIP_H */
