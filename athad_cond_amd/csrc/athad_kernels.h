// Host-side launch wrappers for the ATHAD_COND kernels (defined in athad_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "athad_device.h"

namespace athad {

// One batch of members that share a stream and a storage precision.
struct Batch {
    MemberDesc* d_md = nullptr;   // device array [nmem]
    int nmem = 0;
    long long max_n = 0;          // largest im*jm*km in the batch
    int max_im = 0, max_jm = 0, max_km = 0;
    hipStream_t stream = nullptr;
    bool fp32 = false;
    int turb = kTurbKOmegaSST;    // turbulence closure of every member (selects the compute_sources kernel)
};

// Kernel ids for per-kernel timing.
enum KernelId {
    KID_RK4 = 0, KID_RK4_PLANE, KID_LEVEL_MEAN, KID_POISSON, KID_SATADJ, KID_DAMP, KID_DENSITIES,
    KID_TURB, KID_BC, KID_SOLID, KID_POLAR, KID_RADIAL, KID_STORE, KID_INIT, KID_ICE, KID_CONV, KID_COUNT
};

void launch_advance(const Batch& b);
void launch_level_mean(const Batch& b);
// sum and sum of squares of one field of one member, in a fixed order, left at
// scratch[field_stats_offset(im)], [+1]
void launch_field_stats(const Batch& b, int member, int field);
int field_stats_offset(int im);
// searchMinMax_3D of one field of one member: min, its index, max, its index (indices as
// doubles, row-major, -1 for none) at scratch[diag_offset(im)] + 4 * 256
void launch_minmax(const Batch& b, int member, int field);
// findResiduumAtm of one member: max, its index, sum, count, non-finite count, first
// non-finite index at scratch[diag_offset(im)] + 6 * 256
void launch_residuum(const Batch& b, int member);
int diag_offset(int im);
// write_meridional_streamfunction's arrays of one member (grid im x jm x km): vbar, rho_vbar,
// psi, each im x jm doubles [i][j], at scratch[sf_offset(im)]
void launch_streamfunction(const Batch& b, int member, int im, int jm, double two_pi_a);
int sf_offset(int im);
// run_3D_loop's mode-2 radiation on every member (MultiLayerRadiation and the omega_rad blend of
// t; outputs in the members' radiation, epsilon, albedo), then each member's zonal albedo sums
// per latitude at scratch[sf_offset(im)] (jm doubles). Grids up to radiation_max_levels() levels.
void launch_radiation(const Batch& b);
int radiation_max_levels();
// updateSkinTemperature on the device after launch_radiation, every member; the new t_skin is
// written into both descriptor arrays md0, md1
void launch_skin_step(const Batch& b, MemberDesc* md0, MemberDesc* md1, double relax);
// The two-category ice scheme on every member (initArrays, the column passes, extrapolateBC of
// the fluxes), then the damp_wiggles of P_rain and P_snow through scratch buffers 0-3. Members
// must hold their ice-scheme arrays (MemberDesc::mp). Scratch 6-9 are free for it at the place of
// the ice scheme inside the moist block (the damping has consumed them).
void launch_ice_scheme(const Batch& b, int gate);
// MoistConvection::run on every member (gate: the cadence test of the moist block, 0 = always)
void launch_moist_convection(const Batch& b, int gate = 0);
// precipitationSum, cap_S and the microphysics clamp that follow the ice scheme in the moist
// block; store != 0 also copies the fluxes to their stored arrays
void launch_moist_clamp_ice(const Batch& b, int gate, int store);
// boundary faces of S_c_c and the five sources on members that hold ice-scheme arrays
void launch_bc_sources(const Batch& b);
void launch_bc_conv(const Batch& b);
// P_rainn <- P_rain, P_snown <- P_snow on members that hold ice-scheme arrays.
void launch_ice_store(const Batch& b);
// one plane (n values) of one field, widened to double, at scratch[sf_offset(im)]: axis 0 fixes
// i ([j][k]), 1 fixes j ([i][k]), 2 fixes k ([i][j])
void launch_slice(const Batch& b, int member, int field, int axis, int index, long long n);
// src: add the microphysics sources of the ice-scheme arrays to the tendencies of t, c,
// cloud, ice (RHS_Atm_Turb.cpp:963-966, 1058-1067); every member must hold those arrays
void launch_rk4_jacobi(const Batch& b, bool fused, int src = 0);
void launch_rk4_lex(const Batch& b, int plane, int src = 0);
void launch_poisson(const Batch& b, int gate);
void launch_satadj(const Batch& b, int gate, int to_tmp);
// pass 0,1,2 over the fields of one job
void launch_damp(const Batch& b, int axis, const int* src, const int* dst, int nf, int moist_clamp, int gate);
// Fused three-pass damp_wiggles over nf = 3, 4 or 6 buffers with tile shape `shape` (one of
// damp_tile_shapes()) and ich levels per block; src and dst must differ.
// damp_wiggles of buffer src at level 0 only, into dst (see k_damp_level0)
void launch_damp_level0(const Batch& b, int src, int dst, int gate);
int damp_tile_shapes();
void damp_tile_shape(int shape, int* tj, int* tk);
void launch_damp_tile(const Batch& b, int shape, int ich, const int* src, const int* dst, int nf, int moist_clamp,
                      int gate);
void launch_copy_job(const Batch& b, const int* src, const int* dst, int nf, int gate = 0);
void launch_densities(const Batch& b, int set_water, int evaporate, int gate);
void launch_vel_star(const Batch& b, int gate);
void launch_turb_sources(const Batch& b, int gate);
void launch_turb_init_fields(const Batch& b);
void launch_wall_bc(const Batch& b);
void launch_bc(const Batch& b);
void launch_solid_clamp(const Batch& b);
void launch_polar(const Batch& b);
void launch_radial(const Batch& b, int store);
void launch_store(const Batch& b, int mask);
void launch_copy_field(const Batch& b, int dst, int src);
// project_initial_velocity around the Poisson sweeps: phase 0 seed (aux = velocity, p_dyn = 0),
// 1 gradient correction of the interior velocity, 2 clear p_dyn and aux
void launch_project(const Batch& b, int phase);
void launch_init_column(const Batch& b);
void launch_init_water(const Batch& b);
void launch_init_co2(const Batch& b);
void launch_copy_lid(const Batch& b);
void launch_init_rest(const Batch& b);

}  // namespace athad
