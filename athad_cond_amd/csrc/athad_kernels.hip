// HIP kernels for one ATHAD_COND time step on gfx950 (MI355X).
//
// Memory-bound stencils: fields are SoA, [i][j][k] with k (longitude) fastest, so a
// wavefront of 64 lanes reads 64 consecutive longitudes of one (i,j) row. Every kernel
// takes a device array of MemberDesc and uses blockIdx.z as the member, so the same
// launches serve one grid, an ensemble and the serving batch. Storage type S is double
// (FP64 drop-in) or float (mixed form); p_stat is always double, and the saturation,
// column and Newton arithmetic is always double.
//
// Kernels take the descriptor array as `const MemberDesc* __restrict__`: no kernel writes
// it, and without the qualifier the compiler must assume that a field store may change a
// descriptor, so it reloads descriptor fields after every store with vector loads whose
// wait also drains every load in flight.
//
// Each kernel cites the reference code it restates.
#include <cfloat>

#include "athad_device.h"
#include "athad_kernels.h"

namespace athad {

template <typename S> ATHAD_DEV S* fp(const MemberDesc& M, int f) { return (S*)M.f[f]; }
// ice-scheme array q of a member (allocated once the scheme has run)
template <typename S> ATHAD_DEV S* mpp(const MemberDesc& M, int q) { return (S*)M.mp[q]; }
ATHAD_DEV double* pstat(const MemberDesc& M) { return (double*)M.f[F_P_STAT]; }

// 1-D cell index of this thread for a kernel launched over a member's n cells with the
// XCD-aware block remap; -1 when out of range.
// Threads per block of every kernel that indexes cells through cell_id / cell_id_top_down.
// A constant rather than blockDim.x, which is read from the dispatch packet by a 16-bit vector
// load that every wave waits for before its first cell load can issue.
constexpr int kCellBlock = 256;

ATHAD_DEV long long cell_id(const Consts& K) {
    const long long nb = (K.n + kCellBlock - 1) / kCellBlock;
    const long long b = xcd_block(blockIdx.x, gridDim.x, nb);
    if (b < 0) return -1;
    const long long id = b * kCellBlock + threadIdx.x;
    return (id < K.n) ? id : -1;
}

// 1-D cell index for a pointwise kernel whose cost per cell varies strongly with the level
// (the saturation adjustment: mixed-phase cells of the upper levels run all 20 Newton
// iterations, most others stop after 4 or 5). No XCD remap: the hardware deals consecutive
// blocks round-robin over the XCDs, so every XCD gets an equal share of every level, where
// the remap would hand whole levels, and so all of the slow ones, to one or two XCDs. The
// blocks run from the top level down, so the slow cells start first and the cheap lower
// levels fill in behind them. -1 when out of range.
ATHAD_DEV long long cell_id_top_down(const Consts& K) {
    const long long nb = (K.n + kCellBlock - 1) / kCellBlock;
    if ((long long)blockIdx.x >= nb) return -1;
    const long long id = (nb - 1 - (long long)blockIdx.x) * kCellBlock + threadIdx.x;
    return (id < K.n) ? id : -1;
}

// (i,j,k) of a cell index; a member has fewer than 2^31 cells (checked at creation), so
// 32-bit unsigned division suffices.
ATHAD_DEV void decode(const Consts& K, long long id, int& i, int& j, int& k) {
    const unsigned u = (unsigned)id, km = (unsigned)K.km, jm = (unsigned)K.jm;
    const unsigned ij = u / km;
    k = (int)(u - ij * km);
    const unsigned ii = ij / jm;
    j = (int)(ij - ii * jm);
    i = (int)ii;
}

// A load through the constant address space: at a wave-uniform address it is a scalar load,
// waited for by lgkmcnt, so it does not drain the vector loads in flight the way a late vector
// load waited for by vmcnt(0) does; at a per-lane address it stays a vector load. Only for
// memory that no thread of the running kernel writes: the iteration counters (k_advance, a
// launch of its own, is their only writer) and the per-level tables.
template <typename T> ATHAD_DEV T ro_load(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}

template <typename T> using G1 = __attribute__((address_space(1))) T;   // global-memory object

// Cadence gate: a kernel launched with gate g > 0 does nothing for members whose iteration
// counter is not a multiple of g (Poisson every 4th, moist block every 2nd iteration,
// cAtmosphereModel.cpp:1156-1185), so members at different phases can share launches.
ATHAD_DEV bool gated_off(const MemberDesc& M, int gate) {
    return gate > 0 && (ro_load(M.counters) % gate) != 0;
}

// ------------------------------------------------------------------------------------
// Step counter (device resident so a captured graph can replay steps)
// ------------------------------------------------------------------------------------
__global__ void k_advance(MemberDesc* md) {
    if (threadIdx.x == 0) {
        int* c = md[blockIdx.z].counters;
        c[0] += 1;
        c[1] += 1;
    }
}

// ------------------------------------------------------------------------------------
// Level-mean temperature (RungeKutta_Atm_Turb.cpp:24-41). One block per (level, member);
// fixed reduction tree, so the result is deterministic run to run.
// ------------------------------------------------------------------------------------
constexpr int kLevelChunks = 32;

// pass 1: block (chunk c, level i) sums a fixed contiguous chunk of the level in a fixed tree
template <typename S>
__global__ __launch_bounds__(256) void k_level_mean(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const int i = blockIdx.y, c = blockIdx.x;
    if (i >= K.im) return;
    const S* t = fp<S>(M, F_T);
    __shared__ double ss[256], sw[256];
    const int njk = K.jm * K.km;
    const int per = (njk + kLevelChunks - 1) / kLevelChunks;
    const int q0 = c * per, q1 = min(njk, q0 + per);
    double sum = 0.0, wsum = 0.0;
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const double w = M.sinr[q / K.km];
        sum += w * (double)t[(size_t)i * njk + q];
        wsum += w;
    }
    ss[threadIdx.x] = sum;
    sw[threadIdx.x] = wsum;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) { ss[threadIdx.x] += ss[threadIdx.x + st]; sw[threadIdx.x] += sw[threadIdx.x + st]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        M.scratch[2 * (i * kLevelChunks + c)] = ss[0];
        M.scratch[2 * (i * kLevelChunks + c) + 1] = sw[0];
    }
}

// pass 2: ordered sum of the chunks
__global__ void k_level_mean_final(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i >= M.K.im) return;
    double sum = 0.0, wsum = 0.0;
    for (int c = 0; c < kLevelChunks; c++) {
        sum += M.scratch[2 * (i * kLevelChunks + c)];
        wsum += M.scratch[2 * (i * kLevelChunks + c) + 1];
    }
    M.t_ref[i] = (wsum > 0.0) ? sum / wsum : 1.0;
}

// Sum and sum of squares of one field of member `member` in a fixed order (kStatChunks
// contiguous chunks, each reduced by a fixed tree, then the chunks in order), so equal
// fields give equal bits wherever and whenever they are computed. Results at
// scratch[kStatBase(im) ...].
constexpr int kStatChunks = 256;
ATHAD_DEV int stat_base(const Consts& K) { return 2 * kLevelChunks * K.im; }

template <typename S>
__global__ __launch_bounds__(256) void k_field_stats(const MemberDesc* __restrict__ md, int member, int field) {
    const MemberDesc& M = md[member];
    const Consts& K = M.K;
    const long long n = K.n;
    const long long per = (n + kStatChunks - 1) / kStatChunks;
    const long long q0 = (long long)blockIdx.x * per, q1 = min(n, q0 + per);
    const bool dbl = (field == F_P_STAT) || sizeof(S) == 8;
    __shared__ double ss[256], sq[256];
    double a = 0.0, b = 0.0;
    for (long long q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const double v = dbl ? ((const double*)M.f[field])[q] : (double)((const float*)M.f[field])[q];
        a += v;
        b += v * v;
    }
    ss[threadIdx.x] = a;
    sq[threadIdx.x] = b;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) { ss[threadIdx.x] += ss[threadIdx.x + st]; sq[threadIdx.x] += sq[threadIdx.x + st]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        M.scratch[stat_base(K) + 2 * blockIdx.x] = ss[0];
        M.scratch[stat_base(K) + 2 * blockIdx.x + 1] = sq[0];
    }
}

__global__ void k_field_stats_final(const MemberDesc* __restrict__ md, int member) {
    const MemberDesc& M = md[member];
    double a = 0.0, b = 0.0;
    const double* r = M.scratch + stat_base(M.K);
    for (int c = 0; c < kStatChunks; c++) { a += r[2 * c]; b += r[2 * c + 1]; }
    M.scratch[stat_base(M.K) + 2 * kStatChunks] = a;
    M.scratch[stat_base(M.K) + 2 * kStatChunks + 1] = b;
}

// ------------------------------------------------------------------------------------
// Checkpoint diagnostics on the device (the reference prints them from host loops over the
// fields): searchMinMax_3D (MinMax_Atm.cpp:51-108) and findResiduumAtm (UtilsAtm.h:70-218).
// kStatChunks contiguous chunks, each reduced by a fixed tree, then the chunks in order, in a
// scratch region after the field statistics. Positions travel as doubles (exact below 2^53);
// -1 is "none".
// ------------------------------------------------------------------------------------
ATHAD_DEV int diag_base(const Consts& K) { return stat_base(K) + 2 * kStatChunks + 2; }

// The reference scans in row-major order with strict comparisons against +-DBL_MAX starting
// values, so among equal extremes the first occurrence, the smallest index, is kept, and a value
// equal to the starting bound is never taken. findResiduumAtm keeps a maximum only when it
// exceeds 0 and breaks ties by the smallest (i, j, k), which is the smallest index.
ATHAD_DEV bool keeps_max(double v, double q, double bv, double bq) { return bq < 0.0 || v > bv || (v == bv && q < bq); }
ATHAD_DEV bool keeps_min(double v, double q, double bv, double bq) { return bq < 0.0 || v < bv || (v == bv && q < bq); }

template <typename S>
__global__ __launch_bounds__(256) void k_minmax(const MemberDesc* __restrict__ md, int member, int field) {
    const MemberDesc& M = md[member];
    const Consts& K = M.K;
    const long long n = K.n;
    const long long per = (n + kStatChunks - 1) / kStatChunks;
    const long long q0 = (long long)blockIdx.x * per, q1 = min(n, q0 + per);
    const bool dbl = (field == F_P_STAT) || sizeof(S) == 8;
    __shared__ double s_lo[256], s_lq[256], s_hi[256], s_hq[256];
    double lo = 0.0, lq = -1.0, hi = 0.0, hq = -1.0;
    for (long long q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const double v = dbl ? ((const double*)M.f[field])[q] : (double)((const float*)M.f[field])[q];
        if (!finite_bits(v)) continue;   // searchMinMax_3D skips non-finite values
        if (v < DBL_MAX && keeps_min(v, (double)q, lo, lq)) { lo = v; lq = (double)q; }
        if (v > -DBL_MAX && keeps_max(v, (double)q, hi, hq)) { hi = v; hq = (double)q; }
    }
    s_lo[threadIdx.x] = lo; s_lq[threadIdx.x] = lq; s_hi[threadIdx.x] = hi; s_hq[threadIdx.x] = hq;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        const int t = threadIdx.x;
        if (t < st) {
            if (s_lq[t + st] >= 0.0 && keeps_min(s_lo[t + st], s_lq[t + st], s_lo[t], s_lq[t])) {
                s_lo[t] = s_lo[t + st]; s_lq[t] = s_lq[t + st];
            }
            if (s_hq[t + st] >= 0.0 && keeps_max(s_hi[t + st], s_hq[t + st], s_hi[t], s_hq[t])) {
                s_hi[t] = s_hi[t + st]; s_hq[t] = s_hq[t + st];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* r = M.scratch + diag_base(K) + 4 * blockIdx.x;
        r[0] = s_lo[0]; r[1] = s_lq[0]; r[2] = s_hi[0]; r[3] = s_hq[0];
    }
}

__global__ void k_minmax_final(const MemberDesc* __restrict__ md, int member) {
    const MemberDesc& M = md[member];
    const double* r = M.scratch + diag_base(M.K);
    double lo = 0.0, lq = -1.0, hi = 0.0, hq = -1.0;
    for (int c = 0; c < kStatChunks; c++) {
        if (r[4 * c + 1] >= 0.0 && keeps_min(r[4 * c], r[4 * c + 1], lo, lq)) { lo = r[4 * c]; lq = r[4 * c + 1]; }
        if (r[4 * c + 3] >= 0.0 && keeps_max(r[4 * c + 2], r[4 * c + 3], hi, hq)) { hi = r[4 * c + 2]; hq = r[4 * c + 3]; }
    }
    double* o = M.scratch + diag_base(M.K) + 4 * kStatChunks;
    o[0] = lo; o[1] = lq; o[2] = hi; o[3] = hq;
}

// findResiduumAtm: over i in [10, 30], j in [1, jm-2], k in [1, km-2] the residual
// sqrt((du/dr^2 + dv/dtheta^2 + dw/dphi^2) / 3) of the centred velocity derivatives (the metric
// floor sin(theta) >= 0.55 as in the RHS); its maximum (ties: smallest (i,j,k)), the sum and
// count of the finite residuals and the count and first position of the non-finite ones. The
// reference's sum runs in its thread order; this one in a fixed order.
template <typename S>
__global__ __launch_bounds__(256) void k_residuum(const MemberDesc* __restrict__ md, int member) {
    const MemberDesc& M = md[member];
    const Consts& K = M.K;
    const long long nj = K.jm - 2, nk = K.km - 2, nbox = 21LL * nj * nk;
    const long long per = (nbox + kStatChunks - 1) / kStatChunks;
    const long long b0 = (long long)blockIdx.x * per, b1 = min(nbox, b0 + per);
    const S *u = fp<S>(M, F_U), *v = fp<S>(M, F_V), *w = fp<S>(M, F_W);
    const long long si = (long long)K.jm * K.km, sj = K.km;
    __shared__ double s_mx[256], s_mq[256], s_sum[256], s_cnt[256], s_nan[256], s_nq[256];
    double mx = 0.0, mq = -1.0, sum = 0.0, cnt = 0.0, nan = 0.0, nq = -1.0;
    for (long long b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
        const int i = 10 + (int)(b / (nj * nk));
        const long long r = b % (nj * nk);
        const int j = 1 + (int)(r / nk), k = 1 + (int)(r % nk);
        const long long q = i * si + (long long)j * sj + k;
        const double sinthe = M.sint[j];
        const double rm = M.rad[i];
        const double exp_rm = 1.0 / (rm + 1.0);
        const double rmsinthe = rm * sinthe;
        const double dudr = ((double)u[q + si] - (double)u[q - si]) / (2.0 * K.dr) * exp_rm;
        const double dvdthe = ((double)v[q + sj] - (double)v[q - sj]) / (2.0 * rm * K.dthe);
        const double dwdphi = ((double)w[q + 1] - (double)w[q - 1]) / (2.0 * rmsinthe * K.dphi);
        const double res = sqrt((dudr * dudr + dvdthe * dvdthe + dwdphi * dwdphi) / 3.0);
        if (!finite_bits(res)) {
            if (nq < 0.0) nq = (double)q;   // this thread visits increasing q
            nan += 1.0;
            continue;
        }
        sum += res;
        cnt += 1.0;
        if (res > 0.0 && keeps_max(res, (double)q, mx, mq)) { mx = res; mq = (double)q; }
    }
    const int t = threadIdx.x;
    s_mx[t] = mx; s_mq[t] = mq; s_sum[t] = sum; s_cnt[t] = cnt; s_nan[t] = nan; s_nq[t] = nq;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st) {
            if (s_mq[t + st] >= 0.0 && keeps_max(s_mx[t + st], s_mq[t + st], s_mx[t], s_mq[t])) {
                s_mx[t] = s_mx[t + st]; s_mq[t] = s_mq[t + st];
            }
            s_sum[t] += s_sum[t + st];
            s_cnt[t] += s_cnt[t + st];
            s_nan[t] += s_nan[t + st];
            if (s_nq[t + st] >= 0.0 && (s_nq[t] < 0.0 || s_nq[t + st] < s_nq[t])) s_nq[t] = s_nq[t + st];
        }
        __syncthreads();
    }
    if (t == 0) {
        double* o = M.scratch + diag_base(K) + 6 * blockIdx.x;
        o[0] = s_mx[0]; o[1] = s_mq[0]; o[2] = s_sum[0]; o[3] = s_cnt[0]; o[4] = s_nan[0]; o[5] = s_nq[0];
    }
}

__global__ void k_residuum_final(const MemberDesc* __restrict__ md, int member) {
    const MemberDesc& M = md[member];
    const double* r = M.scratch + diag_base(M.K);
    double mx = 0.0, mq = -1.0, sum = 0.0, cnt = 0.0, nan = 0.0, nq = -1.0;
    for (int c = 0; c < kStatChunks; c++) {
        const double* x = r + 6 * c;
        if (x[1] >= 0.0 && keeps_max(x[0], x[1], mx, mq)) { mx = x[0]; mq = x[1]; }
        sum += x[2];
        cnt += x[3];
        nan += x[4];
        if (x[5] >= 0.0 && (nq < 0.0 || x[5] < nq)) nq = x[5];
    }
    double* o = M.scratch + diag_base(M.K) + 6 * kStatChunks;
    o[0] = mx; o[1] = mq; o[2] = sum; o[3] = cnt; o[4] = nan; o[5] = nq;
}

// write_meridional_streamfunction (MinMax_Atm.cpp:173-227): zonal means of v and r_humid v
// over the fluid cells of each (i, j) (every cell: the configuration has no land), then the
// mass streamfunction integrated down from the lid. The sums run in k order and the integral in
// i order, as the reference's loops, with contraction off, so the three arrays are the
// reference's to the last bit. Arrays [i][j] at sf_base: vbar, rho_vbar, psi.
ATHAD_DEV int sf_base(const Consts& K) { return diag_base(K) + 6 * kStatChunks + 8; }

template <typename S>
__global__ __launch_bounds__(64) void k_zonal_mean_v(const MemberDesc* __restrict__ md, int member) {
#pragma clang fp contract(off)
    const MemberDesc& M = md[member];
    const Consts& K = M.K;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;   // i * jm + j
    if (id >= K.im * K.jm) return;
    const long long row = (long long)id * K.km;
    const S *v = fp<S>(M, F_V), *rh = fp<S>(M, F_R_HUMID);
    double sum = 0.0, rsum = 0.0;
    int n = 0;
    for (int k = 0; k < K.km; k++) {
        const double vv = (double)v[row + k];
        double rho = (double)rh[row + k];
        if (!finite_bits(vv)) continue;
        if (!finite_bits(rho) || rho <= 0.0) rho = K.r_air;   // before the first densities()
        sum += vv;
        rsum += rho * vv;
        n++;
    }
    double* o = M.scratch + sf_base(K);
    o[id] = (n > 0) ? (sum / n) * K.u_0 : 0.0;
    o[K.im * K.jm + id] = (n > 0) ? (rsum / n) * K.u_0 : 0.0;
}

// One thread per latitude: psi(im-1) = 0, psi(i) = psi(i+1) + 2 pi a sin(theta) rvm dz with the
// layer heights' float difference as dz, as get_layer_height returns float.
__global__ void k_streamfunction(const MemberDesc* __restrict__ md, int member, double two_pi_a) {
#pragma clang fp contract(off)
    const MemberDesc& M = md[member];
    const Consts& K = M.K;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= K.jm) return;
    const long long nij = (long long)K.im * K.jm;
    const double* rv = M.scratch + sf_base(K) + nij;
    double* psi = M.scratch + sf_base(K) + 2 * nij;
    const double coeff = two_pi_a * M.sinr[j];
    double acc = 0.0;
    psi[(long long)(K.im - 1) * K.jm + j] = 0.0;
    for (int i = K.im - 2; i >= 0; i--) {
        const double dz = (double)((float)M.hgt[i + 1] - (float)M.hgt[i]);
        const double rvm = 0.5 * (rv[(long long)i * K.jm + j] + rv[(long long)(i + 1) * K.jm + j]);
        acc = acc + coeff * rvm * dz;
        psi[(long long)i * K.jm + j] = acc;
    }
}

// One plane of one field, widened to double, at sf_base: the level i = index ([j][k], the
// ParaView radial slice and the transfer file's surface), the meridian k = index ([i][j], zonal)
// or the latitude j = index ([i][k], longal); Paraview_Atm.cpp:407-1000, FileIO_Atm.cpp:64-120.
template <typename S>
__global__ __launch_bounds__(256) void k_slice(const MemberDesc* __restrict__ md, int member, int field, int axis,
                                               int index) {
    const MemberDesc& M = md[member];
    const Consts& K = M.K;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long n, q;
    if (axis == 0) {
        n = (long long)K.jm * K.km;
        q = (long long)index * n + id;
    } else if (axis == 1) {
        n = (long long)K.im * K.km;
        q = ((id / K.km) * K.jm + index) * K.km + id % K.km;
    } else {
        n = (long long)K.im * K.jm;
        q = id * K.km + index;
    }
    if (id >= n) return;
    const bool dbl = (field == F_P_STAT) || sizeof(S) == 8;
    M.scratch[sf_base(K) + id] = dbl ? ((const double*)M.f[field])[q] : (double)((const float*)M.f[field])[q];
}

// ------------------------------------------------------------------------------------
// Radiation in run_3D_loop's mode 2 (cAtmosphereModel.cpp:799-801, 863-879):
// MultiLayerRadiation::run (MultiLayerRadiation.h:59-384) on the current t, and t blended by
// omega_rad toward its result. A block of 256 threads takes kRadCols consecutive columns of
// one latitude row: the six input fields of those columns are staged through LDS with
// coalesced loads ([field][i][col]), each wave then takes one column at a time with E levels
// per lane (lane l holds levels lE .. lE+E-1), and the outputs go back through the same LDS
// slots with coalesced stores. The two grey flux sweeps
//   up[i] = up[i-1] (1 - eps_i) + eps_i sigma T_i^4,   dn[i] = dn[i+1] (1 - eps_i) + eps_i sigma T_i^4
// are affine recurrences: each level is the map x -> a x + b, and the sweep is the running
// composition of those maps, formed by a Hillis-Steele scan across the 64 lanes (6 steps)
// instead of 41 dependent steps. Every term is non-negative, so the scan agrees with the
// sequential loop to a few ulps. T^4 is (T^2)^2 and x^(1/4) is sqrt(sqrt(x)) (correctly
// rounded square roots) in place of pow. The condensate paths that set the albedo are summed
// in level order, as the reference, so the albedo is its own to the last bit for equal inputs.
// Contraction is off in this section: the cell arithmetic keeps the reference's roundings.
// ------------------------------------------------------------------------------------
constexpr int kRadCols = 16;

// Inclusive composition scan across the 64 lanes: after it, lane l holds the composition of
// the maps of lanes 0..l (upward) or l..63 (downward), the nearer lane's map applied last. A
// step composes B = A B' + B and A = A A' with the farther part (A', B'). The scan runs within
// each 16-lane row with DPP moves (shift by 1, 2, 4, 8 inside the row), then joins the rows
// through the row totals read from lanes 15, 31, 47 (upward) or 16, 32, 48 (downward); lanes
// whose DPP source lies outside their row keep the identity (1, 0), an exact no-op. The A part
// depends only on the emissivities, fixed for a column, so its per-step values and the rows'
// totals are formed once (rad_prep) and each of the five sweeps moves only B. All terms are
// non-negative, so the grouping agrees with the sequential recurrence to a few ulps.
constexpr int kDppRowShr = 0x110, kDppRowShl = 0x100, kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;

// 64-bit DPP move; lanes whose source is invalid (outside the row or the wave) get `old`.
template <int CTRL> ATHAD_DEV double dpp_move(double v, double old) {
    const long long x = __double_as_longlong(v), o = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)x, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
ATHAD_DEV double read_lane(double v, int lane) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)x, lane);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL> ATHAD_DEV float dpp_move(float v, float old) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
ATHAD_DEV float read_lane(float v, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); }

struct RadScan {
    double As_up[4], Af_up, a15, a31, a47;   // upward: own A before each row step, after them, row totals
    double As_dn[4], Af_dn, a16, a32, a48;   // downward
};

ATHAD_DEV void scan_A_up(double A, RadScan& R) {
#pragma clang fp contract(off)
    R.As_up[0] = A; A = A * dpp_move<kDppRowShr + 1>(A, 1.0);
    R.As_up[1] = A; A = A * dpp_move<kDppRowShr + 2>(A, 1.0);
    R.As_up[2] = A; A = A * dpp_move<kDppRowShr + 4>(A, 1.0);
    R.As_up[3] = A; A = A * dpp_move<kDppRowShr + 8>(A, 1.0);
    R.Af_up = A;
    R.a15 = read_lane(A, 15); R.a31 = read_lane(A, 31); R.a47 = read_lane(A, 47);
}
ATHAD_DEV void scan_A_down(double A, RadScan& R) {
#pragma clang fp contract(off)
    R.As_dn[0] = A; A = A * dpp_move<kDppRowShl + 1>(A, 1.0);
    R.As_dn[1] = A; A = A * dpp_move<kDppRowShl + 2>(A, 1.0);
    R.As_dn[2] = A; A = A * dpp_move<kDppRowShl + 4>(A, 1.0);
    R.As_dn[3] = A; A = A * dpp_move<kDppRowShl + 8>(A, 1.0);
    R.Af_dn = A;
    R.a16 = read_lane(A, 16); R.a32 = read_lane(A, 32); R.a48 = read_lane(A, 48);
}
ATHAD_DEV double scan_B_up(double B, const RadScan& R) {
#pragma clang fp contract(off)
    const int row = (threadIdx.x & 63) >> 4;
    B = R.As_up[0] * dpp_move<kDppRowShr + 1>(B, 0.0) + B;
    B = R.As_up[1] * dpp_move<kDppRowShr + 2>(B, 0.0) + B;
    B = R.As_up[2] * dpp_move<kDppRowShr + 4>(B, 0.0) + B;
    B = R.As_up[3] * dpp_move<kDppRowShr + 8>(B, 0.0) + B;
    // prefixes of the rows below: row 0, rows 0-1, rows 0-2
    const double b15 = read_lane(B, 15), b31 = read_lane(B, 31), b47 = read_lane(B, 47);
    const double q1 = b15, q2 = R.a31 * q1 + b31, q3 = R.a47 * q2 + b47;
    if (row > 0) B = R.Af_up * ((row == 1) ? q1 : (row == 2) ? q2 : q3) + B;
    return B;
}
ATHAD_DEV double scan_B_down(double B, const RadScan& R) {
#pragma clang fp contract(off)
    const int row = (threadIdx.x & 63) >> 4;
    B = R.As_dn[0] * dpp_move<kDppRowShl + 1>(B, 0.0) + B;
    B = R.As_dn[1] * dpp_move<kDppRowShl + 2>(B, 0.0) + B;
    B = R.As_dn[2] * dpp_move<kDppRowShl + 4>(B, 0.0) + B;
    B = R.As_dn[3] * dpp_move<kDppRowShl + 8>(B, 0.0) + B;
    // suffixes of the rows above: row 3, rows 2-3, rows 1-3
    const double b16 = read_lane(B, 16), b32 = read_lane(B, 32), b48 = read_lane(B, 48);
    const double q2 = b48, q1 = R.a32 * q2 + b32, q0 = R.a16 * q1 + b16;
    if (row < 3) B = R.Af_dn * ((row == 2) ? q2 : (row == 1) ? q1 : q0) + B;
    return B;
}

// The per-level maps of the two sweeps (MultiLayerRadiation.h:291-303): up from the black
// surface (level 0 is the constant map sigma T_0^4), dn from a dark sky (level 0 is the
// identity, so dn at level 0 equals dn[1], the downwelling flux at the ground). Levels at or
// above im are identities. a_up / a_dn: the slopes; the scans' A parts.
template <int E>
ATHAD_DEV void rad_prep(int lane, int im, const double (&eps)[E], double (&a_up)[E], double (&a_dn)[E], RadScan& R) {
#pragma clang fp contract(off)
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int i = lane * E + e;
        a_up[e] = (i >= im) ? 1.0 : (i == 0) ? 0.0 : 1.0 - eps[e];
        a_dn[e] = (i == 0) ? 1.0 : a_up[e];
    }
    double A = 1.0;
#pragma unroll
    for (int e = 0; e < E; e++) A = a_up[e] * A;
    scan_A_up(A, R);
    A = 1.0;
#pragma unroll
    for (int e = E - 1; e >= 0; e--) A = a_dn[e] * A;
    scan_A_down(A, R);
}

// Both sweeps of one column for the current sigma T^4.
template <int E>
ATHAD_DEV void rad_sweeps(int lane, int im, const double (&eps)[E], const double (&P4)[E], double sigma,
                          const double (&a_up)[E], const double (&a_dn)[E], const RadScan& R,
                          double (&up)[E], double (&dn)[E]) {
#pragma clang fp contract(off)
    double b[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int i = lane * E + e;
        b[e] = (i >= im) ? 0.0 : (i == 0) ? sigma * P4[e] : eps[e] * sigma * P4[e];
    }
    double B = 0.0;
#pragma unroll
    for (int e = 0; e < E; e++) B = a_up[e] * B + b[e];
    B = scan_B_up(B, R);
    double x = dpp_move<kDppWaveShr1>(B, 0.0);   // the lane below's prefix; 0 for lane 0
#pragma unroll
    for (int e = 0; e < E; e++) { x = a_up[e] * x + b[e]; up[e] = x; }
    if (lane == 0) b[0] = 0.0;
    B = 0.0;
#pragma unroll
    for (int e = E - 1; e >= 0; e--) B = a_dn[e] * B + b[e];
    B = scan_B_down(B, R);
    x = dpp_move<kDppWaveShl1>(B, 0.0);          // the lane above's suffix; 0 for lane 63
#pragma unroll
    for (int e = E - 1; e >= 0; e--) { x = a_dn[e] * x + b[e]; dn[e] = x; }
}

template <typename S, int E>
__global__ __launch_bounds__(256) void k_radiation(const MemberDesc* __restrict__ md) {
#pragma clang fp contract(off)
    extern __shared__ double s_col[];
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const int im = K.im, jm = K.jm, km = K.km, top = im - 1;
    const int j = blockIdx.y, k0 = blockIdx.x * kRadCols;
    if (j >= jm || k0 >= km) return;
    // global-memory pointers: through plain pointers the reads are flat loads, which the LDS
    // stores of the staging below would wait for one at a time (both count on lgkmcnt)
    typedef const G1<S>* __restrict__ FieldP;
    const FieldP ft = (FieldP)fp<S>(M, F_T), fc = (FieldP)fp<S>(M, F_C), fco2 = (FieldP)fp<S>(M, F_CO2);
    const FieldP fcl = (FieldP)fp<S>(M, F_CLOUD), fice = (FieldP)fp<S>(M, F_ICE);
    const G1<double>* __restrict__ fpst = (const G1<double>*)M.f[F_P_STAT];
    const long long si = (long long)jm * km, row = (long long)j * km;
    auto slot = [&](int f, int i, int kk) -> double& { return s_col[(f * im + i) * kRadCols + kk]; };
    // stage t, c, co2, p_stat, cloud, ice of the tile's columns: thread (i mod 16, kk) of the
    // block walks the levels 16 apart, so each load covers 16 neighbouring longitudes
    static_assert(kRadCols == 16, "the staging maps 256 threads onto 16 levels x 16 columns");
    const int kk_s = threadIdx.x & 15, i_s = threadIdx.x >> 4;
    const bool col_in = k0 + kk_s < km;
    for (int i = i_s; i < im; i += 16) {
        if (!col_in) break;
        const long long g = i * si + row + k0 + kk_s;
        const double v0 = (double)ft[g], v1 = (double)fc[g], v2 = (double)fco2[g], v3 = fpst[g];
        const double v4 = (double)fcl[g], v5 = (double)fice[g];
        slot(0, i, kk_s) = v0; slot(1, i, kk_s) = v1; slot(2, i, kk_s) = v2;
        slot(3, i, kk_s) = v3; slot(4, i, kk_s) = v4; slot(5, i, kk_s) = v5;
    }
    __syncthreads();
    // incoming short wave of this latitude (MultiLayerRadiation.h:98-104)
    const int jmax = jm - 1, jhalf = jmax / 2;
    const double xs = (double)((j <= jhalf) ? j : jmax - j) / (double)jhalf;
    const double sw = (K.sw_pole - K.sw_equator) * (xs * xs - 2.0 * xs) + K.sw_pole;
    const double sigma = K.sigma, inv_g = 1.0 / K.g, t_0 = K.t_0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int kk = wave; kk < kRadCols; kk += 4) {
        if (k0 + kk >= km) break;
        double T[E], eps[E], P4[E], up[E], dn[E], t_in[E], lw[E], iw[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int i = lane * E + e;
            eps[e] = 0.0; lw[e] = 0.0; iw[e] = 0.0; T[e] = 0.0; t_in[e] = 0.0;
            if (i < im) {
                // optical depth: gas mass with pressure broadening, cloud liquid and ice paths
                // (MultiLayerRadiation.h:156-207)
                t_in[e] = slot(0, i, kk);
                T[e] = t_in[e] * t_0;
                const double p = slot(3, i, kk);
                double dp = (i < top) ? (p - slot(3, i + 1, kk)) * 100.0 : p * 100.0;
                if (dp < 0.0) dp = 0.0;
                const double qv = tmax(0.0, slot(1, i, kk)), qc = tmax(0.0, slot(2, i, kk));
                const double qb = tmax(0.0, 1.0 - qv - qc);
                const double mass = dp * inv_g;
                const double broad = p * 100.0 / 1.0e5;
                const double tau_gas = (K.kappa_h2o * qv + K.kappa_co2 * qc + K.kappa_bg * qb) * mass * broad;
                const int ia = (i < top) ? i : i - 1;
                const double dz = (double)((float)ro_load(M.hgt + ia + 1) - (float)ro_load(M.hgt + ia));
                const double rho = (T[e] > 0.0) ? (p * 100.0) / (th_R_of(qv, qc, K.R_bg) * T[e]) : 0.0;
                lw[e] = tmax(0.0, slot(4, i, kk)) * rho * dz * 1000.0;
                iw[e] = tmax(0.0, slot(5, i, kk)) * rho * dz * 1000.0;
                const double tau = tau_gas + 0.12 * lw[e] + 0.055 * iw[e];
                eps[e] = (tau > 700.0) ? 1.0 : (1.0 - exp(-tau));
            }
            P4[e] = (T[e] * T[e]) * (T[e] * T[e]);
        }
        // condensate paths in level order, then the cloud-raised albedo (MultiLayerRadiation.h:231-238):
        // every lane leaves its levels' paths in the column's cloud and ice slots (read above,
        // not needed again) and lane 0 sums them in level order; the albedo is lane 0's only
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int i = lane * E + e;
            if (i < im) { slot(4, i, kk) = lw[e]; slot(5, i, kk) = iw[e]; }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double alb = K.albedo_surface;
        if (lane == 0) {
            double lwp = 0.0, iwp = 0.0;
            for (int i = 0; i < im; i++) {
                lwp += slot(4, i, kk);
                iwp += slot(5, i, kk);
            }
            if (K.albedo_cloud > alb) {
                const double tau = (lwp + 0.5 * iwp) / 100.0;
                alb = alb + (K.albedo_cloud - alb) * (tau / (tau + 2.0));
            }
        }
        // four Lambda iterations of sigma T_i^4 = (up[i-1] + dn[i+1]) / 2, then final sweeps
        // (MultiLayerRadiation.h:287-333)
        double a_up[E], a_dn[E];
        RadScan R;
        rad_prep<E>(lane, im, eps, a_up, a_dn, R);
        for (int it = 0; it < 4; it++) {
            rad_sweeps<E>(lane, im, eps, P4, sigma, a_up, a_dn, R, up, dn);
            const double up_prev_lane = dpp_move<kDppWaveShr1>(up[E - 1], 0.0);
            const double dn_next_lane = dpp_move<kDppWaveShl1>(dn[0], 0.0);   // 0 for lane 63
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int i = lane * E + e;
                if (i < 1 || i > top) continue;
                const double below = (e > 0) ? up[e > 0 ? e - 1 : 0] : up_prev_lane;
                const double above = (e < E - 1) ? dn[e < E - 1 ? e + 1 : 0] : dn_next_lane;
                const double emit = 0.5 * (below + above);
                if (emit > 0.0 && finite_bits(emit)) {
                    T[e] = sqrt(sqrt(emit / sigma));
                    P4[e] = (T[e] * T[e]) * (T[e] * T[e]);
                }
            }
        }
        rad_sweeps<E>(lane, im, eps, P4, sigma, a_up, a_dn, R, up, dn);
        // surface energy balance with sigma T^4 linearised at the current surface temperature
        // (MultiLayerRadiation.h:338-371); dn at level 0 is dn[1]
        const double T1 = (E > 1) ? T[E > 1 ? 1 : 0] : dpp_move<kDppWaveShl1>(T[0], 0.0);
        double Ts = 0.0;
        if (lane == 0) {
            const double Ts0 = T[0], c_H = 15.0;
            const double d4 = 4.0 * sigma * Ts0 * Ts0 * Ts0;
            Ts = ((1.0 - alb) * sw + K.geothermal + dn[0] - sigma * P4[0] + d4 * Ts0 + c_H * T1) / (d4 + c_H);
            M.albedo[row + k0 + kk] = alb;
        }
        // outputs into the column's LDS slots: blended t, epsilon, upward flux
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int i = lane * E + e;
            if (i >= im) continue;
            const double Tf = (i == 0) ? Ts : T[e];
            slot(0, i, kk) = t_in[e] * (1.0 - K.omega_rad) + (Tf / t_0) * K.omega_rad;
            slot(1, i, kk) = eps[e];
            slot(2, i, kk) = (i == 0) ? sigma * ((Ts * Ts) * (Ts * Ts)) : up[e];
        }
    }
    __syncthreads();
    S* tw = fp<S>(M, F_T);
    for (int i = i_s; i < im; i += 16) {
        if (!col_in) break;
        const long long g = i * si + row + k0 + kk_s;
        tw[g] = (S)slot(0, i, kk_s);
        M.epsilon[g] = slot(1, i, kk_s);
        M.radiation[g] = slot(2, i, kk_s);
    }
}

// Per latitude, the zonal sum of the albedo in k order (planetaryShortWave's inner loop,
// cAtmosphereModel.cpp:248-252), at sf_base. One block per latitude stages the row in LDS with
// coalesced loads, then one thread adds it up in k order.
constexpr int kAlbedoRowMax = 4096;
__global__ __launch_bounds__(256) void k_albedo_rows(const MemberDesc* __restrict__ md) {
#pragma clang fp contract(off)
    __shared__ double s_row[kAlbedoRowMax];
    const MemberDesc& M = md[blockIdx.z];
    const int j = blockIdx.x, km = M.K.km;
    if (j >= M.K.jm) return;
    const double* row = M.albedo + (long long)j * km;
    double a = 0.0;
    for (int k0 = 0; k0 < km; k0 += kAlbedoRowMax) {   // rows longer than the stage in pieces
        const int n = min(kAlbedoRowMax, km - k0);
        for (int q = threadIdx.x; q < n; q += blockDim.x) s_row[q] = row[k0 + q];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < n; q++) a += s_row[q];
        __syncthreads();
    }
    if (threadIdx.x == 0) M.scratch[sf_base(M.K) + j] = a;
}

// With the ice scheme in the moist block, after its flux damping (cAtmosphereModel.cpp:1263-1399):
// precipitationSum (UtilsAtm.h:52-67, graupel and the moist convection's P_conv zero), cap_S
// of the sources and of S_c_c, the microphysics clamp and the water-vapour ceiling.
// store != 0: also P_rainn <- P_rain, P_snown <- P_snow, the iteration's store of the fluxes
// (nothing after this point of the iteration changes them).
template <typename S>
__global__ __launch_bounds__(256) void k_moist_clamp_ice(const MemberDesc* __restrict__ md, int gate, int store) {
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= K.n) return;
    using C = double;
    const S pr = mpp<S>(M, MP_P_RAIN)[id], ps = mpp<S>(M, MP_P_SNOW)[id];
    // precipitationSum (UtilsAtm.h:52-64): P_graupel is zero; P_conv once moist convection runs
    mpp<S>(M, MP_PRECIP)[id] = S(C(pr) + C(ps) + 0.0 + (M.mc[0] ? M.mc[MC_P_CONV][id] : 0.0));
    if (store) { mpp<S>(M, MP_P_RAINN)[id] = pr; mpp<S>(M, MP_P_SNOWN)[id] = ps; }
    auto cap = [](C v) { return !finite_bits(v) ? C(0) : (v > C(1.0e-3)) ? C(1.0e-3) : (v < C(-1.0e-3)) ? C(-1.0e-3) : v; };
    S* scc = fp<S>(M, F_S_C_C);
    scc[id] = S(cap(C(scc[id])));
    for (int q = MP_S_V; q <= MP_S_S; q++) mpp<S>(M, q)[id] = S(cap(C(mpp<S>(M, q)[id])));
    S *pc = fp<S>(M, F_C), *pcl = fp<S>(M, F_CLOUD), *pi = fp<S>(M, F_ICE);
    const C co2 = C(fp<S>(M, F_CO2)[id]);
    C cv = C(pc[id]), cld = C(pcl[id]), ic = C(pi[id]);
    if (!finite_bits(cv)) cv = C(0);
    if (!finite_bits(cld)) cld = C(0);
    if (!finite_bits(ic)) ic = C(0);
    if (cv < C(0)) cv = C(0);
    if (cld < C(0)) cld = C(0);
    if (ic < C(0)) ic = C(0);
    const C cv_max = tmax(C(0), C(1) - co2);
    if (cv > cv_max) cv = cv_max;
    const C sum = cld + ic;
    if (sum > C(K.cloud_cap)) { const C sc = C(K.cloud_cap) / sum; cld *= sc; ic *= sc; }
    if (cv < C(0)) cv = C(0);
    if (cv > cv_max) cv = cv_max;
    pc[id] = S(cv); pcl[id] = S(cld); pi[id] = S(ic);
}

// Boundary faces of the microphysics sources S_c_c, S_v, S_c, S_i, S_r, S_s where the ice
// scheme has run: bcRadius pattern A (cubic at both ends, BC_Atm.h:567-644), then bcTheta's
// copy (:761-790), then bcPhi's seam average (:805-842), in the closed form of the three
// sequential passes (c43 = 1, c13 = 0; the sources are finite, capped on every write). Reads
// only cells with every index interior, writes only boundary cells.
template <typename S>
__global__ __launch_bounds__(256) void k_bc_sources(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= K.n || !M.mp[0]) return;
    int i, j, k;
    decode(K, id, i, j, k);
    const int iml = K.im - 1;
    const bool seam = (k == 0 || k == K.km - 1);
    if (!seam && i > 0 && i < iml && j > 0 && j < K.jm - 1) return;
    const int jj = min(max(j, 1), K.jm - 2);
    for (int q = 0; q < 6; q++) {
        S* x = (q == 0) ? fp<S>(M, F_S_C_C) : mpp<S>(M, MP_S_V + q - 1);
        auto R = [&](int kk) -> double {   // the value after bcRadius at (i, jj, kk), kk interior
            if (i == 0)
                return double(x[ix3(K, 3, jj, kk)]) - 3.0 * double(x[ix3(K, 2, jj, kk)]) + 3.0 * double(x[ix3(K, 1, jj, kk)]);
            if (i == iml)
                return double(x[ix3(K, iml - 3, jj, kk)]) - 3.0 * double(x[ix3(K, iml - 2, jj, kk)])
                     + 3.0 * double(x[ix3(K, iml - 1, jj, kk)]);
            return double(x[ix3(K, i, jj, kk)]);
        };
        x[id] = seam ? S((double(S(R(1))) + double(S(R(K.km - 2)))) * 0.5) : S(R(k));
    }
}

// Boundary faces of the moist-convection arrays (FP64), where they exist. bcRadius: cubic at both
// ends for q_v_u, q_c_u, u_u, v_u, w_u, s, s_u, s_d, M_u, M_d (BC_Atm.h:567-574); x[1] at i=0 and a
// copy at the lid for P_conv (c43 = 1, c13 = 0, :587-590, :679-683); cubic at i=0 and a copy at
// the lid for c_u, e_d, e_l, e_p, g_p (:602-606, :733-737); q_v_u, q_c_u at i=0 floored at zero
// (:741-742). bcTheta's copy (:769-775) and bcPhi's seam average (:810-818) for all but M_u, M_d
// and P_conv, in the closed form of the three passes. Writes only boundary cells, reads only
// cells that no thread writes.
__global__ __launch_bounds__(256) void k_bc_conv(const MemberDesc* __restrict__ md) {
#pragma clang fp contract(off)
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= K.n || !M.mc[0]) return;
    int i, j, k;
    decode(K, id, i, j, k);
    const int iml = K.im - 1;
    const bool radial = (i == 0 || i == iml);
    const bool seam = (k == 0 || K.km - 1 == k);
    if (!radial && !seam && j > 0 && j < K.jm - 1) return;
    if (radial) {   // the radial-only arrays, every (j, k) of both radial faces
        for (int q : {MC_M_U, MC_M_D}) {
            const double* x = M.mc[q];
            M.mc[q][id] = (i == 0) ? x[ix3(K, 3, j, k)] - 3.0 * x[ix3(K, 2, j, k)] + 3.0 * x[ix3(K, 1, j, k)]
                                   : x[ix3(K, iml - 3, j, k)] - 3.0 * x[ix3(K, iml - 2, j, k)] + 3.0 * x[ix3(K, iml - 1, j, k)];
        }
        const double* pc = M.mc[MC_P_CONV];
        M.mc[MC_P_CONV][id] = (i == 0) ? 1.0 * pc[ix3(K, 1, j, k)] - 0.0 * pc[ix3(K, 2, j, k)] : pc[ix3(K, iml - 1, j, k)];
    }
    const int jj = min(max(j, 1), K.jm - 2);
    constexpr int kSides = 13;
    const int sides[kSides] = {MC_Q_V_U, MC_Q_C_U, MC_U_U, MC_V_U, MC_W_U, MC_S, MC_S_U, MC_S_D,
                               MC_C_U, MC_E_D_EVAP, MC_E_L, MC_E_P, MC_G_P};
#pragma unroll
    for (int n = 0; n < kSides; n++) {
        const int q = sides[n];
        const double* x = M.mc[q];
        const bool top_cubic = n < 8;
        const bool floor0 = (q == MC_Q_V_U || q == MC_Q_C_U);
        auto R = [&](int kk) -> double {   // the value after bcRadius at (i, jj, kk), kk interior
            if (i == 0) {
                const double v = x[ix3(K, 3, jj, kk)] - 3.0 * x[ix3(K, 2, jj, kk)] + 3.0 * x[ix3(K, 1, jj, kk)];
                return (floor0 && v < 0.0) ? 0.0 : v;
            }
            if (i == iml)
                return top_cubic ? x[ix3(K, iml - 3, jj, kk)] - 3.0 * x[ix3(K, iml - 2, jj, kk)]
                                       + 3.0 * x[ix3(K, iml - 1, jj, kk)]
                                 : x[ix3(K, iml - 1, jj, kk)];
            return x[ix3(K, i, jj, kk)];
        };
        M.mc[q][id] = seam ? (R(1) + R(K.km - 2)) * 0.5 : R(k);
    }
}

// updateSkinTemperature (cAtmosphereModel.cpp:276-311) with planetaryShortWave (:241-270) on the
// zonal albedo sums k_albedo_rows left at sf_base, on the device for radiation inside the step,
// one block per member: the threads form each latitude's weight, short wave and mean albedo
// in LDS, one thread accumulates them in j order with the host form's arithmetic
// (athad_radiation); the new t_skin goes into both descriptor arrays, where the next densities
// reads it.
constexpr int kSkinChunk = 1024;
__global__ __launch_bounds__(256) void k_skin_step(MemberDesc* md0, MemberDesc* md1, double relax) {
#pragma clang fp contract(off)
    __shared__ double s_w[kSkinChunk], s_sw[kSkinChunk], s_a[kSkinChunk];
    if (!(relax > 0.0)) return;
    const int q = blockIdx.x;
    const Consts& K = md0[q].K;
    const int jm = K.jm, km = K.km, jmax = jm - 1, jhalf = jmax / 2;
    const double* rows = md0[q].scratch + sf_base(K);
    double wsum = 0.0, absorbed = 0.0, amax = 0.0;
    for (int j0 = 0; j0 < jm; j0 += kSkinChunk) {
        const int n = min(kSkinChunk, jm - j0);
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int j = j0 + t;
            const double x = (double)((j <= jhalf) ? j : jmax - j) / (double)jhalf;
            s_sw[t] = (K.sw_pole - K.sw_equator) * (x * x - 2.0 * x) + K.sw_pole;
            s_w[t] = cos((j / (double)(jm - 1) - 0.5) * M_PI);
            s_a[t] = rows[j] / (double)km;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int t = 0; t < n; t++) {
                const double w = s_w[t], a = s_a[t];
                amax = tmax(amax, a);
                wsum += w;
                absorbed += w * (1.0 - a) * s_sw[t];
            }
        __syncthreads();
    }
    if (threadIdx.x != 0 || !(wsum > 0.0) || !(amax > 0.0)) return;
    const double in = absorbed / wsum + K.geothermal;
    if (!(in > 0.0)) return;
    const double target = pow(in / K.sigma, 0.25);
    const double t = K.t_skin + tmin(1.0, relax) * (target - K.t_skin);
    md0[q].K.t_skin = t;
    md1[q].K.t_skin = t;
}

enum { FT = 0, FU, FV, FW, FC, FCL, FI, FG, FCO, FK, FD };

// Solid-ground clamps of bcSolidGround passes 4-5 (BC_Atm.h:475-548) as a value transform,
// used when the Jacobi kernel applies them on read instead of in a separate pass.
template <typename C> ATHAD_DEV C solid_clamp(int f, C v, C tke_max, C vel_max) {
    switch (f) {
        case FC: case FCL: case FI: return (v < C(0)) ? C(0) : v;
        case FK: { C x = (v < C(0)) ? C(0) : v; if (!finite_bits(x)) return C(0);
                   return (x < C(0)) ? C(0) : (x > tke_max) ? tke_max : x; }
        case FD: { C x = (v < C(0)) ? C(0) : v; if (!finite_bits(x)) return C(1.0e-10);
                   return (x < C(1.0e-10)) ? C(1.0e-10) : (x > C(1.0e6)) ? C(1.0e6) : x; }
        case FU: case FV: case FW: if (!finite_bits(v)) return C(0);
                 return (v < -vel_max) ? -vel_max : (v > vel_max) ? vel_max : v;
        default: return v;
    }
}
template <typename C> ATHAD_DEV C solid_clamp_n(int f, C v, C tke_max, C vel_max) {   // tken, disn, un, vn, wn
    switch (f) {
        case FK: if (!finite_bits(v)) return C(0); return (v < C(0)) ? C(0) : (v > tke_max) ? tke_max : v;
        case FD: if (!finite_bits(v)) return C(1.0e-10);
                 return (v < C(1.0e-10)) ? C(1.0e-10) : (v > C(1.0e6)) ? C(1.0e6) : v;
        case FU: case FV: case FW: if (!finite_bits(v)) return C(0);
                 return (v < -vel_max) ? -vel_max : (v > vel_max) ? vel_max : v;
        default: return v;
    }
}

// minmod(a, b) = 0 if a b <= 0, else the argument of smaller magnitude, evaluated as
// max(0, min(a, b)) + min(0, max(a, b)) (equal for finite arguments whose product does not
// underflow to zero).
template <typename C> ATHAD_DEV C minmod(C a, C b) {
    return fmax(C(0), fmin(a, b)) + fmin(C(0), fmax(a, b));
}

// Reciprocal: the hardware estimate refined by Newton steps to within about 1 ulp. Used for
// operands bounded away from zero and infinity (dis >= 1e-10, CD_kw >= 1e-20, sigma blends >= 1).
ATHAD_DEV double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
ATHAD_DEV float rcp_nr(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    const float e = fmaf(-x, r, 1.0f);
    return fmaf(r, e, r);
}

// Natural log, FP64: for a positive normal argument by the reduced-argument method of fdlibm's
// e_log: x = 2^k (1+f) with 1+f in [sqrt(1/2), sqrt(2)) (f exact), s = f/(2+f), and
// log(x) = k ln2_hi - ((f^2/2 - (s (f^2/2 + R) + k ln2_lo)) - f), R a degree-7 polynomial in
// s^2 (fdlibm's Remez coefficients Lg1..Lg7). Within 1 ulp of the library log (98 % of
// results equal over 3e7 points from 1e-300 to 1e300), ~40 instructions instead of ~100.
// Zero, negative, subnormal and non-finite arguments take the library log. FP32: logf.
template <typename C> ATHAD_DEV C log_fast(C x) {
    if constexpr (sizeof(C) == 4) {
        return log(x);
    } else {
        if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308)) return log(x);
        int e;
        double m = frexp(x, &e);   // x = m 2^e, m in [0.5, 1)
        if (m < 0.70710678118654752440) { m = m * 2.0; e -= 1; }
        const double f = m - 1.0;
        const double s = f * rcp_nr(2.0 + f);
        const double z = s * s, w = z * z;
        const double t1 = w * (3.999999999940941908e-01 + w * (2.222219843214978396e-01 + w * 1.531383769920937332e-01));
        const double t2 = z * (6.666666666666735130e-01 + w * (2.857142874366239149e-01
                                + w * (1.818357216161805012e-01 + w * 1.479819860511658591e-01)));
        const double R = t2 + t1;
        const double hfsq = 0.5 * f * f;
        const double k = (double)e;
        return k * 6.93147180369123816490e-01 - ((hfsq - (s * (hfsq + R) + k * 1.90821492927058770002e-10)) - f);
    }
}

// tanh(x) for x >= 0; in FP64 as -m / (2 + m) with m = expm1(-2x) in (-1, 0]: no
// cancellation, within 3 ulp of the library tanh over [1e-9, 25] (mean 0.5 ulp), at about a
// third of its cost (the library's FP64 tanh is ~165 instructions, expm1 ~50). Below 1e-9 (double) or
// 1e-4 (float) the correctly rounded tanh(x) is x itself (x^3/3 is below half an ulp of x).
template <typename C, bool BRANCHLESS = false> ATHAD_DEV C tanh_pos(C x) {
    if constexpr (BRANCHLESS && sizeof(C) == 8) {   // both results, then a select
        const C m = expm1(C(-2) * x);
        const C r = -m * rcp_nr(C(2) + m);
        return (x >= C(1.0e-9)) ? r : x;
    }
    C r = x;
    if (x >= C(sizeof(C) == 8 ? 1.0e-9 : 1.0e-4)) {   // a branch: skipped when no lane needs it
        if constexpr (sizeof(C) == 8) {
            const C m = expm1(C(-2) * x);
            r = -m * rcp_nr(C(2) + m);
        } else {
            r = tanh(x);   // the FP32 library tanh is already short
        }
    }
    return r;
}

// Global-memory access through address space 1 (pointers read from a MemberDesc are
// otherwise generic and compile to flat instructions).
template <typename T> ATHAD_DEV T gld(const T* p, long long i) {
    return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <typename T> ATHAD_DEV void gst(T* p, long long i, T v) {
    ((__attribute__((address_space(1))) T*)p)[i] = v;
}

// Per-cell metric (cAtmosphereModel.h:170-178, 282-284; RHS_Atm_Turb.cpp:86-130), the
// Held-Suarez rate (:976-992), the Rayleigh drag factor and the Coriolis factor.
// ------------------------------------------------------------------------------------
// RHS of one cell (RHS_Atm_Turb.cpp:31-1080, land-free central stencils) and the RK4
// stage loop (RungeKutta_Atm_Turb.cpp:137-263). C is the arithmetic type.
//
// Over the four stages only the cell's own values change; its neighbours do not (they are
// the start-of-RK4 values in the Jacobi order and values on other hyperplanes in the lex
// order). Everything built from neighbours alone -- central gradients, the neighbour sums
// of every Laplacian and of the vector diffusion, the velocity-gradient invariants of the
// production term, the SST cross-diffusion product, the pressure gradient -- is computed
// once per cell at the start of each field group (rk4_dyn, rk4_moist, rk4_turb); each stage
// only adds the centre-value terms, the minmod slopes and the SST closure. The RK4 base and
// accumulator live in LDS.
// ------------------------------------------------------------------------------------
// sigma = p_stat / p_surf of the Held-Suarez rate, zero where p_surf is not positive. The
// division runs for every lane and the zero is selected after it (the same values): written as
// `p_surf > 0 ? ps / p_surf : 0` it compiles to a branch that waits for p_surf before p_stat and
// the cell's later loads are issued.
ATHAD_DEV double hs_sigma(double ps_c, double p_surf) {
    double sigma = ps_c / p_surf;
    if (!(p_surf > 0.0)) sigma = 0.0;
    return sigma;
}

template <typename C> struct CellGeo {
    C aR, aT, aP;          // advection metric / grid step, for the minmod slopes
    C cexp, cirm, cirs;    // advection metric exp_rm, 1/rm, 1/(rm sin theta)
    C wr, wt, wp;          // Laplacian weights of the neighbour sums (r, theta, phi)
    C two_over, crs, Lc2;  // first-derivative Laplacian weights; centre coefficient
    C cu, cvw, cirm2, cirm2s, c2cos, hs_kT, drag, cor;   // dynamics group only
};

template <typename C>
ATHAD_DEV void cell_geo(const MemberDesc& M, int i, int j, double ps_c, double p_surf, CellGeo<C>& g) {
    const Consts& K = M.K;
    // the (i, j) entry of the metric table: 16-byte aligned, read as NGEO / 2 pairs
    typedef double pair __attribute__((ext_vector_type(2)));
    const pair* q2 = (const pair*)(M.geo + ((long long)i * K.jm + j) * NGEO);
    double q[NGEO];
#pragma unroll
    for (int e = 0; e < NGEO / 2; e++) {
        const pair v = gld(q2, e);
        q[2 * e] = v.x;
        q[2 * e + 1] = v.y;
    }
    g.aR = C(q[G_AR]); g.aT = C(q[G_AT]); g.aP = C(q[G_AP]);
    g.cexp = C(q[G_CEXP]); g.cirm = C(q[G_CIRM]); g.cirs = C(q[G_CIRS]);
    g.wr = C(q[G_WR]); g.wt = C(q[G_WT]); g.wp = C(q[G_WP]);
    g.two_over = C(q[G_TWO_OVER]);
    g.crs = C(q[G_CRS]);
    g.Lc2 = C(q[G_LC2]);
    g.cu = C(q[G_CU]);
    g.cvw = C(q[G_CVW]);
    g.cirm2 = C(q[G_CIRM2]); g.cirm2s = C(q[G_CIRM2S]); g.c2cos = C(q[G_C2COS]);
    const double sigma = hs_sigma(ps_c, p_surf);
    double bl = (sigma - 0.7) / (1.0 - 0.7);
    if (bl < 0.0) bl = 0.0;
    const double kT = K.hs_k_a + (K.hs_k_s - K.hs_k_a) * bl * q[G_S4];
    g.hs_kT = C(kT * K.L_atm / K.u_0);
    g.drag = C(q[G_DRAG]);
    g.cor = C(q[G_COR]);
}

// The same coefficients formed in place from the 1-D tables (ten divisions): fewer live
// registers than the table's twenty values, which the FP64 Jacobi kernel, at three waves per
// SIMD on the edge of spilling, needs more than the saved arithmetic.
template <typename C>
ATHAD_DEV void cell_geo_alu(const MemberDesc& M, int i, int j, double ps_c, double p_surf, CellGeo<C>& g) {
    const Consts& K = M.K;
    const double rm = M.rad[i];
    const double exp_rm = 1.0 / (rm + 1.0);
    const double sinthe = M.sint[j], costhe = M.cost[j];
    const double sinthe2 = sinthe * sinthe;
    const double rmet = (K.metric_r0 > 0.0) ? (K.metric_r0 + (rm - K.rad0)) : rm;
    const double inv_rm = 1.0 / rmet;
    const double inv_rm2 = 1.0 / (rmet * rmet);
    const double inv_rms = 1.0 / (rmet * sinthe);
    const double inv_rm2sinthe = inv_rm2 / sinthe;
    const double inv_rm2s2 = inv_rm2 / sinthe2;
    const double v_metric = (1.0 + costhe * costhe / sinthe2) * inv_rm2;
    const double wr = K.inv_dr2 * (exp_rm * exp_rm), wt = K.inv_dthe2 * inv_rm2, wp = K.inv_dphi2 * inv_rm2s2;
    const double Lc2 = 2.0 * (wr + wt + wp);
    g.aR = C(exp_rm * 2.0 * K.inv_2dr);
    g.aT = C(inv_rm * 2.0 * K.inv_2dthe);
    g.aP = C(inv_rms * 2.0 * K.inv_2dphi);
    g.cexp = C(exp_rm); g.cirm = C(inv_rm); g.cirs = C(inv_rms);
    g.wr = C(wr); g.wt = C(wt); g.wp = C(wp);
    g.two_over = C(2.0 * inv_rm * exp_rm);
    g.crs = C(costhe * inv_rm2sinthe);
    g.Lc2 = C(Lc2);
    g.cu = C(2.0 * inv_rm2 - Lc2);
    g.cvw = C(-v_metric - Lc2);
    g.cirm2 = C(inv_rm2); g.cirm2s = C(inv_rm2sinthe); g.c2cos = C(2.0 * costhe * inv_rm2s2);
    const double sigma = hs_sigma(ps_c, p_surf);
    double bl = (sigma - 0.7) / (1.0 - 0.7);
    if (bl < 0.0) bl = 0.0;
    const double kT = K.hs_k_a + (K.hs_k_s - K.hs_k_a) * bl * (sinthe2 * sinthe2);
    g.hs_kT = C(kT * K.L_atm / K.u_0);
    double dp = 1.0 - (double)i / 5.0;
    if (dp < 0.0) dp = 0.0;
    if (dp > 1.0) dp = 1.0;
    g.drag = C(K.drag_coeff * dp);
    g.cor = C(K.force_nd * 2.0 * costhe);
}

// The six neighbours (i+1, i-1, j+1, j-1, k+1, k-1) of field f at id, optionally through the
// solid-ground clamp.
template <typename S, typename C, bool CLAMP>
ATHAD_DEV void load6(const S* x, long long id, long long si, long long sj, int f, C tkm, C vm, C* v) {
    v[0] = C(gld(x, id + si)); v[1] = C(gld(x, id - si)); v[2] = C(gld(x, id + sj));
    v[3] = C(gld(x, id - sj)); v[4] = C(gld(x, id + 1)); v[5] = C(gld(x, id - 1));
    if (CLAMP) {
#pragma unroll
        for (int d = 0; d < 6; d++) v[d] = solid_clamp(f, v[d], tkm, vm);
    }
}

constexpr int kProgF[NPROG] = {F_T, F_U, F_V, F_W, F_C, F_CLOUD, F_ICE, F_GR, F_CO2, F_TKE, F_DIS};
constexpr int kProgN[NPROG] = {F_TN, F_UN, F_VN, F_WN, F_CN, F_CLOUDN, F_ICEN, F_GRN, F_CO2N, F_TKEN, F_DISN};

// Where one cell's RK4 reads its inputs (the 7-point neighbourhoods of the 11 prognostic
// fields and p_dyn, the start and base values, nue, t_eq, p_stat and the surface p_stat, and
// with the ice scheme r_humid and the five sources).
// NbLoad reads each value from global memory where the computation first needs it: the
// Jacobi kernel runs several waves per SIMD, which hide the latency. NbRegs issues every load
// of the cell in one batch into registers before any arithmetic: the lex plane kernel runs
// about one wave per SIMD (a plane holds a few thousand cells), so every dependent memory
// round trip of its cell chain is exposed and one round trip is what it can afford.
// Source flags of the RK4 kernels (template parameter SRC): the ice scheme's microphysics sources
// (RHS_Atm_Turb.cpp:963-966, 1058-1067) and moist convection's tendencies MC_t, MC_v, MC_w, MC_q
// (RHS_Atm_Turb.cpp:370-372, 964, 1023, 1042, 1061), each read once per cell where its arrays exist.
constexpr int SRC_ICE = 1, SRC_CONV = 2;

template <typename C>
ATHAD_DEV void cell_geo_row(const MemberDesc& M, int i, int j, double ps_c, double p_surf, CellGeo<C>& g);

template <typename S, typename C, bool CLAMP> struct NbLoad {
    const MemberDesc& M;
    const S* const* X;
    long long id, si, sj, isurf;
    ATHAD_DEV void six(int f, C* v) const {
        load6<S, C, CLAMP>(X[f], id, si, sj, f, C(M.K.tke_max_nd), C(M.K.vel_max_nd), v);
    }
    ATHAD_DEV void pd6(C* v) const {
        const S* p = fp<S>(M, F_P_DYN);
        v[0] = C(gld(p, id + si)); v[1] = C(gld(p, id - si)); v[2] = C(gld(p, id + sj));
        v[3] = C(gld(p, id - sj)); v[4] = C(gld(p, id + 1)); v[5] = C(gld(p, id - 1));
    }
    ATHAD_DEV C field(int fid) const { return C(gld(fp<S>(M, fid), id)); }   // nue, t_eq, r_humid
    ATHAD_DEV C mp(int q) const { return C(gld((const S*)M.mp[q], id)); }
    ATHAD_DEV double mc(int q) const { return gld((const double*)M.mc[q], id); }
    ATHAD_DEV C start(int f) const {
        C c0 = C(gld(X[f], id));
        if (CLAMP) c0 = solid_clamp(f, c0, C(M.K.tke_max_nd), C(M.K.vel_max_nd));
        return c0;
    }
    ATHAD_DEV C base(int f) const {
        C b0 = C(gld(fp<S>(M, kProgN[f]), id));
        if (CLAMP) b0 = solid_clamp_n(f, b0, C(M.K.tke_max_nd), C(M.K.vel_max_nd));
        return b0;
    }
    ATHAD_DEV double ps() const { return gld(pstat(M), id); }
    ATHAD_DEV double psurf() const { return gld(pstat(M), isurf); }
};

// Buffer-resource access (raw buffers: a 128-bit descriptor in scalar registers, a 32-bit
// vector byte offset, a scalar byte offset). A cell's seven stencil values of a field then
// differ only in the scalar offset, so all of them, and every other field, use the one vector
// offset of the cell and need no 64-bit address arithmetic. Word 3 of the descriptor is the
// gfx9-family raw-buffer format word. aux = 2 marks a load or store as streaming (nt): values
// read or written once per launch, keeping the cache for the stencil fields that neighbouring
// waves read again. A member holds fewer than 2^28 cells (checked at creation), so every byte
// offset below fits 32 bits.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kNT = 2;
ATHAD_DEV Rsrc make_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
}
template <typename T, int AUX = 0> ATHAD_DEV T bld(Rsrc r, unsigned voff, unsigned soff) {
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
    else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
}
template <typename T, int AUX = 0> ATHAD_DEV void bst(Rsrc r, unsigned voff, T v) {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, 0, AUX);
    else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, 0, AUX);
}

// The solid-ground clamps in min / max form (identical results: every branch that the
// comparison form takes for a finite value is the min / max, and non-finite values are
// caught first; the c, cloud, ice clamp keeps its comparison form, which passes a NaN on).
template <typename C> ATHAD_DEV C solid_clamp_mm(int f, C v, C tke_max, C vel_max) {
    switch (f) {
        case FC: case FCL: case FI: return (v < C(0)) ? C(0) : v;
        case FK: return finite_bits(v) ? fmin(fmax(v, C(0)), tke_max) : C(0);
        case FD: return finite_bits(v) ? fmin(fmax(v, C(1.0e-10)), C(1.0e6)) : C(1.0e-10);
        case FU: case FV: case FW: return finite_bits(v) ? fmin(fmax(v, -vel_max), vel_max) : C(0);
        default: return v;
    }
}

// NbLoad for the row kernel. A stencil field is read through a descriptor based one plane and
// one cell before the field (f - si - 1), so the displacements of all seven points are
// non-negative scalar offsets: centre c = (si + 1) E, i -+ 1 at c -+ si E, j -+ 1 at c -+ sj E,
// k -+ 1 at c -+ E (E = bytes per value). off = byte offset of the cell in a field,
// soff = of its surface cell in p_stat (FP64). tkm, vm: the clamp bounds tke_max_nd and
// vel_max_nd, read once by the kernel; the clamps then need no load of their own, which would
// otherwise make each clamp a branch that waits for its value before the next load issues.
template <typename S, typename C, bool CLAMP> struct NbRow {
    static constexpr unsigned E = sizeof(S);
    const MemberDesc& M;
    unsigned off, soff, dsi, dsj, nst, nfld;   // nst: bytes of a stencil descriptor, nfld: of a field
    C tkm, vm;
    ATHAD_DEV Rsrc stencil(int fid) const { return make_rsrc((const char*)M.f[fid] - dsi - E, nst); }
    ATHAD_DEV void six_of(int fid, C* v) const {
        const Rsrc r = stencil(fid);
        const unsigned c = dsi + E;
        v[0] = C(bld<S>(r, off, c + dsi)); v[1] = C(bld<S>(r, off, E));
        v[2] = C(bld<S>(r, off, c + dsj)); v[3] = C(bld<S>(r, off, c - dsj));
        v[4] = C(bld<S>(r, off, c + E)); v[5] = C(bld<S>(r, off, dsi));
    }
    ATHAD_DEV void six(int f, C* v) const {
        six_of(kProgF[f], v);
        if (CLAMP) {
#pragma unroll
            for (int d = 0; d < 6; d++) v[d] = solid_clamp_mm(f, v[d], tkm, vm);
        }
    }
    ATHAD_DEV void pd6(C* v) const { six_of(F_P_DYN, v); }
    ATHAD_DEV C once(const void* p, unsigned n) const { return C(bld<S, kNT>(make_rsrc(p, n), off, 0)); }
    ATHAD_DEV C field(int fid) const { return once(M.f[fid], nfld); }   // nue, t_eq, r_humid
    ATHAD_DEV C mp(int q) const { return once(M.mp[q], nfld); }
    ATHAD_DEV double mc(int q) const {
        return bld<double, kNT>(make_rsrc(M.mc[q], nfld / E * 8u), off / E * 8u, 0);
    }
    ATHAD_DEV C start(int f) const {
        C c0 = C(bld<S>(stencil(kProgF[f]), off, dsi + E));
        if (CLAMP) c0 = solid_clamp_mm(f, c0, tkm, vm);
        return c0;
    }
    ATHAD_DEV C base(int f) const {   // the *n clamps of tken, disn, un, vn, wn
        C b0 = once(M.f[kProgN[f]], nfld);
        if (CLAMP && f != FC && f != FCL && f != FI) b0 = solid_clamp_mm(f, b0, tkm, vm);
        return b0;
    }
    ATHAD_DEV double ps() const { return bld<double, kNT>(make_rsrc(pstat(M), nfld / E * 8u), off / E * 8u, 0); }
    ATHAD_DEV double psurf() const { return bld<double>(make_rsrc(pstat(M), nfld / E * 8u), soff, 0); }
    ATHAD_DEV void put(void* p, C v) const { bst<S, kNT>(make_rsrc(p, nfld), off, S(v)); }
};

template <typename S, typename C, int SRC> struct NbRegs {
    C nb[NPROG][6], pd[6], c0[NPROG], b0[NPROG], nu, te, rh, src[5];
    double psv, pss, mcv[4];
    ATHAD_DEV NbRegs(const MemberDesc& M, const S* const* X, long long id, long long si, long long sj,
                     long long isurf) {
        const C tkm = C(M.K.tke_max_nd), vm = C(M.K.vel_max_nd);
#pragma unroll
        for (int f = 0; f < NPROG; f++) {
            load6<S, C, false>(X[f], id, si, sj, f, tkm, vm, nb[f]);
            c0[f] = C(gld(X[f], id));
            b0[f] = C(gld(fp<S>(M, kProgN[f]), id));
        }
        const S* p = fp<S>(M, F_P_DYN);
        pd[0] = C(gld(p, id + si)); pd[1] = C(gld(p, id - si)); pd[2] = C(gld(p, id + sj));
        pd[3] = C(gld(p, id - sj)); pd[4] = C(gld(p, id + 1)); pd[5] = C(gld(p, id - 1));
        nu = C(gld(fp<S>(M, F_NUE), id));
        te = C(gld(fp<S>(M, F_T_EQ), id));
        psv = gld(pstat(M), id);
        pss = gld(pstat(M), isurf);
        rh = C(0);
#pragma unroll
        for (int q = 0; q < 5; q++) src[q] = C(0);
        if (SRC & SRC_ICE) {
            rh = C(gld(fp<S>(M, F_R_HUMID), id));
#pragma unroll
            for (int q = 0; q < 5; q++) src[q] = C(gld((const S*)M.mp[MP_S_V + q], id));
        }
#pragma unroll
        for (int q = 0; q < 4; q++) mcv[q] = (SRC & SRC_CONV) ? gld((const double*)M.mc[MC_MC_T + q], id) : 0.0;
    }
    ATHAD_DEV void six(int f, C* v) const {
#pragma unroll
        for (int d = 0; d < 6; d++) v[d] = nb[f][d];
    }
    ATHAD_DEV void pd6(C* v) const {
#pragma unroll
        for (int d = 0; d < 6; d++) v[d] = pd[d];
    }
    ATHAD_DEV C field(int fid) const { return fid == F_NUE ? nu : fid == F_T_EQ ? te : rh; }
    ATHAD_DEV C mp(int q) const { return src[q - MP_S_V]; }
    ATHAD_DEV double mc(int q) const { return mcv[q - MC_MC_T]; }
    ATHAD_DEV C start(int f) const { return c0[f]; }
    ATHAD_DEV C base(int f) const { return b0[f]; }
    ATHAD_DEV double ps() const { return psv; }
    ATHAD_DEV double psurf() const { return pss; }
};

template <typename C> ATHAD_DEV C minmod3(const C* nb, C c0, C ua, C va, C wa) {
    return ua * minmod(c0 - nb[1], nb[0] - c0) + va * minmod(c0 - nb[3], nb[2] - c0)
         + wa * minmod(c0 - nb[5], nb[4] - c0);
}

template <typename C> ATHAD_DEV C buoyancy(const MemberDesc& M) {
    const int total = ro_load(M.counters + 1);
    const double ramp = fmin(1.0, (double)total / 300.0);
    return C(ramp * 1.0 * M.K.g * M.K.dt / M.K.u_0);
}

// RK4 stage update of one field (RungeKutta_Atm_Turb.cpp:167-263): b = base (*n value) and
// a = k1+2k2+2k3 accumulator, both in this thread's LDS slots.
// Stage st adds r * kr to the accumulator (kr = 1, 2, 2, 1; the first stage starts it from
// zero) and returns b + r * cdt (cdt = dt/2, dt/2, dt), or in the last stage
// b + (k1 + 2 k2 + 2 k3 + k4) dt/6. The FP32 kernel uses the branch-free form (wave-uniform
// coefficients); the FP64 kernel keeps per-stage branches, which need fewer registers there
// (measured: the branch-free FP64 kernel spills).
template <typename C> ATHAD_DEV C stage_update(int st, C r, C* b, C* a, C dt, C half_dt, C dt_sixth) {
    if constexpr (sizeof(C) == 4) {
        const C kr = (st == 1 || st == 2) ? C(2) : C(1);
        const C cdt = (st == 2) ? dt : half_dt;
        const C acc = ((st == 0) ? C(0) : *a) + r * kr;
        *a = acc;
        const C bv = *b;
        return (st == 3) ? bv + acc * dt_sixth : bv + r * cdt;
    } else {
        const C bv = *b;
        if (st == 0) { *a = r; return bv + r * half_dt; }
        if (st == 1) { *a = *a + C(2) * r; return bv + r * half_dt; }
        if (st == 2) { *a = *a + C(2) * r; return bv + r * dt; }
        return bv + (*a + r) * dt_sixth;
    }
}

constexpr int LB = 256;   // LDS slot stride = threads per block of the RK4 kernels
// SB: stride of a thread's stage slots (base, accumulator, recorded velocities) in LDS. (Private
// slots in registers for the lex plane kernel, stride 1, measured no faster.)

// The four RK4 stages of a field group: a loop (U = false: the code of one stage, the fewest
// registers, for the Jacobi kernel's occupancy) or unrolled (U = true: every stage's branches
// fold, and the independent chains of a group's fields schedule into one another, for the
// lex plane kernel, whose one wave per SIMD exposes every dependent-issue latency).
template <bool U, typename F> ATHAD_DEV void for_stages(F&& f) {
    if constexpr (U) {
        f(0); f(1); f(2); f(3);
    } else {
#pragma unroll 1
        for (int st = 0; st < 4; st++) f(st);
    }
}

// RK4 of the dynamics group T, u, v, w (RHS_Atm_Turb.cpp rhs_t/rhs_u/rhs_v/rhs_w). Their
// tendencies involve no other prognostic field, so the group integrates first and records
// the velocity entering each stage in rec[] for the scalar groups. cen[] = start values in,
// final values out. A, B: velocity-gradient invariants of the production term.
// A per-stage hook of the stage loops (called at the start of stage st, after the dynamics
// group has recorded the stage's velocity): nothing, except where the groups of one cell run
// on different waves and meet at a barrier per stage.
struct NoHook {
    ATHAD_DEV void operator()(int) const {}
};

template <typename S, typename C, bool CLAMP, int SRC, bool U, int SB, typename Nb, typename Hook = NoHook>
ATHAD_DEV void rk4_dyn(const MemberDesc& M, const Nb& src_nb, int i, const CellGeo<C>& g,
                       C* ba, C* rec, C* cen, C* aux, C& A, C& B, Hook hook = Hook{}) {
    const Consts& K = M.K;
    const C i2r = C(K.inv_2dr), i2t = C(K.inv_2dthe), i2p = C(K.inv_2dphi);
    C nb[4][6], g1r[4], g1t[4], g1p[4], Sr[4], St[4], Sp[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        src_nb.six(f, nb[f]);
        g1r[f] = (nb[f][0] - nb[f][1]) * i2r;
        g1t[f] = (nb[f][2] - nb[f][3]) * i2t;
        g1p[f] = (nb[f][4] - nb[f][5]) * i2p;
        Sr[f] = nb[f][0] + nb[f][1]; St[f] = nb[f][2] + nb[f][3]; Sp[f] = nb[f][4] + nb[f][5];
    }
    const C L0T = Sr[FT] * g.wr + g1r[FT] * g.two_over + St[FT] * g.wt + g1t[FT] * g.crs + Sp[FT] * g.wp;
    // vector diffusion of u, v, w without their centre terms
    const C D0u = Sr[FU] * g.wr + St[FU] * g.wt + C(4) * g1r[FU] * g.cirm * g.cexp + g1t[FU] * g.crs + Sp[FU] * g.wp;
    const C D0v = Sr[FV] * g.wr + g1r[FV] * g.two_over + St[FV] * g.wt + g1t[FV] * g.crs + Sp[FV] * g.wp
                + C(2) * g1t[FU] * g.cirm2 - g1p[FW] * g.c2cos;
    const C D0w = Sr[FW] * g.wr + g1r[FW] * g.two_over + St[FW] * g.wt + g1t[FW] * g.crs + Sp[FW] * g.wp
                + C(2) * g1p[FU] * g.cirm2s + g1p[FV] * g.c2cos;
    // production invariants (prod refresh RHS_Atm_Turb.cpp:528-561):
    // prod = max(0, cnue*A - 0.66667*tke*B)
    {
        const C a1 = g1r[FU] * g.cexp, a2 = g1r[FV] * g.cexp, a3 = g1r[FW] * g.cexp;
        const C b1 = g1t[FU] * g.cirm, b2 = g1t[FV] * g.cirm, b3 = g1t[FW] * g.cirm;
        const C c1 = g1p[FU] * g.cirs, c2 = g1p[FV] * g.cirs, c3 = g1p[FW] * g.cirs;
        const C der = C(0.66667) * (a1 + b2 + c3);
        const C s12 = b1 + a2, s13 = c1 + a3, s23 = c2 + b3;
        A = (C(2) * a1 - der) * a1 + (C(2) * b2 - der) * b2 + (C(2) * c3 - der) * c3 + s12 * s12 + s13 * s13
          + s23 * s23;
        B = a1 + b2 + c3;
    }
    C PR, PT, PP;
    {
        C p[6];
        src_nb.pd6(p);
        PR = (p[0] - p[1]) * i2r * g.cexp;
        PT = (p[2] - p[3]) * i2t * g.cirm;
        PP = (p[4] - p[5]) * i2p * g.cirs;
    }
    C dif_t, dif_vel;
    {
        C nv = src_nb.field(F_NUE);
        if (CLAMP) nv = finite_bits(nv) ? tclamp(nv, C(0), C(K.nue_max)) : C(0);
        const C inv_re = C(K.inv_re_turb);
        dif_vel = inv_re + nv;
        dif_t = (inv_re + nv) * C(K.inv_pr_turb);
    }
    const C t_eq = src_nb.field(F_T_EQ);
    const C t_ref = C(ro_load(M.t_ref + i));
    const C buoy = buoyancy<C>(M);
    const C ce = C(K.coeff_energy_p);
    // latent heating of the microphysics sources (RHS_Atm_Turb.cpp:965-966), stage-invariant
    C srcT = C(0);
    if (SRC & SRC_ICE) {
        const C rh = src_nb.field(F_R_HUMID);
        const C Sc = src_nb.mp(MP_S_C), Sr = src_nb.mp(MP_S_R);
        const C Si = src_nb.mp(MP_S_I), Ss = src_nb.mp(MP_S_S);
        const C cen_ = C(K.coeff_energy);
        srcT = cen_ * (Sc + Sr) * C(K.lv) * rh + cen_ * (Si + Ss + C(0)) * C(K.ls) * rh;
    }
    // moist convection's tendencies (RHS_Atm_Turb.cpp:964, 1023, 1042), stage-invariant
    C mcT = C(0), mcV = C(0), mcW = C(0);
    if (SRC & SRC_CONV) {
        mcT = C(K.coeff_mc_t * src_nb.mc(MC_MC_T));
        mcV = C(K.coeff_mc_vel * src_nb.mc(MC_MC_V));
        mcW = C(K.coeff_mc_vel * src_nb.mc(MC_MC_W));
    }
    const C dt = C(K.dt), half_dt = C(0.5 * K.dt), dt_sixth = C(K.dt_sixth);
for_stages<U>([&](const int st) {
        const C T = cen[FT], u = cen[FU], v = cen[FV], w = cen[FW];
        rec[(3 * st + 0) * SB] = u; rec[(3 * st + 1) * SB] = v; rec[(3 * st + 2) * SB] = w;
        hook(st);
        const C ua = u * g.aR, va = v * g.aT, wa = w * g.aP;
        C rT = ce * (u * PR + v * PT + w * PP) - minmod3(nb[FT], T, ua, va, wa) + (L0T - g.Lc2 * T) * dif_t;
        if (SRC & SRC_CONV) rT = rT + mcT;
        if (SRC & SRC_ICE) rT = rT + srcT;
        rT = rT - g.hs_kT * (T - t_eq);
        const C rU = -PR - minmod3(nb[FU], u, ua, va, wa) + (D0u + g.cu * u) * dif_vel + buoy * (T - t_ref);
        C rV = -PT - minmod3(nb[FV], v, ua, va, wa) + (D0v + g.cvw * v) * dif_vel + g.cor * w;
        C rW = -PP - minmod3(nb[FW], w, ua, va, wa) + (D0w + g.cvw * w) * dif_vel - g.cor * v;
        if (SRC & SRC_CONV) { rV = rV + mcV; rW = rW + mcW; }
        rV = rV - g.drag * v;
        rW = rW - g.drag * w;
        if (st == 3) { aux[0] = rU + PR; aux[1] = rV + PT; aux[2] = rW + PP; }
        cen[FT] = stage_update(st, rT, ba + 0 * SB, ba + 4 * SB, dt, half_dt, dt_sixth);
        cen[FU] = stage_update(st, rU, ba + 1 * SB, ba + 5 * SB, dt, half_dt, dt_sixth);
        cen[FV] = stage_update(st, rV, ba + 2 * SB, ba + 6 * SB, dt, half_dt, dt_sixth);
        cen[FW] = stage_update(st, rW, ba + 3 * SB, ba + 7 * SB, dt, half_dt, dt_sixth);
    });
}

// RK4 of moisture fields Q0 .. Q0+NQ-1 of the group c, cloud, ice (q = 0..2, minmod
// advection), gr, co2 (q = 3, 4, central), advected by the recorded stage velocities; each
// field's tendency involves only itself and the velocities, so the group can run whole or in
// parts. cen[0..NQ-1] = start values in, final values out; ba holds NQ bases, then NQ
// accumulators.
// rk4_moist_prep reads the group's inputs and forms its stage-invariant terms, and returns the
// stage loop (taking cen[] and a per-stage hook), so that a caller can read the inputs of two
// groups before running either loop: their reads then go out as one batch.
template <typename S, typename C, bool CLAMP, int SRC, bool U, int SB, int Q0, int NQ, typename Nb, typename Vel>
ATHAD_DEV auto rk4_moist_prep(const MemberDesc& M, const Nb& src_nb, const CellGeo<C>& g, C* ba, Vel vel) {
    const Consts& K = M.K;
    const C i2r = C(K.inv_2dr), i2t = C(K.inv_2dthe), i2p = C(K.inv_2dphi);
    constexpr int NM = (Q0 < 3) ? ((Q0 + NQ < 3 ? Q0 + NQ : 3) - Q0) : 0;   // minmod fields
    C nb[NM > 0 ? NM : 1][6], L0[NQ], gc[NQ - NM > 0 ? NQ - NM : 1][3];
#pragma unroll
    for (int p = 0; p < NQ; p++) {
        const int q = Q0 + p;
        const int f = FC + q;
        C v[6];
        src_nb.six(f, v);
        const C d1r = (v[0] - v[1]) * i2r, d1t = (v[2] - v[3]) * i2t, d1p = (v[4] - v[5]) * i2p;
        L0[p] = (v[0] + v[1]) * g.wr + d1r * g.two_over + (v[2] + v[3]) * g.wt + d1t * g.crs + (v[4] + v[5]) * g.wp;
        if (p < NM) {
#pragma unroll
            for (int d = 0; d < 6; d++) nb[p][d] = v[d];
        } else {
            gc[p - NM][0] = d1r * g.cexp; gc[p - NM][1] = d1t * g.cirm; gc[p - NM][2] = d1p * g.cirs;
        }
    }
    const C dpr = C(K.diff_prec_re_inv), dco = C(K.diff_co2_re_inv);
    // microphysics sources of c, cloud, ice (RHS_Atm_Turb.cpp:1058-1067), stage-invariant
    C src[NM > 0 ? NM : 1];
#pragma unroll
    for (int p = 0; p < (NM > 0 ? NM : 1); p++) src[p] = C(0);
    if (SRC & SRC_ICE) {
        const C rh = src_nb.field(F_R_HUMID);
#pragma unroll
        for (int p = 0; p < NM; p++) src[p] = C(1) * src_nb.mp(MP_S_V + Q0 + p) * rh;
    }
    // moist convection's vapour tendency (RHS_Atm_Turb.cpp:1061), stage-invariant
    const C mcQ = ((SRC & SRC_CONV) && Q0 == 0) ? C(K.coeff_mc_q * src_nb.mc(MC_MC_Q)) : C(0);
    const C dt = C(K.dt), half_dt = C(0.5 * K.dt), dt_sixth = C(K.dt_sixth);
    const CellGeo<C> gg = g;
    return [=](C* cen, auto hook) {
    const CellGeo<C>& g = gg;
for_stages<U>([&](const int st) {
        hook(st);
        C u, v, w;
        vel(st, u, v, w);
        const C ua = u * g.aR, va = v * g.aT, wa = w * g.aP;
        C r[NQ];
#pragma unroll
        for (int p = 0; p < NQ; p++) {
            const int q = Q0 + p;
            if (p < NM) {
                r[p] = -minmod3(nb[p], cen[p], ua, va, wa) + (L0[p] - g.Lc2 * cen[p]) * dpr;
                if (SRC & SRC_ICE) r[p] = r[p] + src[p];
                if ((SRC & SRC_CONV) && q == 0) r[p] = r[p] + mcQ;
            } else {
                const C* gq = gc[p - NM];
                r[p] = -(u * gq[0] + v * gq[1] + w * gq[2]) + (L0[p] - g.Lc2 * cen[p]) * ((q == 4) ? dco : dpr);
            }
        }
#pragma unroll
        for (int p = 0; p < NQ; p++) {
            const C val = stage_update(st, r[p], ba + p * SB, ba + (NQ + p) * SB, dt, half_dt, dt_sixth);
            cen[p] = (Q0 + p < 4) ? tmax(C(0), val) : val;
        }
    });
    };
}

template <typename S, typename C, bool CLAMP, int SRC, bool U, int SB, int Q0, int NQ, typename Nb, typename Vel>
ATHAD_DEV void rk4_moist(const MemberDesc& M, const Nb& src_nb, const CellGeo<C>& g, C* ba,
                         Vel vel, C* cen) {
    rk4_moist_prep<S, C, CLAMP, SRC, U, SB, Q0, NQ>(M, src_nb, g, ba, vel)(cen, NoHook{});
}

// RK4 of the turbulence group tke, dis, advected by the recorded stage velocities, with the
// member's closure: k-omega SST (RHS_Atm_Turb.cpp:633-721), k-epsilon (:563-599) or k-omega
// (:601-631); the production refresh (:528-561) takes nu_t = C_mu k^2 / eps in the k-epsilon
// closure, k / omega otherwise. cen[0..1] = start values in, final values out; prod and the
// sources are those of the fourth stage.
// rk4_turb_prep: the reads and stage-invariant terms; returns the stage loop (taking cen[], the
// production invariants A, B of the dynamics group, prod, tsrc, dsrc and a per-stage hook), as
// rk4_moist_prep does.
template <typename S, typename C, bool CLAMP, bool U, int SB, typename Nb, typename Vel>
ATHAD_DEV auto rk4_turb_prep(const MemberDesc& M, const Nb& src_nb, int i, int j, int k,
                             const CellGeo<C>& g, C* ba, Vel vel) {
    const Consts& K = M.K;
    const C tkm = C(K.tke_max_nd);
    const C i2r = C(K.inv_2dr), i2t = C(K.inv_2dthe), i2p = C(K.inv_2dphi);
    C L0[2], gc[2][3], d1[2][3];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int f = FK + q;
        C v[6];
        src_nb.six(f, v);
        d1[q][0] = (v[0] - v[1]) * i2r; d1[q][1] = (v[2] - v[3]) * i2t; d1[q][2] = (v[4] - v[5]) * i2p;
        L0[q] = (v[0] + v[1]) * g.wr + d1[q][0] * g.two_over + (v[2] + v[3]) * g.wt + d1[q][1] * g.crs
              + (v[4] + v[5]) * g.wp;
        gc[q][0] = d1[q][0] * g.cexp; gc[q][1] = d1[q][1] * g.cirm; gc[q][2] = d1[q][2] * g.cirs;
    }
    const C cross = d1[0][0] * d1[1][0] + (d1[0][1] * g.cirm) * (d1[1][1] * g.cirm)
                  + (d1[0][2] * g.cirs) * (d1[1][2] * g.cirs);
    C nue = src_nb.field(F_NUE);
    if (CLAMP) nue = finite_bits(nue) ? tclamp(nue, C(0), C(K.nue_max)) : C(0);
    // SST wall distance min(i, j, k) (RHS_Atm_Turb.cpp:640-671 without land); the symmetric
    // variant measures j from the nearer pole.
    C K1, K2, K3;
    {
        double dc = (double)i;
        const double dj = K.sym ? (double)min(j, K.jm - 1 - j) : (double)j;
        if (dj <= dc) dc = dj;
        if ((double)k <= dc) dc = (double)k;
        K1 = C(500.0 * K.nue_air_nd / (dc * dc));
        K2 = C(1.0 / (0.09 * dc));
        K3 = C(4.0 * 1.168 / (dc * dc));
    }
    const C sig_k1 = C(1.176), sig_k2 = C(1.0), sig_w1 = C(2.0), sig_w2 = C(1.168), bet_star = C(0.09);
    const C bet1 = C(0.0333), bet2 = C(0.0368), gam1 = C(0.413), gam2 = C(0.2);
    const C inv_re = C(K.inv_re_turb);
    const C dt = C(K.dt), half_dt = C(0.5 * K.dt), dt_sixth = C(K.dt_sixth);
    const int model = K.turb;
    // k-epsilon and k-omega: the diffusion coefficients and the cross-diffusion product
    // sigma_d * cross do not change between stages
    const bool eps = (model == kTurbKEpsilon);
    const C dif_tke_2 = eps ? inv_re + nue / C(1.0) : inv_re + C(0.6) * nue;
    const C dif_dis_2 = eps ? inv_re + nue / C(1.3) : inv_re + C(0.5) * nue;
    const C sig_cross = (eps || cross <= C(0)) ? C(0) : C(0.125) * cross;
    const C Lc2 = g.Lc2, nue_max_c = C(K.nue_max);
    return [=](C* cen, C A, C B, C& prod, C& tsrc, C& dsrc, auto hook) {
for_stages<U>([&](const int st) {
        hook(st);
        C u, v, w;
        vel(st, u, v, w);
        const C tke = cen[0], dis = cen[1];
        // Stage inputs satisfy dis >= 1e-10 and tke >= 0 (start values pass the solid-ground
        // clamp, stage outputs are floored), so max(dis, 1e-10) == dis and one reciprocal
        // serves every 1/dis.
        const C dis_s = tmax(dis, C(1.0e-10));
        const C inv_dis = rcp_nr(dis_s);
        const C tke_s = tmax(tke, C(1.0e-10));
        C dif_tke, dif_dis;
        if (model == kTurbKOmegaSST) {
            const C cnue = tmin(tmax(tke, C(0)) * inv_dis, nue_max_c);
            prod = tmax(C(0), cnue * A - C(0.66667) * tke * B);
            const C CD_kw = tmax(C(2) * sig_w2 * inv_dis * cross, C(1.0e-20));
            const C arg1 = tmin(tmax(C(sqrt(tke)) * inv_dis * K2, K1 * inv_dis), tke * K3 * rcp_nr(CD_kw));
            const C a2 = arg1 * arg1;
            const C F1 = tanh_pos<C, U>(a2 * a2);
            dif_tke = inv_re + nue * rcp_nr(F1 * sig_k1 + (C(1) - F1) * sig_k2);
            dif_dis = inv_re + nue * rcp_nr(F1 * sig_w1 + (C(1) - F1) * sig_w2);
            const C P_k = tmin(prod, C(20) * bet_star * tke_s * dis_s);
            const C Y_k = bet_star * tke_s * dis_s;
            const C P_w = (F1 * gam1 + (C(1) - F1) * gam2) * P_k * dis_s * rcp_nr(tke_s);
            const C Y_w = (F1 * bet1 + (C(1) - F1) * bet2) * dis_s * dis_s;
            const C D_w = C(2) * (C(1) - F1) * sig_w2 * inv_dis * cross;
            const C dsm = C(20) * dis_s * dis_s, tsm = C(20) * tke_s * dis_s;
            tsrc = tclamp(P_k - Y_k, -tsm, tsm);
            dsrc = tclamp(P_w - Y_w + D_w, -dsm, dsm);
        } else {
            const C tke_h = tmax(tke, C(0));
            const C cnue = tmin(eps ? C(0.028) * tke_h * tke_h * inv_dis : tke_h * inv_dis, nue_max_c);
            prod = tmax(C(0), cnue * A - C(0.66667) * tke * B);
            dif_tke = dif_tke_2;
            dif_dis = dif_dis_2;
            const C ratio = dis_s * rcp_nr(tke_s);   // dis / tke
            C P_k, Y_k, P_w, Y_w;
            if (eps) {   // Chien's k-epsilon without its wall terms
                P_k = prod;
                Y_k = dis_s;
                P_w = C(1.35) * ratio * P_k;
                Y_w = C(1.80) * dis_s * ratio;
            } else {     // Wilcox's k-omega
                P_k = tmin(prod, C(20) * C(0.09) * tke_s * dis_s);
                Y_k = C(0.09) * tke_s * dis_s;
                P_w = C(0.52) * ratio * P_k;
                Y_w = C(0.0708) * dis_s * dis_s;
            }
            const C dsm = C(20) * dis_s * dis_s, tsm = C(20) * tke_s * dis_s;
            tsrc = tclamp(P_k - Y_k, -tsm, tsm);
            dsrc = tclamp(P_w - Y_w + sig_cross * inv_dis, -dsm, dsm);
        }
        const C rk = -(u * gc[0][0] + v * gc[0][1] + w * gc[0][2]) + (L0[0] - Lc2 * tke) * dif_tke + tsrc;
        const C rd = -(u * gc[1][0] + v * gc[1][1] + w * gc[1][2]) + (L0[1] - Lc2 * dis) * dif_dis + dsrc;
        C vk = stage_update(st, rk, ba + 0 * SB, ba + 2 * SB, dt, half_dt, dt_sixth);
        vk = finite_bits(vk) ? ((vk < C(0)) ? C(0) : (vk > tkm) ? tkm : vk) : C(0);
        const C vd = stage_update(st, rd, ba + 1 * SB, ba + 3 * SB, dt, half_dt, dt_sixth);
        cen[0] = vk;
        cen[1] = tmax(C(1.0e-10), vd);
    });
    };
}

template <typename S, typename C, bool CLAMP, bool U, int SB, typename Nb, typename Vel>
ATHAD_DEV void rk4_turb(const MemberDesc& M, const Nb& src_nb, int i, int j, int k,
                        const CellGeo<C>& g, C* ba, Vel vel, C A, C B, C* cen, C& prod, C& tsrc, C& dsrc) {
    rk4_turb_prep<S, C, CLAMP, U, SB>(M, src_nb, i, j, k, g, ba, vel)(cen, A, B, prod, tsrc, dsrc, NoHook{});
}

// Waves per SIMD the RK4 kernels are compiled for (register budget 512 / waves).
#ifndef ATHAD_RK4_WAVES_F64
#define ATHAD_RK4_WAVES_F64 3
#endif
#ifndef ATHAD_RK4_WAVES_F32
#define ATHAD_RK4_WAVES_F32 5
#endif
template <typename S> constexpr int rk4_waves() { return sizeof(S) == 8 ? ATHAD_RK4_WAVES_F64 : ATHAD_RK4_WAVES_F32; }
#ifndef ATHAD_RK4_ROW_WAVES_F64
#define ATHAD_RK4_ROW_WAVES_F64 3
#endif
#ifndef ATHAD_RK4_ROW_WAVES_F32
#define ATHAD_RK4_ROW_WAVES_F32 5
#endif
template <typename S> constexpr int rk4_row_waves() { return sizeof(S) == 8 ? ATHAD_RK4_ROW_WAVES_F64 : ATHAD_RK4_ROW_WAVES_F32; }

// Start values (optionally clamped) into cen[] and the bases into the LDS slots of a group.
template <int SB, typename C, typename Nb>
ATHAD_DEV void group_start(const Nb& src_nb, int f0, int nf, C* cen, C* ba) {
#pragma unroll
    for (int q = 0; q < nf; q++) {
        cen[q] = src_nb.start(f0 + q);
        ba[q * SB] = src_nb.base(f0 + q);
    }
}

// RK4 of one interior cell, group by group: dynamics, moisture, turbulence. out(f, value)
// writes a final value; the aux / prod / source outputs are written here.
// GT: where the metric coefficients come from (0: formed from the 1-D tables, 1: the (i, j)
// table read per lane, 2: the (i, j) table read by scalar loads, i and j wave-uniform).
template <typename S, typename C, bool CLAMP, int SRC, bool U, int GT, bool SPLIT, int SB, typename Nb, typename Out, typename OutX>
ATHAD_DEV void rk4_cell(const MemberDesc& M, const Nb& nb, int i, int j, int k, C* ba, C* rec, Out out, OutX outx) {
    CellGeo<C> g;
    if constexpr (GT == 2) cell_geo_row<C>(M, i, j, nb.ps(), nb.psurf(), g);
    else if constexpr (GT == 1) cell_geo<C>(M, i, j, nb.ps(), nb.psurf(), g);
    else cell_geo_alu<C>(M, i, j, nb.ps(), nb.psurf(), g);
    C A, B;
    {
        C cen[4], aux[3];
        group_start<SB>(nb, FT, 4, cen, ba);
        rk4_dyn<S, C, CLAMP, SRC, U, SB>(M, nb, i, g, ba, rec, cen, aux, A, B);
#pragma unroll
        for (int q = 0; q < 4; q++) out(FT + q, cen[q]);
        outx(F_AUX_U, aux[0]);
        outx(F_AUX_V, aux[1]);
        outx(F_AUX_W, aux[2]);
    }
    auto vel = [&](int st, C& u, C& v, C& w) {
        u = rec[(3 * st + 0) * SB]; v = rec[(3 * st + 1) * SB]; w = rec[(3 * st + 2) * SB];
    };
    if constexpr (SPLIT) {
        // c, cloud, ice (six LDS slots), then gr, co2 (slots 0-3) together with tke, dis
        // (slots 4-7): the inputs of both of the last two groups are read before either stage
        // loop runs, one batch of reads instead of two round trips
        {
            C cen[3];
            group_start<SB>(nb, FC, 3, cen, ba);
            rk4_moist<S, C, CLAMP, SRC, U, SB, 0, 3>(M, nb, g, ba, vel, cen);
#pragma unroll
            for (int q = 0; q < 3; q++) out(FC + q, cen[q]);
        }
        C cen_g[2], cen_t[2], prod, ts, ds;
        group_start<SB>(nb, FC + 3, 2, cen_g, ba);
        group_start<SB>(nb, FK, 2, cen_t, ba + 4 * SB);
        auto run_g = rk4_moist_prep<S, C, CLAMP, SRC, U, SB, 3, 2>(M, nb, g, ba, vel);
        auto run_t = rk4_turb_prep<S, C, CLAMP, U, SB>(M, nb, i, j, k, g, ba + 4 * SB, vel);
        run_g(cen_g, NoHook{});
#pragma unroll
        for (int q = 0; q < 2; q++) out(FC + 3 + q, cen_g[q]);
        run_t(cen_t, A, B, prod, ts, ds, NoHook{});
        out(FK, cen_t[0]);
        out(FD, cen_t[1]);
        outx(F_PROD, prod);
        outx(F_TKE_SRC, ts);
        outx(F_DIS_SRC, ds);
    } else {
        C cen[5];
        group_start<SB>(nb, FC, 5, cen, ba);
        rk4_moist<S, C, CLAMP, SRC, U, SB, 0, 5>(M, nb, g, ba, vel, cen);
#pragma unroll
        for (int q = 0; q < 5; q++) out(FC + q, cen[q]);
        C cen_t[2], prod, ts, ds;
        group_start<SB>(nb, FK, 2, cen_t, ba);
        rk4_turb<S, C, CLAMP, U, SB>(M, nb, i, j, k, g, ba, vel, A, B, cen_t, prod, ts, ds);
        out(FK, cen_t[0]);
        out(FD, cen_t[1]);
        outx(F_PROD, prod);
        outx(F_TKE_SRC, ts);
        outx(F_DIS_SRC, ds);
    }
}

// A boundary cell of the Jacobi-order RK4: its prognostic values copied to alt[] (clamped and
// also stored when FUSED), and with FUSED the clamps of nue, prod and the two sources. Every
// value is read before any is written: the compiler cannot tell that the fields do not overlap,
// so a read placed after a write waits for that write, one memory round trip per field.
template <typename S, bool FUSED>
ATHAD_DEV void rk4_boundary_copy(const MemberDesc& M, const S* const* X, long long id, S tkm, S vm) {
    using C = S;
    const Consts& K = M.K;
    C v[NPROG], x[4];
#pragma unroll
    for (int f = 0; f < NPROG; f++) v[f] = C(gld(X[f], id));
    if (FUSED) {
        x[0] = C(gld(fp<S>(M, F_NUE), id));
        x[1] = C(gld(fp<S>(M, F_PROD), id));
        x[2] = C(gld(fp<S>(M, F_TKE_SRC), id));
        x[3] = C(gld(fp<S>(M, F_DIS_SRC), id));
    }
#pragma unroll
    for (int f = 0; f < NPROG; f++) {
        if (FUSED) v[f] = solid_clamp(f, v[f], tkm, vm);
        gst((S*)M.alt[f], id, S(v[f]));
        if (FUSED && f != FU && f != FV && f != FW) gst(fp<S>(M, kProgN[f]), id, S(v[f]));
    }
    if (FUSED) {
        gst(fp<S>(M, F_NUE), id, S(finite_bits(x[0]) ? tclamp(x[0], C(0), C(K.nue_max)) : C(0)));
        gst(fp<S>(M, F_PROD), id, S(finite_bits(x[1]) ? tclamp(x[1], C(0), C(1.0e4)) : C(0)));
        gst(fp<S>(M, F_TKE_SRC), id, S(finite_bits(x[2]) ? tclamp(x[2], C(-1.0e4), C(1.0e4)) : C(-1.0e4)));
        gst(fp<S>(M, F_DIS_SRC), id, S(finite_bits(x[3]) ? tclamp(x[3], C(-1.0e4), C(1.0e4)) : C(-1.0e4)));
    }
}

// Jacobi-order RK4 over the whole grid, one thread per cell in storage order (the FP32 form;
// FP64 runs k_rk4_rows). Interior cells integrate four stages against the start-of-RK4
// neighbourhood read from f[] and write alt[]; boundary cells are copied. The solid-ground
// clamps are applied on read (FUSED=1), and the store of the eight fields that no later filter
// touches is fused (xn = new value).
template <typename S, bool FUSED, int SRC>
__global__ __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(rk4_waves<S>()))) void k_rk4_jacobi(const MemberDesc* __restrict__ md) {
    using C = S;
    __shared__ C s_ba[8 * LB], s_rec[12 * LB];
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long id = cell_id(K);
    if (id < 0) return;
    int i, j, k;
    decode(K, id, i, j, k);
    const S* X[NPROG];
#pragma unroll
    for (int f = 0; f < NPROG; f++) X[f] = fp<S>(M, kProgF[f]);
    const C tkm = C(K.tke_max_nd), vm = C(K.vel_max_nd);
    const bool interior = (i > 0 && i < K.im - 1 && j > 0 && j < K.jm - 1 && k > 0 && k < K.km - 1);
    if (!interior) {
        rk4_boundary_copy<S, FUSED>(M, X, id, tkm, vm);
        return;
    }
    // buffer-resource addressing as in the row kernel; the metric table read per lane in FP32
    // (5 waves per SIMD), formed in place in FP64
    constexpr unsigned E = sizeof(S);
    const unsigned si = (unsigned)(K.jm * K.km), nfld = (unsigned)K.n * E;
    const NbRow<S, C, FUSED> nb{M, (unsigned)id * E, (unsigned)(j * K.km + k) * 8u, si * E, (unsigned)K.km * E,
                                nfld + (2u * si + 2u) * E, nfld, tkm, vm};
    rk4_cell<S, C, FUSED, SRC, false, sizeof(S) == 4, true, LB>(M, nb, i, j, k, s_ba + threadIdx.x, s_rec + threadIdx.x,
        [&](int f, C v) {
            nb.put(M.alt[f], v);
            if (FUSED && f != FU && f != FV && f != FW) nb.put(M.f[kProgN[f]], v);
        },
        [&](int fid, C v) { nb.put(M.f[fid], v); });
}

// The metric coefficients of row (i, j) when i and j are the same for the whole wave: the
// table entry is read through the constant address space, so the values arrive by scalar
// loads into SGPRs and enter the arithmetic as scalar operands instead of occupying twenty
// vector registers per lane. Only the Held-Suarez rate depends on the cell (through p_stat).
template <typename C>
ATHAD_DEV void cell_geo_row(const MemberDesc& M, int i, int j, double ps_c, double p_surf, CellGeo<C>& g) {
    const Consts& K = M.K;
    typedef const __attribute__((address_space(4))) double* cptr;
    const cptr q = (cptr)(M.geo + ((long long)i * K.jm + j) * NGEO);
    g.aR = C(q[G_AR]); g.aT = C(q[G_AT]); g.aP = C(q[G_AP]);
    g.cexp = C(q[G_CEXP]); g.cirm = C(q[G_CIRM]); g.cirs = C(q[G_CIRS]);
    g.wr = C(q[G_WR]); g.wt = C(q[G_WT]); g.wp = C(q[G_WP]);
    g.two_over = C(q[G_TWO_OVER]);
    g.crs = C(q[G_CRS]);
    g.Lc2 = C(q[G_LC2]);
    g.cu = C(q[G_CU]);
    g.cvw = C(q[G_CVW]);
    g.cirm2 = C(q[G_CIRM2]); g.cirm2s = C(q[G_CIRM2S]); g.c2cos = C(q[G_C2COS]);
    const double sigma = hs_sigma(ps_c, p_surf);
    double bl = (sigma - 0.7) / (1.0 - 0.7);
    if (bl < 0.0) bl = 0.0;
    const double kT = K.hs_k_a + (K.hs_k_s - K.hs_k_a) * bl * q[G_S4];
    g.hs_kT = C(kT * K.L_atm / K.u_0);
    g.drag = C(q[G_DRAG]);
    g.cor = C(q[G_COR]);
}

// Jacobi-order RK4 with one wave per row segment (same arithmetic and outputs as
// k_rk4_jacobi): a wave owns longitudes 64 s .. 64 s + 63 of one (i, j) row, so i and j are
// wave-uniform, the metric coefficients sit in scalar registers, the interior test of i and j
// is a uniform branch, and the cell index needs no division. A block's four waves take four
// neighbouring latitude rows j = 4 jb .. 4 jb + 3 at one level and segment, so each row's j - 1
// and j + 1 neighbours are the rows its sibling waves load at the same time; blocks run i
// fastest (the i +- 1 rows are the neighbouring blocks'), then the segment, then jb.
template <typename S, bool FUSED, int SRC>
__global__ __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(rk4_row_waves<S>()))) void k_rk4_rows(const MemberDesc* __restrict__ md) {
    using C = S;
    // 20 slots per thread (40 KB per block in FP64): four blocks of four waves fit a CU's LDS
    __shared__ C s_ba[8 * LB], s_rec[12 * LB];
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const int nseg = (K.km + 63) >> 6;
    const long long nblk = (long long)K.im * ((K.jm + 3) >> 2) * nseg;
    const long long b = xcd_block(blockIdx.x, gridDim.x, nblk);
    if (b < 0) return;
    const int i = (int)(b % K.im);
    const long long rb = b / K.im;
    const int seg = (int)(rb % nseg);
    const int j = (int)(rb / nseg) * 4 + (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (j >= K.jm) return;   // the last group of four rows where jm is not a multiple of 4
    const int k = seg * 64 + (int)(threadIdx.x & 63);
    if (k >= K.km) return;
    const long long row = ((long long)i * K.jm + j) * K.km;
    const long long id = row + k;
    const S* X[NPROG];
#pragma unroll
    for (int f = 0; f < NPROG; f++) X[f] = fp<S>(M, kProgF[f]);
    const C tkm = C(K.tke_max_nd), vm = C(K.vel_max_nd);
    const bool interior = (i > 0 && i < K.im - 1 && j > 0 && j < K.jm - 1 && k > 0 && k < K.km - 1);
    if (!interior) {
        rk4_boundary_copy<S, FUSED>(M, X, id, tkm, vm);
        return;
    }
    constexpr unsigned E = sizeof(S);
    const unsigned si = (unsigned)(K.jm * K.km), nfld = (unsigned)K.n * E;
    const NbRow<S, C, FUSED> nb{M, (unsigned)id * E, (unsigned)(j * K.km + k) * 8u, si * E, (unsigned)K.km * E,
                                nfld + (2u * si + 2u) * E, nfld, tkm, vm};
    rk4_cell<S, C, FUSED, SRC, false, 2, true, LB>(M, nb, i, j, k, s_ba + threadIdx.x, s_rec + threadIdx.x, [&](int f, C v) {
        nb.put(M.alt[f], v);
        if (FUSED && f != FU && f != FV && f != FW) nb.put(M.f[kProgN[f]], v);
    }, [&](int fid, C v) { nb.put(M.f[fid], v); });
}

// Lexicographic-order RK4 restricted to the hyperplane i+j+k = d (in place). Cells of one
// plane only read neighbours on planes d-1 (final) and d+1 (not yet updated), which is
// exactly what the reference's single-threaded sweep sees. A plane's time is one wave's RK4
// chain; LBX threads per block: 64 when a plane has few waves, so each wave has a compute
// unit to itself (15.9 -> 9.7 ms per iteration at 41 x 181 x 361), 256 when many members
// share the launch and fill the device anyway.
template <typename S, int SRC, int LBX>
__global__ __launch_bounds__(LBX) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_rk4_lex_plane(const MemberDesc* __restrict__ md, int d) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const int nj = K.jm - 2;
    const long long qd = (long long)blockIdx.x * LBX + threadIdx.x;
    if (qd >= (long long)(K.im - 2) * nj) return;
    const int i = 1 + (int)(qd / nj), j = 1 + (int)(qd % nj);
    const int k = d - i - j;
    if (k < 1 || k > K.km - 2) return;
    const long long id = (long long)ix3(K, i, j, k);
    S* X[NPROG];
#pragma unroll
    for (int f = 0; f < NPROG; f++) X[f] = fp<S>(M, kProgF[f]);
    const NbRegs<S, C, SRC> nb(M, (const S* const*)X, id, (long long)K.jm * K.km, (long long)K.km,
                               (long long)ix3(K, 0, j, k));
    // one wave per SIMD leaves registers to spare: the RK4 base / accumulator slots and the
    // recorded stage velocities live in registers (stride 1, every index fixed once the stages
    // are unrolled) instead of LDS, off the latency of the serial chain
    C r_ba[10], r_rec[12];
    rk4_cell<S, C, false, SRC, true, true, false, 1>(M, nb, i, j, k, r_ba, r_rec,
                          [&](int f, C v) { gst(X[f], id, S(v)); },
                          [&](int fid, C v) { gst(fp<S>(M, fid), id, S(v)); });
}

// Lexicographic-order RK4 on plane d with the groups of each cell on three waves of one block:
// wave 0 the dynamics group, wave 1 the moisture group, wave 2 the turbulence group, all for the
// same 64 cells. The scalar groups need from the dynamics group only the velocity entering each
// stage and the production invariants A, B; the dynamics wave leaves them in LDS before a
// barrier at the start of every stage, so the turbulence group (half of a cell's arithmetic
// with the SST closure) runs each stage as soon as its velocity exists and the plane's chain is
// about that group's instead of the three groups' one after another. The same arithmetic as
// k_rk4_lex_plane. All three waves hold the same cells, so a lane off the plane is idle in each
// and a wave with no cell on the plane exits with its two siblings: every wave that runs meets
// the same four barriers.
template <typename S, int SRC>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_rk4_lex_roles(const MemberDesc* __restrict__ md, int d) {
    using C = S;
    constexpr int W = 64;
    __shared__ C s_bd[8 * W], s_rec[12 * W], s_ab[2 * W];
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const int role = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    const int nj = K.jm - 2;
    const long long qd = (long long)blockIdx.x * W + lane;
    if (qd >= (long long)(K.im - 2) * nj) return;
    const int i = 1 + (int)(qd / nj), j = 1 + (int)(qd % nj);
    const int k = d - i - j;
    if (k < 1 || k > K.km - 2) return;
    const long long id = (long long)ix3(K, i, j, k);
    S* X[NPROG];
#pragma unroll
    for (int f = 0; f < NPROG; f++) X[f] = fp<S>(M, kProgF[f]);
    // each wave keeps only the reads of its own group (the compiler drops the others)
    const NbRegs<S, C, SRC> nb(M, (const S* const*)X, id, (long long)K.jm * K.km, (long long)K.km,
                               (long long)ix3(K, 0, j, k));
    CellGeo<C> g;
    cell_geo<C>(M, i, j, nb.ps(), nb.psurf(), g);
    auto vel = [&](int st, C& u, C& v, C& w) {
        u = s_rec[(3 * st + 0) * W + lane]; v = s_rec[(3 * st + 1) * W + lane]; w = s_rec[(3 * st + 2) * W + lane];
    };
    if (role == 0) {
        C cen[4], aux[3], A, B;
        group_start<W>(nb, FT, 4, cen, s_bd + lane);
        rk4_dyn<S, C, false, SRC, true, W>(M, nb, i, g, s_bd + lane, s_rec + lane, cen, aux, A, B, [&](int st) {
            if (st == 0) { s_ab[lane] = A; s_ab[W + lane] = B; }
            __syncthreads();
        });
#pragma unroll
        for (int q = 0; q < 4; q++) gst(X[FT + q], id, S(cen[q]));
        gst(fp<S>(M, F_AUX_U), id, S(aux[0]));
        gst(fp<S>(M, F_AUX_V), id, S(aux[1]));
        gst(fp<S>(M, F_AUX_W), id, S(aux[2]));
    } else if (role == 1) {
        C cen[5], ba[10];
        group_start<1>(nb, FC, 5, cen, ba);
        rk4_moist_prep<S, C, false, SRC, true, 1, 0, 5>(M, nb, g, ba, vel)(cen, [&](int) { __syncthreads(); });
#pragma unroll
        for (int q = 0; q < 5; q++) gst(X[FC + q], id, S(cen[q]));
    } else {
        C cen[2], ba[4], prod, ts, ds;
        group_start<1>(nb, FK, 2, cen, ba);
        auto run = rk4_turb_prep<S, C, false, true, 1>(M, nb, i, j, k, g, ba, vel);
        __syncthreads();   // stage 0: A, B and the start velocities are in LDS
        const C A = s_ab[lane], B = s_ab[W + lane];
        run(cen, A, B, prod, ts, ds, [&](int st) {
            if (st > 0) __syncthreads();
        });
        gst(X[FK], id, S(cen[0]));
        gst(X[FD], id, S(cen[1]));
        gst(fp<S>(M, F_PROD), id, S(prod));
        gst(fp<S>(M, F_TKE_SRC), id, S(ts));
        gst(fp<S>(M, F_DIS_SRC), id, S(ds));
    }
}

// ------------------------------------------------------------------------------------
// Poisson red-black sweep (PressureSolverAtm.h:106-121, :225-486, :504-594)
// ------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_aux_bc(const MemberDesc* __restrict__ md, int gate) {
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= (long long)(K.jm - 2) * (K.km - 2)) return;
    const int j = 1 + (int)(q / (K.km - 2)), k = 1 + (int)(q % (K.km - 2));
    // all six values read before any is written (one memory round trip, not one per copy)
    S lo[3], hi[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const S* a = fp<S>(M, F_AUX_U + f);
        lo[f] = a[ix3(K, 1, j, k)];
        hi[f] = a[ix3(K, K.im - 2, j, k)];
    }
#pragma unroll
    for (int f = 0; f < 3; f++) {
        S* a = fp<S>(M, F_AUX_U + f);
        a[ix3(K, 0, j, k)] = lo[f];
        a[ix3(K, K.im - 1, j, k)] = hi[f];
    }
}

template <typename S>
__global__ __launch_bounds__(256) void k_poisson_colour(const MemberDesc* __restrict__ md, int colour, int gate) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long id = cell_id(K);
    if (id < 0) return;
    int i, j, k;
    decode(K, id, i, j, k);
    if (i < 1 || i > K.im - 2 || j < 1 || j > K.jm - 2 || k < 1 || k > K.km - 2) return;
    if (((i + j + k) & 1) != colour) return;
    const double rm = M.rad[i];
    const double exp_rm = 1.0 / (rm + 1.0);
    const double exp_2_rm = exp_rm * exp_rm;
    const double sinthe = M.sint[j];
    const double sinthe2 = sinthe * sinthe;
    const double rmet = (K.metric_r0 > 0.0) ? (K.metric_r0 + (rm - K.rad0)) : rm;
    const double inv_rm = 1.0 / rmet;
    const double inv_rm2 = 1.0 / (rmet * rmet);
    const double inv_rmsinthe = 1.0 / (rmet * sinthe);
    const double m_phi = inv_rm2 / sinthe2;
    const double denom = 2.0 * exp_2_rm * K.inv_dr2 + 2.0 * inv_rm2 * K.inv_dthe2 + 2.0 * m_phi * K.inv_dphi2;
    const C inv_denom = C(1.0 / denom);
    const C num1 = C(exp_2_rm * K.inv_dr2), num2 = C(inv_rm2 * K.inv_dthe2), num3 = C(m_phi * K.inv_dphi2);
    const size_t si = (size_t)K.jm * K.km, sj = K.km;
    const S *au = fp<S>(M, F_AUX_U), *av = fp<S>(M, F_AUX_V), *aw = fp<S>(M, F_AUX_W);
    S* p = fp<S>(M, F_P_DYN);
    const C du_dr = (C(au[id + si]) - C(au[id - si])) * C(K.inv_2dr);
    const C dv_dthe = (C(av[id + sj]) - C(av[id - sj])) * C(K.inv_2dthe);
    const C dw_dphi = (C(aw[id + 1]) - C(aw[id - 1])) * C(K.inv_2dphi);
    C div_src = du_dr * C(exp_rm) + dv_dthe * C(inv_rm) + dw_dphi * C(inv_rmsinthe);
    const C src_max = C(denom * 2.0);
    if (!finite_bits(div_src)) div_src = C(0);
    else if (div_src > src_max) div_src = src_max;
    else if (div_src < -src_max) div_src = -src_max;
    if constexpr (sizeof(S) == 8) {
        p[id] = ((C(p[id + si]) + C(p[id - si])) * num1 + (C(p[id + sj]) + C(p[id - sj])) * num2
               + (C(p[id + 1]) + C(p[id - 1])) * num3 - div_src) * inv_denom;
    } else {
        // The mixed form writes the sweep as a defect correction: the defect of the cell's
        // equation denom p - sum num p_nb = -src is evaluated in FP64 from the FP32 values, and
        // the FP32 correction defect / denom is added to p. Algebraically the same update.
        const double pc = (double)p[id];
        const double defect = -(double)div_src - (denom * pc - ((double)p[id + si] + (double)p[id - si]) * (exp_2_rm * K.inv_dr2)
                                                  - ((double)p[id + sj] + (double)p[id - sj]) * (inv_rm2 * K.inv_dthe2)
                                                  - ((double)p[id + 1] + (double)p[id - 1]) * (m_phi * K.inv_dphi2));
        p[id] = S(pc + (double)(C(defect) * inv_denom));
    }
}

// Boundary values of p_dyn, composed: radial copy, then pole copy, then the seam average
// (PressureSolverAtm.h:504-552). Only boundary cells are written and only interior cells
// are read, so one launch applies all three in the reference's order.
template <typename S>
__global__ void k_p_bc(const MemberDesc* __restrict__ md, int gate) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long id = cell_id(K);
    if (id < 0) return;
    int i, j, k;
    decode(K, id, i, j, k);
    const bool bi = (i == 0 || i == K.im - 1), bj = (j == 0 || j == K.jm - 1), bk = (k == 0 || k == K.km - 1);
    if (!(bi || bj || bk)) return;
    const int ci = min(max(i, 1), K.im - 2), cj = min(max(j, 1), K.jm - 2);
    S* p = fp<S>(M, F_P_DYN);
    C v;
    if (bk) v = (C(p[ix3(K, ci, cj, 1)]) + C(p[ix3(K, ci, cj, K.km - 2)])) / C(2);
    else v = C(p[ix3(K, ci, cj, k)]);
    p[id] = S(v);
}

template <typename S>
__global__ void k_p_ceiling(const MemberDesc* __restrict__ md, int gate) {
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const long long id = cell_id(M.K);
    if (id < 0) return;
    const S c = S((ro_load(M.counters + 1) > 300) ? 3.0 : 10.0);
    S* p = fp<S>(M, F_P_DYN);
    const S v = p[id];
    if (!finite_bits(v)) p[id] = S(0);
    else if (v > c) p[id] = c;
    else if (v < -c) p[id] = -c;
}

// ------------------------------------------------------------------------------------
// SaturationAdjustment (SaturationAdjustment.h:62-451; IceSchemeCommon.h:121-190) and the
// cap on S_c_c (cAtmosphereModel.cpp:1278-1295). Pointwise, FP64 arithmetic.
// ------------------------------------------------------------------------------------
// Newton-loop helpers of the saturation adjustment, templated on the Newton arithmetic type N
// (double in the drop-in form, float in the mixed form). They restate th_* with shared
// sub-expressions evaluated once and divisions by positive bounded operands done through
// refined reciprocals. In the mixed form the loop's state, latent heat and heat capacity are
// FP32 while the saturation vapour pressures and q_sat stay FP64 (sat_newton); the entry test
// and the clamp around it are FP64 too.
template <typename N> ATHAD_DEV void sat_species_cp(N T, N Mb, N& cpH, N& cpC, N& cpB) {
    const N t = tmin(tmax(T, N(298)), N(2000)) * N(1.0e-3);
    const N t2 = t * t, t3 = t2 * t, it2 = rcp_nr(t2);
    cpH = (N(30.09200) + N(6.832514) * t + N(6.793435) * t2 + N(-2.534480) * t3 + N(0.082139) * it2) * N(1.0 / kMH2O);
    cpC = (N(24.99735) + N(55.18696) * t + N(-33.69137) * t2 + N(7.948387) * t3 + N(-0.136638) * it2) * N(1.0 / kMCO2);
    cpB = (N(28.98641) + N(1.853978) * t + N(-9.647459) * t2 + N(16.63537) * t3 + N(0.000117) * it2) / Mb;
}
template <typename N> ATHAD_DEV N sat_mix_cp(N c, N co2, N cpH, N cpC, N cpB) {
    N qv = tmin(tmax(c, N(0)), N(1)), qc = tmin(tmax(co2, N(0)), N(1));
    const N s = qv + qc;
    if (s > N(1)) { qv /= s; qc /= s; }
    const N qb = tmax(N(0), N(1) - qv - qc);
    return qv * cpH + qc * cpC + qb * cpB;
}
template <typename N> ATHAD_DEV N sat_L_vap(N T) {
    if (T >= N(kTcrit)) return N(0);
    return N(2.501e6) * exp(N(0.38) * log_fast((N(kTcrit) - T) * N(1.0 / (kTcrit - 273.15))));
}
template <typename A> ATHAD_DEV A sat_p_sat(A T) {
    if (T >= A(kTcrit)) return A(kNoSat);
    if (T <= A(0)) return A(0);
    const A th = A(1) - T * A(1.0 / kTcrit);
    const A sq = sqrt(th);
    const A t3 = th * th * th;
    const A t4 = t3 * th;
    const A s = A(-7.85951783) * th + A(1.84408259) * (th * sq) + A(-11.7866497) * t3
              + A(22.6807411) * (t3 * sq) + A(-15.9618719) * t4 + A(1.80122502) * (t4 * t3 * sq);
    return A(kPcrit) * exp((A(kTcrit) * rcp_nr(T)) * s);
}
template <typename A> ATHAD_DEV A sat_p_subl(A T) {
    if (T <= A(0)) return A(0);
    const A th = T * A(1.0 / kTtriple);
    const A lth = log_fast(th);
    const A s = A(-0.212144006e2) * exp(A(0.333333333e-2) * lth) + A(0.273203819e2) * exp(A(0.120666667e1) * lth)
              + A(-0.610598130e1) * exp(A(0.170333333e1) * lth);
    return A(kPtriple) * exp(s * rcp_nr(th));
}
// exact q_sat (MixtureAtm.h:152-217) and the factor Af of dq_sat/dT = Af L / (R_v T^2)
template <typename A> ATHAD_DEV void sat_qA(A E, A p, A inv_p, A Mo, A& q, A& Af) {
    Af = A(0);
    if (E == A(kNoSat)) { q = A(1); return; }
    if (!(p > A(0))) { q = A(0); return; }
    if (E >= p) { q = A(1); return; }
    const A x = E * inv_p;
    const A num = x * A(kMH2O);
    const A den = num + (A(1) - x) * Mo;
    if (!(den > A(0))) { q = A(1); return; }
    const A rd = rcp_nr(den);
    q = num * rd;
    Af = num * Mo * rd * rd;
}

// ------------------------------------------------------------------------------------
// Two-category ice scheme (CategoryIceScheme = 2), TwoCatIceScheme::run, TwoCatIceScheme.h:89-111.
// k_ice_columns: one thread per (j,k) column. initArrays (:139-155) over the whole column, then
// for interior columns computeColumns (:160-517): up to three top-down passes of the rain and
// snow fluxes, each level's rates from the flux of the previous pass at that level (Rain, Snow)
// and the flux just formed one level up; the passes stop once the flux at level 1 stays within
// 1e-3 mm/d of its value at entry. Rate variables that the reference leaves unassigned on some
// branches (N_i and m_i outside the mixed-phase band, S_nuc where ice is zero and no nucleation
// condition holds, S_i_dep where c equals q_Ice) carry over from level to level and pass to pass
// in registers, as the reference's column-private variables do. Cells where no condensed phase
// can exist (IceSchemeCommon.h:104-190) evaporate their condensate and falling precipitation.
// Arithmetic in FP64 in both storage precisions, with contraction off so each operation rounds
// as the reference's.
// k_ice_bc: extrapolateBC (IceSchemeCommon.h:28-55) of both fluxes with c43 = 1, c13 = 0
// (ints, cAtmosphereModel.h:158) and the phi-seam average, written as the closed form of the
// three sequential face passes: a boundary cell takes the flux of the nearest cell with
// i <= im-2 and 1 <= j <= jm-2 in its meridian plane, the seam cells the mean of k = 1 and
// k = km-2. The flux values are finite and non-negative (clamped on every write), so c13 * x
// contributes an exact zero.
// ------------------------------------------------------------------------------------

namespace ice {
constexpr double N_i_0 = 1.0e2, m_i_0 = 1.0e-12, m_i_max = 1.0e-9, m_s_0 = 3.0e-9;
constexpr double c_i_dep = 1.3e-5, c_c_au = 1.0e-3, q_c_crit = 5.0e-4, c_i_au = 1.0e-3;
constexpr double c_ac = 0.24, c_rim = 18.6, c_rim_snow = 18.6 / 5.0, c_agg = 10.3;
constexpr double c_i_cri = 0.24, c_r_cri = 3.2e-5, a_ev = 1.0e-3, b_ev = 5.9;
constexpr double c_s_dep = 1.8e-2, b_s_dep = 12.3, c_s_melt = 8.43e-5, b_s_melt = 12.05;
constexpr double c_r_frz = 3.75e-2;
constexpr double t_nuc = 267.15, t_d = 248.15, t_hn = 236.15, t_r_frz = 271.15;
constexpr double P_max = 50.0 / 86400.0;
constexpr int i_check = 1, passes = 3;

// The state of one condensable level that the rates read (TwoCatIceScheme.h:229-266).
struct Level {
    double t_u, cv, cld, ic, rh, scc, q_sw, q_Ice, mass_layer, dt_snow_dim, Rain, Snow;
    double R49, R79, R139;
};

// The three rate variables the reference assigns only on some branches. Each returns whether it
// assigned; otherwise the value carried from the level above (or the previous pass) stands.
ATHAD_DEV bool assign_Ni_mi(const Level& L, double t_0, double& N_i, double& m_i) {   // :273-281
#pragma clang fp contract(off)
    if (!(!(L.t_u > t_0) && !(L.t_u <= t_hn))) return false;
    N_i = N_i_0 * exp(0.2 * (t_0 - L.t_u));
    m_i = tmin((L.rh * L.ic) / N_i, m_i_max);
    if (m_i > m_i_max) m_i = m_i_max;
    if (m_i < m_i_0) m_i = m_i_0;
    return true;
}
ATHAD_DEV bool assign_S_nuc(const Level& L, double N_i, double& S_nuc) {   // :285-292
#pragma clang fp contract(off)
    if (L.ic == 0.0) {
        if (((L.t_u < t_d) && (L.cv >= L.q_Ice)) || (((t_d <= L.t_u) && (L.t_u <= t_nuc)) && (L.cv >= L.q_sw))) {
            S_nuc = m_i_0 / (L.rh * L.dt_snow_dim) * N_i;
            return true;
        }
        return false;
    }
    S_nuc = 0.0;
    return true;
}
ATHAD_DEV bool assign_S_i_dep(const Level& L, double N_i, double m_i, double& S_i_dep) {   // :302-308
#pragma clang fp contract(off)
    bool a = false;
    if (L.cv > L.q_Ice) { S_i_dep = c_i_dep * N_i * pow(m_i, 1.0 / 3.0) * (L.cv - L.q_Ice); a = true; }
    if (L.cv < L.q_Ice) { S_i_dep = tmax((-L.ic / L.dt_snow_dim), ((L.cv - L.q_Ice) / L.dt_snow_dim)); a = true; }
    return a;
}

struct Rates { double Sv, Sc, Si, Sr, Ss, dPr, dPs; };

// Everything else at one level (:312-455) once the carried variables are resolved.
ATHAD_DEV Rates rates(const Level& L, double t_0, double m_i, double S_nuc, double S_i_dep) {
#pragma clang fp contract(off)
    const double t_u = L.t_u, cv = L.cv, cld = L.cld, ic = L.ic, Rain = L.Rain, Snow = L.Snow;
    const double dt_snow_dim = L.dt_snow_dim, mass_layer = L.mass_layer;
    double S_c_frz = (t_u < t_hn) ? cld / dt_snow_dim : 0.0;
    double S_c_au = (cld > q_c_crit) ? c_c_au * (cld - q_c_crit) : 0.0;
    double S_i_au = (ic > 0.0) ? tmax(c_i_au * ic, 0.0) : 0.0;
    double S_d_au = (S_i_dep > 0.0) ? S_i_dep / (1.5 * (pow((m_s_0 / m_i), 2.0 / 3.0) - 1.0)) : 0.0;
    double S_ac = ((t_u >= t_0) && (cld > q_c_crit)) ? c_ac * (cld - q_c_crit) * L.R79 : 0.0;
    double S_rim = (t_u < t_0) ? c_rim_snow * cld * Snow : 0.0;
    double S_shed = (t_u >= t_0) ? c_rim * cld * Snow : 0.0;
    double S_s_melt = 0.0, S_i_melt = 0.0;
    if (t_u > t_0) {
        if (Snow > 1e-12) {
            S_s_melt = c_s_melt * (1.0 + b_s_melt * pow(Snow, 5.0 / 26.0)) * (t_u - t_0) * pow(Snow, 1.0 / 3.0);
            S_s_melt = tmin(S_s_melt, Snow / mass_layer);
        }
        S_i_melt = ic / dt_snow_dim;
    }
    double S_agg = 0.0, S_i_cri = 0.0, S_r_cri = 0.0;
    if (t_u <= t_0) {
        S_agg = c_agg * ic * Snow;
        S_i_cri = c_i_cri * ic * L.R79;
        S_r_cri = c_r_cri * ic / m_i * L.R139;
    }
    double S_ev = 0.0;
    if (t_u > t_0 && Rain > 1e-12 && cv < L.q_sw) {
        S_ev = a_ev * (1.0 + b_ev * pow(Rain, 1.0 / 6.0)) * (L.q_sw - cv) * L.R49;
        S_ev = tmin(S_ev, Rain / mass_layer);
    }
    double S_s_dep = 0.0;
    if (t_u < t_0 && Snow > 1e-12) {
        S_s_dep = c_s_dep * (1.0 + b_s_dep * pow(Snow, 5.0 / 26.0)) * (cv - L.q_Ice) * pow(Snow, 8.0 / 13.0);
        if (S_s_dep < 0.0) S_s_dep = tmax(S_s_dep, -Snow / mass_layer);
    }
    double S_r_frz = 0.0;
    if (t_u <= t_0 && Rain != 0.0) {   // with no rain the product is an exact zero
        const double t_frz = tmax((t_r_frz - t_u), 0.0);
        S_r_frz = c_r_frz * pow(t_frz, 3.0 / 2.0) * pow(Rain, 3.0 / 2.0);
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
    Rates r;
    r.Sv = -L.scc + S_ev - S_i_dep - S_s_dep - S_nuc;
    r.Sc = L.scc - S_c_au - S_ac - S_c_frz + S_i_melt - S_rim - S_shed;
    r.Si = S_nuc + S_c_frz + S_i_dep - S_i_melt - S_i_au - S_d_au - S_agg - S_i_cri;
    r.Sr = S_c_au + S_ac - S_ev + S_shed - S_r_cri - S_r_frz + S_s_melt;
    r.Ss = S_i_au + S_d_au + S_agg + S_rim + S_s_dep + S_i_cri + S_r_cri + S_r_frz - S_s_melt;
    r.dPr = (S_c_au + S_ac + S_shed + S_s_melt - S_ev - S_r_frz) * mass_layer;
    r.dPs = (S_i_au + S_d_au + S_rim + S_agg + S_r_frz - S_s_melt) * mass_layer;
    return r;
}

// The level's inputs; false where no condensed phase can exist (IceSchemeCommon.h:104-107).
// cached: the level's q_sw and q_Ice from an earlier pass in which it could condense (its t and
// c have not changed since), passed in q_cached[2]; else null.
ATHAD_DEV bool load_level(const Consts& K, const double* hgt, int i, double tt, double cv, double cld, double ic,
                          double co2, double p_u, double rh, double scc, double Rain, double Snow, Level& L,
                          const double* q_cached = nullptr) {
#pragma clang fp contract(off)
    const double t_u = tt * K.t_0;
    L.t_u = t_u; L.cv = cv;
    double Mo = 0.0;
    if (q_cached) {
        L.q_sw = q_cached[0];
    } else {
        Mo = th_M_nonwater(cv, co2, K.M_bg);
        L.q_sw = (t_u >= kTcrit) ? 1.0 : th_q_sat(sat_p_sat(t_u), p_u, Mo);
        if (!(L.q_sw < 1.0)) return false;
    }
    L.cld = cld; L.ic = ic; L.rh = rh; L.scc = scc; L.Rain = Rain; L.Snow = Snow;
    L.R49 = 0.0; L.R79 = 0.0; L.R139 = 0.0;
    if (Rain > 1e-12) {
        const double rb = pow(Rain, 1.0 / 9.0);
        const double r2 = rb * rb;
        const double r4 = r2 * r2;
        L.R49 = r4; L.R79 = Rain / r2; L.R139 = Rain * r4;
    }
    L.q_Ice = q_cached ? q_cached[1] : th_q_sat(sat_p_subl(t_u), p_u, Mo);
    const double step = (double)(float)((float)hgt[i + 1] - (float)hgt[i]);
    L.mass_layer = rh * step;
    L.dt_snow_dim = step / 0.96;
    return true;
}

// evaporateWhereImpossible (IceSchemeCommon.h:121-190) on the values of one cell: new t, c and
// whether t changed; condensate, fluxes and sources of the cell become zero.
ATHAD_DEV void evaporate(const Consts& K, double t_u, double cv, double cld, double ic, double gr, double co2,
                         double rh, double Pr, double Ps, double& t_new, bool& t_set, double& c_new) {
#pragma clang fp contract(off)
    const double q_cond = tmax(0.0, cld) + tmax(0.0, ic) + tmax(0.0, gr);
    c_new = cv;
    t_set = false;
    if (q_cond > 0.0) {
        const double c_max = tmax(0.0, 1.0 - co2);
        const double cn = tmin(c_max, cv + q_cond);
        const double q_used = tmax(0.0, cn - cv);
        c_new = cn;
        if (t_u < kTcrit) {
            const double cp_loc = th_cp_of(cn, co2, t_u, K.M_bg);
            if (cp_loc > 0.0) {
                double T_new = t_u - th_L_vap(t_u) * q_used / cp_loc;
                if (T_new < 1.0) T_new = 1.0;
                t_new = T_new / K.t_0;
                t_set = true;
            }
        }
    }
    const double rho = tmax(1.0e-6, rh);
    const double q_p = tmax(0.0, Pr) / (1.6 * rho) + tmax(0.0, Ps) / (0.96 * rho) + 0.0 / (2.0 * rho);
    if (q_p > 0.0) c_new = tmin(tmax(0.0, 1.0 - co2), c_new + q_p);
}
}  // namespace ice

// One column of the scheme, sequentially as the reference runs it: initArrays, then for an
// interior column up to three top-down passes.
template <typename S>
ATHAD_DEV void ice_column_seq(const MemberDesc& M, int j, int k) {
#pragma clang fp contract(off)
    using namespace ice;
    const Consts& K = M.K;
    const int im = K.im;
    const size_t plane = (size_t)K.jm * K.km;
    const size_t base = (size_t)j * K.km + k;   // cell (0, j, k); level i adds i * plane
    // global-memory pointers (descriptor pointers would compile to flat accesses)
    typedef G1<S>* __restrict__ P;
    P ft = (P)M.f[F_T]; P fc = (P)M.f[F_C]; P fcl = (P)M.f[F_CLOUD]; P fi = (P)M.f[F_ICE];
    P fg = (P)M.f[F_GR]; P fscc = (P)M.f[F_S_C_C];
    const G1<S>* __restrict__ fco2 = (const G1<S>*)M.f[F_CO2];
    const G1<S>* __restrict__ frh = (const G1<S>*)M.f[F_R_HUMID];
    const G1<double>* __restrict__ fps = (const G1<double>*)M.f[F_P_STAT];
    P Pr = (P)M.mp[MP_P_RAIN]; P Ps = (P)M.mp[MP_P_SNOW]; P Prn = (P)M.mp[MP_P_RAINN]; P Psn = (P)M.mp[MP_P_SNOWN];
    P Pt = (P)M.mp[MP_PRECIP];
    P Sv = (P)M.mp[MP_S_V]; P Sc = (P)M.mp[MP_S_C]; P Si = (P)M.mp[MP_S_I]; P Sr = (P)M.mp[MP_S_R];
    P Ss = (P)M.mp[MP_S_S];
    auto at = [&](int i) { return base + (size_t)i * plane; };

    // Each level loop reads the next level before it stores this one: a wave's vector-memory
    // counter retires in issue order, so a load issued after a store would wait for the store.
    {   // initArrays
        S c_n = fc[at(0)], cl_n = fcl[at(0)], i_n = fi[at(0)], rn_n = Prn[at(0)], sn_n = Psn[at(0)];
        for (int i = 0; i < im; i++) {
            const size_t id = at(i);
            const S c0 = c_n, cl0 = cl_n, i0 = i_n, rn0 = rn_n, sn0 = sn_n;
            if (i + 1 < im) {
                const size_t nx = at(i + 1);
                c_n = fc[nx]; cl_n = fcl[nx]; i_n = fi[nx]; rn_n = Prn[nx]; sn_n = Psn[nx];
            }
            fc[id] = tmax(S(0), c0); fcl[id] = tmax(S(0), cl0); fi[id] = tmax(S(0), i0);
            const S rn = tmax(S(0), rn0), sn = tmax(S(0), sn0);
            Prn[id] = rn; Psn[id] = sn; Pr[id] = rn; Ps[id] = sn;
        }
    }
    if (j < 1 || j > K.jm - 2 || k < 1 || k > K.km - 2) return;

    const double Rain_check = (double)Pr[at(i_check)];
    double m_i = m_i_max, N_i = 0.0, S_nuc = 0.0, S_i_dep = 0.0;
    // q_sw and q_Ice of the levels that could condense in the previous pass (their t and c are
    // unchanged), kept in scratch 10 and 11 for the later passes: FP64 storage, up to 64 levels
    constexpr bool kCache = sizeof(S) == 8;
    G1<double>* __restrict__ qc0 = (G1<double>*)M.tmp[10];
    G1<double>* __restrict__ qc1 = (G1<double>*)M.tmp[11];
    unsigned long long cached = 0ULL;
    // a level's inputs to one pass
    struct IceL { double t, c, cl, ic, gr, co2, ps, rh, scc, pr, psn, q0, q1; bool hit; };
    auto load_ice = [&](int i) {
        const size_t id = at(i);
        IceL v;
        v.t = (double)ft[id]; v.c = (double)fc[id]; v.cl = (double)fcl[id]; v.ic = (double)fi[id];
        v.gr = (double)fg[id]; v.co2 = (double)fco2[id]; v.ps = fps[id]; v.rh = (double)frh[id];
        v.scc = (double)fscc[id]; v.pr = (double)Pr[id]; v.psn = (double)Ps[id];
        v.hit = kCache && i < 64 && ((cached >> i) & 1ULL);
        v.q0 = 0.0; v.q1 = 0.0;
        if (v.hit) { v.q0 = qc0[id]; v.q1 = qc1[id]; }
        return v;
    };
    for (int pass = 1; pass <= passes; pass++) {
        IceL nx = load_ice(im - 2);
        Pr[at(im - 1)] = S(0);
        Ps[at(im - 1)] = S(0);
        double pr_up = 0.0, ps_up = 0.0;   // fluxes of the level above, as just stored
        unsigned long long can_mask = 0ULL;
        for (int i = im - 2; i >= 0; i--) {
            const size_t id = at(i);
            const IceL V = nx;
            if (i - 1 >= 0) nx = load_ice(i - 1);
            Level L;
            const double cv = V.c, co2 = V.co2;
            double qv[2] = {V.q0, V.q1};
            if (!load_level(K, M.hgt, i, V.t, cv, V.cl, V.ic, co2, V.ps, V.rh, V.scc, V.pr, V.psn, L,
                            V.hit ? qv : nullptr)) {
                double t_new = 0.0, c_new;
                bool t_set;
                evaporate(K, L.t_u, cv, V.cl, V.ic, V.gr, co2, V.rh, V.pr, V.psn, t_new, t_set, c_new);
                if (tmax(0.0, V.cl) + tmax(0.0, V.ic) + tmax(0.0, V.gr) > 0.0) {
                    fcl[id] = S(0); fi[id] = S(0); fg[id] = S(0);
                }
                if (t_set) ft[id] = S(t_new);
                if (c_new != cv) fc[id] = S(c_new);
                Pr[id] = S(0); Ps[id] = S(0); Pt[id] = S(0);
                Sv[id] = S(0); Sc[id] = S(0); Si[id] = S(0); Sr[id] = S(0); Ss[id] = S(0);
                fscc[id] = S(0);
                pr_up = 0.0; ps_up = 0.0;
                continue;
            }
            if (kCache && i < 64 && im <= 64) {
                if (!V.hit && pass < passes) { qc0[id] = L.q_sw; qc1[id] = L.q_Ice; }
                can_mask |= 1ULL << i;
            }
            assign_Ni_mi(L, K.t_0, N_i, m_i);
            assign_S_nuc(L, N_i, S_nuc);
            assign_S_i_dep(L, N_i, m_i, S_i_dep);
            const Rates r = rates(L, K.t_0, m_i, S_nuc, S_i_dep);
            Sv[id] = S(r.Sv); Sc[id] = S(r.Sc); Si[id] = S(r.Si); Sr[id] = S(r.Sr); Ss[id] = S(r.Ss);
            const double pr = tmin(P_max, tmax(0.0, pr_up + r.dPr));
            const double psn = (L.t_u < K.t_0 && L.t_u >= K.t_000) ? tmin(P_max, tmax(0.0, ps_up + r.dPs)) : 0.0;
            Pr[id] = S(pr); Ps[id] = S(psn);
            Pt[id] = S(tmin(P_max, (double)S(pr) + (double)S(psn)));
            pr_up = (double)S(pr); ps_up = (double)S(psn);
        }
        if (fabs((double)Pr[at(i_check)] - Rain_check) * 8.64e4 <= 1.0e-3) break;
        cached = can_mask;
    }
}

// One thread per (j, k) column.
template <typename S>
__global__ __launch_bounds__(256) void k_ice_columns(const MemberDesc* __restrict__ md, int gate) {
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= K.jm * K.km) return;
    const int j = col / K.km;
    ice_column_seq<S>(M, j, col - j * K.km);
}

template <typename S>
__global__ __launch_bounds__(256) void k_ice_bc(const MemberDesc* __restrict__ md, int gate) {
#pragma clang fp contract(off)
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= K.n) return;
    int i, j, k;
    decode(K, id, i, j, k);
    const bool seam = (k == 0 || k == K.km - 1);
    if (!seam && i < K.im - 1 && j > 0 && j < K.jm - 1) return;
    const int ii = min(i, K.im - 2), jj = min(max(j, 1), K.jm - 2);
    for (int q = 0; q < 2; q++) {
        S* P = mpp<S>(M, q == 0 ? MP_P_RAIN : MP_P_SNOW);
        if (!seam) P[id] = P[ix3(K, ii, jj, k)];
        else P[id] = S((double(P[ix3(K, ii, jj, 1)]) + double(P[ix3(K, ii, jj, K.km - 2)])) * 0.5);
    }
}

// P_rainn <- P_rain, P_snown <- P_snow (storeIntermediateData3D, UtilsAtm.h:41-42).
template <typename S>
__global__ __launch_bounds__(256) void k_ice_store(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= M.K.n || !M.mp[0]) return;
    mpp<S>(M, MP_P_RAINN)[id] = mpp<S>(M, MP_P_RAIN)[id];
    mpp<S>(M, MP_P_SNOWN)[id] = mpp<S>(M, MP_P_SNOW)[id];
}

// Column and cell thermodynamics (k_densities, the SatAdj entry and clamp): th_R_of, th_cp_of, th_M_nonwater, th_q_sat and
// th_moist_lapse_E restated with shared sub-expressions (one split, one 1/t^2 for the three
// Shomate fits, reciprocals of repeated denominators) and refined reciprocals, which shortens
// the serial per-level chain of the column recursion; results agree with th_* to a few ulp.
// Templated on the arithmetic type A: double everywhere except the mixed form's column.
template <typename A> struct MixQ { A qv, qc, qb; };
template <typename A> ATHAD_DEV MixQ<A> dens_split(A c, A co2) {
    A qv = tmin(tmax(c, A(0)), A(1)), qc = tmin(tmax(co2, A(0)), A(1));
    const A s = qv + qc;
    if (s > A(1)) { qv /= s; qc /= s; }
    return {qv, qc, tmax(A(0), A(1) - qv - qc)};
}
template <typename A> ATHAD_DEV A dens_R(const MixQ<A>& m, A Rb) { return m.qv * A(kRH2O) + m.qc * A(kRCO2) + m.qb * Rb; }
template <typename A> ATHAD_DEV A dens_cp(const MixQ<A>& m, A T, A inv_Mb) {
    const A t = tmin(tmax(T, A(298)), A(2000)) * A(1.0e-3);
    const A t2 = t * t, t3 = t2 * t, it2 = rcp_nr(t2);
    const A cpH = (A(30.09200) + A(6.832514) * t + A(6.793435) * t2 + A(-2.534480) * t3 + A(0.082139) * it2) * A(1.0 / kMH2O);
    const A cpC = (A(24.99735) + A(55.18696) * t + A(-33.69137) * t2 + A(7.948387) * t3 + A(-0.136638) * it2) * A(1.0 / kMCO2);
    const A cpB = (A(28.98641) + A(1.853978) * t + A(-9.647459) * t2 + A(16.63537) * t3 + A(0.000117) * it2) * inv_Mb;
    return m.qv * cpH + m.qc * cpC + m.qb * cpB;
}
template <typename A> ATHAD_DEV A dens_Mnw(const MixQ<A>& m, A Mb, A inv_Mb) {
    const A s = m.qc + m.qb;
    if (!(s > A(0))) return Mb;
    const A is = rcp_nr(s);
    const A inv = (m.qc * is) * A(1.0 / kMCO2) + (m.qb * is) * inv_Mb;
    return (inv > A(0)) ? rcp_nr(inv) : Mb;
}
template <typename A> ATHAD_DEV A dens_p_auto(A T) {
    if (T >= A(kTcrit)) return A(kNoSat);
    return (T >= A(kTtriple)) ? sat_p_sat(T) : sat_p_subl(T);
}
// moist lapse rate given q_sat's factor Af at (T, p)
template <typename A> ATHAD_DEV A dens_lapse(A T, A Af, A cp, A Rm, A g) {
    if (!(cp > A(0)) || !(T > A(0)) || !(Rm > A(0))) return A(0);
    if (!(Af > A(0))) return g * rcp_nr(cp);
    const A L = (T >= A(kTtriple)) ? sat_L_vap(T) : sat_L_vap(T) + A(3.337e5);
    const A num = A(1) + L * Af * rcp_nr(Rm * T);
    const A den = cp + L * L * Af * rcp_nr(A(kRH2O) * T * T);
    return (den > A(0)) ? g * num * rcp_nr(den) : g * rcp_nr(cp);
}

// Damped Newton iteration of adjustSaturation (SaturationAdjustment.h:177-254) in type N.
// The ice branch (sublimation pressure, its q_sat and derivative) is evaluated only when its
// result is read: deposition fraction > 0 (the gain term), or the target uses q_ice (ice
// present with condensate above 1e-12, or the no-condensate target when T < t_0). Otherwise
// every term it feeds is an exact zero: a zero deposition fraction times a finite factor, or
// zero ice times a finite q_ice. Likewise the liquid branch. (Cells that keep a trace of ice
// below the condensate threshold, above t_0, run all 20 iterations and need only the liquid
// branch.)
template <typename N>
ATHAD_DEV void sat_newton(double& T_io, double& qv_io, double& qc_io, double& qi_io, double co2d, double p_local,
                          double M_other, double Mb_d, double t_00d, double t_0d) {
    const N t_00 = N(t_00d), t_0 = N(t_0d), t_range_inv = N(1.0 / (t_0d - t_00d)), co2 = N(co2d), Mb = N(Mb_d);
    const double inv_p_d = (p_local > 0.0) ? 1.0 / p_local : 0.0;
    N T = N(T_io), qvb = N(qv_io), qcb = N(qc_io), qib = N(qi_io), qv_hyp = qvb;
    // species heat capacities and L(T) of iteration k+1's update are those of iteration k's
    // new temperature
    N cpH, cpC, cpB;
    sat_species_cp(T, Mb, cpH, cpC, cpB);
    N Lv = sat_L_vap(T);
    N T_prev = T, qv_prev = qvb, qc_prev = qcb, qi_prev = qib, CND_prev = N(-1);
    for (int it = 1; it <= 20; it++) {
        const N CND = tmax(N(0), tmin(N(1), (T - t_00) * t_range_inv));
        const N DEP = N(1) - CND;
        N dqv = qv_hyp - qvb;
        if (dqv > N(0)) { const N me = qcb + qib; if (dqv > me) dqv = me; }
        // A zero step with the same phase split as the previous iteration leaves T, qvb, qcb,
        // qib unchanged and recomputes the same qv_hyp from the same inputs: the state is
        // already the fixed point the remaining iterations would keep.
        if (dqv == N(0) && CND == CND_prev) break;
        const N CND_last = CND_prev;
        CND_prev = CND;
        const N dcnd = dqv * CND, ddep = dqv * DEP;
        qvb += dqv;
        qcb = tmax(N(0), qcb - dcnd);
        qib = tmax(N(0), qib - ddep);
        T -= (Lv * dcnd + (Lv + N(3.337e5)) * ddep) * rcp_nr(sat_mix_cp(qvb, co2, cpH, cpC, cpB));
        // An update that leaves (T, qvb, qcb, qib) as the previous iteration left them, under
        // the previous iteration's phase split, feeds the rest of this iteration the same inputs
        // as that iteration: qv_hyp comes out equal to its previous value, the convergence test
        // fails again and the fixed-point test below stops the loop with this state. Stop here.
        if (T == T_prev && qvb == qv_prev && qcb == qc_prev && qib == qi_prev && CND == CND_last) break;
        sat_species_cp(T, Mb, cpH, cpC, cpB);
        Lv = sat_L_vap(T);
        const N qsum = qcb + qib;
        const bool hot = (T >= t_0), some = (qsum > N(1e-12));
        const bool need_l = (CND > N(0)) || (some ? (qcb > N(0)) : hot);
        const bool need_i = (DEP > N(0)) || (some ? (qib > N(0)) : !hot);
        // the saturation vapour pressures and the exact q_sat with its derivative factor in FP64
        // in both instantiations; N = float carries the iteration's state and its latent heat
        // and heat capacity in FP32 (the mixed form)
        N qs = N(0), Ac = N(0), qi = N(0), Ad = N(0);
        if (need_l) {
            double q_, A_;
            sat_qA(sat_p_sat((double)T), p_local, inv_p_d, M_other, q_, A_);
            qs = N(q_); Ac = N(A_);
        }
        if (need_i) {
            double q_, A_;
            sat_qA(sat_p_subl((double)T), p_local, inv_p_d, M_other, q_, A_);
            qi = N(q_); Ad = N(A_);
        }
        const N target = some ? (qcb * qs + qib * qi) * rcp_nr(qsum) : (hot ? qs : qi);
        const N inv_cp = rcp_nr(sat_mix_cp(qvb, co2, cpH, cpC, cpB));
        const N inv_RT2 = (T > N(0)) ? rcp_nr(N(kRH2O) * T * T) : N(0);
        const N Lc = Lv, Ld = Lv + N(3.337e5);
        const N gain = CND * (Lc * inv_cp) * (Ac * Lc * inv_RT2) + DEP * (Ld * inv_cp) * (Ad * Ld * inv_RT2);
        const N om = rcp_nr(N(1) + gain);
        const N hyp_prev = qv_hyp;
        qv_hyp = qvb + om * (target - qvb);
        if (fabs(qvb / qv_hyp - N(1)) <= N(1.0e-6)) break;
        // The iteration is a deterministic map of (T, qvb, qcb, qib, qv_hyp): once an
        // iteration leaves that state unchanged, every later one does too and the break test
        // above keeps failing, so the remaining iterations are skipped with identical results.
        if (T == T_prev && qvb == qv_prev && qcb == qc_prev && qib == qi_prev && qv_hyp == hyp_prev) break;
        T_prev = T; qv_prev = qvb; qc_prev = qcb; qi_prev = qib;
    }
    T_io = (double)T; qv_io = (double)qvb; qc_io = (double)qcb; qi_io = (double)qib;
}

// The saturation adjustment's fade factor 1 / (1 + exp(-(T - t_00) / 5)) and, where condensate
// is released or formed, the mixture's heat capacity and the latent heat: FP64 in the FP64
// form, FP32 arithmetic in the mixed form (storage S = float), where only the saturation
// vapour pressures and q_sat stay FP64.
template <typename S> ATHAD_DEV double fade(double T, double t_00) {
    if constexpr (sizeof(S) == 4) return (double)rcp_nr(1.0f + expf((float)(-(T - t_00) * 0.2)));
    else return rcp_nr(1.0 + exp(-(T - t_00) * 0.2));
}
template <typename S> ATHAD_DEV double heat_cap(double c, double co2, double T, double inv_Mb) {
    if constexpr (sizeof(S) == 4) return (double)dens_cp(dens_split((float)c, (float)co2), (float)T, (float)inv_Mb);
    else return dens_cp(dens_split(c, co2), T, inv_Mb);
}
template <typename S> ATHAD_DEV double latent(double T) {
    if constexpr (sizeof(S) == 4) return (double)sat_L_vap((float)T);
    else return sat_L_vap(T);
}

template <typename S>
__global__ __launch_bounds__(256) void k_satadj(const MemberDesc* __restrict__ md, int gate, int to_tmp) {
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long id = cell_id_top_down(K);
    if (id < 0) return;
    int i, j, k;
    decode(K, id, i, j, k);
    S *pt = fp<S>(M, F_T), *pc = fp<S>(M, F_C), *pcl = fp<S>(M, F_CLOUD), *pi = fp<S>(M, F_ICE), *pg = fp<S>(M, F_GR);
    const double co2 = (double)fp<S>(M, F_CO2)[id];
    const double p_local = pstat(M)[id];
    double t = pt[id], c = pc[id], cloud = pcl[id], ice = pi[id], gr = pg[id];
    double scc = 0.0;
    // the entry test's saturation pressure and fade factor, and the temperature they were taken
    // at: the clamp below reuses them when its temperature is bitwise the same
    double E_entry = 0.0, alpha_entry = 0.0, T_entry = __builtin_nan("");   // NaN: equal to nothing
    const double t_0 = K.t_0, t_00 = K.t_00, Mb = K.M_bg, inv_Mb = 1.0 / K.M_bg;
    const double inv_t_0 = 1.0 / t_0;
    const double inv_p = (p_local > 0.0) ? rcp_nr(p_local) : 0.0;
    if (i >= 1 && i <= K.im - 2) {
        // adjustSaturation :91-284
        const double qv_old = tmax(0.0, c), qc_old = tmax(0.0, cloud), qi_old = tmax(0.0, ice);
        double T = t * t_0;
        if (T < kTcrit) {
            const double M_other = dens_Mnw(dens_split(c, co2), Mb, inv_Mb);
            double qs0, A0;
            E_entry = dens_p_auto(T);
            sat_qA(E_entry, p_local, inv_p, M_other, qs0, A0);
            const double alpha = fade<S>(T, t_00);
            T_entry = T; alpha_entry = alpha;
            if ((qv_old > qs0 && alpha > 0.01) || (qv_old < qs0 && (qc_old > 1e-12 || qi_old > 1e-12))) {
                double qvb = qv_old, qcb = qc_old, qib = qi_old;
                const double T_orig = t * t_0;
                if constexpr (sizeof(S) == 4) sat_newton<float>(T, qvb, qcb, qib, co2, p_local, M_other, Mb, t_00, t_0);
                else sat_newton<double>(T, qvb, qcb, qib, co2, p_local, M_other, Mb, t_00, t_0);
                qcb = tmax(0.0, qcb);
                qib = tmax(0.0, qib);
                if (T > kTcrit) T = kTcrit;
                if (!isnan(T) && !isnan(qvb)) {
                    scc = alpha * (qcb - qc_old) / M.step_dim[i];
                    c = qv_old + alpha * (qvb - qv_old);
                    cloud = qc_old + alpha * (qcb - qc_old);
                    ice = qi_old + alpha * (qib - qi_old);
                    t = (T_orig + alpha * (T - T_orig)) * inv_t_0;
                }
                if (T < t_00) { cloud = 0.0; ice = 0.0; }
            }
        }
    }
    // clampAndFade :303-451
    {
        if (c < 0.0) c = 0.0;
        if (cloud < 0.0) cloud = 0.0;
        if (ice < 0.0) ice = 0.0;
        const double c_max = tmax(0.0, 1.0 - co2);
        if (c > c_max) c = c_max;
        double Td = t * t_0;
        if (Td > K.t_max_phys) { Td = K.t_max_phys; t = K.t_max_phys * inv_t_0; }
        const double M_o = dens_Mnw(dens_split(c, co2), Mb, inv_Mb);
        double q_s, A_s;
        sat_qA((Td == T_entry) ? E_entry : dens_p_auto(Td), p_local, inv_p, M_o, q_s, A_s);
        if (Td >= kTcrit || q_s >= 1.0) {
            const double q_cond = tmax(0.0, cloud) + tmax(0.0, ice) + tmax(0.0, gr);
            if (q_cond > 0.0) {
                const double cm = tmax(0.0, 1.0 - co2);
                const double c_new = tmin(cm, c + q_cond);
                const double q_used = tmax(0.0, c_new - c);
                c = c_new; cloud = 0.0; ice = 0.0; gr = 0.0;
                if (Td < kTcrit) {
                    const double cp_loc = heat_cap<S>(c, co2, Td, inv_Mb);
                    if (cp_loc > 0.0) {
                        double T_new = Td - latent<S>(Td) * q_used * rcp_nr(cp_loc);
                        if (T_new < 1.0) T_new = 1.0;
                        t = T_new / t_0;
                    }
                }
            }
            if (M.mp[0]) {
                // falling rain and snow evaporate too, and the sources clear
                // (IceSchemeCommon.h:175-189); arrays held once the ice scheme has run
                S* Pr = (S*)M.mp[MP_P_RAIN];
                S* Ps = (S*)M.mp[MP_P_SNOW];
                const double rho = tmax(1.0e-6, (double)fp<S>(M, F_R_HUMID)[id]);
                const double q_p = tmax(0.0, (double)Pr[id]) / (1.6 * rho) + tmax(0.0, (double)Ps[id]) / (0.96 * rho)
                                 + 0.0 / (2.0 * rho);
                if (q_p > 0.0) c = tmin(tmax(0.0, 1.0 - co2), c + q_p);
                for (int q = MP_P_RAIN; q < NMP; q++)
                    if (q != MP_P_RAINN && q != MP_P_SNOWN) ((S*)M.mp[q])[id] = S(0);
            }
            scc = 0.0;
        } else {
            const double qs = q_s;
            if (c > qs) {
                const double excess = c - qs;
                c = qs;
                const double cpx = heat_cap<S>(c, co2, Td, inv_Mb);
                if (Td >= t_00) { cloud += excess; Td += latent<S>(Td) * excess * rcp_nr(cpx); }
                else { ice += excess; Td += (latent<S>(Td) + 3.337e5) * excess * rcp_nr(cpx); }
                if (Td > K.t_max_phys) Td = K.t_max_phys;
                t = Td * inv_t_0;
            }
            if (cloud > K.cloud_cap) cloud = K.cloud_cap;
            if (ice > K.cloud_cap) ice = K.cloud_cap;
            const double al = (Td == T_entry) ? alpha_entry : fade<S>(Td, t_00);
            c *= al; cloud *= al; ice *= al;
        }
    }
    // cap_S of S_c_c (cAtmosphereModel.cpp:1290-1305); with the ice scheme in the moist block
    // (bit 1 of to_tmp) the cap comes after it, since the scheme reads S_c_c uncapped
    if (!(to_tmp & 2)) {
        if (!finite_bits(scc)) scc = 0.0;
        else if (scc > 1.0e-3) scc = 1.0e-3;
        else if (scc < -1.0e-3) scc = -1.0e-3;
    }
    if (to_tmp & 1) {   // scratch 6..9 = t, ice, c, cloud for the fused damping that follows
        ((S*)M.tmp[6])[id] = S(t); ((S*)M.tmp[7])[id] = S(ice); ((S*)M.tmp[8])[id] = S(c); ((S*)M.tmp[9])[id] = S(cloud);
    } else {
        pt[id] = S(t); pc[id] = S(c); pcl[id] = S(cloud); pi[id] = S(ice);
    }
    pg[id] = S(gr);
    fp<S>(M, F_S_C_C)[id] = S(scc);
}

// ------------------------------------------------------------------------------------
// damp_wiggles passes (lib/Utils.cpp:429-532): k periodic, j clamped, i clamped, each from
// the previous pass's complete output. A job names up to six (src, dst) pairs, where an
// index < NFIELD is a field and NFIELD + q is scratch buffer q. The i pass can finish with
// the moist clamp of cAtmosphereModel.cpp:1307-1399.
// ------------------------------------------------------------------------------------
struct DampJob { int nf; int src[6]; int dst[6]; int moist_clamp; };

// Buffer q of a damping / copy job: fields, scratch, the ice-scheme arrays, then (FP64 jobs only)
// the convection arrays and their four damping temporaries (conv_qs[0..1], conv_tmp[0..1]).
template <typename S> ATHAD_DEV S* job_ptr(const MemberDesc& M, int q) {
    constexpr int M0 = NFIELD + NTMP + NMP, T0 = M0 + NMC;
    if (q < NFIELD) return (S*)M.f[q];
    if (q < NFIELD + NTMP) return (S*)M.tmp[q - NFIELD];
    if (q < M0) return (S*)M.mp[q - NFIELD - NTMP];
    if (q < T0) return (S*)M.mc[q - M0];
    return (S*)((q - T0 < 2) ? M.conv_qs[q - T0] : M.conv_tmp[q - T0 - 2]);
}

template <typename S, int AXIS>
__global__ __launch_bounds__(256) void k_damp_pass(const MemberDesc* __restrict__ md, DampJob job, int gate) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long id = cell_id(K);
    if (id < 0) return;
    int i, j, k;
    decode(K, id, i, j, k);
    size_t a, b;
    if (AXIS == 0) {   // k, periodic
        const int km1 = (k - 1 + K.km) % K.km, kp1 = (k + 1) % K.km;
        a = id - k + km1; b = id - k + kp1;
    } else if (AXIS == 1) {   // j, clamped
        a = ix3(K, i, max(j - 1, 0), k); b = ix3(K, i, min(j + 1, K.jm - 1), k);
    } else {   // i, clamped
        a = ix3(K, max(i - 1, 0), j, k); b = ix3(K, min(i + 1, K.im - 1), j, k);
    }
    C out[6];
    for (int q = 0; q < job.nf; q++) {
        const S* s = job_ptr<S>(M, job.src[q]);
        const C c0 = C(s[id]);
        // the symmetric form adds the two j neighbours first so mirrored cells round alike
        out[q] = (AXIS == 1 && K.sym) ? c0 + C(0.25) * ((C(s[a]) + C(s[b])) - C(2) * c0)
                                      : c0 + C(0.25) * (C(s[a]) - C(2) * c0 + C(s[b]));
    }
    if (AXIS == 2 && job.moist_clamp) {
        // job order t, ice, c, cloud; co2 from the field
        const C co2 = C(fp<S>(M, F_CO2)[id]);
        C cv = out[2], cld = out[3], ic = out[1];
        if (!finite_bits(cv)) cv = C(0);
        if (!finite_bits(cld)) cld = C(0);
        if (!finite_bits(ic)) ic = C(0);
        if (cv < C(0)) cv = C(0);
        if (cld < C(0)) cld = C(0);
        if (ic < C(0)) ic = C(0);
        const C cv_max = tmax(C(0), C(1) - co2);
        if (cv > cv_max) cv = cv_max;
        const C sum = cld + ic;
        const C cap = C(K.cloud_cap);
        if (sum > cap) { const C sc = cap / sum; cld *= sc; ic *= sc; }
        if (cv < C(0)) cv = C(0);
        if (cv > cv_max) cv = cv_max;
        out[2] = cv; out[3] = cld; out[1] = ic;
    }
    for (int q = 0; q < job.nf; q++) job_ptr<S>(M, job.dst[q])[id] = S(out[q]);
}

// The same three passes fused: a block owns a TJ x TK (latitude x longitude) tile and marches
// it upward through levels [i0, i0 + ich). For each level it stages the tile plus a one-cell
// halo of every field in LDS (j clamped, k periodic, exactly the neighbours the separate
// passes read), forms the k pass and then the j pass at its own cell, and keeps the j-pass
// results of levels i-1, i, i+1 in registers for the i pass. Each thread's staging cells are
// fixed per level, so their offsets are formed once; the next level's values are loaded into
// registers as soon as this level's are in LDS, so the loads overlap the level's passes and
// stores. Reads src, writes dst (distinct buffers: neighbouring tiles still read src). Results
// equal the separate passes bit for bit for every tile shape, so the shape is free for an
// autotuner to choose.
template <typename S, int TJ, int TK, int NF>
__global__ __launch_bounds__(TJ * TK) void k_damp_tile(const MemberDesc* __restrict__ md, DampJob job, int ich, int gate) {
    using C = S;
    constexpr int W = TK + 2, SL = (TJ + 2) * W, NT = TJ * TK, NE = (SL + NT - 1) / NT;
    __shared__ C slab[NF * SL];
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const int ntk = (K.km + TK - 1) / TK, ntj = (K.jm + TJ - 1) / TJ, nch = (K.im + ich - 1) / ich;
    const long long bb = xcd_block(blockIdx.x, gridDim.x, (long long)ntk * ntj * nch);
    if (bb < 0) return;
    int b = (int)bb;
    const int tk = b % ntk; b /= ntk;
    const int tj = b % ntj;
    const int ch = b / ntj;
    const int j0 = tj * TJ, k0 = tk * TK;
    const int i0 = ch * ich, i1 = min(K.im, i0 + ich);
    const int lj = threadIdx.x / TK, lk = threadIdx.x % TK;
    const int j = j0 + lj, k = k0 + lk;
    const bool own = (j < K.jm) && (k < K.km);
    const S* src[NF];
    S* dst[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) { src[q] = job_ptr<S>(M, job.src[q]); dst[q] = job_ptr<S>(M, job.dst[q]); }
    // this thread's staging cells: slab position threadIdx.x + u NT, offset inside a level
    int eoff[NE];
#pragma unroll
    for (int u = 0; u < NE; u++) {
        const int e = threadIdx.x + u * NT;
        const int r = e / W, qk = e - r * W;
        const int jj = min(max(j0 - 1 + r, 0), K.jm - 1);
        const int kk = (k0 - 1 + qk + K.km) % K.km;
        eoff[u] = (e < SL) ? jj * K.km + kk : 0;
    }
    const long long si = (long long)K.jm * K.km;
    C st[NE][NF];   // one level's staging values, loaded a level ahead
    auto fetch = [&](int il) {
        const long long lo = (long long)min(max(il, 0), K.im - 1) * si;
#pragma unroll
        for (int u = 0; u < NE; u++)
            if (u < NE - 1 || threadIdx.x + u * NT < SL) {
#pragma unroll
                for (int q = 0; q < NF; q++) st[u][q] = C(src[q][lo + eoff[u]]);
            }
    };
    // j-pass result at level il (clamped to the grid) for this thread's cell, every field; the
    // staging values of il are in st, and those of il + 1 are requested before the passes
    auto level = [&](int il, C* out) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NE; u++)
            if (u < NE - 1 || threadIdx.x + u * NT < SL) {
#pragma unroll
                for (int q = 0; q < NF; q++) slab[q * SL + threadIdx.x + u * NT] = st[u][q];
            }
        fetch(il + 1);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NF; q++) {
            const C* sb = slab + q * SL + lk;
            C kp[3];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const C* row = sb + (lj + r) * W;
                const C c0 = row[1];
                kp[r] = c0 + C(0.25) * (row[0] - C(2) * c0 + row[2]);
            }
            out[q] = K.sym ? kp[1] + C(0.25) * ((kp[0] + kp[2]) - C(2) * kp[1])
                           : kp[1] + C(0.25) * (kp[0] - C(2) * kp[1] + kp[2]);
        }
    };
    C jm_[NF], jc[NF], jp[NF];
    fetch(i0 - 1);
    level(i0 - 1, jm_);
    level(i0, jc);
    for (int i = i0; i < i1; i++) {
        level(i + 1, jp);
        C out[NF];
#pragma unroll
        for (int q = 0; q < NF; q++) out[q] = jc[q] + C(0.25) * (jm_[q] - C(2) * jc[q] + jp[q]);
        if ((NF == 3 || NF == 4) && job.moist_clamp) {
            // job order [t,] ice, c, cloud; the microphysics clamp and vapour ceiling
            // (cAtmosphereModel.cpp:1307-1399)
            constexpr int o = NF - 3;   // index of ice
            const long long id = ((long long)i * K.jm + min(j, K.jm - 1)) * K.km + min(k, K.km - 1);
            const C co2 = C(fp<S>(M, F_CO2)[id]);
            C cv = out[o + 1], cld = out[o + 2], ic = out[o];
            if (!finite_bits(cv)) cv = C(0);
            if (!finite_bits(cld)) cld = C(0);
            if (!finite_bits(ic)) ic = C(0);
            if (cv < C(0)) cv = C(0);
            if (cld < C(0)) cld = C(0);
            if (ic < C(0)) ic = C(0);
            const C cv_max = tmax(C(0), C(1) - co2);
            if (cv > cv_max) cv = cv_max;
            const C sum = cld + ic;
            const C cap = C(K.cloud_cap);
            if (sum > cap) { const C sc = cap / sum; cld *= sc; ic *= sc; }
            if (cv < C(0)) cv = C(0);
            if (cv > cv_max) cv = cv_max;
            out[o + 1] = cv; out[o + 2] = cld; out[o] = ic;
        }
        if (own) {
            const long long id = ((long long)i * K.jm + j) * K.km + k;
#pragma unroll
            for (int q = 0; q < NF; q++) dst[q][id] = S(out[q]);
        }
#pragma unroll
        for (int q = 0; q < NF; q++) { jm_[q] = jc[q]; jc[q] = jp[q]; }
    }
}

// damp_wiggles of one buffer at the lowest level only: the k and j passes on levels 0 and 1,
// then the i pass with the level below clamped to level 0, the same operations as
// k_damp_tile's for that level. The moist block damps the temperature this way: the
// densities step after it overwrites every level above the surface from the surface value.
template <typename S>
__global__ __launch_bounds__(256) void k_damp_level0(const MemberDesc* __restrict__ md, int src, int dst, int gate) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    const int j = (int)(q / K.km), k = (int)(q % K.km);
    const S* x = job_ptr<S>(M, src);
    const int km1 = (k - 1 + K.km) % K.km, kp1 = (k + 1) % K.km;
    auto jk = [&](int ii) {   // k pass then j pass at level ii, cell (j, k)
        C kp[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int jj = min(max(j - 1 + r, 0), K.jm - 1);
            const S* row = x + ((long long)ii * K.jm + jj) * K.km;
            const C c0 = C(row[k]);
            kp[r] = c0 + C(0.25) * (C(row[km1]) - C(2) * c0 + C(row[kp1]));
        }
        return K.sym ? kp[1] + C(0.25) * ((kp[0] + kp[2]) - C(2) * kp[1])
                     : kp[1] + C(0.25) * (kp[0] - C(2) * kp[1] + kp[2]);
    };
    const C jc = jk(0), jp = jk(min(1, K.im - 1));
    job_ptr<S>(M, dst)[(long long)j * K.km + k] = S(jc + C(0.25) * (jc - C(2) * jc + jp));
}

// Copies up to six buffers (src[q] -> dst[q]).
template <typename S>
__global__ void k_copy_job(const MemberDesc* __restrict__ md, DampJob job, int gate) {
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const long long id = cell_id(M.K);
    if (id < 0) return;
    for (int q = 0; q < job.nf; q++) job_ptr<S>(M, job.dst[q])[id] = job_ptr<S>(M, job.src[q])[id];
}

// ------------------------------------------------------------------------------------
// Column kernel: ThermoAtm::densities (ThermoAtm.h:1436-1614) followed by the sea-surface
// saturation boundary of waterVapourEvaporation (ThermoAtm.h:115-258). One thread per
// (j,k) column, serial upward in i. set_water: initial profile mode. The column arithmetic
// is in the storage type: FP64 in the drop-in form; FP32 in the mixed form, where the
// FP64 library exp/log on the serial per-level chain cost more than the rest of the kernel
// (p_stat is stored in FP64 either way). The evaporation tail is FP64 in both forms.
// ------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(128) void k_densities(const MemberDesc* __restrict__ md, int set_water, int evaporate, int gate) {
    using A = double;   // the hydrostatic column in FP64 in both storage precisions
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    const int j = (int)(q / K.km), k = (int)(q % K.km);
    S *t = fp<S>(M, F_T), *c = fp<S>(M, F_C), *co2 = fp<S>(M, F_CO2), *cl = fp<S>(M, F_CLOUD), *ice = fp<S>(M, F_ICE),
      *rh = fp<S>(M, F_R_HUMID), *rd = fp<S>(M, F_R_DRY);
    double* ps = pstat(M);
    const double Mb = K.M_bg;
    const A t_0 = A(K.t_0), Mb_a = A(K.M_bg), Rb = A(K.R_bg), inv_Mb = A(1.0 / K.M_bg), g = A(K.g);
    const size_t si = (size_t)K.jm * K.km;
    const size_t base = (size_t)j * K.km + k;
    A T_prev = tmax(A(180), A(t[base]) * t_0);
    A p_prev = A(K.p_sl_factor) * T_prev;
    A q_sat_min = A(1);
    A qv_prev = A(0), qc_prev = A(0);
    // the saturation pressure at the previous level's temperature, kept with that temperature:
    // above the cold trap the column is isothermal at t_skin, and a repeated temperature gives
    // the same pressure bit for bit without evaluating it again
    A T_sat = A(-1), E_sat = A(0);
    // this level's inputs; the next level's are loaded before this level's stores so their
    // latency overlaps the column arithmetic
    A qv = A(gld(c, (long long)base)), qc = A(gld(co2, (long long)base));
    A wcl = A(gld(cl, (long long)base)), wic = A(gld(ice, (long long)base));
    for (int i = 0; i < K.im; i++) {
        const long long id = (long long)base + (long long)i * (long long)si;
        A qv_n = A(0), qc_n = A(0), cl_n = A(0), ic_n = A(0);
        if (i + 1 < K.im) {
            qv_n = A(gld(c, id + (long long)si)); qc_n = A(gld(co2, id + (long long)si));
            cl_n = A(gld(cl, id + (long long)si)); ic_n = A(gld(ice, id + (long long)si));
        }
        const A R_loc = dens_R(dens_split(qv, qc), Rb);
        A Ti, pi;
        if (i == 0) { Ti = T_prev; pi = p_prev; }
        else {
            const A dz = A(M.hgt[i] - M.hgt[i - 1]);
            const MixQ<A> mp = dens_split(qv_prev, qc_prev);
            const A cp_loc = dens_cp(mp, T_prev, inv_Mb);
            const A M_nw = dens_Mnw(mp, Mb_a, inv_Mb);
            A q_s, Af;
            if (T_prev != T_sat) { E_sat = dens_p_auto(T_prev); T_sat = T_prev; }
            sat_qA(E_sat, p_prev, (p_prev > A(0)) ? rcp_nr(p_prev) : A(0), M_nw, q_s, Af);
            A gamma;
            if (qv_prev >= A(0.99) * q_s) gamma = dens_lapse(T_prev, Af, cp_loc, R_loc, g);
            else gamma = g * rcp_nr(cp_loc);
            const A T_ad = T_prev - gamma * dz;
            Ti = tmax(A(K.t_skin), T_ad);
            const A T_mean = A(0.5) * (T_prev + Ti);
            pi = p_prev * exp(-g * dz * rcp_nr(R_loc * T_mean));
        }
        gst(t, id, S(Ti / t_0));
        gst(ps, id, (double)pi);
        if (set_water && i > 0) {
            const double q_s_i = th_q_sat_at((double)Ti, (double)pi, th_M_nonwater((double)qv, (double)qc, Mb));
            if (A(q_s_i) < q_sat_min) q_sat_min = A(q_s_i);
            qv = tmax(q_sat_min, A(K.c_h2o_dry_top));
            gst(c, id, S(qv));
        }
        const A wf = tmax(A(0.5), A(1) - A(S(wcl)) - A(S(wic)));
        gst(rh, id, S(A(1e2) * pi * rcp_nr(R_loc * Ti * wf)));
        gst(rd, id, S(A(1e2) * pi * rcp_nr(dens_R(dens_split(A(0), qc), Rb) * Ti)));
        // the next level reads this level's stored water (the reference reads c[i-1] back)
        qv_prev = A(S(qv));
        qc_prev = qc;
        T_prev = Ti; p_prev = pi;
        qv = qv_n; qc = qc_n; wcl = cl_n; wic = ic_n;
    }
    M.t_top[base] = (double)t[base + (size_t)(K.im - 1) * si];
    if (!evaporate) return;
    // waterVapourEvaporation, sea branch
    const double t0s = (double)t[base] * K.t_0;
    if (t0s >= kTcrit) {
        if (M.evap) M.evap[base] = 0.0;   // no evaporation from a supercritical surface (ThermoAtm.h:179-185)
        return;
    }
    const double T_s = tmax(180.0, t0s);
    const double c0 = c[base];
    const double M_nw0 = th_M_nonwater(c0, (double)co2[base], Mb);
    const double E_s = th_p_auto(T_s);
    const double q_sea = th_q_sat(E_s, ps[base], M_nw0);
    if (M.evap) {
        // the Evaporation diagnostic [mm/d] that moist convection's surface-flux trigger reads:
        // Dalton's law on the saturation deficit at the surface (ThermoAtm.h:217-240, with
        // AtomUtils::C_Dalton, lib/Utils.cpp:157-172), computed before the surface update
        const double u0v = (double)fp<S>(M, F_U)[base], v0v = (double)fp<S>(M, F_V)[base];
        const double w0v = (double)fp<S>(M, F_W)[base];
        const double vel = sqrt((u0v * u0v + v0v * v0v + w0v * w0v) / 3.0) * K.u_0;
        const double c_dalton = K.coeff_dalton * sqrt(0.053 * 0.053 * vel / 10.0);
        const double p_s = ps[base];
        const double e_air = (q_sea > 0.0 && q_sea < 1.0) ? c0 * p_s * M_nw0 / (c0 * M_nw0 + (1.0 - c0) * kMH2O) : 0.0;
        const double sd = tmax(0.0, E_s - e_air);
        M.evap[base] = c_dalton * 24.0 * sd;
    }
    c[base] = S(c0 + (q_sea - c0) * K.evap_w_norm);
    double rp = 1.0;
    for (int i = 1; i <= 3 && i < K.im; i++) {
        const size_t id = base + (size_t)i * si;
        rp = pow(K.evap_r, (double)i);
        const double weight = rp * K.evap_w_norm;
        const double T_i = tmax(180.0, (double)t[id] * K.t_0);
        const double ci = c[id];
        const double q_s_i = th_q_sat_at(T_i, ps[id], th_M_nonwater(ci, (double)co2[id], Mb));
        c[id] = S(tmin(ci + (q_sea - c0) * weight, q_s_i));
    }
}

// ------------------------------------------------------------------------------------
// TurbulenceAtm::run (TurbulenceAtm.h:62-79): compute_vel_star, compute_sources with the
// k-omega SST closure, clamp_nue. compute_sources updates tke/dis in place while sweeping
// (j,k,i) in order; here every cell reads its neighbours' raw values and applies, for the
// neighbours a serial sweep would already have visited (i-1, j-1, k-1 inside the interior),
// the same idempotent floors that visit applied, which reproduces the serial sweep exactly
// without ordering. Outputs go to scratch q0..q5 = tke, dis, nue, prod, tke_src, dis_src.
// ------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_vel_star(const MemberDesc* __restrict__ md, int gate) {
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    const int isurf = min(1, K.im - 1);
    const size_t id = (size_t)isurf * K.jm * K.km + q;
    const double vs = (double)fp<S>(M, F_V)[id], ws = (double)fp<S>(M, F_W)[id];
    const double U = K.u_0 * sqrt(vs * vs + ws * ws);
    M.vel_star[q] = tmax(0.05, sqrt(0.002) * U);
}

template <typename S, int MODEL>
__global__ __launch_bounds__(256) void k_turb_sources(const MemberDesc* __restrict__ md, int gate) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    if (gated_off(M, gate)) return;
    const Consts& K = M.K;
    const long long id = cell_id(K);
    if (id < 0) return;
    int i, j, k;
    decode(K, id, i, j, k);
    const S *ptk = fp<S>(M, F_TKE), *pds = fp<S>(M, F_DIS);
    S *otk = (S*)M.tmp[0], *ods = (S*)M.tmp[1], *onu = (S*)M.tmp[2], *opr = (S*)M.tmp[3], *ots = (S*)M.tmp[4],
      *odss = (S*)M.tmp[5];
    const C nue_max = C(K.nue_max);
    const bool interior = (i > 0 && i < K.im - 1 && j > 0 && j < K.jm - 1 && k > 0 && k < K.km - 1);
    // Every value is read before any is written, and the floors below are applied to values
    // read unconditionally: a read inside a conditional, or after a write to a field the
    // compiler cannot tell apart from it, waits for a memory round trip of its own.
    if (!interior) {
        const S tv = ptk[id], dv = pds[id];
        const C nv = C(fp<S>(M, F_NUE)[id]);
        const S pv = fp<S>(M, F_PROD)[id], tsv = fp<S>(M, F_TKE_SRC)[id], dsv = fp<S>(M, F_DIS_SRC)[id];
        const C f_abl = C(M.nue_f[i]);
        otk[id] = tv; ods[id] = dv;
        onu[id] = S(tclamp(nv, C(0), nue_max) * f_abl);
        opr[id] = pv;
        ots[id] = tsv;
        odss[id] = dsv;
        return;
    }
    const size_t si = (size_t)K.jm * K.km, sj = K.km;
    const C u0 = C(K.u_0);
    auto kbg_of = [&](double vsd) {
        const C vs = C(vsd);
        return C(0.01) * vs * vs / (C(K.sqrt_cnue) * u0 * u0);
    };
    const double* vel_star = M.vel_star;
    const size_t q0 = (size_t)j * K.km + k;
    const double vs_c = vel_star[q0], vs_jm = vel_star[q0 - K.km], vs_km = vel_star[q0 - 1];
    const double vs_jp = vel_star[q0 + K.km];
    const C dis_min = C(1.0e-10);
    const C r_tk = C(ptk[id]), r_ds = C(pds[id]);
    const C r_tk_im1 = C(ptk[id - si]), r_ds_im1 = C(pds[id - si]);
    const C r_tk_jm1 = C(ptk[id - sj]), r_ds_jm1 = C(pds[id - sj]);
    const C r_tk_km1 = C(ptk[id - 1]), r_ds_km1 = C(pds[id - 1]);
    const C tk_ip1 = C(ptk[id + si]), ds_ip1 = C(pds[id + si]);
    const C r_tk_jp1 = C(ptk[id + sj]), r_ds_jp1 = C(pds[id + sj]);
    const C tk_kp1 = C(ptk[id + 1]), ds_kp1 = C(pds[id + 1]);
    const C f_abl = C(M.nue_f[i]);
    // visited neighbours: floors of the visit (max(0,.) then max(k_bg,.); dis max(dis_min,.))
    auto tke_vis = [&](C raw, double vs) { return tmax(kbg_of(vs), tmax(C(0), raw)); };
    auto dis_vis = [&](C raw) { return tmax(dis_min, raw); };
    const C tk = tmax(C(0), r_tk);
    const C ds = tmax(dis_min, r_ds);
    const C tk_im1 = (i - 1 >= 1) ? tke_vis(r_tk_im1, vs_c) : r_tk_im1;
    const C ds_im1 = (i - 1 >= 1) ? dis_vis(r_ds_im1) : r_ds_im1;
    const C tk_jm1 = (j - 1 >= 1) ? tke_vis(r_tk_jm1, vs_jm) : r_tk_jm1;
    const C ds_jm1 = (j - 1 >= 1) ? dis_vis(r_ds_jm1) : r_ds_jm1;
    const C tk_km1 = (k - 1 >= 1) ? tke_vis(r_tk_km1, vs_km) : r_tk_km1;
    const C ds_km1 = (k - 1 >= 1) ? dis_vis(r_ds_km1) : r_ds_km1;
    // the symmetric form floors both latitude neighbours (a serial sweep's visit order is
    // not mirror symmetric)
    const bool fl_jp = K.sym && (j + 1 <= K.jm - 2);
    const C tk_jp1 = fl_jp ? tke_vis(r_tk_jp1, vs_jp) : r_tk_jp1;
    const C ds_jp1 = fl_jp ? dis_vis(r_ds_jp1) : r_ds_jp1;

    const double rm = M.rad[i];
    const double exp_rm_d = 1.0 / (rm + 1.0);
    double sinthe = M.sinr[j];
    if (sinthe == 0.0) sinthe = 1.0e-5;
    const double rmsinthe = rm * sinthe;
    const double rmet = (K.metric_r0 > 0.0) ? (K.metric_r0 + (rm - K.rad0)) : rm;
    const C exp_rm = C(exp_rm_d), inv_2dr = C(K.inv_2dr);
    const C inv_rm2dthe = C(K.inv_2dthe / rmet), inv_rms2dphi = C(K.inv_2dphi / (rmet * sinthe));
    const S *u = fp<S>(M, F_U), *v = fp<S>(M, F_V), *w = fp<S>(M, F_W);
    const C dudr = (C(u[id + si]) - C(u[id - si])) * inv_2dr * exp_rm;
    const C dvdr = (C(v[id + si]) - C(v[id - si])) * inv_2dr * exp_rm;
    const C dwdr = (C(w[id + si]) - C(w[id - si])) * inv_2dr * exp_rm;
    const C dudthe = (C(u[id + sj]) - C(u[id - sj])) * inv_rm2dthe;
    const C dvdthe = (C(v[id + sj]) - C(v[id - sj])) * inv_rm2dthe;
    const C dwdthe = (C(w[id + sj]) - C(w[id - sj])) * inv_rm2dthe;
    const C dudphi = (C(u[id + 1]) - C(u[id - 1])) * inv_rms2dphi;
    const C dvdphi = (C(v[id + 1]) - C(v[id - 1])) * inv_rms2dphi;
    const C dwdphi = (C(w[id + 1]) - C(w[id - 1])) * inv_rms2dphi;
    const C dtkedr = (tk_ip1 - tk_im1) * inv_2dr * exp_rm;
    const C ddisdr = (ds_ip1 - ds_im1) * inv_2dr * exp_rm;
    constexpr int model = MODEL;   // one instantiation per closure: the others' registers stay out
    C cnue = (model == kTurbKEpsilon) ? C(0.028) * tmax(tk, C(0)) * tmax(tk, C(0)) / tmax(ds, dis_min)
                                      : tmax(tk, C(0)) / tmax(ds, dis_min);
    cnue = tmin(cnue, nue_max);
    const C der = C(0.66667) * (dudr + dvdthe + dwdphi);
    const C prod = tmax(C(0),
          (cnue * (C(2) * dudr - der) - C(0.66667) * tk) * dudr
        + (cnue * (dudthe + dvdr)) * dudthe
        + (cnue * (dudphi + dwdr)) * dudphi
        + (cnue * (C(2) * dvdthe - der) - C(0.66667) * tk) * dvdthe
        + (cnue * (dvdr + dudthe)) * dvdr
        + (cnue * (dvdphi + dwdthe)) * dvdphi
        + (cnue * (C(2) * dwdphi - der) - C(0.66667) * tk) * dwdphi
        + (cnue * (dwdr + dudphi)) * dwdr
        + (cnue * (dwdthe + dvdphi)) * dwdthe);
    const C W12 = dudthe - dvdr, W13 = dudphi - dwdr, W23 = dvdphi - dwdthe;
    const C Omega = sqrt(W12 * W12 + W13 * W13 + W23 * W23);
    const C rm2dthe = C(rm * 2.0 * K.dthe), rms2dphi = C(rmsinthe * 2.0 * K.dphi);
    const C dtkedthe = (tk_jp1 - tk_jm1) / rm2dthe, ddisdthe = (ds_jp1 - ds_jm1) / rm2dthe;
    const C dtkedphi = (tk_kp1 - tk_km1) / rms2dphi, ddisdphi = (ds_kp1 - ds_km1) / rms2dphi;
    const C cross = dtkedr * ddisdr + dtkedthe * ddisdthe + dtkedphi * ddisdphi;
    C nue, tsrc, dsrc;
    if constexpr (model == kTurbKEpsilon) {
        // compute_k_epsilon (TurbulenceAtm.h:481-530): Chien's damping functions and wall terms,
        // the wall distance from the surface floored at one L_atm
        const double y_phys = tmax(M.hgt[i] - M.hgt[0], K.L_atm);
        const C y_star = C(y_phys / K.L_atm);
        const C d_plus = C(y_phys * M.vel_star[(size_t)j * K.km + k] / 1.5e-5);
        const C f_nue = C(1) - exp(C(-0.0115) * d_plus);
        const C nu_air = C(1.5e-5);
        const C Re_T = tk * tk * C(K.u_0) * C(K.L_atm) / tmax(ds * nu_air, dis_min * nu_air);
        const C f_2 = C(1) - C(0.4 / 1.8) * exp(-Re_T * Re_T / C(36));
        nue = C(0.028) * tk * tk / tmax(ds, dis_min) * f_nue;
        const C y_star2 = y_star * y_star;
        const C L_k = C(-2) * tk / y_star2;
        const C L_w = C(-2) * ds / y_star2 * exp(-d_plus / C(2));
        const C tke_safe = tmax(tk, dis_min);
        const C P_w = C(1.35) * ds / tke_safe * prod;
        const C Y_w = C(1.80) * f_2 * ds * ds / tke_safe;
        tsrc = prod - ds + L_k / C(K.re_turb);
        dsrc = P_w - Y_w + L_w / C(K.re_turb);
    } else if constexpr (model == kTurbKOmega) {
        // compute_k_omega (TurbulenceAtm.h:533-651, grid rm): strain-rate limiter on omega,
        // vorticity-strain correction of beta, cross-diffusion above zero
        const C bet_star = C(0.09);
        const C r2dr = C(2.0 * K.dr), rp1 = C(rm + 1.0);
        const C S11 = (C(u[id + si]) - C(u[id - si])) / r2dr / rp1;
        const C S22 = (C(v[id + sj]) - C(v[id - sj])) / rm2dthe;
        const C S33 = (C(w[id + 1]) - C(w[id - 1])) / rms2dphi;
        const C S12 = C(0.5) * ((C(u[id + sj]) - C(u[id - sj])) / rm2dthe + (C(v[id + si]) - C(v[id - si])) / r2dr / rp1);
        const C S13 = C(0.5) * ((C(u[id + 1]) - C(u[id - 1])) / rms2dphi + (C(w[id + si]) - C(w[id - si])) / r2dr / rp1);
        const C S23 = C(0.5) * ((C(v[id + 1]) - C(v[id - 1])) / rms2dphi + (C(w[id + sj]) - C(w[id - sj])) / rm2dthe);
        const C S_mag = sqrt(C(2) * (S11 * S11 + S22 * S22 + S33 * S33 + C(2) * (S12 * S12 + S13 * S13 + S23 * S23)));
        const C w_hat = tmax(ds, C(0.875) * S_mag / C(0.3));   // sqrt(0.09), which rounds to 0.3
        const C sig_d = (cross <= C(0)) ? C(0) : C(0.125);
        const C bw = bet_star * tmax(ds, C(1.0e-20));
        const C chi_w = fabs(Omega * Omega * S_mag / (bw * bw * bw));
        const C f_bet = (C(1) + C(85) * chi_w) / (C(1) + C(100) * chi_w);
        nue = tk / tmax(w_hat, C(1.0e-20));
        const C P_k = tmin(prod, C(20) * bet_star * tk * ds);
        const C P_w = C(0.52) * ds / tmax(tk, C(1.0e-20)) * P_k;
        tsrc = P_k - bet_star * tk * ds;
        dsrc = P_w - C(0.0708) * f_bet * ds * ds + sig_d / tmax(ds, C(1.0e-20)) * cross;
    } else {
    // compute_k_omega_SST (grid rm in the theta/phi gradients, TurbulenceAtm.h:722-725)
    const C a1 = C(0.31), bet_star = C(0.09), bet1 = C(0.0333), bet2 = C(0.0368), gam1 = C(0.413), gam2 = C(0.2);
    const C sig_w2 = C(1.168);
    const C y_star = C(tmax(M.hgt[i] - M.hgt[0], 1.0e-6) / K.L_atm);
    const C CD_kw = tmax(C(2) * sig_w2 / tmax(ds, C(1.0e-20)) * cross, C(1.0e-20));
    const C dis_safe = tmax(ds, C(1.0e-20));
    const C nu_nd = C(K.nue_air_nd);
    const C arg1 = tmin(tmax(sqrt(tk) / (bet_star * dis_safe * y_star), C(500) * nu_nd / (y_star * y_star * dis_safe)),
                        C(4) * sig_w2 * tk / (CD_kw * y_star * y_star));
    const C arg2 = tmax(C(2) * sqrt(tk) / (bet_star * dis_safe * y_star), C(500) * nu_nd / (y_star * y_star * dis_safe));
    const C a1sq = arg1 * arg1;
    const C F1 = tanh_pos(a1sq * a1sq);
    const C F2 = tanh_pos(arg2 * arg2);
    nue = a1 * tk / tmax(a1 * dis_safe, Omega * F2);
    const C tke_safe = tmax(tk, C(1.0e-20));
    const C P_k = tmin(prod, C(20) * bet_star * tke_safe * dis_safe);
    const C Y_k = bet_star * tke_safe * dis_safe;
    const C P_w = (F1 * gam1 + (C(1) - F1) * gam2) * P_k * dis_safe / tke_safe;
    const C Y_w = (F1 * bet1 + (C(1) - F1) * bet2) * dis_safe * dis_safe;
    const C D_w = C(2) * (C(1) - F1) * sig_w2 / dis_safe * cross;
    tsrc = P_k - Y_k;
    dsrc = P_w - Y_w + D_w;
    }
    const C tk_out = tmax(kbg_of(vs_c), tk);
    const C ds_out = tmax(dis_min, ds);
    nue = tclamp(nue, C(0), nue_max);
    const C dsm = C(20) * ds_out * ds_out;
    dsrc = tclamp(dsrc, -dsm, dsm);
    const C tsm = C(20) * tk_out * ds_out;
    tsrc = tclamp(tsrc, -tsm, tsm);
    otk[id] = S(tk_out);
    ods[id] = S(ds_out);
    onu[id] = S(tclamp(nue, C(0), nue_max) * f_abl);
    opr[id] = S(prod);
    ots[id] = S(tsrc);
    odss[id] = S(dsrc);
}

// ------------------------------------------------------------------------------------
// Boundary conditions (every iteration): bcRadius + apply_wall_bc per (j,k) column,
// bcTheta per (i,k), bcPhi + seam smoothing per (i,j), bcSolidGround clamps per cell.
// BC_Atm.h:556-917, :23-551; TurbulenceAtm.h:109-147.
// ------------------------------------------------------------------------------------
// The boundary kernels load every value they need before storing any (the field pointers
// may alias as far as the compiler knows, so interleaved loads and stores would serialise
// on memory round trips); the final values equal those of the reference's sequence of
// assignments.
template <typename S>
__global__ __launch_bounds__(256) void k_bc_radial(const MemberDesc* __restrict__ md) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    const long long si = (long long)K.jm * K.km;
    const int iml = K.im - 1;
    auto at = [&](int i) { return (long long)i * si + q; };
    // bcRadius (BC_Atm.h:556-745): cubic bottom / log-linear lid for the hydrostatic fields
    double* p = pstat(M);
    const double p1 = gld(p, at(1)), p2 = gld(p, at(2)), p3 = gld(p, at(3));
    const double pl1 = gld(p, at(iml - 1)), pl2 = gld(p, at(iml - 2));
    S* rh = fp<S>(M, F_R_HUMID);
    S* rd = fp<S>(M, F_R_DRY);
    C h[5], d[5];
    h[0] = C(gld(rh, at(1))); h[1] = C(gld(rh, at(2))); h[2] = C(gld(rh, at(3)));
    h[3] = C(gld(rh, at(iml - 1))); h[4] = C(gld(rh, at(iml - 2)));
    d[0] = C(gld(rd, at(1))); d[1] = C(gld(rd, at(2))); d[2] = C(gld(rd, at(3)));
    d[3] = C(gld(rd, at(iml - 1))); d[4] = C(gld(rd, at(iml - 2)));
    // velocities: v, w copied at the bottom, zero at the lid, tapered x1/3, x2/3 below it
    C vw[4][3];   // v, w, vn, wn at i = 1, iml-1, iml-2
    const int fvw[4] = {F_V, F_W, F_VN, F_WN};
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const S* x = fp<S>(M, fvw[n]);
        vw[n][0] = (n < 2) ? C(gld(x, at(1))) : C(0);
        vw[n][1] = C(gld(x, at(iml - 1)));
        vw[n][2] = C(gld(x, at(iml - 2)));
    }
    // scalars: cubic bottom, copied lid
    constexpr int kRad[10] = {F_CLOUD, F_ICE, F_GR, F_CO2, F_TKE, F_DIS, F_NUE, F_PROD, F_TKE_SRC, F_DIS_SRC};
    C sc[10][4];
#pragma unroll
    for (int n = 0; n < 10; n++) {
        const S* x = fp<S>(M, kRad[n]);
        sc[n][0] = C(gld(x, at(1))); sc[n][1] = C(gld(x, at(2))); sc[n][2] = C(gld(x, at(3)));
        sc[n][3] = C(gld(x, at(iml - 1)));
    }
    const double vs = M.vel_star[q];
    // stores
    gst(fp<S>(M, F_T), at(iml), S(M.t_top[q]));
    gst(p, at(0), p3 - 3.0 * p2 + 3.0 * p1);
    gst(p, at(iml), (pl1 > 0.0 && pl2 > 0.0) ? (pl1 * pl1 / pl2) : pl1);
    gst(rh, at(0), S(h[2] - C(3) * h[1] + C(3) * h[0]));
    gst(rh, at(iml), S((h[3] > C(0) && h[4] > C(0)) ? (h[3] * h[3] / h[4]) : h[3]));
    gst(rd, at(0), S(d[2] - C(3) * d[1] + C(3) * d[0]));
    gst(rd, at(iml), S((d[3] > C(0) && d[4] > C(0)) ? (d[3] * d[3] / d[4]) : d[3]));
#pragma unroll
    for (int n = 0; n < 4; n++) {
        S* x = fp<S>(M, fvw[n]);
        if (n < 2) gst(x, at(0), S(vw[n][0]));
        gst(x, at(iml), S(0));
        gst(x, at(iml - 1), S(vw[n][1] * C(1.0 / 3.0)));
        gst(x, at(iml - 2), S(vw[n][2] * C(2.0 / 3.0)));
    }
    for (int f : {F_U, F_UN}) { gst(fp<S>(M, f), at(0), S(0)); gst(fp<S>(M, f), at(iml), S(0)); }
#pragma unroll
    for (int n = 0; n < 10; n++) {
        S* x = fp<S>(M, kRad[n]);
        C bottom = sc[n][2] - C(3) * sc[n][1] + C(3) * sc[n][0];
        // apply_wall_bc (TurbulenceAtm.h:109-147, k-omega SST)
        if (kRad[n] == F_TKE) bottom = C(vs * vs / (K.sqrt_cnue * K.u_0 * K.u_0));
        if (kRad[n] == F_NUE) bottom = C(0);
        if (kRad[n] == F_DIS) bottom = C(K.om_wall);
        gst(x, at(0), S(bottom));
        gst(x, at(iml), S(sc[n][3]));
    }
}

constexpr int kThetaF[18] = {F_T, F_P_STAT, F_R_HUMID, F_R_DRY, F_U, F_V, F_W, F_C, F_CLOUD, F_ICE, F_GR, F_CO2,
                             F_TKE, F_DIS, F_NUE, F_PROD, F_TKE_SRC, F_DIS_SRC};

// bcTheta (BC_Atm.h:746-917): the pole rows copy their neighbours. One thread per
// (field, level, longitude).
template <typename S>
__global__ __launch_bounds__(256) void k_bc_theta(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    // each field's (i, k) range padded to whole waves: the field index is then wave-uniform,
    // so the field table entry and the field pointer are scalar loads
    const long long nik = (long long)K.im * K.km, nikp = (nik + 63) & ~63LL;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= 18 * nikp) return;
    const int n = (int)__builtin_amdgcn_readfirstlane((int)(q / nikp));
    const long long r = q - n * nikp;
    if (r >= nik) return;
    const int i = (int)(r / K.km), k = (int)(r - (long long)i * K.km);
    const long long a0 = ix3(K, i, 0, k), a1 = ix3(K, i, 1, k), b0 = ix3(K, i, K.jm - 1, k), b1 = ix3(K, i, K.jm - 2, k);
    if (kThetaF[n] == F_P_STAT) {
        double* x = pstat(M);
        const double va = gld(x, a1), vb = gld(x, b1);
        gst(x, a0, va); gst(x, b0, vb);
    } else {
        S* x = fp<S>(M, kThetaF[n]);
        const S va = gld(x, a1), vb = gld(x, b1);
        gst(x, a0, va); gst(x, b0, vb);
    }
}

// bcPhi (BC_Atm.h:918-1025): seam columns averaged (t, p_stat, r_humid, r_dry, u, v, w, c,
// cloud, ice, gr, co2) or copied (the six turbulence fields), then two 1-2-1 passes across the
// periodic seam on u, un, v, vn, w, wn (BC_Atm.h:1026-1130). One thread per (field, row); u, v,
// w do their averaging and seam passes in the same thread.
constexpr int kPhiF[21] = {F_T, F_P_STAT, F_R_HUMID, F_R_DRY, F_U, F_V, F_W, F_C, F_CLOUD, F_ICE, F_GR, F_CO2,
                           F_TKE, F_DIS, F_NUE, F_PROD, F_TKE_SRC, F_DIS_SRC, F_UN, F_VN, F_WN};

// e = x[km-3], a = x[km-2], b = x[0] = x[km-1], c = x[1], f = x[2]
template <typename C> ATHAD_DEV void seam_passes(C e, C& a, C& b, C& c, C f) {
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
        const C a1 = a + C(0.25) * (e - C(2) * a + b);
        const C b1 = b + C(0.25) * (a - C(2) * b + c);
        const C c1 = c + C(0.25) * (b - C(2) * c + f);
        a = a1; b = b1; c = c1;
    }
}

template <typename S>
__global__ __launch_bounds__(256) void k_bc_phi(const MemberDesc* __restrict__ md) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    // each field's row range padded to whole waves, so the field index is wave-uniform
    const long long nrow = (long long)K.im * K.jm, nrowp = (nrow + 63) & ~63LL;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= 21 * nrowp) return;
    const int n = (int)__builtin_amdgcn_readfirstlane((int)(q / nrowp));
    if (q - n * nrowp >= nrow) return;
    const long long r = (q - n * nrowp) * K.km;
    const int km = K.km;
    const int f = kPhiF[n];
    if (f == F_P_STAT) {
        double* ps = pstat(M);
        const double v = (gld(ps, r + 1) + gld(ps, r + km - 2)) * 0.5;
        gst(ps, r, v); gst(ps, r + km - 1, v);
        return;
    }
    S* x = fp<S>(M, f);
    if (n < 12) {   // averaged seam
        const C x1 = C(gld(x, r + 1)), x2 = C(gld(x, r + km - 2));
        const C avg = (x1 + x2) * C(0.5);
        if (f == F_U || f == F_V || f == F_W) {
            const C e = C(gld(x, r + km - 3)), ff = C(gld(x, r + 2));
            C a = x2, b = avg, c = x1;
            seam_passes(e, a, b, c, ff);
            gst(x, r + km - 2, S(a)); gst(x, r, S(b)); gst(x, r + km - 1, S(b)); gst(x, r + 1, S(c));
        } else {
            gst(x, r, S(avg)); gst(x, r + km - 1, S(avg));
        }
    } else if (n < 18) {   // turbulence fields: copied
        const S v0 = gld(x, r + 1), v1 = gld(x, r + km - 2);
        gst(x, r, v0); gst(x, r + km - 1, v1);
    } else {   // un, vn, wn: seam passes only
        const C e = C(gld(x, r + km - 3)), a0 = C(gld(x, r + km - 2)), b0 = C(gld(x, r)), c0 = C(gld(x, r + 1)),
                ff = C(gld(x, r + 2));
        C a = a0, b = b0, c = c0;
        seam_passes(e, a, b, c, ff);
        gst(x, r + km - 2, S(a)); gst(x, r, S(b)); gst(x, r + km - 1, S(b)); gst(x, r + 1, S(c));
    }
}

// bcSolidGround passes 4-5 (used by the lexicographic form and the initial state; the
// Jacobi form applies the same clamps inside k_rk4_jacobi).
template <typename S>
__global__ void k_solid_clamp(const MemberDesc* __restrict__ md) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long id = cell_id(K);
    if (id < 0) return;
    const C tkm = C(K.tke_max_nd), vm = C(K.vel_max_nd);
    for (int f = 0; f < NPROG; f++) {
        if (f == FT || f == FG || f == FCO) continue;
        S* x = fp<S>(M, kProgF[f]);
        x[id] = S(solid_clamp(f, C(x[id]), tkm, vm));
    }
    for (int f : {FU, FV, FW, FK, FD}) {
        S* x = fp<S>(M, kProgN[f]);
        x[id] = S(solid_clamp_n(f, C(x[id]), tkm, vm));
    }
    {
        S* x = fp<S>(M, F_NUE); const C v = C(x[id]);
        x[id] = S(finite_bits(v) ? tclamp(v, C(0), C(K.nue_max)) : C(0));
    }
    {
        S* x = fp<S>(M, F_PROD); const C v = C(x[id]);
        x[id] = S(finite_bits(v) ? tclamp(v, C(0), C(1.0e4)) : C(0));
    }
    for (int f : {F_TKE_SRC, F_DIS_SRC}) {
        S* x = fp<S>(M, f); const C v = C(x[id]);
        x[id] = S(finite_bits(v) ? tclamp(v, C(-1.0e4), C(1.0e4)) : C(-1.0e4));
    }
}

// ------------------------------------------------------------------------------------
// Polar zonal filter (lib/Utils.cpp:542-608): per (i,j) row, p_j periodic 1-2-1 passes
// along k, staged in LDS. One block per row.
// ------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_polar(const MemberDesc* __restrict__ md, int fbits) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const int row = blockIdx.x;
    if (row >= K.im * K.jm) return;
    const int j = row % K.jm;
    const int np = ro_load(M.passes + j);
    if (np == 0) return;
    // the selected fields of u, v, w pass together: 2 buffers x 3 fields x km in LDS
    extern __shared__ unsigned char smem_raw[];
    const int km = K.km;
    C* a = (C*)smem_raw;
    C* b = a + 3 * km;
    bool on[3];
#pragma unroll
    for (int f = 0; f < 3; f++) on[f] = (fbits >> f) & 1;
    for (int k = threadIdx.x; k < km; k += blockDim.x) {
#pragma unroll
        for (int f = 0; f < 3; f++)
            if (on[f]) a[f * km + k] = C(fp<S>(M, F_U + f)[(size_t)row * km + k]);
    }
    __syncthreads();
    for (int p = 0; p < np; p++) {
        for (int k = threadIdx.x; k < km; k += blockDim.x) {
            const int km1 = (k == 0) ? km - 1 : k - 1, kp1 = (k == km - 1) ? 0 : k + 1;
#pragma unroll
            for (int f = 0; f < 3; f++) {
                if (!on[f]) continue;
                const C* r = a + f * km;
                b[f * km + k] = r[k] + C(0.25) * (r[km1] - C(2) * r[k] + r[kp1]);
            }
        }
        __syncthreads();
        C* t = a; a = b; b = t;
    }
    for (int k = threadIdx.x; k < km; k += blockDim.x) {
#pragma unroll
        for (int f = 0; f < 3; f++)
            if (on[f]) fp<S>(M, F_U + f)[(size_t)row * km + k] = S(a[f * km + k]);
    }
}

// The same filter with the row in registers: one wave per (row, field); lane l holds the E
// consecutive longitudes E l .. E l + E - 1 (E = ceil(km / 64)), so a pass only needs the edge
// values of the neighbouring lanes (cross-lane moves, no barrier). Same operations as k_polar.
template <typename S, int E>
__global__ __launch_bounds__(256) void k_polar_wave(const MemberDesc* __restrict__ md) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const int lane = threadIdx.x & 63;
    // the wave's index in its block, made wave-uniform for the compiler: the row, the field and
    // its pointer then sit in scalar registers and the pass count is a scalar load
    const int wv = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long pair = (long long)blockIdx.x * (256 >> 6) + wv;
    if (pair >= 3LL * K.im * K.jm) return;   // uniform per wave
    const int row = (int)(pair / 3), f = (int)(pair % 3);
    const int np = ro_load(M.passes + row % K.jm);
    if (np == 0) return;
    const int km = K.km, last = km - 1, l_last = last / E, s_last = last % E;
    S* x = fp<S>(M, F_U + f) + (size_t)row * km;
    // the row is read and written through a descriptor sized to it: the buffer range check
    // gives zero for the lanes' slots past longitude km - 1 and drops their stores, so no
    // element needs a branch (a branch per store made each store wait for the previous one)
    const Rsrc xr = make_rsrc(x, (unsigned)km * (unsigned)sizeof(S));
    C v[E];
#pragma unroll
    for (int s = 0; s < E; s++) v[s] = C(bld<S>(xr, (unsigned)(E * lane + s) * (unsigned)sizeof(S), 0));
    for (int p = 0; p < np; p++) {
        const C up = dpp_move<kDppWaveShr1>(v[E - 1], C(0));   // lane l-1's last longitude
        const C dn = dpp_move<kDppWaveShl1>(v[0], C(0));       // lane l+1's first longitude
        C lv = v[0];
#pragma unroll
        for (int s = 1; s < E; s++)
            if (s == s_last) lv = v[s];
        const C v_last = read_lane(lv, l_last);   // longitude km-1 (l_last is wave-uniform)
        const C v_first = read_lane(v[0], 0);     // longitude 0
        C nv[E];
#pragma unroll
        for (int s = 0; s < E; s++) {
            const int e = E * lane + s;
            const C left = (s > 0) ? v[s - 1] : ((lane > 0) ? up : v_last);
            const C right = (e == last) ? v_first : ((s < E - 1) ? v[s + 1] : dn);
            nv[s] = v[s] + C(0.25) * (left - C(2) * v[s] + right);
        }
#pragma unroll
        for (int s = 0; s < E; s++) v[s] = nv[s];
    }
#pragma unroll
    for (int s = 0; s < E; s++) bst<S>(xr, (unsigned)(E * lane + s) * (unsigned)sizeof(S), S(v[s]));
}

// Radial Shapiro filters (lib/Utils.cpp:831-933): two passes along i of u (1-2-1) and of
// v, w (4th order, 1-2-1 next to the wall/lid), one thread per (j,k) column with the column
// in LDS; optionally the store xn <- x of UtilsAtm.h:25-49 for u, v, w.
template <typename S>
__global__ __launch_bounds__(64) void k_radial_store(const MemberDesc* __restrict__ md, int store) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    extern __shared__ unsigned char smem_raw[];
    C* col = (C*)smem_raw + (size_t)threadIdx.x * K.im;
    const size_t si = (size_t)K.jm * K.km;
    const int im = K.im;
    for (int f = F_U; f <= F_W; f++) {
        S* x = fp<S>(M, f) + q;
        const bool ho = (f != F_U);
        for (int i = 0; i < im; i++) col[i] = C(x[(size_t)i * si]);
        for (int pass = 0; pass < 2; pass++) {
            C o1 = col[0], o2 = C(0);   // old values at i-1 and i-2
            for (int i = 1; i < im - 1; i++) {
                const C c0 = col[i];
                C nv;
                if (ho && i >= 2 && i + 2 <= im - 1) {
                    const C d4 = col[i + 2] - C(4) * col[i + 1] + C(6) * c0 - C(4) * o1 + o2;
                    nv = c0 - C(1.0 / 16.0) * d4;
                } else {
                    nv = c0 + C(0.25) * (o1 - C(2) * c0 + col[i + 1]);
                }
                o2 = o1;
                o1 = c0;
                col[i] = nv;
            }
        }
        S* xn = fp<S>(M, f - F_U + F_UN) + q;
        for (int i = 0; i < im; i++) {
            const S v = S(col[i]);
            x[(size_t)i * si] = v;
            if (store) xn[(size_t)i * si] = v;
        }
    }
}

// The same filters on a tile: a block holds RK consecutive longitudes x all levels of one
// latitude row of u, v, w in LDS. The reference sweeps each pass in place but keeps the old
// values of levels i-1, i-2, so every level of a pass depends only on the previous pass and
// all levels are computed at once (two LDS buffers). Results equal the column kernel's.
template <typename S, int RK>
__global__ __launch_bounds__(256) void k_radial_tile(const MemberDesc* __restrict__ md, int store) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const int nkc = (K.km + RK - 1) / RK;
    if ((int)blockIdx.x >= K.jm * nkc) return;
    const int j = blockIdx.x / nkc, k0 = (blockIdx.x % nkc) * RK;
    const int im = K.im, n = 3 * im * RK;
    extern __shared__ unsigned char smem_raw[];
    C* a = (C*)smem_raw;
    C* b = a + n;
    const size_t si = (size_t)K.jm * K.km;
    const size_t base = (size_t)j * K.km + k0;
    // element (f, i, kk) at (f * im + i) * RK + kk; thread t owns column kk = t % RK and
    // levels i = t / RK, + NR, + 2 NR, ...
    constexpr int NR = 256 / RK;
    const int kk = threadIdx.x % RK, r0 = threadIdx.x / RK;
    const bool in = (k0 + kk < K.km);
#pragma unroll
    for (int f = 0; f < 3; f++)
        for (int i = r0; i < im; i += NR)
            if (in) a[(f * im + i) * RK + kk] = C(fp<S>(M, F_U + f)[base + (size_t)i * si + kk]);
    __syncthreads();
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const C* col = a + (size_t)f * im * RK + kk;
            for (int i = r0; i < im; i += NR) {
                const C c0 = col[i * RK];
                C nv = c0;
                if (i >= 1 && i <= im - 2) {
                    const C o1 = col[(i - 1) * RK];
                    if (f != 0 && i >= 2 && i + 2 <= im - 1) {
                        const C d4 = col[(i + 2) * RK] - C(4) * col[(i + 1) * RK] + C(6) * c0 - C(4) * o1 + col[(i - 2) * RK];
                        nv = c0 - C(1.0 / 16.0) * d4;
                    } else {
                        nv = c0 + C(0.25) * (o1 - C(2) * c0 + col[(i + 1) * RK]);
                    }
                }
                b[(f * im + i) * RK + kk] = nv;
            }
        }
        __syncthreads();
        C* t = a; a = b; b = t;
    }
    if (!in) return;
#pragma unroll
    for (int f = 0; f < 3; f++)
        for (int i = r0; i < im; i += NR) {
            const S v = S(a[(f * im + i) * RK + kk]);
            const size_t g = base + (size_t)i * si + kk;
            fp<S>(M, F_U + f)[g] = v;
            if (store) fp<S>(M, F_UN + f)[g] = v;
        }
}

// The same filters with the column in registers: one thread per (field, j, k) column, so a
// wave's loads of one level are 64 consecutive longitudes of one row, and every level's loads
// are issued before the first use. IMX is a compile-time bound on the levels (all loops are
// unrolled so the column stays in registers); same operations as the kernels above.
template <typename S, int IMX>
__global__ __launch_bounds__(256) void k_radial_regs(const MemberDesc* __restrict__ md, int store) {
    using C = S;
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long plane = (long long)K.jm * K.km;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * plane) return;
    const int f = (int)(q / plane);
    const long long o = q - (long long)f * plane;
    const int im = K.im;
    const size_t si = (size_t)plane;
    S* x = fp<S>(M, F_U + f) + o;
    const bool ho = (f != 0);
    C col[IMX + 2];
#pragma unroll
    for (int i = 0; i < IMX + 2; i++) col[i] = C(0);
#pragma unroll
    for (int i = 0; i < IMX; i++)
        if (i < im) col[i] = C(x[(size_t)i * si]);
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
        C o1 = col[0], o2 = C(0);   // old values at i-1 and i-2
#pragma unroll
        for (int i = 1; i < IMX - 1; i++) {
            if (i < im - 1) {
                const C c0 = col[i];
                C nv;
                if (ho && i >= 2 && i + 2 <= im - 1) {
                    const C d4 = col[i + 2] - C(4) * col[i + 1] + C(6) * c0 - C(4) * o1 + o2;
                    nv = c0 - C(1.0 / 16.0) * d4;
                } else {
                    nv = c0 + C(0.25) * (o1 - C(2) * c0 + col[i + 1]);
                }
                o2 = o1;
                o1 = c0;
                col[i] = nv;
            }
        }
    }
    S* xn = fp<S>(M, F_UN + f) + o;
#pragma unroll
    for (int i = 0; i < IMX; i++)
        if (i < im) {
            const S v = S(col[i]);
            x[(size_t)i * si] = v;
            if (store) xn[(size_t)i * si] = v;
        }
}

// UtilsAtm::storeIntermediateData3D (UtilsAtm.h:25-49): xn <- x for the fields selected by mask.
template <typename S>
__global__ void k_store(const MemberDesc* __restrict__ md, int mask) {
    const MemberDesc& M = md[blockIdx.z];
    const long long id = cell_id(M.K);
    if (id < 0) return;
    for (int f = 0; f < NPROG; f++)
        if (mask & (1 << f)) fp<S>(M, kProgN[f])[id] = fp<S>(M, kProgF[f])[id];
}

// ------------------------------------------------------------------------------------
// Initial state on the device (the recipe the oracle's init follows):
// surface parabola + dry column (InitValues_Atm.cpp:493-632), initWaterWapour (:843-905),
// co2Atmosphere (ThermoAtm.h:873-913), seeded N-S mirrored velocity perturbation.
// ------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(128) void k_init_column(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const double t_eq_K = K.t_surf_eq, t_pole_K = K.t_surf_pole;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    const int j = (int)(q / K.km);
    const size_t si = (size_t)K.jm * K.km;
    S *t = fp<S>(M, F_T), *rh = fp<S>(M, F_R_HUMID), *rd = fp<S>(M, F_R_DRY);
    double* ps = pstat(M);
    const double d_j_half = 0.5 * (K.jm - 1);
    const double ratio = (double)(K.sym ? min(j, K.jm - 1 - j) : j) / d_j_half;
    const double T_s = (t_pole_K - t_eq_K) * (ratio * ratio - 2.0 * ratio) + t_pole_K;
    const double tu = (T_s * (1.0 / K.t_0)) * K.t_0;
    double T_prev = tmax(100.0, tu);
    double p_prev = 1e-2 * (K.r_air * K.R_mix * tu);
    const double R_loc = th_R_of(0.0, 0.0, K.R_bg);
    for (int i = 0; i < K.im; i++) {
        double Tc, pv;
        if (i == 0) { Tc = T_prev; pv = p_prev; }
        else {
            const double dz = (double)(float)((float)M.hgt[i] - (float)M.hgt[i - 1]);
            const double cp_loc = th_cp_of(0.0, 0.0, T_prev, K.M_bg);
            const double T_ad = T_prev - (K.g / cp_loc) * dz;
            Tc = tmax(K.t_skin, T_ad);
            const double Tm = 0.5 * (T_prev + Tc);
            pv = p_prev * exp(-K.g * dz / (R_loc * Tm));
        }
        const size_t id = (size_t)i * si + q;
        t[id] = S(Tc * (1.0 / K.t_0));
        ps[id] = pv;
        rd[id] = S(pv * 100.0 / (th_R_of(0.0, 0.0, K.R_bg) * Tc));
        rh[id] = S(pv * 100.0 / (R_loc * Tc * 1.0));
        T_prev = Tc; p_prev = pv;
    }
}

template <typename S>
__global__ __launch_bounds__(128) void k_init_water(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const double M_nw = K.init_M_nw;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    const size_t si = (size_t)K.jm * K.km;
    S *t = fp<S>(M, F_T), *c = fp<S>(M, F_C);
    const double* ps = pstat(M);
    double q_sat_min = 1.0;
    for (int i = 0; i < K.im; i++) {
        const size_t id = (size_t)i * si + q;
        const double q_s = th_q_sat_at((double)t[id] * K.t_0, ps[id], M_nw);
        if (q_s < q_sat_min) q_sat_min = q_s;
        c[id] = S(tmax(q_sat_min, K.c_h2o_dry_top));
    }
}

template <typename S>
__global__ void k_init_co2(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const long long id = cell_id(M.K);
    if (id < 0) return;
    fp<S>(M, F_CO2)[id] = S((1.0 - (double)fp<S>(M, F_C)[id]) * M.K.init_f_co2 * M.K.co2_scale);
}

template <typename S>
__global__ void k_copy_lid(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    M.t_top[q] = (double)fp<S>(M, F_T)[(size_t)(K.im - 1) * K.jm * K.km + q];
}

ATHAD_DEV uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template <typename S>
__global__ void k_init_rest(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const unsigned long long seed = K.seed;
    const long long id = cell_id(K);
    if (id < 0) return;
    int i, j, k;
    decode(K, id, i, j, k);
    for (int f : {F_CLOUD, F_ICE, F_GR, F_P_DYN, F_NUE, F_PROD, F_TKE_SRC, F_DIS_SRC, F_AUX_U, F_AUX_V, F_AUX_W, F_S_C_C})
        fp<S>(M, f)[id] = S(0);
    const double amp = 1.0e-3 / K.u_0;
    const int jeq = (K.jm - 1) / 2;
    const int jj = (j <= jeq) ? j : (K.jm - 1 - j);
    for (int f = 0; f < 3; f++) {
        const uint64_t ctr = ((((uint64_t)f * K.im + i) * K.jm + jj) * K.km + k);
        const uint64_t r = splitmix64(seed * 0xD1B54A32D192ED03ULL + ctr);
        double val = (2.0 * ((double)(r >> 11) * (1.0 / 9007199254740992.0)) - 1.0) * amp;
        if (f == 1) {
            if (j == jeq) val = 0.0;
            else if (j > jeq) val = -val;
        }
        fp<S>(M, F_U + f)[id] = S(val);
    }
    fp<S>(M, F_TKE)[id] = S(1.0e-4 / (K.u_0 * K.u_0));
    fp<S>(M, F_DIS)[id] = S(1.0 * K.L_atm / K.u_0);
}

// TurbulenceAtm::init_fields (TurbulenceAtm.h:193-255): per interior cell the boundary-layer
// profile k = u_tau^2 / sqrt(C_mu) (1 - z/h_abl)^2 above 1 % of its surface value, and with it
// omega = C_mu^-1/4 sqrt(k) / (kappa z) (k-omega closures) or eps = C_mu^3/4 k^3/2 / (kappa z)
// (k-epsilon) above the same expressions at the background k and z = L_atm; z is the height
// above the surface, written to tke, dis and their stored copies.
template <typename S>
__global__ __launch_bounds__(256) void k_turb_init_fields(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long id = cell_id(K);
    if (id < 0) return;
    int i, j, k;
    decode(K, id, i, j, k);
    if (i < 1 || i > K.im - 2 || j < 1 || j > K.jm - 2 || k < 1 || k > K.km - 2) return;
    constexpr double C_nue = 0.028, Karman = 0.42;
    const double vs = M.vel_star[(size_t)j * K.km + k];
    const double k_bg = 0.01 * vs * vs / sqrt(C_nue);
    const double z_i = M.hgt[i] - M.hgt[0];
    const double z_safe = tmax(z_i, 1.0e-6);
    const double z_frac = tmin(z_i, K.abl_height) / K.abl_height;
    const double k_abl = vs * vs / sqrt(C_nue) * ((1.0 - z_frac) * (1.0 - z_frac));
    const double k_phys = tmax(k_abl, k_bg);
    double dis_nd;
    if (K.turb == kTurbKEpsilon) {
        const double eps_bg = pow(C_nue, 0.75) * pow(k_bg, 1.5) / (Karman * K.L_atm);
        const double eps_phys = pow(C_nue, 0.75) * pow(k_phys, 1.5) / (Karman * z_safe);
        dis_nd = tmax(eps_phys, eps_bg) * (K.L_atm / (K.u_0 * K.u_0 * K.u_0));
    } else {
        const double om_bg = pow(C_nue, -0.25) * sqrt(k_bg) / (Karman * K.L_atm);
        const double om_phys = pow(C_nue, -0.25) * sqrt(k_phys) / (Karman * z_safe);
        dis_nd = tmax(om_phys, om_bg) * (K.L_atm / K.u_0);
    }
    const S tk = S(tmax(0.0, k_phys * (1.0 / (K.u_0 * K.u_0))));
    const S ds = S(tmax(1.0e-10, dis_nd));
    fp<S>(M, F_TKE)[id] = tk;
    fp<S>(M, F_TKEN)[id] = tk;
    fp<S>(M, F_DIS)[id] = ds;
    fp<S>(M, F_DISN)[id] = ds;
}

// TurbulenceAtm::apply_wall_bc (TurbulenceAtm.h:109-147) alone, per (j,k) column: the log-layer
// tke, zero eddy viscosity and the closure's wall value of dis at i = 0.
template <typename S>
__global__ void k_wall_bc(const MemberDesc* __restrict__ md) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)K.jm * K.km) return;
    const double vs = M.vel_star[q];
    fp<S>(M, F_TKE)[q] = S(vs * vs / (K.sqrt_cnue * K.u_0 * K.u_0));
    fp<S>(M, F_NUE)[q] = S(0);
    fp<S>(M, F_DIS)[q] = S(K.om_wall);
}

template <typename S>
__global__ void k_copy_field(const MemberDesc* __restrict__ md, int dst, int src) {
    const MemberDesc& M = md[blockIdx.z];
    const long long id = cell_id(M.K);
    if (id < 0) return;
    fp<S>(M, dst)[id] = fp<S>(M, src)[id];
}

// PressureSolverAtm::project_initial_velocity (PressureSolverAtm.h:686-767), around the
// solver's sweeps (launched between the phases). phase 0: aux = (u, v, w), p_dyn = 0 (:694-706);
// phase 1: the interior gradient correction v -= grad p_dyn with the RHS metric (sin theta
// floored at 0.55, metric radius) (:715-749); phase 2: p_dyn = aux = 0 (:752-763).
template <typename S>
__global__ __launch_bounds__(256) void k_project(const MemberDesc* __restrict__ md, int phase) {
    const MemberDesc& M = md[blockIdx.z];
    const Consts& K = M.K;
    const long long id = cell_id(K);
    if (id < 0) return;
    S *u = fp<S>(M, F_U), *v = fp<S>(M, F_V), *w = fp<S>(M, F_W), *p = fp<S>(M, F_P_DYN);
    S *au = fp<S>(M, F_AUX_U), *av = fp<S>(M, F_AUX_V), *aw = fp<S>(M, F_AUX_W);
    if (phase == 0) {
        au[id] = u[id]; av[id] = v[id]; aw[id] = w[id]; p[id] = S(0);
        return;
    }
    if (phase == 2) {
        p[id] = S(0); au[id] = S(0); av[id] = S(0); aw[id] = S(0);
        return;
    }
    int i, j, k;
    decode(K, id, i, j, k);
    if (i < 1 || i > K.im - 2 || j < 1 || j > K.jm - 2 || k < 1 || k > K.km - 2) return;
    const long long si = (long long)K.jm * K.km, sj = K.km;
    const double rm = M.rad[i];
    const double exp_rm = 1.0 / (rm + 1.0);
    const double rmet = (K.metric_r0 > 0.0) ? (K.metric_r0 + (rm - K.rad0)) : rm;
    const double inv_rm = 1.0 / rmet;
    const double inv_rmsinthe = 1.0 / (rmet * M.sint[j]);
    const double dpdr = ((double)p[id + si] - (double)p[id - si]) * K.inv_2dr;
    const double dpdthe = ((double)p[id + sj] - (double)p[id - sj]) * K.inv_2dthe;
    const double dpdphi = ((double)p[id + 1] - (double)p[id - 1]) * K.inv_2dphi;
    u[id] = S((double)u[id] - dpdr * exp_rm);
    v[id] = S((double)v[id] - dpdthe * inv_rm);
    w[id] = S((double)w[id] - dpdphi * inv_rmsinthe);
}

// ------------------------------------------------------------------------------------
// Host launch wrappers
// ------------------------------------------------------------------------------------
namespace {
inline unsigned cells_grid(const Batch& b, int bs) {
    const long long nb = (b.max_n + bs - 1) / bs;
    return (unsigned)(((nb + 7) / 8) * 8);   // multiple of 8 for the XCD remap
}
inline unsigned cols_grid(const Batch& b, int bs) {
    return (unsigned)(((long long)b.max_jm * b.max_km + bs - 1) / bs);
}
}  // namespace

#define ATHAD_DISPATCH(kern, grid, block, shm, ...)                                              \
    do {                                                                                          \
        if (b.fp32) hipLaunchKernelGGL(kern<float>, grid, block, shm, b.stream, __VA_ARGS__);     \
        else hipLaunchKernelGGL(kern<double>, grid, block, shm, b.stream, __VA_ARGS__);           \
    } while (0)

void launch_advance(const Batch& b) {
    hipLaunchKernelGGL(k_advance, dim3(1, 1, b.nmem), dim3(64), 0, b.stream, b.d_md);
}
void launch_field_stats(const Batch& b, int member, int field) {
    ATHAD_DISPATCH(k_field_stats, dim3(kStatChunks), dim3(256), 0, b.d_md, member, field);
    hipLaunchKernelGGL(k_field_stats_final, dim3(1), dim3(1), 0, b.stream, b.d_md, member);
}
int field_stats_offset(int im) { return 2 * kLevelChunks * im + 2 * kStatChunks; }
int diag_offset(int im) { return 2 * kLevelChunks * im + 2 * kStatChunks + 2; }
int sf_offset(int im) { return diag_offset(im) + 6 * kStatChunks + 8; }
void launch_slice(const Batch& b, int member, int field, int axis, int index, long long n) {
    ATHAD_DISPATCH(k_slice, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b.d_md, member, field, axis, index);
}
void launch_radiation(const Batch& b) {
    const dim3 g((b.max_km + kRadCols - 1) / kRadCols, b.max_jm, b.nmem);
    const size_t lds = (size_t)6 * b.max_im * kRadCols * sizeof(double);
    if (b.max_im <= 64) {
        if (b.fp32) hipLaunchKernelGGL((k_radiation<float, 1>), g, dim3(256), lds, b.stream, b.d_md);
        else hipLaunchKernelGGL((k_radiation<double, 1>), g, dim3(256), lds, b.stream, b.d_md);
    } else {
        if (b.fp32) hipLaunchKernelGGL((k_radiation<float, 2>), g, dim3(256), lds, b.stream, b.d_md);
        else hipLaunchKernelGGL((k_radiation<double, 2>), g, dim3(256), lds, b.stream, b.d_md);
    }
    hipLaunchKernelGGL(k_albedo_rows, dim3(b.max_jm, 1, b.nmem), dim3(256), 0, b.stream, b.d_md);
}
void launch_skin_step(const Batch& b, MemberDesc* md0, MemberDesc* md1, double relax) {
    hipLaunchKernelGGL(k_skin_step, dim3(b.nmem), dim3(256), 0, b.stream, md0, md1, relax);
}
// 80 levels keep the tile's LDS (6 fields x levels x 16 columns) within 64 KB.
int radiation_max_levels() { return 80; }

void launch_ice_scheme(const Batch& b, int gate) {
    const dim3 gc((unsigned)((b.max_jm * b.max_km + 255) / 256), 1, b.nmem);
    const dim3 gb((unsigned)((b.max_n + 255) / 256), 1, b.nmem);
    if (b.fp32) {
        hipLaunchKernelGGL(k_ice_columns<float>, gc, dim3(256), 0, b.stream, b.d_md, gate);
        hipLaunchKernelGGL(k_ice_bc<float>, gb, dim3(256), 0, b.stream, b.d_md, gate);
    } else {
        hipLaunchKernelGGL(k_ice_columns<double>, gc, dim3(256), 0, b.stream, b.d_md, gate);
        hipLaunchKernelGGL(k_ice_bc<double>, gb, dim3(256), 0, b.stream, b.d_md, gate);
    }
    // run_3D_loop's damp_wiggles of the two fluxes (cAtmosphereModel.cpp:1243-1244) through
    // scratch buffers 6-9 (0-5 may hold the turbulence update running beside the moist block)
    const int src[2] = {NFIELD + NTMP + MP_P_RAIN, NFIELD + NTMP + MP_P_SNOW};
    const int t0[2] = {NFIELD + 6, NFIELD + 7}, t1[2] = {NFIELD + 8, NFIELD + 9};
    launch_damp(b, 0, src, t0, 2, 0, gate);
    launch_damp(b, 1, t0, t1, 2, 0, gate);
    launch_damp(b, 2, t1, src, 2, 0, gate);
}

void launch_moist_clamp_ice(const Batch& b, int gate, int store) {
    const dim3 g((unsigned)((b.max_n + 255) / 256), 1, b.nmem);
    if (b.fp32) hipLaunchKernelGGL(k_moist_clamp_ice<float>, g, dim3(256), 0, b.stream, b.d_md, gate, store);
    else hipLaunchKernelGGL(k_moist_clamp_ice<double>, g, dim3(256), 0, b.stream, b.d_md, gate, store);
}

void launch_bc_sources(const Batch& b) {
    const dim3 g((unsigned)((b.max_n + 255) / 256), 1, b.nmem);
    if (b.fp32) hipLaunchKernelGGL(k_bc_sources<float>, g, dim3(256), 0, b.stream, b.d_md);
    else hipLaunchKernelGGL(k_bc_sources<double>, g, dim3(256), 0, b.stream, b.d_md);
}

void launch_bc_conv(const Batch& b) {
    const dim3 g((unsigned)((b.max_n + 255) / 256), 1, b.nmem);
    hipLaunchKernelGGL(k_bc_conv, g, dim3(256), 0, b.stream, b.d_md);
}

void launch_ice_store(const Batch& b) {
    const dim3 g((unsigned)((b.max_n + 255) / 256), 1, b.nmem);
    if (b.fp32) hipLaunchKernelGGL(k_ice_store<float>, g, dim3(256), 0, b.stream, b.d_md);
    else hipLaunchKernelGGL(k_ice_store<double>, g, dim3(256), 0, b.stream, b.d_md);
}
void launch_streamfunction(const Batch& b, int member, int im, int jm, double two_pi_a) {
    ATHAD_DISPATCH(k_zonal_mean_v, dim3((im * jm + 63) / 64), dim3(64), 0, b.d_md, member);
    hipLaunchKernelGGL(k_streamfunction, dim3((jm + 63) / 64), dim3(64), 0, b.stream, b.d_md, member, two_pi_a);
}
void launch_minmax(const Batch& b, int member, int field) {
    ATHAD_DISPATCH(k_minmax, dim3(kStatChunks), dim3(256), 0, b.d_md, member, field);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(1), 0, b.stream, b.d_md, member);
}
void launch_residuum(const Batch& b, int member) {
    ATHAD_DISPATCH(k_residuum, dim3(kStatChunks), dim3(256), 0, b.d_md, member);
    hipLaunchKernelGGL(k_residuum_final, dim3(1), dim3(1), 0, b.stream, b.d_md, member);
}
void launch_level_mean(const Batch& b) {
    ATHAD_DISPATCH(k_level_mean, dim3(kLevelChunks, b.max_im, b.nmem), dim3(256), 0, b.d_md);
    hipLaunchKernelGGL(k_level_mean_final, dim3((b.max_im + 63) / 64, 1, b.nmem), dim3(64), 0, b.stream, b.d_md);
}
namespace {
template <typename S, int SRC> void rk4_jacobi_t(const Batch& b, const dim3& g, bool fused) {
    if (fused) hipLaunchKernelGGL((k_rk4_jacobi<S, true, SRC>), g, dim3(kCellBlock), 0, b.stream, b.d_md);
    else hipLaunchKernelGGL((k_rk4_jacobi<S, false, SRC>), g, dim3(kCellBlock), 0, b.stream, b.d_md);
}
template <typename S> void rk4_jacobi_s(const Batch& b, const dim3& g, bool fused, int src) {
    switch (src & 3) {
        case 0: rk4_jacobi_t<S, 0>(b, g, fused); break;
        case 1: rk4_jacobi_t<S, 1>(b, g, fused); break;
        case 2: rk4_jacobi_t<S, 2>(b, g, fused); break;
        default: rk4_jacobi_t<S, 3>(b, g, fused); break;
    }
}
template <typename S, int SRC> void rk4_rows_t(const Batch& b, const dim3& g, bool fused) {
    if (fused) hipLaunchKernelGGL((k_rk4_rows<S, true, SRC>), g, dim3(256), 0, b.stream, b.d_md);
    else hipLaunchKernelGGL((k_rk4_rows<S, false, SRC>), g, dim3(256), 0, b.stream, b.d_md);
}
template <typename S> void rk4_rows_s(const Batch& b, const dim3& g, bool fused, int src) {
    switch (src & 3) {
        case 0: rk4_rows_t<S, 0>(b, g, fused); break;
        case 1: rk4_rows_t<S, 1>(b, g, fused); break;
        case 2: rk4_rows_t<S, 2>(b, g, fused); break;
        default: rk4_rows_t<S, 3>(b, g, fused); break;
    }
}
template <typename S, int LBX> void rk4_lex_t(const Batch& b, int plane, int src) {
    const long long nq = (long long)(b.max_im - 2) * (b.max_jm - 2);
    const dim3 g((unsigned)((nq + LBX - 1) / LBX), 1, b.nmem);
    switch (src & 3) {
        case 0: hipLaunchKernelGGL((k_rk4_lex_plane<S, 0, LBX>), g, dim3(LBX), 0, b.stream, b.d_md, plane); break;
        case 1: hipLaunchKernelGGL((k_rk4_lex_plane<S, 1, LBX>), g, dim3(LBX), 0, b.stream, b.d_md, plane); break;
        case 2: hipLaunchKernelGGL((k_rk4_lex_plane<S, 2, LBX>), g, dim3(LBX), 0, b.stream, b.d_md, plane); break;
        default: hipLaunchKernelGGL((k_rk4_lex_plane<S, 3, LBX>), g, dim3(LBX), 0, b.stream, b.d_md, plane); break;
    }
}
template <typename S> void rk4_lex_roles_t(const Batch& b, int plane, int src) {
    const long long nq = (long long)(b.max_im - 2) * (b.max_jm - 2);
    const dim3 g((unsigned)((nq + 63) / 64), 1, b.nmem);
    switch (src & 3) {
        case 0: hipLaunchKernelGGL((k_rk4_lex_roles<S, 0>), g, dim3(192), 0, b.stream, b.d_md, plane); break;
        case 1: hipLaunchKernelGGL((k_rk4_lex_roles<S, 1>), g, dim3(192), 0, b.stream, b.d_md, plane); break;
        case 2: hipLaunchKernelGGL((k_rk4_lex_roles<S, 2>), g, dim3(192), 0, b.stream, b.d_md, plane); break;
        default: hipLaunchKernelGGL((k_rk4_lex_roles<S, 3>), g, dim3(192), 0, b.stream, b.d_md, plane); break;
    }
}
template <typename S> void rk4_lex_s(const Batch& b, int plane, int src) {
    // While the waves of a plane fit the device a few times over (one member: ~110 blocks of
    // 64 cells), each cell's groups run on three waves (k_rk4_lex_roles: 9.12 -> 8.15 ms per
    // iteration at 41 x 181 x 361 FP64, against one wave per cell); with many members the device
    // is full anyway and 256-thread blocks of the one-wave kernel run (5 % faster than 64-thread
    // blocks at 32 members).
    const long long waves = (long long)b.nmem * (((long long)(b.max_im - 2) * (b.max_jm - 2) + 63) / 64);
    if (waves <= 1024) rk4_lex_roles_t<S>(b, plane, src);
    else rk4_lex_t<S, 256>(b, plane, src);
}
}  // namespace
// FP64: the row-segment kernel, blocks of four latitude rows taken i fastest, then the
// longitude segment, then the row group (measured at 41 x 181 x 361: 0.447 ms against 0.450
// with single waves taken i fastest, then segment, then j; 0.476 with the storage order, 0.498
// with the segment slowest; the cell-linear FP64 kernel 0.530). FP32: the cell-linear kernel
// at five waves per SIMD (0.214 ms; the row kernel 0.245).
void launch_rk4_jacobi(const Batch& b, bool fused, int src) {
    if (!b.fp32) {
        const int nseg = (b.max_km + 63) / 64;
        const long long nb = (long long)b.max_im * ((b.max_jm + 3) / 4) * nseg;
        const dim3 g((unsigned)(((nb + 7) / 8) * 8), 1, b.nmem);
        rk4_rows_s<double>(b, g, fused, src);
        return;
    }
    const dim3 g(cells_grid(b, 256), 1, b.nmem);
    rk4_jacobi_s<float>(b, g, fused, src);
}
void launch_rk4_lex(const Batch& b, int plane, int src) {
    if (b.fp32) rk4_lex_s<float>(b, plane, src);
    else rk4_lex_s<double>(b, plane, src);
}
void launch_poisson(const Batch& b, int gate) {
    const unsigned na = (unsigned)(((long long)(b.max_jm - 2) * (b.max_km - 2) + 255) / 256);
    ATHAD_DISPATCH(k_aux_bc, dim3(na, 1, b.nmem), dim3(256), 0, b.d_md, gate);
    ATHAD_DISPATCH(k_poisson_colour, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, 0, gate);
    ATHAD_DISPATCH(k_poisson_colour, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, 1, gate);
    ATHAD_DISPATCH(k_p_bc, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, gate);
    ATHAD_DISPATCH(k_p_ceiling, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, gate);
}
void launch_satadj(const Batch& b, int gate, int to_tmp) {
    ATHAD_DISPATCH(k_satadj, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, gate, to_tmp);
}
void launch_damp(const Batch& b, int axis, const int* src, const int* dst, int nf, int moist_clamp, int gate) {
    DampJob job{};
    job.nf = nf;
    for (int q = 0; q < nf; q++) { job.src[q] = src[q]; job.dst[q] = dst[q]; }
    job.moist_clamp = moist_clamp;
    const dim3 g(cells_grid(b, 256), 1, b.nmem);
    if (b.fp32) {
        if (axis == 0) hipLaunchKernelGGL((k_damp_pass<float, 0>), g, dim3(kCellBlock), 0, b.stream, b.d_md, job, gate);
        else if (axis == 1) hipLaunchKernelGGL((k_damp_pass<float, 1>), g, dim3(kCellBlock), 0, b.stream, b.d_md, job, gate);
        else hipLaunchKernelGGL((k_damp_pass<float, 2>), g, dim3(kCellBlock), 0, b.stream, b.d_md, job, gate);
    } else {
        if (axis == 0) hipLaunchKernelGGL((k_damp_pass<double, 0>), g, dim3(kCellBlock), 0, b.stream, b.d_md, job, gate);
        else if (axis == 1) hipLaunchKernelGGL((k_damp_pass<double, 1>), g, dim3(kCellBlock), 0, b.stream, b.d_md, job, gate);
        else hipLaunchKernelGGL((k_damp_pass<double, 2>), g, dim3(kCellBlock), 0, b.stream, b.d_md, job, gate);
    }
}
namespace {
struct TileShape { int tj, tk; };
constexpr TileShape kTiles[] = {{2, 64}, {4, 64}, {8, 64}, {1, 128}, {2, 128}, {4, 128}, {8, 32}, {16, 32}};
constexpr int kNTiles = (int)(sizeof(kTiles) / sizeof(kTiles[0]));

template <typename S, int TJ, int TK>
void damp_tile_launch(const Batch& b, const DampJob& job, int ich, int gate) {
    const long long ntk = (b.max_km + TK - 1) / TK, ntj = (b.max_jm + TJ - 1) / TJ;
    const long long nch = (b.max_im + ich - 1) / ich;
    const unsigned nb = (unsigned)(((ntk * ntj * nch + 7) / 8) * 8);
    if (job.nf == 3) hipLaunchKernelGGL((k_damp_tile<S, TJ, TK, 3>), dim3(nb, 1, b.nmem), dim3(TJ * TK), 0, b.stream, b.d_md, job, ich, gate);
    else if (job.nf == 4) hipLaunchKernelGGL((k_damp_tile<S, TJ, TK, 4>), dim3(nb, 1, b.nmem), dim3(TJ * TK), 0, b.stream, b.d_md, job, ich, gate);
    else hipLaunchKernelGGL((k_damp_tile<S, TJ, TK, 6>), dim3(nb, 1, b.nmem), dim3(TJ * TK), 0, b.stream, b.d_md, job, ich, gate);
}
template <typename S>
void damp_tile_dispatch(const Batch& b, int shape, const DampJob& job, int ich, int gate) {
    switch (shape) {
        case 0: damp_tile_launch<S, 2, 64>(b, job, ich, gate); break;
        case 1: damp_tile_launch<S, 4, 64>(b, job, ich, gate); break;
        case 2: damp_tile_launch<S, 8, 64>(b, job, ich, gate); break;
        case 3: damp_tile_launch<S, 1, 128>(b, job, ich, gate); break;
        case 4: damp_tile_launch<S, 2, 128>(b, job, ich, gate); break;
        case 5: damp_tile_launch<S, 4, 128>(b, job, ich, gate); break;
        case 6: damp_tile_launch<S, 8, 32>(b, job, ich, gate); break;
        default: damp_tile_launch<S, 16, 32>(b, job, ich, gate); break;
    }
}
}  // namespace

int damp_tile_shapes() { return kNTiles; }
void damp_tile_shape(int shape, int* tj, int* tk) {
    const TileShape t = kTiles[(shape >= 0 && shape < kNTiles) ? shape : 0];
    *tj = t.tj; *tk = t.tk;
}
void launch_damp_tile(const Batch& b, int shape, int ich, const int* src, const int* dst, int nf, int moist_clamp,
                      int gate) {
    DampJob job{};
    job.nf = nf;
    for (int q = 0; q < nf; q++) { job.src[q] = src[q]; job.dst[q] = dst[q]; }
    job.moist_clamp = moist_clamp;
    if (!b.d_md || ich < 1) return;
    if (b.fp32) damp_tile_dispatch<float>(b, shape, job, ich, gate);
    else damp_tile_dispatch<double>(b, shape, job, ich, gate);
}
void launch_damp_level0(const Batch& b, int src, int dst, int gate) {
    ATHAD_DISPATCH(k_damp_level0, dim3(cols_grid(b, 256), 1, b.nmem), dim3(256), 0, b.d_md, src, dst, gate);
}
void launch_copy_job(const Batch& b, const int* src, const int* dst, int nf, int gate) {
    DampJob job{};
    job.nf = nf;
    for (int q = 0; q < nf; q++) { job.src[q] = src[q]; job.dst[q] = dst[q]; }
    ATHAD_DISPATCH(k_copy_job, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, job, gate);
}
void launch_densities(const Batch& b, int set_water, int evaporate, int gate) {
    ATHAD_DISPATCH(k_densities, dim3(cols_grid(b, 128), 1, b.nmem), dim3(128), 0, b.d_md, set_water, evaporate, gate);
}
void launch_vel_star(const Batch& b, int gate) {
    ATHAD_DISPATCH(k_vel_star, dim3(cols_grid(b, 256), 1, b.nmem), dim3(256), 0, b.d_md, gate);
}
template <int MODEL> void launch_turb_sources_m(const Batch& b, int gate) {
    const dim3 grid(cells_grid(b, 256), 1, b.nmem);
    if (b.fp32) hipLaunchKernelGGL((k_turb_sources<float, MODEL>), grid, dim3(kCellBlock), 0, b.stream, b.d_md, gate);
    else hipLaunchKernelGGL((k_turb_sources<double, MODEL>), grid, dim3(kCellBlock), 0, b.stream, b.d_md, gate);
}
void launch_turb_sources(const Batch& b, int gate) {
    if (b.turb == kTurbKEpsilon) launch_turb_sources_m<kTurbKEpsilon>(b, gate);
    else if (b.turb == kTurbKOmega) launch_turb_sources_m<kTurbKOmega>(b, gate);
    else launch_turb_sources_m<kTurbKOmegaSST>(b, gate);
}
void launch_bc(const Batch& b) {
    ATHAD_DISPATCH(k_bc_radial, dim3(cols_grid(b, 256), 1, b.nmem), dim3(256), 0, b.d_md);
    // each field's range padded to a multiple of 64 threads, as the kernels index it
    const long long nikp = ((long long)b.max_im * b.max_km + 63) & ~63LL;
    const unsigned nt = (unsigned)((18LL * nikp + 255) / 256);
    ATHAD_DISPATCH(k_bc_theta, dim3(nt, 1, b.nmem), dim3(256), 0, b.d_md);
    const long long nrowp = ((long long)b.max_im * b.max_jm + 63) & ~63LL;
    const unsigned np = (unsigned)((21LL * nrowp + 255) / 256);
    ATHAD_DISPATCH(k_bc_phi, dim3(np, 1, b.nmem), dim3(256), 0, b.d_md);
}
void launch_solid_clamp(const Batch& b) {
    ATHAD_DISPATCH(k_solid_clamp, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md);
}
namespace {
// Rows of up to 768 longitudes run the register kernel (one wave per row and field), longer
// ones the LDS kernel.
constexpr bool kPolarWave = true;
template <int E> void polar_wave_launch(const Batch& b) {
    const dim3 g((unsigned)((3LL * b.max_im * b.max_jm + 3) / 4), 1, b.nmem);
    if (b.fp32) hipLaunchKernelGGL((k_polar_wave<float, E>), g, dim3(256), 0, b.stream, b.d_md);
    else hipLaunchKernelGGL((k_polar_wave<double, E>), g, dim3(256), 0, b.stream, b.d_md);
}
}  // namespace
void launch_polar(const Batch& b) {
    if (kPolarWave && b.max_km <= 768) {
        if (b.max_km <= 192) polar_wave_launch<3>(b);
        else if (b.max_km <= 384) polar_wave_launch<6>(b);
        else polar_wave_launch<12>(b);
        return;
    }
    const size_t shm = 6 * (size_t)b.max_km * (b.fp32 ? 4 : 8);
    ATHAD_DISPATCH(k_polar, dim3((unsigned)(b.max_im * b.max_jm), 1, b.nmem), dim3(256), shm, b.d_md, 7);
}
namespace {
// Up to 64 levels the register-column kernel runs (rows of 64 longitudes per load, no LDS):
// at 41x181x361 FP64 55 us per launch against 81-88 us for the LDS tile kernel. Taller
// columns take the tile kernel.
constexpr bool kRadialRegs = true;
template <int IMX> void radial_regs_launch(const Batch& b, int store) {
    const dim3 g((unsigned)((3LL * b.max_jm * b.max_km + 255) / 256), 1, b.nmem);
    if (b.fp32) hipLaunchKernelGGL((k_radial_regs<float, IMX>), g, dim3(256), 0, b.stream, b.d_md, store);
    else hipLaunchKernelGGL((k_radial_regs<double, IMX>), g, dim3(256), 0, b.stream, b.d_md, store);
}
}  // namespace
void launch_radial(const Batch& b, int store) {
    if (kRadialRegs && b.max_im <= 64) {
        if (b.max_im <= 24) radial_regs_launch<24>(b, store);
        else if (b.max_im <= 48) radial_regs_launch<48>(b, store);
        else radial_regs_launch<64>(b, store);
        return;
    }
    constexpr int RK = 16;
    const size_t tile = 2 * 3 * (size_t)b.max_im * RK * (b.fp32 ? 4 : 8);
    if (tile <= 64 * 1024) {
        const unsigned nb = (unsigned)(b.max_jm * ((b.max_km + RK - 1) / RK));
        if (b.fp32) hipLaunchKernelGGL((k_radial_tile<float, RK>), dim3(nb, 1, b.nmem), dim3(256), tile, b.stream, b.d_md, store);
        else hipLaunchKernelGGL((k_radial_tile<double, RK>), dim3(nb, 1, b.nmem), dim3(256), tile, b.stream, b.d_md, store);
        return;
    }
    const size_t shm = 64 * (size_t)b.max_im * (b.fp32 ? 4 : 8);
    ATHAD_DISPATCH(k_radial_store, dim3(cols_grid(b, 64), 1, b.nmem), dim3(64), shm, b.d_md, store);
}
void launch_store(const Batch& b, int mask) {
    ATHAD_DISPATCH(k_store, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, mask);
}
void launch_project(const Batch& b, int phase) {
    ATHAD_DISPATCH(k_project, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, phase);
}
void launch_copy_field(const Batch& b, int dst, int src) {
    ATHAD_DISPATCH(k_copy_field, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md, dst, src);
}
void launch_init_column(const Batch& b) {
    ATHAD_DISPATCH(k_init_column, dim3(cols_grid(b, 128), 1, b.nmem), dim3(128), 0, b.d_md);
}
void launch_init_water(const Batch& b) {
    ATHAD_DISPATCH(k_init_water, dim3(cols_grid(b, 128), 1, b.nmem), dim3(128), 0, b.d_md);
}
void launch_init_co2(const Batch& b) {
    ATHAD_DISPATCH(k_init_co2, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md);
}
void launch_copy_lid(const Batch& b) {
    ATHAD_DISPATCH(k_copy_lid, dim3(cols_grid(b, 256), 1, b.nmem), dim3(256), 0, b.d_md);
}
void launch_init_rest(const Batch& b) {
    ATHAD_DISPATCH(k_init_rest, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md);
}
void launch_turb_init_fields(const Batch& b) {
    ATHAD_DISPATCH(k_turb_init_fields, dim3(cells_grid(b, kCellBlock), 1, b.nmem), dim3(kCellBlock), 0, b.d_md);
}
void launch_wall_bc(const Batch& b) {
    ATHAD_DISPATCH(k_wall_bc, dim3(cols_grid(b, 256), 1, b.nmem), dim3(256), 0, b.d_md);
}

}  // namespace athad

namespace athad {

// ------------------------------------------------------------------------------------
// Moist convection, MoistConvection::run (MoistConvection.h:106-140), with the model's
// defaults iter_prec = 3, convection_perturbation = 1, convection_mode = 1 and no land.
// One thread per (j, k) column. Every phase of the reference reads only its own column's
// convection arrays and the fields c, cloud, t, u, v, w, p_stat, r_humid, co2 (c also at
// the four horizontal neighbours, which nothing here writes), so a thread runs the phases of
// its column back to back: precompute's clearing, findCloudBaseLFS, the surface-flux
// perturbation, initUpdraft, CloudBaseUpdate, three rounds of updraft entrainment / base
// update / updraft recurrence / downdraft entrainment / downdraft recurrence, rhsForcing.
// computeCAPE is not run (a 1-D diagnostic that ends holding the last column's values and
// feeds nothing). The convection arrays are FP64 in every precision; fields are read in
// their storage precision S. Consecutive threads take consecutive k, so every level's loads
// and stores are coalesced across the wave.
// ------------------------------------------------------------------------------------
namespace conv {
constexpr double t_00 = 236.15, t_00_m2 = 236.15 - 2.0;
constexpr double eps_u = 0.2 / 100.0, del_u = 0.2 / 100.0, b_u = 0.33;   // MoistConvection.h:21-88
constexpr double eps_d = 1.0e-4, del_d = 1.0e-4, gam_d = -0.3, a_u = 3.0e-2, a_d = 3.0e-2, vel = 1.1;
constexpr double a_ev = 1.0e-3, b_ev = 5.9, alf_1 = 0.05, alf_2 = 0.05, bet_p = 2.0e-3;
constexpr double f_beg = 1000.0 / 1013.25, f_end = 970.0 / 1013.25, f_cb = 900.0 / 1013.25;
constexpr double f_diff = 200.0 / 1013.25, f_mid = 700.0 / 1013.25;
constexpr double t_add_u = 0.2, q_v_u_add = 1.0e-4, coeff_recurr = 0.1, scale = 0.98;
constexpr double q_c_u_max = 1.0e-2, w_conv_max = 50.0, M_max = 3.0, c_mb = 0.003, alpha_sfp = 0.2;
constexpr double MCt_max = 0.01, MCq_max = 2.0e-5, MCv_max = 0.01;     // :1323-1339
constexpr double cp_l = 1349.0, s_0 = 274515.75, p_0 = 60000.0;       // cAtmosphereDefaults.cpp.inc:25,99,109

ATHAD_DEV double cap(double v, double mag) {                          // :216-222, :1341-1346
    if (!finite_bits(v)) return 0.0;
    return (v < -mag) ? -mag : (v > mag) ? mag : v;
}
}  // namespace conv

// The saturation mass fractions every column pass needs at each cell, computed beforehand
// with one thread per cell (the column kernel runs one wave per SIMD, where these IAPWS
// evaluations would sit on its serial path): [0] the air lifted by t_add_u (findCloudBaseLFS's
// q_sat_col, :675-677, and initUpdraft's where its perturbation is t_add_u), [1] the air at rest
// (the downdraft's, :935-936, :976-977, and initUpdraft's where its perturbation is zero).
template <typename S>
__global__ __launch_bounds__(256) void k_conv_sat(const MemberDesc* __restrict__ mds, int gate) {
    const MemberDesc& d = mds[blockIdx.z];
    if (gated_off(d, gate)) return;
    const Consts& K = d.K;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= K.n) return;
    const double t_u = (double)((const S*)d.f[F_T])[q] * K.t_0;
    const double p = ((const double*)d.f[F_P_STAT])[q];
    const double Mo = th_M_nonwater((double)((const S*)d.f[F_C])[q], K.co2_0, K.M_bg);
    d.conv_qs[0][q] = th_q_sat(th_p_auto(t_u + conv::t_add_u), p, Mo);
    d.conv_qs[1][q] = th_q_sat(th_p_auto(t_u), p, Mo);
}

// computeCAPE (MoistConvection.h:1259-1299) fills one array of im values, cleared, in the
// column order k outer, j inner, so each level ends holding the last column that reached it.
// Here every column writes its profile to scratch (conv_tmp[0]) and the range of levels it
// reaches (conv_cape_rng); k_cape_select then finds, per level (one block each), the column of
// largest order index k jm + j that reaches it, and takes the level's value from its profile.
__global__ __launch_bounds__(256) void k_cape_select(const MemberDesc* __restrict__ mds, int gate) {
    const MemberDesc& d = mds[blockIdx.z];
    if (gated_off(d, gate)) return;
    const Consts& K = d.K;
    const int i = blockIdx.x;
    if (i >= K.im) return;
    __shared__ int best[256];
    int b = -1;
    for (int c = threadIdx.x; c < K.jm * K.km; c += blockDim.x) {
        const int lo = d.conv_cape_rng[2 * c], hi = d.conv_cape_rng[2 * c + 1];
        if (i >= lo && i <= hi) b = max(b, (c % K.km) * K.jm + c / K.km);
    }
    best[threadIdx.x] = b;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) best[threadIdx.x] = max(best[threadIdx.x], best[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int o = best[0];
        d.conv_cape[i] = (o < 0) ? 0.0 : d.conv_tmp[0][((size_t)i * K.jm + o % K.jm) * K.km + o / K.jm];
    }
}

template <typename S>
__global__ __launch_bounds__(64) void k_moist_convection(const MemberDesc* __restrict__ mds, int gate) {
    using namespace conv;
    const MemberDesc& d = mds[blockIdx.z];
    if (gated_off(d, gate)) return;
    const Consts& K = d.K;
    const int im = K.im, jm = K.jm, km = K.km;
    const int col = blockIdx.x * 64 + threadIdx.x;
    if (col >= jm * km) return;
    const int j = col / km, k = col - j * km;
    const size_t pl = (size_t)jm * km;
    // global-memory pointers (plain pointers read from a descriptor would compile to flat
    // accesses, whose waits also cover the scalar loads)
    typedef const G1<S>* __restrict__ FieldP;
    typedef G1<double>* __restrict__ ArrP;
    FieldP T = (FieldP)d.f[F_T];
    FieldP U = (FieldP)d.f[F_U];
    FieldP V = (FieldP)d.f[F_V];
    FieldP W = (FieldP)d.f[F_W];
    FieldP C = (FieldP)d.f[F_C];
    FieldP CL = (FieldP)d.f[F_CLOUD];
    FieldP CO2 = (FieldP)d.f[F_CO2];
    FieldP RH = (FieldP)d.f[F_R_HUMID];
    const G1<double>* __restrict__ PS = (const G1<double>*)d.f[F_P_STAT];
    ArrP Sx = (ArrP)d.mc[MC_S];   ArrP Mu = (ArrP)d.mc[MC_M_U];
    ArrP Md = (ArrP)d.mc[MC_M_D]; ArrP Su = (ArrP)d.mc[MC_S_U];
    ArrP Qvu = (ArrP)d.mc[MC_Q_V_U]; ArrP Qcu = (ArrP)d.mc[MC_Q_C_U];
    ArrP Cu = (ArrP)d.mc[MC_C_U]; ArrP El = (ArrP)d.mc[MC_E_L];
    ArrP Edv = (ArrP)d.mc[MC_E_D_EVAP]; ArrP Ep = (ArrP)d.mc[MC_E_P];
    ArrP Gp = (ArrP)d.mc[MC_G_P]; ArrP Sd = (ArrP)d.mc[MC_S_D];
    ArrP Qvd = (ArrP)d.mc[MC_Q_V_D]; ArrP Pc = (ArrP)d.mc[MC_P_CONV];
    ArrP Eu = (ArrP)d.mc[MC_E_U]; ArrP Du = (ArrP)d.mc[MC_D_U];
    ArrP Uu = (ArrP)d.mc[MC_U_U]; ArrP Vu = (ArrP)d.mc[MC_V_U];
    ArrP Wu = (ArrP)d.mc[MC_W_U]; ArrP Ed = (ArrP)d.mc[MC_E_D];
    ArrP Dd = (ArrP)d.mc[MC_D_D]; ArrP Ud = (ArrP)d.mc[MC_U_D];
    ArrP Vd = (ArrP)d.mc[MC_V_D]; ArrP Wd = (ArrP)d.mc[MC_W_D];
    // layer heights by ro_load: a scalar load at the wave-uniform level of a level loop, a
    // global (not flat) load at a column's own level
    const double* __restrict__ h = d.hgt;
    const double t0 = K.t_0, Mb = K.M_bg, Rb = K.R_bg, lv = K.lv, ls = K.ls;
    auto id = [&](int i) { return (size_t)i * pl + col; };
    auto step = [&](int i) { return (i + 1 < im) ? ro_load(h + (i + 1)) - ro_load(h + (i)) : ro_load(h + (im - 1)) - ro_load(h + (im - 2)); };   // :306-309
    auto qs_env = [&](double E, double p, double qv) {                // :181-184
        return th_q_sat(E, p, th_M_nonwater(qv, K.co2_0, Mb));
    };
    auto t_of = [&](int i) { return (double)T[id(i)] * t0; };
    const bool inner = j >= 1 && j <= jm - 2 && k >= 1 && k <= km - 2;
    const double p_sfc = PS[id(0)];
    const double inv_a_u = 1.0 / (a_u * K.u_0), inv_a_d = 1.0 / (a_d * K.u_0), uu_cap = w_conv_max / K.u_0;

    // precompute (:336-376)
    // Every level loop below loads the next level's inputs before it stores this level's
    // results: a wave's vector-memory counter retires in issue order, so a load issued after a
    // store could only be waited for together with that store.
    double t_next = (double)T[id(0)];
    for (int i = 0; i < im; i++) {
        const size_t q = id(i);
        const double t_cur = t_next;
        if (i + 1 < im) t_next = (double)T[id(i + 1)];
        Sx[q] = cp_l * t_cur * t0 / s_0;
        Mu[q] = 0.0; Md[q] = 0.0; Su[q] = 0.0; Qvu[q] = 0.0; Qcu[q] = 0.0; Cu[q] = 0.0; El[q] = 0.0;
        Edv[q] = 0.0; Ep[q] = 0.0; Gp[q] = 0.0; Sd[q] = 0.0; Qvd[q] = 0.0; Pc[q] = 0.0;
    }
    int base = 0, lfs = 0, beg = -1, end = -1;
    double MdLFS = 0.0, cape_c = 0.0, dT = 0.0, dq = 0.0;
    // findCloudBaseLFS (:627-777)
    {   // the column's pressures, read once for both searches
        double pcol[4];
        for (int i0 = 0; i0 < im && beg < 0; i0 += 4) {
            for (int l = 0; l < 4; l++) pcol[l] = (i0 + l < im) ? PS[id(i0 + l)] : 0.0;
            for (int l = 0; l < 4 && i0 + l < im; l++)
                if (pcol[l] <= f_beg * p_sfc) { beg = i0 + l; break; }
        }
        if (beg >= 0) d.mc[MC_DEEP_BEG][id(beg)] = ro_load(h + (beg));
        for (int i0 = im - 1; i0 >= 0 && end < 0; i0 -= 4) {
            for (int l = 0; l < 4; l++) pcol[l] = (i0 - l >= 0) ? PS[id(i0 - l)] : 0.0;
            for (int l = 0; l < 4 && i0 - l >= 0; l++)
                if (pcol[l] >= f_end * p_sfc) { end = i0 - l; break; }
        }
        if (end >= 0) d.mc[MC_DEEP_END][id(end)] = ro_load(h + (end));
    }
    const G1<double>* __restrict__ QSL = (const G1<double>*)d.conv_qs[0];
    const G1<double>* __restrict__ QSR = (const G1<double>*)d.conv_qs[1];
    auto qs_lift = [&](int i) { return QSL[id(i)]; };
    if (beg >= 0 && end >= 0) {
        double tn = t_of(beg), qsn = qs_lift(beg), cn = (double)C[id(beg)], psn = PS[id(beg)];
        for (int i = beg; i < im; i++) {
            const size_t q = id(i);
            const double t_u = tn, qs = qsn, cv = cn, psv = psn;
            if (i + 1 < im) { tn = t_of(i + 1); qsn = qs_lift(i + 1); cn = (double)C[id(i + 1)]; psn = PS[id(i + 1)]; }
            const double w = 1.0 / (1.0 + exp(-(t_u - t_00) / 2.0));
            double qv = cv + q_v_u_add, qc = 0.0;
            if (qv >= scale * qs && t_u >= t_00) qc = qv - scale * qs;
            if (qv <= 0.0) qv = 0.0;
            if (qc <= 0.0) qc = 0.0;
            Qvu[q] = qv; Qcu[q] = qc;
            if (qv >= scale * qs && w > 0.01 && psv <= f_cb * p_sfc) {
                base = i;
                d.mc[MC_CLOUD_BASE][q] = ro_load(h + (i)) * w;
                break;
            }
        }
        if (base > 0) {   // level of neutral buoyancy of the theta-e parcel lifted from cloud base
            const size_t qb = id(base);
            const double Tb = t_of(base) + t_add_u;
            const double kap = th_R_of((double)C[qb], (double)CO2[qb], Rb) / th_cp_of((double)C[qb], (double)CO2[qb], Tb, Mb);
            const double Lb = (Tb >= t0) ? lv : ls;
            const double thp = Tb * pow(p_0 / PS[qb], kap) * exp(Lb * qs_lift(base) / (cp_l * Tb));
            bool buoy = false;
            int top = base;
            double cape = 0.0;
            double tn = 0.0, pn = 1.0, qn = 0.0;
            if (base + 1 < im) { tn = t_of(base + 1); pn = PS[id(base + 1)]; qn = qs_lift(base + 1); }
            for (int i = base + 1; i < im; i++) {
                const double Te = tn, pe = pn, qe = qn;
                if (i + 1 < im) { tn = t_of(i + 1); pn = PS[id(i + 1)]; qn = qs_lift(i + 1); }
                const double Le = (Te >= t0) ? lv : ls;
                const double the = Te * pow(p_0 / pe, kap) * exp(Le * qe / (cp_l * Te));
                if (thp - the > 0.0) {
                    buoy = true;
                    top = i;
                    cape += K.g * (thp - the) / the * (ro_load(h + (i)) - ro_load(h + (i - 1)));
                } else if (buoy) {
                    break;
                }
            }
            lfs = tmin(tmax(top, base + 1), im - 2);
            cape_c = tmax(0.0, cape);
            d.mc[MC_LFS][id(lfs)] = ro_load(h + (lfs));
        }
    }
    // surfaceFluxPerturbations, switch 1 (:402-485); Q_sensible_2D is zero in the reference
    if (inner && lfs > base) {
        const double Ts = t_of(0);
        const double rho = tmax((double)RH[id(0)], 1e-6);
        auto Es = [&](double Tx) { return (Tx >= t0) ? th_p_sat(Tx) : th_p_subl(Tx); };
        const double qa = qs_env(Es(Ts), p_sfc, K.c_0), qb = qs_env(Es(288.15), p_sfc, K.c_0);
        const double qcap = q_v_u_add * ((qb > 0.0) ? qa / qb : 1.0);
        if (PS[id(base)] < f_mid * p_sfc) {
            dT = t_add_u;
            dq = qcap;
        } else if (!(PS[id(base)] - PS[id(lfs)] >= f_diff * p_sfc)) {
            const double Hs = 0.0;
            const double E = tmax(0.0, d.evap[col]) / 86400.0;
            const double z = ro_load(h + (base));
            if (!(E <= 0.0 && Hs <= 0.0) && z > 0.0) {
                const double B = (K.g / Ts) * (Hs / (rho * cp_l) + 0.61 * Ts * E / rho);
                const double ws = cbrt(tmax(0.0, B * z));
                if (ws > 0.0) {
                    const double ir = 1.0 / (rho * ws);
                    dT = tmin(alpha_sfp * Hs * ir / cp_l, t_add_u);
                    dq = tmin(alpha_sfp * E * ir, qcap);
                }
            }
        }
    }
    // The inputs of E_u at a level of an inner column (the horizontal neighbours exist there).
    struct EnvL { double t, c, cip, cim, cjp, cjm, ckp, ckm, rh, u, v, w; };
    auto load_env = [&](int i) {
        const size_t q = id(i);
        EnvL e;
        e.t = t_of(i); e.c = (double)C[q];
        e.cip = (double)C[id(tmin(i + 1, im - 1))]; e.cim = (double)C[id(tmax(i - 1, 0))];
        e.cjp = (double)C[q + km]; e.cjm = (double)C[q - km]; e.ckp = (double)C[q + 1]; e.ckm = (double)C[q - 1];
        e.rh = (double)RH[q]; e.u = (double)U[q]; e.v = (double)V[q]; e.w = (double)W[q];
        return e;
    };
    // E_u at level i (:534-543, :841-851) with the entrainment gradients of c (:238-297); deep is
    // the cloud-depth test p(base) - p(lfs) >= f_diff p_sfc, mu the level's M_u
    auto entrain = [&](int i, const EnvL& v, double mu, bool with_m, bool deep) {
        double e = 0.0;
        if (deep) {
            const double rth = (K.r_earth * 1.0e3 + ro_load(h + (i))) * d.conv_cos[j];
            const double Ue = rth * 2.0 * M_PI;
            const double dr = (v.cip - v.cim) * (0.5 / step(i));
            const double dth = (v.cjp - v.cjm) * (0.5 / (Ue / (jm - 1)));
            const double dph = (v.ckp - v.ckm) * (1.0 / (Ue / (km - 1)));
            const double cc = tmax(v.c, 1e-4);
            e = -v.rh / cc * (v.u * dr + v.v * dth + v.w * dph);
            if (with_m) e += eps_u * mu;
        }
        if (v.t <= t_00_m2) e = 0.0;
        return e;
    };
    // a level of the updraft passes: E_u's inputs, the lifted and the resting saturation, p, and
    // the arrays the pass reads
    struct UpL { EnvL e; double qsr, qsl, ps, qc, mu; };
    auto load_up = [&](int i) {
        const size_t q = id(i);
        UpL u;
        u.e = load_env(i); u.qsr = QSR[q]; u.qsl = QSL[q]; u.ps = PS[q]; u.qc = Qcu[q]; u.mu = Mu[q];
        return u;
    };
    // initUpdraft's saturation at t + dT: the precomputed one where dT is 0 or t_add_u
    auto qs_pert_v = [&](const UpL& u, double t_a) {
        return (dT == 0.0) ? u.qsr : (dT == t_add_u) ? u.qsl : qs_env(th_p_auto(t_a), u.ps, u.e.c);
    };
    const double seed_w = c_mb * sqrt(2.0 * cape_c);
    // initUpdraft (:490-622)
    if (inner && beg != -1 && end != -1) {
        const bool deep = PS[id(base)] - PS[id(lfs)] >= f_diff * p_sfc;
        UpL nx;
        if (beg + 1 <= end) nx = load_up(beg + 1);
        const double mbeg = cap((double)RH[id(beg)] * seed_w, M_max);
        Mu[id(beg)] = mbeg;
        double mprev = mbeg;   // M_u one level down, carried in a register
        for (int i = beg + 1; i <= end; i++) {
            const size_t q = id(i);
            const UpL L_ = nx;
            if (i + 1 <= end) nx = load_up(i + 1);
            const double t_u = L_.e.t, t_a = t_u + dT;
            const double qsa = qs_pert_v(L_, t_a);
            double qv = L_.e.c + dq, qc = L_.qc;
            if (qv >= scale * qsa && t_u >= t_00) qc = qv - scale * qsa;
            if (qv <= 0.0) qv = 0.0;
            if (qc <= 0.0) qc = 0.0;
            Qvu[q] = qv; Qcu[q] = qc;
            const double e = entrain(i, L_.e, 0.0, false, deep);
            Eu[q] = e;
            Du[q] = 0.0;
            mprev = cap(mprev + e * step(i), M_max);
            Mu[q] = mprev;
            Sx[q] = cp_l * t_u / s_0;
            Su[q] = cp_l * t_a / s_0;
            Uu[q] = vel * L_.e.u; Vu[q] = vel * L_.e.v; Wu[q] = vel * L_.e.w;
        }
        if (end + 1 <= lfs - 1) nx = load_up(end + 1);
        MdLFS = cap(gam_d * mbeg, M_max);
        Md[id(lfs)] = MdLFS;
        mprev = mbeg;
        for (int i = end + 1; i <= lfs - 1; i++) {
            const size_t q = id(i);
            const UpL L_ = nx;
            if (i + 1 <= lfs - 1) nx = load_up(i + 1);
            const double t_u = L_.e.t, t_a = t_u + dT;
            const double qsa = qs_pert_v(L_, t_a);
            if (i == end + 1) Mu[id(i - 1)] = mbeg;
            Sx[q] = cp_l * t_u / s_0;
            Su[q] = cp_l * t_u / s_0;
            const double qv = L_.e.c + dq;
            Qvu[q] = qv;
            if (qv >= scale * qsa && t_a >= t_00) Qcu[q] = qv - scale * qsa;
            const double e = entrain(i, L_.e, 0.0, false, deep);
            Eu[q] = e;
            Du[q] = 0.0;
            const double m = cap(mprev + e * step(i), M_max);
            mprev = m;
            Mu[q] = m;
            Uu[q] = tmax(-uu_cap, tmin(L_.e.u + m * inv_a_u / tmax(L_.e.rh, 1e-6), uu_cap));
            Vu[q] = L_.e.v; Wu[q] = L_.e.w;
        }
    }
    // CloudBaseUpdate (:1462-1482) with simpsonCxEu (:1426-1455)
    auto cloud_base_update = [&]() {
        if (!inner || beg < 0 || lfs <= beg) return;
        int hit = -1;
        double cum = 0.0;
        int i = beg;
        auto f = [&](int l) { return -((double)C[id(l)] * Eu[id(l)]); };
        // f at i, i + 1, i + 2, the next pair read one step ahead
        double fa = f(i), fb = (i + 1 <= lfs) ? f(i + 1) : 0.0, fc = (i + 2 <= lfs) ? f(i + 2) : 0.0;
        for (; i + 2 <= lfs; i += 2) {
            const double h1 = step(i), h2 = step(i + 1);
            const double f0 = fa, f1 = fb, f2 = fc;
            const double nb = (i + 3 <= lfs) ? f(i + 3) : 0.0, nc = (i + 4 <= lfs) ? f(i + 4) : 0.0;
            const double pc = (h1 + h2) / 6.0 * ((2.0 - h2 / h1) * f0 + (h1 + h2) * (h1 + h2) / (h1 * h2) * f1
                                                 + (2.0 - h1 / h2) * f2);
            if (cum + pc > 0.0) { hit = (cum + 0.5 * h1 * (f0 + f1) > 0.0) ? i + 1 : i + 2; break; }
            cum += pc;
            fa = f2; fb = nb; fc = nc;
        }
        if (hit < 0 && i + 1 <= lfs && cum + 0.5 * step(i) * (fa + fb) > 0.0) hit = i + 1;
        if (hit >= 0 && hit < im && PS[id(hit)] <= f_cb * p_sfc && (base <= 0 || hit <= base)) {
            base = hit;
            d.mc[MC_CLOUD_BASE][id(hit)] = ro_load(h + (hit));
        }
    };
    cloud_base_update();
    for (int ip = 1; ip <= 3; ip++) {
        // updraftEntrainment (:780-908)
        if (inner && lfs > base) {
            const bool shallow = PS[id(base)] - PS[id(lfs)] < f_diff * p_sfc;   // convection_mode 1
            const double dic = shallow ? 1.0e9 : 500.0, hb = ro_load(h + (base));
            UpL nx = load_up(base + 1);
            for (int i = 0; i < base; i++) Mu[id(i)] = 0.0;
            double mprev = cap((double)RH[id(base)] * seed_w, M_max);
            Mu[id(base)] = mprev;
            Su[id(base)] = cp_l * (t_of(base) + dT) / s_0;
            Qvu[id(base)] = tmax((double)C[id(base)] + dq, 0.0);
            MdLFS = cap(gam_d * mprev, M_max);
            for (int i = base + 1; i <= lfs; i++) {
                const size_t q = id(i);
                const UpL L_ = nx;
                if (i + 1 <= lfs) nx = load_up(i + 1);
                const double t_u = L_.e.t, st = step(i), m0 = L_.mu;
                const double e = entrain(i, L_.e, m0, true, !shallow);
                double dd = eps_u * m0;
                if (ip > 1 && i == lfs) dd = del_u * m0 + b_u * m0 / st;
                if (ip > 1 && i == lfs - 1) dd = del_u * m0 + (1.0 - b_u) * m0 / st;
                if (t_u <= t_00_m2) dd = 0.0;
                Eu[q] = e;
                Du[q] = dd;
                mprev = cap(mprev + (e - dd) * st, M_max);
                Mu[q] = mprev;
                Cu[q] = 0.0;
                double el = 0.0;
                if (ip != 1) {
                    el = dd * L_.qc / tmax(L_.e.rh, 0.01);
                    if (el <= 0.0) el = 0.0;
                }
                El[q] = el;
                const double Kp = (ro_load(h + (i)) > hb + dic) ? bet_p : 0.0;
                Gp[q] = (t_u >= t0 && ip != 1) ? Kp * L_.qc : 0.0;
            }
            Mu[id(lfs)] = 0.0;
        }
        cloud_base_update();
        // updraftRecurrence (:1032-1163)
        if (inner && lfs > base) {
            // the level's fields and the arrays this pass reads
            struct RecL { double t, c, cl, rh, u, v, w, sx, ps, co2, eu, du, gp; };
            auto load_rec = [&](int i) {
                const size_t q = id(i);
                RecL r;
                r.t = t_of(i); r.c = (double)C[q]; r.cl = (double)CL[q]; r.rh = (double)RH[q];
                r.u = (double)U[q]; r.v = (double)V[q]; r.w = (double)W[q]; r.sx = Sx[q]; r.ps = PS[q];
                r.co2 = (double)CO2[q]; r.eu = Eu[q]; r.du = Du[q]; r.gp = Gp[q];
                return r;
            };
            RecL lm = load_rec(base), nx;   // lm: the level below
            if (base + 1 <= lfs - 1) nx = load_rec(base + 1);
            // the level below's updraft, carried in registers from one level to the next
            const size_t qb = id(base);
            double Mp = Mu[qb], Em = Eu[qb], Dm = Du[qb], Gm = Gp[qb];
            double Qvm = Qvu[qb], Qcm = Qcu[qb], Vm = Vu[qb], Wm = Wu[qb], Sm = Su[qb];
            for (int i = base + 1; i <= lfs - 1; i++) {
                const size_t q = id(i);
                const RecL L_ = nx;
                if (i + 1 <= lfs - 1) nx = load_rec(i + 1);
                const double t_u = lm.t, sp = step(i - 1);
                double m = cap(Mp + (Em - Dm) * sp, M_max);
                if (t_u <= t_00_m2) m = 0.0;
                const double n_qv = Mp * Qvm + sp * (Em * lm.c - Dm * Qvm);
                const double n_qc = Mp * Qcm + sp * (-Dm * lm.cl - lm.rh * Gm);
                const double n_v = Mp * Vm + sp * Em * lm.v;
                const double n_w = Mp * Wm + sp * Em * lm.w;
                const double n_s = (Mp * Sm + sp * (Em * lm.sx - Dm * Sm)) / s_0;
                double Ei = L_.eu, Di = L_.du, Gi = L_.gp;
                double qv, qc, vu, wu, su;
                if (fabs(m) > coeff_recurr) {
                    const double r = 1.0 / m;
                    qv = n_qv * r; qc = n_qc * r; vu = n_v * r; wu = n_w * r; su = n_s * r;
                } else {
                    su = L_.sx; qv = L_.c; qc = 0.0; vu = L_.v; wu = L_.w;
                }
                double cu = 0.0;
                if (fabs(m) > coeff_recurr) {   // the rising parcel's saturation adjustment, two passes
                    double Tu = su * s_0 / cp_l;
                    const double pu = L_.ps, cq = L_.c;
                    double tot = 0.0;
                    for (int it = 0; it < 2; ++it) {
                        const double ex = qv - qs_env(th_p_auto(Tu), pu, cq);
                        if (ex <= 0.0) break;
                        const double L = (Tu >= t0) ? lv : ls;
                        const double G = (L / cp_l) * th_dqdT(th_p_auto(Tu), th_L_auto(Tu), Tu, pu,
                                                              th_M_nonwater(cq, L_.co2, Mb));
                        const double dc = ex / (1.0 + G);
                        qv -= dc;
                        qc += dc;
                        Tu += (L / cp_l) * dc;
                        tot += dc;
                    }
                    su = cp_l * Tu / s_0;
                    cu = tot * m / (tmax(L_.rh, 0.01) * step(i));
                }
                double uu = tmax(-uu_cap, tmin(L_.u + m * inv_a_u / tmax(L_.rh, 1e-6), uu_cap));
                if (qv <= 0.0) qv = 0.0;
                if (qc <= 0.0) qc = 0.0;
                if (qc > q_c_u_max) qc = q_c_u_max;
                if (su <= 0.0) su = 0.0;
                if (t_u <= t_00) {
                    Eu[q] = 0.0; Du[q] = 0.0; m = 0.0; qv = 0.0; qc = 0.0; uu = 0.0; vu = 0.0; wu = 0.0;
                    su = 1.0; Gp[q] = 0.0; cu = 0.0; El[q] = 0.0;
                    Ei = 0.0; Di = 0.0; Gi = 0.0;
                }
                Mu[q] = m; Qvu[q] = qv; Qcu[q] = qc; Vu[q] = vu; Wu[q] = wu; Su[q] = su; Cu[q] = cu; Uu[q] = uu;
                Mp = m; Em = Ei; Dm = Di; Gm = Gi; Qvm = qv; Qcm = qc; Vm = vu; Wm = wu; Sm = su;
                lm = L_;
            }
        }
        // downdraftEntrainment (:913-1027)
        if (lfs > 0 && lfs > base) {
            struct DnL { double t, qsr, cl, md, rh, u, v, w, ep, pc; };
            auto load_dn = [&](int i) {
                const size_t q = id(i);
                DnL r;
                r.t = t_of(i); r.qsr = QSR[q]; r.cl = (double)CL[q]; r.md = Md[q]; r.rh = (double)RH[q];
                r.u = (double)U[q]; r.v = (double)V[q]; r.w = (double)W[q]; r.ep = Ep[q]; r.pc = Pc[q];
                return r;
            };
            DnL nx = load_dn(lfs - 1);
            const size_t ql = id(lfs);
            const double t_l = t_of(lfs), p_l = PS[ql];
            const double mdl = cap(MdLFS, M_max);
            Md[ql] = mdl;
            Ed[ql] = eps_d * fabs(mdl);
            Dd[ql] = del_d * fabs(mdl);
            Qvd[ql] = 0.5 * ((double)CL[ql] + scale * QSR[ql]);
            const double r_l = 1e2 * p_l / (th_R_of((double)C[ql], (double)CO2[ql], Rb) * t_l);
            Ud[ql] = (double)U[ql] + mdl * inv_a_d / tmax(r_l, 1e-6);
            Vd[ql] = (double)V[ql]; Wd[ql] = (double)W[ql]; Sd[ql] = Sx[ql];
            Pc[ql] = 0.0;
            double sig = 0.0;
            const double wcb = fabs(Wu[id(base)]) * K.u_0;
            if (wcb > 0.0) sig = fabs(Mu[id(base)]) / (tmax((double)RH[id(base)], 1e-6) * wcb);
            sig = tmin(sig, 1.0);
            double mnext = mdl;   // M_d one level up, carried in a register
            for (int i = lfs - 1; i >= 0; i--) {
                const size_t q = id(i);
                const DnL L_ = nx;
                if (i - 1 >= 0) nx = load_dn(i - 1);
                const double t_u = L_.t;
                Sd[q] = cp_l * t_u / s_0;
                const double qsd = L_.qsr;
                const double qvd = 0.5 * (L_.cl + scale * qsd);
                Qvd[q] = qvd;
                const double m0 = L_.md;
                const double e = eps_d * fabs(m0), dd = del_d * fabs(m0);
                Ed[q] = e; Dd[q] = dd;
                const double m = cap(mnext - (e - dd) * step(i), M_max);
                mnext = m;
                Md[q] = m;
                Ud[q] = L_.u + m * inv_a_d / tmax(L_.rh, 1e-6);
                Vd[q] = L_.v; Wd[q] = L_.w;
                if (i <= base) {
                    double ep = 0.0;
                    if (ip != 1) {
                        ep = L_.ep;
                        const double pr = L_.pc;
                        if (pr > 0.0 && sig > 0.0) ep = sig * alf_1 * (qsd - qvd) * sqrt(1e1 / alf_2 * pr / sig);
                        if (ep <= 0.0) ep = 0.0;
                    }
                    Ep[q] = ep;
                }
                double ed = 0.0;
                if (ip != 1) {
                    const double pr = L_.pc;
                    if (pr > 0.0) ed = a_ev * (1.0 + b_ev * pow(pr, 1.0 / 6.0)) * (qsd - qvd) * pow(pr, 4.0 / 9.0);
                    if (ed <= 0.0) ed = 0.0;
                }
                Edv[q] = ed;
            }
        }
        // downdraftRecurrence (:1168-1254), every column
        // the level above, carried in registers from one level to the next
        double Mp = 0, Ep1 = 0, Dp1 = 0, ev = 0, Qvp = 0, Vp = 0, Wp = 0, Sp = 0, Pp = 0;
        struct DrL { double t, rh, c, u, v, w, sx, ed, dd, edv, ep, gp; };
        auto load_dr = [&](int i) {
            const size_t q = id(i);
            DrL r;
            r.t = t_of(i); r.rh = (double)RH[q]; r.c = (double)C[q]; r.u = (double)U[q]; r.v = (double)V[q];
            r.w = (double)W[q]; r.sx = Sx[q]; r.ed = Ed[q]; r.dd = Dd[q]; r.edv = Edv[q]; r.ep = Ep[q]; r.gp = Gp[q];
            return r;
        };
        DrL up, nx;   // up: the level above
        if (lfs > 0) {
            nx = load_dr(lfs - 1);
            up = load_dr(lfs);
            const size_t ql = id(lfs);
            Mp = Md[ql]; Ep1 = Ed[ql]; Dp1 = Dd[ql]; ev = Edv[ql];
            Qvp = Qvd[ql]; Vp = Vd[ql]; Wp = Wd[ql]; Sp = Sd[ql]; Pp = Pc[ql];
        }
        for (int i = lfs - 1; i >= 0; i--) {
            const size_t q = id(i);
            const DrL L_ = nx;
            if (i - 1 >= 0) nx = load_dr(i - 1);
            const double t_u = L_.t, sp = step(i + 1), rp = up.rh;
            double Ei = L_.ed, Di = L_.dd;
            double m = cap(Mp - (Ei - Di) * sp, M_max);
            if (t_u <= t_00_m2) m = 0.0;
            const double L = (t_u >= t0) ? lv : ls;
            const double n_qv = Mp * Qvp - sp * (Ep1 * up.c - Dp1 * Qvp + rp * ev);
            const double n_v = Mp * Vp - sp * (Ep1 * up.v - Dp1 * Vp);
            const double n_w = Mp * Wp - sp * (Ep1 * up.w - Dp1 * Wp);
            const double n_s = (Mp * Sp - sp * (Ep1 * up.sx - Dp1 * Sp - L * rp * ev)) / s_0;
            double evi = L_.edv, epi = L_.ep;
            double qv, vd, wd, sd;
            if (fabs(m) > coeff_recurr) {
                const double r = 1.0 / m;
                qv = n_qv * r; vd = n_v * r; wd = n_w * r; sd = n_s * r;
            } else {
                qv = L_.c; vd = L_.v; wd = L_.w; sd = L_.sx;
            }
            double ud = L_.u + m * inv_a_d / tmax(L_.rh, 1e-6);
            if (qv <= 0.0) qv = 0.0;
            if (sd <= 0.0) sd = 0.0;
            if (t_u <= t_00) {
                Ed[q] = 0.0; Dd[q] = 0.0; m = 0.0; qv = 0.0; ud = 0.0; vd = 0.0; wd = 0.0; sd = 1.0;
                Edv[q] = 0.0; Ep[q] = 0.0;
                Ei = 0.0; Di = 0.0; evi = 0.0; epi = 0.0;
            }
            Md[q] = m; Qvd[q] = qv; Vd[q] = vd; Wd[q] = wd; Sd[q] = sd; Ud[q] = ud;
            if (i > 0) {
                Pp = tmax(0.0, Pp + step(i) * L_.rh * (L_.gp - evi - epi));
                Pc[q] = Pp;
            }
            Mp = m; Ep1 = Ei; Dp1 = Di; ev = evi; Qvp = qv; Vp = vd; Wp = wd; Sp = sd;
            up = L_;
        }
        Pc[id(0)] = Pc[id(1)];
    }
    // rhsForcing (:1304-1417). The caps of M_u and M_d over the column come first in the
    // reference; here each level is capped (and stored back) where its fluxes are formed, which
    // is before any flux reads it.
    ArrP MCt = (ArrP)d.mc[MC_MC_T]; ArrP MCq = (ArrP)d.mc[MC_MC_Q];
    ArrP MCv = (ArrP)d.mc[MC_MC_V]; ArrP MCw = (ArrP)d.mc[MC_MC_W];
    struct FxL { double mu, md, s, su, sd, c, qvu, qvd, v, vu, vd, w, wu, wd, t, rh, cu, edv, el, ep; };
    auto load_fx = [&](int i) {
        const size_t q = id(i);
        FxL x;
        x.mu = Mu[q]; x.md = Md[q]; x.s = Sx[q]; x.su = Su[q]; x.sd = Sd[q]; x.c = (double)C[q]; x.qvu = Qvu[q];
        x.qvd = Qvd[q]; x.v = (double)V[q]; x.vu = Vu[q]; x.vd = Vd[q]; x.w = (double)W[q]; x.wu = Wu[q];
        x.wd = Wd[q]; x.t = (double)T[q]; x.rh = (double)RH[q]; x.cu = Cu[q]; x.edv = Edv[q]; x.el = El[q];
        x.ep = Ep[q];
        return x;
    };
    auto fluxes = [&](int i, const FxL& x, double& fs, double& fq, double& fv, double& fw) {
        const size_t q = id(i);
        const double mu = cap(x.mu, M_max), md = cap(x.md, M_max);
        Mu[q] = mu; Md[q] = md;
        fs = mu * (x.su - x.s) + md * (x.sd - x.s);
        fq = mu * (x.qvu - x.c) + md * (x.qvd - x.c);
        fv = mu * (x.vu - x.v) + md * (x.vd - x.v);
        fw = mu * (x.wu - x.w) + md * (x.wd - x.w);
    };
    FxL a = load_fx(0), b = load_fx(1), c2;
    double fs0 = 0, fq0 = 0, fv0 = 0, fw0 = 0;
    fluxes(0, a, fs0, fq0, fv0, fw0);
    for (int i = 0; i < im - 1; i++) {
        const size_t q = id(i);
        if (i + 2 < im) c2 = load_fx(i + 2);
        double fs1, fq1, fv1, fw1;
        fluxes(i + 1, b, fs1, fq1, fv1, fw1);
        const double L = (a.t * t0 >= t0) ? lv : ls;
        const double isr = 1.0 / (step(i) * tmax(a.rh, 0.01));
        const double src = a.cu - a.edv - a.el - a.ep;
        MCt[q] = cap(-(fs1 - fs0) * isr * t0 + (L / cp_l) * src * t0, MCt_max);
        MCq[q] = cap(-(fq1 - fq0) * isr - src, MCq_max);
        MCv[q] = cap(-(fv1 - fv0) * isr * K.u_0, MCv_max);
        MCw[q] = cap(-(fw1 - fw0) * isr * K.u_0, MCv_max);
        fs0 = fs1; fq0 = fq1; fv0 = fv1; fw0 = fw1;
        a = b; b = c2;
    }
    // computeCAPE's profile of this column (the virtual-temperature excess of the lifted parcel
    // over its environment, integrated from the cloud base to the level of free sinking);
    // alf = (1 - eps) / eps with eps = 1 / (R_WaterVapour / R_Air) (:99-101, defaults 461.5 and
    // 296.8, cAtmosphereDefaults.cpp.inc:93-94)
    {
        const double rwra = 461.5 / 296.8, eps_thermo = 1.0 / rwra, alf = (1.0 - eps_thermo) / eps_thermo;
        G1<double>* __restrict__ prof = (G1<double>*)d.conv_tmp[0];
        auto term = [&](int i) {
            const size_t q = id(i);
            const double t_u = (double)T[q] * t0, t_u_add = t_u + t_add_u;
            const double t_vir = t_u_add * (1.0 + alf * Qvu[q] - Qcu[q]);
            const double t_vir_env = tmax(t_u * (1.0 + alf * (double)C[q] - (double)CL[q]), 1.0);
            const double exp_rm = 1.0 / (d.rad[i] + 1.0);
            return K.g * step(i) / exp_rm * (t_vir - t_vir_env) / t_vir_env;
        };
        double v = term(base);
        prof[id(base)] = v;
        for (int i = base; i < lfs && i < im - 1; i++) {
            v += term(i);
            prof[id(i + 1)] = v;
        }
        // the levels reached: base, and base + 1 .. min(lfs, im - 1) when lfs > base
        d.conv_cape_rng[2 * col] = base;
        d.conv_cape_rng[2 * col + 1] = (lfs > base) ? tmin(lfs, im - 1) : base;
    }
}

void launch_moist_convection(const Batch& b, int gate) {
    const dim3 g((unsigned)((b.max_jm * b.max_km + 63) / 64), 1, b.nmem);
    const dim3 gs((unsigned)((b.max_n + 255) / 256), 1, b.nmem);
    if (b.fp32) {
        hipLaunchKernelGGL(k_conv_sat<float>, gs, dim3(256), 0, b.stream, b.d_md, gate);
        hipLaunchKernelGGL(k_moist_convection<float>, g, dim3(64), 0, b.stream, b.d_md, gate);
    } else {
        hipLaunchKernelGGL(k_conv_sat<double>, gs, dim3(256), 0, b.stream, b.d_md, gate);
        hipLaunchKernelGGL(k_moist_convection<double>, g, dim3(64), 0, b.stream, b.d_md, gate);
    }
    hipLaunchKernelGGL(k_cape_select, dim3((unsigned)b.max_im, 1, b.nmem), dim3(256), 0, b.stream, b.d_md, gate);
}

}  // namespace athad
