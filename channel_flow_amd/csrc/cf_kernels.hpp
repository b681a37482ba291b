// Launchers of the channel-flow kernels (host-callable, stream-ordered).
#pragma once
#include "cf_device.hpp"

namespace cf {

void launch_halo_bc(hipStream_t st, double* U, const double* prof, const Geo& g, int x_periodic);
void launch_flux(hipStream_t st, int dir, const double* U, const double* prof, double* F, const Geo& g);
// rebuilds the derived flux rows of the profile buffer from its metric rows
void launch_flux_tables(hipStream_t st, double* prof, const Geo& g);
// limiter variants; Ep is the sensor on the face layout of F
void launch_flux_limiter(hipStream_t st, int dir, const double* U, const double* Ep, const double* prof, double* F,
                         const Geo& g);
// ADM sensor: Yp = Y_point j-profile, W = FILTER_WORK_FIELDS-field workspace;
// the sensors are written on the face layouts
constexpr int FILTER_WORK_FIELDS = 22;
void launch_filter(hipStream_t st, const double* U, const double* prof, const double* Yp, double* W, double crit,
                   double E_high, double E_low, double* EpX, double* EpY, double* EpZ, const Geo& g);
// fused: one kernel with the primitives and face fluxes in LDS (default);
// otherwise primitives, face fluxes and divergence as three kernels through
// the staging fields Pr (4) and VS (12)
void launch_viscous(hipStream_t st, const double* U, const double* prof, double* Pr, double* VS, double* vF,
                    const Geo& g, bool fused = true);
// RK3NODT stage RK (1..3) on the interior; vF and MR may be null
void launch_rk3(hipStream_t st, int RK, double* Ui, const double* Un, const double* Fx, const double* Fy,
                const double* Fz, const double* vF, const double* prof, double* MR, const Geo& g, double deltaT,
                double f_drive);
// adds the per-j sums of U, V, W, UU, VV, WW, UV of the interior to sums[7][Ym]
void launch_stat(hipStream_t st, const double* U, double* sums, const Geo& g);
// per-i-plane sums of the bulk flux integrand of Driven_force into part[X_out]
void launch_bulk_flux(hipStream_t st, const double* U2, const double* prof, double Jsum, double* part,
                      const Geo& g);
// the wall values of Statistic's y-face viscous x-momentum flux (the terms
// of Shear_stress.dat): out[2][X_out][Z_out], bottom wall then top
void launch_wall_shear(hipStream_t st, const double* U, const double* prof, const Geo& g, double* out);
// out[0] = a[0] + a[1] + ... + a[n-1], added in index order by one thread
void launch_sum_in_order(hipStream_t st, const double* a, int n, double* out);
void launch_rhs(hipStream_t st, const double* Ui, const double* Un, const double* Uq, const double* Fx,
                const double* Fy, const double* Fz, const double* vF, const double* prof, double* MR, double* NL,
                const Geo& g, double deltaT, double f_drive);
int sweep_tiles(const Geo& g, int* nTI, int* nTJ);
long sweep_field_elems(const Geo& g);
long sweep_update_blocks(const Geo& g);
// inputs of the RHS, for the sweep pre-pass that forms MR itself (rin)
struct RhsIn {
  const double *Un, *Uq, *Fx, *Fy, *Fz, *vF;
  double f_drive;
};
// Persistent tiled sweeps (diagonal metrics): the summary pre-pass (forming MR
// with the RHS when rin is given, else reading MR), n_sweeps forward/backward
// pairs, then the primitive update of Ui and er.  Sbuf: G_N + 15 fields of
// sweep_field_elems doubles.  marks (optional, 1 + 3 * n_sweeps events) are
// recorded after the pre-pass, then per sweep after the sync reset, the
// forward and the backward sweep.  fp32: the sweep-order fields, D^-1, the
// split Jacobians and the sweeps in float (pipelined schedule only); the RHS
// and the update stay FP64.
void launch_sweeps(hipStream_t st, double* Ui, const double* MR, double* Sbuf, const double* prof, const Geo& g,
                   double deltaT, int n_sweeps, const int* order, int* sync, double* partial, double* er_dev,
                   double DN, long spin_limit, const hipEvent_t* marks = nullptr, bool pipelined = true,
                   const RhsIn* rin = nullptr, bool fp32 = false);
// NL and NR of the last sweep back in the natural layout
void launch_sweep_outputs(hipStream_t st, const double* Sbuf, double* NL, double* NR, const Geo& g, bool fp32 = false);
// the forward sweep's cross-tile NL slots in Sbuf set to the sentinel
void launch_seed_hand_off(hipStream_t st, double* Sbuf, const Geo& g, bool fp32);

void launch_sweeps_planes(hipStream_t st, const double* Ui, const double* MR, double* NL, double* NR,
                          const double* prof, const Geo& g, double deltaT, int n_sweeps,
                          const hipEvent_t* marks = nullptr);
// DPLUSGS: four point-Jacobi sweeps from NL (= D^-1 MR); NL and NR both end
// with the last sweep's values (ghosts of both must be zero)
void launch_dplusgs(hipStream_t st, const double* Ui, const double* MR, double* NL, double* NR, const double* prof,
                    const Geo& g, double deltaT);
void launch_copy_interior(hipStream_t st, const double* src, double* dst, const Geo& g);
// U_ <- 2 U^n - U^{n-1} (order 1), 3 U^n - 3 U^{n-1} + U^{n-2} (order 2) or
// 4 U^n - 6 U^{n-1} + 4 U^{n-2} - U^{n-3} (order 3, U3 = U^{n-3}), interior
void launch_extrapolate(hipStream_t st, double* Ui, const double* Un, const double* Uq, const double* Uqq,
                        const double* U3, const Geo& g, int order);
long update_blocks(const Geo& g);
void launch_update(hipStream_t st, double* Ui, const double* NR, const double* prof, double* partial, double* er_dev,
                   const Geo& g, double DN);

}  // namespace cf
