/* CPU oracle for the channel-flow dual-time hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's arithmetic (ricklee209/Channel_Flow), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (channel_flow_amd, libchannel_flow.so) never links or calls it.
 *
 * Arrays use the reference layout: one FP64 array per variable,
 * [X_np][Y_m][Z_m], k fastest (Array.h), single rank: X_np = X_out+6,
 * Y_m = Y_out+4, Z_m = Z_out+4, interior i in [3,gend], j in [2,ny],
 * k in [2,nz] with gend = X_out+2, ny = Y_out+1, nz = Z_out+1.
 *
 * Parity status per function.  Pinned bit for bit by golden vectors that the
 * reference itself wrote (built from its own sources by oracle/ref_build.sh,
 * driven by oracle/ref_driver.cpp; fixtures in tests/golden/): cfo_metrics,
 * cfo_hyperplane_counts, cfo_halo_bc, cfo_flux, cfo_viscous, cfo_dual_time
 * (both sweep orders), cfo_rk3nodt (three stages), cfo_grid_points and
 * cfo_filter (at the reference's threshold, with the sensor ratios pinned
 * by threshold runs) on two non-cubic grids, an 18-step trajectory at
 * 16x12x16 and the first inner iteration at 64^3; cfo_flux_limiter on flat
 * and perturbed states except the y near-wall faces, where the reference
 * reads unassigned values (parity unpinned there).  cfo_dplusgs: its
 * right-hand side is golden-pinned; its Jacobi sweeps are pinned through
 * the golden-pinned LU-SGS sweeps, whose matrices tests/test_dplusgs.py
 * reads off cfo_sweeps_apply; the upper neighbours' iterate index (j+1,
 * k+1) is this restatement's choice where the reference is undefined.
 */
#ifndef CF_ORACLE_H
#define CF_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int X_out, Y_out, Z_out;     /* Resolution.h:5-7 */
  double deltaXI, deltaET, deltaZT; /* Resolution.h:17-19 */
  double high;                 /* prm.h:18 */
  double f_drive;              /* prm.h:13 */
  int x_periodic;              /* 0: single-rank self copy (main.cpp:207-238) */
  int n_sweeps;                /* Dual_time_stepping.cpp:514 (2) */
} cfo_grid;

/* Jacobian.cpp, Jacobian_u.cpp, Jacobian_v.cpp, Jacobian_w.cpp.
 * m[9] = xidx,xidy,xidz,etdx,etdy,etdz,ztdx,ztdy,ztdz (same for faces). */
void cfo_metrics(const cfo_grid* g, double gamma, double* J, double* const m[9], double* Ju,
                 double* const mu[9], double* Jv, double* const mv[9], double* Jw, double* const mw[9]);

/* Hyperplane.cpp:50-78: number of interior cells on plane m = i+j+k,
 * counts[m - 7] for m = 7 .. gend+ny+nz. */
int cfo_hyperplane_counts(const cfo_grid* g, int* counts, int ncounts);

/* main.cpp:156-240 (single rank) + Y_boundary_condition.cpp:48-120 +
 * Z_boundary_condition.cpp:48-66 */
void cfo_halo_bc(const cfo_grid* g, double* const U_[5], const double* J);

/* Flux_X.cpp / Flux_Y.cpp / Flux_Z.cpp (ROE == 1), dir 0,1,2 */
void cfo_flux(const cfo_grid* g, int dir, double* const U_[5], const double* J, const double* Jf,
              double* const mf[9], double* const inF[5]);

/* Flux_X_limiter.cpp:104-547, Flux_Y_limiter.cpp:106-685 and
 * Flux_Z_limiter.cpp:113-710, dir 0,1,2: limited MUSCL, mean-state
 * preconditioned dissipation times the sensor Ep (face layout, like inF).
 * Near the y walls the reference scales ML/MR 2..5 by a face J it never
 * assigns there; this restatement uses the face J of the interior branch. */
void cfo_flux_limiter(const cfo_grid* g, int dir, double* const U_[5], const double* J, const double* Jf,
                      double* const mf[9], const double* Ep, double* const inF[5]);

/* Grid.cpp:50-141: Y_point (stretching gamma; the reference hard-codes 2.8)
 * with its ghost planes, and X_point on the interior (may be NULL). */
void cfo_grid_points(const cfo_grid* g, double gamma, double* X_point, double* Y_point);

/* Filter.cpp:12-1720: ADM sensor.  Two chains of six separable filter
 * passes (h and 2h) give E_h, E_2h = |u - u_filtered|^2; EpX/EpY/EpZ on the
 * face layouts are E_high where (E_h + E_h')/(E_2h + E_2h') >= crit, else
 * E_low (a 0/0 ratio gives E_low). */
void cfo_filter(const cfo_grid* g, double crit, double E_high, double E_low, double* const U_[5],
                const double* Y_point, double* EpX, double* EpY, double* EpZ, const double* J);
/* the ratios the sensors threshold (face layouts) */
void cfo_filter_ratios(const cfo_grid* g, double* const U_[5], const double* Y_point, const double* J,
                       double* RX, double* RY, double* RZ);

/* DPLUSGS.cpp:189-1361: the RHS and D^-1 of Dual_time_stepping, then four
 * point-Jacobi sweeps NR = D^-1 (MR - L NL - U NL), NL <- NR, then the
 * same update and norms.  The reference reads the upper neighbours' NL at
 * indices it never assigns (UB); this restatement reads them at j+1, k+1,
 * where it reads their states and metrics.  The RHS, D^-1 and update are
 * those of cfo_dual_time. */
void cfo_dplusgs(const cfo_grid* g, double deltaT, double* const U_[5], double* const U[5], double* const Uq[5],
                 const double* J, double* const m[9], double* const inFx[5], double* const inFy[5],
                 double* const inFz[5], double* const vF[4], double* const MR[5], double* const NL[5],
                 double* const NR[5], double er[5]);

/* RK3NODT.cpp:115-183: stage RK (1..3) of the explicit SSP-RK3 update,
 * U_ <- c1 U + c2 U_ + c3 (dt Rk), c = (1, 0, 1), (3/4, 1/4, 1/4),
 * (1/3, 2/3, 2/3); MR <- dt Rk. */
void cfo_rk3nodt(const cfo_grid* g, int RK, double deltaT, double* const U_[5], double* const U[5], const double* J,
                 double* const inFx[5], double* const inFy[5], double* const inFz[5], double* const vF[4],
                 double* const MR[5]);

/* Driven_force.cpp:38-73: out = {Qall, ER_Q}, Qall the bulk flux
 * sum (U2_ J) 1/J / Jsum / X_out / Z_out in the reference's order. */
void cfo_driven_force(const cfo_grid* g, const double* U2_, const double* J, double Q0, double out[2]);

/* Statistic.cpp:1043-1071: adds the state's sums over i, k of U, V, W, UU,
 * VV, WW, UV to sums[q * Y_m + j] (q = 0..6), in the reference's order. */
void cfo_statistic_accum(const cfo_grid* g, double* const U_[5], double* sums);
void cfo_wall_shear(const cfo_grid* g, double* const U_[5], const double* Jv, double* const mv[9], double* out);

/* The implicit sweeps alone, on a given MR: ghosts of NL/NR zeroed,
 * NL = D^-1 MR, then mode 0 (LU-SGS, lexicographic), 1 (LU-SGS, plane by
 * plane) or 2 (DPLUSGS Jacobi).  A test hook: it lets the tests read the
 * sweeps as linear maps of MR. */
void cfo_sweeps_apply(const cfo_grid* g, double deltaT, int mode, double* const U_[5], const double* J,
                      double* const m[9], double* const MR[5], double* const NL[5], double* const NR[5]);

/* Viscous_terms.cpp:212-883; writes vF[0..3] = vF2..vF5 */
void cfo_viscous(const cfo_grid* g, double* const U_[5], const double* J, const double* Ju,
                 double* const mu[9], const double* Jv, double* const mv[9], const double* Jw,
                 double* const mw[9], double* const vF[4]);

/* Dual_time_stepping.cpp:195-1361: one inner iteration's implicit part.
 * Updates U_ in place, fills MR/NL/NR like the reference, er[5]. */
void cfo_dual_time(const cfo_grid* g, double deltaT, double* const U_[5], double* const U[5],
                   double* const Uq[5], const double* J, double* const m[9], double* const inFx[5],
                   double* const inFy[5], double* const inFz[5], double* const vF[4], double* const MR[5],
                   double* const NL[5], double* const NR[5], double er[5]);

/* Same, with the LU-SGS sweeps visited plane by plane (i+j+k = m ascending /
 * descending) instead of lexicographically: identical arithmetic, used to
 * show that the visiting order inside the dependency DAG does not matter. */
void cfo_dual_time_hyperplane(const cfo_grid* g, double deltaT, double* const U_[5], double* const U[5],
                              double* const Uq[5], const double* J, double* const m[9], double* const inFx[5],
                              double* const inFy[5], double* const inFz[5], double* const vF[4],
                              double* const MR[5], double* const NL[5], double* const NR[5], double er[5]);

#ifdef __cplusplus
}
#endif
#endif
