/* channel_flow.h -- C ABI of the MI355X-native channel-flow dual-time path.
 *
 * Drop-in boundary for the time-step hot path of ricklee209/Channel_Flow.
 * The reference exposes free C++ functions (main.h:10-1069) that take raw
 * `double (*)[Y_m][Z_m]` arrays sized by compile-time macros and are driven by
 * main.cpp:143-498.  This header replaces them with:
 *
 *   - a context that owns one GPU's device memory for a runtime grid;
 *   - per-function entry points that take reference-layout host arrays
 *     ([X_np][Y_m][Z_m] FP64, k fastest; single rank X_np = X_out+6,
 *     Y_m = Y_out+4, Z_m = Z_out+4) with the same meaning and argument order
 *     as the reference function they replace (the int myid and the MUSCL
 *     staging arrays ML/MR are dropped: one context holds the whole domain,
 *     and the staging lives on the device);
 *   - a device-resident path (state stays in HBM between calls).
 *
 * Conventions: every function returns CF_OK (0) or a negative cf_status;
 * cf_last_error() gives the message.  The caller owns host buffers, the
 * context owns device buffers.  A context is not thread-safe; all work runs on
 * the context's one HIP stream and every host-array entry point is
 * synchronous.  No C++ exception crosses this boundary.
 *
 * Host arrays and the device-resident state: an entry point that takes the
 * state as host arrays (U_, U, Uq, U2_) works on stand-in device fields when
 * the context also holds a resident state (cf_load_state / cf_load_initial /
 * cf_load_checkpoint), so that state comes through unchanged.  The metrics
 * such an entry point receives do replace the context's (they are what the
 * next resident iteration uses).
 *
 * Geometry: the reference metrics (Jacobian*.cpp) vary only with j.  Entry
 * points that receive 3-D metric arrays reduce them to j-profiles and return
 * CF_ERR_METRICS if a metric varies along i or k inside the region it is used.
 */
#ifndef CHANNEL_FLOW_H
#define CHANNEL_FLOW_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cf_ctx cf_ctx;

enum cf_status {
  CF_OK = 0,
  CF_ERR_ARG = -1,        /* bad argument (null pointer, bad size) */
  CF_ERR_HIP = -2,        /* a HIP runtime call failed */
  CF_ERR_NONFINITE = -3,  /* residual norm is NaN/Inf */
  CF_ERR_METRICS = -4,    /* metric arrays are not functions of j only */
  CF_ERR_IO = -5,         /* file missing / short */
  CF_ERR_STATE = -6,      /* call order (e.g. no metrics or state loaded) */
  CF_ERR_TIMEOUT = -7     /* a device-side dependency wait exceeded its bound */
};

enum cf_precision {
  CF_FP64 = 0,            /* everything FP64, as the reference */
  CF_FP32_SWEEPS = 1      /* LU-SGS sweeps and split Jacobians in FP32, inside
                             the FP64 loop: the RHS, D^-1's inputs and the
                             update stay FP64, so the converged state is the
                             FP64 one.  Needs diagonal metric rows and sweep
                             schedule 0; other schedules are refused
                             (CF_ERR_ARG).  DPLUSGS stays FP64.  No reference
                             counterpart (the reference is FP64 only). */
};

/* Runtime replacement of Resolution.h / prm.h / main.cpp:47-57 constants. */
typedef struct {
  int X_out, Y_out, Z_out;          /* Resolution.h:5-7 */
  double deltaXI, deltaET, deltaZT; /* Resolution.h:17-19 */
  double high;                      /* prm.h:18 (channel height) */
  double f_drive;                   /* prm.h:13 (constant driving force) */
  double deltaT;                    /* main.cpp:53 */
  int x_periodic;                   /* 0: single-rank zero-gradient x ghosts
                                       (main.cpp:207-238 with np=1);
                                       1: true periodic x */
  int n_sweeps;                     /* Dual_time_stepping.cpp:514 (2) */
  int iteration_end_step;           /* main.cpp:48 (200) */
  double tol;                       /* main.cpp:406 (0.001) */
  int precision;                    /* enum cf_precision */
} cf_params;

/* --- context ------------------------------------------------------------ */
int cf_create(const cf_params* p, cf_ctx** out);
void cf_destroy(cf_ctx* ctx);
const char* cf_last_error(const cf_ctx* ctx);
/* number of FP64 elements of one reference-layout array (X_np*Y_m*Z_m) */
long long cf_field_elems(const cf_ctx* ctx);

/* --- geometry ------------------------------------------------------------ */
/* Jacobian/Jacobian_u/_v/_w (main.h:51-160, Jacobian*.cpp): computes the
 * cell and face metrics for stretching factor gamma on the host, keeps the
 * j-profiles on the device, and optionally returns the reference-layout
 * arrays (any output pointer may be NULL).  m[9] order: xidx,xidy,xidz,
 * etdx,etdy,etdz,ztdx,ztdy,ztdz. */
int cf_jacobian_all(cf_ctx* ctx, double gamma, double* J, double* const m[9], double* J_u,
                    double* const m_u[9], double* J_v, double* const m_v[9], double* J_w,
                    double* const m_w[9]);
/* Installs metrics given in reference layout (reduces them to j-profiles). */
int cf_set_metrics(cf_ctx* ctx, const double* J, const double* const m[9], const double* J_u,
                   const double* const m_u[9], const double* J_v, const double* const m_v[9],
                   const double* J_w, const double* const m_w[9]);
/* Hyperplane (main.h:974-983, Hyperplane.cpp:50-78): per-plane cell counts
 * for m = 7 .. gend+ny+nz; counts[m-7].  Returns the number of planes in
 * *nplanes. */
int cf_hyperplane(cf_ctx* ctx, int* counts, int capacity, int* nplanes);

/* --- per-function entry points (reference-layout host arrays) ------------ */
/* main.cpp:156-250: single-rank x halo, then Y_boundary_condition
 * (main.h:192-207) and Z_boundary_condition (main.h:209-224), in place. */
int cf_boundary_conditions(cf_ctx* ctx, double* const U_[5], const double* J);
/* Flux_X / Flux_Y / Flux_Z (main.h:226-374): 5th-order MUSCL + Roe. */
int cf_flux_x(cf_ctx* ctx, const double* const U_[5], const double* J, const double* J_u,
              const double* const m_u[9], double* const inFx[5]);
int cf_flux_y(cf_ctx* ctx, const double* const U_[5], const double* J, const double* J_v,
              const double* const m_v[9], double* const inFy[5]);
int cf_flux_z(cf_ctx* ctx, const double* const U_[5], const double* J, const double* J_w,
              const double* const m_w[9], double* const inFz[5]);
/* Flux_X/Y/Z_limiter (main.h:376-522; Flux_X_limiter.cpp:104-547,
 * Flux_Y_limiter.cpp:106-685, Flux_Z_limiter.cpp:113-710): limited MUSCL + mean-state preconditioned
 * dissipation scaled by the sensor EpX/EpY/EpZ (face layout of inF, as the
 * reference's).  Near the y walls (faces 1, 2, ny-2, ny-1, ny) the reference
 * multiplies ML/MR 2..5 by a face Jacobian its near-wall branch never
 * assigns; these entry points use the face Jacobian of its interior branch. */
int cf_flux_x_limiter(cf_ctx* ctx, const double* const U_[5], const double* J, const double* J_u,
                      const double* const m_u[9], double* const inFx[5], const double* EpX);
int cf_flux_y_limiter(cf_ctx* ctx, const double* const U_[5], const double* J, const double* J_v,
                      const double* const m_v[9], double* const inFy[5], const double* EpY);
int cf_flux_z_limiter(cf_ctx* ctx, const double* const U_[5], const double* J, const double* J_w,
                      const double* const m_w[9], double* const inFz[5], const double* EpZ);
/* Grid (main.h:10-20, Grid.cpp:50-141): Y_point of the tanh-stretched grid
 * for stretching gamma (the reference hard-codes 2.8) with its ghost planes,
 * and X_point on the interior; either pointer may be NULL.  Also keeps the
 * Y_point profile for cf_filter. */
int cf_grid(cf_ctx* ctx, double gamma, double* X_point, double* Y_point);
/* Filter (main.h:946-972, Filter.cpp:12-1720): the ADM sensor.  EpX/EpY/EpZ
 * become E_high where the ratio of the filtered-velocity energies at scales
 * h and 2h is >= Roe_criterion, else E_low, on the face layouts of inFx/y/z;
 * elements outside those ranges keep their values. */
int cf_filter(cf_ctx* ctx, double Roe_criterion, double E_high, double E_low, const double* const U_[5],
              const double* Y_point, double* EpX, double* EpY, double* EpZ, const double* J);
/* Viscous_terms (main.h:525-651): writes vF[0..3] = vF2..vF5. */
int cf_viscous_terms(cf_ctx* ctx, const double* const U_[5], double* const vF[4], const double* J,
                     const double* const m[9], const double* J_u, const double* const m_u[9],
                     const double* J_v, const double* const m_v[9], const double* J_w,
                     const double* const m_w[9]);
/* Dual_time_stepping (main.h:653-764): one inner iteration's implicit part;
 * U_ updated in place, er[5] = er1..er5, MR/NR/NL as the reference leaves
 * them (NL = NR on exit).  Any of MR/NR/NL may be NULL. */
int cf_dual_time_stepping(cf_ctx* ctx, double deltaT, double er[5], double* const U_[5],
                          const double* const U[5], const double* const Uq[5], const double* J,
                          const double* const m[9], const double* const inFx[5],
                          const double* const inFy[5], const double* const inFz[5],
                          const double* const vF[4], double* const MR[5], double* const NR[5],
                          double* const NL[5]);

/* DPLUSGS (main.h:766-870, DPLUSGS.cpp:189-1361): the RHS and D^-1 of
 * Dual_time_stepping, then four point-Jacobi sweeps NR = D^-1 (MR - L NL -
 * U NL), NL <- NR, then the same update and norms; same arguments as
 * cf_dual_time_stepping.  The reference reads the upper neighbours' NL at
 * indices j_, k_ it never assigns (undefined behaviour); this entry point
 * reads them at j+1, k+1, where the reference reads their states. */
int cf_dplusgs(cf_ctx* ctx, double deltaT, double er[5], double* const U_[5], const double* const U[5],
               const double* const Uq[5], const double* J, const double* const m[9], const double* const inFx[5],
               const double* const inFy[5], const double* const inFz[5], const double* const vF[4],
               double* const MR[5], double* const NR[5], double* const NL[5]);

/* --- device-resident path ------------------------------------------------ */
/* Loads level n and n-1 from interior Q arrays ([X_out][Y_out][Z_out] per
 * variable, rho, rho u, rho v, rho w, rho E; Initial_condition.cpp:87-116)
 * and sets the inner iterate to level n (main.cpp:112-127). */
int cf_load_state(cf_ctx* ctx, const double* const Qn[5], const double* const Qnm1[5]);
/* Initial_condition (main.h:22-49): reads an initial.bin / LES%05d.bin file. */
int cf_load_initial(cf_ctx* ctx, const char* path);
/* Optional: overrides the inner iterate U_ with interior Q arrays. */
int cf_set_iterate(cf_ctx* ctx, const double* const Q[5]);
/* Interior Q of level n, n-1 and of the inner iterate (NULL to skip). */
int cf_read_state(cf_ctx* ctx, double* const Qn[5], double* const Qnm1[5], double* const Qiter[5]);
/* One inner iteration: halo + BC, Flux_X/Y/Z, Viscous_terms,
 * Dual_time_stepping (main.cpp:156-353 with the dual-time calls enabled).
 * er[5] are the raw norms e_k = sqrt(sum dq^4)/N. */
int cf_inner_iteration(cf_ctx* ctx, double er[5]);
/* One physical step: inner iterations until er1..er4 (normalised by their
 * running maxima, main.cpp:387-397) fall below tol or iteration_end_step is
 * reached, then U^{n-1} <- U^n <- U_ (main.cpp:444-464).  Returns the
 * iteration count and the last normalised er[5]. */
int cf_physical_step(cf_ctx* ctx, int* iterations, double er[5]);
/* The time-level shift alone: U^{n-1} <- U^n, U^n <- U_ on the interior
 * (main.cpp:444-464); cf_physical_step calls it on convergence.  While the
 * quadratic initial guess is selected, U^{n-2} <- U^{n-1} first. */
int cf_advance_levels(cf_ctx* ctx);
/* Initial iterate of each cf_physical_step, no counterpart in the reference
 * (which starts every step at U^n, the level shift of main.cpp:444-464):
 * 0 (default) U^n, as the reference; 1 the linear extrapolation
 * 2 U^n - U^{n-1}; 2 the quadratic 3 U^n - 3 U^{n-1} + U^{n-2}, linear until
 * three levels are known (it keeps a U^{n-2} level, 5 more fields); 3 the
 * learned predictor attached with cf_set_predictor (below).  The
 * converged state is the same to the convergence tolerance; the number of
 * inner iterations under the running-max test (main.cpp:387-406) drops. */
int cf_set_initial_guess(cf_ctx* ctx, int order);
/* Highest order of the extrapolated initial iterate, for initial guess 2 and
 * for the learned predictor's fallback; no counterpart in the reference:
 * 2 (default) quadratic; 3 cubic 4 U^n - 6 U^{n-1} + 4 U^{n-2} - U^{n-3}
 * once four levels are known (it keeps a U^{n-3} level, cf_keep_levels(ctx,
 * 4)), quadratic before. */
int cf_set_extrapolation_order(cf_ctx* ctx, int order);
/* Output (main.h:872-896, Output.cpp:56-190): writes <dir>/LES%05d.bin. */
int cf_write_output(cf_ctx* ctx, int step, const char* dir);
/* Exact restart, no counterpart in the reference (whose restart reads
 * LES%05d.bin, Q = U J, and starts E_k and the statistics afresh): the
 * interior of the iterate and of levels n and n-1 as stored (U = Q/J), the
 * running maxima E_k of cf_physical_step and the statistic sums.  Format:
 * "CFCKPT02", int32 X_out Y_out Z_out M, double E[5], double sums[7][Y_m],
 * double wall[4] (the Shear_stress.dat running sums and previous averages),
 * then 15 fields (U_, U, Uq) of X_out*Y_out*Z_out doubles, [x][y][z], then M
 * more: 5 (level n-2) when the quadratic initial guess holds one, 10 (levels
 * n-2 and n-3) when the cubic one holds both, else 0.
 * cf_load_checkpoint also reads "CFCKPT01" files, which have no wall[4]. */
int cf_save_checkpoint(cf_ctx* ctx, const char* path);
int cf_load_checkpoint(cf_ctx* ctx, const char* path);
/* Runs one inner iteration like cf_inner_iteration and times its stages
 * with HIP events recorded on the context's stream.  stage_ms: ghosts,
 * Flux_X, Flux_Y, Flux_Z, Viscous_terms, RHS, sweep summary pre-pass, sweep
 * sync reset, forward sweeps, backward sweeps, update+norms (summed over the
 * n_sweeps sweeps); launches: kernel launches per stage.  No counterpart in
 * the reference (it has no timers, MPI_prm.h:14). */
#define CF_PROFILE_STAGES 11
int cf_profile_inner_iteration(cf_ctx* ctx, double er[5], double stage_ms[CF_PROFILE_STAGES],
                               int launches[CF_PROFILE_STAGES]);
/* Implicit part of cf_inner_iteration / cf_physical_step: 0 (default)
 * Dual_time_stepping's LU-SGS sweeps, 1 DPLUSGS's Jacobi sweeps. */
int cf_set_implicit_solver(cf_ctx* ctx, int solver);
/* Times the variant kernels on the resident iterate (after its ghosts are
 * set) with HIP events on the context's stream: ms[0] Filter (sensor with
 * E_high = 1, E_low = 0 into the device sensor arrays; cf_grid must have
 * set Y_point), ms[1..3] Flux_X/Y/Z_limiter with those sensors.  No
 * counterpart in the reference. */
#define CF_VARIANT_STAGES 4
int cf_profile_variants(cf_ctx* ctx, double Roe_criterion, double ms[CF_VARIANT_STAGES]);
/* Visiting order of the LU-SGS sweeps; all give identical bits because each
 * cell depends only on its three upstream neighbours.  schedule 0 (default):
 * one persistent launch per sweep direction, 8x8-column tiles walking k in
 * dependency order, each tile one compute wave fed by three loader waves;
 * schedule 1: one launch per hyperplane i+j+k = m (the reference order,
 * Hyperplane.cpp:50-78); schedule 2: the tiles of schedule 0 as single waves
 * that load their own inputs.  A CF_FP32_SWEEPS context accepts schedule 0
 * only. */
int cf_set_sweep_schedule(cf_ctx* ctx, int schedule);
/* RK3NODT.cpp:115-183 (main.h RK3NODT; the explicit path of the shipped
 * main.cpp): stage RK (1, 2, 3) of the SSP-RK3 update
 * U_ <- c1 U + c2 U_ + c3 dt Rk with (c1, c2, c3) = (1, 0, 1), (3/4, 1/4,
 * 1/4), (1/3, 2/3, 2/3), Rk the flux divergence + vF + driving force, on
 * reference-layout host arrays.  MR receives dt Rk (may be NULL); er, Uq
 * and the metrics other than J are accepted for the reference's parameter
 * list and not used (RK3NODT.cpp does not touch er1..er5). */
int cf_rk3nodt(cf_ctx* ctx, int RK, double deltaT, double er[5], double* const U_[5], const double* const U[5],
               const double* const Uq[5], const double* J, const double* const m[9], const double* const inFx[5],
               const double* const inFy[5], const double* const inFz[5], const double* const vF[4],
               double* const MR[5]);
/* One pass of main.cpp:152-377 on the device state with the RK3NODT path:
 * three stages of ghosts, Flux_X/Y/Z, Viscous_terms (viscous != 0; with 0
 * the viscous terms are zero, as in the shipped main.cpp, where the call is
 * commented out and vF is never assigned) and RK3NODT, from the iterate
 * with level n as the base.  cf_advance_levels then shifts the levels. */
int cf_rk3_step(cf_ctx* ctx, int viscous);
/* Driven_force.cpp:38-73 (main.h Driven_force): the bulk flux
 * Qall = sum over the interior of (U2_ J) 1/J / Jsum / X_out / Z_out with
 * Jsum = sum_j 1/J, and ER_Q = |Q0 - Qall| / Q0 (the reference prints both
 * and takes Q0 from prm.h:15, 20.464969232814912).  U2_: reference-layout
 * host array, or NULL for the device iterate; J: the cell J, or NULL for the
 * context's.  The sum is formed per i-plane in a fixed tree and the planes
 * are added in order: Qall equals the reference's to rounding. */
int cf_driven_force(cf_ctx* ctx, double Q0, const double* U2_, const double* J, double* Qall, double* ER_Q);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The bulk flux
 * of cf_driven_force with the context's metrics, and no device allocation
 * of any kind: the per-plane partial sums and their sum live in a
 * file-scope device array of CF_NOALLOC_MAX_PLANES + 32 doubles that each
 * GPU holds until the process ends.  U2_dev: a DEVICE pointer to the rho u
 * field in the reference layout [X_np][Y_m][Z_m], or NULL for the resident
 * iterate.  The results are the same bits as cf_driven_force's.  X_out
 * above CF_NOALLOC_MAX_PLANES is refused (CF_ERR_ARG).  All calls on one GPU
 * share that array: calls that overlap on one GPU (other threads, other
 * contexts' streams) race. */
#define CF_NOALLOC_MAX_PLANES 512
int cf_driven_force_noalloc(cf_ctx* ctx, double Q0, const double* U2_dev, double* Qall, double* ER_Q);
/* Statistic.cpp:552-1154 (main.h Statistic): the mean profiles and the
 * wall shear.  Adds the per-j sums over i and k of U, V, W, UU, VV, WW, UV
 * (U = U2_/U1_ ...) of the state to running sums kept in the context; adds
 * the bottom and top wall values of the y-face viscous x-momentum flux
 * (Statistic.cpp:552-828) to running sums in the reference's order and
 * appends to <directory>/Shear_stress.dat the change of their running
 * averages, "%f	%f
" (Statistic.cpp:893-911); when step % statistic_step
 * == 0 writes <directory>/st%05d.dat (the sums / statistic_step / X_out /
 * Z_out, seven blocks of one "%.16f" line per j = 2..ny, the reference's
 * format) and resets all the sums.  U_: reference-layout host arrays (ghost
 * cells included), or NULL for the device iterate (with the ghost cells of
 * its last inner iteration, as in the reference).  The per-j profile sums are
 * combined in a fixed tree: the means equal the reference's to rounding; the
 * wall sums are added in the reference's order, so Shear_stress.dat is the
 * reference's byte for byte.  ND%05d.dat is not produced. */
int cf_statistic(cf_ctx* ctx, int step, int statistic_step, const double* const U_[5], const char* directory);
/* Kernel schedule of Viscous_terms (Viscous_terms.cpp:212-883); both give
 * identical bits.  0 (default): one kernel that keeps the primitives and
 * face fluxes of a tile in LDS; 1: primitives, face fluxes and divergence as
 * three kernels through 16 staging fields in HBM (allocated here on first
 * selection, like the reference's staging arrays). */
int cf_set_viscous_schedule(cf_ctx* ctx, int schedule);
/* Synchronises the context's stream. */
int cf_synchronize(cf_ctx* ctx);
/* LU-SGS sweep pairs of Dual_time_stepping.cpp:514-1277, no counterpart in
 * the reference's parameters.  With enabled != 0 (the default) one
 * forward/backward pair runs for any n_sweeps >= 1: in one domain each later
 * pair reads only MR and values its own pass has formed, so it reproduces
 * the first bit for bit (MR, NL, NR, U_ and er are the same; across MPI
 * ranks they would not be, :521-559).  0 runs all n_sweeps pairs. */
int cf_set_sweep_reuse(cf_ctx* ctx, int enabled);
/* Polls a tiled sweep's halo wave makes for one value from an upstream tile
 * before the sweep gives up (default 4194304; polls >= 1).  A sweep that gives
 * up runs to its end without that value; the iteration then returns
 * CF_ERR_TIMEOUT and its iterate is not usable: reload the state.  No
 * counterpart in the reference, whose ranks wait in MPI receives without a
 * bound. */
int cf_set_spin_limit(cf_ctx* ctx, long long polls);
/* With enabled != 0, cf_inner_iteration (and so cf_physical_step) replays
 * the kernels of one inner iteration as a hipGraph on the context's stream,
 * captured on first use and again when the fields, schedules, precision or
 * parameters it was captured with change; the results are the same bits as
 * the launches issued one by one.  No counterpart in the reference. */
int cf_set_graph(cf_ctx* ctx, int enabled);
/* Number of graph captures made so far. */
int cf_graph_builds(cf_ctx* ctx, long long* builds);
/* With enabled != 0 (the default) the inner iteration runs Flux_X, Flux_Y,
 * Flux_Z and Viscous_terms, which all read the state after its ghost cells
 * and write disjoint outputs, concurrently on the context's stream and three
 * helper streams of default priority (fork and join by events, also inside
 * the graph); 0 runs them one after another.  Same bits either way. */
int cf_set_concurrent_stencils(cf_ctx* ctx, int enabled);

/* --- learned initial iterate ------------------------------------------------
 * No counterpart in the reference, whose physical step starts its inner
 * iterations at U^n (main.cpp:444-464).  A small network predicts
 * Q^{n+1} - Q^n from the last three states:
 *   x'  = ((Q^n - mu) / sig, (Q^n - Q^{n-1}) / s, (Q^{n-1} - Q^{n-2}) / s, 1)
 *   h1  = W1 x'                          pointwise 16 -> 64
 *   z2  = b2 + 27-point conv(W2, h1)      3x3x3, 64 -> 64, zeros outside the domain
 *   y   = b3 + W3 gelu(z2)               pointwise 64 -> 5
 *   Q^{n+1} = Q^n + c1 d1 + c2 d2 + sr y  (d1, d2 the increments above)
 * W1, W2 and the activations are bfloat16 MFMA operands (gfx950), every sum
 * accumulates in FP32; training (forward, backward, Adam) runs on the GPU.
 * Parameters are one FP32 vector in the order W1[64][16], W2[27][64][64]
 * (tap 9 (di+1) + 3 (dj+1) + (dk+1), then out, in), b2[64], W3[5][64], b3[5].
 * The normalisation record is mu[5], sig[5], s[5], sr[5], c1, c2. */
typedef struct cf_pred cf_pred;
#define CF_PRED_NPARAM 112005
#define CF_PRED_NNORM 22
#define CF_PRED_NSTATS 36
#define CF_PRED_PROFILE_STAGES 6
/* A predictor on the current device with room for dataset_crops training
 * crops of 16^3 cells (inner 14^3) and training batches of batch_crops. */
int cf_pred_create(int dataset_crops, int batch_crops, cf_pred** out);
void cf_pred_destroy(cf_pred* p);
const char* cf_pred_last_error(const cf_pred* p);
int cf_pred_init_params(cf_pred* p, unsigned long long seed);
int cf_pred_set_params(cf_pred* p, const float* params, int n);
int cf_pred_get_params(cf_pred* p, float* params, int n);
int cf_pred_set_norm(cf_pred* p, const float norm[CF_PRED_NNORM]);
int cf_pred_get_norm(cf_pred* p, float norm[CF_PRED_NNORM]);
/* Appends ncrops random crops of the context's last physical step (inputs
 * from U^{n-2}, U^{n-1}, U^n, target the converged U^{n+1}); the context
 * must keep four levels (cf_keep_levels) and have run three steps. */
int cf_pred_capture(cf_pred* p, cf_ctx* ctx, int ncrops, unsigned long long seed);
int cf_pred_dataset_size(const cf_pred* p);
/* Appends ncrops crops given as host floats [ncrops][16^3][24]: per cell
 * Q^n, Q^n - Q^{n-1}, Q^{n-1} - Q^{n-2}, Q^{n+1} - Q^n (5 each), the
 * in-domain mask, 3 zeros (the layout cf_pred_capture writes). */
int cf_pred_load_dataset(cf_pred* p, const float* crops, int ncrops);
/* Feature scales and base guess (c1, c2) from crops [0, n); stats (may be
 * NULL) receives the sums they come from. */
int cf_pred_fit_norm(cf_pred* p, int n, double stats[CF_PRED_NSTATS]);
/* steps Adam steps (cosine-decayed step size lr) on batches drawn from crops
 * [0, n_train); losses (may be NULL) receives each step's mean squared error
 * in normalised units. */
int cf_pred_train(cf_pred* p, int n_train, int steps, float lr, unsigned long long seed, float* losses);
/* Mean squared error of the network and of the base guess alone on crops
 * [first, first + n). */
int cf_pred_eval(cf_pred* p, int first, int n, double* loss, double* base_loss);
/* Test entry points on host arrays: y on the inner region (margin cells
 * dropped) of n volumes of features [n][VI][VJ][VK][16]; the loss and
 * gradient of n 16^3 crops with targets [n][14^3][5]. */
int cf_pred_forward(cf_pred* p, int n, int VI, int VJ, int VK, int margin, const float* x, float* y);
int cf_pred_gradient(cf_pred* p, int n, const float* x, const float* target, float* grad, double* loss);
/* HIP-event times (ms): inference over an X x Y x Z domain, then one
 * training batch: pointwise + forward conv, loss head, backward data,
 * backward weight, reduction. */
int cf_pred_profile(cf_pred* p, int X, int Y, int Z, double ms[CF_PRED_PROFILE_STAGES]);
/* Time levels kept besides U^n and U^{n-1}: 3 adds U^{n-2}, 4 adds U^{n-3}. */
int cf_keep_levels(cf_ctx* ctx, int levels);
/* Attaches a predictor (NULL detaches); cf_set_initial_guess(ctx, 3) then
 * starts each cf_physical_step with it once three levels are known.  The
 * network refines the quadratic extrapolation, which is the check: both
 * guesses take one inner iteration and the step continues from the network's
 * unless some er_k of its first iteration exceeds the quadratic's.  On such a
 * fallback the step continues from the quadratic one, or restarts from the
 * cubic one with cf_set_extrapolation_order(ctx, 3) and four known levels
 * (all iterations count).  The converged state is the same; the iteration
 * count changes.  The context keeps the pointer: detach (NULL) or destroy
 * the context before cf_pred_destroy. */
int cf_set_predictor(cf_ctx* ctx, cf_pred* p);
/* Physical steps started by the network, and how many fell back. */
int cf_predictor_stats(cf_ctx* ctx, long long* steps, long long* fallbacks);

#ifdef __cplusplus
}
#endif
#endif
