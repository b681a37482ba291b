// channel_flow_compat.hpp -- the reference's C++ call surface over the C ABI.
//
// Header-only.  Include it after the reference's Resolution.h (which defines
// Y_m and Z_m) in a driver written against main.h of ricklee209/Channel_Flow,
// and link against libchannel_flow.so.  Each function keeps the reference's
// name and parameter list (main.h:10-1069, definitions in the .cpp files
// cited below) and forwards to the cf_* entry points of channel_flow.h, which
// run on one MI355X.  The whole domain belongs to one context (one rank:
// myid is accepted and ignored).
//
//   cf_ctx* ctx; cf_create(&params, &ctx);
//   cf_compat::attach(ctx, params);        // once, before the first call
//   Jacobian(myid, J, xidx, ...);          // as in main.cpp
//
// Differences from the reference, all in outputs nothing downstream reads
// (and one that the reference leaves undefined):
//   - the MUSCL and viscous staging arrays (ML/MR, LR/LL/MR/ML/NR/NL of
//     Viscous_terms), EpX/EpY/EpZ of the non-limiter fluxes and Residual1..5
//     of Dual_time_stepping are accepted and left untouched (the device keeps
//     its own staging);
//   - Y_boundary_condition applies the single-rank x ghosts, the wall ghosts
//     and the periodic z ghosts together (the three are idempotent), and so
//     does Z_boundary_condition;
//   - a failing call prints cf_last_error() and aborts: the reference
//     functions return void;
//   - Flux_Y_limiter near the walls uses the face Jacobian that the
//     reference's interior branch assigns, where the reference reads one it
//     never assigned.
#ifndef CHANNEL_FLOW_COMPAT_HPP
#define CHANNEL_FLOW_COMPAT_HPP

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "channel_flow.h"

#if !defined(Y_m) || !defined(Z_m)
#error "include the reference's Resolution.h (Y_m, Z_m) before channel_flow_compat.hpp"
#endif

namespace cf_compat {

typedef double (*field)[Y_m][Z_m];

struct State {
  cf_ctx* ctx = nullptr;
  cf_params p{};
  double gamma = 1.5;       // Jacobian*.cpp hard-code
  double gamma_grid = 2.8;  // Grid.cpp hard-code
};

inline State& state() {
  static State s;
  return s;
}

// binds the context the reference-named functions forward to; gamma is the
// stretching factor of Jacobian*.cpp (1.5 there)
inline void attach(cf_ctx* ctx, const cf_params& p, double gamma = 1.5) {
  state().ctx = ctx;
  state().p = p;
  state().gamma = gamma;
}

inline cf_ctx* ctx(const char* fn) {
  if (!state().ctx) {
    std::fprintf(stderr, "%s: cf_compat::attach() was not called\n", fn);
    std::abort();
  }
  return state().ctx;
}

inline void check(int rc, const char* fn) {
  if (rc != CF_OK) {
    std::fprintf(stderr, "%s: %s\n", fn, cf_last_error(state().ctx));
    std::abort();
  }
}

inline double* p(field a) { return &a[0][0][0]; }

// X_np, gend, ny, nz of the single-rank layout (Mpi_division.h:27, main.cpp:71)
inline int X_np() { return state().p.X_out + 6; }
inline int gend() { return state().p.X_out + 2; }
inline int ny() { return state().p.Y_out + 1; }
inline int nz() { return state().p.Z_out + 1; }
inline size_t at(int i, int j, int k) { return ((size_t)i * Y_m + j) * Z_m + k; }

}  // namespace cf_compat

#define CF_F(n) cf_compat::field n
#define CF_P(n) cf_compat::p(n)

// Jacobian*.cpp: cell and face metrics (all four compute the same host
// geometry; each returns its own arrays)
inline void Jacobian(int, CF_F(J), CF_F(xidx), CF_F(xidy), CF_F(xidz), CF_F(etdx), CF_F(etdy), CF_F(etdz),
                     CF_F(ztdx), CF_F(ztdy), CF_F(ztdz)) {
  double* const m[9] = {CF_P(xidx), CF_P(xidy), CF_P(xidz), CF_P(etdx), CF_P(etdy),
                        CF_P(etdz), CF_P(ztdx), CF_P(ztdy), CF_P(ztdz)};
  cf_compat::check(cf_jacobian_all(cf_compat::ctx("Jacobian"), cf_compat::state().gamma, CF_P(J), m, nullptr,
                                   nullptr, nullptr, nullptr, nullptr, nullptr),
                   "Jacobian");
}

inline void Jacobian_u(int, CF_F(J_u), CF_F(xidx_u), CF_F(xidy_u), CF_F(xidz_u), CF_F(etdx_u), CF_F(etdy_u),
                       CF_F(etdz_u), CF_F(ztdx_u), CF_F(ztdy_u), CF_F(ztdz_u)) {
  double* const m[9] = {CF_P(xidx_u), CF_P(xidy_u), CF_P(xidz_u), CF_P(etdx_u), CF_P(etdy_u),
                        CF_P(etdz_u), CF_P(ztdx_u), CF_P(ztdy_u), CF_P(ztdz_u)};
  cf_compat::check(cf_jacobian_all(cf_compat::ctx("Jacobian_u"), cf_compat::state().gamma, nullptr, nullptr,
                                   CF_P(J_u), m, nullptr, nullptr, nullptr, nullptr),
                   "Jacobian_u");
}

// Jacobian_v.cpp also receives the cell metrics (its wall-face override
// copies them); they are inputs only
inline void Jacobian_v(int, CF_F(J_v), CF_F(xidx_v), CF_F(xidy_v), CF_F(xidz_v), CF_F(etdx_v), CF_F(etdy_v),
                       CF_F(etdz_v), CF_F(ztdx_v), CF_F(ztdy_v), CF_F(ztdz_v), CF_F(J), CF_F(xidx), CF_F(xidy),
                       CF_F(xidz), CF_F(etdx), CF_F(etdy), CF_F(etdz), CF_F(ztdx), CF_F(ztdy), CF_F(ztdz)) {
  (void)J, (void)xidx, (void)xidy, (void)xidz, (void)etdx, (void)etdy, (void)etdz, (void)ztdx, (void)ztdy,
      (void)ztdz;
  double* const m[9] = {CF_P(xidx_v), CF_P(xidy_v), CF_P(xidz_v), CF_P(etdx_v), CF_P(etdy_v),
                        CF_P(etdz_v), CF_P(ztdx_v), CF_P(ztdy_v), CF_P(ztdz_v)};
  cf_compat::check(cf_jacobian_all(cf_compat::ctx("Jacobian_v"), cf_compat::state().gamma, nullptr, nullptr,
                                   nullptr, nullptr, CF_P(J_v), m, nullptr, nullptr),
                   "Jacobian_v");
}

inline void Jacobian_w(int, CF_F(J_w), CF_F(xidx_w), CF_F(xidy_w), CF_F(xidz_w), CF_F(etdx_w), CF_F(etdy_w),
                       CF_F(etdz_w), CF_F(ztdx_w), CF_F(ztdy_w), CF_F(ztdz_w)) {
  double* const m[9] = {CF_P(xidx_w), CF_P(xidy_w), CF_P(xidz_w), CF_P(etdx_w), CF_P(etdy_w),
                        CF_P(etdz_w), CF_P(ztdx_w), CF_P(ztdy_w), CF_P(ztdz_w)};
  cf_compat::check(cf_jacobian_all(cf_compat::ctx("Jacobian_w"), cf_compat::state().gamma, nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nullptr, CF_P(J_w), m),
                   "Jacobian_w");
}

// Hyperplane.cpp:50-78: Nijk[m] = {cumulative end, count}, ijk[1..] the cells
// of plane m in (i, j, k) lexicographic order.  The device sweeps do not read
// these lists; they are filled for drivers that do.
inline void Hyperplane(int, int (*Nijk)[2], int (*ijk)[3]) {
  (void)cf_compat::ctx("Hyperplane");
  const int ge = cf_compat::gend(), ny = cf_compat::ny(), nz = cf_compat::nz();
  int ii = 0;
  for (int m = 7; m <= ge + ny + nz; m++) {
    int mi = 0;
    for (int i = 3; i <= ge; i++)
      for (int j = 2; j <= ny; j++) {
        const int k = m - i - j;
        if (k < 2 || k > nz) continue;
        ii++;
        mi++;
        ijk[ii][0] = i;
        ijk[ii][1] = j;
        ijk[ii][2] = k;
      }
    Nijk[m][0] = ii;
    Nijk[m][1] = mi;
  }
}

// Grid.cpp:50-141 (Z_point is not written, as in the reference)
inline void Grid(int, CF_F(X_point), CF_F(Y_point), CF_F(Z_point)) {
  (void)Z_point;
  cf_compat::check(cf_grid(cf_compat::ctx("Grid"), cf_compat::state().gamma_grid, CF_P(X_point), CF_P(Y_point)),
                   "Grid");
}

// Filter.cpp:12-1720
inline void Filter(int, double Roe_criterion, double E_high, double E_low, CF_F(U1_), CF_F(U2_), CF_F(U3_),
                   CF_F(U4_), CF_F(U5_), CF_F(Y_point), CF_F(EpX), CF_F(EpY), CF_F(EpZ), CF_F(J)) {
  const double* const U[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};
  cf_compat::check(cf_filter(cf_compat::ctx("Filter"), Roe_criterion, E_high, E_low, U, CF_P(Y_point), CF_P(EpX),
                             CF_P(EpY), CF_P(EpZ), CF_P(J)),
                   "Filter");
}

// main.cpp:156-250 x ghosts, Y_boundary_condition.cpp:48-120 and
// Z_boundary_condition.cpp:48-66, applied together
inline void Y_boundary_condition(int, CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_), CF_F(U5_), CF_F(J)) {
  double* const U[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};
  cf_compat::check(cf_boundary_conditions(cf_compat::ctx("Y_boundary_condition"), U, CF_P(J)),
                   "Y_boundary_condition");
}

inline void Z_boundary_condition(int myid, CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_), CF_F(U5_), CF_F(J)) {
  Y_boundary_condition(myid, U1_, U2_, U3_, U4_, U5_, J);
}

#define CF_FLUX(NAME, FN, S, F, EP)                                                                              \
  inline void NAME(int, CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_), CF_F(U5_), CF_F(MR1), CF_F(MR2), CF_F(MR3), \
                   CF_F(MR4), CF_F(MR5), CF_F(ML1), CF_F(ML2), CF_F(ML3), CF_F(ML4), CF_F(ML5), CF_F(J),        \
                   CF_F(J##S), CF_F(xidx##S), CF_F(xidy##S), CF_F(xidz##S), CF_F(etdx##S), CF_F(etdy##S),       \
                   CF_F(etdz##S), CF_F(ztdx##S), CF_F(ztdy##S), CF_F(ztdz##S), CF_F(F##1), CF_F(F##2),          \
                   CF_F(F##3), CF_F(F##4), CF_F(F##5), CF_F(EP)) {                                             \
    (void)MR1, (void)MR2, (void)MR3, (void)MR4, (void)MR5, (void)ML1, (void)ML2, (void)ML3, (void)ML4, (void)ML5, \
        (void)EP;                                                                                              \
    const double* const U[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};                        \
    const double* const m[9] = {CF_P(xidx##S), CF_P(xidy##S), CF_P(xidz##S), CF_P(etdx##S), CF_P(etdy##S),     \
                                CF_P(etdz##S), CF_P(ztdx##S), CF_P(ztdy##S), CF_P(ztdz##S)};                   \
    double* const out[5] = {CF_P(F##1), CF_P(F##2), CF_P(F##3), CF_P(F##4), CF_P(F##5)};                       \
    cf_compat::check(FN(cf_compat::ctx(#NAME), U, CF_P(J), CF_P(J##S), m, out), #NAME);                        \
  }

// Flux_X.cpp:92-424, Flux_Y.cpp:91-465, Flux_Z.cpp:96-512
CF_FLUX(Flux_X, cf_flux_x, _u, inFx, EpX)
CF_FLUX(Flux_Y, cf_flux_y, _v, inFy, EpY)
CF_FLUX(Flux_Z, cf_flux_z, _w, inFz, EpZ)
#undef CF_FLUX

// the limiter variants take the same parameters and read the sensor
#define CF_FLUX_LIM(NAME, FN, S, F, EP)                                                                          \
  inline void NAME(int, CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_), CF_F(U5_), CF_F(MR1), CF_F(MR2), CF_F(MR3), \
                   CF_F(MR4), CF_F(MR5), CF_F(ML1), CF_F(ML2), CF_F(ML3), CF_F(ML4), CF_F(ML5), CF_F(J),        \
                   CF_F(J##S), CF_F(xidx##S), CF_F(xidy##S), CF_F(xidz##S), CF_F(etdx##S), CF_F(etdy##S),       \
                   CF_F(etdz##S), CF_F(ztdx##S), CF_F(ztdy##S), CF_F(ztdz##S), CF_F(F##1), CF_F(F##2),          \
                   CF_F(F##3), CF_F(F##4), CF_F(F##5), CF_F(EP)) {                                             \
    (void)MR1, (void)MR2, (void)MR3, (void)MR4, (void)MR5, (void)ML1, (void)ML2, (void)ML3, (void)ML4, (void)ML5; \
    const double* const U[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};                        \
    const double* const m[9] = {CF_P(xidx##S), CF_P(xidy##S), CF_P(xidz##S), CF_P(etdx##S), CF_P(etdy##S),     \
                                CF_P(etdz##S), CF_P(ztdx##S), CF_P(ztdy##S), CF_P(ztdz##S)};                   \
    double* const out[5] = {CF_P(F##1), CF_P(F##2), CF_P(F##3), CF_P(F##4), CF_P(F##5)};                       \
    cf_compat::check(FN(cf_compat::ctx(#NAME), U, CF_P(J), CF_P(J##S), m, out, CF_P(EP)), #NAME);              \
  }

// Flux_X_limiter.cpp:104-547, Flux_Y_limiter.cpp:106-685, Flux_Z_limiter.cpp:113-710
CF_FLUX_LIM(Flux_X_limiter, cf_flux_x_limiter, _u, inFx, EpX)
CF_FLUX_LIM(Flux_Y_limiter, cf_flux_y_limiter, _v, inFy, EpY)
CF_FLUX_LIM(Flux_Z_limiter, cf_flux_z_limiter, _w, inFz, EpZ)
#undef CF_FLUX_LIM

#define CF_METRIC_ARGS(S)                                                                                   \
  CF_F(J##S), CF_F(xidx##S), CF_F(xidy##S), CF_F(xidz##S), CF_F(etdx##S), CF_F(etdy##S), CF_F(etdz##S), \
      CF_F(ztdx##S), CF_F(ztdy##S), CF_F(ztdz##S)
#define CF_METRIC_PTRS(S)                                                                                \
  {CF_P(xidx##S), CF_P(xidy##S), CF_P(xidz##S), CF_P(etdx##S), CF_P(etdy##S), CF_P(etdz##S), CF_P(ztdx##S), \
   CF_P(ztdy##S), CF_P(ztdz##S)}
#define CF_FIVE(n) CF_F(n##1), CF_F(n##2), CF_F(n##3), CF_F(n##4), CF_F(n##5)
#define CF_FIVE_P(n) {CF_P(n##1), CF_P(n##2), CF_P(n##3), CF_P(n##4), CF_P(n##5)}

// Viscous_terms.cpp:212-883
inline void Viscous_terms(int, CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_), CF_F(U5_), CF_F(vF2), CF_F(vF3),
                          CF_F(vF4), CF_F(vF5), CF_METRIC_ARGS(), CF_METRIC_ARGS(_u), CF_METRIC_ARGS(_v),
                          CF_METRIC_ARGS(_w), CF_FIVE(LR), CF_FIVE(LL), CF_FIVE(MR), CF_FIVE(ML), CF_FIVE(NR),
                          CF_FIVE(NL)) {
  const double* const U[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};
  double* const vF[4] = {CF_P(vF2), CF_P(vF3), CF_P(vF4), CF_P(vF5)};
  const double* const m[9] = CF_METRIC_PTRS();
  const double* const mu[9] = CF_METRIC_PTRS(_u);
  const double* const mv[9] = CF_METRIC_PTRS(_v);
  const double* const mw[9] = CF_METRIC_PTRS(_w);
  (void)LR1, (void)LL1, (void)MR1, (void)ML1, (void)NR1, (void)NL1;
  cf_compat::check(cf_viscous_terms(cf_compat::ctx("Viscous_terms"), U, vF, CF_P(J), m, CF_P(J_u), mu, CF_P(J_v), mv,
                                    CF_P(J_w), mw),
                   "Viscous_terms");
}

// Dual_time_stepping.cpp:18-1361: one inner iteration's implicit part
inline void Dual_time_stepping(int, double deltaT, double* er1, double* er2, double* er3, double* er4, double* er5,
                               int (*Nijk)[2], int (*ijk)[3], CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_),
                               CF_F(U5_), CF_F(U1), CF_F(U2), CF_F(U3), CF_F(U4), CF_F(U5), CF_F(U1q), CF_F(U2q),
                               CF_F(U3q), CF_F(U4q), CF_F(U5q), CF_METRIC_ARGS(), CF_FIVE(inFx), CF_FIVE(inFy),
                               CF_FIVE(inFz), CF_F(vF2), CF_F(vF3), CF_F(vF4), CF_F(vF5), CF_FIVE(MR),
                               CF_FIVE(NR), CF_FIVE(NL), CF_FIVE(Residual)) {
  (void)Nijk, (void)ijk, (void)Residual1;
  double* const Ui[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};
  const double* const Un[5] = {CF_P(U1), CF_P(U2), CF_P(U3), CF_P(U4), CF_P(U5)};
  const double* const Uq[5] = {CF_P(U1q), CF_P(U2q), CF_P(U3q), CF_P(U4q), CF_P(U5q)};
  const double* const m[9] = CF_METRIC_PTRS();
  const double* const fx[5] = CF_FIVE_P(inFx);
  const double* const fy[5] = CF_FIVE_P(inFy);
  const double* const fz[5] = CF_FIVE_P(inFz);
  const double* const vF[4] = {CF_P(vF2), CF_P(vF3), CF_P(vF4), CF_P(vF5)};
  double* const MRp[5] = CF_FIVE_P(MR);
  double* const NRp[5] = CF_FIVE_P(NR);
  double* const NLp[5] = CF_FIVE_P(NL);
  double er[5];
  cf_compat::check(cf_dual_time_stepping(cf_compat::ctx("Dual_time_stepping"), deltaT, er, Ui, Un, Uq, CF_P(J), m,
                                         fx, fy, fz, vF, MRp, NRp, NLp),
                   "Dual_time_stepping");
  *er1 = er[0], *er2 = er[1], *er3 = er[2], *er4 = er[3], *er5 = er[4];
}

// RK3NODT.cpp:16-183 (main.h): stage RK of the explicit SSP-RK3 update;
// er1..er5 are left untouched, as the reference leaves them
inline void RK3NODT(int, int RK, double deltaT, double* er1, double* er2, double* er3, double* er4, double* er5,
                    CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_), CF_F(U5_), CF_F(U1), CF_F(U2), CF_F(U3), CF_F(U4),
                    CF_F(U5), CF_F(U1q), CF_F(U2q), CF_F(U3q), CF_F(U4q), CF_F(U5q), CF_METRIC_ARGS(),
                    CF_FIVE(inFx), CF_FIVE(inFy), CF_FIVE(inFz), CF_F(vF2), CF_F(vF3), CF_F(vF4), CF_F(vF5),
                    CF_FIVE(MR)) {
  (void)er1, (void)er2, (void)er3, (void)er4, (void)er5;
  double* const Ui[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};
  const double* const Un[5] = {CF_P(U1), CF_P(U2), CF_P(U3), CF_P(U4), CF_P(U5)};
  const double* const Uq[5] = {CF_P(U1q), CF_P(U2q), CF_P(U3q), CF_P(U4q), CF_P(U5q)};
  const double* const m[9] = CF_METRIC_PTRS();
  const double* const fx[5] = CF_FIVE_P(inFx);
  const double* const fy[5] = CF_FIVE_P(inFy);
  const double* const fz[5] = CF_FIVE_P(inFz);
  const double* const vF[4] = {CF_P(vF2), CF_P(vF3), CF_P(vF4), CF_P(vF5)};
  double* const MRp[5] = CF_FIVE_P(MR);
  double er[5];
  cf_compat::check(cf_rk3nodt(cf_compat::ctx("RK3NODT"), RK, deltaT, er, Ui, Un, Uq, CF_P(J), m, fx, fy, fz, vF, MRp),
                   "RK3NODT");
}

// Driven_force.cpp:38-73 (main.h): prints the bulk flux line as the
// reference does, with prm.h's Q0
inline void Driven_force(int myid, CF_F(U2_), CF_F(J)) {
  const double Q0 = 20.464969232814912;  // prm.h:15
  double Q = 0, E = 0;
  cf_compat::check(cf_driven_force(cf_compat::ctx("Driven_force"), Q0, CF_P(U2_), CF_P(J), &Q, &E), "Driven_force");
  if (myid == 0) std::printf("Q0=%f\tQall=%f\tER_Q=%f\n", Q0, Q, E);
}

// Statistic.cpp:1043-1154 (main.h): the mean profiles and st%05d.dat in the
// working directory, like the reference; the numerical-dissipation and
// shear-stress statistics of Statistic.cpp:100-1040 are not produced
inline void Statistic(int, int step, int iteration, int statistic_step, CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_),
                      CF_F(U5_), CF_F(J), CF_F(J_v), CF_F(xidx_v), CF_F(xidy_v), CF_F(xidz_v), CF_F(etdx_v),
                      CF_F(etdy_v), CF_F(etdz_v), CF_F(ztdx_v), CF_F(ztdy_v), CF_F(ztdz_v), CF_FIVE(MR), CF_FIVE(ML),
                      CF_F(EpY)) {
  (void)iteration, (void)J, (void)J_v, (void)xidx_v, (void)xidy_v, (void)xidz_v, (void)etdx_v, (void)etdy_v;
  (void)etdz_v, (void)ztdx_v, (void)ztdy_v, (void)ztdz_v, (void)MR1, (void)MR2, (void)MR3, (void)MR4, (void)MR5;
  (void)ML1, (void)ML2, (void)ML3, (void)ML4, (void)ML5, (void)EpY;
  const double* const U[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};
  cf_compat::check(cf_statistic(cf_compat::ctx("Statistic"), step, statistic_step, U, "."), "Statistic");
}

// DPLUSGS.cpp:189-1361 (the Dual_time_stepping parameters without Nijk/ijk);
// the upper neighbours' NL is read at j+1, k+1 (the reference's j_, k_ are
// never assigned)
inline void DPLUSGS(int, double deltaT, double* er1, double* er2, double* er3, double* er4, double* er5,
                    CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_), CF_F(U5_), CF_F(U1), CF_F(U2), CF_F(U3), CF_F(U4),
                    CF_F(U5), CF_F(U1q), CF_F(U2q), CF_F(U3q), CF_F(U4q), CF_F(U5q), CF_METRIC_ARGS(),
                    CF_FIVE(inFx), CF_FIVE(inFy), CF_FIVE(inFz), CF_F(vF2), CF_F(vF3), CF_F(vF4), CF_F(vF5),
                    CF_FIVE(MR), CF_FIVE(NR), CF_FIVE(NL), CF_FIVE(Residual)) {
  (void)Residual1;
  double* const Ui[5] = {CF_P(U1_), CF_P(U2_), CF_P(U3_), CF_P(U4_), CF_P(U5_)};
  const double* const Un[5] = {CF_P(U1), CF_P(U2), CF_P(U3), CF_P(U4), CF_P(U5)};
  const double* const Uq[5] = {CF_P(U1q), CF_P(U2q), CF_P(U3q), CF_P(U4q), CF_P(U5q)};
  const double* const m[9] = CF_METRIC_PTRS();
  const double* const fx[5] = CF_FIVE_P(inFx);
  const double* const fy[5] = CF_FIVE_P(inFy);
  const double* const fz[5] = CF_FIVE_P(inFz);
  const double* const vF[4] = {CF_P(vF2), CF_P(vF3), CF_P(vF4), CF_P(vF5)};
  double* const MRp[5] = CF_FIVE_P(MR);
  double* const NRp[5] = CF_FIVE_P(NR);
  double* const NLp[5] = CF_FIVE_P(NL);
  double er[5];
  cf_compat::check(cf_dplusgs(cf_compat::ctx("DPLUSGS"), deltaT, er, Ui, Un, Uq, CF_P(J), m, fx, fy, fz, vF, MRp,
                              NRp, NLp),
                   "DPLUSGS");
  *er1 = er[0], *er2 = er[1], *er3 = er[2], *er4 = er[3], *er5 = er[4];
}

// Initial_condition.cpp:64-116: reads initial.bin (10 blocks of
// X_out*Y_out*Z_out FP64: rho, rho u, rho v, rho w, rho E at n, then n-1)
// and stores Q/J; the inner iterate starts at level n (main.cpp:112-127)
inline void Initial_condition(int, CF_F(U1_), CF_F(U2_), CF_F(U3_), CF_F(U4_), CF_F(U5_), CF_F(U1), CF_F(U2),
                              CF_F(U3), CF_F(U4), CF_F(U5), CF_F(U1q), CF_F(U2q), CF_F(U3q), CF_F(U4q), CF_F(U5q),
                              CF_F(J)) {
  (void)cf_compat::ctx("Initial_condition");
  const cf_params& p = cf_compat::state().p;
  const size_t n = (size_t)p.X_out * p.Y_out * p.Z_out;
  std::vector<double> buf(10 * n);
  std::FILE* f = std::fopen("initial.bin", "rb");
  if (!f || std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) {
    std::fprintf(stderr, "Initial_condition: cannot read initial.bin\n");
    std::abort();
  }
  std::fclose(f);
  cf_compat::field lv[10] = {U1, U2, U3, U4, U5, U1q, U2q, U3q, U4q, U5q};
  cf_compat::field it[5] = {U1_, U2_, U3_, U4_, U5_};
  for (int v = 0; v < 10; v++)
    for (int i = 0; i < p.X_out; i++)
      for (int j = 0; j < p.Y_out; j++)
        for (int k = 0; k < p.Z_out; k++) {
          const double q = buf[v * n + ((size_t)i * p.Y_out + j) * p.Z_out + k];
          lv[v][i + 3][j + 2][k + 2] = q / J[i + 3][j + 2][k + 2];
          if (v < 5) it[v][i + 3][j + 2][k + 2] = lv[v][i + 3][j + 2][k + 2];
        }
}

// Output.cpp:56-190: LES%05d.bin with Q = U*J of levels n and n-1
inline void Output(int step, int, CF_F(U1), CF_F(U2), CF_F(U3), CF_F(U4), CF_F(U5), CF_F(U1q), CF_F(U2q),
                   CF_F(U3q), CF_F(U4q), CF_F(U5q), CF_F(J)) {
  (void)cf_compat::ctx("Output");
  const cf_params& p = cf_compat::state().p;
  const size_t n = (size_t)p.X_out * p.Y_out * p.Z_out;
  std::vector<double> buf(10 * n);
  cf_compat::field lv[10] = {U1, U2, U3, U4, U5, U1q, U2q, U3q, U4q, U5q};
  for (int v = 0; v < 10; v++)
    for (int i = 0; i < p.X_out; i++)
      for (int j = 0; j < p.Y_out; j++)
        for (int k = 0; k < p.Z_out; k++)
          buf[v * n + ((size_t)i * p.Y_out + j) * p.Z_out + k] = lv[v][i + 3][j + 2][k + 2] * J[i + 3][j + 2][k + 2];
  char name[32];
  std::snprintf(name, sizeof name, "LES%05d.bin", step);
  std::FILE* f = std::fopen(name, "wb");
  if (!f || std::fwrite(buf.data(), sizeof(double), buf.size(), f) != buf.size()) {
    std::fprintf(stderr, "Output: cannot write %s\n", name);
    std::abort();
  }
  std::fclose(f);
}

#undef CF_F
#undef CF_P
#undef CF_METRIC_ARGS
#undef CF_METRIC_PTRS
#undef CF_FIVE
#undef CF_FIVE_P

#endif  // CHANNEL_FLOW_COMPAT_HPP
