/* CPU oracle for the channel-flow dual-time hot path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference's FP64 arithmetic, written for
 * checking the HIP path (tests/, smoke(), bench.py cpu_baseline).  Every
 * expression keeps the reference's operation order (compile with
 * -ffp-contract=off) so that results agree with the reference bit for bit;
 * the code structure (recipe tables, one Roe routine for three directions,
 * per-point primitive records) is this project's own.
 *
 * Reference citations are file:line in ricklee209/Channel_Flow.
 */
#include "cf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KG 1.4            /* prm.h:9   K   */
#define RG 287.0          /* prm.h:10  R   */
#define EPS_PRE 0.0001    /* prm.h:7   e   */
#define CV 717.5          /* Viscous_terms.h:3 */
#define MU_L 0.0000185    /* Viscous_terms.h:4 */
#define PR_L 0.72         /* Viscous_terms.h:5 */
#define T_WALL2 (2 * 298.0592) /* Y_boundary_condition.cpp:87 */

typedef struct {
  int Ym, Zm, gend, ny, nz;
} dims_t;

static dims_t dims(const cfo_grid* g) {
  dims_t d;
  d.Ym = g->Y_out + 4;
  d.Zm = g->Z_out + 4;
  d.gend = g->X_out + 2;
  d.ny = g->Y_out + 1;
  d.nz = g->Z_out + 1;
  return d;
}

#define IX(d, i, j, k) ((((size_t)(i)) * (size_t)(d).Ym + (size_t)(j)) * (size_t)(d).Zm + (size_t)(k))

static double dmax(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* metrics: Jacobian.cpp:60-244, Jacobian_u.cpp:57-97, Jacobian_v.cpp:77-141,
 * Jacobian_w.cpp:57-98 (tanh stretching in y, Cartesian in x and z)       */
/* ------------------------------------------------------------------------ */

static void stretch(const cfo_grid* g, double gamma, double jpos, double* etdy, double* J) {
  /* Jacobian.cpp:74-99 with (j-1.5) replaced by the caller's offset */
  double half = 0.5 * g->high;
  double a = gamma * (1 - 2 * jpos * g->deltaET);
  double temp = half * (1 - 1. / tanh(gamma) * tanh(a));
  double temp1 = (1 - temp / half) * tanh(gamma);
  *etdy = tanh(gamma) / (2 * half * gamma * (1 - temp1 * temp1));
  double temp2 = 1 / cosh(a);
  double ydet = g->high * gamma * 1. / tanh(gamma) * temp2 * temp2;
  *J = 1. / (1 * ydet * 1);
}

static void set_cart(double* const m[9], size_t p, double etdy) {
  const double c[9] = {1, 0, 0, 0, 0, 0, 0, 0, 1};
  for (int q = 0; q < 9; q++) m[q][p] = c[q];
  m[4][p] = etdy;
}

void cfo_metrics(const cfo_grid* g, double gamma, double* J, double* const m[9], double* Ju,
                 double* const mu[9], double* Jv, double* const mv[9], double* Jw, double* const mw[9]) {
  dims_t d = dims(g);
  int gend = d.gend, ny = d.ny, nz = d.nz, nyy = ny + 1, nzz = nz + 1;
  /* cell centres, interior (Jacobian.cpp:60-108): eta = (j-1.5) deltaET */
  for (int i = 3; i <= gend; i++)
    for (int j = 2; j <= ny; j++)
      for (int k = 2; k <= nz; k++) {
        double ey, jj;
        stretch(g, gamma, j - 1.5, &ey, &jj);
        set_cart(m, IX(d, i, j, k), ey);
        J[IX(d, i, j, k)] = jj;
      }
  /* x ghosts (Jacobian.cpp:111-160): all but ztdy/ztdz copied (ztdx is
   * written three times there), before the y/z ghosts exist */
  const int xg[2][2] = {{2, 3}, {gend + 1, gend}};
  for (int s = 0; s < 2; s++) {
    int ig = xg[s][0], is = xg[s][1];
    for (int j = 1; j <= nyy; j++)
      for (int k = 1; k <= nzz; k++) {
        size_t p = IX(d, ig, j, k), q = IX(d, is, j, k);
        for (int c = 0; c < 7; c++) m[c][p] = m[c][q];
        J[p] = J[q];
      }
  }
  /* y ghosts j=1, nyy (Jacobian.cpp:162-203): ztdy/ztdz take ztdx */
  for (int i = 2; i <= gend + 1; i++)
    for (int k = 2; k <= nz; k++) {
      const int yg[2][2] = {{1, 2}, {nyy, ny}};
      for (int s = 0; s < 2; s++) {
        size_t p = IX(d, i, yg[s][0], k), q = IX(d, i, yg[s][1], k);
        for (int c = 0; c < 7; c++) m[c][p] = m[c][q];
        m[7][p] = m[6][q];
        m[8][p] = m[6][q];
        J[p] = J[q];
      }
    }
  /* z ghosts k=1, nzz (Jacobian.cpp:205-244) */
  for (int i = 2; i <= gend + 1; i++)
    for (int j = 1; j <= nyy; j++) {
      const int zg[2][2] = {{1, 2}, {nzz, nz}};
      for (int s = 0; s < 2; s++) {
        size_t p = IX(d, i, j, zg[s][0]), q = IX(d, i, j, zg[s][1]);
        for (int c = 0; c < 7; c++) m[c][p] = m[c][q];
        m[7][p] = m[6][q];
        m[8][p] = m[6][q];
        J[p] = J[q];
      }
    }
  /* x faces (Jacobian_u.cpp:57-97): i in [2,gend], j in [1,ny-1], eta = (j-0.5) */
  for (int i = 2; i <= gend; i++)
    for (int j = 1; j < ny; j++)
      for (int k = 1; k < nz; k++) {
        double ey, jj;
        stretch(g, gamma, j - 0.5, &ey, &jj);
        set_cart(mu, IX(d, i, j, k), ey);
        Ju[IX(d, i, j, k)] = jj;
      }
  /* y faces (Jacobian_v.cpp:77-141): eta = (j-1.0); wall faces take the
   * cell-centre ghost values */
  for (int i = 2; i <= gend - 1; i++)
    for (int j = 1; j <= ny; j++)
      for (int k = 1; k < nz; k++) {
        double ey, jj;
        stretch(g, gamma, j - 1.0, &ey, &jj);
        set_cart(mv, IX(d, i, j, k), ey);
        Jv[IX(d, i, j, k)] = jj;
      }
  for (int i = 2; i <= gend - 1; i++)
    for (int k = 1; k < nz; k++) {
      size_t w0 = IX(d, i, 1, k), w1 = IX(d, i, ny, k);
      size_t c0 = IX(d, i, 1, k), c1 = IX(d, i, nyy, k);
      for (int c = 3; c < 6; c++) {
        mv[c][w0] = m[c][c0];
        mv[c][w1] = m[c][c1];
      }
      Jv[w1] = J[c1];
      Jv[w0] = J[c0];
    }
  /* z faces (Jacobian_w.cpp:57-98): j in [1,ny-1], k in [1,nz], eta = (j-0.5) */
  for (int i = 2; i <= gend - 1; i++)
    for (int j = 1; j < ny; j++)
      for (int k = 1; k <= nz; k++) {
        double ey, jj;
        stretch(g, gamma, j - 0.5, &ey, &jj);
        set_cart(mw, IX(d, i, j, k), ey);
        Jw[IX(d, i, j, k)] = jj;
      }
}

int cfo_hyperplane_counts(const cfo_grid* g, int* counts, int ncounts) {
  dims_t d = dims(g);
  int mmax = d.gend + d.ny + d.nz;
  int n = mmax - 7 + 1;
  if (ncounts < n) return -1;
  memset(counts, 0, sizeof(int) * (size_t)n);
  for (int i = 3; i <= d.gend; i++)
    for (int j = 2; j <= d.ny; j++)
      for (int k = 2; k <= d.nz; k++) counts[i + j + k - 7]++;
  return n;
}

/* ------------------------------------------------------------------------ */
/* ghosts: main.cpp:203-240, Y_boundary_condition.cpp:48-120,               */
/* Z_boundary_condition.cpp:48-66                                           */
/* ------------------------------------------------------------------------ */

void cfo_halo_bc(const cfo_grid* g, double* const U_[5], const double* J) {
  dims_t d = dims(g);
  int gend = d.gend, ny = d.ny, nz = d.nz, nyy = ny + 1, nzz = nz + 1, nyyy = nyy + 1;
  size_t plane = (size_t)d.Ym * d.Zm;
  /* x: one plane each side; a single rank receives its own plane back */
  int src2 = g->x_periodic ? gend : 3;
  int srcE = g->x_periodic ? 3 : gend;
  for (int v = 0; v < 5; v++) {
    memmove(U_[v] + IX(d, 2, 0, 0), U_[v] + IX(d, src2, 0, 0), plane * sizeof(double));
    memmove(U_[v] + IX(d, gend + 1, 0, 0), U_[v] + IX(d, srcE, 0, 0), plane * sizeof(double));
  }
  /* isothermal no-slip walls: P extrapolated, T mirrored about T_w, u reflected */
  for (int i = 2; i <= gend + 1; i++)
    for (int k = 2; k <= nz; k++) {
      const int wall[2][2] = {{2, 1}, {ny, nyy}};
      for (int s = 0; s < 2; s++) {
        size_t a = IX(d, i, wall[s][0], k), b = IX(d, i, wall[s][1], k);
        double rho = U_[0][a] * J[a];
        double U = U_[1][a] / U_[0][a];
        double V = U_[2][a] / U_[0][a];
        double W = U_[3][a] / U_[0][a];
        double VV = U * U + V * V + W * W;
        double P = (U_[4][a] * J[a] - 0.5 * rho * VV) * (KG - 1);
        double temp = P / rho / RG;
        double T = T_WALL2 - temp;
        rho = P / RG / T;
        U_[0][b] = P / RG / T / J[b];
        U_[1][b] = -P / RG / T * U / J[b];
        U_[2][b] = -P / RG / T * V / J[b];
        U_[3][b] = -P / RG / T * W / J[b];
        U_[4][b] = (rho * RG * T / (KG - 1) + 0.5 * P / RG / T * VV) / J[b];
      }
    }
  /* periodic z, all rows including the wall ghosts */
  for (int i = 2; i <= gend + 1; i++)
    for (int j = 0; j <= nyyy; j++)
      for (int v = 0; v < 5; v++) {
        U_[v][IX(d, i, j, 1)] = U_[v][IX(d, i, j, nz)];
        U_[v][IX(d, i, j, nzz)] = U_[v][IX(d, i, j, 2)];
      }
}

/* ------------------------------------------------------------------------ */
/* MUSCL reconstruction recipes (Flux_X.cpp:92-215, Flux_Y.cpp:91-240,      */
/* Flux_Z.cpp:96-300).  A face f separates cells f and f+1 along the         */
/* direction.  Each state is sum(w*Q/J) / sum(w/J) over consecutive cells    */
/* starting at `first`.                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  int first;    /* first cell of the stencil (offset from face index f) */
  int n;        /* number of weights */
  double w[5];
} recipe_t;

static const recipe_t R_O5L = {-2, 5, {2, -13, 47, 27, -3}};  /* left,  centred on f   */
static const recipe_t R_O5R = {-1, 5, {-3, 27, 47, -13, 2}};  /* right, centred on f+1 */
static const recipe_t R_O3L = {-1, 3, {-1, 5, 2}};
static const recipe_t R_O3R = {0, 3, {2, 5, -1}};
static const recipe_t R_AVG = {0, 2, {1, 1}};                 /* cells f, f+1 */
static const recipe_t R_AVGM = {-1, 2, {1, 1}};               /* cells f-1, f */
static const recipe_t R_AVG1 = {1, 2, {1, 1}};                /* cells f+1, f+2 */

/* face range and recipe choice per direction (single rank) */
static void face_range(const dims_t* d, int dir, int* f0, int* f1) {
  if (dir == 0) { *f0 = 2; *f1 = d->gend; }
  else if (dir == 1) { *f0 = 1; *f1 = d->ny; }
  else { *f0 = 1; *f1 = d->nz; }
}

static const recipe_t* recipe_left(const dims_t* d, int dir, int f) {
  if (dir == 0) {
    int ge = d->gend;
    if (f == 2) return &R_AVG;
    if (f == 3 || f == ge) return &R_O3L;
    return &R_O5L;
  } else if (dir == 1) {
    int ny = d->ny;
    if (f == 1 || f == 2) return &R_AVG;
    if (f == 3 || f == ny - 1) return &R_O3L;
    if (f == ny) return &R_AVGM;
    return &R_O5L;
  } else {
    int nz = d->nz;
    if (f == 1) return &R_AVG;
    if (f == 2 || f == nz) return &R_O3L;
    return &R_O5L;
  }
}

static const recipe_t* recipe_right(const dims_t* d, int dir, int f) {
  if (dir == 0) {
    int ge = d->gend;
    if (f == 2 || f == ge - 1) return &R_O3R;
    if (f == ge) return &R_AVG;
    return &R_O5R;
  } else if (dir == 1) {
    int ny = d->ny;
    if (f == 1) return &R_AVG1;
    if (f == 2 || f == ny - 2) return &R_O3R;
    if (f == ny - 1 || f == ny) return &R_AVG;
    return &R_O5R;
  } else {
    int nz = d->nz;
    if (f == 1 || f == nz - 1) return &R_O3R;
    if (f == nz) return &R_AVG;
    return &R_O5R;
  }
}

static void reconstruct(const recipe_t* r, double* const U_[5], const double* J, size_t base, size_t stride,
                        int f, double q[5]) {
  size_t p0 = base + (size_t)((long)(f + r->first)) * stride;
  double s = r->w[0] / J[p0];
  for (int t = 1; t < r->n; t++) s += r->w[t] / J[p0 + (size_t)t * stride];
  double temp = 1. / s;
  for (int v = 0; v < 5; v++) {
    double a = r->w[0] * U_[v][p0];
    for (int t = 1; t < r->n; t++) a += r->w[t] * U_[v][p0 + (size_t)t * stride];
    q[v] = temp * a;
  }
}

/* Roe flux, ROE == 1 branch (Flux_X.cpp:221-420): n = contravariant metric
 * row of the face direction.  The energy dissipation uses deltaP times the
 * un-normalised contravariant velocity in x and z (Flux_X.cpp:412,
 * Flux_Z.cpp:499) and the normalised one in y (Flux_Y.cpp:441). */
static void roe_flux(const double L[5], const double Rr[5], const double n[3], double Jf, int energy_norm,
                     double F[5], double* Fav3_out, double* Pavg_out) {
  double insqr = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double N0 = n[0] * insqr, N1 = n[1] * insqr, N2 = n[2] * insqr;
  double _rho = L[0], _u = L[1] / _rho, _v = L[2] / _rho, _w = L[3] / _rho;
  double _U = n[0] * _u + n[1] * _v + n[2] * _w;
  double _VV = _u * _u + _v * _v + _w * _w;
  double _P = (L[4] - 0.5 * _rho * _VV) * (KG - 1);
  double _C = KG * _P / _rho;
  double _H = 0.5 * _VV + _C / (KG - 1);
  double rho_ = Rr[0], u_ = Rr[1] / rho_, v_ = Rr[2] / rho_, w_ = Rr[3] / rho_;
  double U_ = n[0] * u_ + n[1] * v_ + n[2] * w_;
  double VV_ = u_ * u_ + v_ * v_ + w_ * w_;
  double P_ = (Rr[4] - 0.5 * rho_ * VV_) * (KG - 1);
  double C_ = KG * P_ / rho_;
  double H_ = 0.5 * VV_ + C_ / (KG - 1);
  double t5 = sqrt(_rho), t6 = sqrt(rho_), t4 = t5 + t6;
  double rho = sqrt(_rho * rho_);
  double u = (t5 * _u + t6 * u_) / t4;
  double v = (t5 * _v + t6 * v_) / t4;
  double w = (t5 * _w + t6 * w_) / t4;
  double U = (t5 * _U + t6 * U_) / t4;
  double VV = u * u + v * v + w * w;
  double H = (t5 * _H + t6 * H_) / t4;
  double C = (H - 0.5 * VV) * (KG - 1);
  double dU1 = rho_ - _rho;
  double dU2 = rho_ * u_ - _rho * _u;
  double dU3 = rho_ * v_ - _rho * _v;
  double dU4 = rho_ * w_ - _rho * _w;
  double dU5 = (P_ / (KG - 1) + 0.5 * rho_ * VV_) - (_P / (KG - 1) + 0.5 * _rho * _VV);
  double S = sqrt(C);
  double temp1 = (P_ - _P) / rho / C;
  double temp2 = U * insqr / S * (U_ - _U) * insqr;
  double deltaU = (S - fabs(U * insqr)) * temp1 + temp2;
  double deltaP = U * insqr / S * (P_ - _P) + (S - fabs(U * insqr)) * rho * (U_ - _U) * insqr;
  double a = fabs(U * insqr);
  double Fav1 = a * dU1 + deltaU * rho;
  double Fav2 = a * dU2 + deltaU * rho * u + deltaP * N0;
  double Fav3 = a * dU3 + deltaU * rho * v + deltaP * N1;
  double Fav4 = a * dU4 + deltaU * rho * w + deltaP * N2;
  double Fav5 = energy_norm ? a * dU5 + deltaU * rho * H + deltaP * U * insqr
                            : a * dU5 + deltaU * rho * H + deltaP * U;
  F[0] = 0.5 * ((_rho * _U + rho_ * U_) - Fav1 / insqr) / Jf;
  F[1] = 0.5 * ((_rho * _u * _U + rho_ * u_ * U_ + n[0] * (_P + P_)) - Fav2 / insqr) / Jf;
  F[2] = 0.5 * ((_rho * _v * _U + rho_ * v_ * U_ + n[1] * (_P + P_)) - Fav3 / insqr) / Jf;
  F[3] = 0.5 * ((_rho * _w * _U + rho_ * w_ * U_ + n[2] * (_P + P_)) - Fav4 / insqr) / Jf;
  F[4] = 0.5 * ((_U * (3.5 * _P + 0.5 * _rho * _VV) + U_ * (3.5 * P_ + 0.5 * rho_ * VV_)) - Fav5 / insqr) / Jf;
  *Fav3_out = Fav3 / insqr;
  *Pavg_out = _P + P_;
}

void cfo_flux(const cfo_grid* g, int dir, double* const U_[5], const double* J, const double* Jf,
              double* const mf[9], double* const inF[5]) {
  dims_t d = dims(g);
  int f0, f1;
  face_range(&d, dir, &f0, &f1);
  size_t sI = (size_t)d.Ym * d.Zm, sJ = (size_t)d.Zm, sK = 1;
  /* the two transverse loops: cells (a,b) of the face's plane */
  int a0, a1, b0, b1;
  size_t stride, sa, sb;
  if (dir == 0) { a0 = 2; a1 = d.ny; b0 = 2; b1 = d.nz; stride = sI; sa = sJ; sb = sK; }
  else if (dir == 1) { a0 = 3; a1 = d.gend; b0 = 2; b1 = d.nz; stride = sJ; sa = sI; sb = sK; }
  else { a0 = 3; a1 = d.gend; b0 = 2; b1 = d.ny; stride = sK; sa = sI; sb = sJ; }
  const int mrow = 3 * dir; /* xi, eta or zeta row of the metric */
  for (int f = f0; f <= f1; f++) {
    const recipe_t* rl = recipe_left(&d, dir, f);
    const recipe_t* rr = recipe_right(&d, dir, f);
    for (int a = a0; a <= a1; a++)
      for (int b = b0; b <= b1; b++) {
        size_t base = (size_t)a * sa + (size_t)b * sb;
        double L[5], Rr[5], F[5], fav3, psum;
        reconstruct(rl, U_, J, base, stride, f, L);
        reconstruct(rr, U_, J, base, stride, f, Rr);
        /* face arrays carry the transverse cell indices minus one */
        size_t o = (size_t)f * stride + (size_t)(a - 1) * sa + (size_t)(b - 1) * sb;
        double n[3] = {mf[mrow][o], mf[mrow + 1][o], mf[mrow + 2][o]};
        roe_flux(L, Rr, n, Jf[o], dir == 1, F, &fav3, &psum);
        if (dir == 1 && (f == 1 || f == d.ny)) {
          /* impermeable wall faces: only the wall-normal momentum flux
           * (pressure + dissipation) survives (Flux_Y.cpp:452-460) */
          F[0] = 0;
          F[1] = 0;
          F[2] = 0.5 * (n[1] * psum - fav3) / Jf[o];
          F[3] = 0;
          F[4] = 0;
        }
        for (int v = 0; v < 5; v++) inF[v][o] = F[v];
      }
  }
}

/* ------------------------------------------------------------------------ */
/* viscous terms: Viscous_terms.cpp:212-883                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
  double U, V, W, T;
} prim_t;

static prim_t prim_visc(double* const U_[5], const double* J, size_t p) {
  /* Viscous_terms.cpp:261-395: rho here is rho/J */
  prim_t q;
  double rho = U_[0][p];
  q.U = U_[1][p] / rho;
  q.V = U_[2][p] / rho;
  q.W = U_[3][p] / rho;
  double VV = q.U * q.U + q.V * q.V + q.W * q.W;
  double P = (U_[4][p] - 0.5 * rho * VV) * (KG - 1) * J[p];
  q.T = P / (rho * RG) / J[p];
  return q;
}

/* derivative set of one face: d/dxi, d/deta, d/dzeta of u,v,w, of the
 * products uu,vv,ww,uv,vw,uw and of T */
typedef struct {
  double u, v, w, uu, vv, ww, uv, vw, uw, T;
} dset_t;

static dset_t d_two(prim_t a, prim_t b, double inv) {
  /* (a - b) * inv, a = face's upper cell */
  dset_t r;
  r.u = (a.U - b.U) * inv;
  r.v = (a.V - b.V) * inv;
  r.w = (a.W - b.W) * inv;
  r.uu = (a.U * a.U - b.U * b.U) * inv;
  r.vv = (a.V * a.V - b.V * b.V) * inv;
  r.ww = (a.W * a.W - b.W * b.W) * inv;
  r.uv = (a.U * a.V - b.U * b.V) * inv;
  r.vw = (a.V * a.W - b.V * b.W) * inv;
  r.uw = (a.U * a.W - b.U * b.W) * inv;
  r.T = (a.T - b.T) * inv;
  return r;
}

static dset_t d_four(prim_t p1, prim_t p2, prim_t m1, prim_t m2, double inv4) {
  /* (p1 + p2 - m1 - m2) * inv4: the tangential derivative averaged over the
   * two cells of the face */
  dset_t r;
  r.u = (p1.U + p2.U - m1.U - m2.U) * inv4;
  r.v = (p1.V + p2.V - m1.V - m2.V) * inv4;
  r.w = (p1.W + p2.W - m1.W - m2.W) * inv4;
  r.uu = (p1.U * p1.U + p2.U * p2.U - m1.U * m1.U - m2.U * m2.U) * inv4;
  r.vv = (p1.V * p1.V + p2.V * p2.V - m1.V * m1.V - m2.V * m2.V) * inv4;
  r.ww = (p1.W * p1.W + p2.W * p2.W - m1.W * m1.W - m2.W * m2.W) * inv4;
  r.uv = (p1.U * p1.V + p2.U * p2.V - m1.U * m1.V - m2.U * m2.V) * inv4;
  r.vw = (p1.V * p1.W + p2.V * p2.W - m1.V * m1.W - m2.V * m2.W) * inv4;
  r.uw = (p1.U * p1.W + p2.U * p2.W - m1.U * m1.W - m2.U * m2.W) * inv4;
  r.T = (p1.T + p2.T - m1.T - m2.T) * inv4;
  return r;
}

/* metric coefficient groups of one face (Viscous_terms.cpp:471-526) */
typedef struct {
  double a[8], b[8], c[8], d[11], e[11], f[11]; /* 1-based like the reference */
} coef_t;

static void coefs(const double xi[3], const double et[3], const double zt[3], coef_t* q) {
  double xix = xi[0], xiy = xi[1], xiz = xi[2];
  double etx = et[0], ety = et[1], etz = et[2];
  double ztx = zt[0], zty = zt[1], ztz = zt[2];
  q->a[1] = (4. / 3) * xix * xix + xiy * xiy + xiz * xiz;
  q->a[2] = xix * xix + (4. / 3) * xiy * xiy + xiz * xiz;
  q->a[3] = xix * xix + xiy * xiy + (4. / 3) * xiz * xiz;
  q->a[4] = xix * xix + xiy * xiy + xiz * xiz;
  q->a[5] = (1. / 3) * xix * xiy;
  q->a[6] = (1. / 3) * xiy * xiz;
  q->a[7] = (1. / 3) * xix * xiz;
  q->b[1] = (4. / 3) * etx * etx + ety * ety + etz * etz;
  q->b[2] = etx * etx + (4. / 3) * ety * ety + etz * etz;
  q->b[3] = etx * etx + ety * ety + (4. / 3) * etz * etz;
  q->b[4] = etx * etx + ety * ety + etz * etz;
  q->b[5] = (1. / 3) * etx * ety;
  q->b[6] = (1. / 3) * ety * etz;
  q->b[7] = (1. / 3) * etx * etz;
  q->c[1] = (4. / 3) * ztx * ztx + zty * zty + ztz * ztz;
  q->c[2] = ztx * ztx + (4. / 3) * zty * zty + ztz * ztz;
  q->c[3] = ztx * ztx + zty * zty + (4. / 3) * ztz * ztz;
  q->c[4] = ztx * ztx + zty * zty + ztz * ztz;
  q->c[5] = (1. / 3) * ztx * zty;
  q->c[6] = (1. / 3) * zty * ztz;
  q->c[7] = (1. / 3) * ztx * ztz;
  /* mixed groups share one pattern: g(p,q) for metric rows p and q */
  const double* P[3] = {xi, xi, et};
  const double* Q[3] = {et, zt, zt};
  double* G[3] = {q->d, q->e, q->f};
  for (int s = 0; s < 3; s++) {
    double px = P[s][0], py = P[s][1], pz = P[s][2];
    double qx = Q[s][0], qy = Q[s][1], qz = Q[s][2];
    double* o = G[s];
    o[1] = (4. / 3) * px * qx + py * qy + pz * qz;
    o[2] = px * qx + (4. / 3) * py * qy + pz * qz;
    o[3] = px * qx + py * qy + (4. / 3) * pz * qz;
    o[4] = px * qx + py * qy + pz * qz;
    o[5] = px * qy - (2. / 3) * py * qx;
    o[6] = px * qz - (2. / 3) * pz * qx;
    o[7] = py * qx - (2. / 3) * px * qy;
    o[8] = py * qz - (2. / 3) * pz * qy;
    o[9] = pz * qx - (2. / 3) * px * qz;
    o[10] = pz * qy - (2. / 3) * py * qz;
  }
}

static void load_face_metric(double* const mf[9], size_t o, double xi[3], double et[3], double zt[3]) {
  for (int c = 0; c < 3; c++) {
    xi[c] = mf[c][o];
    et[c] = mf[3 + c][o];
    zt[c] = mf[6 + c][o];
  }
}

#define CVK_PR(g4, dT) (CV * KG * (g4) * (dT) / (PR_L))

void cfo_viscous(const cfo_grid* g, double* const U_[5], const double* J, const double* Ju,
                 double* const mu[9], const double* Jv, double* const mv[9], const double* Jw,
                 double* const mw[9], double* const vF[4]) {
  dims_t d = dims(g);
  int gend = d.gend, ny = d.ny, nz = d.nz;
  double invXI = 1. / (g->deltaXI), invET = 1. / (g->deltaET), invZT = 1. / (g->deltaZT);
  double inv4XI = 1. / (4 * g->deltaXI), inv4ET = 1. / (4 * g->deltaET), inv4ZT = 1. / (4 * g->deltaZT);
  size_t N = (size_t)(g->X_out + 6) * d.Ym * d.Zm;
  /* face fluxes: LL on the x face (i-1,i), ML on the y face (j-1,j), NL on
   * the z face (k-1,k), all stored at the upper cell (i,j,k) */
  double* LL[4];
  double* ML[4];
  double* NL[4];
  for (int v = 0; v < 4; v++) {
    LL[v] = (double*)calloc(N, sizeof(double));
    ML[v] = (double*)calloc(N, sizeof(double));
    NL[v] = (double*)calloc(N, sizeof(double));
  }
  size_t sI = (size_t)d.Ym * d.Zm, sJ = (size_t)d.Zm;
  for (int i = 3; i <= gend + 1; i++)
    for (int j = 2; j <= ny + 1; j++)
      for (int k = 2; k <= nz + 1; k++) {
        size_t c = IX(d, i, j, k);
        size_t o = IX(d, i - 1, j - 1, k - 1); /* face metric index */
        prim_t P0 = prim_visc(U_, J, c);
        /* x face between i-1 and i: needed for i in [3,gend+1], j,k interior */
        if (j <= ny && k <= nz) {
          prim_t iP = prim_visc(U_, J, c - sI);
          prim_t Pj = prim_visc(U_, J, c + sJ), iPj = prim_visc(U_, J, c - sI + sJ);
          prim_t jP = prim_visc(U_, J, c - sJ), ijP = prim_visc(U_, J, c - sI - sJ);
          prim_t Pk = prim_visc(U_, J, c + 1), iPk = prim_visc(U_, J, c - sI + 1);
          prim_t kP = prim_visc(U_, J, c - 1), ikP = prim_visc(U_, J, c - sI - 1);
          dset_t dx = d_two(P0, iP, invXI);
          dset_t dy = d_four(Pj, iPj, jP, ijP, inv4ET);
          dset_t dz = d_four(Pk, iPk, kP, ikP, inv4ZT);
          double xi[3], et[3], zt[3];
          load_face_metric(mu, o, xi, et, zt);
          coef_t q;
          coefs(xi, et, zt, &q);
          double Ux = 0.5 * (P0.U + iP.U), Vx = 0.5 * (P0.V + iP.V), Wx = 0.5 * (P0.W + iP.W);
          double Jf = Ju[o];
          const double *a = q.a, *dd = q.d, *e = q.e;
          LL[0][c] = MU_L * (a[1] * dx.u + a[5] * dx.v + a[7] * dx.w + dd[1] * dy.u + dd[7] * dy.v + dd[9] * dy.w +
                             e[1] * dz.u + e[7] * dz.v + e[9] * dz.w) / Jf;
          LL[1][c] = MU_L * (a[5] * dx.u + a[2] * dx.v + a[6] * dx.w + dd[5] * dy.u + dd[2] * dy.v + dd[10] * dy.w +
                             e[5] * dz.u + e[2] * dz.v + e[10] * dz.w) / Jf;
          LL[2][c] = MU_L * (a[7] * dx.u + a[6] * dx.v + a[3] * dx.w + dd[6] * dy.u + dd[8] * dy.v + dd[3] * dy.w +
                             e[6] * dz.u + e[8] * dz.v + e[3] * dz.w) / Jf;
          LL[3][c] = MU_L * (0.5 * a[1] * dx.uu + 0.5 * a[2] * dx.vv + 0.5 * a[3] * dx.ww + a[5] * dx.uv +
                             a[6] * dx.vw + a[7] * dx.uw + CVK_PR(a[4], dx.T) +
                             0.5 * dd[1] * dy.uu + 0.5 * dd[2] * dy.vv + 0.5 * dd[3] * dy.ww + dd[5] * Vx * dy.u +
                             dd[6] * Wx * dy.u + dd[7] * Ux * dy.v + dd[8] * Wx * dy.v + dd[9] * Ux * dy.w +
                             dd[10] * Vx * dy.w + CVK_PR(dd[4], dy.T) +
                             0.5 * e[1] * dz.uu + 0.5 * e[2] * dz.vv + 0.5 * e[3] * dz.ww + e[5] * Vx * dz.u +
                             e[6] * Wx * dz.u + e[7] * Ux * dz.v + e[8] * Wx * dz.v + e[9] * Ux * dz.w +
                             e[10] * Vx * dz.w + CVK_PR(e[4], dz.T)) / Jf;
        }
        /* y face between j-1 and j: i interior, j in [2,ny+1], k interior */
        if (i <= gend && k <= nz) {
          prim_t Pi = prim_visc(U_, J, c + sI), jPi = prim_visc(U_, J, c + sI - sJ);
          prim_t iP = prim_visc(U_, J, c - sI), ijP = prim_visc(U_, J, c - sI - sJ);
          prim_t jP = prim_visc(U_, J, c - sJ);
          prim_t Pk = prim_visc(U_, J, c + 1), jPk = prim_visc(U_, J, c - sJ + 1);
          prim_t kP = prim_visc(U_, J, c - 1), jkP = prim_visc(U_, J, c - sJ - 1);
          dset_t dx = d_four(Pi, jPi, iP, ijP, inv4XI);
          dset_t dy = d_two(P0, jP, invET);
          dset_t dz = d_four(Pk, jPk, kP, jkP, inv4ZT);
          double xi[3], et[3], zt[3];
          load_face_metric(mv, o, xi, et, zt);
          coef_t q;
          coefs(xi, et, zt, &q);
          double Uy = 0.5 * (P0.U + jP.U), Vy = 0.5 * (P0.V + jP.V), Wy = 0.5 * (P0.W + jP.W);
          double Jf = Jv[o];
          const double *b = q.b, *dd = q.d, *f = q.f;
          ML[0][c] = MU_L * (dd[1] * dx.u + dd[5] * dx.v + dd[6] * dx.w + b[1] * dy.u + b[5] * dy.v + b[7] * dy.w +
                             f[1] * dz.u + f[7] * dz.v + f[9] * dz.w) / Jf;
          ML[1][c] = MU_L * (dd[7] * dx.u + dd[2] * dx.v + dd[8] * dx.w + b[5] * dy.u + b[2] * dy.v + b[6] * dy.w +
                             f[5] * dz.u + f[2] * dz.v + f[10] * dz.w) / Jf;
          ML[2][c] = MU_L * (dd[9] * dx.u + dd[10] * dx.v + dd[3] * dx.w + b[7] * dy.u + b[6] * dy.v + b[3] * dy.w +
                             f[6] * dz.u + f[8] * dz.v + f[3] * dz.w) / Jf;
          ML[3][c] = MU_L * (0.5 * b[1] * dy.uu + 0.5 * b[2] * dy.vv + 0.5 * b[3] * dy.ww + b[5] * dy.uv +
                             b[6] * dy.vw + b[7] * dy.uw + CVK_PR(b[4], dy.T) +
                             0.5 * dd[1] * dx.uu + 0.5 * dd[2] * dx.vv + 0.5 * dd[3] * dx.ww + dd[5] * Uy * dx.v +
                             dd[6] * Uy * dx.w + dd[7] * Vy * dx.u + dd[8] * Vy * dx.w + dd[9] * Wy * dx.u +
                             dd[10] * Wy * dx.v + CVK_PR(dd[4], dx.T) +
                             0.5 * f[1] * dz.uu + 0.5 * f[2] * dz.vv + 0.5 * f[3] * dz.ww + f[5] * Vy * dz.u +
                             f[6] * Wy * dz.u + f[7] * Uy * dz.v + f[8] * Wy * dz.v + f[9] * Uy * dz.w +
                             f[10] * Vy * dz.w + CVK_PR(f[4], dz.T)) / Jf;
        }
        /* z face between k-1 and k: i, j interior, k in [2,nz+1] */
        if (i <= gend && j <= ny) {
          prim_t Pi = prim_visc(U_, J, c + sI), kPi = prim_visc(U_, J, c + sI - 1);
          prim_t iP = prim_visc(U_, J, c - sI), ikP = prim_visc(U_, J, c - sI - 1);
          prim_t Pj = prim_visc(U_, J, c + sJ), kPj = prim_visc(U_, J, c + sJ - 1);
          prim_t jP = prim_visc(U_, J, c - sJ), jkP = prim_visc(U_, J, c - sJ - 1);
          prim_t kP = prim_visc(U_, J, c - 1);
          prim_t iPj = prim_visc(U_, J, c - sI + sJ);
          dset_t dx = d_four(Pi, kPi, iP, ikP, inv4XI);
          dset_t dy = d_four(Pj, kPj, jP, jkP, inv4ET);
          /* Viscous_terms.cpp:728-730: the second product of duv/dvw/duw
           * d/deta takes its V or W from cell (i-1,j+1) */
          dy.uv = (Pj.U * Pj.V + kPj.U * iPj.V - jP.U * jP.V - jkP.U * jkP.V) * inv4ET;
          dy.vw = (Pj.V * Pj.W + kPj.V * iPj.W - jP.V * jP.W - jkP.V * jkP.W) * inv4ET;
          dy.uw = (Pj.U * Pj.W + kPj.U * iPj.W - jP.U * jP.W - jkP.U * jkP.W) * inv4ET;
          dset_t dz = d_two(P0, kP, invZT);
          double xi[3], et[3], zt[3];
          load_face_metric(mw, o, xi, et, zt);
          coef_t q;
          coefs(xi, et, zt, &q);
          double Uz = 0.5 * (P0.U + kP.U), Vz = 0.5 * (P0.V + kP.V), Wz = 0.5 * (P0.W + kP.W);
          double Jf = Jw[o];
          const double *cc = q.c, *e = q.e, *f = q.f;
          NL[0][c] = MU_L * (e[1] * dx.u + e[5] * dx.v + e[6] * dx.w + f[1] * dy.u + f[5] * dy.v + f[6] * dy.w +
                             cc[1] * dz.u + cc[5] * dz.v + cc[7] * dz.w) / Jf;
          NL[1][c] = MU_L * (e[7] * dx.u + e[2] * dx.v + e[8] * dx.w + f[7] * dy.u + f[2] * dy.v + f[8] * dy.w +
                             cc[5] * dz.u + cc[2] * dz.v + cc[6] * dz.w) / Jf;
          NL[2][c] = MU_L * (e[9] * dx.u + e[10] * dx.v + e[3] * dx.w + f[9] * dy.u + f[10] * dy.v + f[3] * dy.w +
                             cc[7] * dz.u + cc[6] * dz.v + cc[3] * dz.w) / Jf;
          NL[3][c] = MU_L * (0.5 * cc[1] * dz.uu + 0.5 * cc[2] * dz.vv + 0.5 * cc[3] * dz.ww + cc[5] * dz.uv +
                             cc[6] * dz.vw + cc[7] * dz.uw + CVK_PR(cc[4], dz.T) +
                             0.5 * e[1] * dx.uu + 0.5 * e[2] * dx.vv + 0.5 * e[3] * dx.ww + e[5] * Uz * dx.v +
                             e[6] * Uz * dx.w + e[7] * Vz * dx.u + e[8] * Vz * dx.w + e[9] * Wz * dx.u +
                             e[10] * Wz * dx.v + CVK_PR(e[4], dx.T) +
                             0.5 * f[1] * dy.uu + 0.5 * f[2] * dy.vv + 0.5 * f[3] * dy.ww + f[5] * Uz * dy.v +
                             f[6] * Uz * dy.w + f[7] * Vz * dy.u + f[8] * Vz * dy.w + f[9] * Wz * dy.u +
                             f[10] * Wz * dy.v + CVK_PR(f[4], dy.T)) / Jf;
        }
      }
  /* divergence (Viscous_terms.cpp:849-883) */
  for (int i = 3; i <= gend; i++)
    for (int j = 2; j <= ny; j++)
      for (int k = 2; k <= nz; k++) {
        size_t c = IX(d, i, j, k);
        for (int v = 0; v < 4; v++)
          vF[v][c] = (LL[v][c + sI] - LL[v][c]) * invXI + (ML[v][c + sJ] - ML[v][c]) * invET +
                     (NL[v][c + 1] - NL[v][c]) * invZT;
      }
  for (int v = 0; v < 4; v++) {
    free(LL[v]);
    free(ML[v]);
    free(NL[v]);
  }
}

/* ------------------------------------------------------------------------ */
/* dual time stepping: Dual_time_stepping.cpp                               */
/* ------------------------------------------------------------------------ */

typedef struct {
  double rho, u, v, w, VV, P, C, T, H, beta;
} pstate_t;

static pstate_t prim_dt(double* const U_[5], const double* J, size_t p) {
  /* Dual_time_stepping.cpp:343-372 */
  pstate_t s;
  s.rho = U_[0][p] * J[p];
  s.u = U_[1][p] / U_[0][p];
  s.v = U_[2][p] / U_[0][p];
  s.w = U_[3][p] / U_[0][p];
  s.VV = s.u * s.u + s.v * s.v + s.w * s.w;
  s.P = (U_[4][p] * J[p] - 0.5 * s.rho * s.VV) * (KG - 1);
  s.C = KG * s.P / s.rho;
  s.T = s.P / s.rho;
  s.H = 0.5 * s.VV + s.C / (KG - 1);
  s.beta = dmax(s.VV / s.C, EPS_PRE);
  return s;
}

/* spectral radius of the preconditioned flux Jacobian along metric row n */
static double lambda_dir(const pstate_t* s, const double n[3], double* Uc) {
  double U = n[0] * s->u + n[1] * s->v + n[2] * s->w;
  double S = sqrt(U * U * (s->beta - 1) * (s->beta - 1) + 4 * s->beta * s->C * (n[0] * n[0] + n[1] * n[1] + n[2] * n[2]));
  *Uc = U;
  return ((s->beta + 1) * fabs(U) + S) / 2;
}

/* diagonal D^-1 entries d11, d22(=d33=d44=d55), d51 at a cell
 * (Dual_time_stepping.cpp:464-489) */
static void dinv(const pstate_t* s, const double m9[9], const cfo_grid* g, double deltaT, double* d11, double* d22,
                 double* d51) {
  double U, V, W;
  double lx = lambda_dir(s, m9 + 0, &U);
  double ly = lambda_dir(s, m9 + 3, &V);
  double lz = lambda_dir(s, m9 + 6, &W);
  double lambda = lx / g->deltaXI + ly / g->deltaET + lz / g->deltaZT;
  double beta = s->beta;
  *d11 = 2 * deltaT / (3 * beta + 2 * lambda * deltaT);
  *d22 = 2 * deltaT / (3 + 2 * lambda * deltaT);
  double temp1 = 6 * (KG - 1) * (beta - 1) * deltaT;
  double temp2 = KG * (3 + 2 * lambda * deltaT) * (3 * beta + 2 * lambda * deltaT) * s->rho;
  *d51 = -temp1 / temp2;
}

/* split flux Jacobian of one neighbour along metric row n
 * (Dual_time_stepping.cpp:608-632): entries 11..55 of A+ and the diagonal
 * of A- (backward sweep, :968-972) */
typedef struct {
  double p11, p12, p13, p14, p21, p22, p31, p33, p41, p44, p51, p52, p53, p54, p55;
  double m11, m22, m33, m44, m55;
} split_t;

static split_t split_jac(const pstate_t* s, const double n[3]) {
  split_t a;
  double U;
  double lam = lambda_dir(s, n, &U);
  double dd = fabs(lam);
  double beta = s->beta, P = s->P, rho = s->rho, T = s->T;
  a.p11 = 0.5 * (beta * U + dd);
  a.p12 = 0.5 * P * KG * beta * n[0];
  a.p13 = 0.5 * P * KG * beta * n[1];
  a.p14 = 0.5 * P * KG * beta * n[2];
  a.p21 = 0.5 * n[0] / rho;
  a.p22 = 0.5 * (U + dd);
  a.p31 = 0.5 * n[1] / rho;
  a.p33 = 0.5 * (U + dd);
  a.p41 = 0.5 * n[2] / rho;
  a.p44 = 0.5 * (U + dd);
  a.p51 = 0.5 * U * (KG - 1) * (beta - 1) / rho / KG;
  a.p55 = 0.5 * (U + dd);
  a.m11 = 0.5 * (beta * U - dd);
  a.m22 = 0.5 * (U - dd);
  a.m33 = 0.5 * (U - dd);
  a.m44 = 0.5 * (U - dd);
  a.m55 = 0.5 * (U - dd);
  return a;
}

/* A+ entries 52..54 are written xi*beta*T*(K-1) in x (:624-626) and
 * beta*T*(K-1)*eta in y and z (:680-682, :737-739) */
static void split_p5(split_t* a, const pstate_t* s, const double n[3], int dir) {
  double beta = s->beta, T = s->T;
  if (dir == 0) {
    a->p52 = 0.5 * n[0] * beta * T * (KG - 1);
    a->p53 = 0.5 * n[1] * beta * T * (KG - 1);
    a->p54 = 0.5 * n[2] * beta * T * (KG - 1);
  } else {
    a->p52 = 0.5 * beta * T * (KG - 1) * n[0];
    a->p53 = 0.5 * beta * T * (KG - 1) * n[1];
    a->p54 = 0.5 * beta * T * (KG - 1) * n[2];
  }
}

static void load_m9(double* const m[9], size_t p, double o[9]) {
  for (int c = 0; c < 9; c++) o[c] = m[c][p];
}

typedef struct {
  const cfo_grid* g;
  dims_t d;
  double deltaT;
  double* const* U_;
  const double* J;
  double* const* m;
  double* const* MR;
  double* const* NL;
  double* const* NR;
} sweep_ctx;

static void fwd_cell(const sweep_ctx* x, int i, int j, int k) {
  const dims_t d = x->d;
  size_t c = IX(d, i, j, k);
  size_t nb[3] = {c - (size_t)d.Ym * d.Zm, c - (size_t)d.Zm, c - 1};
  double sum[3][5];
  for (int dir = 0; dir < 3; dir++) {
    pstate_t s = prim_dt(x->U_, x->J, nb[dir]);
    double n[3] = {x->m[3 * dir][nb[dir]], x->m[3 * dir + 1][nb[dir]], x->m[3 * dir + 2][nb[dir]]};
    split_t a = split_jac(&s, n);
    split_p5(&a, &s, n, dir);
    double N1 = x->NL[0][nb[dir]], N2 = x->NL[1][nb[dir]], N3 = x->NL[2][nb[dir]], N4 = x->NL[3][nb[dir]],
           N5 = x->NL[4][nb[dir]];
    sum[dir][0] = -a.p11 * N1 - a.p12 * N2 - a.p13 * N3 - a.p14 * N4;
    sum[dir][1] = -a.p21 * N1 - a.p22 * N2;
    sum[dir][2] = -a.p31 * N1 - a.p33 * N3;
    sum[dir][3] = -a.p41 * N1 - a.p44 * N4;
    sum[dir][4] = -a.p51 * N1 - a.p52 * N2 - a.p53 * N3 - a.p54 * N4 - a.p55 * N5;
  }
  double LLv[5];
  for (int v = 0; v < 5; v++)
    LLv[v] = sum[0][v] / x->g->deltaXI + sum[1][v] / x->g->deltaET + sum[2][v] / x->g->deltaZT;
  pstate_t s0 = prim_dt(x->U_, x->J, c);
  double m9[9], d11, d22, d51;
  load_m9(x->m, c, m9);
  dinv(&s0, m9, x->g, x->deltaT, &d11, &d22, &d51);
  x->NL[0][c] = d11 * (x->MR[0][c] - LLv[0]);
  x->NL[1][c] = d22 * (x->MR[1][c] - LLv[1]);
  x->NL[2][c] = d22 * (x->MR[2][c] - LLv[2]);
  x->NL[3][c] = d22 * (x->MR[3][c] - LLv[3]);
  x->NL[4][c] = d51 * (x->MR[0][c] - LLv[0]) + d22 * (x->MR[4][c] - LLv[4]);
}

static void bwd_cell(const sweep_ctx* x, int i, int j, int k) {
  const dims_t d = x->d;
  size_t c = IX(d, i, j, k);
  size_t nb[3] = {c + (size_t)d.Ym * d.Zm, c + (size_t)d.Zm, c + 1};
  double sum[3][5];
  for (int dir = 0; dir < 3; dir++) {
    pstate_t s = prim_dt(x->U_, x->J, nb[dir]);
    double n[3] = {x->m[3 * dir][nb[dir]], x->m[3 * dir + 1][nb[dir]], x->m[3 * dir + 2][nb[dir]]};
    split_t a = split_jac(&s, n);
    split_p5(&a, &s, n, dir);
    double N1 = x->NR[0][nb[dir]], N2 = x->NR[1][nb[dir]], N3 = x->NR[2][nb[dir]], N4 = x->NR[3][nb[dir]],
           N5 = x->NR[4][nb[dir]];
    sum[dir][0] = +a.m11 * N1 + a.p12 * N2 + a.p13 * N3 + a.p14 * N4;
    sum[dir][1] = +a.p21 * N1 + a.m22 * N2;
    sum[dir][2] = +a.p31 * N1 + a.m33 * N3;
    sum[dir][3] = +a.p41 * N1 + a.m44 * N4;
    sum[dir][4] = +a.p51 * N1 + a.p52 * N2 + a.p53 * N3 + a.p54 * N4 + a.m55 * N5;
  }
  double UUv[5];
  for (int v = 0; v < 5; v++)
    UUv[v] = sum[0][v] / x->g->deltaXI + sum[1][v] / x->g->deltaET + sum[2][v] / x->g->deltaZT;
  pstate_t s0 = prim_dt(x->U_, x->J, c);
  double m9[9], d11, d22, d51;
  load_m9(x->m, c, m9);
  dinv(&s0, m9, x->g, x->deltaT, &d11, &d22, &d51);
  x->NR[0][c] = x->NL[0][c] - d11 * UUv[0];
  x->NR[1][c] = x->NL[1][c] - d22 * UUv[1];
  x->NR[2][c] = x->NL[2][c] - d22 * UUv[2];
  x->NR[3][c] = x->NL[3][c] - d22 * UUv[3];
  x->NR[4][c] = x->NL[4][c] - (d51 * UUv[0] + d22 * UUv[4]);
}

/* DPLUSGS.cpp:573-1121: one point-Jacobi update of a cell from the previous
 * iterate NL on both sides, NR = D^-1 (MR - LL - UU).  The upper neighbours'
 * states and metrics are read at j+1 and k+1 (:874-887, :931-938); their NL
 * is read at j_ and k_, which the reference never assigns (:995-1051), and
 * is taken at j+1 and k+1 here. */
static void jacobi_cell(const sweep_ctx* x, int i, int j, int k) {
  const dims_t d = x->d;
  size_t c = IX(d, i, j, k);
  size_t lo[3] = {c - (size_t)d.Ym * d.Zm, c - (size_t)d.Zm, c - 1};
  size_t hi[3] = {c + (size_t)d.Ym * d.Zm, c + (size_t)d.Zm, c + 1};
  double lsum[3][5], usum[3][5];
  for (int dir = 0; dir < 3; dir++) {
    pstate_t s = prim_dt(x->U_, x->J, lo[dir]);
    double n[3] = {x->m[3 * dir][lo[dir]], x->m[3 * dir + 1][lo[dir]], x->m[3 * dir + 2][lo[dir]]};
    split_t a = split_jac(&s, n);
    split_p5(&a, &s, n, dir);
    double N1 = x->NL[0][lo[dir]], N2 = x->NL[1][lo[dir]], N3 = x->NL[2][lo[dir]], N4 = x->NL[3][lo[dir]],
           N5 = x->NL[4][lo[dir]];
    lsum[dir][0] = -a.p11 * N1 - a.p12 * N2 - a.p13 * N3 - a.p14 * N4;
    lsum[dir][1] = -a.p21 * N1 - a.p22 * N2;
    lsum[dir][2] = -a.p31 * N1 - a.p33 * N3;
    lsum[dir][3] = -a.p41 * N1 - a.p44 * N4;
    lsum[dir][4] = -a.p51 * N1 - a.p52 * N2 - a.p53 * N3 - a.p54 * N4 - a.p55 * N5;
  }
  for (int dir = 0; dir < 3; dir++) {
    pstate_t s = prim_dt(x->U_, x->J, hi[dir]);
    double n[3] = {x->m[3 * dir][hi[dir]], x->m[3 * dir + 1][hi[dir]], x->m[3 * dir + 2][hi[dir]]};
    split_t a = split_jac(&s, n);
    split_p5(&a, &s, n, dir);
    double N1 = x->NL[0][hi[dir]], N2 = x->NL[1][hi[dir]], N3 = x->NL[2][hi[dir]], N4 = x->NL[3][hi[dir]],
           N5 = x->NL[4][hi[dir]];
    usum[dir][0] = +a.m11 * N1 + a.p12 * N2 + a.p13 * N3 + a.p14 * N4;
    usum[dir][1] = +a.p21 * N1 + a.m22 * N2;
    usum[dir][2] = +a.p31 * N1 + a.m33 * N3;
    usum[dir][3] = +a.p41 * N1 + a.m44 * N4;
    usum[dir][4] = +a.p51 * N1 + a.p52 * N2 + a.p53 * N3 + a.p54 * N4 + a.m55 * N5;
  }
  double LLv[5], UUv[5];
  for (int v = 0; v < 5; v++) {
    LLv[v] = lsum[0][v] / x->g->deltaXI + lsum[1][v] / x->g->deltaET + lsum[2][v] / x->g->deltaZT;
    UUv[v] = usum[0][v] / x->g->deltaXI + usum[1][v] / x->g->deltaET + usum[2][v] / x->g->deltaZT;
  }
  pstate_t s0 = prim_dt(x->U_, x->J, c);
  double m9[9], d11, d22, d51;
  load_m9(x->m, c, m9);
  dinv(&s0, m9, x->g, x->deltaT, &d11, &d22, &d51);
  x->NR[0][c] = d11 * (x->MR[0][c] - LLv[0] - UUv[0]);
  x->NR[1][c] = d22 * (x->MR[1][c] - LLv[1] - UUv[1]);
  x->NR[2][c] = d22 * (x->MR[2][c] - LLv[2] - UUv[2]);
  x->NR[3][c] = d22 * (x->MR[3][c] - LLv[3] - UUv[3]);
  x->NR[4][c] = d51 * (x->MR[0][c] - LLv[0] - UUv[0]) + d22 * (x->MR[4][c] - LLv[4] - UUv[4]);
}

/* zero implicit coupling through every ghost layer
 * (Dual_time_stepping.cpp:195-316) */
static void zero_sweep_ghosts(dims_t d, double* const NL[5], double* const NR[5]) {
  int gend = d.gend, nyy = d.ny + 1, nzz = d.nz + 1;
  for (int v = 0; v < 5; v++) {
    for (int j = 1; j <= nyy; j++)
      for (int k = 1; k <= nzz; k++) {
        NL[v][IX(d, 2, j, k)] = NR[v][IX(d, 2, j, k)] = 0;
        NL[v][IX(d, gend + 1, j, k)] = NR[v][IX(d, gend + 1, j, k)] = 0;
      }
    for (int i = 2; i <= gend + 1; i++) {
      for (int k = 1; k <= nzz; k++) {
        NL[v][IX(d, i, 1, k)] = NR[v][IX(d, i, 1, k)] = 0;
        NL[v][IX(d, i, nyy, k)] = NR[v][IX(d, i, nyy, k)] = 0;
      }
      for (int j = 1; j <= nyy; j++) {
        NL[v][IX(d, i, j, 1)] = NR[v][IX(d, i, j, 1)] = 0;
        NL[v][IX(d, i, j, nzz)] = NR[v][IX(d, i, j, nzz)] = 0;
      }
    }
  }
}

/* NL = D^-1 MR at one cell (Dual_time_stepping.cpp:480-507); NR = NL for
 * the LU-SGS sweeps (mode 0, 1), left alone for the Jacobi sweeps (mode 2) */
static void init_nl_cell(const cfo_grid* g, dims_t d, double deltaT, double* const U_[5], const double* J,
                         double* const m[9], double* const MR[5], double* const NL[5], double* const NR[5],
                         size_t c, int mode) {
  pstate_t s = prim_dt(U_, J, c);
  double m9[9], e11, e22, e51;
  load_m9(m, c, m9);
  dinv(&s, m9, g, deltaT, &e11, &e22, &e51);
  NL[0][c] = e11 * MR[0][c];
  NL[1][c] = e22 * MR[1][c];
  NL[2][c] = e22 * MR[2][c];
  NL[3][c] = e22 * MR[3][c];
  NL[4][c] = e51 * MR[0][c] + e22 * MR[4][c];
  if (mode != 2)
    for (int q = 0; q < 5; q++) NR[q][c] = NL[q][c];
  (void)d;
}

/* symmetric Gauss-Seidel sweeps (Dual_time_stepping.cpp:514-1277): mode 0
 * lexicographic, mode 1 plane by plane; mode 2: the four Jacobi sweeps of
 * DPLUSGS.cpp:488-1140.  Each sweep ends with NL <- NR. */
static void run_sweeps(const sweep_ctx* x, int mode) {
  const dims_t d = x->d;
  const int gend = d.gend, ny = d.ny, nz = d.nz, mmax = gend + ny + nz;
  const int nsw = mode == 2 ? 4 : x->g->n_sweeps;
  double* const* NL = x->NL;
  double* const* NR = x->NR;
  for (int sweep = 1; sweep <= nsw; sweep++) {
    if (mode == 2) {
      for (int i = 3; i <= gend; i++)
        for (int j = 2; j <= ny; j++)
          for (int k = 2; k <= nz; k++) jacobi_cell(x, i, j, k);
    } else if (mode == 0) {
      for (int i = 3; i <= gend; i++)
        for (int j = 2; j <= ny; j++)
          for (int k = 2; k <= nz; k++) fwd_cell(x, i, j, k);
      for (int i = gend; i >= 3; i--)
        for (int j = ny; j >= 2; j--)
          for (int k = nz; k >= 2; k--) bwd_cell(x, i, j, k);
    } else {
      for (int mm = 7; mm <= mmax; mm++)
        for (int i = 3; i <= gend; i++)
          for (int j = 2; j <= ny; j++) {
            int k = mm - i - j;
            if (k >= 2 && k <= nz) fwd_cell(x, i, j, k);
          }
      for (int mm = mmax; mm >= 7; mm--)
        for (int i = 3; i <= gend; i++)
          for (int j = 2; j <= ny; j++) {
            int k = mm - i - j;
            if (k >= 2 && k <= nz) bwd_cell(x, i, j, k);
          }
    }
    for (int q = 0; q < 5; q++)
      for (int i = 3; i <= gend; i++)
        for (int j = 2; j <= ny; j++)
          for (int k = 2; k <= nz; k++) NL[q][IX(d, i, j, k)] = NR[q][IX(d, i, j, k)];
  }
}

static void dual_time_impl(const cfo_grid* g, double deltaT, double* const U_[5], double* const U[5],
                           double* const Uq[5], const double* J, double* const m[9], double* const inFx[5],
                           double* const inFy[5], double* const inFz[5], double* const vF[4],
                           double* const MR[5], double* const NL[5], double* const NR[5], double er[5],
                           int hyperplane) {
  dims_t d = dims(g);
  int gend = d.gend, ny = d.ny, nz = d.nz;
  size_t sI = (size_t)d.Ym * d.Zm, sJ = (size_t)d.Zm;
  zero_sweep_ghosts(d, NL, NR);
  /* unsteady residual, preconditioned primitive transform and D^-1 (:338-507) */
  for (int i = 3; i <= gend; i++)
    for (int j = 2; j <= ny; j++)
      for (int k = 2; k <= nz; k++) {
        size_t c = IX(d, i, j, k);
        size_t fx = IX(d, i, j - 1, k - 1), fxm = fx - sI;   /* inFx[i][_j][_k], [i-1] */
        size_t fy = IX(d, i - 1, j, k - 1), fym = fy - sJ;   /* inFy[i-1][j][_k], [_j] */
        size_t fz = IX(d, i - 1, j - 1, k), fzm = fz - 1;    /* inFz[i-1][_j][k], [_k] */
        pstate_t s = prim_dt(U_, J, c);
        double beta = s.beta, rho = s.rho, u = s.u, v = s.v, w = s.w, VV = s.VV, T = s.T, H = s.H;
        double temp = rho * KG;
        double d11 = -(H * (KG - 1) + VV - KG * (T + VV)) * beta;
        double d12 = -(KG - 1) * u * beta, d13 = -(KG - 1) * v * beta, d14 = -(KG - 1) * w * beta;
        double d15 = (KG - 1) * beta;
        double d21 = -u / rho, d22 = 1 / rho, d31 = -v / rho, d41 = -w / rho;
        double d51 = (KG - 1) * (VV + KG * T * beta + (KG - 1) * VV * beta + H * (-1 + beta - KG * beta)) / temp;
        double d52 = -(KG - 1) * u * (1 + (KG - 1) * beta) / temp;
        double d53 = -(KG - 1) * v * (1 + (KG - 1) * beta) / temp;
        double d54 = -(KG - 1) * w * (1 + (KG - 1) * beta) / temp;
        double d55 = (KG - 1) * (1 + (KG - 1) * beta) / temp;
        double Rk[5];
        for (int q = 0; q < 5; q++) {
          double Rp = -(3 * U_[q][c] - 4 * U[q][c] + Uq[q][c]) / (2 * deltaT);
          double Rf = -((inFx[q][fx] - inFx[q][fxm]) / g->deltaXI + (inFy[q][fy] - inFy[q][fym]) / g->deltaET +
                        (inFz[q][fz] - inFz[q][fzm]) / g->deltaZT);
          Rk[q] = Rp + Rf;
        }
        Rk[1] = Rk[1] + vF[0][c] + g->f_drive / J[c];
        Rk[2] = Rk[2] + vF[1][c];
        Rk[3] = Rk[3] + vF[2][c];
        Rk[4] = Rk[4] + vF[3][c] + g->f_drive * u / J[c];
        /* the zero entries of Gamma^-1 are kept as explicit products so that
         * -0/NaN propagate as in the reference */
        MR[0][c] = (d11 * Rk[0] + d12 * Rk[1] + d13 * Rk[2] + d14 * Rk[3] + d15 * Rk[4]);
        MR[1][c] = (d21 * Rk[0] + d22 * Rk[1] + 0 * Rk[2] + 0 * Rk[3] + 0 * Rk[4]);
        MR[2][c] = (d31 * Rk[0] + 0 * Rk[1] + d22 * Rk[2] + 0 * Rk[3] + 0 * Rk[4]);
        MR[3][c] = (d41 * Rk[0] + 0 * Rk[1] + 0 * Rk[2] + d22 * Rk[3] + 0 * Rk[4]);
        MR[4][c] = (d51 * Rk[0] + d52 * Rk[1] + d53 * Rk[2] + d54 * Rk[3] + d55 * Rk[4]);
        init_nl_cell(g, d, deltaT, U_, J, m, MR, NL, NR, c, hyperplane);
      }
  sweep_ctx x = {g, d, deltaT, U_, J, m, MR, NL, NR};
  run_sweeps(&x, hyperplane);
  /* primitive update and residual norms (:1290-1352) */
  double e[5] = {0, 0, 0, 0, 0};
  double DN = 1. / ((double)g->X_out * g->Y_out * g->Z_out);
  for (int i = 3; i <= gend; i++)
    for (int j = 2; j <= ny; j++)
      for (int k = 2; k <= nz; k++) {
        size_t c = IX(d, i, j, k);
        double rhoold = U_[0][c] * J[c];
        double Uold = U_[1][c] / U_[0][c];
        double Vold = U_[2][c] / U_[0][c];
        double Wold = U_[3][c] / U_[0][c];
        double VVold = Uold * Uold + Vold * Vold + Wold * Wold;
        double Pold = (U_[4][c] * J[c] - 0.5 * rhoold * VVold) * (KG - 1);
        double Told = Pold / rhoold;
        double P = Pold + NR[0][c] * J[c];
        double Uu = Uold + NR[1][c] * J[c];
        double V = Vold + NR[2][c] * J[c];
        double W = Wold + NR[3][c] * J[c];
        double T = Told + NR[4][c] * J[c];
        double rho = P / T;
        U_[0][c] = rho / J[c];
        U_[1][c] = rho * Uu / J[c];
        U_[2][c] = rho * V / J[c];
        U_[3][c] = rho * W / J[c];
        U_[4][c] = (P / (KG - 1) + 0.5 * rho * (Uu * Uu + V * V + W * W)) / J[c];
        double r[5] = {(P - Pold) * (P - Pold), (Uu - Uold) * (Uu - Uold), (V - Vold) * (V - Vold),
                       (W - Wold) * (W - Wold), (T - Told) * (T - Told)};
        for (int q = 0; q < 5; q++) e[q] = e[q] + r[q] * r[q];
      }
  for (int q = 0; q < 5; q++) er[q] = sqrt(e[q]) * DN;
  (void)sJ;
}

void cfo_dual_time(const cfo_grid* g, double deltaT, double* const U_[5], double* const U[5],
                   double* const Uq[5], const double* J, double* const m[9], double* const inFx[5],
                   double* const inFy[5], double* const inFz[5], double* const vF[4], double* const MR[5],
                   double* const NL[5], double* const NR[5], double er[5]) {
  dual_time_impl(g, deltaT, U_, U, Uq, J, m, inFx, inFy, inFz, vF, MR, NL, NR, er, 0);
}

void cfo_dual_time_hyperplane(const cfo_grid* g, double deltaT, double* const U_[5], double* const U[5],
                              double* const Uq[5], const double* J, double* const m[9], double* const inFx[5],
                              double* const inFy[5], double* const inFz[5], double* const vF[4],
                              double* const MR[5], double* const NL[5], double* const NR[5], double er[5]) {
  dual_time_impl(g, deltaT, U_, U, Uq, J, m, inFx, inFy, inFz, vF, MR, NL, NR, er, 1);
}

void cfo_dplusgs(const cfo_grid* g, double deltaT, double* const U_[5], double* const U[5], double* const Uq[5],
                 const double* J, double* const m[9], double* const inFx[5], double* const inFy[5],
                 double* const inFz[5], double* const vF[4], double* const MR[5], double* const NL[5],
                 double* const NR[5], double er[5]) {
  dual_time_impl(g, deltaT, U_, U, Uq, J, m, inFx, inFy, inFz, vF, MR, NL, NR, er, 2);
}

/* RK3NODT.cpp:115-183: one stage RK (1..3) of the explicit three-stage
 * update U_ <- c1 U + c2 U_ + c3 dt Rk, Rk the flux divergence plus viscous
 * terms and the driving force; MR receives dt Rk. */
void cfo_rk3nodt(const cfo_grid* g, int RK, double deltaT, double* const U_[5], double* const U[5], const double* J,
                 double* const inFx[5], double* const inFy[5], double* const inFz[5], double* const vF[4],
                 double* const MR[5]) {
  const double c1[3] = {1.0, 0.75, 1.0 / 3.0};
  const double c2[3] = {0.0, 0.25, 2.0 / 3.0};
  const double c3[3] = {1.0, 0.25, 2.0 / 3.0};
  dims_t d = dims(g);
  size_t sI = (size_t)d.Ym * d.Zm, sJ = (size_t)d.Zm;
  for (int i = 3; i <= d.gend; i++)
    for (int j = 2; j <= d.ny; j++)
      for (int k = 2; k <= d.nz; k++) {
        size_t c = IX(d, i, j, k);
        size_t fx = IX(d, i, j - 1, k - 1), fxm = fx - sI;  /* inFx[i][j-1][_k], [i-1] */
        size_t fy = IX(d, i - 1, j, k - 1), fym = fy - sJ;  /* inFy[i-1][j][_k], [j-1] */
        size_t fz = IX(d, i - 1, j - 1, k), fzm = fz - 1;   /* inFz[i-1][j-1][k], [_k] */
        double tmp = U_[1][c] / U_[0][c];
        double Rk[5];
        for (int q = 0; q < 5; q++)
          Rk[q] = -((inFx[q][fx] - inFx[q][fxm]) / g->deltaXI + (inFy[q][fy] - inFy[q][fym]) / g->deltaET +
                    (inFz[q][fz] - inFz[q][fzm]) / g->deltaZT);
        Rk[1] = Rk[1] + vF[0][c] + g->f_drive / J[c];
        Rk[2] = Rk[2] + vF[1][c];
        Rk[3] = Rk[3] + vF[2][c];
        Rk[4] = Rk[4] + vF[3][c] + g->f_drive * tmp / J[c];
        for (int q = 0; q < 5; q++) {
          MR[q][c] = deltaT * Rk[q];
          U_[q][c] = c1[RK - 1] * U[q][c] + c2[RK - 1] * U_[q][c] + c3[RK - 1] * MR[q][c];
        }
      }
  (void)sJ;
}

/* Driven_force.cpp:38-73: bulk flux Qall = sum over the interior of
 * (U2_ J) 1/J / Jsum, Jsum = sum_j 1/J (at i = k = 3), divided by X_out and
 * Z_out; ER_Q = |Q0 - Qall| / Q0.  out = {Qall, ER_Q}. */
void cfo_driven_force(const cfo_grid* g, const double* U2_, const double* J, double Q0, double out[2]) {
  dims_t d = dims(g);
  double Jsum = 0, Q = 0;
  for (int j = 2; j <= d.ny; j++) Jsum = Jsum + 1. / J[IX(d, 3, j, 3)];
  for (int i = 3; i <= d.gend; i++)
    for (int j = 2; j <= d.ny; j++)
      for (int k = 2; k <= d.nz; k++) {
        size_t c = IX(d, i, j, k);
        double temp = (U2_[c] * J[c]) * 1. / J[c] / Jsum;
        Q = Q + temp;
      }
  Q = Q / g->X_out / g->Z_out;
  out[0] = Q;
  out[1] = fabs(Q0 - Q) / Q0;
}

/* Statistic.cpp:1043-1071: adds one state's sums over i and k of
 * U, V, W, UU, VV, WW, UV (velocities U_2/U_1 ...) to sums[q * Y_m + j],
 * q = 0..6, in the reference's order (j, then i, then k). */
void cfo_statistic_accum(const cfo_grid* g, double* const U_[5], double* sums) {
  dims_t d = dims(g);
  for (int j = 2; j <= d.ny; j++)
    for (int i = 3; i <= d.gend; i++)
      for (int k = 2; k <= d.nz; k++) {
        size_t c = IX(d, i, j, k);
        double U = U_[1][c] / U_[0][c];
        double V = U_[2][c] / U_[0][c];
        double W = U_[3][c] / U_[0][c];
        sums[0 * d.Ym + j] = sums[0 * d.Ym + j] + U;
        sums[1 * d.Ym + j] = sums[1 * d.Ym + j] + V;
        sums[2 * d.Ym + j] = sums[2 * d.Ym + j] + W;
        sums[3 * d.Ym + j] = sums[3 * d.Ym + j] + U * U;
        sums[4 * d.Ym + j] = sums[4 * d.Ym + j] + V * V;
        sums[5 * d.Ym + j] = sums[5 * d.Ym + j] + W * W;
        sums[6 * d.Ym + j] = sums[6 * d.Ym + j] + U * V;
      }
}

/* Statistic.cpp:552-828, the part behind Shear_stress.dat: the y-face
 * viscous flux of x-momentum ML2 (Statistic.cpp:818-824) at the bottom wall
 * face (j = 2, between the wall ghost and the first cell: SXYm[1]) and the
 * top one (j = ny + 1: SXYm[ny]), for i = 3..gend and k = 2..nz, into
 * out[(w * X_out + i - 3) * Z_out + k - 2].  Velocities are U_2/U_1 ...
 * (Statistic.cpp:603-712), derivatives the 4-point averages along x and z
 * and the 2-point difference along y (:734-786), metrics those of the y
 * faces at [i-1][j-1][k-1] (:716-724). */
void cfo_wall_shear(const cfo_grid* g, double* const U_[5], const double* Jv, double* const mv[9], double* out) {
  dims_t d = dims(g);
  const double inv4XI = 1. / (4 * g->deltaXI), invET = 1. / (g->deltaET), inv4ZT = 1. / (4 * g->deltaZT);
  const int X = d.gend - 2, Z = d.nz - 1;
  for (int w = 0; w < 2; w++) {
    const int j = w == 0 ? 2 : d.ny + 1;
    for (int i = 3; i <= d.gend; i++)
      for (int k = 2; k <= d.nz; k++) {
        double u[3][3][3], v[3][3][3], wv[3][3][3]; /* [di+1][dj+1][dk+1], dj in {-1, 0} */
        const int off[10][3] = {{0, 0, 0},  {-1, 0, 0}, {1, 0, 0},  {0, -1, 0}, {-1, -1, 0},
                                {1, -1, 0}, {0, 0, 1},  {0, 0, -1}, {0, -1, 1}, {0, -1, -1}};
        for (int q = 0; q < 10; q++) {
          const size_t c = IX(d, i + off[q][0], j + off[q][1], k + off[q][2]);
          const double rho = U_[0][c];
          u[off[q][0] + 1][off[q][1] + 1][off[q][2] + 1] = U_[1][c] / rho;
          v[off[q][0] + 1][off[q][1] + 1][off[q][2] + 1] = U_[2][c] / rho;
          wv[off[q][0] + 1][off[q][1] + 1][off[q][2] + 1] = U_[3][c] / rho;
        }
#define AT(a, di, dj, dk) a[(di) + 1][(dj) + 1][(dk) + 1]
        const double du_dx = (AT(u, 1, 0, 0) + AT(u, 1, -1, 0) - AT(u, -1, 0, 0) - AT(u, -1, -1, 0)) * inv4XI;
        const double dv_dx = (AT(v, 1, 0, 0) + AT(v, 1, -1, 0) - AT(v, -1, 0, 0) - AT(v, -1, -1, 0)) * inv4XI;
        const double dw_dx = (AT(wv, 1, 0, 0) + AT(wv, 1, -1, 0) - AT(wv, -1, 0, 0) - AT(wv, -1, -1, 0)) * inv4XI;
        const double du_dy = (AT(u, 0, 0, 0) - AT(u, 0, -1, 0)) * invET;
        const double dv_dy = (AT(v, 0, 0, 0) - AT(v, 0, -1, 0)) * invET;
        const double dw_dy = (AT(wv, 0, 0, 0) - AT(wv, 0, -1, 0)) * invET;
        const double du_dz = (AT(u, 0, 0, 1) + AT(u, 0, -1, 1) - AT(u, 0, 0, -1) - AT(u, 0, -1, -1)) * inv4ZT;
        const double dv_dz = (AT(v, 0, 0, 1) + AT(v, 0, -1, 1) - AT(v, 0, 0, -1) - AT(v, 0, -1, -1)) * inv4ZT;
        const double dw_dz = (AT(wv, 0, 0, 1) + AT(wv, 0, -1, 1) - AT(wv, 0, 0, -1) - AT(wv, 0, -1, -1)) * inv4ZT;
#undef AT
        const size_t f = IX(d, i - 1, j - 1, k - 1);
        const double xix = mv[0][f], xiy = mv[1][f], xiz = mv[2][f], etx = mv[3][f], ety = mv[4][f],
                     etz = mv[5][f], ztx = mv[6][f], zty = mv[7][f], ztz = mv[8][f];
        const double b1 = (4. / 3) * etx * etx + ety * ety + etz * etz;
        const double b5 = (1. / 3) * etx * ety;
        const double b7 = (1. / 3) * etx * etz;
        const double d1 = (4. / 3) * xix * etx + xiy * ety + xiz * etz;
        const double d5 = xix * ety - (2. / 3) * xiy * etx;
        const double d6 = xix * etz - (2. / 3) * xiz * etx;
        const double f1 = (4. / 3) * etx * ztx + ety * zty + etz * ztz;
        const double f7 = ety * ztx - (2. / 3) * etx * zty;
        const double f9 = etz * ztx - (2. / 3) * etx * ztz;
        out[((size_t)w * X + i - 3) * Z + k - 2] =
            MU_L * (d1 * du_dx + d5 * dv_dx + d6 * dw_dx + b1 * du_dy + b5 * dv_dy + b7 * dw_dy + f1 * du_dz +
                    f7 * dv_dz + f9 * dw_dz) / Jv[f];
      }
  }
}

void cfo_sweeps_apply(const cfo_grid* g, double deltaT, int mode, double* const U_[5], const double* J,
                      double* const m[9], double* const MR[5], double* const NL[5], double* const NR[5]) {
  dims_t d = dims(g);
  zero_sweep_ghosts(d, NL, NR);
  for (int i = 3; i <= d.gend; i++)
    for (int j = 2; j <= d.ny; j++)
      for (int k = 2; k <= d.nz; k++) init_nl_cell(g, d, deltaT, U_, J, m, MR, NL, NR, IX(d, i, j, k), mode);
  sweep_ctx x = {g, d, deltaT, U_, J, m, MR, NL, NR};
  run_sweeps(&x, mode);
}

/* ------------------------------------------------------------------------ */
/* limiter fluxes: Flux_X_limiter.cpp, Flux_Y_limiter.cpp, Flux_Z_limiter.cpp */
/* ------------------------------------------------------------------------ */

/* the reference's min/max macros (Flux_X_limiter.cpp:13-14): (a<b)?a:b and
 * (a>b)?a:b, so a NaN argument in the second place is returned as is */
static double mac_min(double a, double b) { return a < b ? a : b; }
static double mac_max(double a, double b) { return a > b ? a : b; }

/* limited 3-point slopes of one cell (Flux_X_limiter.cpp:149-160):
 * lim3_up = value at the cell's upper face before the face-J scaling,
 * lim3_dn = value at its lower face.  sgn = +1 (x, y) or -1 (z, B9). */
static double lim3_up(double um, double u, double up, double sgn) {
  double rL = (up - u) / (u - um);
  double bL = 1. / 3 * (1 + 2 * rL);
  double m = mac_min(2, mac_min(2 * rL, bL));
  return sgn > 0 ? u + 0.5 * mac_max(0, m) * (u - um) : u - 0.5 * mac_max(0, m) * (u - um);
}
static double lim3_dn(double um, double u, double up) {
  double rR = (u - um) / (up - u);
  double bR = 1. / 3 * (1 + 2 * rR);
  double m = mac_min(2, mac_min(2 * rR, bR));
  return u - 0.5 * mac_max(0, m) * (up - u);
}
/* limited 5-point slopes (Flux_X_limiter.cpp:250-265) */
static double lim5_up(double umm, double um, double u, double up, double upp, double sgn) {
  double irL = (u - um) / (um - umm);
  double rL = (up - u) / (u - um);
  double rLi = (upp - up) / (up - u);
  double bL = 1. / 30 * (-2. / irL + 11 + 24 * rL - 3 * rL * rLi);
  double m = mac_min(2, mac_min(2 * rL, bL));
  return sgn > 0 ? u + 0.5 * mac_max(0, m) * (u - um) : u - 0.5 * mac_max(0, m) * (u - um);
}
static double lim5_dn(double umm, double um, double u, double up, double upp) {
  double rRi = (up - u) / (upp - up);
  double rR = (u - um) / (up - u);
  double irR = (um - umm) / (u - um);
  double bR = 1. / 30 * (-2. / rRi + 11 + 24 * rR - 3 * rR * irR);
  double m = mac_min(2, mac_min(2 * rR, bR));
  return u - 0.5 * mac_max(0, m) * (up - u);
}
/* y, near-wall density (Flux_Y_limiter.cpp:146-168): 1e-7 added to the
 * denominators and the cell value scaled by J while the slope is not (B8) */
static double lim3y_up(double um, double u, double up, double Jc) {
  double rL = (up - u) / ((u - um) + 0.0000001);
  double bL = 1. / 3 * (1 + 2 * rL);
  double m = mac_min(2, mac_min(2 * rL, bL));
  return u * Jc + 0.5 * mac_max(0, m) * (u - um);
}
static double lim3y_dn(double um, double u, double up, double Jc) {
  double rR = (u - um) / ((up - u) + 0.0000001);
  double bR = 1. / 3 * (1 + 2 * rR);
  double m = mac_min(2, mac_min(2 * rR, bR));
  return u * Jc - 0.5 * mac_max(0, m) * (up - u);
}

/* preconditioned mean-state dissipation and flux of one face
 * (Flux_X_limiter.cpp:361-531): n = the face's metric row along the flux
 * direction, Ep = sensor, Jf = face Jacobian. */
static void limiter_face_flux(const double L[5], const double Rr[5], const double n[3], double Jf, double Ep,
                              double F[5], double* Fav3_out, double* Psum_out) {
  const double K = KG;
  double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double Nx = n[0] / nn, Ny = n[1] / nn, Nz = n[2] / nn;
  double _rho = L[0], _u = L[1] / _rho, _v = L[2] / _rho, _w = L[3] / _rho;
  double _U = n[0] * _u + n[1] * _v + n[2] * _w;
  double __U = Nx * _u + Ny * _v + Nz * _w;
  double _VV = _u * _u + _v * _v + _w * _w;
  double _P = (L[4] - 0.5 * _rho * _VV) * (K - 1);
  double _T = _P / _rho;
  double _C = K * _P / _rho;
  double _H = 0.5 * _VV + _C / (K - 1);
  double rho_ = Rr[0], u_ = Rr[1] / rho_, v_ = Rr[2] / rho_, w_ = Rr[3] / rho_;
  double U_ = n[0] * u_ + n[1] * v_ + n[2] * w_;
  double U__ = Nx * u_ + Ny * v_ + Nz * w_;
  double VV_ = u_ * u_ + v_ * v_ + w_ * w_;
  double P_ = (Rr[4] - 0.5 * rho_ * VV_) * (K - 1);
  double T_ = P_ / rho_;
  double C_ = K * P_ / rho_;
  double H_ = 0.5 * VV_ + C_ / (K - 1);
  double rho = 0.5 * (_rho + rho_);
  double u = 0.5 * (_u + u_), v = 0.5 * (_v + v_), w = 0.5 * (_w + w_);
  double U = 0.5 * (_U + U_);
  double _U_ = 0.5 * (__U + U__);
  double VV = u * u + v * v + w * w;
  double H = 0.5 * (_H + H_);
  double C = (H - 0.5 * VV) * (K - 1);
  double P = rho * C / K;
  double T = P / rho;
  double dU1 = P_ - _P, dU2 = u_ - _u, dU3 = v_ - _v, dU4 = w_ - _w, dU5 = T_ - _T;
  double beta = mac_max(VV / C, EPS_PRE);
  double S = sqrt(U * U * (beta - 1) * (beta - 1) + 4 * beta * C * (n[0] * n[0] + n[1] * n[1] + n[2] * n[2]));
  double _S_ = sqrt(_U_ * _U_ * (beta - 1) * (beta - 1) + 4 * beta * C * (Nx * Nx + Ny * Ny + Nz * Nz));
  double temp = 4 * K * _S_ * T * beta;
  double temp1 = S + U + U * beta;
  double temp2 = -S + U + U * beta;
  double temp3 = _S_ - _U_ + _U_ * beta;
  double temp4 = _S_ + _U_ - _U_ * beta;
  double temp5 = _S_ * _S_ - (beta - 1) * (beta - 1) * _U_ * _U_;
  double tu1 = 2 * T * Nx * K * beta + u * _S_ + u * (beta - 1) * _U_;
  double tu2 = 2 * T * Nx * K * beta - u * _S_ + u * (beta - 1) * _U_;
  double tv1 = 2 * T * Ny * K * beta + v * _S_ + v * (beta - 1) * _U_;
  double tv2 = 2 * T * Ny * K * beta - v * _S_ + v * (beta - 1) * _U_;
  double tw1 = 2 * T * Nz * K * beta + w * _S_ + w * (beta - 1) * _U_;
  double tw2 = 2 * T * Nz * K * beta - w * _S_ + w * (beta - 1) * _U_;
  double tuvw = 2 * T * K * beta * (u * Nx + v * Ny + w * Nz);
  double th1 = tuvw + H * _S_ + H * _U_ * beta - H * _U_;
  double th2 = tuvw - H * _S_ + H * _U_ * beta - H * _U_;
  double aU = fabs(U), a1 = fabs(temp1), a2 = fabs(temp2);
  double d11 = 1 / temp * (4 * (K - 1) * _S_ * beta * aU + temp3 * a1 + temp4 * a2);
  double d12 = -1 / (2 * temp) * (Nx * rho * (a2 - a1) * temp5);
  double d13 = -1 / (2 * temp) * (Ny * rho * (a2 - a1) * temp5);
  double d14 = -1 / (2 * temp) * (Nz * rho * (a2 - a1) * temp5);
  double d15 = -rho * aU / T;
  double d21 = 1 / temp * (u * _S_ * (4 * (K - 1) * beta * aU + a2 + a1) - (a2 - a1) * (2 * T * Nx * K * beta + u * _U_ * (beta - 1)));
  double d22 = 1 / (2 * temp) * (rho * (8 * T * K * _S_ * beta * (Ny * Ny + Nz * Nz) * aU + Nx * (temp3 * tu2 * a2 + temp4 * tu1 * a1)));
  double d23 = 1 / (2 * temp) * (rho * Ny * (-8 * T * K * _S_ * beta * Nx * aU + temp3 * tu2 * a2 + temp4 * tu1 * a1));
  double d24 = 1 / (2 * temp) * (rho * Nz * (-8 * T * K * _S_ * beta * Nx * aU + temp3 * tu2 * a2 + temp4 * tu1 * a1));
  double d25 = -u * rho * aU / T;
  double d31 = 1 / temp * (v * _S_ * (4 * (K - 1) * beta * aU + a2 + a1) - (a2 - a1) * (2 * T * Ny * K * beta + v * _U_ * (beta - 1)));
  double d32 = 1 / (2 * temp) * (rho * Nx * (-8 * T * K * _S_ * beta * Ny * aU + temp3 * tv2 * a2 + temp4 * tv1 * a1));
  double d33 = 1 / (2 * temp) * (rho * (8 * T * K * _S_ * beta * (Nx * Nx + Nz * Nz) * aU + Ny * (temp3 * tv2 * a2 + temp4 * tv1 * a1)));
  double d34 = 1 / (2 * temp) * (rho * Nz * (-8 * T * K * _S_ * beta * Ny * aU + temp3 * tv2 * a2 + temp4 * tv1 * a1));
  double d35 = -v * rho * aU / T;
  double d41 = 1 / temp * (w * _S_ * (4 * (K - 1) * beta * aU + a2 + a1) - (a2 - a1) * (2 * T * Nz * K * beta + w * _U_ * (beta - 1)));
  double d42 = 1 / (2 * temp) * (rho * Nx * (-8 * T * K * _S_ * beta * Nz * aU + temp3 * tw2 * a2 + temp4 * tw1 * a1));
  double d43 = 1 / (2 * temp) * (rho * Ny * (-8 * T * K * _S_ * beta * Nz * aU + temp3 * tw2 * a2 + temp4 * tw1 * a1));
  double d44 = 1 / (2 * temp) * (rho * (8 * T * K * _S_ * beta * (Nx * Nx + Ny * Ny) * aU + Nz * (temp3 * tw2 * a2 + temp4 * tw1 * a1)));
  double d45 = -w * rho * aU / T;
  double d51 = 1 / temp * (_S_ * (4 * beta * (H * K - H - T * K) * aU + H * (a2 + a1)) - (a2 - a1) * (tuvw + H * _U_ * beta - H * _U_));
  double d52 = 1 / (2 * temp) * (rho * (8 * T * K * _S_ * beta * (-v * Nx * Ny - w * Nx * Nz + u * (Ny * Ny + Nz * Nz) * aU) + Nx * (th2 * temp3 * a2 + th1 * temp4 * a1)));
  double d53 = 1 / (2 * temp) * (rho * (8 * T * K * _S_ * beta * (-u * Nx * Ny - w * Ny * Nz + v * (Nx * Nx + Nz * Nz) * aU) + Ny * (th2 * temp3 * a2 + th1 * temp4 * a1)));
  double d54 = 1 / (2 * temp) * (rho * (8 * T * K * _S_ * beta * (-u * Nx * Nz - v * Ny * Nz + w * (Nx * Nx + Ny * Ny) * aU) + Nz * (th2 * temp3 * a2 + th1 * temp4 * a1)));
  double d55 = (-H / T + K / (K - 1)) * rho * aU;
  double Fav1 = d11 * dU1 + d12 * dU2 + d13 * dU3 + d14 * dU4 + d15 * dU5;
  double Fav2 = d21 * dU1 + d22 * dU2 + d23 * dU3 + d24 * dU4 + d25 * dU5;
  double Fav3 = d31 * dU1 + d32 * dU2 + d33 * dU3 + d34 * dU4 + d35 * dU5;
  double Fav4 = d41 * dU1 + d42 * dU2 + d43 * dU3 + d44 * dU4 + d45 * dU5;
  double Fav5 = d51 * dU1 + d52 * dU2 + d53 * dU3 + d54 * dU4 + d55 * dU5;
  F[0] = 0.5 * ((_rho * _U + rho_ * U_) - Ep * Fav1) / Jf;
  F[1] = 0.5 * ((_rho * _u * _U + rho_ * u_ * U_ + n[0] * (_P + P_)) - Ep * Fav2) / Jf;
  F[2] = 0.5 * ((_rho * _v * _U + rho_ * v_ * U_ + n[1] * (_P + P_)) - Ep * Fav3) / Jf;
  F[3] = 0.5 * ((_rho * _w * _U + rho_ * w_ * U_ + n[2] * (_P + P_)) - Ep * Fav4) / Jf;
  F[4] = 0.5 * ((_U * (3.5 * _P + 0.5 * _rho * _VV) + U_ * (3.5 * P_ + 0.5 * rho_ * VV_)) - Ep * Fav5) / Jf;
  *Fav3_out = Fav3;
  *Psum_out = _P + P_;
}

/* left (ML) and right (MR) states of face f along dir for the transverse
 * cells at `base` (Flux_*_limiter.cpp MUSCL sections and seam overwrites).
 * Jfl/Jfr: face Jacobian used to scale the left/right value (x, y). */
static void limiter_states(const dims_t* d, int dir, double* const U_[5], const double* J, size_t base,
                           size_t stride, int f, double Jfl, double Jfr, double L[5], double Rr[5]) {
#define UV(v, c) U_[v][base + (size_t)(c) * stride]
#define JC(c) J[base + (size_t)(c) * stride]
  if (dir == 0) {
    /* cell c = f gives ML, cell c = f+1 gives MR; 3-point near the x ends
     * (Flux_X_limiter.cpp:133), never-filled ghosts read as zero */
    int ge = d->gend;
    for (int v = 0; v < 5; v++) {
      int c = f;
      L[v] = (c <= 3 || c >= ge) ? lim3_up(UV(v, c - 1), UV(v, c), UV(v, c + 1), 1) * Jfl
                                 : lim5_up(UV(v, c - 2), UV(v, c - 1), UV(v, c), UV(v, c + 1), UV(v, c + 2), 1) * Jfl;
      c = f + 1;
      Rr[v] = (c <= 3 || c >= ge) ? lim3_dn(UV(v, c - 1), UV(v, c), UV(v, c + 1)) * Jfr
                                  : lim5_dn(UV(v, c - 2), UV(v, c - 1), UV(v, c), UV(v, c + 1), UV(v, c + 2)) * Jfr;
    }
  } else if (dir == 1) {
    int ny = d->ny;
    for (int v = 0; v < 5; v++) {
      /* left state: the wall face f = 1 is the J-weighted average of cells
       * 1, 2 (Flux_Y_limiter.cpp:368-374); otherwise cell c = f */
      int c = f;
      if (f == 1) {
        double temp = 1. / (1 / JC(1) + 1 / JC(2));
        L[v] = temp * (UV(v, 1) + UV(v, 2));
      } else if (c <= 2 || c >= ny - 1) {
        L[v] = v == 0 ? lim3y_up(UV(v, c - 1), UV(v, c), UV(v, c + 1), JC(c))
                      : lim3_up(UV(v, c - 1), UV(v, c), UV(v, c + 1), 1) * Jfl;
      } else {
        L[v] = lim5_up(UV(v, c - 2), UV(v, c - 1), UV(v, c), UV(v, c + 1), UV(v, c + 2), 1) * Jfl;
      }
      /* right state: face f = ny is the average of cells ny, ny+1
       * (Flux_Y_limiter.cpp:434-440); otherwise cell c = f+1 */
      c = f + 1;
      if (f == ny) {
        double temp = 1. / (1 / JC(ny) + 1 / JC(ny + 1));
        Rr[v] = temp * (UV(v, ny) + UV(v, ny + 1));
      } else if (c <= 2 || c >= ny - 1) {
        Rr[v] = v == 0 ? lim3y_dn(UV(v, c - 1), UV(v, c), UV(v, c + 1), JC(c))
                       : lim3_dn(UV(v, c - 1), UV(v, c), UV(v, c + 1)) * Jfr;
      } else {
        Rr[v] = lim5_dn(UV(v, c - 2), UV(v, c - 1), UV(v, c), UV(v, c + 1), UV(v, c + 2)) * Jfr;
      }
    }
  } else {
    int nz = d->nz, ny = d->ny;
    const recipe_t* rl = (f == 1) ? &R_AVG : (f == 2 || f == nz) ? &R_O3L : NULL;
    const recipe_t* rr = (f == 1 || f == nz - 1) ? &R_O3R : (f == nz) ? &R_AVG : NULL;
    if (rl) reconstruct(rl, U_, J, base, stride, f, L);
    if (rr) reconstruct(rr, U_, J, base, stride, f, Rr);
    /* MUSCL cells k in [3, nz-1]; the 3-point branch is taken for k >= ny
     * (Flux_Z_limiter.cpp:130 compares k with ny, B9), the left state uses
     * u - 0.5 phi du and the face weights are 1/sum(w/J) unnormalised */
    for (int side = 0; side < 2; side++) {
      if (side == 0 && rl) continue;
      if (side == 1 && rr) continue;
      int c = side == 0 ? f : f + 1;
      double jw;
      int o3 = (c <= 2 || c >= ny);
      if (o3) {
        jw = side == 0 ? 1. / (-1 / JC(c - 1) + 5 / JC(c) + 2 / JC(c + 1))
                       : 1. / (2 / JC(c - 1) + 5 / JC(c) - 1 / JC(c + 1));
      } else {
        jw = side == 0 ? 1. / (2 / JC(c - 2) - 13 / JC(c - 1) + 47 / JC(c) + 27 / JC(c + 1) - 3 / JC(c + 2))
                       : 1. / (-3 / JC(c - 2) + 27 / JC(c - 1) + 47 / JC(c) - 13 / JC(c + 1) + 2 / JC(c + 2));
      }
      for (int v = 0; v < 5; v++) {
        double s;
        if (o3) s = side == 0 ? lim3_up(UV(v, c - 1), UV(v, c), UV(v, c + 1), -1) : lim3_dn(UV(v, c - 1), UV(v, c), UV(v, c + 1));
        else s = side == 0 ? lim5_up(UV(v, c - 2), UV(v, c - 1), UV(v, c), UV(v, c + 1), UV(v, c + 2), -1)
                           : lim5_dn(UV(v, c - 2), UV(v, c - 1), UV(v, c), UV(v, c + 1), UV(v, c + 2));
        if (side == 0) L[v] = s * jw; else Rr[v] = s * jw;
      }
    }
  }
#undef UV
#undef JC
}

void cfo_flux_limiter(const cfo_grid* g, int dir, double* const U_[5], const double* J, const double* Jf,
                      double* const mf[9], const double* Ep, double* const inF[5]) {
  dims_t d = dims(g);
  size_t sI = (size_t)d.Ym * d.Zm, sJ = (size_t)d.Zm, sK = 1;
  /* faces and transverse cells per direction (Flux_*_limiter.cpp flux loops) */
  int f0, f1, a0, a1, b0, b1;
  size_t stride, sa, sb;
  if (dir == 0) { f0 = 2; f1 = d.gend; a0 = 2; a1 = d.ny; b0 = 2; b1 = d.nz; stride = sI; sa = sJ; sb = sK; }
  else if (dir == 1) { f0 = 1; f1 = d.ny; a0 = 3; a1 = d.gend; b0 = 2; b1 = d.nz; stride = sJ; sa = sI; sb = sK; }
  else { f0 = 1; f1 = d.nz; a0 = 3; a1 = d.gend; b0 = 2; b1 = d.ny; stride = sK; sa = sI; sb = sJ; }
  const int mrow = 3 * dir;
  for (int f = f0; f <= f1; f++)
    for (int a = a0; a <= a1; a++)
      for (int b = b0; b <= b1; b++) {
        size_t base = (size_t)a * sa + (size_t)b * sb;
        size_t o = (size_t)f * stride + (size_t)(a - 1) * sa + (size_t)(b - 1) * sb;
        /* face J that scales ML / MR: J_u[c][j-1][k-1] and J_u[c-1][j-1][k-1]
         * (x, Flux_X_limiter.cpp:125-126); J_v at faces f and f (y, the
         * interior assignment of Flux_Y_limiter.cpp:231-232, also used near
         * the walls where the reference leaves it unassigned) */
        double Jfl = 0, Jfr = 0;
        if (dir == 0) { Jfl = Jf[o]; Jfr = Jf[o]; }
        else if (dir == 1) {
          Jfl = Jf[o];
          Jfr = Jf[o];
        }
        double L[5], Rr[5], F[5], fav3, psum;
        limiter_states(&d, dir, U_, J, base, stride, f, Jfl, Jfr, L, Rr);
        double n[3] = {mf[mrow][o], mf[mrow + 1][o], mf[mrow + 2][o]};
        limiter_face_flux(L, Rr, n, Jf[o], Ep[o], F, &fav3, &psum);
        if (dir == 1 && (f == 1 || f == d.ny)) {
          /* wall faces (Flux_Y_limiter.cpp:658-666) */
          F[0] = 0;
          F[1] = 0;
          F[2] = 0.5 * (n[1] * psum - Ep[o] * fav3) / Jf[o];
          F[3] = 0;
          F[4] = 0;
        }
        for (int v = 0; v < 5; v++) inF[v][o] = F[v];
      }
}

/* ------------------------------------------------------------------------ */
/* Grid (Grid.cpp:50-141) and the ADM sensor Filter (Filter.cpp:12-1720)    */
/* ------------------------------------------------------------------------ */

void cfo_grid_points(const cfo_grid* g, double gamma, double* X_point, double* Y_point) {
  dims_t d = dims(g);
  for (int i = 0; i <= d.gend + 3; i++)
    for (int j = 2; j <= d.ny; j++)
      for (int k = 2; k <= d.nz; k++)
        Y_point[IX(d, i, j, k)] = (0.5 * g->high) * (1 - 1. / tanh(gamma) * tanh(gamma * (1 - 2 * (j - 1.5) * g->deltaET)));
  for (int j = 2; j <= d.ny; j++)
    for (int k = 2; k <= d.nz; k++) {
      Y_point[IX(d, 2, j, k)] = Y_point[IX(d, 3, j, k)];
      Y_point[IX(d, d.gend + 1, j, k)] = Y_point[IX(d, d.gend, j, k)];
    }
  for (int i = 2; i <= d.gend + 1; i++)
    for (int k = 2; k <= d.nz; k++) {
      Y_point[IX(d, i, 1, k)] = -Y_point[IX(d, i, 2, k)];
      Y_point[IX(d, i, d.ny + 1, k)] = Y_point[IX(d, i, d.ny, k)] + 2 * Y_point[IX(d, i, 2, k)];
    }
  for (int i = 2; i <= d.gend + 1; i++)
    for (int j = 1; j <= d.ny + 1; j++) {
      Y_point[IX(d, i, j, 1)] = Y_point[IX(d, i, j, 2)];
      Y_point[IX(d, i, j, d.nz + 1)] = Y_point[IX(d, i, j, d.nz)];
    }
  if (X_point)
    for (int i = 3; i <= d.gend; i++)
      for (int j = 2; j <= d.ny; j++)
        for (int k = 2; k <= d.nz; k++) X_point[IX(d, i, j, k)] = g->deltaXI * (i - 1.5);
}

/* one separable pass of the filter on components 0..3 over i in [3,gend],
 * j in [2,ny], k in [2,nz]; the work arrays start as copies of U_, so every
 * value read outside that box is U_'s own (Filter.cpp:121-135).
 * wide = 0: the h filter (1,6,1)/8 (Filter.cpp:156-240); wide = 1: the 2h
 * filter (1,2,1)/4 (Filter.cpp:831-915). */
static void filter_pass(const dims_t* d, int wide, double* const U_[5], const double* src[4], double* Gx[4],
                        double* Gy[4], double* out[4], const double* J, const double* Yp) {
  for (int v = 0; v < 4; v++) {
    for (int i = 3; i <= d->gend; i++)
      for (int j = 2; j <= d->ny; j++)
        for (int k = 2; k <= d->nz; k++) {
          size_t p = IX(*d, i, j, k), pm = IX(*d, i - 1, j, k), pp = IX(*d, i + 1, j, k);
          double a = (i - 1 < 3) ? U_[v][pm] : src[v][pm];
          double b = src[v][p];
          double c = (i + 1 > d->gend) ? U_[v][pp] : src[v][pp];
          Gx[v][p] = wide ? 0.25 * (a + 2 * b + c) : 0.125 * (a + 6 * b + c);
        }
    for (int i = 3; i <= d->gend; i++)
      for (int j = 2; j <= d->ny; j++)
        for (int k = 2; k <= d->nz; k++) {
          size_t p = IX(*d, i, j, k), pm = IX(*d, i, j - 1, k), pp = IX(*d, i, j + 1, k);
          double A = Yp[pp] - Yp[pm];
          double B = Yp[pp] - Yp[p];
          double Cc = Yp[p] - Yp[pm];
          double gm = (j - 1 < 2) ? U_[v][pm] : Gx[v][pm];
          double gp = (j + 1 > d->ny) ? U_[v][pp] : Gx[v][pp];
          Gy[v][p] = wide ? 0.5 * (B / A * gm * J[pm] + Gx[v][p] * J[p] + Cc / A * gp * J[pp]) / J[p]
                          : 0.125 * (2 * B / A * gm * J[pm] + 6 * Gx[v][p] * J[p] + 2 * Cc / A * gp * J[pp]) / J[p];
        }
    for (int i = 3; i <= d->gend; i++)
      for (int j = 2; j <= d->ny; j++)
        for (int k = 2; k <= d->nz; k++) {
          size_t p = IX(*d, i, j, k), pm = IX(*d, i, j, k - 1), pp = IX(*d, i, j, k + 1);
          double a = (k - 1 < 2) ? U_[v][pm] : Gy[v][pm];
          double c = (k + 1 > d->nz) ? U_[v][pp] : Gy[v][pp];
          out[v][p] = wide ? 0.25 * (a + 2 * Gy[v][p] + c) : 0.125 * (a + 6 * Gy[v][p] + c);
        }
  }
}

/* E = |u - u_f|^2 with u_f from 6 G - 15 G^2 + 20 G^3 - 15 G^4 + 6 G^5 - G^6
 * (Filter.cpp:763-800 and 1438-1475) */
static void filter_energy(const dims_t* d, int wide, double* const U_[5], const double* J, const double* Yp,
                          double* E, double* work) {
  size_t nf = (size_t)(d->gend + 4) * d->Ym * d->Zm;
  double *G[4], *Gn[4], *Gx[4], *Gy[4], *acc[4];
  for (int v = 0; v < 4; v++) {
    G[v] = work + (0 * 4 + v) * nf;
    Gn[v] = work + (1 * 4 + v) * nf;
    Gx[v] = work + (2 * 4 + v) * nf;
    Gy[v] = work + (3 * 4 + v) * nf;
    acc[v] = work + (4 * 4 + v) * nf;
  }
  static const double coef[6] = {6, -15, 20, -15, 6, -1};
  const double* src[4] = {U_[0], U_[1], U_[2], U_[3]};
  for (int pass = 0; pass < 6; pass++) {
    filter_pass(d, wide, U_, src, Gx, Gy, Gn, J, Yp);
    for (int v = 0; v < 4; v++)
      for (int i = 3; i <= d->gend; i++)
        for (int j = 2; j <= d->ny; j++)
          for (int k = 2; k <= d->nz; k++) {
            size_t p = IX(*d, i, j, k);
            double t = Gn[v][p];
            if (pass == 0) acc[v][p] = 6 * t;
            else if (pass == 5) acc[v][p] = acc[v][p] - t;
            else if (coef[pass] > 0) acc[v][p] = acc[v][p] + coef[pass] * t;
            else acc[v][p] = acc[v][p] - (-coef[pass]) * t;
            G[v][p] = t;
          }
    for (int v = 0; v < 4; v++) src[v] = G[v];
  }
  for (int i = 3; i <= d->gend; i++)
    for (int j = 2; j <= d->ny; j++)
      for (int k = 2; k <= d->nz; k++) {
        size_t p = IX(*d, i, j, k);
        double Ue = U_[1][p] / U_[0][p] - acc[1][p] / acc[0][p];
        double Ve = U_[2][p] / U_[0][p] - acc[2][p] / acc[0][p];
        double We = U_[3][p] / U_[0][p] - acc[3][p] / acc[0][p];
        E[p] = Ue * Ue + Ve * Ve + We * We;
      }
}

/* E at a cell of the sensor stencils: the planes next to the box copy the
 * box's first/last cells (Filter.cpp:1490-1540) */
static double e_at(const dims_t* d, const double* E, int i, int j, int k) {
  if (i < 3) i = 3;
  if (i > d->gend) i = d->gend;
  if (j < 2) j = 2;
  if (j > d->ny) j = d->ny;
  if (k < 2) k = 2;
  if (k > d->nz) k = d->nz;
  return E[IX(*d, i, j, k)];
}

/* the sensor ratios themselves (test infrastructure: thresholds for the
 * pinning fixtures are chosen from them) */
static void filter_ratios(const dims_t* d, const double* Eh, const double* E2h, double* R[3]);

void cfo_filter_ratios(const cfo_grid* g, double* const U_[5], const double* Y_point, const double* J,
                       double* RX, double* RY, double* RZ) {
  dims_t d = dims(g);
  size_t nf = (size_t)(d.gend + 4) * d.Ym * d.Zm;
  double* work = (double*)calloc(20 * nf, sizeof(double));
  double* Eh = (double*)calloc(nf, sizeof(double));
  double* E2h = (double*)calloc(nf, sizeof(double));
  filter_energy(&d, 0, U_, J, Y_point, Eh, work);
  filter_energy(&d, 1, U_, J, Y_point, E2h, work);
  double* R[3] = {RX, RY, RZ};
  filter_ratios(&d, Eh, E2h, R);
  free(work);
  free(Eh);
  free(E2h);
}

void cfo_filter(const cfo_grid* g, double crit, double E_high, double E_low, double* const U_[5],
                const double* Y_point, double* EpX, double* EpY, double* EpZ, const double* J) {
  dims_t d = dims(g);
  size_t nf = (size_t)(d.gend + 4) * d.Ym * d.Zm;
  double* work = (double*)calloc(20 * nf, sizeof(double));
  double* Eh = (double*)calloc(nf, sizeof(double));
  double* E2h = (double*)calloc(nf, sizeof(double));
  filter_energy(&d, 0, U_, J, Y_point, Eh, work);
  filter_energy(&d, 1, U_, J, Y_point, E2h, work);
  /* sensors on the face layouts (Filter.cpp:1550-1645); 0/0 compares false */
  for (int i = 2; i <= d.gend; i++)
    for (int j = 1; j < d.ny; j++)
      for (int k = 1; k < d.nz; k++) {
        double Em = (e_at(&d, Eh, i, j + 1, k + 1) + e_at(&d, Eh, i + 1, j + 1, k + 1)) /
                    (e_at(&d, E2h, i, j + 1, k + 1) + e_at(&d, E2h, i + 1, j + 1, k + 1));
        EpX[IX(d, i, j, k)] = Em >= crit ? E_high : E_low;
      }
  for (int i = 2; i <= d.gend - 1; i++)
    for (int j = 1; j <= d.ny; j++)
      for (int k = 1; k < d.nz; k++) {
        double Em = (e_at(&d, Eh, i + 1, j, k + 1) + e_at(&d, Eh, i + 1, j + 1, k + 1)) /
                    (e_at(&d, E2h, i + 1, j, k + 1) + e_at(&d, E2h, i + 1, j + 1, k + 1));
        EpY[IX(d, i, j, k)] = Em >= crit ? E_high : E_low;
      }
  for (int i = 2; i <= d.gend - 1; i++)
    for (int j = 1; j < d.ny; j++)
      for (int k = 1; k <= d.nz; k++) {
        double Em = (e_at(&d, Eh, i + 1, j + 1, k) + e_at(&d, Eh, i + 1, j + 1, k + 1)) /
                    (e_at(&d, E2h, i + 1, j + 1, k) + e_at(&d, E2h, i + 1, j + 1, k + 1));
        EpZ[IX(d, i, j, k)] = Em >= crit ? E_high : E_low;
      }
  free(work);
  free(Eh);
  free(E2h);
}

static void filter_ratios(const dims_t* d, const double* Eh, const double* E2h, double* R[3]) {
  for (int i = 2; i <= d->gend; i++)
    for (int j = 1; j < d->ny; j++)
      for (int k = 1; k < d->nz; k++)
        R[0][IX(*d, i, j, k)] = (e_at(d, Eh, i, j + 1, k + 1) + e_at(d, Eh, i + 1, j + 1, k + 1)) /
                                (e_at(d, E2h, i, j + 1, k + 1) + e_at(d, E2h, i + 1, j + 1, k + 1));
  for (int i = 2; i <= d->gend - 1; i++)
    for (int j = 1; j <= d->ny; j++)
      for (int k = 1; k < d->nz; k++)
        R[1][IX(*d, i, j, k)] = (e_at(d, Eh, i + 1, j, k + 1) + e_at(d, Eh, i + 1, j + 1, k + 1)) /
                                (e_at(d, E2h, i + 1, j, k + 1) + e_at(d, E2h, i + 1, j + 1, k + 1));
  for (int i = 2; i <= d->gend - 1; i++)
    for (int j = 1; j < d->ny; j++)
      for (int k = 1; k <= d->nz; k++)
        R[2][IX(*d, i, j, k)] = (e_at(d, Eh, i + 1, j + 1, k) + e_at(d, Eh, i + 1, j + 1, k + 1)) /
                                (e_at(d, E2h, i + 1, j + 1, k) + e_at(d, E2h, i + 1, j + 1, k + 1));
}
