// Device-side arithmetic of the channel-flow dual-time path (gfx950).
//
// Every expression follows the reference's FP64 operation order
// (ricklee209/Channel_Flow, cited per function) and the library is compiled
// with -ffp-contract=off, so each add/mul/div/sqrt is one correctly rounded
// IEEE operation in the same order as the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cf {

constexpr double KG = 1.4;         // prm.h:9
constexpr double RG = 287.0;       // prm.h:10
constexpr double EPS_PRE = 0.0001; // prm.h:7
constexpr double CV = 717.5;       // Viscous_terms.h:3
constexpr double MU_L = 0.0000185; // Viscous_terms.h:4
constexpr double PR_L = 0.72;      // Viscous_terms.h:5
constexpr double T_WALL2 = 2 * 298.0592;  // Y_boundary_condition.cpp:87

// Device layout of one field: the reference's [X_np][Y_m][Z_m] with the k
// row padded to Zp doubles (64-byte multiple).  All fields of the context
// share it; `nf` is the element stride between consecutive fields.
struct Geo {
  int Xnp, Ym, Zp, Zm;
  int gend, ny, nz;
  long plane;   // Ym * Zp
  long nf;      // Xnp * plane
  double dXI, dET, dZT;
  double rXI, rET, rZT;  // rcp_checked of the spacings, for div_prof
};

__device__ __forceinline__ long at(const Geo& g, int i, int j, int k) {
  return (long)i * g.plane + (long)j * g.Zp + k;
}

// j-profiles of the metrics: prof[(set*10 + comp) * Ym + j]; set 0 = cell
// centres, 1/2/3 = x/y/z faces; comp 0 = J, 1..9 = xidx..ztdz.
enum { SET_C = 0, SET_U = 1, SET_V = 2, SET_W = 3 };
// The profile buffer holds PROF_ROWS rows of Ym values: 4 sets x (J + 9
// metrics), then rows derived from them for the fluxes (FT_*).
constexpr int PROF_METRIC_ROWS = 40;
constexpr int FT_REC = PROF_METRIC_ROWS;      // 3 directions x 7 recipes: 1 / sum(w/J)
constexpr int FT_NRM = FT_REC + 21;           // 3 directions x (insqr, N0, N1, N2) of the face normal
constexpr int FT_RCP = FT_NRM + 12;          // 3 directions x (rcp_checked of Jf, of insqr)
constexpr int PROF_ROWS = FT_RCP + 6;
__device__ __forceinline__ double pf(const double* __restrict__ prof, int Ym, int set, int comp, int j) {
  return prof[(set * 10 + comp) * Ym + j];
}

// row `row` of the profile buffer (the derived FT_* rows), at j
__device__ __forceinline__ double prow(const double* __restrict__ prof, int Ym, int row, int j) {
  return prof[(long)row * Ym + j];
}

// XCD-aware block order.  Consecutive workgroups are dealt to the 8 XCDs in
// turn, and each XCD has its own L2; this bijection of [0, n) gives XCD x
// (the physical blocks b with b % 8 == x) one contiguous range of logical
// blocks, so neighbouring logical blocks share an L2.
constexpr int N_XCD = 8;
__device__ __forceinline__ long xcd_block(long b, long n) {
  const long x = b % N_XCD, i = b / N_XCD, per = n / N_XCD, rem = n % N_XCD;
  return x * per + (x < rem ? x : rem) + i;
}

// the logical (x, y, z) block of a 3-D grid after xcd_block
__device__ __forceinline__ void xcd_block3(int* bx, int* by, int* bz) {
  const long gx = gridDim.x, gy = gridDim.y;
  const long n = gx * gy * gridDim.z;
  const long L = xcd_block((blockIdx.z * gy + blockIdx.y) * gx + blockIdx.x, n);
  *bx = (int)(L % gx);
  *by = (int)((L / gx) % gy);
  *bz = (int)(L / (gx * gy));
}

// x/d for a divisor d fixed by the grid, with r = RN(1/d):
//   q0 = x r,  e = fma(q0, d, -x) = q0 d - x,  q = fma(-e, r, q0)
// (each step one rounding).  When |r d - 1| <= 2^-54, q0 is within one ulp of
// x/d for every x, the remainder e is exact, and Markstein's theorem (IBM J.
// Res. Dev. 34(1), 1990) makes q the correctly rounded x/d: the same bits as
// the division.  Forming e as q0 d - x keeps the sign of a zero quotient
// (x = -0 gives e = +0 and q = -0 + -0).
__device__ __forceinline__ double div_const(double x, double d, double r) {
  const double q0 = x * r;
  const double e = fma(q0, d, -x);
  return fma(-e, r, q0);
}

// RN(1/d) when it meets the bound above (r d - 1 is exact in one fma), else 0
__host__ __device__ __forceinline__ double rcp_checked(double d) {
  const double r = 1.0 / d;
  const double e = fma(r, d, -1.0);
  return fabs(e) <= 0x1p-54 ? r : 0.0;
}

// x/d with a reciprocal from rcp_checked (0: plain division)
__device__ __forceinline__ double div_prof(double x, double d, double r) {
  return r != 0.0 ? div_const(x, d, r) : x / d;
}

// 1/(K-1); |RKG (K-1) - 1| = 2^-54.3 meets the bound (checked on the host
// by tests/test_division_kat.py)
constexpr double RKG = 1.0 / (KG - 1);

// ------------------------------------------------------------------------
// MUSCL recipes (Flux_X.cpp:92-215, Flux_Y.cpp:91-240, Flux_Z.cpp:96-300)
// ------------------------------------------------------------------------
enum Recipe : int { O5L = 0, O5R, O3L, O3R, AVG, AVGM, AVG1 };

__device__ __forceinline__ int recipe_left(int dir, int f, int gend, int ny, int nz) {
  if (dir == 0) {
    if (f == 2) return AVG;
    if (f == 3 || f == gend) return O3L;
    return O5L;
  } else if (dir == 1) {
    if (f == 1 || f == 2) return AVG;
    if (f == 3 || f == ny - 1) return O3L;
    if (f == ny) return AVGM;
    return O5L;
  }
  if (f == 1) return AVG;
  if (f == 2 || f == nz) return O3L;
  return O5L;
}

__device__ __forceinline__ int recipe_right(int dir, int f, int gend, int ny, int nz) {
  if (dir == 0) {
    if (f == 2 || f == gend - 1) return O3R;
    if (f == gend) return AVG;
    return O5R;
  } else if (dir == 1) {
    if (f == 1) return AVG1;
    if (f == 2 || f == ny - 2) return O3R;
    if (f == ny - 1 || f == ny) return AVG;
    return O5R;
  }
  if (f == 1 || f == nz - 1) return O3R;
  if (f == nz) return AVG;
  return O5R;
}

// Weighted state sum(w*U)/sum(w/J) over the recipe's cells; u[v][t] and
// jj[t] hold U and J of cells f-2 .. f+3 (t = cell - f + 2).  Each recipe is
// a compile-time instance so every register index is a constant.
template <int FIRST, int N>
__device__ __forceinline__ void rec_w(const double (&u)[5][6], const double (&jj)[6], const double (&w)[N],
                                      double q[5]) {
  constexpr int t0 = FIRST + 2;
  double s = w[0] / jj[t0];
#pragma unroll
  for (int t = 1; t < N; t++) s += w[t] / jj[t0 + t];
  double temp = 1. / s;
#pragma unroll
  for (int v = 0; v < 5; v++) {
    double a = w[0] * u[v][t0];
#pragma unroll
    for (int t = 1; t < N; t++) a += w[t] * u[v][t0 + t];
    q[v] = temp * a;
  }
}

// the weighted sums of rec_w with temp = 1/sum(w/J) taken from the profile
// rows (it depends only on the recipe and the J of the stencil, hence on j)
template <int FIRST, int N>
__device__ __forceinline__ void rec_w_t(const double (&u)[5][6], const double (&w)[N], double temp, double q[5]) {
  constexpr int t0 = FIRST + 2;
#pragma unroll
  for (int v = 0; v < 5; v++) {
    double a = w[0] * u[v][t0];
#pragma unroll
    for (int t = 1; t < N; t++) a += w[t] * u[v][t0 + t];
    q[v] = temp * a;
  }
}

__device__ __forceinline__ void reconstruct_t(int r, const double (&u)[5][6], double temp, double q[5]) {
  switch (r) {
    case O5L: { const double w[5] = {2, -13, 47, 27, -3}; rec_w_t<-2, 5>(u, w, temp, q); break; }
    case O5R: { const double w[5] = {-3, 27, 47, -13, 2}; rec_w_t<-1, 5>(u, w, temp, q); break; }
    case O3L: { const double w[3] = {-1, 5, 2}; rec_w_t<-1, 3>(u, w, temp, q); break; }
    case O3R: { const double w[3] = {2, 5, -1}; rec_w_t<0, 3>(u, w, temp, q); break; }
    case AVG: { const double w[2] = {1, 1}; rec_w_t<0, 2>(u, w, temp, q); break; }
    case AVGM: { const double w[2] = {1, 1}; rec_w_t<-1, 2>(u, w, temp, q); break; }
    default: { const double w[2] = {1, 1}; rec_w_t<1, 2>(u, w, temp, q); break; }  // AVG1
  }
}

// 1/sum(w/J) of recipe r, formed exactly as rec_w forms it
template <int FIRST, int N>
__device__ __forceinline__ double rec_temp_w(const double (&jj)[6], const double (&w)[N]) {
  constexpr int t0 = FIRST + 2;
  double s = w[0] / jj[t0];
#pragma unroll
  for (int t = 1; t < N; t++) s += w[t] / jj[t0 + t];
  return 1. / s;
}

__device__ __forceinline__ double rec_temp(int r, const double (&jj)[6]) {
  switch (r) {
    case O5L: { const double w[5] = {2, -13, 47, 27, -3}; return rec_temp_w<-2, 5>(jj, w); }
    case O5R: { const double w[5] = {-3, 27, 47, -13, 2}; return rec_temp_w<-1, 5>(jj, w); }
    case O3L: { const double w[3] = {-1, 5, 2}; return rec_temp_w<-1, 3>(jj, w); }
    case O3R: { const double w[3] = {2, 5, -1}; return rec_temp_w<0, 3>(jj, w); }
    case AVG: { const double w[2] = {1, 1}; return rec_temp_w<0, 2>(jj, w); }
    case AVGM: { const double w[2] = {1, 1}; return rec_temp_w<-1, 2>(jj, w); }
    default: { const double w[2] = {1, 1}; return rec_temp_w<1, 2>(jj, w); }  // AVG1
  }
}

__device__ __forceinline__ void reconstruct(int r, const double (&u)[5][6], const double (&jj)[6], double q[5]) {
  switch (r) {
    case O5L: { const double w[5] = {2, -13, 47, 27, -3}; rec_w<-2, 5>(u, jj, w, q); break; }
    case O5R: { const double w[5] = {-3, 27, 47, -13, 2}; rec_w<-1, 5>(u, jj, w, q); break; }
    case O3L: { const double w[3] = {-1, 5, 2}; rec_w<-1, 3>(u, jj, w, q); break; }
    case O3R: { const double w[3] = {2, 5, -1}; rec_w<0, 3>(u, jj, w, q); break; }
    case AVG: { const double w[2] = {1, 1}; rec_w<0, 2>(u, jj, w, q); break; }
    case AVGM: { const double w[2] = {1, 1}; rec_w<-1, 2>(u, jj, w, q); break; }
    default: { const double w[2] = {1, 1}; rec_w<1, 2>(u, jj, w, q); break; }  // AVG1
  }
}

// Roe flux, ROE == 1 (Flux_X.cpp:221-420, Flux_Y.cpp:247-460,
// Flux_Z.cpp:307-507).  energy_norm selects deltaP*V*insqr (Flux_Y.cpp:441)
// over deltaP*U (Flux_X.cpp:412, Flux_Z.cpp:499); wall turns on the
// impermeable-wall face of Flux_Y.cpp:452-460.
// nt = (insqr, N0, N1, N2) = (1/|n|, n/|n|) of the face, a function of j
// formed once in the profile rows as 1.0 / sqrt(n.n) and n * insqr
// FAST: divide by Jf and insqr through their reciprocals rJf, rin
// (rcp_checked, both non-zero); otherwise plainly
template <bool FAST>
__device__ __forceinline__ double div_face(double x, double d, double r) {
  if (FAST) return div_const(x, d, r);
  return x / d;
}

template <bool FAST>
__device__ __forceinline__ void roe_flux(const double L[5], const double Rr[5], const double n[3], const double nt[4],
                                         double Jf, double rJf, double rin, bool energy_norm, bool wall,
                                         double F[5]) {
  const double insqr = nt[0], N0 = nt[1], N1 = nt[2], N2 = nt[3];
  double _rho = L[0], _u = L[1] / _rho, _v = L[2] / _rho, _w = L[3] / _rho;
  double _U = n[0] * _u + n[1] * _v + n[2] * _w;
  double _VV = _u * _u + _v * _v + _w * _w;
  double _P = (L[4] - 0.5 * _rho * _VV) * (KG - 1);
  double _C = KG * _P / _rho;
  double _H = 0.5 * _VV + div_const(_C, KG - 1, RKG);
  double rho_ = Rr[0], u_ = Rr[1] / rho_, v_ = Rr[2] / rho_, w_ = Rr[3] / rho_;
  double U_ = n[0] * u_ + n[1] * v_ + n[2] * w_;
  double VV_ = u_ * u_ + v_ * v_ + w_ * w_;
  double P_ = (Rr[4] - 0.5 * rho_ * VV_) * (KG - 1);
  double C_ = KG * P_ / rho_;
  double H_ = 0.5 * VV_ + div_const(C_, KG - 1, RKG);
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
  double dU5 = (div_const(P_, KG - 1, RKG) + 0.5 * rho_ * VV_) - (div_const(_P, KG - 1, RKG) + 0.5 * _rho * _VV);
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
  double Fav5 = energy_norm ? a * dU5 + deltaU * rho * H + deltaP * U * insqr : a * dU5 + deltaU * rho * H + deltaP * U;
  if (wall) {
    F[0] = 0;
    F[1] = 0;
    F[2] = div_face<FAST>(0.5 * (n[1] * (_P + P_) - div_face<FAST>(Fav3, insqr, rin)), Jf, rJf);
    F[3] = 0;
    F[4] = 0;
    return;
  }
  F[0] = div_face<FAST>(0.5 * ((_rho * _U + rho_ * U_) - div_face<FAST>(Fav1, insqr, rin)), Jf, rJf);
  F[1] = div_face<FAST>(0.5 * ((_rho * _u * _U + rho_ * u_ * U_ + n[0] * (_P + P_)) - div_face<FAST>(Fav2, insqr, rin)), Jf, rJf);
  F[2] = div_face<FAST>(0.5 * ((_rho * _v * _U + rho_ * v_ * U_ + n[1] * (_P + P_)) - div_face<FAST>(Fav3, insqr, rin)), Jf, rJf);
  F[3] = div_face<FAST>(0.5 * ((_rho * _w * _U + rho_ * w_ * U_ + n[2] * (_P + P_)) - div_face<FAST>(Fav4, insqr, rin)), Jf, rJf);
  F[4] = div_face<FAST>(0.5 * ((_U * (3.5 * _P + 0.5 * _rho * _VV) + U_ * (3.5 * P_ + 0.5 * rho_ * VV_)) - div_face<FAST>(Fav5, insqr, rin)),
                  Jf, rJf);
}

// ------------------------------------------------------------------------
// dual time stepping (Dual_time_stepping.cpp)
// ------------------------------------------------------------------------
struct PState {
  double rho, u, v, w, VV, P, C, T, H, beta;
};

// Dual_time_stepping.cpp:343-372
__device__ __forceinline__ PState prim_dt(double U1, double U2, double U3, double U4, double U5, double J) {
  PState s;
  s.rho = U1 * J;
  s.u = U2 / U1;
  s.v = U3 / U1;
  s.w = U4 / U1;
  s.VV = s.u * s.u + s.v * s.v + s.w * s.w;
  s.P = (U5 * J - 0.5 * s.rho * s.VV) * (KG - 1);
  s.C = KG * s.P / s.rho;
  s.T = s.P / s.rho;
  s.H = 0.5 * s.VV + s.C / (KG - 1);
  double r = s.VV / s.C;
  s.beta = r > EPS_PRE ? r : EPS_PRE;  // max(VV/C, e)
  return s;
}

__device__ __forceinline__ double lambda_dir(const PState& s, double n0, double n1, double n2, double* Uc) {
  double U = n0 * s.u + n1 * s.v + n2 * s.w;
  double S = sqrt(U * U * (s.beta - 1) * (s.beta - 1) + 4 * s.beta * s.C * (n0 * n0 + n1 * n1 + n2 * n2));
  *Uc = U;
  return ((s.beta + 1) * fabs(U) + S) / 2;
}

// D^-1 = {d11, d22 (= d33 = d44 = d55), d51}, Dual_time_stepping.cpp:464-489,
// from the three spectral radii of the cell
__device__ __forceinline__ void dinv_lam(const PState& s, double lx, double ly, double lz, double dXI, double dET,
                                         double dZT, double deltaT, double* d11, double* d22, double* d51) {
  double lambda = lx / dXI + ly / dET + lz / dZT;
  double beta = s.beta;
  *d11 = 2 * deltaT / (3 * beta + 2 * lambda * deltaT);
  *d22 = 2 * deltaT / (3 + 2 * lambda * deltaT);
  double temp1 = 6 * (KG - 1) * (beta - 1) * deltaT;
  double temp2 = KG * (3 + 2 * lambda * deltaT) * (3 * beta + 2 * lambda * deltaT) * s.rho;
  *d51 = -temp1 / temp2;
}

__device__ __forceinline__ void dinv(const PState& s, const double m9[9], double dXI, double dET, double dZT,
                                     double deltaT, double* d11, double* d22, double* d51) {
  double U, V, W;
  double lx = lambda_dir(s, m9[0], m9[1], m9[2], &U);
  double ly = lambda_dir(s, m9[3], m9[4], m9[5], &V);
  double lz = lambda_dir(s, m9[6], m9[7], m9[8], &W);
  dinv_lam(s, lx, ly, lz, dXI, dET, dZT, deltaT, d11, d22, d51);
}


// Split flux Jacobian of one cell along metric row n (Dual_time_stepping.cpp
// :608-632 forward, :968-972 backward diagonals).  Entries 52..54 are
// written n*beta*T*(K-1) in x and beta*T*(K-1)*n in y and z.
struct Split {
  double p11, p12, p13, p14, p21, p22, p31, p41, p51, p52, p53, p54;
  double m11, m22;
};

__device__ __forceinline__ Split split_jac(const PState& s, double n0, double n1, double n2, int dir) {
  Split a;
  double U;
  double lam = lambda_dir(s, n0, n1, n2, &U);
  double dd = fabs(lam);
  double beta = s.beta, P = s.P, rho = s.rho, T = s.T;
  a.p11 = 0.5 * (beta * U + dd);
  a.p12 = 0.5 * P * KG * beta * n0;
  a.p13 = 0.5 * P * KG * beta * n1;
  a.p14 = 0.5 * P * KG * beta * n2;
  a.p21 = 0.5 * n0 / rho;
  a.p22 = 0.5 * (U + dd);
  a.p31 = 0.5 * n1 / rho;
  a.p41 = 0.5 * n2 / rho;
  a.p51 = 0.5 * U * (KG - 1) * (beta - 1) / rho / KG;
  if (dir == 0) {
    a.p52 = 0.5 * n0 * beta * T * (KG - 1);
    a.p53 = 0.5 * n1 * beta * T * (KG - 1);
    a.p54 = 0.5 * n2 * beta * T * (KG - 1);
  } else {
    a.p52 = 0.5 * beta * T * (KG - 1) * n0;
    a.p53 = 0.5 * beta * T * (KG - 1) * n1;
    a.p54 = 0.5 * beta * T * (KG - 1) * n2;
  }
  a.m11 = 0.5 * (beta * U - dd);
  a.m22 = 0.5 * (U - dd);
  return a;
}

// Sweep summary for metrics whose rows are diagonal (xi = (xix,0,0),
// eta = (0,ety,0), zeta = (0,0,ztz): every grid Jacobian*.cpp produces).
// The off-diagonal split-Jacobian entries are then exact zeros.  Per cell:
// P, beta, rho and, per direction d, the contravariant velocity (G_UX + 2d)
// and the spectral radius (G_LX + 2d).  T = P/rho (prim_dt's own
// expression), D^-1 (diag_dinv) and the two split-Jacobian entries that
// divide by rho (split_pd_p51) are formed from them where they are used.
enum DiagField : int {
  G_P = 0, G_BETA, G_RHO,
  G_UX, G_LX, G_UY, G_LY, G_UZ, G_LZ,
  G_N
};

// x / rho and x / K: correctly rounded in FP64; in the FP32 sweeps, products
// with the hardware reciprocal (1 ulp) and with 1/K, which the inexact inner
// solve tolerates
__device__ __forceinline__ double over_rho(double x, double rho) { return x / rho; }
__device__ __forceinline__ float over_rho(float x, float rho) { return x * __builtin_amdgcn_rcpf(rho); }
__device__ __forceinline__ double over_k(double x) { return x / KG; }
__device__ __forceinline__ float over_k(float x) { return x * float(1.0 / KG); }

// x / y in the sweep's type: correctly rounded in FP64, reciprocal product in
// the FP32 sweeps
__device__ __forceinline__ double sweep_div(double x, double y) { return x / y; }
__device__ __forceinline__ float sweep_div(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }

// D^-1 = {d11, d22 (= d33 = d44), d51} of one cell from its spectral radii
// (Dual_time_stepping.cpp:464-489), dinv_lam's expressions in its order;
// rdXI.. are 1/dXI.. for the FP32 sweeps (unused in FP64)
template <class R>
__device__ __forceinline__ void diag_dinv(R beta, R rho, R lx, R ly, R lz, double dXI, double dET, double dZT,
                                          double deltaT, R* d11, R* d22, R* d51) {
  R lambda;
  if constexpr (sizeof(R) == sizeof(double))
    lambda = lx / dXI + ly / dET + lz / dZT;
  else
    lambda = lx * R(1.0 / dXI) + ly * R(1.0 / dET) + lz * R(1.0 / dZT);
  const R dt = R(deltaT);
  *d11 = sweep_div(R(2) * dt, R(3) * beta + R(2) * lambda * dt);
  *d22 = sweep_div(R(2) * dt, R(3) + R(2) * lambda * dt);
  const R temp1 = R(6 * (KG - 1)) * (beta - R(1)) * dt;
  const R temp2 = R(KG) * (R(3) + R(2) * lambda * dt) * (R(3) * beta + R(2) * lambda * dt) * rho;
  *d51 = -sweep_div(temp1, temp2);
}

// PD = 0.5*n_dd/rho (A+21, B+31, C+41) and P51 = 0.5*U*(K-1)*(beta-1)/rho/K
// of one direction (Dual_time_stepping.cpp:751-811), in the reference's order
template <class R>
__device__ __forceinline__ void split_pd_p51(R n, R U, R beta, R rho, R* pd, R* p51) {
  *pd = over_rho(R(0.5) * n, rho);
  *p51 = over_k(over_rho(R(0.5) * U * R(KG - 1) * (beta - R(1)), rho));
}

struct DNb {  // one neighbour as seen along one direction
  double P, T, beta, U, lam, pd, p51;
};

// Split-Jacobian entries of one neighbour along direction D with diagonal
// metric n: forward A+ (c11 = p11, c22 = p22) or backward A- (m11, m22);
// c1q, c5q, pd, p51 are shared.  They depend on the neighbour's state and
// the metric at its j only.
// R is the sweep's arithmetic type: double (the reference's arithmetic, in its
// order) or float (the FP32 inner solve)
template <class R>
struct SplitCoefT {
  R c11, c1q, c22, c5q, pd, p51;
};
using SplitCoef = SplitCoefT<double>;

template <int D, bool FWD, class R>
__device__ __forceinline__ SplitCoefT<R> diag_coef(R P, R T, R beta, R U, R lam, R pd, R p51, R n) {
  const R h = R(0.5), kg = R(KG), kg1 = R(KG - 1);
  const R dd = fabs(lam);
  SplitCoefT<R> c;
  c.c11 = FWD ? h * (beta * U + dd) : h * (beta * U - dd);
  c.c1q = h * P * kg * beta * n;
  c.c22 = FWD ? h * (U + dd) : h * (U - dd);
  c.c5q = D == 0 ? h * n * beta * T * kg1 : h * beta * T * kg1 * n;
  c.pd = pd;
  c.p51 = p51;
  return c;
}

// the neighbour's contribution: forward -A+ N (Dual_time_stepping.cpp:751-811),
// backward A- N (Dual_time_stepping.cpp:1116-1177), without the exact-zero terms
template <int D, bool FWD, class R>
__device__ __forceinline__ void diag_apply(const SplitCoefT<R>& c, const R N[5], R o[5]) {
  constexpr int q = D + 1;
  if (FWD) {
    o[0] = -c.c11 * N[0] - c.c1q * N[q];
#pragma unroll
    for (int m = 1; m <= 3; m++) o[m] = (m == q) ? -c.pd * N[0] - c.c22 * N[m] : -(c.c22 * N[m]);
    o[4] = -c.p51 * N[0] - c.c5q * N[q] - c.c22 * N[4];
  } else {
    o[0] = +c.c11 * N[0] + c.c1q * N[q];
#pragma unroll
    for (int m = 1; m <= 3; m++) o[m] = (m == q) ? +c.pd * N[0] + c.c22 * N[m] : c.c22 * N[m];
    o[4] = +c.p51 * N[0] + c.c5q * N[q] + c.c22 * N[4];
  }
}

template <int D>
__device__ __forceinline__ void lower_diag(const DNb& c, double n, const double N[5], double o[5]) {
  diag_apply<D, true>(diag_coef<D, true>(c.P, c.T, c.beta, c.U, c.lam, c.pd, c.p51, n), N, o);
}

template <int D>
__device__ __forceinline__ void upper_diag(const DNb& c, double n, const double N[5], double o[5]) {
  diag_apply<D, false>(diag_coef<D, false>(c.P, c.T, c.beta, c.U, c.lam, c.pd, c.p51, n), N, o);
}


// lower (forward) contribution -A+ N of one neighbour, Dual_time_stepping.cpp:751-811
__device__ __forceinline__ void lower_sum(const Split& a, const double N[5], double o[5]) {
  o[0] = -a.p11 * N[0] - a.p12 * N[1] - a.p13 * N[2] - a.p14 * N[3];
  o[1] = -a.p21 * N[0] - a.p22 * N[1];
  o[2] = -a.p31 * N[0] - a.p22 * N[2];
  o[3] = -a.p41 * N[0] - a.p22 * N[3];
  o[4] = -a.p51 * N[0] - a.p52 * N[1] - a.p53 * N[2] - a.p54 * N[3] - a.p22 * N[4];
}

// upper (backward) contribution of one neighbour, Dual_time_stepping.cpp:1116-1177
__device__ __forceinline__ void upper_sum(const Split& a, const double N[5], double o[5]) {
  o[0] = +a.m11 * N[0] + a.p12 * N[1] + a.p13 * N[2] + a.p14 * N[3];
  o[1] = +a.p21 * N[0] + a.m22 * N[1];
  o[2] = +a.p31 * N[0] + a.m22 * N[2];
  o[3] = +a.p41 * N[0] + a.m22 * N[3];
  o[4] = +a.p51 * N[0] + a.p52 * N[1] + a.p53 * N[2] + a.p54 * N[3] + a.m22 * N[4];
}

// ------------------------------------------------------------------------
// viscous terms (Viscous_terms.cpp)
// ------------------------------------------------------------------------
struct Prim {
  double U, V, W, T;
};

// Viscous_terms.cpp:261-395 (rho here is rho/J)
__device__ __forceinline__ Prim prim_visc(double U1, double U2, double U3, double U4, double U5, double J) {
  Prim q;
  double rho = U1;
  q.U = U2 / rho;
  q.V = U3 / rho;
  q.W = U4 / rho;
  double VV = q.U * q.U + q.V * q.V + q.W * q.W;
  double P = (U5 - 0.5 * rho * VV) * (KG - 1) * J;
  q.T = P / (rho * RG) / J;
  return q;
}

struct DSet {
  double u, v, w, uu, vv, ww, uv, vw, uw, T;
};

__device__ __forceinline__ DSet d_two(const Prim& a, const Prim& b, double inv) {
  DSet r;
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

__device__ __forceinline__ DSet d_four(const Prim& p1, const Prim& p2, const Prim& m1, const Prim& m2, double inv4) {
  DSet r;
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

// coefficient groups a..f of one face (Viscous_terms.cpp:471-526), 1-based
struct Coef {
  double a[8], b[8], c[8], d[11], e[11], f[11];
};

__device__ __forceinline__ void diag_group(const double* p, double* o) {
  double x = p[0], y = p[1], z = p[2];
  o[1] = (4. / 3) * x * x + y * y + z * z;
  o[2] = x * x + (4. / 3) * y * y + z * z;
  o[3] = x * x + y * y + (4. / 3) * z * z;
  o[4] = x * x + y * y + z * z;
  o[5] = (1. / 3) * x * y;
  o[6] = (1. / 3) * y * z;
  o[7] = (1. / 3) * x * z;
}

__device__ __forceinline__ void mixed_group(const double* p, const double* q, double* o) {
  double px = p[0], py = p[1], pz = p[2], qx = q[0], qy = q[1], qz = q[2];
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

__device__ __forceinline__ void coefs(const double m[9], Coef& q) {
  diag_group(m + 0, q.a);
  diag_group(m + 3, q.b);
  diag_group(m + 6, q.c);
  mixed_group(m + 0, m + 3, q.d);
  mixed_group(m + 0, m + 6, q.e);
  mixed_group(m + 3, m + 6, q.f);
}

#define CF_CVK_PR(g4, dT) (CV * KG * (g4) * (dT) / (PR_L))

// x-face flux LL2..LL5 (Viscous_terms.cpp:533-547)
__device__ __forceinline__ void visc_flux_x(const Coef& q, const DSet& dx, const DSet& dy, const DSet& dz, double Ux,
                                            double Vx, double Wx, double Jf, double o[4]) {
  const double *a = q.a, *d = q.d, *e = q.e;
  o[0] = MU_L * (a[1] * dx.u + a[5] * dx.v + a[7] * dx.w + d[1] * dy.u + d[7] * dy.v + d[9] * dy.w + e[1] * dz.u +
                 e[7] * dz.v + e[9] * dz.w) / Jf;
  o[1] = MU_L * (a[5] * dx.u + a[2] * dx.v + a[6] * dx.w + d[5] * dy.u + d[2] * dy.v + d[10] * dy.w + e[5] * dz.u +
                 e[2] * dz.v + e[10] * dz.w) / Jf;
  o[2] = MU_L * (a[7] * dx.u + a[6] * dx.v + a[3] * dx.w + d[6] * dy.u + d[8] * dy.v + d[3] * dy.w + e[6] * dz.u +
                 e[8] * dz.v + e[3] * dz.w) / Jf;
  o[3] = MU_L * (0.5 * a[1] * dx.uu + 0.5 * a[2] * dx.vv + 0.5 * a[3] * dx.ww + a[5] * dx.uv + a[6] * dx.vw +
                 a[7] * dx.uw + CF_CVK_PR(a[4], dx.T) + 0.5 * d[1] * dy.uu + 0.5 * d[2] * dy.vv + 0.5 * d[3] * dy.ww +
                 d[5] * Vx * dy.u + d[6] * Wx * dy.u + d[7] * Ux * dy.v + d[8] * Wx * dy.v + d[9] * Ux * dy.w +
                 d[10] * Vx * dy.w + CF_CVK_PR(d[4], dy.T) + 0.5 * e[1] * dz.uu + 0.5 * e[2] * dz.vv +
                 0.5 * e[3] * dz.ww + e[5] * Vx * dz.u + e[6] * Wx * dz.u + e[7] * Ux * dz.v + e[8] * Wx * dz.v +
                 e[9] * Ux * dz.w + e[10] * Vx * dz.w + CF_CVK_PR(e[4], dz.T)) /
         Jf;
}

// y-face flux ML2..ML5 (Viscous_terms.cpp:681-695)
__device__ __forceinline__ void visc_flux_y(const Coef& q, const DSet& dx, const DSet& dy, const DSet& dz, double Uy,
                                            double Vy, double Wy, double Jf, double o[4]) {
  const double *b = q.b, *d = q.d, *f = q.f;
  o[0] = MU_L * (d[1] * dx.u + d[5] * dx.v + d[6] * dx.w + b[1] * dy.u + b[5] * dy.v + b[7] * dy.w + f[1] * dz.u +
                 f[7] * dz.v + f[9] * dz.w) / Jf;
  o[1] = MU_L * (d[7] * dx.u + d[2] * dx.v + d[8] * dx.w + b[5] * dy.u + b[2] * dy.v + b[6] * dy.w + f[5] * dz.u +
                 f[2] * dz.v + f[10] * dz.w) / Jf;
  o[2] = MU_L * (d[9] * dx.u + d[10] * dx.v + d[3] * dx.w + b[7] * dy.u + b[6] * dy.v + b[3] * dy.w + f[6] * dz.u +
                 f[8] * dz.v + f[3] * dz.w) / Jf;
  o[3] = MU_L * (0.5 * b[1] * dy.uu + 0.5 * b[2] * dy.vv + 0.5 * b[3] * dy.ww + b[5] * dy.uv + b[6] * dy.vw +
                 b[7] * dy.uw + CF_CVK_PR(b[4], dy.T) + 0.5 * d[1] * dx.uu + 0.5 * d[2] * dx.vv + 0.5 * d[3] * dx.ww +
                 d[5] * Uy * dx.v + d[6] * Uy * dx.w + d[7] * Vy * dx.u + d[8] * Vy * dx.w + d[9] * Wy * dx.u +
                 d[10] * Wy * dx.v + CF_CVK_PR(d[4], dx.T) + 0.5 * f[1] * dz.uu + 0.5 * f[2] * dz.vv +
                 0.5 * f[3] * dz.ww + f[5] * Vy * dz.u + f[6] * Wy * dz.u + f[7] * Uy * dz.v + f[8] * Wy * dz.v +
                 f[9] * Uy * dz.w + f[10] * Vy * dz.w + CF_CVK_PR(f[4], dz.T)) /
         Jf;
}

// z-face flux NL2..NL5 (Viscous_terms.cpp:827-841)
__device__ __forceinline__ void visc_flux_z(const Coef& q, const DSet& dx, const DSet& dy, const DSet& dz, double Uz,
                                            double Vz, double Wz, double Jf, double o[4]) {
  const double *c = q.c, *e = q.e, *f = q.f;
  o[0] = MU_L * (e[1] * dx.u + e[5] * dx.v + e[6] * dx.w + f[1] * dy.u + f[5] * dy.v + f[6] * dy.w + c[1] * dz.u +
                 c[5] * dz.v + c[7] * dz.w) / Jf;
  o[1] = MU_L * (e[7] * dx.u + e[2] * dx.v + e[8] * dx.w + f[7] * dy.u + f[2] * dy.v + f[8] * dy.w + c[5] * dz.u +
                 c[2] * dz.v + c[6] * dz.w) / Jf;
  o[2] = MU_L * (e[9] * dx.u + e[10] * dx.v + e[3] * dx.w + f[9] * dy.u + f[10] * dy.v + f[3] * dy.w + c[7] * dz.u +
                 c[6] * dz.v + c[3] * dz.w) / Jf;
  o[3] = MU_L * (0.5 * c[1] * dz.uu + 0.5 * c[2] * dz.vv + 0.5 * c[3] * dz.ww + c[5] * dz.uv + c[6] * dz.vw +
                 c[7] * dz.uw + CF_CVK_PR(c[4], dz.T) + 0.5 * e[1] * dx.uu + 0.5 * e[2] * dx.vv + 0.5 * e[3] * dx.ww +
                 e[5] * Uz * dx.v + e[6] * Uz * dx.w + e[7] * Vz * dx.u + e[8] * Vz * dx.w + e[9] * Wz * dx.u +
                 e[10] * Wz * dx.v + CF_CVK_PR(e[4], dx.T) + 0.5 * f[1] * dy.uu + 0.5 * f[2] * dy.vv +
                 0.5 * f[3] * dy.ww + f[5] * Uz * dy.v + f[6] * Uz * dy.w + f[7] * Vz * dy.u + f[8] * Vz * dy.w +
                 f[9] * Wz * dy.u + f[10] * Wz * dy.v + CF_CVK_PR(f[4], dy.T)) /
         Jf;
}

// ------------------------------------------------------------------------
// Limiter fluxes (Flux_X_limiter.cpp, Flux_Y_limiter.cpp, Flux_Z_limiter.cpp)
// ------------------------------------------------------------------------
// the reference's min/max macros: (a<b)?a:b and (a>b)?a:b (a NaN in the
// second place is returned as is)
__device__ __forceinline__ double mac_min(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double mac_max(double a, double b) { return a > b ? a : b; }

// limited 3-point value at a cell's upper face (before the face-J scaling);
// SGN -1 is the z left state u - 0.5 phi du (Flux_Z_limiter.cpp:163)
template <int SGN>
__device__ __forceinline__ double lim3_up(double um, double u, double up) {
  const double rL = (up - u) / (u - um);
  const double bL = 1. / 3 * (1 + 2 * rL);
  const double m = mac_min(2, mac_min(2 * rL, bL));
  return SGN > 0 ? u + 0.5 * mac_max(0, m) * (u - um) : u - 0.5 * mac_max(0, m) * (u - um);
}
__device__ __forceinline__ double lim3_dn(double um, double u, double up) {
  const double rR = (u - um) / (up - u);
  const double bR = 1. / 3 * (1 + 2 * rR);
  const double m = mac_min(2, mac_min(2 * rR, bR));
  return u - 0.5 * mac_max(0, m) * (up - u);
}
template <int SGN>
__device__ __forceinline__ double lim5_up(double umm, double um, double u, double up, double upp) {
  const double irL = (u - um) / (um - umm);
  const double rL = (up - u) / (u - um);
  const double rLi = (upp - up) / (up - u);
  const double bL = 1. / 30 * (-2. / irL + 11 + 24 * rL - 3 * rL * rLi);
  const double m = mac_min(2, mac_min(2 * rL, bL));
  return SGN > 0 ? u + 0.5 * mac_max(0, m) * (u - um) : u - 0.5 * mac_max(0, m) * (u - um);
}
__device__ __forceinline__ double lim5_dn(double umm, double um, double u, double up, double upp) {
  const double rRi = (up - u) / (upp - up);
  const double rR = (u - um) / (up - u);
  const double irR = (um - umm) / (u - um);
  const double bR = 1. / 30 * (-2. / rRi + 11 + 24 * rR - 3 * rR * irR);
  const double m = mac_min(2, mac_min(2 * rR, bR));
  return u - 0.5 * mac_max(0, m) * (up - u);
}
// near-wall y density (Flux_Y_limiter.cpp:146-168): 1e-7 in the denominators,
// cell value scaled by J, slope not
__device__ __forceinline__ double lim3y_up(double um, double u, double up, double Jc) {
  const double rL = (up - u) / ((u - um) + 0.0000001);
  const double bL = 1. / 3 * (1 + 2 * rL);
  const double m = mac_min(2, mac_min(2 * rL, bL));
  return u * Jc + 0.5 * mac_max(0, m) * (u - um);
}
__device__ __forceinline__ double lim3y_dn(double um, double u, double up, double Jc) {
  const double rR = (u - um) / ((up - u) + 0.0000001);
  const double bR = 1. / 3 * (1 + 2 * rR);
  const double m = mac_min(2, mac_min(2 * rR, bR));
  return u * Jc - 0.5 * mac_max(0, m) * (up - u);
}

// mean-state preconditioned dissipation D (5x5 on dP, du, dv, dw, dT) and
// the face flux (Flux_X_limiter.cpp:361-531).  n = the face's metric row
// along the flux direction; N = n/|n|.
__device__ __forceinline__ void limiter_face_flux(const double L[5], const double R[5], const double n[3],
                                                  double Jf, double Ep, double F[5], double* Fav3_out,
                                                  double* Psum_out) {
  constexpr double K = KG;
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const double Nx = n[0] / nn, Ny = n[1] / nn, Nz = n[2] / nn;
  const double _rho = L[0], _u = L[1] / _rho, _v = L[2] / _rho, _w = L[3] / _rho;
  const double _U = n[0] * _u + n[1] * _v + n[2] * _w;
  const double __U = Nx * _u + Ny * _v + Nz * _w;
  const double _VV = _u * _u + _v * _v + _w * _w;
  const double _P = (L[4] - 0.5 * _rho * _VV) * (K - 1);
  const double _T = _P / _rho;
  const double _C = K * _P / _rho;
  const double _H = 0.5 * _VV + _C / (K - 1);
  const double rho_ = R[0], u_ = R[1] / rho_, v_ = R[2] / rho_, w_ = R[3] / rho_;
  const double U_ = n[0] * u_ + n[1] * v_ + n[2] * w_;
  const double U__ = Nx * u_ + Ny * v_ + Nz * w_;
  const double VV_ = u_ * u_ + v_ * v_ + w_ * w_;
  const double P_ = (R[4] - 0.5 * rho_ * VV_) * (K - 1);
  const double T_ = P_ / rho_;
  const double C_ = K * P_ / rho_;
  const double H_ = 0.5 * VV_ + C_ / (K - 1);
  const double rho = 0.5 * (_rho + rho_);
  const double u = 0.5 * (_u + u_), v = 0.5 * (_v + v_), w = 0.5 * (_w + w_);
  const double U = 0.5 * (_U + U_);
  const double _U_ = 0.5 * (__U + U__);
  const double VV = u * u + v * v + w * w;
  const double H = 0.5 * (_H + H_);
  const double C = (H - 0.5 * VV) * (K - 1);
  const double P = rho * C / K;
  const double T = P / rho;
  const double dU1 = P_ - _P, dU2 = u_ - _u, dU3 = v_ - _v, dU4 = w_ - _w, dU5 = T_ - _T;
  const double beta = mac_max(VV / C, EPS_PRE);
  const double S = sqrt(U * U * (beta - 1) * (beta - 1) + 4 * beta * C * (n[0] * n[0] + n[1] * n[1] + n[2] * n[2]));
  const double _S_ = sqrt(_U_ * _U_ * (beta - 1) * (beta - 1) + 4 * beta * C * (Nx * Nx + Ny * Ny + Nz * Nz));
  const double temp = 4 * K * _S_ * T * beta;
  const double temp1 = S + U + U * beta;
  const double temp2 = -S + U + U * beta;
  const double temp3 = _S_ - _U_ + _U_ * beta;
  const double temp4 = _S_ + _U_ - _U_ * beta;
  const double temp5 = _S_ * _S_ - (beta - 1) * (beta - 1) * _U_ * _U_;
  const double tu1 = 2 * T * Nx * K * beta + u * _S_ + u * (beta - 1) * _U_;
  const double tu2 = 2 * T * Nx * K * beta - u * _S_ + u * (beta - 1) * _U_;
  const double tv1 = 2 * T * Ny * K * beta + v * _S_ + v * (beta - 1) * _U_;
  const double tv2 = 2 * T * Ny * K * beta - v * _S_ + v * (beta - 1) * _U_;
  const double tw1 = 2 * T * Nz * K * beta + w * _S_ + w * (beta - 1) * _U_;
  const double tw2 = 2 * T * Nz * K * beta - w * _S_ + w * (beta - 1) * _U_;
  const double tuvw = 2 * T * K * beta * (u * Nx + v * Ny + w * Nz);
  const double th1 = tuvw + H * _S_ + H * _U_ * beta - H * _U_;
  const double th2 = tuvw - H * _S_ + H * _U_ * beta - H * _U_;
  const double aU = fabs(U), a1 = fabs(temp1), a2 = fabs(temp2);
  const double it = 1 / temp, i2t = 1 / (2 * temp), mi2t = -1 / (2 * temp);
  const double d11 = it * (4 * (K - 1) * _S_ * beta * aU + temp3 * a1 + temp4 * a2);
  const double d12 = mi2t * (Nx * rho * (a2 - a1) * temp5);
  const double d13 = mi2t * (Ny * rho * (a2 - a1) * temp5);
  const double d14 = mi2t * (Nz * rho * (a2 - a1) * temp5);
  const double d15 = -rho * aU / T;
  const double d21 = it * (u * _S_ * (4 * (K - 1) * beta * aU + a2 + a1) - (a2 - a1) * (2 * T * Nx * K * beta + u * _U_ * (beta - 1)));
  const double d22 = i2t * (rho * (8 * T * K * _S_ * beta * (Ny * Ny + Nz * Nz) * aU + Nx * (temp3 * tu2 * a2 + temp4 * tu1 * a1)));
  const double d23 = i2t * (rho * Ny * (-8 * T * K * _S_ * beta * Nx * aU + temp3 * tu2 * a2 + temp4 * tu1 * a1));
  const double d24 = i2t * (rho * Nz * (-8 * T * K * _S_ * beta * Nx * aU + temp3 * tu2 * a2 + temp4 * tu1 * a1));
  const double d25 = -u * rho * aU / T;
  const double d31 = it * (v * _S_ * (4 * (K - 1) * beta * aU + a2 + a1) - (a2 - a1) * (2 * T * Ny * K * beta + v * _U_ * (beta - 1)));
  const double d32 = i2t * (rho * Nx * (-8 * T * K * _S_ * beta * Ny * aU + temp3 * tv2 * a2 + temp4 * tv1 * a1));
  const double d33 = i2t * (rho * (8 * T * K * _S_ * beta * (Nx * Nx + Nz * Nz) * aU + Ny * (temp3 * tv2 * a2 + temp4 * tv1 * a1)));
  const double d34 = i2t * (rho * Nz * (-8 * T * K * _S_ * beta * Ny * aU + temp3 * tv2 * a2 + temp4 * tv1 * a1));
  const double d35 = -v * rho * aU / T;
  const double d41 = it * (w * _S_ * (4 * (K - 1) * beta * aU + a2 + a1) - (a2 - a1) * (2 * T * Nz * K * beta + w * _U_ * (beta - 1)));
  const double d42 = i2t * (rho * Nx * (-8 * T * K * _S_ * beta * Nz * aU + temp3 * tw2 * a2 + temp4 * tw1 * a1));
  const double d43 = i2t * (rho * Ny * (-8 * T * K * _S_ * beta * Nz * aU + temp3 * tw2 * a2 + temp4 * tw1 * a1));
  const double d44 = i2t * (rho * (8 * T * K * _S_ * beta * (Nx * Nx + Ny * Ny) * aU + Nz * (temp3 * tw2 * a2 + temp4 * tw1 * a1)));
  const double d45 = -w * rho * aU / T;
  const double d51 = it * (_S_ * (4 * beta * (H * K - H - T * K) * aU + H * (a2 + a1)) - (a2 - a1) * (tuvw + H * _U_ * beta - H * _U_));
  const double d52 = i2t * (rho * (8 * T * K * _S_ * beta * (-v * Nx * Ny - w * Nx * Nz + u * (Ny * Ny + Nz * Nz) * aU) + Nx * (th2 * temp3 * a2 + th1 * temp4 * a1)));
  const double d53 = i2t * (rho * (8 * T * K * _S_ * beta * (-u * Nx * Ny - w * Ny * Nz + v * (Nx * Nx + Nz * Nz) * aU) + Ny * (th2 * temp3 * a2 + th1 * temp4 * a1)));
  const double d54 = i2t * (rho * (8 * T * K * _S_ * beta * (-u * Nx * Nz - v * Ny * Nz + w * (Nx * Nx + Ny * Ny) * aU) + Nz * (th2 * temp3 * a2 + th1 * temp4 * a1)));
  const double d55 = (-H / T + K / (K - 1)) * rho * aU;
  const double Fav1 = d11 * dU1 + d12 * dU2 + d13 * dU3 + d14 * dU4 + d15 * dU5;
  const double Fav2 = d21 * dU1 + d22 * dU2 + d23 * dU3 + d24 * dU4 + d25 * dU5;
  const double Fav3 = d31 * dU1 + d32 * dU2 + d33 * dU3 + d34 * dU4 + d35 * dU5;
  const double Fav4 = d41 * dU1 + d42 * dU2 + d43 * dU3 + d44 * dU4 + d45 * dU5;
  const double Fav5 = d51 * dU1 + d52 * dU2 + d53 * dU3 + d54 * dU4 + d55 * dU5;
  F[0] = 0.5 * ((_rho * _U + rho_ * U_) - Ep * Fav1) / Jf;
  F[1] = 0.5 * ((_rho * _u * _U + rho_ * u_ * U_ + n[0] * (_P + P_)) - Ep * Fav2) / Jf;
  F[2] = 0.5 * ((_rho * _v * _U + rho_ * v_ * U_ + n[1] * (_P + P_)) - Ep * Fav3) / Jf;
  F[3] = 0.5 * ((_rho * _w * _U + rho_ * w_ * U_ + n[2] * (_P + P_)) - Ep * Fav4) / Jf;
  F[4] = 0.5 * ((_U * (3.5 * _P + 0.5 * _rho * _VV) + U_ * (3.5 * P_ + 0.5 * rho_ * VV_)) - Ep * Fav5) / Jf;
  *Fav3_out = Fav3;
  *Psum_out = _P + P_;
}


}  // namespace cf
