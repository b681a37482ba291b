// Golden-vector driver for the reference solver (test infrastructure only).
//
// Linked against the reference's own objects by oracle/ref_build.sh; it calls
// the reference functions exactly as main.cpp does with the dual-time path
// enabled (main.cpp:143-498 with Viscous_terms + Dual_time_stepping in place
// of RK3NODT) and dumps every array the hot path produces, so that the
// restatement in oracle/cf_oracle.c and the HIP path can be compared with the
// reference on identical inputs.
//
//   ref_driver unit <dir> <deltaT>
//       reads <dir>/initial.bin (Initial_condition.cpp:64-79 format) and the
//       optional <dir>/iterate.bin (5 fields, same layout) for the current
//       inner iterate U_; runs halo+BC, Flux_X/Y/Z, Viscous_terms,
//       Dual_time_stepping once, then the variants (Filter, limiter fluxes,
//       DPLUSGS, RK3NODT) on the same post-BC state; writes <name>.bin files.
//   ref_driver traj <dir> <deltaT> <steps> <iteration_end_step>
//       the main.cpp step loop; writes er.txt (one line per inner iteration:
//       step it er1..er5 raw and normalised) and LES%05d.bin via Output().
//   ref_driver force <dir> <deltaT>
//       Driven_force's printed line for the iterate and for level n.
//   ref_driver stat <dir> <deltaT>
//       Statistic's mean profiles over two steps and over one (st*.dat).
//   ref_driver shear <dir> <deltaT>
//       Statistic on the post-BC iterates iterate.bin, iterate2.bin,
//       iterate.bin, iterate2.bin (steps 1-4, statistic_step 3): the
//       Shear_stress.dat lines, including the reset after step 3.
//   ref_driver time <dir> <deltaT> <inner_iterations>
//       times the inner iteration (halo+BC, fluxes, viscous, dual time).
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <omp.h>
#include <mpi.h>
#include <unistd.h>
#include <string>

#include "main.h"

int X_np;  // main.h declares it extern for its default arguments

typedef double (*F3)[Y_m][Z_m];

static int g_gend;
static size_t g_n;  // doubles per field

static F3 fld() { return (F3)calloc(g_n, sizeof(double)); }

static void dump(const std::string& dir, const char* name, F3 a) {
  std::string p = dir + "/" + name + ".bin";
  FILE* f = fopen(p.c_str(), "wb");
  if (!f) { perror(p.c_str()); exit(3); }
  fwrite(a, sizeof(double), g_n, f);
  fclose(f);
}

static void load5(const std::string& path, F3 u[5], F3 J) {
  // 5 blocks of X_out*Y_out*Z_out Q values, stored as Q/J like the state
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return;
  size_t n = (size_t)X_out * Y_out * Z_out;
  double* buf = (double*)malloc(n * sizeof(double));
  for (int v = 0; v < 5; v++) {
    if (fread(buf, sizeof(double), n, f) != n) { fprintf(stderr, "short read %s\n", path.c_str()); exit(3); }
    for (int i = 0; i < X_out; i++)
      for (int j = 0; j < Y_out; j++)
        for (int k = 0; k < Z_out; k++)
          u[v][i + 3][j + 2][k + 2] = buf[((size_t)i * Y_out + j) * Z_out + k] / J[i + 3][j + 2][k + 2];
  }
  free(buf);
  fclose(f);
}

struct State {
  F3 U_[5], U[5], Uq[5];
  F3 J, m[9], Ju, mu[9], Jv, mv[9], Jw, mw[9];
  F3 Xp, Yp, Zp;
  F3 inFx[5], inFy[5], inFz[5], vF[4];
  F3 MR[5], ML[5], NR[5], NL[5], LR[5], LL[5], Res[5];
  F3 EpX, EpY, EpZ;
  int (*Nijk)[2];
  int (*ijk)[3];
};

static void alloc_state(State& s) {
  for (int v = 0; v < 5; v++) {
    s.U_[v] = fld(); s.U[v] = fld(); s.Uq[v] = fld();
    s.inFx[v] = fld(); s.inFy[v] = fld(); s.inFz[v] = fld();
    s.MR[v] = fld(); s.ML[v] = fld(); s.NR[v] = fld(); s.NL[v] = fld();
    s.LR[v] = fld(); s.LL[v] = fld(); s.Res[v] = fld();
  }
  for (int v = 0; v < 4; v++) s.vF[v] = fld();
  s.J = fld(); s.Ju = fld(); s.Jv = fld(); s.Jw = fld();
  for (int v = 0; v < 9; v++) { s.m[v] = fld(); s.mu[v] = fld(); s.mv[v] = fld(); s.mw[v] = fld(); }
  s.Xp = fld(); s.Yp = fld(); s.Zp = fld();
  s.EpX = fld(); s.EpY = fld(); s.EpZ = fld();
  s.Nijk = (int(*)[2])calloc((size_t)((X_np - 3) + ny + nz + 7) * 2, sizeof(int));
  s.ijk = (int(*)[3])calloc((size_t)(X_np - 6) * Y_out * Z_out * 3 * 3 + 3, sizeof(int));
}

static void setup(State& s, int myid) {
  Grid(myid, s.Xp, s.Yp, s.Zp);
  Hyperplane(myid, s.Nijk, s.ijk);
  Jacobian(myid, s.J, s.m[0], s.m[1], s.m[2], s.m[3], s.m[4], s.m[5], s.m[6], s.m[7], s.m[8]);
  Jacobian_u(myid, s.Ju, s.mu[0], s.mu[1], s.mu[2], s.mu[3], s.mu[4], s.mu[5], s.mu[6], s.mu[7], s.mu[8]);
  Jacobian_v(myid, s.Jv, s.mv[0], s.mv[1], s.mv[2], s.mv[3], s.mv[4], s.mv[5], s.mv[6], s.mv[7], s.mv[8],
             s.J, s.m[0], s.m[1], s.m[2], s.m[3], s.m[4], s.m[5], s.m[6], s.m[7], s.m[8]);
  Jacobian_w(myid, s.Jw, s.mw[0], s.mw[1], s.mw[2], s.mw[3], s.mw[4], s.mw[5], s.mw[6], s.mw[7], s.mw[8]);
  Initial_condition(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.U[0], s.U[1], s.U[2], s.U[3], s.U[4],
                    s.Uq[0], s.Uq[1], s.Uq[2], s.Uq[3], s.Uq[4], s.J);
  // main.cpp:112-127: the first inner iterate is level n
  for (int v = 0; v < 5; v++)
    for (int i = 3; i <= g_gend; i++)
      for (int j = 2; j < nyy; j++)
        for (int k = 2; k < nzz; k++) s.U_[v][i][j][k] = s.U[v][i][j][k];
}

// main.cpp:156-240 (halo + periodic x) with the same MPI calls, then the
// wall and periodic-z ghosts (main.cpp:247-250)
static void halo_bc(State& s, int myid) {
  MPI_Comm comm = MPI_COMM_WORLD;
  MPI_Status istat[8];
  int nproc = 1, istart = 3, iend = g_gend;
  int l_nbr = myid - 1, r_nbr = myid + 1;
  if (myid == 0) l_nbr = MPI_PROC_NULL;
  if (myid == nproc - 1) r_nbr = MPI_PROC_NULL;
  int icount = 3 * Y_m * Z_m;
  for (int v = 0; v < 5; v++) {
    MPI_Sendrecv((void*)&s.U_[v][istart][0][0], icount, MPI_DOUBLE, l_nbr, 110 + 10 * v,
                 (void*)&s.U_[v][iend + 1][0][0], icount, MPI_DOUBLE, r_nbr, 110 + 10 * v, comm, istat);
  }
  for (int v = 0; v < 5; v++) {
    MPI_Sendrecv((void*)&s.U_[v][iend - 2][0][0], icount, MPI_DOUBLE, r_nbr, 160 + 10 * v,
                 (void*)&s.U_[v][istart - 3][0][0], icount, MPI_DOUBLE, l_nbr, 160 + 10 * v, comm, istat);
  }
  MPI_Barrier(comm);
  icount = Y_m * Z_m;
  if (myid == 0)
    for (int v = 0; v < 5; v++)
      MPI_Sendrecv((void*)&s.U_[v][3][0][0], icount, MPI_DOUBLE, nproc - 1, 310 + 10 * v,
                   (void*)&s.U_[v][2][0][0], icount, MPI_DOUBLE, nproc - 1, 310 + 10 * v, comm, istat);
  if (myid == nproc - 1)
    for (int v = 0; v < 5; v++)
      MPI_Sendrecv((void*)&s.U_[v][iend][0][0], icount, MPI_DOUBLE, 0, 310 + 10 * v,
                   (void*)&s.U_[v][iend + 1][0][0], icount, MPI_DOUBLE, 0, 310 + 10 * v, comm, istat);
  MPI_Barrier(comm);
  Y_boundary_condition(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.J);
  Z_boundary_condition(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.J);
}

static void fluxes(State& s, int myid) {
  Flux_X(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.MR[0], s.MR[1], s.MR[2], s.MR[3], s.MR[4],
         s.ML[0], s.ML[1], s.ML[2], s.ML[3], s.ML[4], s.J, s.Ju, s.mu[0], s.mu[1], s.mu[2], s.mu[3], s.mu[4],
         s.mu[5], s.mu[6], s.mu[7], s.mu[8], s.inFx[0], s.inFx[1], s.inFx[2], s.inFx[3], s.inFx[4], s.EpX);
  Flux_Y(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.MR[0], s.MR[1], s.MR[2], s.MR[3], s.MR[4],
         s.ML[0], s.ML[1], s.ML[2], s.ML[3], s.ML[4], s.J, s.Jv, s.mv[0], s.mv[1], s.mv[2], s.mv[3], s.mv[4],
         s.mv[5], s.mv[6], s.mv[7], s.mv[8], s.inFy[0], s.inFy[1], s.inFy[2], s.inFy[3], s.inFy[4], s.EpY);
  Flux_Z(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.MR[0], s.MR[1], s.MR[2], s.MR[3], s.MR[4],
         s.ML[0], s.ML[1], s.ML[2], s.ML[3], s.ML[4], s.J, s.Jw, s.mw[0], s.mw[1], s.mw[2], s.mw[3], s.mw[4],
         s.mw[5], s.mw[6], s.mw[7], s.mw[8], s.inFz[0], s.inFz[1], s.inFz[2], s.inFz[3], s.inFz[4], s.EpZ);
}

static void fluxes_limiter(State& s, int myid) {
  Flux_X_limiter(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.MR[0], s.MR[1], s.MR[2], s.MR[3], s.MR[4],
                 s.ML[0], s.ML[1], s.ML[2], s.ML[3], s.ML[4], s.J, s.Ju, s.mu[0], s.mu[1], s.mu[2], s.mu[3],
                 s.mu[4], s.mu[5], s.mu[6], s.mu[7], s.mu[8], s.inFx[0], s.inFx[1], s.inFx[2], s.inFx[3],
                 s.inFx[4], s.EpX);
  Flux_Y_limiter(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.MR[0], s.MR[1], s.MR[2], s.MR[3], s.MR[4],
                 s.ML[0], s.ML[1], s.ML[2], s.ML[3], s.ML[4], s.J, s.Jv, s.mv[0], s.mv[1], s.mv[2], s.mv[3],
                 s.mv[4], s.mv[5], s.mv[6], s.mv[7], s.mv[8], s.inFy[0], s.inFy[1], s.inFy[2], s.inFy[3],
                 s.inFy[4], s.EpY);
  Flux_Z_limiter(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.MR[0], s.MR[1], s.MR[2], s.MR[3], s.MR[4],
                 s.ML[0], s.ML[1], s.ML[2], s.ML[3], s.ML[4], s.J, s.Jw, s.mw[0], s.mw[1], s.mw[2], s.mw[3],
                 s.mw[4], s.mw[5], s.mw[6], s.mw[7], s.mw[8], s.inFz[0], s.inFz[1], s.inFz[2], s.inFz[3],
                 s.inFz[4], s.EpZ);
}

static void viscous(State& s, int myid) {
  Viscous_terms(myid, s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.vF[0], s.vF[1], s.vF[2], s.vF[3], s.J,
                s.m[0], s.m[1], s.m[2], s.m[3], s.m[4], s.m[5], s.m[6], s.m[7], s.m[8], s.Ju, s.mu[0], s.mu[1],
                s.mu[2], s.mu[3], s.mu[4], s.mu[5], s.mu[6], s.mu[7], s.mu[8], s.Jv, s.mv[0], s.mv[1], s.mv[2],
                s.mv[3], s.mv[4], s.mv[5], s.mv[6], s.mv[7], s.mv[8], s.Jw, s.mw[0], s.mw[1], s.mw[2], s.mw[3],
                s.mw[4], s.mw[5], s.mw[6], s.mw[7], s.mw[8], s.LR[0], s.LR[1], s.LR[2], s.LR[3], s.LR[4],
                s.LL[0], s.LL[1], s.LL[2], s.LL[3], s.LL[4], s.MR[0], s.MR[1], s.MR[2], s.MR[3], s.MR[4],
                s.ML[0], s.ML[1], s.ML[2], s.ML[3], s.ML[4], s.NR[0], s.NR[1], s.NR[2], s.NR[3], s.NR[4],
                s.NL[0], s.NL[1], s.NL[2], s.NL[3], s.NL[4]);
}

static void dual_time(State& s, int myid, double deltaT, double er[5]) {
  Dual_time_stepping(myid, deltaT, &er[0], &er[1], &er[2], &er[3], &er[4], s.Nijk, s.ijk, s.U_[0], s.U_[1],
                     s.U_[2], s.U_[3], s.U_[4], s.U[0], s.U[1], s.U[2], s.U[3], s.U[4], s.Uq[0], s.Uq[1],
                     s.Uq[2], s.Uq[3], s.Uq[4], s.J, s.m[0], s.m[1], s.m[2], s.m[3], s.m[4], s.m[5], s.m[6],
                     s.m[7], s.m[8], s.inFx[0], s.inFx[1], s.inFx[2], s.inFx[3], s.inFx[4], s.inFy[0],
                     s.inFy[1], s.inFy[2], s.inFy[3], s.inFy[4], s.inFz[0], s.inFz[1], s.inFz[2], s.inFz[3],
                     s.inFz[4], s.vF[0], s.vF[1], s.vF[2], s.vF[3], s.MR[0], s.MR[1], s.MR[2], s.MR[3],
                     s.MR[4], s.NR[0], s.NR[1], s.NR[2], s.NR[3], s.NR[4], s.NL[0], s.NL[1], s.NL[2], s.NL[3],
                     s.NL[4], s.Res[0], s.Res[1], s.Res[2], s.Res[3], s.Res[4]);
}

static void dplusgs(State& s, int myid, double deltaT, double er[5]) {
  DPLUSGS(myid, deltaT, &er[0], &er[1], &er[2], &er[3], &er[4], s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4],
          s.U[0], s.U[1], s.U[2], s.U[3], s.U[4], s.Uq[0], s.Uq[1], s.Uq[2], s.Uq[3], s.Uq[4], s.J, s.m[0],
          s.m[1], s.m[2], s.m[3], s.m[4], s.m[5], s.m[6], s.m[7], s.m[8], s.inFx[0], s.inFx[1], s.inFx[2],
          s.inFx[3], s.inFx[4], s.inFy[0], s.inFy[1], s.inFy[2], s.inFy[3], s.inFy[4], s.inFz[0], s.inFz[1],
          s.inFz[2], s.inFz[3], s.inFz[4], s.vF[0], s.vF[1], s.vF[2], s.vF[3], s.MR[0], s.MR[1], s.MR[2],
          s.MR[3], s.MR[4], s.NR[0], s.NR[1], s.NR[2], s.NR[3], s.NR[4], s.NL[0], s.NL[1], s.NL[2], s.NL[3],
          s.NL[4], s.Res[0], s.Res[1], s.Res[2], s.Res[3], s.Res[4]);
}

static void statistic(State& s, int myid, F3 u[5], int step, int statistic_step) {
  Statistic(myid, step, 1, statistic_step, u[0], u[1], u[2], u[3], u[4], s.J, s.Jv, s.mv[0], s.mv[1], s.mv[2],
            s.mv[3], s.mv[4], s.mv[5], s.mv[6], s.mv[7], s.mv[8], s.MR[0], s.MR[1], s.MR[2], s.MR[3], s.MR[4],
            s.ML[0], s.ML[1], s.ML[2], s.ML[3], s.ML[4], s.EpY);
}

static void copy5(F3 dst[5], F3 src[5]) {
  for (int v = 0; v < 5; v++) memcpy(dst[v], src[v], g_n * sizeof(double));
}

static void dump5(const std::string& dir, const char* base, F3 a[5], int n = 5) {
  char name[64];
  for (int v = 0; v < n; v++) {
    snprintf(name, sizeof name, "%s%d", base, v + 1);
    dump(dir, name, a[v]);
  }
}

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  int myid = 0;
  MPI_Comm_rank(MPI_COMM_WORLD, &myid);
  if (argc < 4) { fprintf(stderr, "usage: ref_driver unit|traj|filter|limiter|stat|force|time <dir> <deltaT> ...\n"); return 2; }
  std::string mode = argv[1], dir = argv[2];
  double deltaT = atof(argv[3]);
  X_np = X_out + 6;
  g_gend = X_out + 2;
  g_n = (size_t)X_np * Y_m * Z_m;
  if (chdir(dir.c_str()) != 0) { perror("chdir"); return 3; }
  dir = ".";
  State s;
  alloc_state(s);
  setup(s, myid);

  if (mode == "unit") {
    load5("iterate.bin", s.U_, s.J);
    dump(dir, "J", s.J); dump(dir, "J_u", s.Ju); dump(dir, "J_v", s.Jv); dump(dir, "J_w", s.Jw);
    const char* mn[9] = {"xidx", "xidy", "xidz", "etdx", "etdy", "etdz", "ztdx", "ztdy", "ztdz"};
    char name[64];
    for (int v = 0; v < 9; v++) {
      dump(dir, mn[v], s.m[v]);
      snprintf(name, sizeof name, "%s_u", mn[v]); dump(dir, name, s.mu[v]);
      snprintf(name, sizeof name, "%s_v", mn[v]); dump(dir, name, s.mv[v]);
      snprintf(name, sizeof name, "%s_w", mn[v]); dump(dir, name, s.mw[v]);
    }
    dump(dir, "Y_point", s.Yp);
    {
      FILE* f = fopen("hyperplane.txt", "w");
      for (int m = 7; m <= g_gend + ny + nz; m++) fprintf(f, "%d %d %d\n", m, s.Nijk[m][0], s.Nijk[m][1]);
      fclose(f);
    }
    dump5(dir, "U_pre", s.U_);
    dump5(dir, "U", s.U);
    dump5(dir, "Uq", s.Uq);
    halo_bc(s, myid);
    dump5(dir, "U_bc", s.U_);
    F3 Ubc[5];
    for (int v = 0; v < 5; v++) Ubc[v] = fld();
    copy5(Ubc, s.U_);
    fluxes(s, myid);
    dump5(dir, "inFx", s.inFx); dump5(dir, "inFy", s.inFy); dump5(dir, "inFz", s.inFz);
    viscous(s, myid);
    dump5(dir, "vF", s.vF, 4);
    double er[5];
    dual_time(s, myid, deltaT, er);
    dump5(dir, "U_dt", s.U_);
    dump5(dir, "MR_dt", s.MR); dump5(dir, "NL_dt", s.NL); dump5(dir, "NR_dt", s.NR);
    FILE* f = fopen("scalars.txt", "w");
    fprintf(f, "er_dt %.17g %.17g %.17g %.17g %.17g\n", er[0], er[1], er[2], er[3], er[4]);

    // variants on the same post-BC state (fluxes/viscous recomputed because
    // Dual_time_stepping reuses MR/NL/NR as work space)
    copy5(s.U_, Ubc);
    Filter(myid, 0.007, 1., 0., s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.Yp, s.EpX, s.EpY, s.EpZ, s.J);
    dump(dir, "EpX", s.EpX); dump(dir, "EpY", s.EpY); dump(dir, "EpZ", s.EpZ);
    copy5(s.U_, Ubc);
    fluxes_limiter(s, myid);
    dump5(dir, "inFxL", s.inFx); dump5(dir, "inFyL", s.inFy); dump5(dir, "inFzL", s.inFz);
    // limiter fluxes again with a smooth sensor field in [0,1] so that the
    // dissipation matrix is not multiplied away
    for (int i = 0; i < X_np; i++)
      for (int j = 0; j < Y_m; j++)
        for (int k = 0; k < Z_m; k++) {
          double t = 0.5 + 0.5 * sin(0.7 * i + 1.3 * j + 0.4 * k);
          s.EpX[i][j][k] = t; s.EpY[i][j][k] = 1.0 - 0.5 * t; s.EpZ[i][j][k] = 0.25 + 0.5 * t;
        }
    dump(dir, "EpX_s", s.EpX); dump(dir, "EpY_s", s.EpY); dump(dir, "EpZ_s", s.EpZ);
    copy5(s.U_, Ubc);
    fluxes_limiter(s, myid);
    dump5(dir, "inFxS", s.inFx); dump5(dir, "inFyS", s.inFy); dump5(dir, "inFzS", s.inFz);
    // RK3NODT stage 1 on the MUSCL/Roe fluxes and viscous terms of the state
    copy5(s.U_, Ubc);
    fluxes(s, myid);
    viscous(s, myid);
    for (int RK = 1; RK <= 3; RK++) {
      copy5(s.U_, Ubc);
      RK3NODT(myid, RK, deltaT, &er[0], &er[1], &er[2], &er[3], &er[4], s.U_[0], s.U_[1], s.U_[2], s.U_[3],
              s.U_[4], s.U[0], s.U[1], s.U[2], s.U[3], s.U[4], s.Uq[0], s.Uq[1], s.Uq[2], s.Uq[3], s.Uq[4], s.J,
              s.m[0], s.m[1], s.m[2], s.m[3], s.m[4], s.m[5], s.m[6], s.m[7], s.m[8], s.inFx[0], s.inFx[1],
              s.inFx[2], s.inFx[3], s.inFx[4], s.inFy[0], s.inFy[1], s.inFy[2], s.inFy[3], s.inFy[4], s.inFz[0],
              s.inFz[1], s.inFz[2], s.inFz[3], s.inFz[4], s.vF[0], s.vF[1], s.vF[2], s.vF[3], s.MR[0], s.MR[1],
              s.MR[2], s.MR[3], s.MR[4]);
      char nm[32];
      snprintf(nm, sizeof nm, "U_rk%d_", RK);
      dump5(dir, nm, s.U_);
    }
    fclose(f);
  } else if (mode == "traj" && argc >= 6) {
    int count = atoi(argv[4]);
    int iteration_end_step = atoi(argv[5]);
    double E[5] = {0, 0, 0, 0, 0};
    FILE* f = fopen("er.txt", "w");
    for (int step = 1; step <= count; step++) {
      for (int iteration = 1; iteration < 50000; iteration++) {
        double er[5];
        halo_bc(s, myid);
        fluxes(s, myid);
        viscous(s, myid);
        dual_time(s, myid, deltaT, er);
        double raw[5];
        for (int v = 0; v < 5; v++) { raw[v] = er[v]; E[v] = fmax(E[v], er[v]); er[v] = er[v] / E[v]; }
        fprintf(f, "%d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", step, iteration,
                raw[0], raw[1], raw[2], raw[3], raw[4], er[0], er[1], er[2], er[3], er[4]);
        // main.cpp:406 (bitwise & of the comparisons, as written)
        if ((er[0] < 0.001 & er[1] < 0.001 & er[2] < 0.001 & er[3] < 0.001) | iteration == iteration_end_step) {
          for (int v = 0; v < 5; v++)
            for (int i = 3; i <= g_gend; i++)
              for (int j = 2; j < nyy; j++)
                for (int k = 2; k < nzz; k++) {
                  s.Uq[v][i][j][k] = s.U[v][i][j][k];
                  s.U[v][i][j][k] = s.U_[v][i][j][k];
                }
          break;
        }
      }
      Output(step, myid, s.U[0], s.U[1], s.U[2], s.U[3], s.U[4], s.Uq[0], s.Uq[1], s.Uq[2], s.Uq[3], s.Uq[4], s.J);
    }
    fclose(f);
  } else if (mode == "filter" && argc >= 5) {
    // the reference Filter on the post-BC iterate once per threshold listed
    // in <crit file> (one %la per line); EpX/EpY/EpZ of run r go to
    // EpX_c<r>.bin ...  A threshold equal to a ratio computed elsewhere and
    // its successor show whether the reference computes that same ratio.
    load5("iterate.bin", s.U_, s.J);
    halo_bc(s, myid);
    FILE* cf = fopen(argv[4], "r");
    if (!cf) { perror("crit file"); return 3; }
    double crit;
    int r = 0;
    char name[64];
    while (fscanf(cf, "%la", &crit) == 1) {
      Filter(myid, crit, 1., 0., s.U_[0], s.U_[1], s.U_[2], s.U_[3], s.U_[4], s.Yp, s.EpX, s.EpY, s.EpZ, s.J);
      snprintf(name, sizeof name, "EpX_c%d", r); dump(dir, name, s.EpX);
      snprintf(name, sizeof name, "EpY_c%d", r); dump(dir, name, s.EpY);
      snprintf(name, sizeof name, "EpZ_c%d", r); dump(dir, name, s.EpZ);
      r++;
    }
    fclose(cf);
  } else if (mode == "limiter") {
    // the limiter fluxes on the post-BC iterate with the smooth sensor field
    // of the unit mode; the sensors are dumped as well (EpX_s ..)
    load5("iterate.bin", s.U_, s.J);
    halo_bc(s, myid);
    for (int i = 0; i < X_np; i++)
      for (int j = 0; j < Y_m; j++)
        for (int k = 0; k < Z_m; k++) {
          double t = 0.5 + 0.5 * sin(0.7 * i + 1.3 * j + 0.4 * k);
          s.EpX[i][j][k] = t; s.EpY[i][j][k] = 1.0 - 0.5 * t; s.EpZ[i][j][k] = 0.25 + 0.5 * t;
        }
    dump(dir, "EpX_s", s.EpX); dump(dir, "EpY_s", s.EpY); dump(dir, "EpZ_s", s.EpZ);
    fluxes_limiter(s, myid);
    dump5(dir, "inFxS", s.inFx); dump5(dir, "inFyS", s.inFy); dump5(dir, "inFzS", s.inFz);
  } else if (mode == "stat") {
    // Statistic.cpp:1043-1154, the mean profiles: step 1 accumulates the
    // post-BC iterate and step 2 the level-n state with statistic_step = 2
    // (st00002.dat, their average); step 3 with statistic_step = 1 writes the
    // iterate alone (st00003.dat)
    load5("iterate.bin", s.U_, s.J);
    halo_bc(s, myid);
    statistic(s, myid, s.U_, 1, 2);
    statistic(s, myid, s.U, 2, 2);
    statistic(s, myid, s.U_, 3, 1);
  } else if (mode == "shear") {
    // Statistic.cpp:552-911 (Shear_stress.dat) and its reset at the
    // statistic step (Statistic.cpp:1008-1018)
    const char* files[4] = {"iterate.bin", "iterate2.bin", "iterate.bin", "iterate2.bin"};
    for (int step = 1; step <= 4; step++) {
      load5(files[step - 1], s.U_, s.J);
      halo_bc(s, myid);
      statistic(s, myid, s.U_, step, 3);
    }
  } else if (mode == "force") {
    // Driven_force.cpp:38-73 on the post-BC iterate, then on level n: each
    // call prints its Q0 / Qall / ER_Q line
    load5("iterate.bin", s.U_, s.J);
    halo_bc(s, myid);
    Driven_force(myid, s.U_[1], s.J);
    Driven_force(myid, s.U[1], s.J);
  } else if (mode == "time" && argc >= 5) {
    int nit = atoi(argv[4]);
    double er[5];
    double t0 = omp_get_wtime();
    for (int it = 0; it < nit; it++) {
      halo_bc(s, myid);
      fluxes(s, myid);
      viscous(s, myid);
      dual_time(s, myid, deltaT, er);
    }
    double dt = omp_get_wtime() - t0;
    printf("{\"threads\": %d, \"inner_iterations\": %d, \"seconds\": %.6f, \"er1\": %.17g}\n",
           omp_get_max_threads(), nit, dt, er[0]);
  } else {
    fprintf(stderr, "bad mode\n");
    return 2;
  }
  MPI_Finalize();
  return 0;
}
