// C ABI of the channel-flow dual-time path (include/channel_flow.h).
//
// Host side of the MI355X build: owns one GPU's device memory for a runtime
// grid, moves reference-layout host arrays in and out, and sequences the
// kernels of one inner iteration on the context's stream.  Replaces the
// driver in main.cpp:143-498 and the call surface of main.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/channel_flow.h"
#include "cf_kernels.hpp"
#include "cf_pred_internal.hpp"

using cf::Geo;

namespace {
constexpr int NPROF = cf::PROF_METRIC_ROWS;  // 4 sets x (J + 9 metrics), uploaded from the host
}

// ---------------------------------------------------------------------------
// Scratch of cf_driven_force_noalloc: per-i-plane partial sums of the bulk
// flux in scratch[0, CF_NOALLOC_MAX_PLANES) and their in-order sum in
// scratch[CF_NOALLOC_MAX_PLANES].  Each GPU holds its own copy of the array
// until the process ends (4.3 KB).  Host code never passes `scratch` itself
// to a kernel: its device address is looked up once per device (in
// cf_create, with that device current) and kept in scratch_dev[device].
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[CF_NOALLOC_MAX_PLANES + 32];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_dev[MAX_DEVICES] = {};

// device address of `scratch` on the current device `device`, looked up once
hipError_t scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_dev[device]) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  scratch_dev[device] = static_cast<double*>(p);
  return hipSuccess;
}

double* scratch_of(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(scratch_mu);
  return scratch_dev[device];
}
}  // namespace

struct cf_ctx {
  cf_params p{};
  Geo g{};
  hipStream_t stream = nullptr;
  int device = 0;
  // device fields (each nf doubles, padded layout)
  double* Ui = nullptr;   // U_ : inner iterate (5)
  double* Un = nullptr;   // U  : level n (5)
  double* Uq = nullptr;   // Uq : level n-1 (5)
  double* Fx = nullptr;   // inFx (5)
  double* Fy = nullptr;   // inFy (5)
  double* Fz = nullptr;   // inFz (5)
  double* vF = nullptr;   // vF2..vF5 (4)
  double* MR = nullptr;   // (5)
  double* NL = nullptr;   // (5)
  double* NR = nullptr;   // (5)
  double* Pr = nullptr;   // viscous primitives U,V,W,T (4)
  double* VS = nullptr;   // viscous face fluxes LL/ML/NL (12)
  double* Ep = nullptr;   // limiter sensors EpX, EpY, EpZ on the face layouts (3)
  double* Fw = nullptr;   // Filter workspace (FILTER_WORK_FIELDS), allocated on first use
  double* ypt = nullptr;  // Y_point j-profile (Y_m)
  double* Sm = nullptr;   // sweep-order fields: the per-cell summary (G_N) and MR, NL, NR
  int* order = nullptr;   // sweep tiles in dependency order
  int* sync = nullptr;    // ticket, error flag, per-tile progress (x2 directions)
  int ntiles = 0;
  int schedule = 0;       // 0: persistent tiles, 1: one hyperplane per launch
  int solver = 0;         // implicit part: 0 LU-SGS (Dual_time_stepping), 1 DPLUSGS
  int visc_schedule = 0;  // 0: fused viscous kernel, 1: three kernels through Pr/VS (allocated on first use)
  long prof_gen = 0;  // metric uploads so far
  bool fp32 = false;      // CF_FP32_SWEEPS: LU-SGS sweeps and split Jacobians in float
  int guess_order = 0;    // initial iterate of cf_physical_step: 0 U^n, 1 linear, 2 quadratic extrapolation
  double* Uqq = nullptr;  // U^{n-2} (5 fields), kept while guess_order is 2 (allocated on first use)
  double* Hs = nullptr;   // 15 fields that stand in for U_, U, Uq while a host-array entry point runs on a
                          // context that holds a device-resident state (allocated on first such call)
  bool uqq_valid = false; // Uqq holds the level before Uq
  double* U3 = nullptr;   // U^{n-3} (5 fields), kept while four levels are kept (cf_keep_levels)
  bool u3_valid = false;  // U3 holds the level before Uqq
  // highest order of the extrapolated initial iterate (guess order 2 and the
  // predictor's fallback): 2 quadratic, 3 cubic (keeps U3)
  int extrap_order = 2;
  // learned initial iterate (cf_set_predictor; guess_order 3)
  cf_pred* pred = nullptr;
  unsigned short* px = nullptr;  // features of the whole domain (bf16, 16 per cell)
  float* py = nullptr;           // network outputs (5 per cell)
  double* Ualt = nullptr;        // the learned guess after its first inner iteration (5 fields)
  long pred_steps = 0, pred_fallbacks = 0;
  // LU-SGS sweep pairs executed: one for any n_sweeps >= 1 while sweep_reuse
  // is on (a single domain's later pairs reproduce the first bit for bit)
  bool sweep_reuse = true;
  // hipGraph of one inner iteration's kernels (cf_set_graph), rebuilt when
  // what it was captured with changes
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  struct GraphKey {
    const double *Ui, *Un, *Uq;
    int schedule, visc_schedule, solver, x_periodic, pairs;
    bool fp32, fork;
    double deltaT, f_drive;
    long prof_gen;  // launch choices that follow the metrics (the fused viscous kernel)
    long spin_limit;
    bool operator==(const GraphKey& o) const {
      return Ui == o.Ui && Un == o.Un && Uq == o.Uq && schedule == o.schedule && visc_schedule == o.visc_schedule &&
             solver == o.solver && x_periodic == o.x_periodic && pairs == o.pairs && fp32 == o.fp32 &&
             fork == o.fork && deltaT == o.deltaT && f_drive == o.f_drive && prof_gen == o.prof_gen &&
             spin_limit == o.spin_limit;
    }
  } gkey{};
  long graph_builds = 0;
  // the flux and viscous kernels of an inner iteration are independent: with
  // `fork` on they run on the context's stream and three helper streams
  // (default priority), joined by events, eagerly and inside the graph
  bool fork = true;
  hipStream_t aux[3] = {nullptr, nullptr, nullptr};
  hipEvent_t fork_ev = nullptr, join_ev[3] = {nullptr, nullptr, nullptr};
  long spin_limit = 1L << 22;
  // what the sweep-order NL slots that the forward sweep hands across tiles
  // hold between iterations: 1 the double sentinel, 2 the float one, 0 unknown
  int hand_off = 0;
  double* partial = nullptr;
  double* er_dev = nullptr;   // the norms, then the two sweep error flags
  double* er_host = nullptr;  // pinned copy of er_dev, read after each iteration
  double* stat = nullptr;  // running per-j sums of Statistic's mean profiles (7 * Ym)
  double* shear = nullptr; // wall values of Statistic's y-face viscous flux (2 * X_out * Z_out)
  // Shear_stress.dat: running sums SXYm[1], SXYm[ny] and the previous running
  // averages SXY_bottom, SXY_up (Statistic.cpp:92-97, 879-911)
  double sxy_sum[2] = {0, 0}, sxy_prev[2] = {0, 0};
  double* prof = nullptr;  // metric j-profiles (NPROF * Ym)
  std::vector<double> hprof;
  std::vector<double> staging;  // reference-layout staging for 2D copies
  bool have_metrics = false, have_state = false;
  double E[5] = {0, 0, 0, 0, 0};  // running maxima of er (main.cpp:63-67)
  std::string err;
};

static int fail(cf_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail((c), CF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static long host_elems(const cf_ctx* c) { return (long)c->g.Xnp * c->g.Ym * c->g.Zm; }

// reference layout (Z_m pitch) <-> device layout (Zp pitch), one field
static int h2d(cf_ctx* c, double* dst, const double* src) {
  HIPCHK(c, hipMemcpy2DAsync(dst, (size_t)c->g.Zp * 8, src, (size_t)c->g.Zm * 8, (size_t)c->g.Zm * 8,
                             (size_t)c->g.Xnp * c->g.Ym, hipMemcpyHostToDevice, c->stream));
  return CF_OK;
}
static int d2h(cf_ctx* c, double* dst, const double* src) {
  HIPCHK(c, hipMemcpy2DAsync(dst, (size_t)c->g.Zm * 8, src, (size_t)c->g.Zp * 8, (size_t)c->g.Zm * 8,
                             (size_t)c->g.Xnp * c->g.Ym, hipMemcpyDeviceToHost, c->stream));
  return CF_OK;
}
static int h2d_n(cf_ctx* c, double* dst, const double* const* src, int n) {
  for (int v = 0; v < n; v++) {
    if (!src[v]) return fail(c, CF_ERR_ARG, "null input array %d", v);
    int r = h2d(c, dst + v * c->g.nf, src[v]);
    if (r) return r;
  }
  return CF_OK;
}
static int d2h_n(cf_ctx* c, double* const* dst, const double* src, int n) {
  for (int v = 0; v < n; v++) {
    if (!dst[v]) continue;
    int r = d2h(c, dst[v], src + v * c->g.nf);
    if (r) return r;
  }
  return CF_OK;
}
static int sync(cf_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipGetLastError());
  return CF_OK;
}

// ---------------------------------------------------------------------------
// metric j-profiles: the reference metrics (Jacobian*.cpp) depend only on j.
// A 3-D array is reduced to its j-profile after checking that it is constant
// along i and k over the region where the kernels read it.
// ---------------------------------------------------------------------------
struct Region {
  int i0, i1, j0, j1, k0, k1;
};

static Region region_of(const cf_ctx* c, int set, int comp) {
  const Geo& g = c->g;
  int gend = g.gend, ny = g.ny, nz = g.nz;
  switch (set) {
    case cf::SET_C:
      if (comp == 0) return {2, gend + 1, 1, ny + 1, 1, nz + 1};
      return {3, gend, 2, ny, 2, nz};
    case cf::SET_U: return {2, gend, 1, ny - 1, 1, nz - 1};
    case cf::SET_V: return {2, gend - 1, 1, ny, 1, nz - 1};
    default: return {2, gend - 1, 1, ny - 1, 1, nz};
  }
}

static int ingest(cf_ctx* c, const double* A, int set, int comp, const char* name) {
  if (!A) return fail(c, CF_ERR_ARG, "null metric array %s", name);
  const Geo& g = c->g;
  Region r = region_of(c, set, comp);
  double* prof = &c->hprof[(size_t)(set * 10 + comp) * g.Ym];
  for (int j = r.j0; j <= r.j1; j++) {
    double v0 = A[((size_t)r.i0 * g.Ym + j) * g.Zm + r.k0];
    for (int i = r.i0; i <= r.i1; i++)
      for (int k = r.k0; k <= r.k1; k++) {
        double v = A[((size_t)i * g.Ym + j) * g.Zm + k];
        if (!(v == v0 || (std::isnan(v) && std::isnan(v0))))
          return fail(c, CF_ERR_METRICS, "%s varies along i/k at (%d,%d,%d): %.17g vs %.17g", name, i, j, k, v, v0);
      }
    prof[j] = v0;
  }
  return CF_OK;
}

static int upload_prof(cf_ctx* c) {
  c->prof_gen++;
  HIPCHK(c, hipMemcpyAsync(c->prof, c->hprof.data(), c->hprof.size() * 8, hipMemcpyHostToDevice, c->stream));
  cf::launch_flux_tables(c->stream, c->prof, c->g);
  HIPCHK(c, hipGetLastError());
  return CF_OK;
}

static int ingest_set(cf_ctx* c, int set, const double* J, const double* const m[9]) {
  static const char* nm[9] = {"xidx", "xidy", "xidz", "etdx", "etdy", "etdz", "ztdx", "ztdy", "ztdz"};
  int r = ingest(c, J, set, 0, "J");
  if (r) return r;
  if (m)
    for (int q = 0; q < 9; q++) {
      r = ingest(c, m[q], set, 1 + q, nm[q]);
      if (r) return r;
    }
  return CF_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int cf_create(const cf_params* p, cf_ctx** out) {
  if (!p || !out) return CF_ERR_ARG;
  *out = nullptr;
  if (p->X_out < 8 || p->Y_out < 8 || p->Z_out < 8 || p->X_out > 1000 || p->Y_out > 1000 || p->Z_out > 1000)
    return CF_ERR_ARG;
  if (!(p->deltaXI > 0) || !(p->deltaET > 0) || !(p->deltaZT > 0) || !(p->deltaT > 0)) return CF_ERR_ARG;
  if (p->n_sweeps < 1 || p->iteration_end_step < 1) return CF_ERR_ARG;
  if (p->precision != CF_FP64 && p->precision != CF_FP32_SWEEPS) return CF_ERR_ARG;
  cf_ctx* c = new cf_ctx();
  c->p = *p;
  c->fp32 = p->precision == CF_FP32_SWEEPS;
  Geo& g = c->g;
  g.Xnp = p->X_out + 6;
  g.Ym = p->Y_out + 4;
  g.Zm = p->Z_out + 4;
  g.Zp = (g.Zm + 7) / 8 * 8;
  g.gend = p->X_out + 2;
  g.ny = p->Y_out + 1;
  g.nz = p->Z_out + 1;
  g.plane = (long)g.Ym * g.Zp;
  g.nf = (long)g.Xnp * g.plane;
  g.dXI = p->deltaXI;
  g.dET = p->deltaET;
  g.dZT = p->deltaZT;
  // 0 where the residual-corrected division is not guaranteed: plain division
  g.rXI = cf::rcp_checked(p->deltaXI);
  g.rET = cf::rcp_checked(p->deltaET);
  g.rZT = cf::rcp_checked(p->deltaZT);
  c->hprof.assign((size_t)NPROF * g.Ym, 0.0);
  auto bail = [&](int code) {
    cf_destroy(c);
    return code;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipError_t e = scratch_init(c->device); e != hipSuccess) {
    fprintf(stderr, "cf_create: hipGetSymbolAddress(scratch) on device %d: %s\n", c->device, hipGetErrorString(e));
    return bail(CF_ERR_HIP);
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(CF_ERR_HIP);
  size_t nf = (size_t)g.nf;
  struct {
    double** ptr;
    int n;
  } fields[] = {{&c->Ui, 5}, {&c->Un, 5}, {&c->Uq, 5}, {&c->Fx, 5}, {&c->Fy, 5}, {&c->Fz, 5},
                {&c->vF, 4}, {&c->MR, 5}, {&c->NL, 5}, {&c->NR, 5}, {&c->Ep, 3}};
  for (auto& f : fields) {
    if (hipMalloc(f.ptr, nf * f.n * sizeof(double)) != hipSuccess) return bail(CF_ERR_HIP);
    // the reference leaves work arrays uninitialised; here they start at zero
    if (hipMemsetAsync(*f.ptr, 0, nf * f.n * sizeof(double), c->stream) != hipSuccess) return bail(CF_ERR_HIP);
  }
  if (hipMalloc(&c->Sm, (size_t)cf::sweep_field_elems(g) * (cf::G_N + 15) * sizeof(double)) != hipSuccess)
    return bail(CF_ERR_HIP);
  {
    int nTI, nTJ;
    c->ntiles = cf::sweep_tiles(g, &nTI, &nTJ);
    std::vector<int> ord;
    ord.reserve(c->ntiles);
    for (int d = 0; d <= nTI + nTJ - 2; d++)
      for (int tx = 0; tx < nTI; tx++) {
        int ty = d - tx;
        if (ty >= 0 && ty < nTJ) ord.push_back(tx + ty * nTI);
      }
    if (hipMalloc(&c->order, ord.size() * sizeof(int)) != hipSuccess) return bail(CF_ERR_HIP);
    if (hipMemcpy(c->order, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return bail(CF_ERR_HIP);
    if (hipMalloc(&c->sync, (2 * c->ntiles + 4) * sizeof(int)) != hipSuccess) return bail(CF_ERR_HIP);
    if (hipMemsetAsync(c->sync, 0, (2 * c->ntiles + 4) * sizeof(int), c->stream) != hipSuccess)
      return bail(CF_ERR_HIP);
  }
  long nb = std::max(cf::update_blocks(g), cf::sweep_update_blocks(g));
  if (hipMalloc(&c->partial, (size_t)nb * 5 * sizeof(double)) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipMalloc(&c->er_dev, 8 * sizeof(double)) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipHostMalloc(&c->er_host, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipMalloc(&c->stat, (size_t)7 * g.Ym * sizeof(double)) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipMalloc(&c->shear, (size_t)2 * p->X_out * p->Z_out * sizeof(double)) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipMemsetAsync(c->stat, 0, (size_t)7 * g.Ym * sizeof(double), c->stream) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipMalloc(&c->prof, (size_t)cf::PROF_ROWS * g.Ym * sizeof(double)) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipMalloc(&c->ypt, (size_t)g.Ym * sizeof(double)) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipMemsetAsync(c->ypt, 0, (size_t)g.Ym * sizeof(double), c->stream) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipMemsetAsync(c->prof, 0, (size_t)cf::PROF_ROWS * g.Ym * sizeof(double), c->stream) != hipSuccess)
    return bail(CF_ERR_HIP);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(CF_ERR_HIP);
  *out = c;
  return CF_OK;
}

extern "C" void cf_destroy(cf_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  for (int q = 0; q < 3; q++) {
    if (c->aux[q]) (void)hipStreamSynchronize(c->aux[q]);
    if (c->aux[q]) (void)hipStreamDestroy(c->aux[q]);
    if (c->join_ev[q]) (void)hipEventDestroy(c->join_ev[q]);
  }
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  double* ptrs[] = {c->Ui, c->Un, c->Uq, c->Fx, c->Fy, c->Fz, c->vF, c->MR,  c->NL,     c->NR,    c->Pr,
                    c->VS, c->Ep, c->Fw, c->ypt, c->Sm, c->partial, c->er_dev, c->prof, c->stat, c->Uqq, c->Hs, c->U3, c->Ualt, c->shear};
  for (double* q : ptrs)
    if (q) (void)hipFree(q);
  if (c->order) (void)hipFree(c->order);
  if (c->px) (void)hipFree(c->px);
  if (c->py) (void)hipFree(c->py);
  if (c->sync) (void)hipFree(c->sync);
  if (c->er_host) (void)hipHostFree(c->er_host);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* cf_last_error(const cf_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" long long cf_field_elems(const cf_ctx* c) { return c ? host_elems(c) : 0; }

extern "C" int cf_synchronize(cf_ctx* c) {
  if (!c) return CF_ERR_ARG;
  return sync(c);
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
// Jacobian.cpp:74-99 for one wall-normal position eta = jpos*deltaET
static void stretch(double high, double gamma, double deltaET, double jpos, double* etdy, double* J) {
  double half = 0.5 * high;
  double a = gamma * (1 - 2 * jpos * deltaET);
  double temp = half * (1 - 1. / tanh(gamma) * tanh(a));
  double temp1 = (1 - temp / half) * tanh(gamma);
  *etdy = tanh(gamma) / (2 * half * gamma * (1 - temp1 * temp1));
  double temp2 = 1 / cosh(a);
  double ydet = high * gamma * 1. / tanh(gamma) * temp2 * temp2;
  *J = 1. / (1 * ydet * 1);
}

extern "C" int cf_jacobian_all(cf_ctx* c, double gamma, double* J, double* const m[9], double* J_u,
                               double* const m_u[9], double* J_v, double* const m_v[9], double* J_w,
                               double* const m_w[9]) {
  if (!c || !(gamma > 0)) return fail(c, CF_ERR_ARG, "bad gamma");
  const Geo& g = c->g;
  int gend = g.gend, ny = g.ny, nz = g.nz, Ym = g.Ym;
  std::fill(c->hprof.begin(), c->hprof.end(), 0.0);
  auto P = [&](int set, int comp) { return &c->hprof[(size_t)(set * 10 + comp) * Ym]; };
  const double cart[9] = {1, 0, 0, 0, 0, 0, 0, 0, 1};
  // cell centres j in [2,ny]: (j-1.5); ghosts j=1, nyy copy the first interior
  for (int j = 2; j <= ny; j++) {
    double ey, jj;
    stretch(c->p.high, gamma, g.dET, j - 1.5, &ey, &jj);
    P(cf::SET_C, 0)[j] = jj;
    for (int q = 0; q < 9; q++) P(cf::SET_C, 1 + q)[j] = cart[q];
    P(cf::SET_C, 5)[j] = ey;
  }
  for (int q = 0; q < 10; q++) {
    P(cf::SET_C, q)[1] = P(cf::SET_C, q)[2];
    P(cf::SET_C, q)[ny + 1] = P(cf::SET_C, q)[ny];
  }
  // faces: x and z at (j-0.5) for j in [1,ny-1]; y at (j-1.0) for j in [1,ny]
  for (int j = 1; j <= ny; j++) {
    double ey, jj;
    if (j <= ny - 1) {
      stretch(c->p.high, gamma, g.dET, j - 0.5, &ey, &jj);
      for (int set : {cf::SET_U, cf::SET_W}) {
        P(set, 0)[j] = jj;
        for (int q = 0; q < 9; q++) P(set, 1 + q)[j] = cart[q];
        P(set, 5)[j] = ey;
      }
    }
    stretch(c->p.high, gamma, g.dET, j - 1.0, &ey, &jj);
    P(cf::SET_V, 0)[j] = jj;
    for (int q = 0; q < 9; q++) P(cf::SET_V, 1 + q)[j] = cart[q];
    P(cf::SET_V, 5)[j] = ey;
  }
  // wall faces take the cell-centre ghost values (Jacobian_v.cpp:124-141)
  for (int q = 0; q < 4; q++) {
    int comp = q == 0 ? 0 : 3 + q;  // J, etdx, etdy, etdz
    P(cf::SET_V, comp)[1] = P(cf::SET_C, comp)[1];
    P(cf::SET_V, comp)[ny] = P(cf::SET_C, comp)[ny + 1];
  }
  c->have_metrics = true;
  int r = upload_prof(c);
  if (r) return r;
  // optional reference-layout expansion, ghost copies as Jacobian.cpp:111-244
  auto fill = [&](double* A, int set, int comp, Region rg) {
    if (!A) return;
    const double* pr = P(set, comp);
    for (int i = rg.i0; i <= rg.i1; i++)
      for (int j = rg.j0; j <= rg.j1; j++)
        for (int k = rg.k0; k <= rg.k1; k++) A[((size_t)i * Ym + j) * g.Zm + k] = pr[j];
  };
  Region cellall{2, gend + 1, 1, ny + 1, 1, nz + 1};
  fill(J, cf::SET_C, 0, cellall);
  if (m)
    for (int q = 0; q < 9; q++) {
      if (!m[q]) continue;
      if (q < 7) {
        fill(m[q], cf::SET_C, 1 + q, cellall);
      } else {
        // ztdy/ztdz: interior, y and z ghosts take ztdx (0); x-ghost planes
        // are left untouched as in the reference
        fill(m[q], cf::SET_C, 1 + q, Region{3, gend, 2, ny, 2, nz});
        for (int i = 2; i <= gend + 1; i++)
          for (int j = 1; j <= ny + 1; j++)
            for (int k = 1; k <= nz + 1; k++) {
              bool ghost = j == 1 || j == ny + 1 || k == 1 || k == nz + 1;
              if (ghost) m[q][((size_t)i * Ym + j) * g.Zm + k] = P(cf::SET_C, 7)[j];
            }
      }
    }
  fill(J_u, cf::SET_U, 0, region_of(c, cf::SET_U, 0));
  fill(J_v, cf::SET_V, 0, region_of(c, cf::SET_V, 0));
  fill(J_w, cf::SET_W, 0, region_of(c, cf::SET_W, 0));
  for (int q = 0; q < 9; q++) {
    if (m_u && m_u[q]) fill(m_u[q], cf::SET_U, 1 + q, region_of(c, cf::SET_U, 0));
    if (m_v && m_v[q]) fill(m_v[q], cf::SET_V, 1 + q, region_of(c, cf::SET_V, 0));
    if (m_w && m_w[q]) fill(m_w[q], cf::SET_W, 1 + q, region_of(c, cf::SET_W, 0));
  }
  return sync(c);
}

extern "C" int cf_set_metrics(cf_ctx* c, const double* J, const double* const m[9], const double* J_u,
                              const double* const m_u[9], const double* J_v, const double* const m_v[9],
                              const double* J_w, const double* const m_w[9]) {
  if (!c) return CF_ERR_ARG;
  int r;
  if ((r = ingest_set(c, cf::SET_C, J, m))) return r;
  if ((r = ingest_set(c, cf::SET_U, J_u, m_u))) return r;
  if ((r = ingest_set(c, cf::SET_V, J_v, m_v))) return r;
  if ((r = ingest_set(c, cf::SET_W, J_w, m_w))) return r;
  c->have_metrics = true;
  if ((r = upload_prof(c))) return r;
  return sync(c);
}

extern "C" int cf_hyperplane(cf_ctx* c, int* counts, int capacity, int* nplanes) {
  if (!c || !counts || !nplanes) return fail(c, CF_ERR_ARG, "null argument");
  const Geo& g = c->g;
  int mmax = g.gend + g.ny + g.nz, n = mmax - 7 + 1;
  *nplanes = n;
  if (capacity < n) return fail(c, CF_ERR_ARG, "capacity %d < %d planes", capacity, n);
  // plane m holds the (i,j) pairs with 2 <= m-i-j <= nz: O(planes * X) count
  for (int m = 7; m <= mmax; m++) {
    long cnt = 0;
    for (int i = 3; i <= g.gend; i++) {
      int jlo = std::max(2, m - i - g.nz), jhi = std::min(g.ny, m - i - 2);
      if (jhi >= jlo) cnt += jhi - jlo + 1;
    }
    counts[m - 7] = (int)cnt;
  }
  return CF_OK;
}

// ---------------------------------------------------------------------------
// per-function entry points
// ---------------------------------------------------------------------------
// The per-function entry points upload the caller's arrays into the state
// fields U_, U, Uq and run the kernels on them.  When the context also holds
// a device-resident state (cf_load_state, cf_load_initial, a checkpoint),
// these fields are swapped for the Hs stand-ins for the length of the call,
// so that the resident iterate and levels come through unchanged.  The work
// fields (fluxes, vF, MR, NL, NR, sensors) are rewritten by every inner
// iteration and are shared.
struct HostArrayScope {
  cf_ctx* c;
  double *Ui = nullptr, *Un = nullptr, *Uq = nullptr;
  explicit HostArrayScope(cf_ctx* ctx) : c(ctx) {}
  int enter() {
    if (!c->have_state) return CF_OK;
    const size_t nf = (size_t)c->g.nf;
    if (!c->Hs) {
      if (hipMalloc(&c->Hs, nf * 15 * sizeof(double)) != hipSuccess) {
        c->Hs = nullptr;
        return fail(c, CF_ERR_HIP, "out of device memory for the host-array state fields");
      }
      HIPCHK(c, hipMemsetAsync(c->Hs, 0, nf * 15 * sizeof(double), c->stream));
    }
    Ui = c->Ui;
    Un = c->Un;
    Uq = c->Uq;
    c->Ui = c->Hs;
    c->Un = c->Hs + 5 * nf;
    c->Uq = c->Hs + 10 * nf;
    return CF_OK;
  }
  ~HostArrayScope() {
    if (!Ui) return;
    c->Ui = Ui;
    c->Un = Un;
    c->Uq = Uq;
  }
};

extern "C" int cf_boundary_conditions(cf_ctx* c, double* const U_[5], const double* J) {
  if (!c || !U_) return CF_ERR_ARG;
  HostArrayScope hs(c);
  int r;
  if ((r = hs.enter())) return r;
  if ((r = ingest(c, J, cf::SET_C, 0, "J")) || (r = upload_prof(c))) return r;
  if ((r = h2d_n(c, c->Ui, (const double* const*)U_, 5))) return r;
  cf::launch_halo_bc(c->stream, c->Ui, c->prof, c->g, c->p.x_periodic);
  if ((r = d2h_n(c, U_, c->Ui, 5))) return r;
  return sync(c);
}

static int flux_entry(cf_ctx* c, int dir, const double* const U_[5], const double* J, const double* Jf,
                      const double* const mf[9], double* const inF[5]) {
  if (!c || !U_ || !inF) return CF_ERR_ARG;
  HostArrayScope hs(c);
  int r;
  if ((r = hs.enter())) return r;
  const int set = dir == 0 ? cf::SET_U : dir == 1 ? cf::SET_V : cf::SET_W;
  if ((r = ingest(c, J, cf::SET_C, 0, "J")) || (r = ingest_set(c, set, Jf, mf)) || (r = upload_prof(c))) return r;
  if ((r = h2d_n(c, c->Ui, U_, 5))) return r;
  double* F = dir == 0 ? c->Fx : dir == 1 ? c->Fy : c->Fz;
  // outputs outside the face range come back as zeros
  HIPCHK(c, hipMemsetAsync(F, 0, (size_t)c->g.nf * 5 * 8, c->stream));
  cf::launch_flux(c->stream, dir, c->Ui, c->prof, F, c->g);
  if ((r = d2h_n(c, inF, F, 5))) return r;
  return sync(c);
}

extern "C" int cf_flux_x(cf_ctx* c, const double* const U_[5], const double* J, const double* J_u,
                         const double* const m_u[9], double* const inFx[5]) {
  return flux_entry(c, 0, U_, J, J_u, m_u, inFx);
}
extern "C" int cf_flux_y(cf_ctx* c, const double* const U_[5], const double* J, const double* J_v,
                         const double* const m_v[9], double* const inFy[5]) {
  return flux_entry(c, 1, U_, J, J_v, m_v, inFy);
}
extern "C" int cf_flux_z(cf_ctx* c, const double* const U_[5], const double* J, const double* J_w,
                         const double* const m_w[9], double* const inFz[5]) {
  return flux_entry(c, 2, U_, J, J_w, m_w, inFz);
}

static int flux_limiter_entry(cf_ctx* c, int dir, const double* const U_[5], const double* J, const double* Jf,
                              const double* const mf[9], const double* Ep, double* const inF[5]) {
  if (!c || !U_ || !inF || !Ep) return CF_ERR_ARG;
  HostArrayScope hs(c);
  int r;
  if ((r = hs.enter())) return r;
  const int set = dir == 0 ? cf::SET_U : dir == 1 ? cf::SET_V : cf::SET_W;
  if ((r = ingest(c, J, cf::SET_C, 0, "J")) || (r = ingest_set(c, set, Jf, mf)) || (r = upload_prof(c))) return r;
  if ((r = h2d_n(c, c->Ui, U_, 5))) return r;
  double* E = c->Ep + (long)dir * c->g.nf;
  if ((r = h2d_n(c, E, &Ep, 1))) return r;
  double* F = dir == 0 ? c->Fx : dir == 1 ? c->Fy : c->Fz;
  HIPCHK(c, hipMemsetAsync(F, 0, (size_t)c->g.nf * 5 * 8, c->stream));
  cf::launch_flux_limiter(c->stream, dir, c->Ui, E, c->prof, F, c->g);
  if ((r = d2h_n(c, inF, F, 5))) return r;
  return sync(c);
}

extern "C" int cf_flux_x_limiter(cf_ctx* c, const double* const U_[5], const double* J, const double* J_u,
                                 const double* const m_u[9], double* const inFx[5], const double* EpX) {
  return flux_limiter_entry(c, 0, U_, J, J_u, m_u, EpX, inFx);
}
extern "C" int cf_flux_y_limiter(cf_ctx* c, const double* const U_[5], const double* J, const double* J_v,
                                 const double* const m_v[9], double* const inFy[5], const double* EpY) {
  return flux_limiter_entry(c, 1, U_, J, J_v, m_v, EpY, inFy);
}
extern "C" int cf_flux_z_limiter(cf_ctx* c, const double* const U_[5], const double* J, const double* J_w,
                                 const double* const m_w[9], double* const inFz[5], const double* EpZ) {
  return flux_limiter_entry(c, 2, U_, J, J_w, m_w, EpZ, inFz);
}

// ---------------------------------------------------------------------------
// Grid (Grid.cpp:50-141) and Filter (Filter.cpp:12-1720)
// ---------------------------------------------------------------------------
extern "C" int cf_grid(cf_ctx* c, double gamma, double* X_point, double* Y_point) {
  if (!c) return CF_ERR_ARG;
  const Geo& g = c->g;
  const int gend = g.gend, ny = g.ny, nz = g.nz, Ym = g.Ym, Zm = g.Zm;
  // cell-centre heights of the tanh-stretched grid, a function of j only
  std::vector<double> yp(Ym, 0.0);
  for (int j = 2; j <= ny; j++)
    yp[j] = (0.5 * c->p.high) * (1 - 1. / tanh(gamma) * tanh(gamma * (1 - 2 * (j - 1.5) * c->p.deltaET)));
  const double ylo = -yp[2], yhi = yp[ny] + 2 * yp[2];  // wall ghosts
  auto A = [&](int i, int j, int k) -> double& { return Y_point[((size_t)i * Ym + j) * Zm + k]; };
  if (Y_point) {
    for (int i = 0; i <= gend + 3; i++)
      for (int j = 2; j <= ny; j++)
        for (int k = 2; k <= nz; k++) A(i, j, k) = yp[j];
    for (int i = 2; i <= gend + 1; i++)
      for (int k = 2; k <= nz; k++) {
        A(i, 1, k) = ylo;
        A(i, ny + 1, k) = yhi;
      }
    for (int i = 2; i <= gend + 1; i++)
      for (int j = 1; j <= ny + 1; j++) {
        A(i, j, 1) = A(i, j, 2);
        A(i, j, nz + 1) = A(i, j, nz);
      }
  }
  if (X_point)
    for (int i = 3; i <= gend; i++)
      for (int j = 2; j <= ny; j++)
        for (int k = 2; k <= nz; k++) X_point[((size_t)i * Ym + j) * Zm + k] = c->p.deltaXI * (i - 1.5);
  yp[1] = ylo;
  yp[ny + 1] = yhi;
  HIPCHK(c, hipMemcpyAsync(c->ypt, yp.data(), Ym * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return sync(c);
}

extern "C" int cf_filter(cf_ctx* c, double Roe_criterion, double E_high, double E_low, const double* const U_[5],
                         const double* Y_point, double* EpX, double* EpY, double* EpZ, const double* J) {
  if (!c || !U_ || !Y_point || !EpX || !EpY || !EpZ) return CF_ERR_ARG;
  const Geo& g = c->g;
  HostArrayScope hs(c);
  int r;
  if ((r = hs.enter())) return r;
  if ((r = ingest(c, J, cf::SET_C, 0, "J")) || (r = upload_prof(c))) return r;
  // Y_point as read by the filter: j in [1, ny+1] over the interior i, k
  std::vector<double> yp(g.Ym, 0.0);
  for (int j = 1; j <= g.ny + 1; j++) {
    const double v0 = Y_point[((size_t)3 * g.Ym + j) * g.Zm + 2];
    for (int i = 3; i <= g.gend; i++)
      for (int k = 2; k <= g.nz; k++)
        if (Y_point[((size_t)i * g.Ym + j) * g.Zm + k] != v0)
          return fail(c, CF_ERR_METRICS, "Y_point varies along i/k at (%d,%d,%d)", i, j, k);
    yp[j] = v0;
  }
  HIPCHK(c, hipMemcpyAsync(c->ypt, yp.data(), g.Ym * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (!c->Fw) {
    if (hipMalloc(&c->Fw, (size_t)g.nf * cf::FILTER_WORK_FIELDS * sizeof(double)) != hipSuccess)
      return fail(c, CF_ERR_HIP, "hipMalloc of the filter workspace failed");
  }
  if ((r = h2d_n(c, c->Ui, U_, 5))) return r;
  // the sensors keep the caller's values outside the face ranges
  double* const ep[3] = {EpX, EpY, EpZ};
  if ((r = h2d_n(c, c->Ep, (const double* const*)ep, 3))) return r;
  cf::launch_filter(c->stream, c->Ui, c->prof, c->ypt, c->Fw, Roe_criterion, E_high, E_low, c->Ep, c->Ep + g.nf,
                    c->Ep + 2 * g.nf, g);
  if ((r = d2h_n(c, ep, c->Ep, 3))) return r;
  return sync(c);
}

// the fused viscous kernel drops the coefficient products of the off-diagonal
// face metrics, which must then be exactly zero in all three face sets
static bool diagonal_face_metrics(const cf_ctx* c) {
  const double* P = c->hprof.data();
  const int off[6] = {2, 3, 4, 6, 7, 8};  // xidy, xidz, etdx, etdz, ztdx, ztdy
  for (int set = cf::SET_U; set <= cf::SET_W; set++)
    for (int q = 0; q < 6; q++)
      for (int j = 0; j < c->g.Ym; j++)
        if (P[(size_t)(set * 10 + off[q]) * c->g.Ym + j] != 0.0) return false;
  return true;
}

static int ensure_viscous_staging(cf_ctx* c) {
  if (c->Pr) return CF_OK;
  const size_t nf = (size_t)c->g.nf;
  if (hipMalloc(&c->Pr, nf * 4 * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->VS, nf * 12 * sizeof(double)) != hipSuccess)
    return fail(c, CF_ERR_HIP, "out of device memory for the viscous staging fields");
  HIPCHK(c, hipMemsetAsync(c->Pr, 0, nf * 4 * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(c->VS, 0, nf * 12 * sizeof(double), c->stream));
  return CF_OK;
}

// Viscous_terms.cpp:212-883 on the device state: the fused kernel for
// diagonal metrics (schedule 0), otherwise the three staged kernels
static bool viscous_fused(const cf_ctx* c) { return c->visc_schedule == 0 && diagonal_face_metrics(c); }

static int run_viscous(cf_ctx* c) {
  const bool fused = viscous_fused(c);
  if (!fused) {
    int r = ensure_viscous_staging(c);
    if (r) return r;
  }
  cf::launch_viscous(c->stream, c->Ui, c->prof, c->Pr, c->VS, c->vF, c->g, fused);
  return CF_OK;
}

extern "C" int cf_viscous_terms(cf_ctx* c, const double* const U_[5], double* const vF[4], const double* J,
                                const double* const m[9], const double* J_u, const double* const m_u[9],
                                const double* J_v, const double* const m_v[9], const double* J_w,
                                const double* const m_w[9]) {
  if (!c || !U_ || !vF) return CF_ERR_ARG;
  HostArrayScope hs(c);
  int r;
  if ((r = hs.enter())) return r;
  if ((r = ingest(c, J, cf::SET_C, 0, "J")) || (r = ingest_set(c, cf::SET_U, J_u, m_u)) ||
      (r = ingest_set(c, cf::SET_V, J_v, m_v)) || (r = ingest_set(c, cf::SET_W, J_w, m_w)) || (r = upload_prof(c)))
    return r;
  (void)m;
  if ((r = h2d_n(c, c->Ui, U_, 5))) return r;
  HIPCHK(c, hipMemsetAsync(c->vF, 0, (size_t)c->g.nf * 4 * 8, c->stream));
  if (int rv = run_viscous(c)) return rv;
  if ((r = d2h_n(c, vF, c->vF, 4))) return r;
  return sync(c);
}

// the tiled sweep takes the split Jacobians of diagonal metric rows; other
// metrics use the one-hyperplane-per-launch schedule
static bool diagonal_metrics(const cf_ctx* c) {
  const double* P = c->hprof.data();
  const int off[6] = {2, 3, 4, 6, 7, 8};  // xidy, xidz, etdx, etdz, ztdx, ztdy
  for (int q = 0; q < 6; q++)
    for (int j = 1; j <= c->g.ny + 1; j++)
      if (P[(size_t)(cf::SET_C * 10 + off[q]) * c->g.Ym + j] != 0.0) return false;
  return true;
}

// 0: tiled (pipelined or, for schedule 2, single-wave tiles), 1: per hyperplane
static int effective_schedule(const cf_ctx* c) { return (c->schedule != 1 && diagonal_metrics(c)) ? 0 : 1; }

// LU-SGS forward/backward sweeps (Dual_time_stepping.cpp:514-1277)
// fuse_rhs: the tiled sweeps' pre-pass forms MR itself (the RHS kernel is
// not launched and MR is not written in the natural layout)
// Sweep pairs to execute.  Dual_time_stepping.cpp:514-1277 runs n_sweeps
// forward/backward pairs with NL <- NR between them; in one domain the
// forward sweep reads only MR and values its own pass has already formed
// (upstream neighbours in hyperplane order, zero ghosts), and the backward
// sweep only that forward result and its own pass's values, so every later
// pair reproduces the first bit for bit.  Across MPI ranks the pairs differ
// (lagged neighbour planes, :521-559); this library holds one domain.
static int sweep_pairs(const cf_ctx* c) { return c->sweep_reuse ? std::min(c->p.n_sweeps, 1) : c->p.n_sweeps; }

// The forward tiled sweep needs its cross-tile NL slots at the sentinel, and
// every backward tiled sweep leaves them so: they are seeded here only when
// that is not known (a new context, a sweep that gave up), eagerly and never
// inside a graph capture (cf_inner_iteration seeds before it captures)
static void seed_hand_off(cf_ctx* c) {
  if (c->solver != 0 || effective_schedule(c) != 0) return;
  const int want = c->fp32 ? 2 : 1;
  if (c->hand_off == want) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  cf::launch_seed_hand_off(c->stream, c->Sm, c->g, c->fp32);
  c->hand_off = want;
}

static void sweeps(cf_ctx* c, double deltaT, const hipEvent_t* marks = nullptr, bool fuse_rhs = false) {
  const Geo& g = c->g;
  if (effective_schedule(c) == 1) {
    cf::launch_sweeps_planes(c->stream, c->Ui, c->MR, c->NL, c->NR, c->prof, g, deltaT, sweep_pairs(c), marks);
    return;
  }
  double DN = 1. / ((double)c->p.X_out * c->p.Y_out * c->p.Z_out);
  const cf::RhsIn rin{c->Un, c->Uq, c->Fx, c->Fy, c->Fz, c->vF, c->p.f_drive};
  seed_hand_off(c);
  cf::launch_sweeps(c->stream, c->Ui, c->MR, c->Sm, c->prof, g, deltaT, sweep_pairs(c), c->order, c->sync,
                    c->partial, c->er_dev, DN, c->spin_limit, marks, c->schedule != 2, fuse_rhs ? &rin : nullptr,
                    c->fp32);
}

// FP32 sweeps exist for the pipelined tiled schedule, which needs diagonal
// metric rows; any other combination is refused rather than run in FP64
static int check_precision(cf_ctx* c) {
  if (c->fp32 && (c->schedule != 0 || effective_schedule(c) != 0))
    return fail(c, CF_ERR_ARG, "FP32 sweeps need diagonal metrics and the pipelined sweep schedule (0)");
  return CF_OK;
}

// the RHS runs inside the tiled sweeps' pre-pass when the caller needs
// neither MR nor NL in the natural layout
static bool rhs_fused(const cf_ctx* c, bool want_nl) { return !want_nl && effective_schedule(c) == 0; }

// primitive update + norms after the sweeps; the tiled schedule applies them
// inside its last backward sweep
static void update(cf_ctx* c) {
  double DN = 1. / ((double)c->p.X_out * c->p.Y_out * c->p.Z_out);
  if (effective_schedule(c) == 1) cf::launch_update(c->stream, c->Ui, c->NR, c->prof, c->partial, c->er_dev, c->g, DN);
}

// er to the host and the sweep error flags checked
// (one copy into pinned memory: the norms and the flags, which the last
// reduction kernel placed after them)
static int finish_iteration(cf_ctx* c, double er[5]) {
  HIPCHK(c, hipMemcpyAsync(c->er_host, c->er_dev, 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  int r = sync(c);
  if (r) {
    c->hand_off = 0;
    return r;
  }
  for (int v = 0; v < 5; v++) er[v] = c->er_host[v];
  if (c->solver == 0 && effective_schedule(c) == 0 && (c->er_host[5] != 0. || c->er_host[6] != 0.)) {
    c->hand_off = 0;  // a sweep that gave up may leave slots without the sentinel
    return fail(c, CF_ERR_TIMEOUT, "LU-SGS sweep: a tile waited past its bound for an upstream tile");
  }
  for (int v = 0; v < 5; v++)
    if (!std::isfinite(er[v])) return fail(c, CF_ERR_NONFINITE, "residual norm er%d is not finite", v + 1);
  return CF_OK;
}

// the implicit part of one inner iteration on device-resident fields
// DPLUSGS on device-resident fields: the same RHS, four Jacobi sweeps, then
// the update and norms (DPLUSGS.cpp:189-1361)
static int dplusgs_device(cf_ctx* c, double deltaT, double er[5]) {
  const Geo& g = c->g;
  cf::launch_rhs(c->stream, c->Ui, c->Un, c->Uq, c->Fx, c->Fy, c->Fz, c->vF, c->prof, c->MR, c->NL, g, deltaT,
                 c->p.f_drive);
  cf::launch_dplusgs(c->stream, c->Ui, c->MR, c->NL, c->NR, c->prof, g, deltaT);
  double DN = 1. / ((double)c->p.X_out * c->p.Y_out * c->p.Z_out);
  cf::launch_update(c->stream, c->Ui, c->NR, c->prof, c->partial, c->er_dev, g, DN);
  HIPCHK(c, hipMemcpyAsync(c->er_host, c->er_dev, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  int r = sync(c);
  if (r) return r;
  for (int v = 0; v < 5; v++) er[v] = c->er_host[v];
  for (int v = 0; v < 5; v++)
    if (!std::isfinite(er[v])) return fail(c, CF_ERR_NONFINITE, "residual norm er%d is not finite", v + 1);
  return CF_OK;
}

static int dual_time_device(cf_ctx* c, double deltaT, double er[5], bool want_nl) {
  const Geo& g = c->g;
  if (c->solver == 1) return dplusgs_device(c, deltaT, er);
  if (int r = check_precision(c)) return r;
  const bool fuse = rhs_fused(c, want_nl);
  if (!fuse)
    cf::launch_rhs(c->stream, c->Ui, c->Un, c->Uq, c->Fx, c->Fy, c->Fz, c->vF, c->prof, c->MR,
                   want_nl ? c->NL : nullptr, g, deltaT, c->p.f_drive);
  sweeps(c, deltaT, nullptr, fuse);
  if (want_nl) {
    if (effective_schedule(c) == 0) cf::launch_sweep_outputs(c->stream, c->Sm, c->NL, c->NR, g, c->fp32);
    cf::launch_copy_interior(c->stream, c->NR, c->NL, g);
  }
  update(c);
  return finish_iteration(c, er);
}

static int implicit_entry(cf_ctx* c, int solver, double deltaT, double er[5], double* const U_[5],
                          const double* const U[5], const double* const Uq[5], const double* J,
                          const double* const m[9], const double* const inFx[5], const double* const inFy[5],
                          const double* const inFz[5], const double* const vF[4], double* const MR[5],
                          double* const NR[5], double* const NL[5]);

extern "C" int cf_dplusgs(cf_ctx* c, double deltaT, double er[5], double* const U_[5], const double* const U[5],
                          const double* const Uq[5], const double* J, const double* const m[9],
                          const double* const inFx[5], const double* const inFy[5], const double* const inFz[5],
                          const double* const vF[4], double* const MR[5], double* const NR[5], double* const NL[5]) {
  return implicit_entry(c, 1, deltaT, er, U_, U, Uq, J, m, inFx, inFy, inFz, vF, MR, NR, NL);
}

extern "C" int cf_dual_time_stepping(cf_ctx* c, double deltaT, double er[5], double* const U_[5],
                                     const double* const U[5], const double* const Uq[5], const double* J,
                                     const double* const m[9], const double* const inFx[5],
                                     const double* const inFy[5], const double* const inFz[5],
                                     const double* const vF[4], double* const MR[5], double* const NR[5],
                                     double* const NL[5]) {
  return implicit_entry(c, 0, deltaT, er, U_, U, Uq, J, m, inFx, inFy, inFz, vF, MR, NR, NL);
}

static int implicit_entry(cf_ctx* c, int solver, double deltaT, double er[5], double* const U_[5],
                                     const double* const U[5], const double* const Uq[5], const double* J,
                                     const double* const m[9], const double* const inFx[5],
                                     const double* const inFy[5], const double* const inFz[5],
                                     const double* const vF[4], double* const MR[5], double* const NR[5],
                                     double* const NL[5]) {
  if (!c || !er || !U_ || !U || !Uq || !inFx || !inFy || !inFz || !vF) return fail(c, CF_ERR_ARG, "null argument");
  if (!(deltaT > 0)) return fail(c, CF_ERR_ARG, "deltaT must be positive");
  HostArrayScope hs(c);
  int r;
  if ((r = hs.enter())) return r;
  if ((r = ingest_set(c, cf::SET_C, J, m)) || (r = upload_prof(c))) return r;
  if ((r = h2d_n(c, c->Ui, (const double* const*)U_, 5)) || (r = h2d_n(c, c->Un, U, 5)) ||
      (r = h2d_n(c, c->Uq, Uq, 5)) || (r = h2d_n(c, c->Fx, inFx, 5)) || (r = h2d_n(c, c->Fy, inFy, 5)) ||
      (r = h2d_n(c, c->Fz, inFz, 5)) || (r = h2d_n(c, c->vF, vF, 4)))
    return r;
  // ghost layers of the implicit work arrays carry no coupling (:195-316)
  HIPCHK(c, hipMemsetAsync(c->NL, 0, (size_t)c->g.nf * 5 * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(c->NR, 0, (size_t)c->g.nf * 5 * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(c->MR, 0, (size_t)c->g.nf * 5 * 8, c->stream));
  const int keep = c->solver;
  c->solver = solver;
  r = dual_time_device(c, deltaT, er, true);
  c->solver = keep;
  if (r) {
    d2h_n(c, U_, c->Ui, 5);
    sync(c);
    return r;
  }
  if ((r = d2h_n(c, U_, c->Ui, 5))) return r;
  if (MR && (r = d2h_n(c, MR, c->MR, 5))) return r;
  if (NR && (r = d2h_n(c, NR, c->NR, 5))) return r;
  if (NL && (r = d2h_n(c, NL, c->NL, 5))) return r;
  return sync(c);
}

extern "C" int cf_rk3nodt(cf_ctx* c, int RK, double deltaT, double er[5], double* const U_[5],
                          const double* const U[5], const double* const Uq[5], const double* J,
                          const double* const m[9], const double* const inFx[5], const double* const inFy[5],
                          const double* const inFz[5], const double* const vF[4], double* const MR[5]) {
  if (!c || !U_ || !U || !inFx || !inFy || !inFz || !vF) return fail(c, CF_ERR_ARG, "null argument");
  if (RK < 1 || RK > 3) return fail(c, CF_ERR_ARG, "RK stage must be 1, 2 or 3");
  (void)er;  // RK3NODT.cpp leaves er1..er5 untouched
  (void)Uq;
  HostArrayScope hs(c);
  int r;
  if ((r = hs.enter())) return r;
  if ((r = ingest_set(c, cf::SET_C, J, m)) || (r = upload_prof(c))) return r;
  if ((r = h2d_n(c, c->Ui, (const double* const*)U_, 5)) || (r = h2d_n(c, c->Un, U, 5)) ||
      (r = h2d_n(c, c->Fx, inFx, 5)) || (r = h2d_n(c, c->Fy, inFy, 5)) || (r = h2d_n(c, c->Fz, inFz, 5)) ||
      (r = h2d_n(c, c->vF, vF, 4)))
    return r;
  cf::launch_rk3(c->stream, RK, c->Ui, c->Un, c->Fx, c->Fy, c->Fz, c->vF, c->prof, c->MR, c->g, deltaT,
                 c->p.f_drive);
  HIPCHK(c, hipGetLastError());
  if ((r = d2h_n(c, U_, c->Ui, 5))) return r;
  if (MR && (r = d2h_n(c, MR, c->MR, 5))) return r;
  return sync(c);
}

// ---------------------------------------------------------------------------
// device-resident path
// ---------------------------------------------------------------------------
// interior Q [X_out][Y_out][Z_out] -> U = Q/J in the device layout
static int upload_Q(cf_ctx* c, double* dst, const double* const Q[5]) {
  const Geo& g = c->g;
  const double* Jp = &c->hprof[0];
  std::vector<double> buf((size_t)g.nf, 0.0);
  size_t X = c->p.X_out, Y = c->p.Y_out, Z = c->p.Z_out;
  for (int v = 0; v < 5; v++) {
    if (!Q[v]) return fail(c, CF_ERR_ARG, "null state array %d", v);
    std::fill(buf.begin(), buf.end(), 0.0);
    for (size_t i = 0; i < X; i++)
      for (size_t j = 0; j < Y; j++)
        for (size_t k = 0; k < Z; k++)
          buf[(i + 3) * g.plane + (j + 2) * g.Zp + (k + 2)] = Q[v][(i * Y + j) * Z + k] / Jp[j + 2];
    HIPCHK(c, hipMemcpy(dst + v * g.nf, buf.data(), (size_t)g.nf * 8, hipMemcpyHostToDevice));
  }
  return CF_OK;
}

static int download_Q(cf_ctx* c, double* const Q[5], const double* src) {
  const Geo& g = c->g;
  const double* Jp = &c->hprof[0];
  std::vector<double> buf((size_t)g.nf);
  size_t X = c->p.X_out, Y = c->p.Y_out, Z = c->p.Z_out;
  for (int v = 0; v < 5; v++) {
    if (!Q[v]) continue;
    HIPCHK(c, hipMemcpy(buf.data(), src + v * g.nf, (size_t)g.nf * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < X; i++)
      for (size_t j = 0; j < Y; j++)
        for (size_t k = 0; k < Z; k++)
          Q[v][(i * Y + j) * Z + k] = buf[(i + 3) * g.plane + (j + 2) * g.Zp + (k + 2)] * Jp[j + 2];
  }
  return CF_OK;
}

extern "C" int cf_load_state(cf_ctx* c, const double* const Qn[5], const double* const Qnm1[5]) {
  if (!c || !Qn || !Qnm1) return CF_ERR_ARG;
  if (!c->have_metrics) return fail(c, CF_ERR_STATE, "metrics not set (cf_jacobian_all / cf_set_metrics)");
  int r;
  if ((r = sync(c))) return r;
  if ((r = upload_Q(c, c->Un, Qn)) || (r = upload_Q(c, c->Uq, Qnm1)) || (r = upload_Q(c, c->Ui, Qn))) return r;
  for (int v = 0; v < 5; v++) c->E[v] = 0;
  c->have_state = true;
  c->uqq_valid = false;
  c->u3_valid = false;
  return CF_OK;
}

extern "C" int cf_set_iterate(cf_ctx* c, const double* const Q[5]) {
  if (!c || !Q) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  int r;
  if ((r = sync(c))) return r;
  return upload_Q(c, c->Ui, Q);
}

extern "C" int cf_load_initial(cf_ctx* c, const char* path) {
  if (!c || !path) return CF_ERR_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(c, CF_ERR_IO, "cannot open %s", path);
  size_t n = (size_t)c->p.X_out * c->p.Y_out * c->p.Z_out;
  std::vector<double> data(10 * n);
  size_t got = fread(data.data(), sizeof(double), 10 * n, f);
  fclose(f);
  if (got != 10 * n) return fail(c, CF_ERR_IO, "%s: expected %zu doubles, read %zu", path, 10 * n, got);
  const double* Qn[5];
  const double* Qm[5];
  for (int v = 0; v < 5; v++) {
    Qn[v] = data.data() + v * n;
    Qm[v] = data.data() + (5 + v) * n;
  }
  return cf_load_state(c, Qn, Qm);
}

extern "C" int cf_read_state(cf_ctx* c, double* const Qn[5], double* const Qnm1[5], double* const Qiter[5]) {
  if (!c) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  int r;
  if ((r = sync(c))) return r;
  if (Qn && (r = download_Q(c, Qn, c->Un))) return r;
  if (Qnm1 && (r = download_Q(c, Qnm1, c->Uq))) return r;
  if (Qiter && (r = download_Q(c, Qiter, c->Ui))) return r;
  return CF_OK;
}

// Everything one inner iteration of the device state launches: kernels
// only, no copies, no memsets and no synchronisation, so that the sequence
// can be captured into a graph.  The checks and allocations it could need
// happen first (prepare_iteration).
static int prepare_iteration(cf_ctx* c) {
  if (!viscous_fused(c))
    if (int r = ensure_viscous_staging(c)) return r;
  if (c->fork && !c->fork_ev) {
    for (int q = 0; q < 3; q++) {
      HIPCHK(c, hipStreamCreateWithFlags(&c->aux[q], hipStreamNonBlocking));
      HIPCHK(c, hipEventCreateWithFlags(&c->join_ev[q], hipEventDisableTiming));
    }
    HIPCHK(c, hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
  }
  if (c->solver == 0)
    if (int r = check_precision(c)) return r;
  seed_hand_off(c);
  return CF_OK;
}

static void enqueue_iteration(cf_ctx* c) {
  const Geo& g = c->g;
  cf::launch_halo_bc(c->stream, c->Ui, c->prof, g, c->p.x_periodic);
  if (c->fork && c->fork_ev) {
    // Flux_X on the context's stream, Flux_Y, Flux_Z and Viscous_terms on
    // the helper streams, all after the ghost cells, all before the RHS
    (void)hipEventRecord(c->fork_ev, c->stream);
    for (int q = 0; q < 3; q++) (void)hipStreamWaitEvent(c->aux[q], c->fork_ev, 0);
    cf::launch_flux(c->stream, 0, c->Ui, c->prof, c->Fx, g);
    cf::launch_flux(c->aux[0], 1, c->Ui, c->prof, c->Fy, g);
    cf::launch_flux(c->aux[1], 2, c->Ui, c->prof, c->Fz, g);
    cf::launch_viscous(c->aux[2], c->Ui, c->prof, c->Pr, c->VS, c->vF, g, viscous_fused(c));
    for (int q = 0; q < 3; q++) {
      (void)hipEventRecord(c->join_ev[q], c->aux[q]);
      (void)hipStreamWaitEvent(c->stream, c->join_ev[q], 0);
    }
  } else {
    cf::launch_flux(c->stream, 0, c->Ui, c->prof, c->Fx, g);
    cf::launch_flux(c->stream, 1, c->Ui, c->prof, c->Fy, g);
    cf::launch_flux(c->stream, 2, c->Ui, c->prof, c->Fz, g);
    cf::launch_viscous(c->stream, c->Ui, c->prof, c->Pr, c->VS, c->vF, g, viscous_fused(c));
  }
  if (c->solver == 1) {
    // DPLUSGS.cpp:189-1361: the RHS, four Jacobi sweeps, the update and norms
    cf::launch_rhs(c->stream, c->Ui, c->Un, c->Uq, c->Fx, c->Fy, c->Fz, c->vF, c->prof, c->MR, c->NL, g,
                   c->p.deltaT, c->p.f_drive);
    cf::launch_dplusgs(c->stream, c->Ui, c->MR, c->NL, c->NR, c->prof, g, c->p.deltaT);
    const double DN = 1. / ((double)c->p.X_out * c->p.Y_out * c->p.Z_out);
    cf::launch_update(c->stream, c->Ui, c->NR, c->prof, c->partial, c->er_dev, g, DN);
    return;
  }
  const bool fuse = rhs_fused(c, false);
  if (!fuse)
    cf::launch_rhs(c->stream, c->Ui, c->Un, c->Uq, c->Fx, c->Fy, c->Fz, c->vF, c->prof, c->MR, nullptr, g,
                   c->p.deltaT, c->p.f_drive);
  sweeps(c, c->p.deltaT, nullptr, fuse);
  update(c);
}

// the iteration's kernels as one graph launch, captured on first use and
// again whenever the fields, schedules or parameters it was captured with
// have changed
static int launch_graph(cf_ctx* c) {
  const cf_ctx::GraphKey key{c->Ui,     c->Un,          c->Uq,          c->schedule, c->visc_schedule,
                             c->solver, c->p.x_periodic, sweep_pairs(c), c->fp32,     c->fork,
                             c->p.deltaT, c->p.f_drive,  c->prof_gen, c->spin_limit};
  if (!c->gexec || !(key == c->gkey)) {
    if (c->gexec) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_iteration(c);
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamEndCapture(c->stream, &graph));
    hipError_t e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      c->gexec = nullptr;
      return fail(c, CF_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    c->gkey = key;
    c->graph_builds++;
  }
  HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
  return CF_OK;
}

extern "C" int cf_inner_iteration(cf_ctx* c, double er[5]) {
  if (!c || !er) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  if (int r = prepare_iteration(c)) return r;
  if (c->use_graph) {
    if (int r = launch_graph(c)) return r;
  } else {
    enqueue_iteration(c);
  }
  return finish_iteration(c, er);
}

extern "C" int cf_set_graph(cf_ctx* c, int enabled) {
  if (!c) return CF_ERR_ARG;
  c->use_graph = enabled != 0;
  if (!c->use_graph && c->gexec) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  return CF_OK;
}

extern "C" int cf_graph_builds(cf_ctx* c, long long* builds) {
  if (!c || !builds) return CF_ERR_ARG;
  *builds = c->graph_builds;
  return CF_OK;
}

extern "C" int cf_set_concurrent_stencils(cf_ctx* c, int enabled) {
  if (!c) return CF_ERR_ARG;
  for (int q = 0; q < 3; q++)
    if (c->aux[q]) HIPCHK(c, hipStreamSynchronize(c->aux[q]));
  c->fork = enabled != 0;
  return CF_OK;
}

extern "C" int cf_set_sweep_reuse(cf_ctx* c, int enabled) {
  if (!c) return CF_ERR_ARG;
  c->sweep_reuse = enabled != 0;
  return CF_OK;
}

extern "C" int cf_set_spin_limit(cf_ctx* c, long long polls) {
  if (!c || polls < 1) return CF_ERR_ARG;
  c->spin_limit = (long)polls;
  return CF_OK;
}

extern "C" int cf_advance_levels(cf_ctx* c) {
  if (!c) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  // U^{n-1} <- U^n, U^n <- U_ on the interior (main.cpp:444-464); U^{n-2}
  // <- U^{n-1} first while the quadratic initial guess keeps it
  if (c->U3) {
    cf::launch_copy_interior(c->stream, c->Uqq, c->U3, c->g);
    c->u3_valid = c->uqq_valid;
  }
  if (c->Uqq) {
    cf::launch_copy_interior(c->stream, c->Uq, c->Uqq, c->g);
    c->uqq_valid = true;
  }
  cf::launch_copy_interior(c->stream, c->Un, c->Uq, c->g);
  cf::launch_copy_interior(c->stream, c->Ui, c->Un, c->g);
  return sync(c);
}

// HIP events created together and destroyed on every exit path
struct EventSet {
  std::vector<hipEvent_t> ev;
  hipError_t create(size_t n) {
    ev.reserve(n);
    for (size_t q = 0; q < n; q++) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      ev.push_back(e);
    }
    return hipSuccess;
  }
  ~EventSet() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

extern "C" int cf_profile_inner_iteration(cf_ctx* c, double er[5], double stage_ms[CF_PROFILE_STAGES],
                                          int launches[CF_PROFILE_STAGES]) {
  if (!c || !er || !stage_ms || !launches) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  if (int r = check_precision(c)) return r;
  seed_hand_off(c);  // outside the timed stages
  const Geo& g = c->g;
  const int ns = sweep_pairs(c);
  // ev[0..6]: stage ends up to the RHS; then the sweep marks; then the end
  const int nm = 1 + 3 * ns;
  EventSet es;
  HIPCHK(c, es.create(7 + nm + 1));
  std::vector<hipEvent_t>& ev = es.ev;
  hipEvent_t* mk = ev.data() + 7;
  HIPCHK(c, hipEventRecord(ev[0], c->stream));
  cf::launch_halo_bc(c->stream, c->Ui, c->prof, g, c->p.x_periodic);
  HIPCHK(c, hipEventRecord(ev[1], c->stream));
  cf::launch_flux(c->stream, 0, c->Ui, c->prof, c->Fx, g);
  HIPCHK(c, hipEventRecord(ev[2], c->stream));
  cf::launch_flux(c->stream, 1, c->Ui, c->prof, c->Fy, g);
  HIPCHK(c, hipEventRecord(ev[3], c->stream));
  cf::launch_flux(c->stream, 2, c->Ui, c->prof, c->Fz, g);
  HIPCHK(c, hipEventRecord(ev[4], c->stream));
  if (int rv = run_viscous(c)) return rv;
  HIPCHK(c, hipEventRecord(ev[5], c->stream));
  const bool fuse = rhs_fused(c, false);
  if (!fuse)
    cf::launch_rhs(c->stream, c->Ui, c->Un, c->Uq, c->Fx, c->Fy, c->Fz, c->vF, c->prof, c->MR, nullptr, g,
                   c->p.deltaT, c->p.f_drive);
  HIPCHK(c, hipEventRecord(ev[6], c->stream));
  sweeps(c, c->p.deltaT, mk, fuse);
  update(c);
  HIPCHK(c, hipEventRecord(ev[7 + nm], c->stream));
  int r = finish_iteration(c, er);
  if (r) return r;
  auto span = [&](hipEvent_t a, hipEvent_t b, double* acc) -> int {
    float ms = 0;
    hipError_t e = hipEventElapsedTime(&ms, a, b);
    if (e != hipSuccess) return fail(c, CF_ERR_HIP, "hipEventElapsedTime: %s", hipGetErrorString(e));
    *acc += ms;
    return CF_OK;
  };
  for (int q = 0; q < CF_PROFILE_STAGES; q++) stage_ms[q] = 0;
  for (int q = 0; q < 6; q++)
    if ((r = span(ev[q], ev[q + 1], &stage_ms[q]))) return r;
  if ((r = span(ev[6], mk[0], &stage_ms[6]))) return r;
  for (int s = 0; s < ns; s++) {
    if ((r = span(mk[3 * s], mk[1 + 3 * s], &stage_ms[7]))) return r;
    if ((r = span(mk[1 + 3 * s], mk[2 + 3 * s], &stage_ms[8]))) return r;
    if ((r = span(mk[2 + 3 * s], mk[3 + 3 * s], &stage_ms[9]))) return r;
  }
  if ((r = span(mk[nm - 1], ev[7 + nm], &stage_ms[10]))) return r;
  const bool tiled = effective_schedule(c) == 0;
  const int nplanes = g.gend + g.ny + g.nz - 6;
  const int nl[CF_PROFILE_STAGES] = {3, 1, 1, 1, viscous_fused(c) ? 1 : 3, fuse ? 0 : 1, tiled ? 1 : 0, tiled ? ns : 0, tiled ? ns : nplanes * ns,
                                     tiled ? ns : nplanes * ns, 2};
  for (int q = 0; q < CF_PROFILE_STAGES; q++) launches[q] = nl[q];
  return CF_OK;
}

extern "C" int cf_profile_variants(cf_ctx* c, double Roe_criterion, double ms[CF_VARIANT_STAGES]) {
  if (!c || !ms) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  const Geo& g = c->g;
  if (!c->Fw) {
    if (hipMalloc(&c->Fw, (size_t)g.nf * cf::FILTER_WORK_FIELDS * sizeof(double)) != hipSuccess)
      return fail(c, CF_ERR_HIP, "hipMalloc of the filter workspace failed");
  }
  EventSet es;
  HIPCHK(c, es.create(CF_VARIANT_STAGES + 1));
  hipEvent_t* ev = es.ev.data();
  // on the resident iterate with its ghosts set, as in an inner iteration
  cf::launch_halo_bc(c->stream, c->Ui, c->prof, g, c->p.x_periodic);
  HIPCHK(c, hipEventRecord(ev[0], c->stream));
  cf::launch_filter(c->stream, c->Ui, c->prof, c->ypt, c->Fw, Roe_criterion, 1.0, 0.0, c->Ep, c->Ep + g.nf,
                    c->Ep + 2 * g.nf, g);
  HIPCHK(c, hipEventRecord(ev[1], c->stream));
  double* F[3] = {c->Fx, c->Fy, c->Fz};
  for (int d = 0; d < 3; d++) {
    cf::launch_flux_limiter(c->stream, d, c->Ui, c->Ep + (long)d * g.nf, c->prof, F[d], g);
    HIPCHK(c, hipEventRecord(ev[2 + d], c->stream));
  }
  int r = sync(c);
  if (r) return r;
  for (int q = 0; q < CF_VARIANT_STAGES; q++) {
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, ev[q], ev[q + 1]));
    ms[q] = t;
  }
  return CF_OK;
}

extern "C" int cf_set_implicit_solver(cf_ctx* c, int solver) {
  if (!c || solver < 0 || solver > 1) return fail(c, CF_ERR_ARG, "bad solver");
  c->solver = solver;
  return CF_OK;
}

extern "C" int cf_rk3_step(cf_ctx* c, int viscous) {
  if (!c) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  const Geo& g = c->g;
  for (int RK = 1; RK <= 3; RK++) {
    cf::launch_halo_bc(c->stream, c->Ui, c->prof, g, c->p.x_periodic);
    cf::launch_flux(c->stream, 0, c->Ui, c->prof, c->Fx, g);
    cf::launch_flux(c->stream, 1, c->Ui, c->prof, c->Fy, g);
    cf::launch_flux(c->stream, 2, c->Ui, c->prof, c->Fz, g);
    if (viscous)
      if (int rv = run_viscous(c)) return rv;
    cf::launch_rk3(c->stream, RK, c->Ui, c->Un, c->Fx, c->Fy, c->Fz, viscous ? c->vF : nullptr, c->prof, nullptr, g,
                   c->p.deltaT, c->p.f_drive);
  }
  HIPCHK(c, hipGetLastError());
  return sync(c);
}

// ---------------------------------------------------------------------------
// exact checkpoint: the device state as stored (U = Q/J), the running maxima
// of er and the statistic sums, so that a restarted run continues bit for bit
// ---------------------------------------------------------------------------
static const char CKPT_MAGIC[8] = {'C', 'F', 'C', 'K', 'P', 'T', '0', '2'};
static const char CKPT_MAGIC_V1[8] = {'C', 'F', 'C', 'K', 'P', 'T', '0', '1'};  // no Shear_stress.dat sums

static int ckpt_field(cf_ctx* c, FILE* f, double* dev, bool write) {
  const Geo& g = c->g;
  std::vector<double> buf((size_t)g.nf, 0.0), in((size_t)c->p.Y_out * c->p.Z_out);
  const size_t X = c->p.X_out, Y = c->p.Y_out, Z = c->p.Z_out;
  // loading overwrites the interior only; the ghosts are rebuilt by every iteration
  HIPCHK(c, hipMemcpy(buf.data(), dev, buf.size() * 8, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < X; i++) {
    if (!write && fread(in.data(), 8, in.size(), f) != in.size()) return fail(c, CF_ERR_IO, "short checkpoint");
    for (size_t j = 0; j < Y; j++)
      for (size_t k = 0; k < Z; k++) {
        double& v = buf[(i + 3) * g.plane + (j + 2) * g.Zp + (k + 2)];
        if (write) in[j * Z + k] = v;
        else v = in[j * Z + k];
      }
    if (write && fwrite(in.data(), 8, in.size(), f) != in.size()) return fail(c, CF_ERR_IO, "checkpoint write failed");
  }
  if (!write) HIPCHK(c, hipMemcpy(dev, buf.data(), buf.size() * 8, hipMemcpyHostToDevice));
  return CF_OK;
}

extern "C" int cf_save_checkpoint(cf_ctx* c, const char* path) {
  if (!c || !path) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  int r;
  if ((r = sync(c))) return r;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(c, CF_ERR_IO, "cannot open %s", path);
  // dims[3]: fields after the 15 levels: 5 when U^{n-2} is kept and valid,
  // 10 when U^{n-3} is as well (the cubic initial iterate restarts exactly)
  const int nextra = c->Uqq && c->uqq_valid ? (c->U3 && c->u3_valid ? 10 : 5) : 0;
  const int32_t dims[4] = {c->p.X_out, c->p.Y_out, c->p.Z_out, nextra};
  std::vector<double> st((size_t)7 * c->g.Ym);
  HIPCHK(c, hipMemcpy(st.data(), c->stat, st.size() * 8, hipMemcpyDeviceToHost));
  const double sxy[4] = {c->sxy_sum[0], c->sxy_sum[1], c->sxy_prev[0], c->sxy_prev[1]};
  bool ok = fwrite(CKPT_MAGIC, 1, 8, f) == 8 && fwrite(dims, 4, 4, f) == 4 && fwrite(c->E, 8, 5, f) == 5 &&
            fwrite(st.data(), 8, st.size(), f) == st.size() && fwrite(sxy, 8, 4, f) == 4;
  double* fields[25];
  for (int v = 0; v < 5; v++) {
    fields[v] = c->Ui + v * c->g.nf;
    fields[5 + v] = c->Un + v * c->g.nf;
    fields[10 + v] = c->Uq + v * c->g.nf;
    fields[15 + v] = c->Uqq ? c->Uqq + v * c->g.nf : nullptr;
    fields[20 + v] = c->U3 ? c->U3 + v * c->g.nf : nullptr;
  }
  for (int q = 0; ok && q < 15 + nextra; q++)
    if ((r = ckpt_field(c, f, fields[q], true))) {
      fclose(f);
      return r;
    }
  if (fclose(f) != 0 || !ok) return fail(c, CF_ERR_IO, "checkpoint write failed: %s", path);
  return CF_OK;
}

extern "C" int cf_load_checkpoint(cf_ctx* c, const char* path) {
  if (!c || !path) return CF_ERR_ARG;
  // the context's stream does not synchronise with the blocking copies below
  if (int r0 = sync(c)) return r0;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(c, CF_ERR_IO, "cannot open %s", path);
  char magic[8];
  int32_t dims[4];
  double E[5];
  std::vector<double> st((size_t)7 * c->g.Ym);
  bool ok = fread(magic, 1, 8, f) == 8 && (memcmp(magic, CKPT_MAGIC, 8) == 0 || memcmp(magic, CKPT_MAGIC_V1, 8) == 0) &&
            fread(dims, 4, 4, f) == 4;
  const bool v1 = memcmp(magic, CKPT_MAGIC_V1, 8) == 0;
  if (!ok || dims[0] != c->p.X_out || dims[1] != c->p.Y_out || dims[2] != c->p.Z_out ||
      (dims[3] != 0 && dims[3] != 5 && dims[3] != 10)) {
    fclose(f);
    return fail(c, CF_ERR_IO, "%s is not a checkpoint of this grid", path);
  }
  double sxy[4] = {0, 0, 0, 0};
  if (fread(E, 8, 5, f) != 5 || fread(st.data(), 8, st.size(), f) != st.size() || (!v1 && fread(sxy, 8, 4, f) != 4)) {
    fclose(f);
    return fail(c, CF_ERR_IO, "short checkpoint");
  }
  int r;
  const int nextra = dims[3];
  // the file keeps U^{n-2} (and U^{n-3}): hold them even if the initial
  // guess that uses them is not selected yet
  auto hold = [&](double** level, const char* name) -> int {
    if (*level) return CF_OK;
    const size_t bytes = (size_t)c->g.nf * 5 * sizeof(double);
    if (hipMalloc(level, bytes) != hipSuccess) {
      *level = nullptr;
      return fail(c, CF_ERR_HIP, "hipMalloc of the %s level failed", name);
    }
    if (hipMemset(*level, 0, bytes) != hipSuccess) return fail(c, CF_ERR_HIP, "hipMemset of the %s level failed", name);
    return CF_OK;
  };
  if ((nextra >= 5 && (r = hold(&c->Uqq, "U^{n-2}"))) || (nextra >= 10 && (r = hold(&c->U3, "U^{n-3}")))) {
    fclose(f);
    return r;
  }
  double* fields[25];
  for (int v = 0; v < 5; v++) {
    fields[v] = c->Ui + v * c->g.nf;
    fields[5 + v] = c->Un + v * c->g.nf;
    fields[10 + v] = c->Uq + v * c->g.nf;
    fields[15 + v] = c->Uqq ? c->Uqq + v * c->g.nf : nullptr;
    fields[20 + v] = c->U3 ? c->U3 + v * c->g.nf : nullptr;
  }
  for (int q = 0; q < 15 + nextra; q++)
    if ((r = ckpt_field(c, f, fields[q], false))) {
      fclose(f);
      return r;
    }
  fclose(f);
  HIPCHK(c, hipMemcpy(c->stat, st.data(), st.size() * 8, hipMemcpyHostToDevice));
  for (int v = 0; v < 5; v++) c->E[v] = E[v];
  c->sxy_sum[0] = sxy[0];
  c->sxy_sum[1] = sxy[1];
  c->sxy_prev[0] = sxy[2];
  c->sxy_prev[1] = sxy[3];
  c->have_state = true;
  c->uqq_valid = nextra >= 5;
  c->u3_valid = nextra == 10;
  return CF_OK;
}

extern "C" int cf_driven_force(cf_ctx* c, double Q0, const double* U2_, const double* J, double* Qall,
                               double* ER_Q) {
  if (!c || !Qall || !ER_Q) return fail(c, CF_ERR_ARG, "null argument");
  const Geo& g = c->g;
  HostArrayScope hs(c);
  int r;
  if (U2_ && (r = hs.enter())) return r;
  if (J && ((r = ingest(c, J, cf::SET_C, 0, "J")) || (r = upload_prof(c)))) return r;
  const double* U2 = c->Ui + g.nf;
  if (U2_) {
    // reference layout [X_np][Y_m][Z_m] -> the device rows of Zp
    if ((r = h2d(c, c->Ui + g.nf, U2_))) return r;
  } else if (!c->have_state) {
    return fail(c, CF_ERR_STATE, "no state loaded");
  }
  // Driven_force.cpp:45-47, in the reference's order
  double Jsum = 0;
  for (int j = 2; j <= g.ny; j++) Jsum = Jsum + 1. / c->hprof[j];
  const int X = g.gend - 2;
  if ((size_t)X > (size_t)cf::update_blocks(g) * 5) return fail(c, CF_ERR_ARG, "grid too large for the partials");
  cf::launch_bulk_flux(c->stream, U2, c->prof, Jsum, c->partial, g);
  HIPCHK(c, hipGetLastError());
  std::vector<double> part((size_t)X);
  HIPCHK(c, hipMemcpyAsync(part.data(), c->partial, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
  if ((r = sync(c))) return r;
  double Q = 0;
  for (int i = 0; i < X; i++) Q = Q + part[i];
  Q = Q / c->p.X_out / c->p.Z_out;
  *Qall = Q;
  *ER_Q = fabs(Q0 - Q) / Q0;
  return CF_OK;
}

// Driven_force.cpp:38-73 like cf_driven_force, with no device allocation:
// the per-plane partials and their in-order sum live in `scratch`
// (scratch_of(device)), the plane sum is formed on the device in the order
// the host forms it in cf_driven_force, so the results are the same bits.
extern "C" int cf_driven_force_noalloc(cf_ctx* c, double Q0, const double* U2_dev, double* Qall, double* ER_Q) {
  if (!c || !Qall || !ER_Q) return fail(c, CF_ERR_ARG, "null argument");
  const Geo& g = c->g;
  const int X = g.gend - 2;
  if (X > CF_NOALLOC_MAX_PLANES)
    return fail(c, CF_ERR_ARG, "cf_driven_force_noalloc: X_out = %d exceeds the %d planes of its scratch", X,
                CF_NOALLOC_MAX_PLANES);
  double* scr = scratch_of(c->device);
  if (!scr) return fail(c, CF_ERR_STATE, "cf_driven_force_noalloc: no scratch address for device %d", c->device);
  Geo gg = g;
  const double* U2 = c->Ui + g.nf;
  if (U2_dev) {
    // a reference-layout device array: rows of Z_m, no padding
    gg.Zp = g.Zm;
    gg.plane = (long)g.Ym * g.Zm;
    gg.nf = (long)g.Xnp * gg.plane;
    U2 = U2_dev;
  } else if (!c->have_state) {
    return fail(c, CF_ERR_STATE, "no state loaded");
  }
  double Jsum = 0;  // Driven_force.cpp:45-47, in the reference's order
  for (int j = 2; j <= g.ny; j++) Jsum = Jsum + 1. / c->hprof[j];
  cf::launch_bulk_flux(c->stream, U2, c->prof, Jsum, scr, gg);
  cf::launch_sum_in_order(c->stream, scr, X, scr + CF_NOALLOC_MAX_PLANES);
  HIPCHK(c, hipGetLastError());
  double Q = 0;
  HIPCHK(c, hipMemcpyAsync(&Q, scr + CF_NOALLOC_MAX_PLANES, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  int r;
  if ((r = sync(c))) return r;
  Q = Q / c->p.X_out / c->p.Z_out;
  *Qall = Q;
  *ER_Q = fabs(Q0 - Q) / Q0;
  return CF_OK;
}

extern "C" int cf_statistic(cf_ctx* c, int step, int statistic_step, const double* const U_[5],
                            const char* directory) {
  if (!c || statistic_step < 1) return fail(c, CF_ERR_ARG, "statistic_step must be >= 1");
  const Geo& g = c->g;
  HostArrayScope hs(c);
  int r;
  if (U_) {
    if ((r = hs.enter())) return r;
    if ((r = h2d_n(c, c->Ui, U_, 5))) return r;
  } else if (!c->have_state) {
    return fail(c, CF_ERR_STATE, "no state loaded");
  }
  cf::launch_stat(c->stream, c->Ui, c->stat, g);
  // Statistic.cpp:552-911: the wall values of the y-face viscous flux,
  // added to the running sums in the reference's order (i, then k), and
  // the change of their running averages appended to Shear_stress.dat
  cf::launch_wall_shear(c->stream, c->Ui, c->prof, g, c->shear);
  HIPCHK(c, hipGetLastError());
  const size_t XZ = (size_t)c->p.X_out * c->p.Z_out;
  std::vector<double> wall(2 * XZ);
  HIPCHK(c, hipMemcpyAsync(wall.data(), c->shear, wall.size() * 8, hipMemcpyDeviceToHost, c->stream));
  if ((r = sync(c))) return r;
  {
    double avg[2];
    for (int w = 0; w < 2; w++) {
      double acc = c->sxy_sum[w];
      for (size_t q = 0; q < XZ; q++) acc = acc + wall[w * XZ + q];
      c->sxy_sum[w] = acc;
      avg[w] = acc / c->p.X_out / c->p.Z_out;
    }
    char spath[4096];
    snprintf(spath, sizeof spath, "%s/Shear_stress.dat", directory ? directory : ".");
    FILE* sf = fopen(spath, "a");
    if (!sf) return fail(c, CF_ERR_IO, "cannot open %s", spath);
    fprintf(sf, "%f\t%f\n", avg[0] - c->sxy_prev[0], avg[1] - c->sxy_prev[1]);
    if (fclose(sf) != 0) return fail(c, CF_ERR_IO, "cannot write %s", spath);
    c->sxy_prev[0] = avg[0];
    c->sxy_prev[1] = avg[1];
  }
  if (step % statistic_step != 0) return CF_OK;
  // Statistic.cpp:1008-1018: the statistic step resets the wall sums
  c->sxy_sum[0] = c->sxy_sum[1] = c->sxy_prev[0] = c->sxy_prev[1] = 0;
  std::vector<double> h((size_t)7 * g.Ym);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->stat, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemsetAsync(c->stat, 0, h.size() * 8, c->stream));
  if ((r = sync(c))) return r;
  // Statistic.cpp:1117-1146: average, one "%.16f" line per j = 2..ny for
  // U, V, W, UU, VV, WW, UV, into st%05d.dat
  const double inv = 1. / statistic_step / c->p.X_out / c->p.Z_out;
  char path[4096];
  snprintf(path, sizeof path, "%s/st%05d.dat", directory ? directory : ".", step);
  FILE* f = fopen(path, "w");
  if (!f) return fail(c, CF_ERR_IO, "cannot open %s", path);
  for (int q = 0; q < 7; q++)
    for (int j = 2; j <= g.ny; j++) fprintf(f, "%.16f\n", h[(size_t)q * g.Ym + j] * inv);
  if (fclose(f) != 0) return fail(c, CF_ERR_IO, "cannot write %s", path);
  return CF_OK;
}

extern "C" int cf_set_viscous_schedule(cf_ctx* c, int schedule) {
  if (!c || schedule < 0 || schedule > 1) return fail(c, CF_ERR_ARG, "bad viscous schedule");
  c->visc_schedule = schedule;
  return CF_OK;
}

extern "C" int cf_set_sweep_schedule(cf_ctx* c, int schedule) {
  if (!c || schedule < 0 || schedule > 2) return fail(c, CF_ERR_ARG, "bad schedule");
  if (c->fp32 && schedule != 0) return fail(c, CF_ERR_ARG, "FP32 sweeps run on the pipelined schedule (0) only");
  c->schedule = schedule;
  return CF_OK;
}

static int ensure_level(cf_ctx* c, double** level) {
  if (*level) return CF_OK;
  const size_t bytes = (size_t)c->g.nf * 5 * sizeof(double);
  if (hipMalloc(level, bytes) != hipSuccess) {
    *level = nullptr;
    return fail(c, CF_ERR_HIP, "hipMalloc of a time level failed");
  }
  HIPCHK(c, hipMemsetAsync(*level, 0, bytes, c->stream));
  return CF_OK;
}

extern "C" int cf_keep_levels(cf_ctx* c, int levels) {
  if (!c || levels < 2 || levels > 4) return fail(c, CF_ERR_ARG, "levels must be 2, 3 or 4");
  int r;
  if (levels >= 3 && !c->Uqq) {
    if ((r = ensure_level(c, &c->Uqq))) return r;
    c->uqq_valid = false;
  }
  if (levels >= 4 && !c->U3) {
    if ((r = ensure_level(c, &c->U3))) return r;
    c->u3_valid = false;
  }
  return sync(c);
}

extern "C" int cf_set_extrapolation_order(cf_ctx* c, int order) {
  if (!c || order < 2 || order > 3) return fail(c, CF_ERR_ARG, "extrapolation order must be 2 or 3");
  if (order == 3)
    if (int r = cf_keep_levels(c, 4)) return r;
  c->extrap_order = order;
  return CF_OK;
}

extern "C" int cf_set_initial_guess(cf_ctx* c, int order) {
  if (!c || order < 0 || order > 3) return fail(c, CF_ERR_ARG, "initial-guess order must be 0, 1, 2 or 3");
  if (order == 3 && !c->pred) return fail(c, CF_ERR_STATE, "no predictor attached (cf_set_predictor)");
  if (order >= 2 && !c->Uqq) {
    const size_t bytes = (size_t)c->g.nf * 5 * sizeof(double);
    if (hipMalloc(&c->Uqq, bytes) != hipSuccess) {
      c->Uqq = nullptr;
      return fail(c, CF_ERR_HIP, "hipMalloc of the U^{n-2} level failed");
    }
    HIPCHK(c, hipMemsetAsync(c->Uqq, 0, bytes, c->stream));
    c->uqq_valid = false;
  }
  c->guess_order = order;
  return sync(c);
}

// the extrapolation order once three levels are known
static int extrapolation_order(const cf_ctx* c) { return c->extrap_order == 3 && c->U3 && c->u3_valid ? 3 : 2; }

// The learned initial iterate with its fallback.  The network refines the
// quadratic extrapolation, so its guess and the quadratic extrapolation each
// take one inner iteration, and the step continues from the network's unless
// the quadratic's first residual is smaller in every er1..er4 ratio taken
// together (max_k er_k(net) / er_k(quadratic) <= 1 keeps the network's).  On
// a fallback with the cubic extrapolation selected (and four levels known),
// the step restarts from the cubic one, which is the better start once the
// flow changes smoothly.  *its receives the inner iterations run (2 or 3);
// all of them count.
static int learned_start(cf_ctx* c, double er[5], int* its) {
  const Geo& g = c->g;
  const cfp::Norm& nm = c->pred->norm;
  cfp::launch_feat(c->stream, c->Un, c->Uq, c->Uqq, c->prof, g, nm, c->px);
  const cfp::Vol v = cfp::make_vol(1, g.gend - 2, g.ny - 1, g.nz - 1, 0);
  cfp::launch_fwd(c->stream, v, c->px, true, c->pred->W1b, c->pred->W2b, c->pred->P, c->py, false);
  cfp::launch_apply(c->stream, c->Ui, c->Un, c->Uq, c->Uqq, c->prof, g, nm, c->py);
  HIPCHK(c, hipGetLastError());
  double er_net[5], er_ext[5];
  int r = cf_inner_iteration(c, er_net);
  if (r) return r;
  cf::launch_copy_interior(c->stream, c->Ui, c->Ualt, g);
  cf::launch_extrapolate(c->stream, c->Ui, c->Un, c->Uq, c->Uqq, c->U3, g, 2);
  if ((r = cf_inner_iteration(c, er_ext))) return r;
  double worst = 0;
  for (int k = 0; k < 4; k++) worst = std::max(worst, er_net[k] / er_ext[k]);
  c->pred_steps++;
  *its = 2;
  if (worst <= 1.0) {
    cf::launch_copy_interior(c->stream, c->Ualt, c->Ui, g);
    for (int k = 0; k < 5; k++) er[k] = er_net[k];
  } else {
    c->pred_fallbacks++;
    if (extrapolation_order(c) == 3) {
      cf::launch_extrapolate(c->stream, c->Ui, c->Un, c->Uq, c->Uqq, c->U3, g, 3);
      if ((r = cf_inner_iteration(c, er_ext))) return r;
      *its = 3;
    }
    for (int k = 0; k < 5; k++) er[k] = er_ext[k];
  }
  return sync(c);
}

extern "C" int cf_physical_step(cf_ctx* c, int* iterations, double er[5]) {
  if (!c || !iterations || !er) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  const bool learned = c->guess_order == 3 && c->pred && c->uqq_valid;
  if (c->guess_order > 0 && !learned) {
    // linear with two known levels, then quadratic (cubic with four levels
    // when asked for)
    const int order = c->guess_order >= 2 && c->uqq_valid ? extrapolation_order(c) : 1;
    cf::launch_extrapolate(c->stream, c->Ui, c->Un, c->Uq, c->Uqq, c->U3, c->g, order);
    HIPCHK(c, hipGetLastError());
  }
  for (int it = 1;; it++) {
    int r;
    if (learned && it == 1) {
      r = learned_start(c, er, &it);  // two or three inner iterations were run
    } else {
      r = cf_inner_iteration(c, er);
    }
    if (r) return r;
    for (int v = 0; v < 5; v++) {
      c->E[v] = std::max(c->E[v], er[v]);
      er[v] = er[v] / c->E[v];
    }
    bool conv = er[0] < c->p.tol && er[1] < c->p.tol && er[2] < c->p.tol && er[3] < c->p.tol;
    if (conv || it >= c->p.iteration_end_step) {
      *iterations = it;
      return cf_advance_levels(c);
    }
  }
}

extern "C" int cf_set_predictor(cf_ctx* c, cf_pred* p) {
  if (!c) return CF_ERR_ARG;
  if (!p) {
    c->pred = nullptr;
    if (c->guess_order == 3) c->guess_order = 2;
    return CF_OK;
  }
  if (!p->have_norm) return fail(c, CF_ERR_STATE, "the predictor has no normalisation");
  if (p->device != c->device) return fail(c, CF_ERR_ARG, "the predictor lives on another device");
  const Geo& g = c->g;
  const size_t cells = (size_t)(g.gend - 2) * (g.ny - 1) * (g.nz - 1);
  int r;
  if ((r = ensure_level(c, &c->Uqq)) || (r = ensure_level(c, &c->Ualt))) return r;
  if (!c->px && hipMalloc(&c->px, cells * cfp::CIN * 2) != hipSuccess) {
    c->px = nullptr;
    return fail(c, CF_ERR_HIP, "hipMalloc of the predictor features failed");
  }
  if (!c->py && hipMalloc(&c->py, cells * cfp::COUT * 4) != hipSuccess) {
    c->py = nullptr;
    return fail(c, CF_ERR_HIP, "hipMalloc of the predictor outputs failed");
  }
  // the predictor's weights are written on its own stream
  HIPCHK(c, hipStreamSynchronize(p->stream));
  c->pred = p;
  return sync(c);
}

extern "C" int cf_predictor_stats(cf_ctx* c, long long* steps, long long* fallbacks) {
  if (!c || !steps || !fallbacks) return CF_ERR_ARG;
  *steps = c->pred_steps;
  *fallbacks = c->pred_fallbacks;
  return CF_OK;
}

// crops of the last physical step: inputs from U^{n-2}, U^{n-1}, U^n and the
// converged U^{n+1}, which after cf_physical_step's level shift sit in U3,
// Uqq, Uq and U; the inner (VC-2)^3 cells of a crop lie in the domain, its
// margin may not (zero features, mask 0)
extern "C" int cf_pred_capture(cf_pred* p, cf_ctx* c, int ncrops, unsigned long long seed) {
  if (!p || !c || ncrops < 1) return CF_ERR_ARG;
  if (!c->have_state || !c->U3 || !c->u3_valid || !c->uqq_valid)
    return fail(c, CF_ERR_STATE, "capture needs four time levels (cf_keep_levels(ctx, 4) and three physical steps)");
  if (p->n_crops + ncrops > p->cap) {
    p->err = "the crop dataset is full";
    return CF_ERR_ARG;
  }
  const Geo& g = c->g;
  const int X = g.gend - 2, Y = g.ny - 1, Z = g.nz - 1, VC = p->VC;
  if (X < VC - 2 || Y < VC - 2 || Z < VC - 2) return fail(c, CF_ERR_ARG, "grid smaller than a crop");
  std::vector<int> org((size_t)3 * ncrops);
  unsigned long long s = seed;
  auto next = [&]() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const int ext[3] = {X, Y, Z};
  for (int q = 0; q < ncrops; q++)
    for (int d = 0; d < 3; d++) org[3 * q + d] = -1 + (int)(next() % (unsigned long long)(ext[d] - (VC - 2) + 1));
  int* dorg = nullptr;
  HIPCHK(c, hipMalloc(&dorg, org.size() * sizeof(int)));
  hipError_t e = hipMemcpyAsync(dorg, org.data(), org.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    float* D = p->D + p->n_crops * (long)VC * VC * VC * cfp::NDATA;
    cfp::launch_capture(c->stream, c->U3, c->Uqq, c->Uq, c->Un, c->prof, g, dorg, ncrops, VC, D);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dorg);
  if (e != hipSuccess) return fail(c, CF_ERR_HIP, "capture: %s", hipGetErrorString(e));
  p->n_crops += ncrops;
  return CF_OK;
}

extern "C" int cf_write_output(cf_ctx* c, int step, const char* dir) {
  if (!c || !dir) return CF_ERR_ARG;
  if (!c->have_state) return fail(c, CF_ERR_STATE, "no state loaded");
  size_t n = (size_t)c->p.X_out * c->p.Y_out * c->p.Z_out;
  std::vector<double> data(10 * n);
  double* Qn[5];
  double* Qm[5];
  for (int v = 0; v < 5; v++) {
    Qn[v] = data.data() + v * n;
    Qm[v] = data.data() + (5 + v) * n;
  }
  int r = cf_read_state(c, Qn, Qm, nullptr);
  if (r) return r;
  char path[4096];
  snprintf(path, sizeof path, "%s/LES%05d.bin", dir, step);
  FILE* f = fopen(path, "wb");
  if (!f) return fail(c, CF_ERR_IO, "cannot open %s", path);
  size_t put = fwrite(data.data(), sizeof(double), 10 * n, f);
  fclose(f);
  if (put != 10 * n) return fail(c, CF_ERR_IO, "short write %s", path);
  return CF_OK;
}
