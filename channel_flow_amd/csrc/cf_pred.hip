// C ABI of the learned initial iterate: the predictor object, its crop
// dataset, normalisation, training on the GPU (forward, backward and Adam
// all in this library's kernels, cf_predictor.hip) and test entry points.
// No counterpart in the reference (main.cpp:444-464 starts every physical
// step at U^n).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/channel_flow.h"
#include "cf_pred_internal.hpp"

static_assert(CF_PRED_NPARAM == cfp::NPARAM, "parameter count of the header");
static_assert(CF_PRED_NNORM * sizeof(float) == sizeof(cfp::Norm), "normalisation record of the header");

static int pfail(cf_pred* p, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (p) p->err = buf;
  return code;
}

#define PCHK(p, expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return pfail((p), CF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <class T>
static hipError_t dalloc(T** ptr, size_t n) {
  hipError_t e = hipMalloc(ptr, n * sizeof(T));
  if (e != hipSuccess) *ptr = nullptr;
  return e;
}

extern "C" int cf_pred_create(int dataset_crops, int batch_crops, cf_pred** out) {
  if (!out || dataset_crops < 1 || batch_crops < 1 || batch_crops > 256) return CF_ERR_ARG;
  *out = nullptr;
  cf_pred* p = new cf_pred();
  auto bail = [&](int code) {
    cf_pred_destroy(p);
    return code;
  };
  if (hipGetDevice(&p->device) != hipSuccess) return bail(CF_ERR_HIP);
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) return bail(CF_ERR_HIP);
  p->cap = dataset_crops;
  p->B = batch_crops;
  const int VC = p->VC;
  const long n3 = (long)VC * VC * VC, ni3 = (long)(VC - 2) * (VC - 2) * (VC - 2);
  const cfp::Vol v = pred_batch_vol(p->B, VC);
  const long hb = (v.inner_cells() + 63) / 64, db = cfp::make_vol(p->B, VC, VC, VC, 0).bricks();
  if (dalloc(&p->P, cfp::NPARAM) || dalloc(&p->G, cfp::NPARAM + 1) || dalloc(&p->M, cfp::NPARAM) ||
      dalloc(&p->V, cfp::NPARAM) || dalloc(&p->W1b, cfp::NW1) || dalloc(&p->W2b, cfp::NW2) ||
      dalloc(&p->W2Tb, cfp::NW2) || dalloc(&p->D, (size_t)p->cap * n3 * cfp::NDATA) ||
      dalloc(&p->x, (size_t)p->B * n3 * cfp::CIN) || dalloc(&p->h1, (size_t)p->B * n3 * cfp::CH) ||
      dalloc(&p->dz2, (size_t)p->B * n3 * cfp::CH) || dalloc(&p->z2, (size_t)p->B * ni3 * cfp::CH) ||
      dalloc(&p->tgt, (size_t)p->B * ni3 * cfp::COUT) || dalloc(&p->part_h, (size_t)hb * cfp::HEAD_PART) ||
      dalloc(&p->part_d, (size_t)db * cfp::NW1) || dalloc(&p->part_w, (size_t)p->splits * cfp::NW2))
    return bail(CF_ERR_HIP);
  // dz2 is written on the inner cells only: its margin stays zero
  if (hipMemsetAsync(p->dz2, 0, (size_t)p->B * n3 * cfp::CH * 2, p->stream) != hipSuccess ||
      hipMemsetAsync(p->x, 0, (size_t)p->B * n3 * cfp::CIN * 2, p->stream) != hipSuccess ||
      hipMemsetAsync(p->tgt, 0, (size_t)p->B * ni3 * cfp::COUT * 4, p->stream) != hipSuccess ||
      hipMemsetAsync(p->P, 0, cfp::NPARAM * 4, p->stream) != hipSuccess ||
      hipMemsetAsync(p->M, 0, cfp::NPARAM * 4, p->stream) != hipSuccess ||
      hipMemsetAsync(p->V, 0, cfp::NPARAM * 4, p->stream) != hipSuccess)
    return bail(CF_ERR_HIP);
  cfp::launch_sync_bf16(p->stream, p->P, p->W1b, p->W2b, p->W2Tb);
  if (hipStreamSynchronize(p->stream) != hipSuccess) return bail(CF_ERR_HIP);
  *out = p;
  return CF_OK;
}

extern "C" void cf_pred_destroy(cf_pred* p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  void* ptrs[] = {p->P,  p->G,   p->M,   p->V,   p->W1b,    p->W2b,    p->W2Tb,   p->D,     p->x,
                  p->h1, p->dz2, p->z2, p->tgt, p->part_h, p->part_d, p->part_w, p->losses};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

extern "C" const char* cf_pred_last_error(const cf_pred* p) { return p ? p->err.c_str() : "null predictor"; }

extern "C" int cf_pred_set_params(cf_pred* p, const float* params, int n) {
  if (!p || !params || n != cfp::NPARAM) return pfail(p, CF_ERR_ARG, "expected %d parameters", cfp::NPARAM);
  PCHK(p, hipMemcpyAsync(p->P, params, cfp::NPARAM * 4, hipMemcpyHostToDevice, p->stream));
  PCHK(p, hipMemsetAsync(p->M, 0, cfp::NPARAM * 4, p->stream));
  PCHK(p, hipMemsetAsync(p->V, 0, cfp::NPARAM * 4, p->stream));
  p->adam_step = 0;
  cfp::launch_sync_bf16(p->stream, p->P, p->W1b, p->W2b, p->W2Tb);
  PCHK(p, hipGetLastError());
  PCHK(p, hipStreamSynchronize(p->stream));
  return CF_OK;
}

extern "C" int cf_pred_get_params(cf_pred* p, float* params, int n) {
  if (!p || !params || n != cfp::NPARAM) return pfail(p, CF_ERR_ARG, "expected %d parameters", cfp::NPARAM);
  PCHK(p, hipMemcpyAsync(params, p->P, cfp::NPARAM * 4, hipMemcpyDeviceToHost, p->stream));
  PCHK(p, hipStreamSynchronize(p->stream));
  return CF_OK;
}

// He-style initial weights from a SplitMix64 stream (Box-Muller normals);
// the output layer starts at zero, so the untrained network adds nothing to
// the base guess
extern "C" int cf_pred_init_params(cf_pred* p, unsigned long long seed) {
  if (!p) return CF_ERR_ARG;
  std::vector<float> h(cfp::NPARAM, 0.f);
  unsigned long long s = seed;
  auto next = [&]() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto normal = [&]() { return std::sqrt(-2.0 * std::log(next())) * std::cos(6.283185307179586 * next()); };
  for (int q = 0; q < cfp::NW1; q++) h[cfp::OFF_W1 + q] = (float)(normal() * std::sqrt(1.0 / cfp::CIN));
  for (int q = 0; q < cfp::NW2; q++) h[cfp::OFF_W2 + q] = (float)(normal() * std::sqrt(2.0 / (27.0 * cfp::CH)));
  return cf_pred_set_params(p, h.data(), cfp::NPARAM);
}

extern "C" int cf_pred_set_norm(cf_pred* p, const float norm[CF_PRED_NNORM]) {
  if (!p || !norm) return CF_ERR_ARG;
  std::memcpy(&p->norm, norm, sizeof(cfp::Norm));
  for (int v = 0; v < 5; v++)
    if (!(p->norm.sig[v] > 0) || !(p->norm.s[v] > 0) || !(p->norm.sr[v] > 0))
      return pfail(p, CF_ERR_ARG, "normalisation scales must be positive");
  p->have_norm = true;
  return CF_OK;
}

extern "C" int cf_pred_get_norm(cf_pred* p, float norm[CF_PRED_NNORM]) {
  if (!p || !norm) return CF_ERR_ARG;
  if (!p->have_norm) return pfail(p, CF_ERR_STATE, "no normalisation (cf_pred_fit_norm / cf_pred_set_norm)");
  std::memcpy(norm, &p->norm, sizeof(cfp::Norm));
  return CF_OK;
}

extern "C" int cf_pred_dataset_size(const cf_pred* p) { return p ? (int)p->n_crops : 0; }

extern "C" int cf_pred_load_dataset(cf_pred* p, const float* crops, int ncrops) {
  if (!p || !crops || ncrops < 1) return CF_ERR_ARG;
  if (p->n_crops + ncrops > p->cap) return pfail(p, CF_ERR_ARG, "the crop dataset is full");
  const size_t per = (size_t)p->VC * p->VC * p->VC * cfp::NDATA;
  PCHK(p, hipMemcpyAsync(p->D + p->n_crops * per, crops, (size_t)ncrops * per * 4, hipMemcpyHostToDevice, p->stream));
  PCHK(p, hipStreamSynchronize(p->stream));
  p->n_crops += ncrops;
  return CF_OK;
}

// feature scales from the crops [0, n): mean and standard deviation of Q^n,
// RMS of the increments, and the base guess (0, d1 or 2 d1 - d2) with the
// smallest error relative to the increment, whose RMS error becomes sr
extern "C" int cf_pred_fit_norm(cf_pred* p, int n, double stats[CF_PRED_NSTATS]) {
  if (!p || n < 1 || n > p->n_crops) return pfail(p, CF_ERR_ARG, "crop count out of range");
  const long cells = (long)n * p->VC * p->VC * p->VC;
  double* dpart = nullptr;
  PCHK(p, dalloc(&dpart, (size_t)256 * cfp::NSTAT));
  const int nb = cfp::launch_stats(p->stream, p->D, cells, dpart);
  std::vector<double> h((size_t)nb * cfp::NSTAT);
  hipError_t e = hipMemcpyAsync(h.data(), dpart, h.size() * 8, hipMemcpyDeviceToHost, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  (void)hipFree(dpart);
  if (e != hipSuccess) return pfail(p, CF_ERR_HIP, "statistics: %s", hipGetErrorString(e));
  double S[cfp::NSTAT] = {};
  for (int b = 0; b < nb; b++)
    for (int q = 0; q < cfp::NSTAT; q++) S[q] += h[(size_t)b * cfp::NSTAT + q];
  const double N = S[35];
  if (!(N > 0)) return pfail(p, CF_ERR_STATE, "no in-domain cells in the crops");
  cfp::Norm nm{};
  double rel[3] = {0, 0, 0};
  for (int v = 0; v < 5; v++)
    for (int b = 0; b < 3; b++) rel[b] += S[20 + 5 * b + v] / std::max(S[20 + v], 1e-300);
  const int best = rel[2] < rel[1] ? (rel[2] < rel[0] ? 2 : 0) : (rel[1] < rel[0] ? 1 : 0);
  const float c1[3] = {0.f, 1.f, 2.f}, c2[3] = {0.f, 0.f, -1.f};
  nm.c1 = c1[best];
  nm.c2 = c2[best];
  for (int v = 0; v < 5; v++) {
    const double mu = S[v] / N, var = S[5 + v] / N - mu * mu;
    nm.mu[v] = (float)mu;
    nm.sig[v] = (float)(var > 0 ? std::sqrt(var) : std::max(std::fabs(mu), 1.0));
    const double s = std::sqrt(S[10 + v] / N), sr = std::sqrt(S[20 + 5 * best + v] / N);
    nm.s[v] = (float)(s > 0 ? s : 1.0);
    nm.sr[v] = (float)(sr > 0 ? sr : nm.s[v]);
  }
  p->norm = nm;
  p->have_norm = true;
  if (stats)
    for (int q = 0; q < cfp::NSTAT; q++) stats[q] = S[q];
  return CF_OK;
}

// loss and gradient of one batch already packed in x / tgt (n crops): the
// pointwise layer, the forward conv, the loss head, both backward convs and
// the ordered reduction into G (summed squared error at G[NPARAM])
static void batch_gradient(cf_pred* p, int n, unsigned short* x, unsigned short* h1, float* z2, float* tgt,
                           unsigned short* dz2, float* ph, float* pd, float* pw) {
  const cfp::Vol v = pred_batch_vol(n, p->VC);
  const float inv_norm = 2.0f / (5.0f * (float)v.inner_cells());
  cfp::launch_l1(p->stream, v, x, p->W1b, h1);
  cfp::launch_fwd(p->stream, v, h1, false, p->W1b, p->W2b, p->P, z2, true);
  const int hb = cfp::launch_head(p->stream, v, z2, tgt, p->P, inv_norm, dz2, ph);
  const int db = cfp::launch_bwd_data(p->stream, v, dz2, p->W2Tb, x, pd);
  cfp::launch_bwd_weight(p->stream, v, dz2, h1, p->splits, pw);
  cfp::launch_reduce(p->stream, ph, hb, pd, db, pw, p->splits, p->G);
}

extern "C" int cf_pred_train(cf_pred* p, int n_train, int steps, float lr, unsigned long long seed, float* losses) {
  if (!p || steps < 0 || !(lr > 0)) return CF_ERR_ARG;
  if (n_train < 1 || n_train > p->n_crops) return pfail(p, CF_ERR_ARG, "n_train out of range");
  if (!p->have_norm) return pfail(p, CF_ERR_STATE, "no normalisation (cf_pred_fit_norm / cf_pred_set_norm)");
  if (losses && p->losses_cap < steps) {
    if (p->losses) (void)hipFree(p->losses);
    PCHK(p, dalloc(&p->losses, (size_t)steps));
    p->losses_cap = steps;
  }
  for (int s = 0; s < steps; s++) {
    cfp::launch_pack(p->stream, p->D, p->VC, n_train, 0, p->B, seed, s, p->norm, p->x, p->tgt);
    batch_gradient(p, p->B, p->x, p->h1, p->z2, p->tgt, p->dz2, p->part_h, p->part_d, p->part_w);
    // cosine decay of the step size over the run
    const float lr_s = lr * 0.5f * (1.0f + std::cos(3.14159265f * (float)s / (float)steps));
    cfp::launch_adam(p->stream, p->P, p->G, p->M, p->V, ++p->adam_step, lr_s, p->W1b, p->W2b, p->W2Tb);
    if (losses)
      PCHK(p, hipMemcpyAsync(p->losses + s, p->G + cfp::NPARAM, 4, hipMemcpyDeviceToDevice, p->stream));
    PCHK(p, hipGetLastError());
  }
  if (losses && steps > 0) {
    PCHK(p, hipMemcpyAsync(losses, p->losses, (size_t)steps * 4, hipMemcpyDeviceToHost, p->stream));
  }
  PCHK(p, hipStreamSynchronize(p->stream));
  if (losses) {
    const double N = 5.0 * p->B * (double)(p->VC - 2) * (p->VC - 2) * (p->VC - 2);
    for (int s = 0; s < steps; s++) losses[s] = (float)(losses[s] / N);
  }
  return CF_OK;
}

// mean squared error of the network over crops [first, first + n) and of the
// base guess alone (the network predicting zero), in normalised units
extern "C" int cf_pred_eval(cf_pred* p, int first, int n, double* loss, double* base_loss) {
  if (!p || !loss || !base_loss) return CF_ERR_ARG;
  if (first < 0 || n < 1 || first + n > p->n_crops) return pfail(p, CF_ERR_ARG, "crop range out of the dataset");
  if (!p->have_norm) return pfail(p, CF_ERR_STATE, "no normalisation");
  const long ni3 = (long)(p->VC - 2) * (p->VC - 2) * (p->VC - 2);
  double se = 0, sb = 0, cnt = 0;
  std::vector<float> t((size_t)p->B * ni3 * cfp::COUT);
  for (int c0 = 0; c0 < n; c0 += p->B) {
    const int nb = std::min(p->B, n - c0);
    cfp::launch_pack(p->stream, p->D, p->VC, p->n_crops, first + c0, nb, 0, -1, p->norm, p->x, p->tgt);
    batch_gradient(p, nb, p->x, p->h1, p->z2, p->tgt, p->dz2, p->part_h, p->part_d, p->part_w);
    float g = 0;
    PCHK(p, hipMemcpyAsync(&g, p->G + cfp::NPARAM, 4, hipMemcpyDeviceToHost, p->stream));
    PCHK(p, hipMemcpyAsync(t.data(), p->tgt, (size_t)nb * ni3 * cfp::COUT * 4, hipMemcpyDeviceToHost, p->stream));
    PCHK(p, hipStreamSynchronize(p->stream));
    se += g;
    for (long q = 0; q < (long)nb * ni3 * cfp::COUT; q++) sb += (double)t[q] * t[q];
    cnt += (double)nb * ni3 * cfp::COUT;
  }
  *loss = se / cnt;
  *base_loss = sb / cnt;
  return CF_OK;
}

// the network's outputs y on the inner region of n volumes given as host
// float features ([n][VI][VJ][VK][16], rounded to bf16 here)
extern "C" int cf_pred_forward(cf_pred* p, int n, int VI, int VJ, int VK, int margin, const float* x, float* y) {
  if (!p || !x || !y || n < 1 || margin < 0 || VI - 2 * margin < 1 || VJ - 2 * margin < 1 || VK - 2 * margin < 1)
    return pfail(p, CF_ERR_ARG, "bad volume");
  const cfp::Vol v = cfp::make_vol(n, VI, VJ, VK, margin);
  float* xf = nullptr;
  unsigned short* xb = nullptr;
  float* yd = nullptr;
  int rc = CF_OK;
  if (dalloc(&xf, (size_t)v.cells() * cfp::CIN) || dalloc(&xb, (size_t)v.cells() * cfp::CIN) ||
      dalloc(&yd, (size_t)v.inner_cells() * cfp::COUT)) {
    rc = pfail(p, CF_ERR_HIP, "out of device memory");
  } else {
    hipError_t e = hipMemcpyAsync(xf, x, (size_t)v.cells() * cfp::CIN * 4, hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) {
      cfp::launch_to_bf16(p->stream, xf, xb, v.cells() * cfp::CIN);
      cfp::launch_fwd(p->stream, v, xb, true, p->W1b, p->W2b, p->P, yd, false);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(y, yd, (size_t)v.inner_cells() * cfp::COUT * 4, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) rc = pfail(p, CF_ERR_HIP, "forward: %s", hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(p->stream);
  if (xf) (void)hipFree(xf);
  if (xb) (void)hipFree(xb);
  if (yd) (void)hipFree(yd);
  return rc;
}

// loss and gradient of n crops of VC^3 cells given as host floats (features
// [n][VC^3][16], targets [n][(VC-2)^3][5]); grad receives NPARAM values, loss
// the mean squared error
extern "C" int cf_pred_gradient(cf_pred* p, int n, const float* x, const float* target, float* grad, double* loss) {
  if (!p || !x || !target || !grad || !loss || n < 1) return CF_ERR_ARG;
  const int VC = p->VC;
  const cfp::Vol v = pred_batch_vol(n, VC);
  const long n3 = v.cells(), ni3 = v.inner_cells();
  const cfp::Vol full = cfp::make_vol(n, VC, VC, VC, 0);
  float *xf = nullptr, *z2 = nullptr, *tg = nullptr, *ph = nullptr, *pd = nullptr, *pw = nullptr;
  unsigned short *xb = nullptr, *h1 = nullptr, *dz2 = nullptr;
  int rc = CF_OK;
  if (dalloc(&xf, (size_t)n3 * cfp::CIN) || dalloc(&xb, (size_t)n3 * cfp::CIN) || dalloc(&h1, (size_t)n3 * cfp::CH) ||
      dalloc(&dz2, (size_t)n3 * cfp::CH) || dalloc(&z2, (size_t)ni3 * cfp::CH) || dalloc(&tg, (size_t)ni3 * cfp::COUT) ||
      dalloc(&ph, (size_t)((ni3 + 63) / 64) * cfp::HEAD_PART) || dalloc(&pd, (size_t)full.bricks() * cfp::NW1) ||
      dalloc(&pw, (size_t)p->splits * cfp::NW2)) {
    rc = pfail(p, CF_ERR_HIP, "out of device memory");
  } else {
    hipError_t e = hipMemcpyAsync(xf, x, (size_t)n3 * cfp::CIN * 4, hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tg, target, (size_t)ni3 * cfp::COUT * 4, hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) e = hipMemsetAsync(dz2, 0, (size_t)n3 * cfp::CH * 2, p->stream);
    if (e == hipSuccess) {
      cfp::launch_to_bf16(p->stream, xf, xb, n3 * cfp::CIN);
      batch_gradient(p, n, xb, h1, z2, tg, dz2, ph, pd, pw);
      e = hipGetLastError();
    }
    float sse = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(grad, p->G, cfp::NPARAM * 4, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&sse, p->G + cfp::NPARAM, 4, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) rc = pfail(p, CF_ERR_HIP, "gradient: %s", hipGetErrorString(e));
    *loss = sse / (5.0 * (double)ni3);
  }
  (void)hipStreamSynchronize(p->stream);
  void* ptrs[] = {xf, z2, tg, ph, pd, pw, xb, h1, dz2};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  return rc;
}

// times one forward (inference) of the network over a domain of X x Y x Z
// cells of random features with HIP events on the predictor's stream:
// ms[0] the pointwise + conv + head kernel; also one training step on a
// batch: ms[1] forward (pointwise + conv), ms[2] loss head, ms[3] backward
// data, ms[4] backward weight, ms[5] reduction + Adam
extern "C" int cf_pred_profile(cf_pred* p, int X, int Y, int Z, double ms[CF_PRED_PROFILE_STAGES]) {
  if (!p || !ms || X < 1 || Y < 1 || Z < 1) return CF_ERR_ARG;
  const cfp::Vol v = cfp::make_vol(1, X, Y, Z, 0);
  unsigned short* xb = nullptr;
  float* yd = nullptr;
  hipEvent_t ev[8] = {};
  int rc = CF_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = CF_ERR_HIP;
  if (rc == CF_OK && (dalloc(&xb, (size_t)v.cells() * cfp::CIN) || dalloc(&yd, (size_t)v.inner_cells() * cfp::COUT)))
    rc = pfail(p, CF_ERR_HIP, "out of device memory");
  if (rc == CF_OK) {
    // features: the bf16 bit patterns of a counter hash, kept finite and small
    std::vector<unsigned short> h((size_t)v.cells() * cfp::CIN);
    unsigned long long s = 12345;
    for (auto& q : h) {
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      q = (unsigned short)(0x3C00u + ((s >> 40) & 0x3FFu));  // 2^-7 .. 2^-6
    }
    hipError_t e = hipMemcpyAsync(xb, h.data(), h.size() * 2, hipMemcpyHostToDevice, p->stream);
    const cfp::Vol bv = pred_batch_vol(p->B, p->VC);
    const float inv_norm = 2.0f / (5.0f * (float)bv.inner_cells());
    if (e == hipSuccess) {
      cfp::launch_fwd(p->stream, v, xb, true, p->W1b, p->W2b, p->P, yd, false);  // warm-up
      (void)hipEventRecord(ev[0], p->stream);
      cfp::launch_fwd(p->stream, v, xb, true, p->W1b, p->W2b, p->P, yd, false);
      (void)hipEventRecord(ev[1], p->stream);
      cfp::launch_l1(p->stream, bv, p->x, p->W1b, p->h1);
      cfp::launch_fwd(p->stream, bv, p->h1, false, p->W1b, p->W2b, p->P, p->z2, true);
      (void)hipEventRecord(ev[2], p->stream);
      const int hb = cfp::launch_head(p->stream, bv, p->z2, p->tgt, p->P, inv_norm, p->dz2, p->part_h);
      (void)hipEventRecord(ev[3], p->stream);
      const int db = cfp::launch_bwd_data(p->stream, bv, p->dz2, p->W2Tb, p->x, p->part_d);
      (void)hipEventRecord(ev[4], p->stream);
      cfp::launch_bwd_weight(p->stream, bv, p->dz2, p->h1, p->splits, p->part_w);
      (void)hipEventRecord(ev[5], p->stream);
      cfp::launch_reduce(p->stream, p->part_h, hb, p->part_d, db, p->part_w, p->splits, p->G);
      (void)hipEventRecord(ev[6], p->stream);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    for (int q = 0; e == hipSuccess && q < 6; q++) {
      float t = 0;
      e = hipEventElapsedTime(&t, ev[q], ev[q + 1]);
      ms[q] = t;
    }
    if (e != hipSuccess) rc = pfail(p, CF_ERR_HIP, "profile: %s", hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(p->stream);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (xb) (void)hipFree(xb);
  if (yd) (void)hipFree(yd);
  return rc;
}
