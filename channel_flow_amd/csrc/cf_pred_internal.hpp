// The predictor object behind the cf_pred_* entry points (shared by
// cf_pred.hip, which trains it, and cf_api.hip, which captures crops from a
// solver context and runs it inside cf_physical_step).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "cf_predictor.hpp"

struct cf_pred {
  int device = 0;
  hipStream_t stream = nullptr;
  // parameters: FP32 masters, Adam moments, gradient (+ summed squared error)
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
  unsigned short *W1b = nullptr, *W2b = nullptr, *W2Tb = nullptr;  // bf16 operand copies
  int adam_step = 0;
  cfp::Norm norm{};
  bool have_norm = false;
  // training crops: [cap][VC^3][NDATA]
  int VC = 16;
  long cap = 0, n_crops = 0;
  float* D = nullptr;
  // one training batch of B crops
  int B = 0, splits = 16;
  unsigned short *x = nullptr, *h1 = nullptr, *dz2 = nullptr;
  float *z2 = nullptr, *tgt = nullptr, *part_h = nullptr, *part_d = nullptr, *part_w = nullptr, *losses = nullptr;
  long losses_cap = 0;
  std::string err;
};

// the one place the conv's batch buffers are sized: a batch of `n` crops of
// VC^3 cells (inner (VC - 2)^3)
inline cfp::Vol pred_batch_vol(int n, int VC) { return cfp::make_vol(n, VC, VC, VC, 1); }
