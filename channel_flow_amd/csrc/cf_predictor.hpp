// Learned initial iterate of a physical step: kernels and their launchers.
//
// The network (no counterpart in the reference, whose physical step starts
// at U^n, main.cpp:444-464) maps 16 features of the last three states to 5
// outputs per cell:
//   h1 = W1 x'                        pointwise 16 -> 64 (x'[15] = in-domain mask)
//   z2 = b2 + sum_t W2[t] h1(p + o_t)  27-point 3x3x3 conv 64 -> 64
//   y  = b3 + W3 gelu(z2)             pointwise 64 -> 5
// W1, W2 and the activations h1 (and, in training, dz2) are bfloat16 MFMA
// operands; every sum accumulates in FP32.  Volumes are channel-last,
// [vol][VI][VJ][VK][C], with zeros taken outside a volume.
#pragma once
#include <hip/hip_runtime.h>

#include "cf_device.hpp"

namespace cfp {

constexpr int CIN = 16, CH = 64, COUT = 5;
constexpr int NW1 = CH * CIN, NW2 = 27 * CH * CH, NB2 = CH, NW3 = COUT * CH, NB3 = COUT;
constexpr int OFF_W1 = 0, OFF_W2 = OFF_W1 + NW1, OFF_B2 = OFF_W2 + NW2, OFF_W3 = OFF_B2 + NB2, OFF_B3 = OFF_W3 + NW3;
constexpr int NPARAM = OFF_B3 + NB3;  // 112005 parameters, FP32 masters
constexpr int NDATA = 24;             // dataset channels per crop cell: Q^n 5, d1 5, d2 5, target 5, mask, 3 zero
constexpr int HEAD_PART = NW3 + NB2 + NB3 + 1;  // per-block partials of the loss head: dW3, db2, db3, loss
// output bricks of the conv kernels: 4 x 8 x 8 cells (i, j, k), one i-plane per wave
constexpr int BI = 4, BJ = 8, BK = 8, BCELLS = BI * BJ * BK;

// n volumes of VI x VJ x VK cells; outputs on the inner region [m, V - m)
struct Vol {
  int n, VI, VJ, VK, m;
  int I, J, K;         // inner sizes
  int nbi, nbj, nbk;   // bricks over the inner region
  __host__ __device__ long bricks() const { return (long)n * nbi * nbj * nbk; }
  __host__ __device__ long cells() const { return (long)n * VI * VJ * VK; }
  __host__ __device__ long inner_cells() const { return (long)n * I * J * K; }
};
Vol make_vol(int n, int VI, int VJ, int VK, int m);

// Feature scales: x' = ((Q^n - mu) / sig, d1 / s, d2 / s, mask) with
// d1 = Q^n - Q^{n-1}, d2 = Q^{n-1} - Q^{n-2}; the prediction is
// Q^{n+1} = Q^n + c1 d1 + c2 d2 + sr y.
struct Norm {
  float mu[5], sig[5], s[5], sr[5];
  float c1, c2;
};

// forward: x' (bf16, 16 channels) with the pointwise layer fused, or a stored
// h1 (bf16, 64 channels); y (5 per inner cell, infer) or z2 (64, train)
void launch_fwd(hipStream_t st, const Vol& v, const unsigned short* src, bool src_is_x, const unsigned short* W1b,
                const unsigned short* W2b, const float* P, float* out, bool train);
// h1 = bf16(W1 x') on every cell of the volumes
void launch_l1(hipStream_t st, const Vol& v, const unsigned short* x, const unsigned short* W1b, unsigned short* h1);
// loss head on the inner cells: dz2 (bf16, written into a volume whose margin
// stays zero) and per-block partials [block][HEAD_PART]; returns the block count
int launch_head(hipStream_t st, const Vol& v, const float* z2, const float* target, const float* P, float inv_norm,
                unsigned short* dz2, float* part);
// dh1 = transposed conv of dz2 on every cell, and per-brick partials of dW1
// [brick][NW1]; returns the brick count
int launch_bwd_data(hipStream_t st, const Vol& v, const unsigned short* dz2, const unsigned short* W2Tb,
                    const unsigned short* x, float* part);
// per-split partials of dW2 [split][27][64][64]
void launch_bwd_weight(hipStream_t st, const Vol& v, const unsigned short* dz2, const unsigned short* h1, int splits,
                       float* part);
// gradient (NPARAM) and loss (at [NPARAM]) from the partials, in a fixed order
void launch_reduce(hipStream_t st, const float* head_part, int head_blocks, const float* data_part, int data_blocks,
                   const float* w_part, int splits, float* grad);
// Adam on the FP32 masters, then the bf16 operand copies
void launch_adam(hipStream_t st, float* P, const float* G, float* M, float* V, int step, float lr,
                 unsigned short* W1b, unsigned short* W2b, unsigned short* W2Tb);
void launch_sync_bf16(hipStream_t st, const float* P, unsigned short* W1b, unsigned short* W2b, unsigned short* W2Tb);
// host float arrays -> bf16 (round to nearest even)
void launch_to_bf16(hipStream_t st, const float* src, unsigned short* dst, long n);

// training crops: dataset [crop][VC^3][NDATA] from four solver levels (U =
// Q/J in the solver's device layout); origins[3 c ..] = interior cell of the
// crop's first cell (may lie outside the domain)
void launch_capture(hipStream_t st, const double* Um3, const double* Um2, const double* Um1, const double* Un,
                    const double* prof, const cf::Geo& g, const int* origins, int ncrops, int VC, float* D);
// a batch: crops picked by a counter hash of (seed, step) from [0, n_avail)
// (or the crops first .. first + n - 1 when step < 0) -> x' and targets
void launch_pack(hipStream_t st, const float* D, int VC, long n_avail, int first, int nb, unsigned long long seed,
                 int step, const Norm& nm, unsigned short* x, float* target);
// x' of the whole domain from the solver's levels n, n-1, n-2
void launch_feat(hipStream_t st, const double* Un, const double* Uq, const double* Uqq, const double* prof,
                 const cf::Geo& g, const Norm& nm, unsigned short* x);
// U_ <- (Q^n + c1 d1 + c2 d2 + sr y) / J on the interior
void launch_apply(hipStream_t st, double* Ui, const double* Un, const double* Uq, const double* Uqq,
                  const double* prof, const cf::Geo& g, const Norm& nm, const float* y);
// sums over the dataset's in-domain cells for the feature scales, per
// variable: sum Q^n, sum (Q^n)^2, sum d1^2, sum d2^2, and the sums of the
// squared target minus each base guess (0, d1, 2 d1 - d2); then the cell
// count: NSTAT FP64 sums per block; returns the block count
constexpr int NSTAT = 36;
int launch_stats(hipStream_t st, const float* D, long cells, double* part);

}  // namespace cfp
