// HIP kernels of the learned initial iterate (gfx950, bfloat16 MFMA with FP32
// accumulation).  No counterpart in the reference: its physical step starts
// the inner iterations at U^n (main.cpp:444-464).
//
// The 27-point convolution is an implicit GEMM per output brick of 4 x 8 x 8
// cells: the brick's 6 x 10 x 10 halo of 64-channel bf16 activations sits in
// LDS (one 128-byte row per cell, its eight 16-byte chunks XOR-swizzled by
// the cell index so that eight consecutive cells' reads of one chunk hit
// eight different bank groups; 76.8 KB, two blocks per CU), each of the four waves owns one i-plane of
// 64 output cells x 64 channels (16 accumulator tiles of 16 x 16), and the
// K loop runs over 27 taps x 2 chunks of 32 input channels with
// __builtin_amdgcn_mfma_f32_16x16x32_bf16: the A operand is a shifted row
// gather from the LDS tile, the B operand a 16-byte read of the weights.
// The pointwise 16 -> 64 layer that fills the tile uses
// __builtin_amdgcn_mfma_f32_16x16x16bf16_1k.
//
// MFMA lane layouts (16 x 16 output tile, lane l, g = l / 16, r = l % 16):
//   16x16x32: A[m = r][k = 8 g + e], B[k = 8 g + e][n = r], e < 8
//   16x16x16: A[m = r][k = 4 g + e], B[k = 4 g + e][n = r], e < 4
//   D[m = 4 g + q][n = r], q < 4
#include "cf_predictor.hpp"

namespace cfp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int HI = BI + 2, HJ = BJ + 2, HK = BK + 2, HN = HI * HJ * HK;  // halo brick: 600 cells
constexpr int AS = CH + 1;   // floats per LDS row of the per-cell epilogue arrays
constexpr int TS = BCELLS + 8;  // bf16 per LDS row of the transposed tiles of the weight gradient
static_assert(BCELLS == 256 && BI == 4, "one i-plane of 64 cells per wave, 256 threads");
static_assert(BCELLS * AS * 4 + BCELLS * CIN * 2 <= HN * CH * 2, "epilogue arrays fit in the tile");

// element index of 16-byte chunk `chunk` (0..7) of halo cell h in the tile
__device__ __forceinline__ int tix(int h, int chunk) { return h * CH + ((chunk ^ (h & 7)) << 3); }

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite values)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ f32x4 mma16(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// eight bf16 from a 16-byte aligned address (global or LDS)
__device__ __forceinline__ bf16x8 ld8(const unsigned short* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

__device__ __forceinline__ long vcell(const Vol& v, int vol, int i, int j, int k) {
  return (((long)vol * v.VI + i) * v.VJ + j) * v.VK + k;
}
__device__ __forceinline__ bool inside(const Vol& v, int i, int j, int k) {
  return i >= 0 && i < v.VI && j >= 0 && j < v.VJ && k >= 0 && k < v.VK;
}
// brick b of the inner region: its volume and the volume coordinates of its first cell
__device__ __forceinline__ void brick_of(const Vol& v, long b, int* vol, int* i0, int* j0, int* k0) {
  const int bk = (int)(b % v.nbk);
  b /= v.nbk;
  const int bj = (int)(b % v.nbj);
  b /= v.nbj;
  const int bi = (int)(b % v.nbi);
  *vol = (int)(b / v.nbi);
  *i0 = v.m + bi * BI;
  *j0 = v.m + bj * BJ;
  *k0 = v.m + bk * BK;
}
// brick-local output cell o = (ii * BJ + jj) * BK + kk
__device__ __forceinline__ void cell_of(int o, int* ii, int* jj, int* kk) {
  *ii = o / (BJ * BK);
  *jj = (o / BK) % BJ;
  *kk = o % BK;
}

// halo tile from x' with the pointwise layer: h1 = bf16(W1 x'), 0 outside the volumes
__device__ void fill_tile_x(unsigned short* tile, const Vol& v, int vol, int i0, int j0, int k0,
                            const unsigned short* __restrict__ x, const unsigned short* __restrict__ W1b) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  s16x4 bw[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) bw[nt] = *reinterpret_cast<const s16x4*>(W1b + (nt * 16 + r) * CIN + 4 * g);
  for (int rt = w; rt < (HN + 15) / 16; rt += 4) {
    const int h = rt * 16 + r;
    s16x4 a = {0, 0, 0, 0};
    if (h < HN) {
      const int i = i0 - 1 + h / (HJ * HK), j = j0 - 1 + (h / HK) % HJ, k = k0 - 1 + h % HK;
      if (inside(v, i, j, k)) a = *reinterpret_cast<const s16x4*>(x + vcell(v, vol, i, j, k) * CIN + 4 * g);
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const f32x4 d = mma16(a, bw[nt], f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int hc = rt * 16 + 4 * g + q;
        if (hc < HN) tile[tix(hc, nt * 2 + (r >> 3)) + (r & 7)] = f2bf(d[q]);
      }
    }
  }
}

// halo tile from a stored 64-channel bf16 volume, 0 outside the volumes
__device__ void fill_tile_h(unsigned short* tile, const Vol& v, int vol, int i0, int j0, int k0,
                            const unsigned short* __restrict__ h1) {
  for (int q = threadIdx.x; q < HN * 8; q += blockDim.x) {
    const int h = q >> 3, c = q & 7;
    const int i = i0 - 1 + h / (HJ * HK), j = j0 - 1 + (h / HK) % HJ, k = k0 - 1 + h % HK;
    u32x4 val = {0u, 0u, 0u, 0u};
    if (inside(v, i, j, k)) val = *reinterpret_cast<const u32x4*>(h1 + vcell(v, vol, i, j, k) * CH + c * 8);
    *reinterpret_cast<u32x4*>(tile + tix(h, c)) = val;
  }
}

// acc[mt][nt] += sum over taps and input channels of W[t][n][k] tile(cell + o_t)[k]
// for the wave's output cells (ii = wave, jj = 2 mt + m / 8, kk = m % 8) and
// output channels nt * 16 + n.  W: [27][64 out][64 in] bf16.
__device__ __forceinline__ void conv27(const unsigned short* tile, const unsigned short* __restrict__ W,
                                       f32x4 (&acc)[4][4]) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  for (int t = 0; t < 27; t++) {
    const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      bf16x8 bv[4];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) bv[nt] = ld8(W + ((t * CH + nt * 16 + r) * CH + c * 32 + 8 * g));
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const int jj = 2 * mt + (r >> 3), kk = r & 7;
        const int h = ((w + 1 + di) * HJ + (jj + 1 + dj)) * HK + (kk + 1 + dk);
        const bf16x8 av = ld8(tile + tix(h, c * 4 + g));
#pragma unroll
        for (int nt = 0; nt < 4; nt++) acc[mt][nt] = mma32(av, bv[nt], acc[mt][nt]);
      }
    }
  }
}

// forward over the inner bricks: infer writes y [inner cell][5], train writes
// z2 = conv + b2 [inner cell][64]
template <bool SRC_X, bool TRAIN>
__global__ void __launch_bounds__(256, 2) k_fwd(Vol v, const unsigned short* __restrict__ src,
                                             const unsigned short* __restrict__ W1b,
                                             const unsigned short* __restrict__ W2b, const float* __restrict__ P,
                                             float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) unsigned short tile[HN * CH];
  int vol, i0, j0, k0;
  brick_of(v, blockIdx.x, &vol, &i0, &j0, &k0);
  if (SRC_X)
    fill_tile_x(tile, v, vol, i0, j0, k0, src, W1b);
  else
    fill_tile_h(tile, v, vol, i0, j0, k0, src);
  __syncthreads();
  f32x4 acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  conv27(tile, W2b, acc);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  const float* b2 = P + OFF_B2;
  const int ib = i0 - v.m, jb = j0 - v.m, kb = k0 - v.m;  // inner coordinates of the brick
  if (TRAIN) {
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int m = 4 * g + q, ii = ib + w, jj = jb + 2 * mt + (m >> 3), kk = kb + (m & 7);
        if (ii >= v.I || jj >= v.J || kk >= v.K) continue;
        float* z = out + ((((long)vol * v.I + ii) * v.J + jj) * v.K + kk) * CH;
#pragma unroll
        for (int nt = 0; nt < 4; nt++) z[nt * 16 + r] = acc[mt][nt][q] + b2[nt * 16 + r];
      }
    return;
  }
  __syncthreads();  // every wave is done with the tile
  float* aT = reinterpret_cast<float*>(tile);  // [256][AS]
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int o = w * 64 + mt * 16 + 4 * g + q, co = nt * 16 + r;
        aT[o * AS + co] = gelu(acc[mt][nt][q] + b2[co]);
      }
  __syncthreads();
  const int o = threadIdx.x;
  int ii, jj, kk;
  cell_of(o, &ii, &jj, &kk);
  ii += ib;
  jj += jb;
  kk += kb;
  if (ii >= v.I || jj >= v.J || kk >= v.K) return;
  const float* W3 = P + OFF_W3;
  const float* b3 = P + OFF_B3;
  float* y = out + ((((long)vol * v.I + ii) * v.J + jj) * v.K + kk) * COUT;
#pragma unroll
  for (int c = 0; c < COUT; c++) {
    float s = b3[c];
    for (int co = 0; co < CH; co++) s += W3[c * CH + co] * aT[o * AS + co];
    y[c] = s;
  }
}

// h1 = bf16(W1 x') on every cell: 256 cells (16 row tiles) per block
__global__ void __launch_bounds__(256) k_l1(Vol v, const unsigned short* __restrict__ x,
                                            const unsigned short* __restrict__ W1b, unsigned short* __restrict__ h1) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  const long ncell = v.cells();
  s16x4 bw[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) bw[nt] = *reinterpret_cast<const s16x4*>(W1b + (nt * 16 + r) * CIN + 4 * g);
  for (int q = 0; q < 4; q++) {
    const long c0 = ((long)blockIdx.x * 16 + w * 4 + q) * 16;
    s16x4 a = {0, 0, 0, 0};
    if (c0 + r < ncell) a = *reinterpret_cast<const s16x4*>(x + (c0 + r) * CIN + 4 * g);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const f32x4 d = mma16(a, bw[nt], f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const long cc = c0 + 4 * g + e;
        if (cc < ncell) h1[cc * CH + nt * 16 + r] = f2bf(d[e]);
      }
    }
  }
}

// loss head, one inner cell per thread: y, dy = inv_norm (y - target),
// dz2 = (W3^T dy) gelu'(z2) into the dz2 volume; block partials of dW3, db2,
// db3 and the sum of squared errors
__global__ void __launch_bounds__(64) k_head(Vol v, const float* __restrict__ z2, const float* __restrict__ target,
                                             const float* __restrict__ P, float inv_norm,
                                             unsigned short* __restrict__ dz2, float* __restrict__ part) {
  __shared__ float aT[64][AS], dzT[64][AS], dyT[64][COUT + 1], lossT[64];
  const int t = threadIdx.x;
  const long cell = (long)blockIdx.x * 64 + t;
  const float* W3 = P + OFF_W3;
  const float* b3 = P + OFF_B3;
  float dy[COUT] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float ls = 0.f;
  if (cell < v.inner_cells()) {
    const float* z = z2 + cell * CH;
    for (int co = 0; co < CH; co++) aT[t][co] = gelu(z[co]);
    for (int c = 0; c < COUT; c++) {
      float y = b3[c];
      for (int co = 0; co < CH; co++) y += W3[c * CH + co] * aT[t][co];
      const float e = y - target[cell * COUT + c];
      dy[c] = inv_norm * e;
      ls += e * e;
    }
    long rest = cell;
    const int kk = (int)(rest % v.K);
    rest /= v.K;
    const int jj = (int)(rest % v.J);
    rest /= v.J;
    const int ii = (int)(rest % v.I);
    const int vol = (int)(rest / v.I);
    unsigned int* dst = reinterpret_cast<unsigned int*>(dz2 + vcell(v, vol, ii + v.m, jj + v.m, kk + v.m) * CH);
    for (int co = 0; co < CH; co += 2) {
      float d2[2];
      for (int e = 0; e < 2; e++) {
        float da = 0.f;
        for (int c = 0; c < COUT; c++) da += W3[c * CH + co + e] * dy[c];
        d2[e] = da * gelu_grad(z[co + e]);
        dzT[t][co + e] = d2[e];
      }
      dst[co >> 1] = (unsigned)f2bf(d2[0]) | ((unsigned)f2bf(d2[1]) << 16);
    }
  } else {
    for (int co = 0; co < CH; co++) {
      aT[t][co] = 0.f;
      dzT[t][co] = 0.f;
    }
  }
  for (int c = 0; c < COUT; c++) dyT[t][c] = dy[c];
  lossT[t] = ls;
  __syncthreads();
  float* pp = part + (long)blockIdx.x * HEAD_PART;
  for (int c = 0; c < COUT; c++) {
    float s = 0.f;
    for (int u = 0; u < 64; u++) s += dyT[u][c] * aT[u][t];
    pp[c * CH + t] = s;
  }
  float s = 0.f;
  for (int u = 0; u < 64; u++) s += dzT[u][t];
  pp[NW3 + t] = s;
  if (t < COUT) {
    float sb = 0.f;
    for (int u = 0; u < 64; u++) sb += dyT[u][t];
    pp[NW3 + NB2 + t] = sb;
  }
  if (t == COUT) {
    float sl = 0.f;
    for (int u = 0; u < 64; u++) sl += lossT[u];
    pp[NW3 + NB2 + NB3] = sl;
  }
}

// dh1 on the bricks of the whole volume (v.m = 0): the transposed conv of
// dz2 with W2T[t][ci][co] = W2[26 - t][co][ci]; per-brick partials of
// dW1 = sum over cells of dh1 x'^T
__global__ void __launch_bounds__(256) k_bwd_data(Vol v, const unsigned short* __restrict__ dz2,
                                                  const unsigned short* __restrict__ W2Tb,
                                                  const unsigned short* __restrict__ x, float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) unsigned short tile[HN * CH];
  int vol, i0, j0, k0;
  brick_of(v, blockIdx.x, &vol, &i0, &j0, &k0);
  fill_tile_h(tile, v, vol, i0, j0, k0, dz2);
  __syncthreads();
  f32x4 acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  conv27(tile, W2Tb, acc);
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  float* dT = reinterpret_cast<float*>(tile);                         // [256][AS]
  unsigned short* xT = reinterpret_cast<unsigned short*>(dT + BCELLS * AS);  // [256][CIN] bf16
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int q = 0; q < 4; q++) dT[(w * 64 + mt * 16 + 4 * g + q) * AS + nt * 16 + r] = acc[mt][nt][q];
  {
    const int o = threadIdx.x;
    int ii, jj, kk;
    cell_of(o, &ii, &jj, &kk);
    const int i = i0 + ii, j = j0 + jj, k = k0 + kk;
    const bool in = inside(v, i, j, k);
    const unsigned short* xs = x + (in ? vcell(v, vol, i, j, k) * CIN : 0);
    for (int f = 0; f < CIN; f++) xT[o * CIN + f] = in ? xs[f] : (unsigned short)0;
  }
  __syncthreads();
  const int co = threadIdx.x & 63, f0 = (threadIdx.x >> 6) * 4;
  float* pp = part + (long)blockIdx.x * NW1;
  for (int q = 0; q < 4; q++) {
    float s = 0.f;
    for (int o = 0; o < BCELLS; o++) s += dT[o * AS + co] * bf2f(xT[o * CIN + f0 + q]);
    pp[co * CIN + f0 + q] = s;
  }
}

// dW2[t][co][ci] = sum over cells of dz2(p)[co] h1(p + o_t)[ci]; block
// (split s, tap t) walks its share of the inner bricks with both operands
// transposed in LDS (cells along K), then writes its partial
__global__ void __launch_bounds__(256) k_bwd_weight(Vol v, const unsigned short* __restrict__ dz2,
                                                    const unsigned short* __restrict__ h1, int splits,
                                                    float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) unsigned short T1[CH * TS];
  __shared__ __attribute__((aligned(16))) unsigned short T2[CH * TS];
  const int t = blockIdx.x % 27, s = blockIdx.x / 27;
  const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
  const long nb = v.bricks(), b0 = nb * s / splits, b1 = nb * (s + 1) / splits;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  f32x4 acc[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long b = b0; b < b1; b++) {
    int vol, i0, j0, k0;
    brick_of(v, b, &vol, &i0, &j0, &k0);
    __syncthreads();  // the previous brick's operands have been read
    {
      const int o = threadIdx.x;
      int ii, jj, kk;
      cell_of(o, &ii, &jj, &kk);
      const int i = i0 + ii, j = j0 + jj, k = k0 + kk;
      const bool in_d = inside(v, i, j, k), in_h = inside(v, i + di, j + dj, k + dk);
      const u32x4* pd = reinterpret_cast<const u32x4*>(dz2 + (in_d ? vcell(v, vol, i, j, k) * CH : 0));
      const u32x4* ph = reinterpret_cast<const u32x4*>(h1 + (in_h ? vcell(v, vol, i + di, j + dj, k + dk) * CH : 0));
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const u32x4 dv = in_d ? pd[c] : u32x4{0u, 0u, 0u, 0u};
        const u32x4 hv = in_h ? ph[c] : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
        for (int e = 0; e < 8; e++) {
          T1[(c * 8 + e) * TS + o] = (unsigned short)(dv[e >> 1] >> (16 * (e & 1)));
          T2[(c * 8 + e) * TS + o] = (unsigned short)(hv[e >> 1] >> (16 * (e & 1)));
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < BCELLS / 32; c++) {
      const bf16x8 av = ld8(T1 + (16 * w + r) * TS + 32 * c + 8 * g);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[nt] = mma32(av, ld8(T2 + (16 * nt + r) * TS + 32 * c + 8 * g), acc[nt]);
    }
  }
  float* pp = part + ((long)s * 27 + t) * (CH * CH);
#pragma unroll
  for (int nt = 0; nt < 4; nt++)
#pragma unroll
    for (int q = 0; q < 4; q++) pp[(16 * w + 4 * g + q) * CH + 16 * nt + r] = acc[nt][q];
}

// the gradient in parameter order (W1, W2, b2, W3, b3) and the summed squared
// error at [NPARAM]; partials added in index order
__global__ void k_reduce(const float* __restrict__ hp, int hb, const float* __restrict__ dp, int db,
                         const float* __restrict__ wp, int splits, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NPARAM) return;
  float s = 0.f;
  if (p < OFF_W2) {
    for (int b = 0; b < db; b++) s += dp[(long)b * NW1 + p];
  } else if (p < OFF_B2) {
    for (int b = 0; b < splits; b++) s += wp[(long)b * NW2 + (p - OFF_W2)];
  } else if (p < OFF_W3) {
    for (int b = 0; b < hb; b++) s += hp[(long)b * HEAD_PART + NW3 + (p - OFF_B2)];
  } else if (p < OFF_B3) {
    for (int b = 0; b < hb; b++) s += hp[(long)b * HEAD_PART + (p - OFF_W3)];
  } else if (p < NPARAM) {
    for (int b = 0; b < hb; b++) s += hp[(long)b * HEAD_PART + NW3 + NB2 + (p - OFF_B3)];
  } else {
    for (int b = 0; b < hb; b++) s += hp[(long)b * HEAD_PART + NW3 + NB2 + NB3];
  }
  grad[p] = s;
}

__device__ __forceinline__ void bf16_copy(int p, float val, unsigned short* W1b, unsigned short* W2b,
                                          unsigned short* W2Tb) {
  if (p < OFF_W2) {
    W1b[p] = f2bf(val);
  } else if (p < OFF_B2) {
    const int q = p - OFF_W2, t = q / (CH * CH), co = (q / CH) % CH, ci = q % CH;
    const unsigned short b = f2bf(val);
    W2b[q] = b;
    W2Tb[((26 - t) * CH + ci) * CH + co] = b;
  }
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                       float* __restrict__ V, float b1c, float b2c, float lr, unsigned short* __restrict__ W1b,
                       unsigned short* __restrict__ W2b, unsigned short* __restrict__ W2Tb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float g = G[p];
  const float m = 0.9f * M[p] + 0.1f * g;
  const float v = 0.999f * V[p] + 0.001f * g * g;
  M[p] = m;
  V[p] = v;
  const float val = P[p] - lr * (m / b1c) / (sqrtf(v / b2c) + 1e-8f);
  P[p] = val;
  bf16_copy(p, val, W1b, W2b, W2Tb);
}

__global__ void k_sync_bf16(const float* __restrict__ P, unsigned short* __restrict__ W1b,
                            unsigned short* __restrict__ W2b, unsigned short* __restrict__ W2Tb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < OFF_B2) bf16_copy(p, P[p], W1b, W2b, W2Tb);
}

__global__ void k_to_bf16(const float* __restrict__ src, unsigned short* __restrict__ dst, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f2bf(src[i]);
}

// ---------------------------------------------------------------------------
// solver-side kernels: crops for training, features and the guess
// ---------------------------------------------------------------------------
// crop cell c of crop blockIdx.y from levels L0 = U^{n-2}, L1 = U^{n-1},
// L2 = U^n, L3 = U^{n+1} (U = Q / J): Q^n, d1, d2, Q^{n+1} - Q^n, mask
__global__ void k_capture(const double* __restrict__ L0, const double* __restrict__ L1,
                          const double* __restrict__ L2, const double* __restrict__ L3,
                          const double* __restrict__ prof, cf::Geo g, const int* __restrict__ origins, int VC,
                          float* __restrict__ D) {
  const long n3 = (long)VC * VC * VC;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n3) return;
  const int crop = blockIdx.y;
  const int ci = origins[3 * crop] + (int)(c / (VC * VC)), cj = origins[3 * crop + 1] + (int)((c / VC) % VC),
            ck = origins[3 * crop + 2] + (int)(c % VC);
  float* d = D + ((long)crop * n3 + c) * NDATA;
  const int X = g.gend - 2, Y = g.ny - 1, Z = g.nz - 1;
  if (ci < 0 || ci >= X || cj < 0 || cj >= Y || ck < 0 || ck >= Z) {
    for (int q = 0; q < NDATA; q++) d[q] = 0.f;
    return;
  }
  const long a = cf::at(g, ci + 3, cj + 2, ck + 2);
  const double J = cf::pf(prof, g.Ym, cf::SET_C, 0, cj + 2);
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const long o = v * g.nf + a;
    const double q0 = L0[o] * J, q1 = L1[o] * J, q2 = L2[o] * J, q3 = L3[o] * J;
    d[v] = (float)q2;
    d[5 + v] = (float)(q2 - q1);
    d[10 + v] = (float)(q1 - q0);
    d[15 + v] = (float)(q3 - q2);
  }
  d[20] = 1.f;
  d[21] = d[22] = d[23] = 0.f;
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_pack(const float* __restrict__ D, int VC, long n_avail, int first, unsigned long long seed,
                       int step, Norm nm, unsigned short* __restrict__ x, float* __restrict__ target) {
  const long n3 = (long)VC * VC * VC;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n3) return;
  const int b = blockIdx.y;
  const long crop = step < 0 ? (long)first + b
                             : (long)(splitmix64(seed ^ splitmix64(((unsigned long long)step << 20) + b)) %
                                      (unsigned long long)n_avail);
  const float* d = D + (crop * n3 + c) * NDATA;
  unsigned int* xo = reinterpret_cast<unsigned int*>(x + ((long)b * n3 + c) * CIN);
  float f[CIN];
  const bool in = d[20] != 0.f;
#pragma unroll
  for (int v = 0; v < 5; v++) {
    f[v] = in ? (d[v] - nm.mu[v]) / nm.sig[v] : 0.f;
    f[5 + v] = in ? d[5 + v] / nm.s[v] : 0.f;
    f[10 + v] = in ? d[10 + v] / nm.s[v] : 0.f;
  }
  f[15] = in ? 1.f : 0.f;
#pragma unroll
  for (int q = 0; q < CIN / 2; q++) xo[q] = (unsigned)f2bf(f[2 * q]) | ((unsigned)f2bf(f[2 * q + 1]) << 16);
  const int li = (int)(c / (VC * VC)), lj = (int)((c / VC) % VC), lk = (int)(c % VC), VI = VC - 2;
  if (li < 1 || li > VI || lj < 1 || lj > VI || lk < 1 || lk > VI) return;
  float* tg = target + ((((long)b * VI + li - 1) * VI + lj - 1) * VI + lk - 1) * COUT;
#pragma unroll
  for (int v = 0; v < 5; v++) tg[v] = (d[15 + v] - (nm.c1 * d[5 + v] + nm.c2 * d[10 + v])) / nm.sr[v];
}

__global__ void k_feat(const double* __restrict__ Un, const double* __restrict__ Uq, const double* __restrict__ Uqq,
                       const double* __restrict__ prof, cf::Geo g, Norm nm, unsigned short* __restrict__ x) {
  const int Y = g.ny - 1, Z = g.nz - 1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
  if (k >= Z) return;
  const long a = cf::at(g, i + 3, j + 2, k + 2);
  const double J = cf::pf(prof, g.Ym, cf::SET_C, 0, j + 2);
  float f[CIN];
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const long o = v * g.nf + a;
    const double q0 = Un[o] * J, q1 = Uq[o] * J, q2 = Uqq[o] * J;
    f[v] = ((float)q0 - nm.mu[v]) / nm.sig[v];
    f[5 + v] = (float)(q0 - q1) / nm.s[v];
    f[10 + v] = (float)(q1 - q2) / nm.s[v];
  }
  f[15] = 1.f;
  unsigned int* xo = reinterpret_cast<unsigned int*>(x + (((long)i * Y + j) * Z + k) * CIN);
#pragma unroll
  for (int q = 0; q < CIN / 2; q++) xo[q] = (unsigned)f2bf(f[2 * q]) | ((unsigned)f2bf(f[2 * q + 1]) << 16);
}

__global__ void k_apply(double* __restrict__ Ui, const double* __restrict__ Un, const double* __restrict__ Uq,
                        const double* __restrict__ Uqq, const double* __restrict__ prof, cf::Geo g, Norm nm,
                        const float* __restrict__ y) {
  const int Y = g.ny - 1, Z = g.nz - 1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
  if (k >= Z) return;
  const long a = cf::at(g, i + 3, j + 2, k + 2);
  const double J = cf::pf(prof, g.Ym, cf::SET_C, 0, j + 2);
  const float* yc = y + (((long)i * Y + j) * Z + k) * COUT;
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const long o = v * g.nf + a;
    const double q0 = Un[o] * J, q1 = Uq[o] * J, q2 = Uqq[o] * J;
    const double dq = (double)nm.c1 * (q0 - q1) + (double)nm.c2 * (q1 - q2) + (double)nm.sr[v] * (double)yc[v];
    Ui[o] = (q0 + dq) / J;
  }
}

__global__ void __launch_bounds__(256) k_stats(const float* __restrict__ D, long cells, double* __restrict__ part) {
  __shared__ double red[256];
  double acc[NSTAT];
  for (int q = 0; q < NSTAT; q++) acc[q] = 0.0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long)gridDim.x * blockDim.x) {
    const float* d = D + c * NDATA;
    if (d[20] == 0.f) continue;
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const double q = d[v], d1 = d[5 + v], d2 = d[10 + v], tg = d[15 + v];
      acc[v] += q;
      acc[5 + v] += q * q;
      acc[10 + v] += d1 * d1;
      acc[15 + v] += d2 * d2;
      acc[20 + v] += tg * tg;
      acc[25 + v] += (tg - d1) * (tg - d1);
      acc[30 + v] += (tg - 2 * d1 + d2) * (tg - 2 * d1 + d2);
    }
    acc[35] += 1.0;
  }
  for (int q = 0; q < NSTAT; q++) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.x * NSTAT + q] = red[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
Vol make_vol(int n, int VI, int VJ, int VK, int m) {
  Vol v;
  v.n = n;
  v.VI = VI;
  v.VJ = VJ;
  v.VK = VK;
  v.m = m;
  v.I = VI - 2 * m;
  v.J = VJ - 2 * m;
  v.K = VK - 2 * m;
  v.nbi = (v.I + BI - 1) / BI;
  v.nbj = (v.J + BJ - 1) / BJ;
  v.nbk = (v.K + BK - 1) / BK;
  return v;
}

void launch_fwd(hipStream_t st, const Vol& v, const unsigned short* src, bool src_is_x, const unsigned short* W1b,
                const unsigned short* W2b, const float* P, float* out, bool train) {
  const dim3 grid((unsigned)v.bricks()), blk(256);
  if (src_is_x && !train)
    hipLaunchKernelGGL((k_fwd<true, false>), grid, blk, 0, st, v, src, W1b, W2b, P, out);
  else if (src_is_x)
    hipLaunchKernelGGL((k_fwd<true, true>), grid, blk, 0, st, v, src, W1b, W2b, P, out);
  else if (!train)
    hipLaunchKernelGGL((k_fwd<false, false>), grid, blk, 0, st, v, src, W1b, W2b, P, out);
  else
    hipLaunchKernelGGL((k_fwd<false, true>), grid, blk, 0, st, v, src, W1b, W2b, P, out);
}

void launch_l1(hipStream_t st, const Vol& v, const unsigned short* x, const unsigned short* W1b, unsigned short* h1) {
  hipLaunchKernelGGL(k_l1, dim3((unsigned)((v.cells() + 255) / 256)), dim3(256), 0, st, v, x, W1b, h1);
}

int launch_head(hipStream_t st, const Vol& v, const float* z2, const float* target, const float* P, float inv_norm,
                unsigned short* dz2, float* part) {
  const int nb = (int)((v.inner_cells() + 63) / 64);
  hipLaunchKernelGGL(k_head, dim3(nb), dim3(64), 0, st, v, z2, target, P, inv_norm, dz2, part);
  return nb;
}

int launch_bwd_data(hipStream_t st, const Vol& v, const unsigned short* dz2, const unsigned short* W2Tb,
                    const unsigned short* x, float* part) {
  const Vol full = make_vol(v.n, v.VI, v.VJ, v.VK, 0);
  const int nb = (int)full.bricks();
  hipLaunchKernelGGL(k_bwd_data, dim3(nb), dim3(256), 0, st, full, dz2, W2Tb, x, part);
  return nb;
}

void launch_bwd_weight(hipStream_t st, const Vol& v, const unsigned short* dz2, const unsigned short* h1, int splits,
                       float* part) {
  hipLaunchKernelGGL(k_bwd_weight, dim3(27 * splits), dim3(256), 0, st, v, dz2, h1, splits, part);
}

void launch_reduce(hipStream_t st, const float* head_part, int head_blocks, const float* data_part, int data_blocks,
                   const float* w_part, int splits, float* grad) {
  hipLaunchKernelGGL(k_reduce, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, st, head_part, head_blocks, data_part,
                     data_blocks, w_part, splits, grad);
}

void launch_adam(hipStream_t st, float* P, const float* G, float* M, float* V, int step, float lr,
                 unsigned short* W1b, unsigned short* W2b, unsigned short* W2Tb) {
  const float b1c = 1.0f - powf(0.9f, (float)step), b2c = 1.0f - powf(0.999f, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, st, P, G, M, V, b1c, b2c, lr, W1b, W2b, W2Tb);
}

void launch_sync_bf16(hipStream_t st, const float* P, unsigned short* W1b, unsigned short* W2b, unsigned short* W2Tb) {
  hipLaunchKernelGGL(k_sync_bf16, dim3((OFF_B2 + 255) / 256), dim3(256), 0, st, P, W1b, W2b, W2Tb);
}

void launch_to_bf16(hipStream_t st, const float* src, unsigned short* dst, long n) {
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
}

void launch_capture(hipStream_t st, const double* L0, const double* L1, const double* L2, const double* L3,
                    const double* prof, const cf::Geo& g, const int* origins, int ncrops, int VC, float* D) {
  const long n3 = (long)VC * VC * VC;
  hipLaunchKernelGGL(k_capture, dim3((unsigned)((n3 + 255) / 256), ncrops), dim3(256), 0, st, L0, L1, L2, L3, prof, g,
                     origins, VC, D);
}

void launch_pack(hipStream_t st, const float* D, int VC, long n_avail, int first, int nb, unsigned long long seed,
                 int step, const Norm& nm, unsigned short* x, float* target) {
  const long n3 = (long)VC * VC * VC;
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((n3 + 255) / 256), nb), dim3(256), 0, st, D, VC, n_avail, first, seed,
                     step, nm, x, target);
}

void launch_feat(hipStream_t st, const double* Un, const double* Uq, const double* Uqq, const double* prof,
                 const cf::Geo& g, const Norm& nm, unsigned short* x) {
  const int X = g.gend - 2, Y = g.ny - 1, Z = g.nz - 1;
  hipLaunchKernelGGL(k_feat, dim3((Z + 63) / 64, Y, X), dim3(64), 0, st, Un, Uq, Uqq, prof, g, nm, x);
}

void launch_apply(hipStream_t st, double* Ui, const double* Un, const double* Uq, const double* Uqq,
                  const double* prof, const cf::Geo& g, const Norm& nm, const float* y) {
  const int X = g.gend - 2, Y = g.ny - 1, Z = g.nz - 1;
  hipLaunchKernelGGL(k_apply, dim3((Z + 63) / 64, Y, X), dim3(64), 0, st, Ui, Un, Uq, Uqq, prof, g, nm, y);
}

int launch_stats(hipStream_t st, const float* D, long cells, double* part) {
  const int nb = 256;
  hipLaunchKernelGGL(k_stats, dim3(nb), dim3(256), 0, st, D, cells, part);
  return nb;
}

}  // namespace cfp
