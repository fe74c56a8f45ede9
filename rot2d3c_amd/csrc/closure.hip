// closure.hip -- gfx950 kernels of the learned subgrid closure: a 6-layer
// periodic 3x3 CNN (32 channels, GELU) with BF16 weights and activations and
// FP32 accumulation on the matrix cores (gfx950's v_mfma_f32_16x16x32_bf16
// for the 32-channel inputs, v_mfma_f32_16x16x16_bf16 for the 16-channel
// ones), its backward pass, an MSE loss and Adam.  No reference counterpart.
//
// Layouts (all per sample, pixels row-major, channels fastest, "NHWC"):
//   activations  bf16 [B][H][W][C], C = 16 (input, 3 used) or 32
//   pre-activations Z  fp32 [B][H][W][C] (kept for GELU' in the backward pass)
//   weights      fp32 master W[9][CO][CI] + b[CO] per layer, flat; tap t =
//                3 (dy+1) + (dx+1) reads X[y+dy][x+dx] (periodic)
//   bf16 copies  Wf[t][co][ci] (forward) and Wb[t][ci][co] = W[8-t][co][ci]
//                (the backward data pass is the forward conv of dZ with Wb)
// Layer shapes (CI, CO): (16, 32) (32, 32) x4 (32, 16); the input's channels
// 3..15 and the output's channels 2..15 are zero padding (their weights stay
// zero: their gradients are zero).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/rot2d3c.h"
#include "fft_core.h"

int r2d3c_internal_set_error(int code, const char *msg);

namespace {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

constexpr int kLayers = 6;
constexpr int kCI[kLayers] = {16, 32, 32, 32, 32, 32};
constexpr int kCO[kLayers] = {32, 32, 32, 32, 32, 16};

__host__ __device__ constexpr int wsize(int l) { return 9 * kCI[l] * kCO[l]; }
__host__ __device__ constexpr int psize(int l) { return wsize(l) + kCO[l]; }
__host__ __device__ constexpr int poff(int l) { return l == 0 ? 0 : poff(l - 1) + psize(l - 1); }
constexpr int kParams = poff(kLayers);
__host__ __device__ constexpr int woff(int l) { return l == 0 ? 0 : woff(l - 1) + wsize(l - 1); }
constexpr int kWeights = woff(kLayers);

// float -> bf16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((unsigned)h << 16); }

// GELU(z) = z/2 (1 + erf(z/sqrt 2)) and its derivative
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { CONV_GELU = 0, CONV_LINEAR = 1, CONV_BWD = 2 };

// A block of kConvWaves waves computes kConvWaves * kTilesPerWave tiles of 16
// consecutive pixels of one row for all CO channels: D[p][co] = sum_t sum_ci X[p + d(t)][ci]
// Wt[t][co][ci], K = 9 CI as 9*CI/16 MFMA steps.  The block first stages the
// layer's weights in LDS (rows padded by 4 to spread the banks); each weight
// fragment is then read once per wave and used for its kTilesPerWave tiles.  Operand
// layout of v_mfma_f32_16x16x16_bf16 (lane l, m = l%16, q = l/16):
// A[m][4q..4q+3] = pixel m's channels, B[4q..4q+3][m] = weights of output
// channel m; D[4q+i][m] = pixel 4q+i, channel m.
//   CONV_GELU / CONV_LINEAR: Z = D + bias (stored when Zout), Y = bf16(gelu(Z) or Z)
//   CONV_BWD: Y = bf16(D * gelu'(Zprev))  (the next lower layer's dZ)
#ifndef R2D_CONV_TPW
#define R2D_CONV_TPW 1
#endif
#ifndef R2D_CONV_WPB
#define R2D_CONV_WPB 16
#endif
constexpr int kTilesPerWave = R2D_CONV_TPW, kConvWaves = R2D_CONV_WPB;

// Epilogue of one tile of 16 pixels (img, y, x0 .. x0+15): lane (m, q) holds
// pixels 4q..4q+3 of channel 16n + m in acc[n].
//   CONV_GELU / CONV_LINEAR: Z = D + bias (stored when Zout), Y = bf16(gelu(Z) or Z)
//   CONV_BWD: Y = bf16(D * gelu'(Zprev))
template <int CO, int MODE>
__device__ __forceinline__ void conv_store(const f32x4 (&acc)[CO / 16], int img, int y, int x0,
                                           int m, int q, int H, int W,
                                           const float *__restrict__ bias,
                                           const float *__restrict__ Zprev,
                                           float *__restrict__ Zout, bf16 *__restrict__ Y,
                                           float clamp, const bf16 *__restrict__ X0, float keep) {
#pragma unroll
  for (int n = 0; n < CO / 16; ++n) {
    const int co = 16 * n + m;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const size_t o = ((size_t)((size_t)img * H + y) * W + x0 + 4 * q + r) * CO + co;
      const float v = acc[n][r];
      if constexpr (MODE == CONV_BWD) {
        Y[o] = f2bf(v * gelu_grad(Zprev[o]));
      } else {
        float z = v + bias[co];
        if constexpr (MODE == CONV_LINEAR) {
          if (clamp > 0.f) z = fminf(fmaxf(z, -clamp), clamp);
          // sign clipping: where output channel co < 2 shares the sign of
          // input channel co (X0 is the network input, 16 channels) it is
          // scaled by keep = 1 - strength
          if (X0 && co < 2 && z * bf2f(X0[o / CO * 16 + co]) > 0.f) z *= keep;
        }
        if (Zout) Zout[o] = z;
        Y[o] = f2bf(MODE == CONV_GELU ? gelu(z) : z);
      }
    }
  }
}

template <int CI, int CO, int MODE>
__global__ __launch_bounds__(64 * kConvWaves) void k_conv3x3(const bf16 *__restrict__ X,
                                                  const bf16 *__restrict__ Wt,
                                                  const float *__restrict__ bias,
                                                  const float *__restrict__ Zprev,
                                                  float *__restrict__ Zout, bf16 *__restrict__ Y,
                                                  int B, int H, int W, float clamp,
                                                  const bf16 *__restrict__ X0, float keep) {
  // weight rows padded by 4 (16x16x16 path) or 8 (16x16x32 path: keeps the
  // 16-byte operand reads aligned) bf16 to spread the LDS banks
  constexpr int CI16 = CI / 16, CO16 = CO / 16, WS = CI + (CI % 32 == 0 ? 8 : 4),
                TW = kTilesPerWave;
  __shared__ bf16 ws[9 * CO * WS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, q = lane >> 4;
  const int tpr = W / 16;
  const long ntiles = (long)B * H * tpr;
  int tb[TW], ty[TW], tx[TW];
  bool ok[TW];
#pragma unroll
  for (int i = 0; i < TW; ++i) {
    const long tile = ((long)blockIdx.x * kConvWaves + wave) * TW + i;
    ok[i] = tile < ntiles;
    const long tt = ok[i] ? tile : 0;
    tb[i] = (int)(tt / ((long)H * tpr));
    const int rem = (int)(tt - (long)tb[i] * H * tpr);
    ty[i] = rem / tpr;
    tx[i] = (rem - ty[i] * tpr) * 16;
  }
  // The input fragments of all nine taps are loaded before the weights are
  // staged, so the two round trips to memory overlap.  16x16x32 path: lane l
  // holds 8 consecutive channels 8(l/16)..8(l/16)+7 of pixel l%16 (A) and the
  // same channels of the weights of output channel l%16 (B); A and B use one
  // channel order, so the reduction over K is the layer's channel sum.
  constexpr bool K32 = CI % 32 == 0;
  constexpr int NC = K32 ? CI / 32 : CI16;
  using frag = typename std::conditional<K32, s16x8, s16x4>::type;
  frag a[9][NC][TW];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int dy = t / 3 - 1, dx = t % 3 - 1;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int i = 0; i < TW; ++i) {
        int yy = ty[i] + dy, xx = tx[i] + m + dx;
        yy += yy < 0 ? H : (yy >= H ? -H : 0);
        xx += xx < 0 ? W : (xx >= W ? -W : 0);
        const bf16 *src = X + ((size_t)((size_t)tb[i] * H + yy) * W + xx) * CI +
                          (K32 ? 32 * c + 8 * q : 16 * c + 4 * q);
        a[t][c][i] = ok[i] ? *(const frag *)src : frag{};
      }
  }
  for (int e = threadIdx.x; e < 9 * CO * (CI / 4); e += 64 * kConvWaves) {
    const int row = e / (CI / 4), c4 = e - row * (CI / 4);
    *(uint2 *)(ws + row * WS + 4 * c4) = *(const uint2 *)(Wt + (size_t)row * CI + 4 * c4);
  }
  __syncthreads();
  f32x4 acc[TW][CO16];
#pragma unroll
  for (int i = 0; i < TW; ++i)
#pragma unroll
    for (int n = 0; n < CO16; ++n) acc[i][n] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 9; ++t) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int n = 0; n < CO16; ++n) {
        if constexpr (K32) {
          const s16x8 w = *(const s16x8 *)(ws + (t * CO + 16 * n + m) * WS + 32 * c + 8 * q);
#pragma unroll
          for (int i = 0; i < TW; ++i)
            acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                __builtin_bit_cast(bf16x8, a[t][c][i]), __builtin_bit_cast(bf16x8, w), acc[i][n],
                0, 0, 0);
        } else {
          const s16x4 w = *(const s16x4 *)(ws + (t * CO + 16 * n + m) * WS + 16 * c + 4 * q);
#pragma unroll
          for (int i = 0; i < TW; ++i)
            acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t][c][i], w, acc[i][n], 0, 0,
                                                                   0);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < TW; ++i)
    if (ok[i])
      conv_store<CO, MODE>(acc[i], tb[i], ty[i], tx[i], m, q, H, W, bias, Zprev, Zout, Y, clamp,
                           X0, keep);
}

// The first layer with the network's input pass folded in (a closure run's
// steady steps): the block of image row y forms the input rows y-1, y, y+1
// itself from the column transforms of omega, vz and psi in a step
// workspace (tiled mixed fields, pitch kp, columns kx < nc) into an LDS
// tile, with the row transforms and roundings of the input pass
// (k_ci_row: omega + i vz of one row per transform, psi of the row pairs
// (2m, 2m+1)), then runs the layer on it as k_conv3x3<16, 32, GELU> does on
// the input image; it also writes row y of the input image.  Each transform
// is one group of n/8 lanes inside a wavefront, ordered within its wave, so
// the other waves meet it only at the one block barrier.
template <int N>
__global__ __launch_bounds__(64 * kConvWaves) void k_conv0_mixed(
    const r2d::cplx *__restrict__ mw, const r2d::cplx *__restrict__ mv,
    const r2d::cplx *__restrict__ mp, int kp, int nc, const r2d::cplx *__restrict__ tw,
    double s0, double s1, double s2, bf16 *__restrict__ X0, const bf16 *__restrict__ Wt,
    const float *__restrict__ bias, float *__restrict__ Zout, bf16 *__restrict__ Y) {
  constexpr int CI = 16, CO = 32, CO16 = CO / 16, WS = CI + 4, T = N / 8,
                LN = r2d::Shape<N>::LDSN;
  static_assert(N == 16 * kConvWaves && kTilesPerWave == 1, "one image row per block");
  static_assert(6 * T <= 64 * kConvWaves && 64 % T == 0, "transform groups inside waves");
  __shared__ bf16 ws[9 * CO * WS];
  __shared__ bf16 xt[3 * N * CI];   // input rows y-1, y, y+1: [r][x][channel]
  __shared__ r2d::cplx fl[5 * LN];  // the five transforms' LDS areas
  const int y = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, q = lane >> 4;
  for (int e = threadIdx.x; e < 9 * CO * (CI / 4); e += 64 * kConvWaves) {
    const int row = e / (CI / 4), c4 = e - row * (CI / 4);
    *(uint2 *)(ws + row * WS + 4 * c4) = *(const uint2 *)(Wt + (size_t)row * CI + 4 * c4);
  }
  // channels 3..15 of the tile are zero
  for (int e = threadIdx.x; e < 3 * N; e += 64 * kConvWaves) {
    bf16 *px = xt + e * CI;
    px[3] = 0;
    *(uint2 *)(px + 4) = make_uint2(0u, 0u);
    *(uint4 *)(px + 8) = make_uint4(0u, 0u, 0u, 0u);
  }
  const int g = threadIdx.x / T, b = threadIdx.x - g * T;
  if (g < 5) {
    // groups 0-2: omega + i vz of rows y-1, y, y+1; groups 3, 4: psi of the
    // row pairs holding y-1 and y+1
    const int ym = (y + N - 1) % N, yp = (y + 1) % N;
    const int ra = g < 3 ? (g == 0 ? ym : (g == 1 ? y : yp)) : ((g == 3 ? ym : yp) & ~1);
    const int rb = g < 3 ? ra : ra + 1;
    const r2d::cplx *fa = g < 3 ? mw : mp, *fb = g < 3 ? mv : mp;
    r2d::cplx v[8];
#pragma unroll
    for (int k8 = 0; k8 < 8; ++k8) {
      const int k = b + T * k8;
      const bool mir = k > N / 2;
      const int kk = mir ? N - k : k;
      const bool in = kk < nc;
      v[k8] = r2d::packed_z<N>(k, in ? fa[r2d::tix(ra, kk, kp)] : r2d::mk(0.0, 0.0),
                               in ? fb[r2d::tix(rb, kk, kp)] : r2d::mk(0.0, 0.0), mir);
    }
    r2d::fft_canon<N, +1, 1, 1, 0, false, true, true>(v, fl + g * LN, b, tw);
#pragma unroll
    for (int k8 = 0; k8 < 8; ++k8) {
      const int x = b + T * k8;
      if (g < 3) {
        bf16 *px = xt + (g * N + x) * CI;
        px[0] = f2bf((float)(v[k8].x * s0));
        px[1] = f2bf((float)(v[k8].y * s1));
      } else {
        // tile row of image row ra (real part) and ra + 1 (imaginary part)
        const int da = (ra - ym + N) % N, db = (ra + 1 - ym + N) % N;
        if (da < 3) xt[(da * N + x) * CI + 2] = f2bf((float)(v[k8].x * s2));
        if (db < 3) xt[(db * N + x) * CI + 2] = f2bf((float)(v[k8].y * s2));
      }
    }
  }
  __syncthreads();
  // row y of the input image (the last layer's sign clipping reads it)
  for (int e = threadIdx.x; e < N * CI / 8; e += 64 * kConvWaves)
    *(uint4 *)(X0 + (size_t)y * N * CI + 8 * e) = *(const uint4 *)(xt + N * CI + 8 * e);
  // the layer: one tile of 16 pixels per wave, A fragments from the tile
  const int tx = 16 * wave;
  s16x4 a[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    int xx = tx + m + dx;
    xx += xx < 0 ? N : (xx >= N ? -N : 0);
    a[t] = *(const s16x4 *)(xt + ((dy + 1) * N + xx) * CI + 4 * q);
  }
  f32x4 acc[1][CO16];
#pragma unroll
  for (int n = 0; n < CO16; ++n) acc[0][n] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int n = 0; n < CO16; ++n) {
      const s16x4 w = *(const s16x4 *)(ws + (t * CO + 16 * n + m) * WS + 4 * q);
      acc[0][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], w, acc[0][n], 0, 0, 0);
    }
  conv_store<CO, CONV_GELU>(acc[0], 0, y, tx, m, q, N, N, bias, nullptr, Zout, Y, 0.f, nullptr,
                            1.f);
}

// Weight gradient of one layer over a slice of rows: dW[t][co][ci] =
// sum_p G[p][co] X[p + d(t)][ci], db[co] = sum_p G[p][co].  A block stages a
// 64-pixel segment of G and the 3 x 66 pixels of X around it in LDS, then each
// wave runs the MFMA reduction (K = pixels) for taps wave, wave + 4, wave + 8:
// A = G^T (m = co), B = the shifted X (n = ci).  Partial sums per block go to
// part[block][kW + CO]; k_reduce_cols adds the blocks.
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const bf16 *__restrict__ X,
                                                        const bf16 *__restrict__ G,
                                                        float *__restrict__ part, int B, int H,
                                                        int W, int rows_per_block) {
  constexpr int CI16 = CI / 16, CO16 = CO / 16, SEG = 64;
  __shared__ bf16 gs[SEG * CO];
  __shared__ bf16 xs[3 * (SEG + 2) * CI];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = lane & 15, q = lane >> 4;
  f32x4 acc[3][CO16][CI16];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int a = 0; a < CO16; ++a)
#pragma unroll
      for (int c = 0; c < CI16; ++c) acc[k][a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  float dbacc = 0.f;
  const long rows = (long)B * H;
  const long r0 = (long)blockIdx.x * rows_per_block;
  for (long r = r0; r < r0 + rows_per_block && r < rows; ++r) {
    const int b = (int)(r / H), y = (int)(r - (long)b * H);
    for (int x0 = 0; x0 < W; x0 += SEG) {
      __syncthreads();
      for (int e = tid; e < SEG * CO; e += 256) {
        const int p = e / CO, c = e - p * CO;
        gs[e] = G[((size_t)r * W + x0 + p) * CO + c];
      }
      for (int e = tid; e < 3 * (SEG + 2) * CI; e += 256) {
        const int c = e % CI, rest = e / CI, p = rest % (SEG + 2), j = rest / (SEG + 2);
        int yy = y + j - 1, xx = x0 + p - 1;
        yy += yy < 0 ? H : (yy >= H ? -H : 0);
        xx += xx < 0 ? W : (xx >= W ? -W : 0);
        xs[e] = X[((size_t)((size_t)b * H + yy) * W + xx) * CI + c];
      }
      __syncthreads();
      if (tid < CO)
        for (int p = 0; p < SEG; ++p) dbacc += bf2f(gs[p * CO + tid]);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int t = wave + 4 * k;
        if (t < 9) {
          const int dy = t / 3 - 1, dx = t % 3 - 1;
          for (int p0 = 0; p0 < SEG; p0 += 16) {
            s16x4 av[CO16], bv[CI16];
#pragma unroll
            for (int a = 0; a < CO16; ++a) {
              const bf16 *g = gs + (p0 + 4 * q) * CO + 16 * a + m;
              av[a] = s16x4{(short)g[0], (short)g[CO], (short)g[2 * CO], (short)g[3 * CO]};
            }
#pragma unroll
            for (int c = 0; c < CI16; ++c) {
              const bf16 *xv = xs + ((dy + 1) * (SEG + 2) + p0 + 4 * q + dx + 1) * CI + 16 * c + m;
              bv[c] = s16x4{(short)xv[0], (short)xv[CI], (short)xv[2 * CI], (short)xv[3 * CI]};
            }
#pragma unroll
            for (int a = 0; a < CO16; ++a)
#pragma unroll
              for (int c = 0; c < CI16; ++c)
                acc[k][a][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[a], bv[c], acc[k][a][c], 0, 0, 0);
          }
        }
      }
    }
  }
  // D[4q+i][m] = co 16a + 4q + i, ci 16c + m
  float *dst = part + (size_t)blockIdx.x * (9 * CO * CI + CO);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int t = wave + 4 * k;
    if (t < 9) {
#pragma unroll
      for (int a = 0; a < CO16; ++a)
#pragma unroll
        for (int c = 0; c < CI16; ++c)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            dst[((size_t)t * CO + 16 * a + 4 * q + i) * CI + 16 * c + m] = acc[k][a][c][i];
    }
  }
  if (tid < CO) dst[9 * CO * CI + tid] = dbacc;
}

// out[j] = sum_b part[b][j], j < len (deterministic order)
__global__ void k_reduce_cols(const float *__restrict__ part, int nb, int len,
                              float *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * len + j];
  out[j] = s;
}

// fp32 [P][3] -> bf16 [P][16] (channels 3..15 zero)
__global__ void k_pack_input(const float *__restrict__ in, long P, bf16 *__restrict__ X) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P * 16) return;
  const long p = e >> 4;
  const int c = (int)(e & 15);
  X[e] = c < 3 ? f2bf(in[p * 3 + c]) : (bf16)0;
}

// prediction = channels 0, 1 of the last layer's Z
__global__ void k_unpack_output(const float *__restrict__ Z, long P, float *__restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P * 2) return;
  out[e] = Z[(e >> 1) * 16 + (e & 1)];
}

// MSE loss L = mean over pixels and both channels of (z - t)^2: dZ =
// bf16(2 (z - t) / (2P)) on channels 0, 1 and 0 on the padding; per-block
// partial sums of (z - t)^2 in part[block].
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Z,
                                              const float *__restrict__ tgt, long P,
                                              bf16 *__restrict__ dZ, double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  const float scale = 1.0f / (float)P;  // 2 / (2 P)
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < P * 16;
       e += (long)gridDim.x * blockDim.x) {
    const long p = e >> 4;
    const int c = (int)(e & 15);
    float g = 0.f;
    if (c < 2) {
      const float d = Z[e] - tgt[p * 2 + c];
      s += (double)d * d;
      g = d * scale;
    }
    dZ[e] = f2bf(g);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_mse_finish(const double *__restrict__ part, int nb, double inv, double *out) {
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[b];
  out[0] = s * inv;
}

// Adam on the fp32 master parameters
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  w[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

struct LayerPtrs {
  int woff[kLayers], poff[kLayers], ci[kLayers], co[kLayers];
};

// bf16 copies of the weights: Wf[t][co][ci] = W, Wb[t][ci][co] = W[8-t][co][ci]
__global__ void k_emit_bf16(const float *__restrict__ params, bf16 *__restrict__ wf,
                            bf16 *__restrict__ wb, LayerPtrs L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kWeights) return;
  int l = 0;
  while (l + 1 < kLayers && i >= L.woff[l + 1]) ++l;
  const int j = i - L.woff[l], CI = L.ci[l], CO = L.co[l];
  const int t = j / (CO * CI), r = j - t * CO * CI, co = r / CI, ci = r - co * CI;
  const bf16 h = f2bf(params[L.poff[l] + j]);
  wf[i] = h;
  wb[L.woff[l] + ((8 - t) * CI + ci) * CO + co] = h;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct r2d3c_cnn {
  int capB = 0, H = 0, W = 0;
  float *params = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
  bf16 *wf = nullptr, *wb = nullptr;
  bf16 *act[kLayers + 1] = {};  // act[0] input (16 ch), act[l+1] output of layer l
  float *z[kLayers] = {};       // pre-activations
  bf16 *g = nullptr, *g2 = nullptr;  // dZ ping-pong (32 ch)
  float *part = nullptr;        // weight-gradient partials
  double *lpart = nullptr;
  int wblocks = 0;
  long step = 0;
  bool weights_ready = false;
  float out_clamp = 0.f;        // inference output saturation (0: none)
  float sign_clip = 0.f;        // inference sign clipping strength of outputs 0, 1
  double *fgain = nullptr;      // closure forcing gain per shell (NULL: none)
  int nfgain = 0;
};

namespace {
int err(int code, const char *m) { return r2d3c_internal_set_error(code, m); }
int herr(hipError_t e, const char *where) {
  return r2d3c_internal_set_error(-(R2D3C_EHIP_BASE + (int)e),
                                  (std::string(where) + ": " + hipGetErrorString(e)).c_str());
}
#define CHK(expr)                                     \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return herr(e_, #expr);     \
  } while (0)
#define KCHK(where)                                   \
  do {                                                \
    hipError_t e_ = hipGetLastError();                \
    if (e_ != hipSuccess) return herr(e_, where);     \
  } while (0)

LayerPtrs layer_ptrs() {
  LayerPtrs L;
  for (int l = 0; l < kLayers; ++l) {
    L.woff[l] = woff(l);
    L.poff[l] = poff(l);
    L.ci[l] = kCI[l];
    L.co[l] = kCO[l];
  }
  return L;
}

int emit(r2d3c_cnn *c, hipStream_t s) {
  k_emit_bf16<<<(kWeights + 255) / 256, 256, 0, s>>>(c->params, c->wf, c->wb, layer_ptrs());
  KCHK("cnn weights");
  c->weights_ready = true;
  return 0;
}

template <int CI, int CO, int MODE>
void conv(const r2d3c_cnn *c, int l, const bf16 *X, const bf16 *Wt, const float *Zprev, float *Zout,
          bf16 *Y, int B, hipStream_t s, float clamp = 0.f, const bf16 *X0 = nullptr,
          float keep = 1.f) {
  const long tiles = (long)B * c->H * (c->W / 16), per = kConvWaves * kTilesPerWave;
  k_conv3x3<CI, CO, MODE><<<(unsigned)((tiles + per - 1) / per), 64 * kConvWaves, 0, s>>>(
      X, Wt, MODE == CONV_BWD ? nullptr : c->params + poff(l) + wsize(l), Zprev, Zout, Y, B, c->H,
      c->W, clamp, X0, keep);
}

// inference options: clamp > 0 saturates the outputs to [-clamp, clamp]; with
// sign_clip > 0 the output channels 0, 1 are scaled by 1 - sign_clip where
// they share the sign of input channels 0, 1
// keep_z: store the pre-activations of layers 0-4 (the backward pass needs
// them for GELU'); the last layer's Z (the prediction) is always stored.
// first_done: layer 0 already ran (k_conv0_mixed)
int forward(r2d3c_cnn *c, int B, hipStream_t s, float clamp = 0.f, float sign_clip = 0.f,
            bool keep_z = false, bool first_done = false) {
  const bf16 *wf = c->wf;
  if (!first_done)
    conv<16, 32, CONV_GELU>(c, 0, c->act[0], wf + woff(0), nullptr, keep_z ? c->z[0] : nullptr,
                            c->act[1], B, s);
  for (int l = 1; l < 5; ++l)
    conv<32, 32, CONV_GELU>(c, l, c->act[l], wf + woff(l), nullptr, keep_z ? c->z[l] : nullptr,
                            c->act[l + 1], B, s);
  conv<32, 16, CONV_LINEAR>(c, 5, c->act[5], wf + woff(5), nullptr, c->z[5], c->act[6], B, s,
                            clamp, sign_clip > 0.f ? c->act[0] : nullptr, 1.f - sign_clip);
  KCHK("cnn forward");
  return 0;
}

template <int CI, int CO>
int wgrad(r2d3c_cnn *c, int l, const bf16 *X, const bf16 *G, int B, hipStream_t s) {
  const long rows = (long)B * c->H;
  const int rpb = (int)((rows + c->wblocks - 1) / c->wblocks);
  const int nb = (int)((rows + rpb - 1) / rpb);
  k_conv3x3_wgrad<CI, CO><<<nb, 256, 0, s>>>(X, G, c->part, B, c->H, c->W, rpb);
  const int len = 9 * CO * CI + CO;
  k_reduce_cols<<<(len + 255) / 256, 256, 0, s>>>(c->part, nb, len, c->grad + poff(l));
  KCHK("cnn weight gradient");
  return 0;
}

// dZ of the last layer is in c->g (16 channels); writes every layer's gradient
int backward(r2d3c_cnn *c, int B, hipStream_t s) {
  const bf16 *wb = c->wb;
  bf16 *g = c->g, *g2 = c->g2;
  int rc;
  if ((rc = wgrad<32, 16>(c, 5, c->act[5], g, B, s))) return rc;
  // dZ4 = conv(dZ5, Wb5) * gelu'(Z4): in 16 channels, out 32
  conv<16, 32, CONV_BWD>(c, 5, g, wb + woff(5), c->z[4], nullptr, g2, B, s);
  std::swap(g, g2);
  for (int l = 4; l >= 1; --l) {
    if ((rc = wgrad<32, 32>(c, l, c->act[l], g, B, s))) return rc;
    conv<32, 32, CONV_BWD>(c, l, g, wb + woff(l), c->z[l - 1], nullptr, g2, B, s);
    std::swap(g, g2);
  }
  if ((rc = wgrad<16, 32>(c, 0, c->act[0], g, B, s))) return rc;
  KCHK("cnn backward");
  return 0;
}

int pack(r2d3c_cnn *c, const float *in, int B, hipStream_t s) {
  const long P = (long)B * c->H * c->W;
  k_pack_input<<<(unsigned)((P * 16 + 255) / 256), 256, 0, s>>>(in, P, c->act[0]);
  KCHK("cnn input");
  return 0;
}

int check_batch(const r2d3c_cnn *c, int B) {
  if (B < 1 || B > c->capB) return err(R2D3C_EINVAL, "cnn: batch must be in 1..capacity");
  if (!c->weights_ready) return err(R2D3C_EINVAL, "cnn: parameters not set");
  return 0;
}

int loss_grad(r2d3c_cnn *c, const float *tgt, int B, double *d_loss, hipStream_t s) {
  const long P = (long)B * c->H * c->W;
  k_mse<<<256, 256, 0, s>>>(c->z[5], tgt, P, c->g, c->lpart);
  k_mse_finish<<<1, 1, 0, s>>>(c->lpart, 256, 1.0 / (2.0 * (double)P), d_loss);
  KCHK("cnn loss");
  return 0;
}

void cnn_free(r2d3c_cnn *c) {
  if (!c) return;
  float *f[] = {c->params, c->grad, c->m1, c->m2, c->part};
  for (float *p : f) if (p) (void)hipFree(p);
  if (c->wf) (void)hipFree(c->wf);
  if (c->wb) (void)hipFree(c->wb);
  if (c->g) (void)hipFree(c->g);
  if (c->g2) (void)hipFree(c->g2);
  if (c->lpart) (void)hipFree(c->lpart);
  if (c->fgain) (void)hipFree(c->fgain);
  for (auto *a : c->act) if (a) (void)hipFree(a);
  for (auto *z : c->z) if (z) (void)hipFree(z);
  delete c;
}

}  // namespace

extern "C" {

int r2d3c_cnn_param_count(void) { return kParams; }

int r2d3c_cnn_layer_offsets(int *woffs, int *boffs, int *cin, int *cout) {
  for (int l = 0; l < kLayers; ++l) {
    if (woffs) woffs[l] = poff(l);
    if (boffs) boffs[l] = poff(l) + wsize(l);
    if (cin) cin[l] = kCI[l];
    if (cout) cout[l] = kCO[l];
  }
  return kLayers;
}

int r2d3c_cnn_create(r2d3c_cnn **out, int batch, int height, int width) {
  if (!out) return err(R2D3C_EINVAL, "cnn pointer is NULL");
  *out = nullptr;
  if (batch < 1 || height < 3 || width < 64 || width % 64 != 0)
    return err(R2D3C_EINVAL, "cnn: batch >= 1, height >= 3, width a multiple of 64");
  r2d3c_cnn *c = new r2d3c_cnn();
  c->capB = batch;
  c->H = height;
  c->W = width;
  const size_t P = (size_t)batch * height * width;
  hipError_t e = hipSuccess;
  auto al = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  al((void **)&c->params, sizeof(float) * kParams);
  al((void **)&c->grad, sizeof(float) * kParams);
  al((void **)&c->m1, sizeof(float) * kParams);
  al((void **)&c->m2, sizeof(float) * kParams);
  al((void **)&c->wf, sizeof(bf16) * kWeights);
  al((void **)&c->wb, sizeof(bf16) * kWeights);
  al((void **)&c->act[0], sizeof(bf16) * P * 16);
  for (int l = 0; l < kLayers; ++l) {
    al((void **)&c->act[l + 1], sizeof(bf16) * P * kCO[l]);
    al((void **)&c->z[l], sizeof(float) * P * kCO[l]);
  }
  al((void **)&c->g, sizeof(bf16) * P * 32);
  al((void **)&c->g2, sizeof(bf16) * P * 32);
  c->wblocks = 256;
  al((void **)&c->part, sizeof(float) * (size_t)c->wblocks * (9 * 32 * 32 + 32));
  al((void **)&c->lpart, sizeof(double) * 256);
  if (e == hipSuccess) e = hipMemset(c->m1, 0, sizeof(float) * kParams);
  if (e == hipSuccess) e = hipMemset(c->m2, 0, sizeof(float) * kParams);
  if (e != hipSuccess) {
    cnn_free(c);
    return herr(e, "cnn allocation");
  }
  *out = c;
  return 0;
}

int r2d3c_cnn_destroy(r2d3c_cnn *c) {
  cnn_free(c);
  return 0;
}

int r2d3c_cnn_set_params(r2d3c_cnn *c, const float *d_params, void *stream) {
  if (!c || !d_params) return err(R2D3C_EINVAL, "cnn: NULL argument");
  hipStream_t s = (hipStream_t)stream;
  CHK(hipMemcpyAsync(c->params, d_params, sizeof(float) * kParams, hipMemcpyDeviceToDevice, s));
  CHK(hipMemsetAsync(c->m1, 0, sizeof(float) * kParams, s));
  CHK(hipMemsetAsync(c->m2, 0, sizeof(float) * kParams, s));
  c->step = 0;
  return emit(c, s);
}

int r2d3c_cnn_get_params(r2d3c_cnn *c, float *d_params, void *stream) {
  if (!c || !d_params) return err(R2D3C_EINVAL, "cnn: NULL argument");
  CHK(hipMemcpyAsync(d_params, c->params, sizeof(float) * kParams, hipMemcpyDeviceToDevice,
                     (hipStream_t)stream));
  return 0;
}

int r2d3c_cnn_shape(const r2d3c_cnn *c, int *batch, int *height, int *width) {
  if (!c) return err(R2D3C_EINVAL, "cnn: NULL argument");
  if (batch) *batch = c->capB;
  if (height) *height = c->H;
  if (width) *width = c->W;
  return 0;
}

int r2d3c_cnn_buffers(r2d3c_cnn *c, void **input_bf16, const float **output_f32) {
  if (!c) return err(R2D3C_EINVAL, "cnn: NULL argument");
  if (input_bf16) *input_bf16 = c->act[0];
  if (output_f32) *output_f32 = c->z[5];
  return 0;
}

int r2d3c_cnn_forward_buffers(r2d3c_cnn *c, int batch, void *stream) {
  if (!c) return err(R2D3C_EINVAL, "cnn: NULL argument");
  int rc = check_batch(c, batch);
  if (rc) return rc;
  return forward(c, batch, (hipStream_t)stream, c->out_clamp, c->sign_clip);
}


int r2d3c_cnn_set_sign_clip(r2d3c_cnn *c, double strength) {
  if (!c) return err(R2D3C_EINVAL, "cnn: NULL argument");
  if (!(strength >= 0.0 && strength <= 1.0)) return err(R2D3C_EINVAL, "cnn: strength in [0, 1]");
  c->sign_clip = (float)strength;
  return 0;
}

int r2d3c_cnn_set_forcing_gain(r2d3c_cnn *c, const double *gain, int nshells) {
  if (!c) return err(R2D3C_EINVAL, "cnn: NULL argument");
  if (!gain || nshells == 0) {
    c->nfgain = 0;
    return 0;
  }
  const int cap = (c->H > c->W ? c->H : c->W) / 2 + 2;
  if (nshells < 0 || nshells > cap) return err(R2D3C_EINVAL, "cnn: nshells in [0, max(h, w)/2 + 2]");
  for (int k = 0; k < nshells; ++k)
    if (!(gain[k] >= 0.0 && gain[k] <= 16.0)) return err(R2D3C_EINVAL, "cnn: gains in [0, 16]");
  if (!c->fgain) CHK(hipMalloc((void **)&c->fgain, sizeof(double) * cap));
  CHK(hipMemcpy(c->fgain, gain, sizeof(double) * nshells, hipMemcpyHostToDevice));
  c->nfgain = nshells;
  return 0;
}

int r2d3c_cnn_set_output_clamp(r2d3c_cnn *c, double clamp) {
  if (!c) return err(R2D3C_EINVAL, "cnn: NULL argument");
  if (!(clamp >= 0.0)) return err(R2D3C_EINVAL, "cnn: clamp must be >= 0");
  c->out_clamp = (float)clamp;
  return 0;
}

int r2d3c_cnn_forward(r2d3c_cnn *c, const float *d_in, int batch, float *d_out, void *stream) {
  if (!c || !d_in || !d_out) return err(R2D3C_EINVAL, "cnn: NULL argument");
  int rc = check_batch(c, batch);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = pack(c, d_in, batch, s)) || (rc = forward(c, batch, s, c->out_clamp, c->sign_clip)))
    return rc;
  const long P = (long)batch * c->H * c->W;
  k_unpack_output<<<(unsigned)((P * 2 + 255) / 256), 256, 0, s>>>(c->z[5], P, d_out);
  KCHK("cnn output");
  return 0;
}

int r2d3c_cnn_gradients(r2d3c_cnn *c, const float *d_in, const float *d_target, int batch,
                        float *d_grad, double *d_loss, void *stream) {
  if (!c || !d_in || !d_target || !d_loss) return err(R2D3C_EINVAL, "cnn: NULL argument");
  int rc = check_batch(c, batch);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = pack(c, d_in, batch, s)) || (rc = forward(c, batch, s, 0.f, 0.f, true)) ||
      (rc = loss_grad(c, d_target, batch, d_loss, s)) || (rc = backward(c, batch, s)))
    return rc;
  if (d_grad)
    CHK(hipMemcpyAsync(d_grad, c->grad, sizeof(float) * kParams, hipMemcpyDeviceToDevice, s));
  return 0;
}

int r2d3c_cnn_train_step(r2d3c_cnn *c, const float *d_in, const float *d_target, int batch,
                         float lr, double *d_loss, void *stream) {
  int rc = r2d3c_cnn_gradients(c, d_in, d_target, batch, nullptr, d_loss, stream);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  ++c->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)c->step), c2 = 1.f - std::pow(b2, (float)c->step);
  k_adam<<<(kParams + 255) / 256, 256, 0, s>>>(c->params, c->grad, c->m1, c->m2, kParams, lr, b1,
                                               b2, 1e-8f, c1, c2);
  KCHK("cnn adam");
  return emit(c, s);
}

}  // extern "C"

// Inference of one image whose input pass is folded into the first layer
// (k_conv0_mixed): M is a closure run's step workspace (mixed fields 0, 2, 4:
// psi, omega, vz; field stride fs, column pitch kp, columns kx < nc), tw the
// grid's twiddle table, s0..s2 the input scales (times 1/n^2).  The input
// image buffer receives the same values as r2d3c_closure_net_inputs would
// write.  Networks of 256 x 256 images only (one image row per block of the
// first layer); others return R2D3C_ENOTSUP.
int r2d3c_internal_cnn_forward_mixed(r2d3c_cnn *c, const double *M, size_t fs, int kp, int nc,
                                     const double *tw, double s0, double s1, double s2,
                                     void *stream) {
  constexpr int N = 16 * kConvWaves;
  if (!c || !M || !tw) return err(R2D3C_EINVAL, "cnn: NULL argument");
  if (c->H != N || c->W != N) return err(R2D3C_ENOTSUP, "cnn: folded input pass needs 256 x 256");
  int rc = check_batch(c, 1);
  if (rc) return rc;
  const hipStream_t s = (hipStream_t)stream;
  const r2d::cplx *m = (const r2d::cplx *)M;
  k_conv0_mixed<N><<<N, 64 * kConvWaves, 0, s>>>(
      m + 2 * fs, m + 4 * fs, m, kp, nc, (const r2d::cplx *)tw, s0, s1, s2, c->act[0],
      c->wf + woff(0), c->params + poff(0) + wsize(0), nullptr, c->act[1]);
  KCHK("cnn first layer (folded input pass)");
  return forward(c, 1, s, c->out_clamp, c->sign_clip, false, true);
}

// The closure forcing gain table of a network (rot2d3c.hip's forcing epilogue).
void r2d3c_internal_cnn_forcing_gain(const r2d3c_cnn *c, const double **gain, int *n) {
  *gain = c && c->nfgain ? c->fgain : nullptr;
  *n = c ? c->nfgain : 0;
}
