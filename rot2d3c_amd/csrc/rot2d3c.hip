// rot2d3c.hip -- gfx950 kernels and C ABI of the Rot2D3C pseudospectral step.
//
// Layouts
//   spectral field  S[i][j]   i = kx column 0..n/2, j = ky in FFT order
//                             (reference layout, src/fftp2D.f90:153)
//   real field      R[y][x]
//   mixed field     M(kx, y): the column transform of a spectral field, kept
//                   y-tiled as M[(y/4)][kx][y%4] with a column pitch kp.  A
//                   column kernel touches 64-byte chunks (4 y of one kx) and a
//                   row kernel reads or writes 4 whole rows as one contiguous
//                   span.
//
// The fused step (r2d3c_step) uses only kx <= K = floor(n/3): every field that
// enters a Poisson bracket is dealiased (derivk2 masks kx^2+ky^2 > (n/3)^2,
// src/pseudospec2D_Rot2D3C.f90:35,48,58) and the stage update keeps only
// tiny < k^2 <= kmax (src/Rot2D3C.para.f90:580), so columns kx > K are zero in
// every mixed field and in the result.
//
// Per step: one column-inverse kernel, then per stage one row kernel and one
// fused kernel (column forward + semi-implicit update + next stage's column
// inverse, or the state write after the last stage).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/rot2d3c.h"
#include "fft_core.h"

using namespace r2d;

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int kyof(int j, int n) { return j < n / 2 ? j : j - n; }
// half-plane index i -> ka(i): 0..n/2-1, and -n/2 for i = n/2 (src/Rot2D3C.para.f90:220-223)
__device__ __forceinline__ int kxof(int i, int n) { return i < n / 2 ? i : i - n; }
// tix as a byte offset in 32 bits, for n <= 4096 (a tiled field then holds
// fewer than 2^24 16-byte slots): the access takes the field's uniform base
// address plus a 32-bit lane offset instead of a 64-bit address per lane.
__device__ __forceinline__ unsigned tix_bytes(int y, int kx, int kp) {
  return ((unsigned)(y >> 2) * (unsigned)(kp * 4) + (unsigned)kx * 4u + (unsigned)(y & 3)) * 16u;
}
template <class Tp> __device__ __forceinline__ Tp *at_bytes(Tp *base, unsigned off) {
  return reinterpret_cast<Tp *>(reinterpret_cast<typename std::conditional<
      std::is_const<Tp>::value, const char, char>::type *>(base) + off);
}
// integer power by binary exponentiation (Fortran x**n with integer n >= 0)
__host__ __device__ __forceinline__ double ipow(double x, int e) {
  double r = 1.0;
  while (e > 0) {
    if (e & 1) r *= x;
    x *= x;
    e >>= 1;
  }
  return r;
}
// Workgroup ids are dispatched round-robin over the 8 XCDs.  Map hardware id
// hb to logical id lb so that logical pairs (2p, 2p+1) -- neighbouring kx
// columns whose 64-byte chunks share 128-byte lines -- run on the same XCD in
// consecutive slots.  A bijection when gridDim.x % 16 == 0 (the launches
// round their grids up with round16); other grids keep the identity.
__device__ __forceinline__ int xcd_pairs(int hb) {
  if (gridDim.x & 15) return hb;
  const int x = hb & 7, s = hb >> 3;
  return ((s >> 1) << 4) + (x << 1) + (s & 1);
}
__host__ __forceinline__ int round16(int x) { return (x + 15) & ~15; }

// Sizes whose step kernels load each thread's last-pass twiddles once and
// share them across its transforms (a bit mask of log2 n; 0: every transform
// loads them).  The 2048 row kernel stays at 4 waves per SIMD with them and
// gains 1%; the 2048 column kernels (also 4 waves per SIMD with them) measured
// no gain, and at 512 and 4096 the extra registers cost occupancy or spill.
// The column inverse of a call's first step and the fused column kernel's
// emitted transforms take the same setting, so they emit the same bits.
#ifndef R2D_ROW_TWLAST
#define R2D_ROW_TWLAST (1 << 11)
#endif
#ifndef R2D_COL_TWLAST
#define R2D_COL_TWLAST 0
#endif
template <int N> constexpr bool row_twlast() { return ((R2D_ROW_TWLAST) >> ilog2(N)) & 1; }
template <int N> constexpr bool col_twlast() { return ((R2D_COL_TWLAST) >> ilog2(N)) & 1; }

// Latency-bound fused column launches (the emission split over slices) of
// grids up to this size load the state and the dense forcing before the
// column transform (colfused_body EARLY); 0: never.
#ifndef R2D_COL_EARLY_MAXN
#define R2D_COL_EARLY_MAXN 512
#endif

// Lone closure runs fold the closure forcing's passes into the first stage's
// launches (1) or run them as launches of their own (0); at 256 the
// network's input pass also runs inside its first layer (R2D_CL_FOLD0).
#ifndef R2D_CL_FOLD
#define R2D_CL_FOLD 1
#endif
#ifndef R2D_CL_FOLD0
#define R2D_CL_FOLD0 1
#endif

// Steady ensemble steps run as graphs of about this many steps (a multiple of
// the closure period); 1: one graph launch per step.
#ifndef R2D_ENS_BLOCK
#define R2D_ENS_BLOCK 8
#endif

// Latency-bound ensemble column launches (emission split over 4 slices): the
// slices as thread groups of one block in the last stage (1) or in every
// stage (2); 0: the last stage unsplit and the others' slices as blocks.
#ifndef R2D_ENS_GSLICES
#define R2D_ENS_GSLICES 1
#endif

// Ensemble column kernels below n = 512 pack 2 or 4 columns per block so that
// a block fills a wavefront (1: on, 0: one column per block).
#ifndef R2D_ENS_COL_GROUPS
#define R2D_ENS_COL_GROUPS 1
#endif

// spectral operators applied on load (derivk2 / laplak2 / inv_laplak2)
enum { OP_RAW = 0, OP_DX = 1, OP_DY = 2, OP_MASK = 3, OP_LAP = 4, OP_INVLAP = 5, OP_LDX = 6,
       OP_LDY = 7 };

struct Geo {
  int n;
  int nc, kp;    // fused step: columns kx <= K, pitch (even)
  int ncf, kpf;  // full transforms: n/2+1 columns, pitch
  double kmax, tiny;
};

__device__ __forceinline__ cplx apply_op(int op, cplx a, int kx, int ky, double kmax,
                                         double tiny) {
  const double ka2 = (double)kx * kx + (double)ky * ky;
  switch (op) {
    case OP_DX: return ka2 <= kmax ? mk(-(double)kx * a.y, (double)kx * a.x) : mk(0.0, 0.0);
    case OP_DY: return ka2 <= kmax ? mk(-(double)ky * a.y, (double)ky * a.x) : mk(0.0, 0.0);
    case OP_MASK: return ka2 <= kmax ? a : mk(0.0, 0.0);
    case OP_LAP: return mk(-ka2 * a.x, -ka2 * a.y);
    case OP_INVLAP: return ka2 > tiny ? mk(-a.x / ka2, -a.y / ka2) : mk(0.0, 0.0);
    // derivatives of the laplacian: OP_DX / OP_DY applied to OP_LAP's value
    case OP_LDX: {
      const cplx l = mk(-ka2 * a.x, -ka2 * a.y);
      return ka2 <= kmax ? mk(-(double)kx * l.y, (double)kx * l.x) : mk(0.0, 0.0);
    }
    case OP_LDY: {
      const cplx l = mk(-ka2 * a.x, -ka2 * a.y);
      return ka2 <= kmax ? mk(-(double)ky * l.y, (double)ky * l.x) : mk(0.0, 0.0);
    }
    default: return a;
  }
}


// ---------------------------------------------------------------------------
// generic transforms (full kx range, for r2d3c_c2r / r2c / poisson / cfl)
// ---------------------------------------------------------------------------

// column inverse: S[i][:] (with op) -> M(i, y)
template <int N>
__global__ __launch_bounds__(N / 8) void k_col_inv_full(const cplx *__restrict__ spec,
                                                         cplx *__restrict__ mix, int op,
                                                         double scale, Geo g,
                                                         const cplx *__restrict__ tw) {
  constexpr int T = N / 8;
  __shared__ cplx lds[Shape<N>::LDSN];
  const int i = blockIdx.x, b = threadIdx.x;
  const int kx = kxof(i, N);
  cplx v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int j = b + T * q;
    cplx a = spec[(size_t)i * N + j];
    a = apply_op(op, a, kx, kyof(j, N), g.kmax, g.tiny);
    v[q] = cscale(a, scale);
  }
  fft_canon<N, +1, 1>(v, lds, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) mix[tix(b + T * q, i, g.kpf)] = v[q];
}

// column forward: M(i, y) -> S[i][:]
template <int N>
__global__ __launch_bounds__(N / 8) void k_col_fwd_full(const cplx *__restrict__ mix,
                                                         cplx *__restrict__ spec, Geo g,
                                                         const cplx *__restrict__ tw) {
  constexpr int T = N / 8;
  __shared__ cplx lds[Shape<N>::LDSN];
  const int i = blockIdx.x, b = threadIdx.x;
  cplx v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = mix[tix(b + T * q, i, g.kpf)];
  fft_canon<N, -1, 1>(v, lds, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) spec[(size_t)i * N + b + T * q] = v[q];
}

// rows per block of the full-range row kernels: 2 packed transforms = 4 rows
template <int N> struct RowFull {
  static constexpr int P = (N / 8) * 2 <= 1024 ? 2 : 1;  // packed transforms per block
  static constexpr int ROWS = 2 * P;
};

// row c2r: M(:, y), M(:, y+1) -> R[y][:], R[y+1][:]
template <int N>
__global__ __launch_bounds__(N / 8 * RowFull<N>::P) void k_row_c2r(const cplx *__restrict__ mix,
                                                                    double *__restrict__ out,
                                                                    Geo g,
                                                                    const cplx *__restrict__ tw) {
  constexpr int T = N / 8, P = RowFull<N>::P;
  __shared__ cplx lds[P * Shape<N>::LDSN];
  const int r = threadIdx.x % P, b = threadIdx.x / P;
  const int y = blockIdx.x * RowFull<N>::ROWS + 2 * r;
  cplx v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = b + T * q;
    const bool mir = k > N / 2;
    const int kk = mir ? N - k : k;
    const cplx A = mix[tix(y, kk, g.kpf)], B = mix[tix(y + 1, kk, g.kpf)];
    v[q] = packed_z<N>(k, A, B, mir);
  }
  fft_canon<N, +1, P>(v, lds + r, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = b + T * q;
    out[(size_t)y * N + x] = v[q].x;
    out[(size_t)(y + 1) * N + x] = v[q].y;
  }
}

// Unpack a packed forward row transform held in LDS (canonical k order) into
// the half spectra of its two real rows, writing kx = k for k < ncols.
// KC: the column count ncols when it is known at compile time (n/3 + 1 for
// the step's kept columns, n/2 + 1 for full rows), 0 otherwise; whole slots
// past it are then skipped and whole slots below it stored without a test.
template <int N, int IL, int KC = 0>
__device__ __forceinline__ void unpack_store(const cplx *lds, int b, int ncols, int kp,
                                             cplx *__restrict__ mixa, int ya,
                                             cplx *__restrict__ mixb, int yb_) {
  constexpr int T = N / 8;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (KC > 0 && T * q >= KC) continue;
    const int k = b + T * q;
    if ((KC > 0 && T * q + T - 1 < KC) || k < ncols) {
      const cplx w = lds[pad<N>(k) * IL], wm = lds[pad<N>((N - k) & (N - 1)) * IL];
      const cplx A = mk(0.5 * (w.x + wm.x), 0.5 * (w.y - wm.y));
      const cplx B = mk(0.5 * (w.y + wm.y), -0.5 * (w.x - wm.x));
      *at_bytes(mixa, tix_bytes(ya, k, kp)) = A;
      *at_bytes(mixb, tix_bytes(yb_, k, kp)) = B;
    }
  }
}

// row r2c: R[y][:], R[y+1][:] -> M(:, y), M(:, y+1)
template <int N>
__global__ __launch_bounds__(N / 8 * RowFull<N>::P) void k_row_r2c(const double *__restrict__ in,
                                                                    cplx *__restrict__ mix, Geo g,
                                                                    const cplx *__restrict__ tw) {
  constexpr int T = N / 8, P = RowFull<N>::P;
  __shared__ cplx lds[P * Shape<N>::LDSN];
  const int r = threadIdx.x % P, b = threadIdx.x / P;
  const int y = blockIdx.x * RowFull<N>::ROWS + 2 * r;
  cplx v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = b + T * q;
    v[q] = mk(in[(size_t)y * N + x], in[(size_t)(y + 1) * N + x]);
  }
  fft_canon<N, -1, P>(v, lds + r, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) lds[pad<N>(b + T * q) * P + r] = v[q];
  __syncthreads();
  unpack_store<N, P, N / 2 + 1>(lds + r, b, g.ncf, g.kpf, mix, y, mix, y + 1);
}

// rows per block of the step row kernel and the single-field full-range
// bracket/max kernels: up to R2D_ROW_RW (a divisor of the 4-row y tile)
#ifndef R2D_ROW_RW
#define R2D_ROW_RW 2
#endif
template <int N> struct RowStep {
  static constexpr int ROWS = (N / 8) * R2D_ROW_RW <= 1024 ? R2D_ROW_RW
                              : ((N / 8) * 2 <= 1024 ? 2 : 1);
  static constexpr int G = 4 / ROWS;  // blocks sharing one y tile
};
// Blocks of one y tile read interleaved 16-byte slots of the same 64-byte
// chunks; map G consecutive logical blocks onto one XCD so they share L2.
// The map is a bijection only when gridDim.x % (8 G) == 0; smaller grids keep
// the identity.
template <int G> __device__ __forceinline__ int xcd_group(int hb) {
  if constexpr (G == 1) return hb;
  if (gridDim.x % (8 * G) != 0) return hb;
  const int x = hb & 7, s = hb >> 3;
  return (s / G) * (8 * G) + x * G + (s % G);
}
// the same map over the first n blocks of a launch (n % (8 G) == 0: a
// bijection of [0, n); otherwise the identity)
template <int G> __device__ __forceinline__ int xcd_group_n(int hb, int n) {
  if constexpr (G == 1) return hb;
  if (n % (8 * G) != 0) return hb;
  const int x = hb & 7, s = hb >> 3;
  return (s / G) * (8 * G) + x * G + (s % G);
}

// Poisson bracket rows (full range): z1 = ax + i ay, z2 = bx + i by from four
// mixed fields; out[y][x] = (ax*by - ay*bx)/n^4  (src/pseudospec2D_Rot2D3C.f90:195-216)
template <int N>
__global__ __launch_bounds__(N / 8 * RowStep<N>::ROWS) void k_row_bracket(
    const cplx *__restrict__ m0, const cplx *__restrict__ m1, const cplx *__restrict__ m2,
    const cplx *__restrict__ m3, double *__restrict__ out, double n4, Geo g,
    const cplx *__restrict__ tw) {
  constexpr int T = N / 8, RW = RowStep<N>::ROWS;
  __shared__ cplx lds[RW * Shape<N>::LDSN];
  const int r = threadIdx.x % RW, b = threadIdx.x / RW;
  const int y = blockIdx.x * RW + r;
  cplx z1[8], z2[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = b + T * q;
    const bool mir = k > N / 2;
    const int kk = mir ? N - k : k;
    z1[q] = packed_z<N>(k, m0[tix(y, kk, g.kpf)], m1[tix(y, kk, g.kpf)], mir);
    z2[q] = packed_z<N>(k, m2[tix(y, kk, g.kpf)], m3[tix(y, kk, g.kpf)], mir);
  }
  fft_canon<N, +1, RW>(z1, lds + r, b, tw);
  fft_canon<N, +1, RW>(z2, lds + r, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    double r3 = z1[q].x * z2[q].y;
    r3 = (r3 - z1[q].y * z2[q].x) / n4;
    out[(size_t)y * N + b + T * q] = r3;
  }
}

// max over the grid of |z|^2 with z = c2r(m0) + i c2r(m1)  (CFL, full range)
template <int N>
__global__ __launch_bounds__(N / 8 * RowStep<N>::ROWS) void k_row_gradmax(
    const cplx *__restrict__ m0, const cplx *__restrict__ m1, unsigned long long *maxbits, Geo g,
    const cplx *__restrict__ tw) {
  constexpr int T = N / 8, RW = RowStep<N>::ROWS;
  __shared__ cplx lds[RW * Shape<N>::LDSN];
  const int r = threadIdx.x % RW, b = threadIdx.x / RW;
  const int y = blockIdx.x * RW + r;
  cplx z[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = b + T * q;
    const bool mir = k > N / 2;
    const int kk = mir ? N - k : k;
    z[q] = packed_z<N>(k, m0[tix(y, kk, g.kpf)], m1[tix(y, kk, g.kpf)], mir);
  }
  fft_canon<N, +1, RW>(z, lds + r, b, tw);
  double m = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) m = fmax(m, z[q].x * z[q].x + z[q].y * z[q].y);
  for (int off = 32; off > 0; off >>= 1)
    if (off < (int)blockDim.x) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
}

// ---------------------------------------------------------------------------
// elementwise spectral operators and reductions
// ---------------------------------------------------------------------------
__global__ void k_spec_op(const cplx *__restrict__ a, cplx *__restrict__ b, int op, int n,
                          double kmax, double tiny) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    b[e] = apply_op(op, a[e], kxof(i, n), kyof(j, n), kmax, tiny);
  }
}

// mode 0: energy(a, kin); mode 1: inerprod(a, b, kin).  Per-block partial
// sums in a fixed order (deterministic), finished by k_reduce_finish.
__global__ __launch_bounds__(256) void k_reduce_partial(const cplx *__restrict__ a,
                                                         const cplx *__restrict__ bb, int mode,
                                                         int kin, int n, double kmax, double tiny,
                                                         double *__restrict__ partial) {
  __shared__ double red[256];
  const size_t total = (size_t)(n / 2 + 1) * n;
  const double inv_n4 = 1.0 / ((double)n * n * (double)n * n);
  double acc = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double kx = kxof(i, n), ky = kyof(j, n);
    const double ka2 = kx * kx + ky * ky;
    const double two = (i == 0) ? 1.0 : 2.0;
    const double kp = kin >= 0 ? ipow(ka2, kin) : (ka2 > 0.0 ? 1.0 / ipow(ka2, -kin) : 0.0);
    const cplx x = a[e];
    if (mode == 0) {
      if (kin >= 0 || ka2 >= tiny) acc += two * (x.x * x.x + x.y * x.y) * kp * inv_n4;
    } else {
      const cplx y = bb[e];
      const bool in = ka2 <= kmax && (kin >= 0 || ka2 >= tiny);
      if (in) acc += two * kp * (y.x * x.x + y.y * x.y) * inv_n4;  // Re(b conj(a))
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_reduce_finish(const double *__restrict__ partial, int np,
                                                        double *__restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int e = threadIdx.x; e < np; e += 256) acc += partial[e];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

struct DtConst {
  double cfl, kcut, nrm, a, bq, two_omega;
};
// dt = cfl/(kcut*sqrt(max)/nrm + nu*kcut^(2nn) + nuv*kcut^(2nnv) + 2*omega)
// (src/pseudospec2D_Rot2D3C.f90:892-895), summed left to right as written.
__device__ __forceinline__ double dt_from_max(const DtConst &c, unsigned long long bits) {
  const double tmp = sqrt(__longlong_as_double((long long)bits)) / c.nrm;
  return c.cfl / (((c.kcut * tmp + c.a) + c.bq) + c.two_omega);
}
__global__ void k_dt_from_max(const unsigned long long *maxbits, DtConst c, double *dt) {
  *dt = dt_from_max(c, *maxbits);
}

// dense forcing fill (rand_force, src/pseudospec2D_Rot2D3C.f90:1100-1122)
struct ForceEnt {
  int i, j;         // half-plane column, full index
  double sgn;       // +1, or -1 for the conjugate partner
  double cph, sph;  // cos(phase), sin(phase)
};
struct ForceSet {
  int nent;
  ForceEnt e[2];
  double num, kupk;  // sqrt(2*f0), kup**kin
};
__device__ __forceinline__ cplx force_value(const ForceSet &f, int e, double dt, double n2) {
  const double tmp = f.num / (sqrt(dt) * f.kupk);
  const double t1 = tmp * f.e[e].cph, t2 = tmp * f.e[e].sph;
  return mk(t1 * n2, f.e[e].sgn * (t2 * n2));
}
__global__ void k_force_fill(cplx *__restrict__ fp, ForceSet f, double dt, int n) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  const double n2 = (double)n * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    cplx v = mk(0.0, 0.0);
    for (int q = 0; q < f.nent; ++q)
      if (f.e[q].i == i && f.e[q].j == j) v = force_value(f, q, dt, n2);
    fp[e] = v;
  }
}

// ---------------------------------------------------------------------------
// diagnostics: cond_check, spectrum, transfers (src/pseudospec2D_Rot2D3C.f90:455-837)
// ---------------------------------------------------------------------------
// ka2^kin as the reference's sums use it; false when the mode is excluded
// (kin < 0 skips ka2 < tiny, src/pseudospec2D_Rot2D3C.f90:377-386)
__device__ __forceinline__ bool kpow(double ka2, int kin, double tiny, double *w) {
  if (kin >= 0) { *w = ipow(ka2, kin); return true; }
  if (ka2 < tiny) return false;
  *w = 1.0 / ipow(ka2, -kin);
  return true;
}

struct CondConst {
  int n, nn, mm, nnv, mmv;
  double kmax, tiny, band_lo, band_hi;  // Efp band: kup^2 <= ka2 <= 2.01f*kup^2
};
constexpr int kCondSums = 22;

// The 22 sums of cond_check in one pass, one partial per block and sum.
// Order: Ep1 injp1 Dsp Hds Epv injz Dspv Hdsv coup Ep2 injp2 Dspw Hdsw Eph
// injh1 injh2 Dsph1 Hdsh1 Dsph2 Hdsh2 Efp Efpz (:484-568).
__global__ __launch_bounds__(256) void k_cond_partial(const cplx *__restrict__ ps,
                                                       const cplx *__restrict__ vz,
                                                       const cplx *__restrict__ fp,
                                                       const cplx *__restrict__ fz, CondConst c,
                                                       double *__restrict__ partial) {
  __shared__ double red[256];
  double acc[kCondSums];
#pragma unroll
  for (int q = 0; q < kCondSums; ++q) acc[q] = 0.0;
  const int n = c.n;
  const size_t total = (size_t)(n / 2 + 1) * n;
  const double inv_n4 = 1.0 / ((double)n * n * (double)n * n);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double kx = kxof(i, n), ky = kyof(j, n);
    const double ka2 = kx * kx + ky * ky;
    const double two = (i == 0) ? 1.0 : 2.0;
    const cplx a = ps[e], v = vz[e], f = fp[e], g = fz[e];
    const double pp = a.x * a.x + a.y * a.y, vv = v.x * v.x + v.y * v.y;
    const bool in = ka2 <= c.kmax;
    double w;
    // energy(a, kin): every mode for kin >= 0
    auto en = [&](int q, double sq, int kin) { if (kpow(ka2, kin, c.tiny, &w)) acc[q] += two * sq * w * inv_n4; };
    // inerprod(a, b, kin): Re(b conj a) inside the dealiasing mask
    auto ip = [&](int q, cplx x, cplx y, int kin) {
      if (in && kpow(ka2, kin, c.tiny, &w)) acc[q] += two * w * (y.x * x.x + y.y * x.y) * inv_n4;
    };
    en(0, pp, 1);
    ip(1, a, f, 1);
    en(2, pp, c.nn + 1);
    en(3, pp, 1 - c.mm);
    en(4, vv, 0);
    ip(5, v, g, 0);
    en(6, vv, c.nnv);
    en(7, vv, -c.mmv);
    ip(8, v, mk(-ky * a.y, ky * a.x), 0);  // C1 = derivk2(ps, 2)
    en(9, pp, 2);
    ip(10, a, f, 2);
    en(11, pp, c.nn + 2);
    en(12, pp, 2 - c.mm);
    ip(13, v, a, 1);
    ip(14, v, f, 1);
    ip(15, a, g, 1);
    ip(16, v, a, c.nn + 1);
    ip(17, v, a, 1 - c.mm);
    ip(18, a, v, c.nnv + 1);
    ip(19, a, v, 1 - c.mmv);
    if (ka2 <= c.band_hi && ka2 >= c.band_lo) {
      acc[20] += two * ka2 * pp * inv_n4;
      acc[21] += two * vv * inv_n4;
    }
  }
#pragma unroll
  for (int q = 0; q < kCondSums; ++q) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[q * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ void k_cond_finish(const double *__restrict__ partial, int np, double *__restrict__ out) {
  const int q = threadIdx.x;
  if (q >= kCondSums) return;
  double s = 0.0;
  for (int b = 0; b < np; ++b) s += partial[q * np + b];
  out[q] = s;
}

// max and min of a real field: per-block partials, then one block
__global__ __launch_bounds__(256) void k_minmax_partial(const double *__restrict__ r, size_t total,
                                                         double *__restrict__ partial) {
  __shared__ double mx[256], mn[256];
  double a = -INFINITY, b = INFINITY;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    a = fmax(a, r[e]);
    b = fmin(b, r[e]);
  }
  mx[threadIdx.x] = a;
  mn[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + st]);
      mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = mx[0];
    partial[2 * blockIdx.x + 1] = mn[0];
  }
}
__global__ void k_minmax_finish(const double *__restrict__ partial, int np, double *__restrict__ out) {
  double a = -INFINITY, b = INFINITY;
  for (int k = 0; k < np; ++k) {
    a = fmax(a, partial[2 * k]);
    b = fmin(b, partial[2 * k + 1]);
  }
  out[0] = a;
  out[1] = b;
}

// Extrema of an nprocs-rank run: block r reduces the rows [lo, hi) of rank r's
// slab (the split of `range`, src/fftp.f90:17-47, over rows 1..n); the finish
// combines maxima and minima alike with a max, as the reference's MPI_MAX
// does for both (src/pseudospec2D_Rot2D3C.f90:578,586).
__global__ __launch_bounds__(256) void k_minmax_slabs(const double *__restrict__ r, int n,
                                                       double *__restrict__ partial) {
  __shared__ double mx[256], mn[256];
  const int np = gridDim.x, rk = blockIdx.x, w1 = n / np, w2 = n % np;
  const int lo = rk * w1 + min(rk, w2), hi = lo + w1 + (rk < w2 ? 1 : 0);
  double a = -INFINITY, b = INFINITY;
  for (size_t e = (size_t)lo * n + threadIdx.x; e < (size_t)hi * n; e += blockDim.x) {
    a = fmax(a, r[e]);
    b = fmin(b, r[e]);
  }
  mx[threadIdx.x] = a;
  mn[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + st]);
      mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * rk] = mx[0];
    partial[2 * rk + 1] = mn[0];
  }
}
__global__ void k_minmax_finish_ranks(const double *__restrict__ partial, int np,
                                      double *__restrict__ out) {
  double a = -INFINITY, b = -INFINITY;
  for (int k = 0; k < np; ++k) {
    a = fmax(a, partial[2 * k]);
    b = fmax(b, partial[2 * k + 1]);
  }
  out[0] = a;
  out[1] = b;
}

// Shell sums, one block per shell kmn = blockIdx.x + 1, kmn = int(sqrt(ka2) + 0.5)
// (src/pseudospec2D_Rot2D3C.f90:652-656).  kmn is monotone in |ky| at fixed kx, so
// each thread walks its columns over the contiguous |ky| range of the shell,
// in a fixed order (deterministic).
// mode 0 (spectrum): out0 = sum two*ka2*|ps|^2, out1 = sum two*|vz|^2
// mode 1 (transfers): out0 = sum two*ka2*Re(ps conj c2), out1 = sum two*Re(ps conj c2),
//                     out2 = -sum two*Re(vz conj c1)  (all /n^4)
__device__ __forceinline__ int shell_of(double ka2) { return (int)(sqrt(ka2) + 0.5); }

__global__ __launch_bounds__(256) void k_shells(const cplx *__restrict__ ps, const cplx *__restrict__ vz,
                                                 const cplx *__restrict__ c1, const cplx *__restrict__ c2,
                                                 int mode, int n, double *__restrict__ out) {
  __shared__ double red[3][256];
  const int shell = blockIdx.x + 1, nsh = n / 2 + 1;
  const double inv_n4 = 1.0 / ((double)n * n * (double)n * n);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const int imax = shell < n / 2 ? shell : n / 2;  // kmn >= kx
  for (int i = threadIdx.x; i <= imax; i += blockDim.x) {
    const double kx = kxof(i, n);
    const double two = (i == 0) ? 1.0 : 2.0;
    const double lo2 = ((double)shell - 0.5) * ((double)shell - 0.5) - kx * kx;
    int a0 = lo2 > 0.0 ? (int)sqrt(lo2) - 1 : 0;
    if (a0 < 0) a0 = 0;
    for (int side = 0; side < 2; ++side) {  // ky >= 0, then ky < 0
      int a = side == 0 ? a0 : (a0 > 1 ? a0 : 1);
      const int amax = side == 0 ? n / 2 - 1 : n / 2;
      while (a <= amax && shell_of(kx * kx + (double)a * a) < shell) ++a;
      for (; a <= amax; ++a) {
        const double ka2 = kx * kx + (double)a * a;
        if (shell_of(ka2) != shell) break;
        const int j = side == 0 ? a : n - a;
        const size_t e = (size_t)i * n + j;
        const cplx p = ps[e];
        if (mode == 0) {
          const cplx v = vz[e];
          s0 += two * ka2 * (p.x * p.x + p.y * p.y) * inv_n4;
          s1 += two * (v.x * v.x + v.y * v.y) * inv_n4;
        } else {
          const cplx b2 = c2[e], b1 = c1[e], v = vz[e];
          const double re = p.x * b2.x + p.y * b2.y;
          s0 += two * ka2 * re * inv_n4;
          s1 += two * re * inv_n4;
          s2 -= two * (v.x * b1.x + v.y * b1.y) * inv_n4;
        }
      }
    }
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < (mode == 0 ? 2 : 3); ++q) out[q * nsh + blockIdx.x] = red[q][0];
}

// ---------------------------------------------------------------------------
// the fused step
// ---------------------------------------------------------------------------
struct StepConst {
  Geo g;
  size_t fs;  // complex elements per mixed field
  double nu, hnu, nuv, hnuv, two_omega;
  int nn, mm, nnv, mmv;
  int cfl_mode;
  double dt_fixed;
  DtConst dc;
  ForceSet fp, fz;
  const cplx *fpd, *fzd;  // static dense forcing (NULL = none)
  double two_omega_n2, inv_n4, n2;
};

// ---------------------------------------------------------------------------
// iflow = 2: const_inj and the stage-by-stage step
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool ci_band(double ka2, double lo2, double hi2) {
  return ka2 <= hi2 && ka2 >= lo2;
}
// fpre = ps/(|ps|^2+1) on the band; kx = 0 column for j = 1..n/2-1 mirrored by
// conjugation, j = 0 and n/2 zero (src/pseudospec2D_Rot2D3C.f90:1005-1028)
__global__ void k_ci_pre(const cplx *__restrict__ ps, cplx *__restrict__ fpre, double lo2, double hi2,
                         int n) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    int jj = j;
    bool conj = false, zero = false;
    if (i == 0) {
      if (j == 0 || j == n / 2) zero = true;
      else if (j > n / 2) { jj = n - j; conj = true; }
    }
    const double kx = kxof(i, n), ky = kyof(jj, n), ka2 = kx * kx + ky * ky;
    cplx v = mk(0.0, 0.0);
    if (!zero && ci_band(ka2, lo2, hi2)) {
      const cplx a = ps[(size_t)i * n + jj];
      const double d = a.x * a.x + a.y * a.y + 1.0;
      v = mk(a.x / d, a.y / d);
      if (conj) v.y = -v.y;
    }
    fpre[e] = v;
  }
}
// fp = fpre*fp0/Efp + i*(r*|k|)*ps on the band (:1033-1057); r: the randu
// value of the mode (0 off the band), Efp read from device memory
__global__ void k_ci_post(const cplx *__restrict__ ps, const cplx *__restrict__ fpre,
                          const float *__restrict__ rtab, const double *__restrict__ efp, double fp0,
                          double lo2, double hi2, cplx *__restrict__ fp, int n) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  const double ef = *efp;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    int jj = j;
    bool conj = false, zero = false;
    if (i == 0) {
      if (j == 0 || j == n / 2) zero = true;
      else if (j > n / 2) { jj = n - j; conj = true; }
    }
    const double kx = kxof(i, n), ky = kyof(jj, n), ka2 = kx * kx + ky * ky;
    cplx v = mk(0.0, 0.0);
    if (!zero && ci_band(ka2, lo2, hi2)) {
      const size_t ee = (size_t)i * n + jj;
      const cplx f = fpre[ee], a = ps[ee];
      const double tmp3 = (double)rtab[ee] * sqrt(ka2);
      v = mk(f.x * fp0 / ef + (-tmp3 * a.y), f.y * fp0 / ef + tmp3 * a.x);
      if (conj) v.y = -v.y;
    }
    fp[e] = v;
  }
}
// One stage update of src/Rot2D3C.para.f90:578-599 from spectral C3..C6 and
// dense fp, fz; dt from memory (CFL) or the argument.
__global__ void k_update(const cplx *__restrict__ ps, const cplx *__restrict__ vz, cplx *c1, cplx *c2,
                         const cplx *__restrict__ c3, const cplx *__restrict__ c4,
                         const cplx *__restrict__ c5, const cplx *__restrict__ c6,
                         const cplx *__restrict__ fp, const cplx *__restrict__ fz, StepConst c,
                         const double *__restrict__ dtp, int o) {
  const int n = c.g.n;
  const size_t total = (size_t)(n / 2 + 1) * n;
  const double dt = dtp ? *dtp : c.dt_fixed;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double kx = kxof(i, n), ky = kyof(j, n), ka2 = kx * kx + ky * ky;
    cplx a = mk(0.0, 0.0), b = mk(0.0, 0.0);
    if (ka2 <= c.g.kmax && ka2 > c.g.tiny) {
      const double tmp1 = dt / (double)o, tmp2 = 1.0 / ka2;
      double tmp3 = 1.0 + (c.nu * ipow(ka2, c.nn) + c.hnu * ipow(tmp2, c.mm)) * tmp1;
      const cplx t5 = c5[e], t4 = c4[e], f = fp[e];
      cplx t = mk((-t5.x + c.two_omega * t4.x) * tmp2 + f.x, (-t5.y + c.two_omega * t4.y) * tmp2 + f.y);
      t = cadd(ps[e], cscale(t, tmp1));
      a = mk(t.x / tmp3, t.y / tmp3);
      tmp3 = 1.0 + (c.nuv * ipow(ka2, c.nnv) + c.hnuv * ipow(tmp2, c.mmv)) * tmp1;
      const cplx t6 = c6[e], t3 = c3[e], g = fz[e];
      t = mk(t6.x + c.two_omega * t3.x + g.x, t6.y + c.two_omega * t3.y + g.y);
      t = cadd(vz[e], cscale(t, tmp1));
      b = mk(t.x / tmp3, t.y / tmp3);
    }
    c1[e] = a;
    c2[e] = b;
  }
}
__global__ void k_set_f64(double *p, double v) { *p = v; }

// Column inverse of the stage-1 state.  Logical blocks [0, nc): psi column kx
// -> M0 = psi~, M1 = d_y psi~, M2 = w~, M3 = d_y w~ with w = -k^2 psi
// (laplak2, src/Rot2D3C.para.f90:571); blocks [ncp, ncp+nc): vz column kx ->
// M4 = vz~, M5 = d_y vz~.  All inputs dealiased as derivk2 does.
template <int N, bool CS = false>
__device__ __forceinline__ void colinv_body(const cplx *__restrict__ psi,
                                            const cplx *__restrict__ vz, cplx *__restrict__ M,
                                            const StepConst &c, unsigned long long *maxbits,
                                            const cplx *__restrict__ tw, int lb) {
  constexpr int T = N / 8;
  __shared__ cplx lds[Shape<N>::LDSN];
  const int b = threadIdx.x;
  if (maxbits && lb == 0 && b == 0) *maxbits = 0ull;
  const int ncp = (c.g.nc + 1) & ~1;
  int field, kx;
  if (lb < c.g.nc) { field = 0; kx = lb; }
  else if (lb >= ncp && lb < ncp + c.g.nc) { field = 1; kx = lb - ncp; }
  else return;
  const cplx *src = (field == 0 ? psi : vz) + (size_t)kx * N;
  cplx s[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {  // inverse transforms take the interleaved order
    const int j = ilv<N, CS>(b, q), ky = kyof(j, N);
    const double ka2 = (double)kx * kx + (double)ky * ky;
    s[q] = ka2 <= c.g.kmax ? src[j] : mk(0.0, 0.0);
  }
  const int nf = field == 0 ? 4 : 2;
  cplx *dst0 = M + (size_t)(field == 0 ? 0 : 4) * c.fs;
  // the same twiddles as the fused column kernel's emitted transforms, so the
  // first step of a call and the later ones emit the same bits
  const TwLast tls = tw_last<N>(tw, b);
  const TwLast *tl = col_twlast<N>() ? &tls : nullptr;
  for (int f = 0; f < nf; ++f) {
    // opaque copy of the thread index: keeps the per-q index and address math
    // from being hoisted out of this loop (it would pin ~120 VGPRs and halve
    // the occupancy)
    int bb = b;
    asm volatile("" : "+v"(bb));
    cplx v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int ky = kyof(ilv<N, CS>(bb, q), N);
      const double ka2 = (double)kx * kx + (double)ky * ky;
      cplx a = s[q];
      // products rounded on their own (never fused into the transform's
      // first additions), so every kernel that emits these fields, here and
      // in the fused column kernel, produces the same bits
      if (f >= 2) a = mk(mul_rn(-ka2, a.x), mul_rn(-ka2, a.y));
      if (f & 1) a = mk(mul_rn(-(double)ky, a.y), mul_rn((double)ky, a.x));
      v[q] = a;
    }
    fft_dit<N, +1, 1, CS>(v, lds, bb, tw, tl);
    cplx *dst = dst0 + (size_t)f * c.fs;
#pragma unroll
    for (int q = 0; q < 8; ++q) dst[tix(bb + T * q, kx, c.g.kp)] = v[q];
  }
}

template <int N, bool CS = false>
__global__ __launch_bounds__(N / 8) void k_step_colinv(const cplx *__restrict__ psi,
                                                        const cplx *__restrict__ vz,
                                                        cplx *__restrict__ M, StepConst c,
                                                        unsigned long long *maxbits,
                                                        const cplx *__restrict__ tw) {
  colinv_body<N, CS>(psi, vz, M, c, maxbits, tw, xcd_pairs(blockIdx.x));
}

// Row kernel of one stage: for RW rows, z1 = psi_x + i psi_y,
// z2 = w_x + i w_y, z3 = vz_x + i vz_y (x derivatives as i*kx on the mixed
// fields), then
//   P1 = (-(psi_x w_y - psi_y w_x) + 2 Omega n^2 vz_y)/n^4
//   P2 = ( (psi_x vz_y - psi_y vz_x) + 2 Omega n^2 psi_y)/n^4
// whose forward transforms are -C5 + 2 Omega C4 and C6 + 2 Omega C3 of
// src/Rot2D3C.para.f90:571-576,586,591.  P1 + i P2 is forward transformed as
// one packed row and unpacked into Q0, Q1 for kx < nc.  With maxbits set, the
// max of psi_x^2 + psi_y^2 is accumulated (CFL on the pre-step state).
// Packed row inputs of one transform: the i*kx derivative of field A and
// field B at the kept columns kx < nc = n/3 + 1 (and their mirrors), zero
// elsewhere.  The loads of a group of LG slots are issued before any of its
// arithmetic, without branches (slots in the zero band read column 0 and are
// discarded), so they are in flight together; slots that no thread can keep
// (by nc = n/3 + 1) are not loaded at all.  Smaller groups hold fewer
// registers: from n = 2048 one slot per group keeps the row kernel at 4
// waves per SIMD (groups of 4 or 8 slots take 136-142 VGPRs) and measured
// fastest there.
#ifndef R2D_ROW_LG
#define R2D_ROW_LG 1
#endif
template <int N> constexpr int row_load_group() { return N <= 1024 ? 8 : R2D_ROW_LG; }
template <int N>
__device__ __forceinline__ void load_packed_step(cplx v[8], const cplx *__restrict__ FA,
                                                 const cplx *__restrict__ FB, int y, int b, int nc,
                                                 int kp) {
  // nc = n/3 + 1 (checked when the plan is made): with 0 <= b < T, slot q
  // holds k in [T q, T q + T - 1], so whole slots are known to be kept
  // columns (all_lo), their mirrors (all_hi) or zero at compile time
  constexpr int T = N / 8, NC = N / 3 + 1, LG = row_load_group<N>();
  (void)nc;
#pragma unroll
  for (int g = 0; g < 8; g += LG) {
    cplx fa[LG], fb[LG];
#pragma unroll
    for (int u = 0; u < LG; ++u) {
      const int q = g + u;
      if (T * q >= NC && T * q + T - 1 <= N - NC) continue;  // only zero-band slots
      const int k = b + T * q;
      const bool all_lo = T * q + T - 1 < NC, all_hi = T * q > N - NC;
      const bool lo = all_lo || (!all_hi && k < NC), hi = all_hi || (!all_lo && k > N - NC);
      const int kk = lo ? k : (hi ? N - k : 0);
      fa[u] = *at_bytes(FA, tix_bytes(y, kk, kp));
      fb[u] = *at_bytes(FB, tix_bytes(y, kk, kp));
    }
#pragma unroll
    for (int u = 0; u < LG; ++u) {
      const int q = g + u;
      if (T * q >= NC && T * q + T - 1 <= N - NC) { v[q] = mk(0.0, 0.0); continue; }
      const int k = b + T * q;
      const bool all_lo = T * q + T - 1 < NC, all_hi = T * q > N - NC;
      const bool lo = all_lo || (!all_hi && k < NC), hi = all_hi || (!all_lo && k > N - NC);
      // z = packed_z(A = i kx f, B) with its roundings spelled out: the
      // real part of A fused into its add, the imaginary part rounded on
      // its own where it is zeroed at k = 0 and k = n/2
      const double dk = (double)(lo ? k : N - k);
      const cplx f = fa[u], B = fb[u];
      cplx z;
      if (lo) {  // Z[k] = A + i B
        const bool edge = (q == 0 && k == 0) || (T * q <= N / 2 && N / 2 < T * q + T && k == N / 2);
        const double ay = edge ? 0.0 : mul_rn(dk, f.x), by = edge ? 0.0 : B.y;
        z = mk(fma(-f.y, dk, -by), ay + B.x);
      } else {  // Z[n-k] = conj(A) + i conj(B)
        z = mk(fma(-f.y, dk, B.y), fma(-f.x, dk, B.x));
      }
      v[q] = (lo || hi) ? z : mk(0.0, 0.0);
    }
    if (LG < 8) __builtin_amdgcn_sched_barrier(0);  // next group's loads stay below
  }
}

// Slots q of a packed row transform's canonical register set that are zero
// for every thread (all their columns lie in the band past n/3).
template <int N> constexpr int zero_row_slots() {
  constexpr int T = N / 8, NC = N / 3 + 1;
  int z = 0;
  for (int q = 0; q < 8; ++q)
    if (T * q >= NC && T * q + T - 1 <= N - NC) z |= 1 << q;
  return z;
}

// Unpack of a packed forward row transform whose radix-4 last pass was run
// mirrored (fft_canon MIR): thread b holds X[j_u + 2T m] in v[u + 2m] with
// j_0 = b, j_1 = 2T - b (T for b = 0), so each output's mirror X[N - k] is in
// its own registers.  Writes the half spectra of the two real rows,
// A = (X[k] + conj X[N-k])/2 to QA and B = (X[k] - conj X[N-k])/(2i) to QB,
// at the kept columns k < n/3 + 1 of row y: the arithmetic of unpack_store
// without the trip through LDS.
template <int N>
__device__ __forceinline__ void mirror_store(const cplx v[8], int b, int kp,
                                             cplx *__restrict__ QA, cplx *__restrict__ QB,
                                             int y) {
  constexpr int T = N / 8, H = 2 * T, NC = N / 3 + 1;
  const bool b0 = b == 0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int lo = (u ? T : 0) + H * m, hi = (u ? H - 1 : T - 1) + H * m;
      if (lo >= NC) continue;  // wholly past the kept columns
      const int k = (u ? (b0 ? T : H - b) : b) + H * m;
      if (hi >= NC && k >= NC) continue;
      const cplx w = v[u + 2 * m];
      // X[N - k]: (1 - u, 3 - m), and for b = 0 (j_0 = 0, j_1 = T) the
      // self-mirrored columns (0, -m mod 4) and (1, 3 - m)
      const cplx wm = u ? (b0 ? v[1 + 2 * (3 - m)] : v[2 * (3 - m)])
                        : (b0 ? v[2 * ((4 - m) & 3)] : v[1 + 2 * (3 - m)]);
      const cplx A = mk(0.5 * (w.x + wm.x), 0.5 * (w.y - wm.y));
      const cplx B = mk(0.5 * (w.y + wm.y), -0.5 * (w.x - wm.x));
      *at_bytes(QA, tix_bytes(y, k, kp)) = A;
      *at_bytes(QB, tix_bytes(y, k, kp)) = B;
    }
  }
}

// Rows y (this thread's, r = y % RW) of one run: three packed inverse
// transforms, the brackets and the rotation terms, one forward transform,
// unpacked into the bracket fields Q.  b: the thread's index in its group of
// T*RW threads; lds: the group's RW * LDSN slots.  Every thread of the block
// must call it together (block barriers).  LAST: nothing follows in the
// kernel (the forward transform skips its closing barrier).
template <int N, bool LAST = false>
__device__ __forceinline__ void step_row_body(const cplx *__restrict__ M, cplx *__restrict__ Q,
                                              const StepConst &c, unsigned long long *maxbits,
                                              const cplx *__restrict__ tw, double two_omega_n2,
                                              int y, int b, int r, cplx *lds) {
  constexpr int T = N / 8, RW = RowStep<N>::ROWS;
  // Each transform below takes an opaque copy of the thread index (b0..b7):
  // the compiler cannot hoist one transform's index and address math over
  // the others, which keeps the kernel at 128 VGPRs (4 waves per SIMD)
  // instead of ~245.
  const int nc = c.g.nc, kp = c.g.kp;
  const size_t fs = c.fs;
  // the last-pass twiddles of this thread, shared by the four transforms
  const TwLast tls = tw_last<N>(tw, b);
  const TwLast *tl = row_twlast<N>() ? &tls : nullptr;
  cplx v[8];
  double px[8], py[8], j1[8];
  int b0 = b;
  asm volatile("" : "+v"(b0));
  load_packed_step<N>(v, M, M + fs, y, b0, nc, kp);
  int b1 = b;
  asm volatile("" : "+v"(b1));
  constexpr int ZQ = zero_row_slots<N>();
  fft_canon<N, +1, RW, 1, ZQ>(v, lds + r, b1, tw, tl);
#pragma unroll
  for (int q = 0; q < 8; ++q) { px[q] = v[q].x; py[q] = v[q].y; }
  if (maxbits) {
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) m = fmax(m, px[q] * px[q] + py[q] * py[q]);
    for (int off = 32; off > 0; off >>= 1)
      if (off < (int)blockDim.x) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
  }
  int b2 = b;
  asm volatile("" : "+v"(b2));
  load_packed_step<N>(v, M + 2 * fs, M + 3 * fs, y, b2, nc, kp);
  int b3 = b;
  asm volatile("" : "+v"(b3));
  fft_canon<N, +1, RW, 1, ZQ>(v, lds + r, b3, tw, tl);
#pragma unroll
  for (int q = 0; q < 8; ++q) j1[q] = px[q] * v[q].y - py[q] * v[q].x;
  int b4 = b;
  asm volatile("" : "+v"(b4));
  load_packed_step<N>(v, M + 4 * fs, M + 5 * fs, y, b4, nc, kp);
  int b5 = b;
  asm volatile("" : "+v"(b5));
  fft_canon<N, +1, RW, 1, ZQ>(v, lds + r, b5, tw, tl);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double vx = v[q].x, vy = v[q].y;
    const double p1 = (-j1[q] + two_omega_n2 * vy) * c.inv_n4;
    const double p2 = ((px[q] * vy - py[q] * vx) + two_omega_n2 * py[q]) * c.inv_n4;
    v[q] = mk(p1, p2);
  }
  int b6 = b;
  asm volatile("" : "+v"(b6));
  if constexpr (Shape<N>::RL == 4) {
    // mirrored last pass: the unpack needs no exchange through LDS
    fft_canon<N, -1, RW, 1, 0, true, !LAST>(v, lds + r, b6, tw, tl);
    int b7 = b;
    asm volatile("" : "+v"(b7));
    mirror_store<N>(v, b7, kp, Q, Q + fs, y);
  } else {
    fft_canon<N, -1, RW>(v, lds + r, b6, tw, tl);
#pragma unroll
    for (int q = 0; q < 8; ++q) lds[pad<N>(b + T * q) * RW + r] = v[q];
    __syncthreads();
    int b7 = b;
    asm volatile("" : "+v"(b7));
    unpack_store<N, RW, N / 3 + 1>(lds + r, b7, nc, kp, Q, y, Q + fs, y);
  }
}

// tick (ensembles): a step counter the launch advances (block 0, thread 0);
// the launch is the first of a time step, and the step's readers of the
// counter (the column kernels) run after it.
__device__ __forceinline__ void row_tick(int *tick) {
  if (tick && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) tick[0] = tick[0] + 1;
}

template <int N>
__global__ __launch_bounds__(N / 8 * RowStep<N>::ROWS) void k_step_row(
    const cplx *__restrict__ M, cplx *__restrict__ Q, StepConst c, unsigned long long *maxbits,
    const cplx *__restrict__ tw, size_t mstride, const double *__restrict__ om_tab, int *tick) {
  constexpr int RW = RowStep<N>::ROWS;
  __shared__ cplx lds[RW * Shape<N>::LDSN];
  row_tick(tick);
  const int r = threadIdx.x % RW, b = threadIdx.x / RW;
  const int y = xcd_group<RowStep<N>::G>(blockIdx.x) * RW + r;
  M += blockIdx.y * mstride;
  Q += blockIdx.y * mstride;
  const double two_omega_n2 = om_tab ? om_tab[blockIdx.y] : c.two_omega_n2;
  step_row_body<N, true>(M, Q, c, maxbits, tw, two_omega_n2, y, b, r, lds);
}


// Row kernel with the three field pairs in parallel, for small grids (n <=
// 512), where a launch is latency-bound: its chain is one load -> inverse
// transform -> products -> forward transform instead of three loads and
// inverse transforms in turn.  Thread group g = threadIdx.x / (T RW) loads
// field pair g (psi, w = -k^2 psi, vz) of the RW rows, runs its inverse
// transform in LDS region g and publishes the result there; every group then
// forms the products from the three regions and runs the forward transform
// in its own region (the same arithmetic in every group, so no thread waits
// at a barrier the others skip); group 0 stores.  The arithmetic is
// k_step_row's (results agree within the rounding of multiply-add fusion).
template <int N>
__device__ __forceinline__ void row3_body(const cplx *__restrict__ M, cplx *__restrict__ Q,
                                          const StepConst &c, unsigned long long *maxbits,
                                          const cplx *__restrict__ tw, size_t mstride,
                                          const double *__restrict__ om_tab, int lb, cplx *lds) {
  constexpr int T = N / 8, RW = RowStep<N>::ROWS, GT = T * RW, LN = RW * Shape<N>::LDSN;
  const int g = threadIdx.x / GT, t = threadIdx.x - g * GT;
  const int r = t % RW, b = t / RW;
  const int y = lb * RW + r;
  M += blockIdx.y * mstride;
  Q += blockIdx.y * mstride;
  const double two_omega_n2 = om_tab ? om_tab[blockIdx.y] : c.two_omega_n2;
  const int nc = c.g.nc, kp = c.g.kp;
  const size_t fs = c.fs;
  cplx *P = lds + g * LN + r;
  cplx v[8];
  int b0 = b;
  asm volatile("" : "+v"(b0));
  load_packed_step<N>(v, M + (size_t)(2 * g) * fs, M + (size_t)(2 * g + 1) * fs, y, b0, nc, kp);
  int b1 = b;
  asm volatile("" : "+v"(b1));
  fft_canon<N, +1, RW>(v, P, b1, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) P[pad<N>(b + T * q) * RW] = v[q];
  __syncthreads();
  double m = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int o = pad<N>(b + T * q) * RW + r;
    const cplx z1 = lds[o], z2 = lds[LN + o], z3 = lds[2 * LN + o];
    const double px = z1.x, py = z1.y, vx = z3.x, vy = z3.y;
    m = fmax(m, px * px + py * py);
    const double j1 = px * z2.y - py * z2.x;
    const double p1 = (-j1 + two_omega_n2 * vy) * c.inv_n4;
    const double p2 = ((px * vy - py * vx) + two_omega_n2 * py) * c.inv_n4;
    v[q] = mk(p1, p2);
  }
  if (maxbits) {  // every group holds the same values: a redundant, exact max
    for (int off = 32; off > 0; off >>= 1)
      if (off < (int)blockDim.x) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
  }
  __syncthreads();  // every region read before the forward transforms reuse them
  int b6 = b;
  asm volatile("" : "+v"(b6));
  fft_canon<N, -1, RW>(v, P, b6, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) P[pad<N>(b + T * q) * RW] = v[q];
  __syncthreads();
  if (g == 0) {
    int b7 = b;
    asm volatile("" : "+v"(b7));
    unpack_store<N, RW, N / 3 + 1>(P, b7, nc, kp, Q, y, Q + fs, y);
  }
}
template <int N>
__global__ __launch_bounds__(3 * (N / 8) * RowStep<N>::ROWS) void k_step_row3(
    const cplx *__restrict__ M, cplx *__restrict__ Q, StepConst c, unsigned long long *maxbits,
    const cplx *__restrict__ tw, size_t mstride, const double *__restrict__ om_tab, int *tick) {
  __shared__ cplx lds[3 * RowStep<N>::ROWS * Shape<N>::LDSN];
  row_tick(tick);
  row3_body<N>(M, Q, c, maxbits, tw, mstride, om_tab, xcd_group<RowStep<N>::G>(blockIdx.x), lds);
}

// Row kernel of a stage: the parallel-pair form for small grids when the
// plan's row split allows it (-1: automatic, n <= 512; 1: always; 0: never),
// else the non-persistent kernel.  blockIdx.y: the run (count runs).  cfr:
// the closure forcing's row pass rides along (k_step_row3_cf; only where the
// parallel-pair form runs, row3_takes_cf).
struct CfRowArgs {
  const float *z;  // network output, fp32 [y][x][16]
  cplx *mix;       // the plan's mixed fields 0, 1
  double s0, s1;
  int ncf, kpf, member, nrb;  // nrb: the row launch's own blocks
};
template <int N>
static void launch_row_np(const r2d3c_plan *p, const cplx *M, cplx *Q, const StepConst &c,
                          unsigned long long *macc, const cplx *tw, size_t ms, const double *om,
                          int count, hipStream_t s, hipEvent_t ea, hipEvent_t eb,
                          int *tick = nullptr, const CfRowArgs *cfr = nullptr);


// Persistent variant of the row kernel for n >= 256.  A block walks its y
// tiles; the inputs of each packed transform (a field pair of RW rows) are
// copied straight into LDS with global_load_lds_dwordx4 while the previous
// transform runs, so the block never waits on HBM between its four transforms.
// LDS: RW*LDSN complex exchange slots + one staging buffer of 2*nc*RW.
// RWQ (rows per block, 1 or 2) and the grid size are schedule choices: they
// change neither the arithmetic nor its order, only the thread and LDS
// layout.  The grid size leaves the compiled kernel unchanged, so it gives
// bit-identical results; the RWQ variants are compiled separately and agree
// within rounding (multiply-add fusion may differ).
template <int N, int RWQ> struct RowV2 {
  static constexpr int T = N / 8;
  static constexpr int NCMAX = N / 3 + 2;
  static constexpr int fits(int rw) {
    return ((2 * NCMAX * rw + 63) / 64 * 64 + rw * Shape<N>::LDSN) * 16 <= 160 * 1024 &&
           T * rw <= 1024;
  }
  static constexpr int RW = (RWQ == 2 && fits(2)) ? 2 : 1;
  static constexpr int THREADS = T * RW;
  static constexpr int SBN = (2 * NCMAX * RW + 63) / 64 * 64;
  static constexpr int G = 4 / RW;  // blocks per y tile
};

__device__ __forceinline__ void wait_vm() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Stage field pair p of the RW rows (tile t, sub-block sub) into sb as
// sb[f2*nc*RW + kx*RW + rr] = M_{2p+f2}(kx, 4t + sub*RW + rr).  Each wave
// instruction fills 64 consecutive slots (LDS address = wave base + lane*16).
template <int N, int RWQ>
__device__ __forceinline__ void stage_pair(cplx *sb, const cplx *__restrict__ M, size_t fs, int p,
                                           int t, int sub, int nc, int kp) {
  using RV = RowV2<N, RWQ>;
  constexpr int RW = RV::RW, NW = RV::THREADS / 64;
  // opaque copy of the thread index: keeps the per-lane index math below from
  // being hoisted out of the caller's tile loop (it would pin ~40 VGPRs)
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int wave = tid >> 6, lane = tid & 63;
  const int per = nc * RW, total = 2 * per;
  const cplx *base = M + (size_t)(2 * p) * fs + (size_t)t * (size_t)(kp * 4) + sub * RW;
  for (int w0 = wave * 64; w0 < total; w0 += NW * 64) {
    const int e = w0 + lane;
    if (e < total) {
      const int f2 = e >= per, rem = e - f2 * per;
      const int kx = rem / RW, rr = rem - kx * RW;
      const cplx *g = base + (size_t)f2 * fs + kx * 4 + rr;
      __builtin_amdgcn_global_load_lds((const void *)g,
                                       (__attribute__((address_space(3))) void *)(sb + w0), 16, 0,
                                       0);
    }
  }
}

// z = c2r(i kx A) + i c2r(B) of row r from the staged pair (see load_packed_step)
template <int N, int RWQ>
__device__ __forceinline__ void build_packed(cplx v[8], const cplx *sb, int r, int b, int nc) {
  constexpr int T = N / 8, RW = RowV2<N, RWQ>::RW;
  const int per = nc * RW;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = b + T * q;
    cplx z = mk(0.0, 0.0);
    const bool lo = k < nc, hi = k > N - nc;
    if (lo || hi) {
      const int kk = lo ? k : N - k;
      const int o = kk * RW + r;
      const cplx f = sb[o], B = sb[per + o];
      const cplx A = mk(-(double)kk * f.y, (double)kk * f.x);
      z = packed_z<N>(k, A, B, hi);
    }
    v[q] = z;
  }
}

// The persistent row kernel.  blockIdx.x walks the y tiles of one run;
// blockIdx.y selects the run of an ensemble (its mixed fields lie mstride
// elements after the previous run's, and om_tab, when given, holds each run's
// 2 Omega n^2), 0 for a single run.
template <int N, int RWQ>
__global__ __launch_bounds__((RowV2<N, RWQ>::THREADS)) void k_step_row2(
    const cplx *__restrict__ M, cplx *__restrict__ Q, StepConst c, unsigned long long *maxbits,
    const cplx *__restrict__ tw, size_t mstride, const double *__restrict__ om_tab, int *tick) {
  using RV = RowV2<N, RWQ>;
  constexpr int T = N / 8, RW = RV::RW, G = RV::G;
  __shared__ cplx xb[RW * Shape<N>::LDSN];
  __shared__ cplx sb[RV::SBN];
  row_tick(tick);
  const int tid = threadIdx.x;
  const int r = tid % RW, b = tid / RW;
  cplx *P = xb + r;
  M += blockIdx.y * mstride;
  Q += blockIdx.y * mstride;
  const double two_omega_n2 = om_tab ? om_tab[blockIdx.y] : c.two_omega_n2;
  const int nc = c.g.nc, kp = c.g.kp;
  const size_t fs = c.fs;
  // tile of iteration k: G consecutive slots of one XCD share a tile
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int tslots = (int)(gridDim.x >> 3) / G, sub = slot % G;
  auto tile_of = [&](int k) { return (k * tslots + slot / G) * 8 + xcd; };
  constexpr int NTILES = N / 4;
  int k = 0, t = tile_of(0);
  if (t >= NTILES) return;
  double m = 0.0;
  stage_pair<N, RWQ>(sb, M, fs, 0, t, sub, nc, kp);
  wait_vm();
  __syncthreads();
  for (;;) {
    const int y = 4 * t + sub * RW + r;
    cplx v[8];
    double px[8], py[8], j1[8];
    build_packed<N, RWQ>(v, sb, r, b, nc);
    __syncthreads();
    stage_pair<N, RWQ>(sb, M, fs, 1, t, sub, nc, kp);
    fft_canon<N, +1, RW>(v, P, b, tw);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      px[q] = v[q].x;
      py[q] = v[q].y;
      m = fmax(m, px[q] * px[q] + py[q] * py[q]);
    }
    wait_vm();
    __syncthreads();
    build_packed<N, RWQ>(v, sb, r, b, nc);
    __syncthreads();
    stage_pair<N, RWQ>(sb, M, fs, 2, t, sub, nc, kp);
    fft_canon<N, +1, RW>(v, P, b, tw);
#pragma unroll
    for (int q = 0; q < 8; ++q) j1[q] = px[q] * v[q].y - py[q] * v[q].x;
    wait_vm();
    __syncthreads();
    build_packed<N, RWQ>(v, sb, r, b, nc);
    __syncthreads();
    const int tn = tile_of(k + 1);
    if (tn < NTILES) stage_pair<N, RWQ>(sb, M, fs, 0, tn, sub, nc, kp);
    fft_canon<N, +1, RW>(v, P, b, tw);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double vx = v[q].x, vy = v[q].y;
      const double p1 = (-j1[q] + two_omega_n2 * vy) * c.inv_n4;
      const double p2 = ((px[q] * vy - py[q] * vx) + two_omega_n2 * py[q]) * c.inv_n4;
      v[q] = mk(p1, p2);
    }
    fft_canon<N, -1, RW>(v, P, b, tw);
#pragma unroll
    for (int q = 0; q < 8; ++q) P[pad<N>(b + T * q) * RW] = v[q];
    __syncthreads();
    unpack_store<N, RW, N / 3 + 1>(P, b, nc, kp, Q, y, Q + fs, y);
    if (tn >= NTILES) break;
    wait_vm();
    __syncthreads();
    t = tn;
    ++k;
  }
  if (maxbits) {
    for (int off = 32; off > 0; off >>= 1)
      if (off < (int)blockDim.x) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
  }
}

// Fused column kernel of stage o: column forward of Q (psi blocks: Q0, vz
// blocks: Q1), the semi-implicit update of src/Rot2D3C.para.f90:578-599 with
// the forcing, then, by `mode`:
//   FM_EMIT   the inverse column transforms of the updated fields (the next
//             stage's, or the next step's first stage's, mixed fields);
//   FM_STATE  the state write psi = C1, vz = C2 (src/Rot2D3C.para.f90:610-611),
//             every ky of the column, or with FM_DISK only the entries inside
//             the update mask (the rest is known to be zero already);
//   FM_ZERO   extra blocks zero the columns kx > K of both fields.
// The forward transform is always the same (DIF, leaving the interleaved order
// the inverse transforms take), so every path rounds alike: a run of chained
// steps equals single steps bit for bit.  The state goes back through one LDS
// exchange to the canonical order for coalesced stores.  Blocks with o == 4
// write dt and clear the CFL slot of the next step (mreset).
enum { FM_EMIT = 1, FM_STATE = 2, FM_DISK = 4, FM_ZERO = 8 };

// b: the thread's index in its group of N/8 threads; lds: the group's LDSN
// slots.  Split emission (nslices > 1, never with FM_STATE): slice z of a
// column repeats the forward transform and the update and emits only field
// z, so a small grid's blocks each run two transforms instead of up to five.
// With act false the group runs every transform and barrier on its (valid)
// column but stores nothing: a group of a block that has no column of its
// own keeps the block's barrier sequence.
// The closure forcing's column pass folded into the first stage's fused
// column kernel (latency-bound closure runs, colfused_body EARLY): mix holds
// the forcing's row transforms (channels 0 and 1 as mixed fields of pitch
// kpf, all columns); the kernel transforms its column as k_cf_col does,
// writes the column's forcing to fp or fz (slice 0) and uses it in the
// update.
struct CfFuse {
  const cplx *mix;
  const double *gain;
  cplx *fp, *fz;
  double kmin2;
  int kpf, ngain, member;
};

// Emission slices (nslices > 1): slice s emits field s only.  As separate
// blocks (GS false), slices past the column's field count return at once and
// the state is never written (every slice reads it); as thread groups of one
// block (GS true), every slice runs the same transforms and barriers (slices
// past the field count redo field s mod count without storing it), and slice
// 0 writes the state once the block barrier after the update has ordered it
// after every group's reads.
template <int N, bool CS = false, bool GS = false, bool EARLY = false>
__device__ __forceinline__ void colfused_body(const cplx *__restrict__ Q, cplx *psi, cplx *vz,
                                              cplx *__restrict__ M, const StepConst &c, int o,
                                              int mode, const unsigned long long *mread,
                                              unsigned long long *mreset, double *dt_out,
                                              const cplx *__restrict__ tw, int lb, int b,
                                              cplx *lds, int slice, int nslices, bool act,
                                              const CfFuse &cf = CfFuse{}, bool cfon = false) {
  auto pos = [](int b, int q) { return ilv<N, CS>(b, q); };
  // the folded closure forcing pass: early loads and the canonical order
  // (pos = b + T q); other instantiations ignore cfon
  constexpr bool CFOK = EARLY && (N < 512 || CS);
  constexpr int T = N / 8;
  const int nc = c.g.nc, ncp = (nc + 1) & ~1;
  int field, kx;
  if (lb < nc) { field = 0; kx = lb; }
  else if (lb >= ncp && lb < ncp + nc) { field = 1; kx = lb - ncp; }
  else if ((mode & FM_ZERO) && lb >= 2 * ncp && lb < 2 * ncp + (N / 2 + 1 - nc)) {
    if (slice) return;
    kx = nc + (lb - 2 * ncp);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      psi[(size_t)kx * N + b + T * q] = mk(0.0, 0.0);
      vz[(size_t)kx * N + b + T * q] = mk(0.0, 0.0);
    }
    return;
  } else return;
  if (!GS && nslices > 1 && slice >= (field == 0 ? 4 : 2)) return;  // no field left to emit

  const double dt = c.cfl_mode ? dt_from_max(c.dc, *mread) : c.dt_fixed;
  if (lb == 0 && b == 0 && o == 4 && slice == 0) {
    if (dt_out) *dt_out = dt;
    if (mreset) *mreset = 0ull;
  }

  // the last-pass twiddles of this thread, shared by the column's transforms
  const TwLast tls = tw_last<N>(tw, b);
  const TwLast *tl = col_twlast<N>() ? &tls : nullptr;
  cplx v[8];
  const cplx *Qf = Q + (size_t)field * c.fs;
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = Qf[tix(b + T * q, kx, c.g.kp)];
  cplx *state = field == 0 ? psi : vz;
  const cplx *fd = field == 0 ? c.fpd : c.fzd;
  // EARLY (latency-bound launches): the state and dense-forcing loads of the
  // update are issued with the column's loads, before the transform, so the
  // chain has one round trip to memory instead of two.  Otherwise they follow
  // it (about 50 fewer live registers: more waves per SIMD).
  cplx s0e[EARLY ? 8 : 1], fde[EARLY ? 8 : 1], cfv[EARLY ? 8 : 1];
  if constexpr (EARLY) {
    const cplx *cfm = (CFOK && cfon) ? cf.mix + (size_t)field * N * cf.kpf : nullptr;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int j = pos(b, q), ky = kyof(j, N);
      const double ka2 = (double)kx * kx + (double)ky * ky;
      const bool in = ka2 <= c.g.kmax && ka2 > c.g.tiny;
      s0e[q] = in ? state[(size_t)kx * N + j] : mk(0.0, 0.0);
      fde[q] = (in && fd && !(CFOK && cfon)) ? fd[(size_t)kx * N + j] : mk(0.0, 0.0);
      cfv[q] = (CFOK && cfon) ? cfm[tix(b + T * q, kx, cf.kpf)] : mk(0.0, 0.0);
    }
  }
  fft_dif<N, -1, 1, CS>(v, lds, b, tw, tl);  // v[q] = spectrum at ky index pos(b, q)
  if constexpr (CFOK) {
    if (cfon) {  // uniform over the block (one run per block)
      // the closure forcing of this column: k_cf_col's transform and arithmetic
      fft_canon<N, -1, 1>(cfv, lds, b, tw);
      cplx *dst = field == 0 ? cf.fp : cf.fz;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int j = b + T * q, ky = kyof(j, N);
        const double ka2 = (double)kx * kx + (double)ky * ky;
        cplx o = mk(0.0, 0.0);
        if (ka2 <= c.g.kmax && ka2 > c.g.tiny && ka2 >= cf.kmin2) {
          o = field == 0 ? mk(-cfv[q].x / ka2, -cfv[q].y / ka2) : cfv[q];
          const int shell = (int)(sqrt(ka2) + 0.5);
          if (shell < cf.ngain) o = mk(o.x * cf.gain[shell], o.y * cf.gain[shell]);
        }
        if (slice == 0 && act) dst[(size_t)kx * N + j] = o;
        fde[q] = o;
      }
    }
  }

  const ForceSet &fo = field == 0 ? c.fp : c.fz;
  // at most two forced modes per field (scalars: no private-memory array)
  const bool f0 = fo.nent > 0 && fo.e[0].i == kx, f1 = fo.nent > 1 && fo.e[1].i == kx;
  const cplx fv0 = f0 ? force_value(fo, 0, dt, c.n2) : mk(0.0, 0.0);
  const cplx fv1 = f1 ? force_value(fo, 1, dt, c.n2) : mk(0.0, 0.0);
  const int fj0 = f0 ? fo.e[0].j : -1, fj1 = f1 ? fo.e[1].j : -1;

  const double tmp1 = dt / (double)o;
  const double nu = field == 0 ? c.nu : c.nuv, hnu = field == 0 ? c.hnu : c.hnuv;
  const int nn = field == 0 ? c.nn : c.nnv, mm = field == 0 ? c.mm : c.mmv;
  // the state on the disc, in two groups of four slots: each group's loads
  // are issued before its arithmetic, so they are in flight together
  cplx s[8];
#pragma unroll
  for (int g = 0; g < 8; g += 4) {
    cplx s0v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = pos(b, g + u), ky = kyof(j, N);
      const double ka2 = (double)kx * kx + (double)ky * ky;
      s0v[u] = mk(0.0, 0.0);
      if constexpr (EARLY) s0v[u] = s0e[g + u];
      else if (ka2 <= c.g.kmax && ka2 > c.g.tiny) s0v[u] = state[(size_t)kx * N + j];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = g + u;
      const int j = pos(b, q), ky = kyof(j, N);
      const double ka2 = (double)kx * kx + (double)ky * ky;
      cplx cc = mk(0.0, 0.0);
      if (ka2 <= c.g.kmax && ka2 > c.g.tiny) {
        const double tmp2 = 1.0 / ka2;
        const double tmp3 = 1.0 + (nu * ipow(ka2, nn) + hnu * ipow(tmp2, mm)) * tmp1;
        cplx f = j == fj0 ? fv0 : (j == fj1 ? fv1 : mk(0.0, 0.0));
        if (fd) f = cadd(f, EARLY ? fde[q] : fd[(size_t)kx * N + j]);
        const cplx s0 = s0v[u];
        cplx t = field == 0 ? cscale(v[q], tmp2) : v[q];
        t = cadd(t, f);
        t = cadd(s0, cscale(t, tmp1));
        cc = mk(t.x / tmp3, t.y / tmp3);
      }
      s[q] = cc;
    }
    __builtin_amdgcn_sched_barrier(0);  // the next group's loads stay below
  }
  if (mode & FM_STATE) {
    // interleaved -> canonical through LDS, then coalesced stores
#pragma unroll
    for (int q = 0; q < 8; ++q) lds[pad<N>(pos(b, q))] = s[q];
    __syncthreads();
    const bool disk = (mode & FM_DISK) != 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int j = b + T * q, ky = kyof(j, N);
      const double ka2 = (double)kx * kx + (double)ky * ky;
      if (act && slice == 0 && (!disk || (ka2 <= c.g.kmax && ka2 > c.g.tiny)))
        state[(size_t)kx * N + j] = lds[pad<N>(j)];
    }
    __syncthreads();
  }
  {
    if (!(mode & FM_EMIT)) return;
    const int nf = field == 0 ? 4 : 2;
    cplx *dst0 = M + (size_t)(field == 0 ? 0 : 4) * c.fs;
    for (int f = 0; f < nf; ++f) {
      if (nslices > 1 && f != (GS ? slice % nf : slice)) continue;
      int bb = b;  // opaque thread index, as in k_step_colinv
      asm volatile("" : "+v"(bb));
      cplx w[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int ky = kyof(ilv<N, CS>(bb, q), N);
        const double ka2 = (double)kx * kx + (double)ky * ky;
        cplx a = s[q];
        // products rounded on their own, as in colinv_body
        if (f >= 2) a = mk(mul_rn(-ka2, a.x), mul_rn(-ka2, a.y));
        if (f & 1) a = mk(mul_rn(-(double)ky, a.y), mul_rn((double)ky, a.x));
        w[q] = a;
      }
      fft_dit<N, +1, 1, CS>(w, lds, bb, tw, tl);
      cplx *dst = dst0 + (size_t)f * c.fs;
      if (act && (!GS || slice < nf)) {
#pragma unroll
        for (int q = 0; q < 8; ++q) dst[tix(bb + T * q, kx, c.g.kp)] = w[q];
      }
    }
  }
}

template <int N, bool CS = false, bool EARLY = false>
__global__ __launch_bounds__(N / 8) void k_step_colfused(
    const cplx *__restrict__ Q, cplx *psi, cplx *vz, cplx *__restrict__ M, StepConst c, int o,
    int mode, const unsigned long long *mread, unsigned long long *mreset, double *dt_out,
    const cplx *__restrict__ tw) {
  __shared__ cplx lds[Shape<N>::LDSN];
  colfused_body<N, CS, false, EARLY>(Q, psi, vz, M, c, o, mode, mread, mreset, dt_out, tw,
                       xcd_pairs(blockIdx.x), threadIdx.x, lds, blockIdx.z, gridDim.z, true);
}

// ---------------------------------------------------------------------------
// ensembles: many independent runs of one size per launch (blockIdx.y = run)
// ---------------------------------------------------------------------------
// One run of an ensemble: its constants (parameters, fixed dt, geometry; the
// forcing of each step comes from the forcing table), its state and its
// mixed-field workspace (8 fields: M0..M5, then the bracket fields Q0, Q1).
struct EnsMember {
  StepConst c;
  cplx *psi, *vz, *M;
};
struct EnsArgs {
  const EnsMember *mem;   // the runs of this launch's size class
  const ForceSet *ftab;   // [step][run][field], NULL = no forcing
  const int *step;        // current step of the forcing table
  int stride, base;       // runs per table step, index of this class's first run
  CfFuse cf;              // folded closure forcing pass of run cf.member (mix NULL: none)
};
__device__ __forceinline__ StepConst ens_const(const EnsArgs &a, int m) {
  StepConst c = a.mem[m].c;
  if (a.ftab) {
    const ForceSet *f = a.ftab + ((size_t)(*a.step) * a.stride + a.base + m) * 2;
    c.fp = f[0];
    c.fz = f[1];
  }
  return c;
}

template <int N, bool CS = false>
__global__ __launch_bounds__(N / 8) void k_ens_colinv(EnsArgs a, const cplx *__restrict__ tw) {
  const EnsMember &e = a.mem[blockIdx.y];
  const StepConst c = ens_const(a, blockIdx.y);
  colinv_body<N, CS>(e.psi, e.vz, e.M, c, nullptr, tw, xcd_pairs(blockIdx.x));
}
template <int N, bool CS = false, bool EARLY = false>
__global__ __launch_bounds__(N / 8) void k_ens_colfused(EnsArgs a, int o, int mode,
                                                         const cplx *__restrict__ tw) {
  const EnsMember &e = a.mem[blockIdx.y];
  const StepConst c = ens_const(a, blockIdx.y);
  __shared__ cplx lds[Shape<N>::LDSN];
  colfused_body<N, CS, false, EARLY>(e.M + 6 * c.fs, e.psi, e.vz, e.M, c, o, mode, nullptr, nullptr, nullptr, tw,
                       xcd_pairs(blockIdx.x), threadIdx.x, lds, blockIdx.z, gridDim.z, true, a.cf,
                       a.cf.mix && (int)blockIdx.y == a.cf.member);
}
// Emission slices as NS thread groups of one block (colfused_body, GS):
// one column per block, each group emitting one field; unlike slices as
// separate blocks this also splits the state-writing last stage.  For
// latency-bound launches (a few small runs).
template <int N, bool CS, int NS>
__global__ __launch_bounds__(NS * (N / 8)) void k_ens_colfused_s(EnsArgs a, int o, int mode,
                                                                 const cplx *__restrict__ tw) {
  constexpr int T = N / 8;
  const EnsMember &e = a.mem[blockIdx.y];
  const StepConst c = ens_const(a, blockIdx.y);
  __shared__ cplx lds[NS * Shape<N>::LDSN];
  const int g = threadIdx.x / T, b = threadIdx.x - g * T;
  colfused_body<N, CS, true, (N <= R2D_COL_EARLY_MAXN)>(
      e.M + 6 * c.fs, e.psi, e.vz, e.M, c, o, mode, nullptr, nullptr, nullptr, tw,
      xcd_pairs(blockIdx.x), b, lds + g * Shape<N>::LDSN, g, NS, true, a.cf,
      a.cf.mix && (int)blockIdx.y == a.cf.member);
}
// The same for n < 512, whose n/8-thread column groups fill less than a
// wavefront: GC neighbouring columns per block (GC n/8 <= 64 threads), group g
// = threadIdx.x / (n/8) taking logical column blockIdx.x GC + g.  GC divides
// ncp = nc rounded up to even, so a block's groups are all psi columns, all vz
// columns or all zeroed columns and follow one barrier sequence; a group past
// nc in its range runs its field's last column without storing (act false).
// Each column's arithmetic is k_ens_colfused's, so the results are the same
// bits.
template <int N, bool CS, int GC, bool EARLY = false>
__global__ __launch_bounds__(GC * (N / 8)) void k_ens_colfused_g(EnsArgs a, int o, int mode,
                                                                 const cplx *__restrict__ tw) {
  constexpr int T = N / 8;
  static_assert(GC * T <= 64, "one wavefront per block");
  const EnsMember &e = a.mem[blockIdx.y];
  const StepConst c = ens_const(a, blockIdx.y);
  __shared__ cplx lds[GC * Shape<N>::LDSN];
  const int g = threadIdx.x / T, b = threadIdx.x - g * T;
  const int nc = c.g.nc, ncp = (nc + 1) & ~1;
  const int lg = xcd_pairs(blockIdx.x) * GC + g;
  int lb = lg;
  bool act = true;
  if (lg < ncp) {
    act = lg < nc;
    lb = act ? lg : nc - 1;
  } else if (lg < 2 * ncp) {
    act = lg - ncp < nc;
    lb = act ? lg : ncp + nc - 1;
  }
  colfused_body<N, CS, false, EARLY>(e.M + 6 * c.fs, e.psi, e.vz, e.M, c, o, mode, nullptr, nullptr, nullptr, tw,
                       lb, b, lds + g * Shape<N>::LDSN, blockIdx.z, gridDim.z, act,
                       a.cf, a.cf.mix && (int)blockIdx.y == a.cf.member);
}
// ---------------------------------------------------------------------------
// ensembles: one persistent launch per RK4 stage for every size class
// ---------------------------------------------------------------------------
// The launch's blocks (kPsT threads, all resident: the grid is at most the
// occupancy limit) walk a work-list of units.  Row phase: a unit is G = 1024/n
// row pairs of one run, one group of (n/8)*2 threads each (step_row_body);
// then a grid barrier; column phase: a unit is G = 2048/n columns of one
// field of one run, one group of n/8 threads each (colfused_body).  Every
// unit holds 2048 points, so the units of all sizes cost about the same, and
// every group of a block runs the same size's body, so the block's barriers
// line up.  The arithmetic is the per-size kernels' (row: k_step_row; columns:
// the radix-8 Stockham form for n >= 512), so each run equals the same run
// advanced by those kernels, bit for bit.
constexpr int kPsT = 256;
constexpr int kPsMaxClasses = 5;           // sizes 64..1024
// the block's LDS, passed to the unit bodies as an LDS-space pointer
typedef __attribute__((address_space(3))) cplx *PsLds;
constexpr unsigned kPsSpin = 1u << 22;     // barrier polls before the time-out
// complex slots: the largest of G * 2 * LDSN (row units) and G * LDSN (column
// units) over the sizes 64..1024
template <int N> constexpr int ps_lds() {
  constexpr int row = kPsT / ((N / 8) * 2) * 2 * Shape<N>::LDSN, col = kPsT / (N / 8) * Shape<N>::LDSN;
  return row > col ? row : col;
}
constexpr int ps_lds_max(int a, int b) { return a > b ? a : b; }
constexpr int kPsLds = ps_lds_max(ps_lds_max(ps_lds<64>(), ps_lds<128>()),
                                  ps_lds_max(ps_lds_max(ps_lds<256>(), ps_lds<512>()), ps_lds<1024>()));
struct PsClass {
  EnsArgs a;            // the class's runs, forcing table and step
  const cplx *tw;       // twiddles of size n
  const double *om;     // 2 Omega n^2 per run
  int n, count, ucol;   // size, runs, column units per (run, field)
  int row_end, col_end; // units of the classes up to and including this one
};

// Grid-wide barrier of a launch whose blocks are all resident.  bar[0] counts
// arrivals, bar[1] is the generation: the last block to arrive resets the
// count and advances the generation.  Every thread releases its writes at
// device (agent) scope before the barrier and acquires after it, which
// writes back and invalidates the XCDs' non-coherent L2s.  The wait is
// bounded: after `spin` polls a block records the time-out in *err and goes
// on, so the launch always ends (its results are then invalid; later launches
// see *err and do nothing, and r2d3c_ensemble_sync reports it).
__device__ __forceinline__ void ps_grid_sync(unsigned *bar, int *err, unsigned spin) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned gen = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived =
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (arrived == gridDim.x) {
      __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(bar + 1, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned k = 0;
      while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(2);
        if (++k > spin) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}


template <int N>
__device__ __forceinline__ void ps_row_unit(const PsClass *k, int u, int tid, PsLds lds3) {
  cplx *lds = (cplx *)lds3;
  constexpr int RW = RowStep<N>::ROWS, GT = (N / 8) * RW, G = kPsT / GT, UR = (N / 2) / G;
  static_assert(RW == 2 && G * GT == kPsT && G * RW * Shape<N>::LDSN <= kPsLds, "row unit");
  const int m = u / UR, up = u - m * UR;
  const int g = tid / GT, t = tid - g * GT;
  const int r = t % RW, b = t / RW;
  const EnsMember *e = k->a.mem + m;
  const StepConst &c = e->c;
  step_row_body<N>(e->M, e->M + 6 * c.fs, c, nullptr, k->tw, k->om[m], (up * G + g) * RW + r, b,
                   r, lds + g * RW * Shape<N>::LDSN);
}

template <int N>
__device__ __forceinline__ void ps_col_unit(const PsClass *k, int u, int o, int mode, int tid, PsLds lds3) {
  cplx *lds = (cplx *)lds3;
  constexpr int GT = N / 8, G = kPsT / GT;
  static_assert(G * GT == kPsT && G * Shape<N>::LDSN <= kPsLds, "column unit");
  const int ucol = k->ucol, per_run = 2 * ucol;
  const int m = u / per_run, rem = u - m * per_run;
  const int field = rem >= ucol, j = rem - field * ucol;
  const int g = tid / GT, b = tid - g * GT;
  const EnsMember *e = k->a.mem + m;
  const StepConst c = ens_const(k->a, m);
  const int nc = c.g.nc, ncp = (nc + 1) & ~1;
  const int kx = j * G + g;
  const bool act = kx < nc;
  const int lb = (act ? kx : nc - 1) + (field ? ncp : 0);
  colfused_body<N, (N >= 512)>(e->M + 6 * c.fs, e->psi, e->vz, e->M, c, o, mode, nullptr,
                                nullptr, nullptr, k->tw, lb, b, lds + g * Shape<N>::LDSN, 0, 1,
                                act);
}

// Logical block of hardware block hb: pairs (2p, 2p+1) on one XCD in
// consecutive slots (the grid is a multiple of 16), so the two units of a
// 4-row tile at n = 1024 share that XCD's L2.
__device__ __forceinline__ int ps_logical_block(int hb) {
  const int x = hb & 7, s = hb >> 3;
  return ((s >> 1) << 4) + (x << 1) + (s & 1);
}

__global__ __launch_bounds__(kPsT) __attribute__((amdgpu_waves_per_eu(3))) void k_ens_stage(
    const PsClass *__restrict__ cls, int nk, unsigned *bar, int *err, unsigned spin, int o,
    int mode) {
  __shared__ cplx lds_arr[kPsLds];
  const PsLds lds = (PsLds)lds_arr;
  if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const int lb = ps_logical_block(blockIdx.x);
  const int row_total = cls[nk - 1].row_end, col_total = cls[nk - 1].col_end;
  for (int u = lb; u < row_total; u += gridDim.x) {
    int ci = 0;
    while (cls[ci].row_end <= u) ++ci;
    const PsClass *k = cls + ci;
    const int v = u - (ci ? cls[ci - 1].row_end : 0);
    // an opaque copy of the thread index per unit: no body's index math is
    // hoisted out of the loop (ten bodies' worth would not fit the registers)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    switch (k->n) {
      case 64: ps_row_unit<64>(k, v, tid, lds); break;
      case 128: ps_row_unit<128>(k, v, tid, lds); break;
      case 256: ps_row_unit<256>(k, v, tid, lds); break;
      case 512: ps_row_unit<512>(k, v, tid, lds); break;
      case 1024: ps_row_unit<1024>(k, v, tid, lds); break;
      default: break;
    }
  }
  ps_grid_sync(bar, err, spin);
  for (int u = lb; u < col_total; u += gridDim.x) {
    int ci = 0;
    while (cls[ci].col_end <= u) ++ci;
    const PsClass *k = cls + ci;
    const int v = u - (ci ? cls[ci - 1].col_end : 0);
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    switch (k->n) {
      case 64: ps_col_unit<64>(k, v, o, mode, tid, lds); break;
      case 128: ps_col_unit<128>(k, v, o, mode, tid, lds); break;
      case 256: ps_col_unit<256>(k, v, o, mode, tid, lds); break;
      case 512: ps_col_unit<512>(k, v, o, mode, tid, lds); break;
      case 1024: ps_col_unit<1024>(k, v, o, mode, tid, lds); break;
      default: break;
    }
  }
}

// The same work-list as k_ens_stage, one unit per block and one launch per
// phase (the kernel boundary is the barrier): rows of every size, then
// columns of every size.
__global__ __launch_bounds__(kPsT) __attribute__((amdgpu_waves_per_eu(3))) void k_ens_rows(
    const PsClass *__restrict__ cls, int nk) {
  __shared__ cplx lds_arr[kPsLds];
  const PsLds lds = (PsLds)lds_arr;
  const int u = ps_logical_block(blockIdx.x);
  if (u >= cls[nk - 1].row_end) return;
  int ci = 0;
  while (cls[ci].row_end <= u) ++ci;
  const PsClass *k = cls + ci;
  const int v = u - (ci ? cls[ci - 1].row_end : 0), tid = threadIdx.x;
  switch (k->n) {
    case 64: ps_row_unit<64>(k, v, tid, lds); break;
    case 128: ps_row_unit<128>(k, v, tid, lds); break;
    case 256: ps_row_unit<256>(k, v, tid, lds); break;
    case 512: ps_row_unit<512>(k, v, tid, lds); break;
    case 1024: ps_row_unit<1024>(k, v, tid, lds); break;
    default: break;
  }
}
__global__ __launch_bounds__(kPsT) __attribute__((amdgpu_waves_per_eu(3))) void k_ens_cols(
    const PsClass *__restrict__ cls, int nk, int o, int mode) {
  __shared__ cplx lds_arr[kPsLds];
  const PsLds lds = (PsLds)lds_arr;
  const int u = ps_logical_block(blockIdx.x);
  if (u >= cls[nk - 1].col_end) return;
  int ci = 0;
  while (cls[ci].col_end <= u) ++ci;
  const PsClass *k = cls + ci;
  const int v = u - (ci ? cls[ci - 1].col_end : 0), tid = threadIdx.x;
  switch (k->n) {
    case 64: ps_col_unit<64>(k, v, o, mode, tid, lds); break;
    case 128: ps_col_unit<128>(k, v, o, mode, tid, lds); break;
    case 256: ps_col_unit<256>(k, v, o, mode, tid, lds); break;
    case 512: ps_col_unit<512>(k, v, o, mode, tid, lds); break;
    case 1024: ps_col_unit<1024>(k, v, o, mode, tid, lds); break;
    default: break;
  }
}

// advances the forcing-table step at the end of a time step
__global__ void k_ens_tick(int *step) { step[0] = step[0] + 1; }
__global__ void k_set_i32(int *p, int v) { p[0] = v; }

// ---------------------------------------------------------------------------
// host side: plan, dispatch, C ABI
// ---------------------------------------------------------------------------
struct r2d3c_plan {
  int n = 0, device = 0, cus = 256;
  Geo g{};
  size_t fs = 0, fsf = 0;         // complex elements per step / full mixed field
  cplx *tw = nullptr;             // n twiddles exp(-2 pi i t/n)
  cplx *mstep = nullptr;          // 6 mixed fields + 2 bracket fields (fused step)
  cplx *mfull = nullptr;          // 4 full-range mixed fields
  double *real = nullptr;         // n*n real scratch
  double *real2 = nullptr;        // second n*n real scratch (extrema)
  double *real3 = nullptr;        // third n*n real scratch (n = 16384 brackets)
  cplx *spec3 = nullptr;          // three spectral scratch fields (transfers)
  double *dpart = nullptr;        // diagnostics partials
  cplx *ci = nullptr;             // iflow = 2: C1..C6, fpre, fp, fz
  float *ci_rtab = nullptr;       // const_inj random values per mode
  double ci_key[3] = {-1.0, -1.0, -1.0};  // (kdn, kup, nprocs) of ci_rtab
  std::vector<float> ci_host;
  double *ci_efp = nullptr;
  unsigned long long *maxbits = nullptr;
  double *dtslot = nullptr;
  double *partial = nullptr;      // reduction partials
  long long bytes = 0;
  // schedule of the row kernel (r2d3c_plan_set_schedule / r2d3c_autotune):
  // rows per block and blocks per CU of the persistent kernel, or bpc = 0
  // for the non-persistent kernel (2 rows per block)
  int row_rw = 2, row_bpc = 0;
  // the non-persistent row kernel's form for n <= 512: -1 automatic (the
  // parallel field pairs of k_step_row3), 0 sequential (k_step_row), 1 parallel
  int row_split = -1;
  // radix schedule of the column transforms for n >= 512: 0 the four-step
  // form, 1 radix-8 Stockham passes (fft_canon)
  int col_cs = 0;
  // split emission of the fused column kernel's stages 1-3 (-1: when the
  // launch holds fewer threads than one wave per SIMD; 0 off; 1 on)
  int col_split = -1;
  // n = 16384 step: 1 the fused bracket pass (k_big_row_jac2), 0 stage by
  // stage through six real fields
  int big_fused = 1;
  // kernel timing of r2d3c_step (r2d3c_timing_enable / r2d3c_timing_read)
  int timing = 0;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  struct Rec { int kind; hipEvent_t a, b; };
  std::vector<Rec> recs;
};

// Start/stop events of one timed launch (null events when timing is off);
// hipExtLaunchKernelGGL stamps them at the kernel's own start and end.
static void ev_pair(r2d3c_plan *p, int kind, hipEvent_t *a, hipEvent_t *b) {
  *a = *b = nullptr;
  if (!p->timing) return;
  while (p->evused + 2 > p->evpool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    p->evpool.push_back(e);
  }
  *a = p->evpool[p->evused++];
  *b = p->evpool[p->evused++];
  p->recs.push_back({kind, *a, *b});
}

static thread_local std::string g_err;
static int set_err(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
// for the library's other translation units (closure.hip)
int r2d3c_internal_set_error(int code, const char *msg) { return set_err(code, msg); }
static int hip_err(hipError_t e, const char *where) {
  return set_err(-(R2D3C_EHIP_BASE + (int)e), std::string(where) + ": " + hipGetErrorString(e));
}
#define HIPCHK(expr)                                  \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hip_err(e_, #expr);  \
  } while (0)
#define LAUNCHCHK(where)                              \
  do {                                                \
    hipError_t e_ = hipGetLastError();                \
    if (e_ != hipSuccess) return hip_err(e_, where);  \
  } while (0)

static constexpr int kReduceBlocks = 256;

static int launch_grid(size_t total) {
  size_t g = (total + 255) / 256;
  return (int)(g > 2048 ? 2048 : (g == 0 ? 1 : g));
}


#define R2D_DISPATCH(n, FN, ...)                                   \
  switch (n) {                                                     \
    case 16: return FN<16>(__VA_ARGS__);                           \
    case 32: return FN<32>(__VA_ARGS__);                           \
    case 64: return FN<64>(__VA_ARGS__);                           \
    case 128: return FN<128>(__VA_ARGS__);                         \
    case 256: return FN<256>(__VA_ARGS__);                         \
    case 512: return FN<512>(__VA_ARGS__);                         \
    case 1024: return FN<1024>(__VA_ARGS__);                       \
    case 2048: return FN<2048>(__VA_ARGS__);                       \
    case 4096: return FN<4096>(__VA_ARGS__);                       \
    default: return set_err(R2D3C_ENOTSUP, "unsupported grid size"); \
  }

template <int N> static int run_c2r(r2d3c_plan *p, const cplx *in, double *out, hipStream_t s) {
  k_col_inv_full<N><<<N / 2 + 1, N / 8, 0, s>>>(in, p->mfull, OP_RAW, 1.0, p->g, p->tw);
  k_row_c2r<N><<<N / RowFull<N>::ROWS, N / 8 * RowFull<N>::P, 0, s>>>(p->mfull, out, p->g, p->tw);
  LAUNCHCHK("c2r");
  return 0;
}

template <int N> static int run_r2c(r2d3c_plan *p, const double *in, cplx *out, hipStream_t s) {
  k_row_r2c<N><<<N / RowFull<N>::ROWS, N / 8 * RowFull<N>::P, 0, s>>>(in, p->mfull, p->g, p->tw);
  k_col_fwd_full<N><<<N / 2 + 1, N / 8, 0, s>>>(p->mfull, out, p->g, p->tw);
  LAUNCHCHK("r2c");
  return 0;
}

template <int N>
static int run_poisson(r2d3c_plan *p, const cplx *a, const cplx *b, cplx *c, hipStream_t s) {
  cplx *m0 = p->mfull, *m1 = m0 + p->fsf, *m2 = m1 + p->fsf, *m3 = m2 + p->fsf;
  k_col_inv_full<N><<<N / 2 + 1, N / 8, 0, s>>>(a, m0, OP_DX, 1.0, p->g, p->tw);
  k_col_inv_full<N><<<N / 2 + 1, N / 8, 0, s>>>(a, m1, OP_DY, 1.0, p->g, p->tw);
  k_col_inv_full<N><<<N / 2 + 1, N / 8, 0, s>>>(b, m2, OP_DX, 1.0, p->g, p->tw);
  k_col_inv_full<N><<<N / 2 + 1, N / 8, 0, s>>>(b, m3, OP_DY, 1.0, p->g, p->tw);
  const double n4 = (double)N * N * (double)N * N;
  k_row_bracket<N><<<N / RowStep<N>::ROWS, N / 8 * RowStep<N>::ROWS, 0, s>>>(m0, m1, m2, m3,
                                                                            p->real, n4, p->g,
                                                                            p->tw);
  LAUNCHCHK("poisson");
  return run_r2c<N>(p, p->real, c, s);
}

static DtConst make_dt_const(int n, const r2d3c_params *q, double cfl) {
  DtConst d;
  d.cfl = cfl;
  d.kcut = (double)n / 3.0;
  d.nrm = (double)n * (double)n;
  d.a = q->nu * ipow(d.kcut, 2 * q->nn);
  d.bq = q->nuv * ipow(d.kcut, 2 * q->nnv);
  d.two_omega = 2 * q->omega;
  return d;
}

template <int N>
static int run_cfl(r2d3c_plan *p, const r2d3c_params *q, double cfl, const cplx *psi, double *d_dt,
                   hipStream_t s) {
  cplx *m0 = p->mfull, *m1 = m0 + p->fsf;
  k_set_u64<<<1, 1, 0, s>>>(p->maxbits, 0ull);
  k_col_inv_full<N><<<N / 2 + 1, N / 8, 0, s>>>(psi, m0, OP_DX, 1.0, p->g, p->tw);
  k_col_inv_full<N><<<N / 2 + 1, N / 8, 0, s>>>(psi, m1, OP_DY, 1.0, p->g, p->tw);
  k_row_gradmax<N><<<N / RowStep<N>::ROWS, N / 8 * RowStep<N>::ROWS, 0, s>>>(m0, m1, p->maxbits,
                                                                            p->g, p->tw);
  k_dt_from_max<<<1, 1, 0, s>>>(p->maxbits, make_dt_const(N, q, cfl), d_dt);
  LAUNCHCHK("cfl");
  return 0;
}

// Placement of one rand_force draw, bug-compatible with
// src/pseudospec2D_Rot2D3C.f90:1100-1122: in columns i >= 1 the half-plane
// wavenumber ka(i) is compared with +-ky and the full index with kx; only the
// kx = 0 column (i = 1 in Fortran) gets a conjugate partner.
static int make_force_set(const r2d3c_force_draw *d, int n, ForceSet *f) {
  std::memset(f, 0, sizeof(*f));
  if (d == nullptr || !d->enabled) return 0;
  f->num = std::sqrt(2.0 * d->f0);
  f->kupk = d->kupk;
  const double cph = std::cos(d->phase), sph = std::sin(d->phase);
  auto ka = [n](int i) { return (double)(i < n / 2 ? i : i - n); };
  auto add = [&](int i, int j, double sgn) {
    if (f->nent >= 2) return false;
    f->e[f->nent++] = ForceEnt{i, j, sgn, cph, sph};
    return true;
  };
  for (int i = 1; i <= n / 2; ++i)
    if (ka(i) == d->ky || ka(i) == -d->ky)
      for (int j = 0; j < n; ++j)
        if (ka(j) == d->kx && !add(i, j, 1.0))
          return set_err(R2D3C_EINVAL, "forcing placement overflow");
  if (ka(0) == d->ky)
    for (int j = 1; j < n / 2; ++j)
      if (ka(j) == d->kx && !(add(0, j, 1.0) && add(0, n - j, -1.0)))
        return set_err(R2D3C_EINVAL, "forcing placement overflow");
  return 0;
}

// Persistent row grid: bpc blocks per CU, a multiple of 8 G (the tile map),
// at most one block per tile slot.
// gridDim.z of a fused column launch of stages 1-3 holding `threads` threads:
// 4 emission slices when split (a latency-bound launch), else 1
static int col_slices(const r2d3c_plan *p, long long threads) {
  if (p->col_split == 0) return 1;
  if (p->col_split == 1) return 4;
  return threads < (long long)p->cus * 256 ? 4 : 1;
}

template <int N, int RWQ> static int row2_grid(int cus, int bpc) {
  constexpr int G = RowV2<N, RWQ>::G, Q8 = 8 * G;
  int g = cus * bpc / Q8 * Q8;
  const int need = (N / 4) * G;
  if (g > need) g = need;
  if (g < Q8) g = Q8;
  return g;
}

static int make_step_const(r2d3c_plan *p, const r2d3c_params *q, int dt_mode, double dt_or_cfl,
                           StepConst *c) {
  std::memset(c, 0, sizeof(*c));
  const int n = p->n;
  c->g = p->g;
  c->fs = p->fs;
  c->nu = q->nu; c->hnu = q->hnu; c->nuv = q->nuv; c->hnuv = q->hnuv;
  c->nn = q->nn; c->mm = q->mm; c->nnv = q->nnv; c->mmv = q->mmv;
  c->two_omega = 2 * q->omega;
  c->cfl_mode = dt_mode == R2D3C_DT_CFL;
  c->dt_fixed = dt_or_cfl;
  c->dc = make_dt_const(n, q, dt_or_cfl);
  c->n2 = (double)n * (double)n;
  c->two_omega_n2 = c->two_omega * c->n2;
  c->inv_n4 = 1.0 / (c->n2 * c->n2);
  return 0;
}

template <int N> static bool row3_takes_cf(const r2d3c_plan *p) {
  return N <= 512 && p->row_split != 0;
}
template <int N>
static void launch_row_np_impl(const r2d3c_plan *p, const cplx *M, cplx *Q, const StepConst &c,
                               unsigned long long *macc, const cplx *tw, size_t ms,
                               const double *om, int count, hipStream_t s, hipEvent_t ea,
                               hipEvent_t eb, int *tick) {
  constexpr int RW = RowStep<N>::ROWS;
  if (N <= 512 && p->row_split != 0)
    hipExtLaunchKernelGGL(k_step_row3<(N <= 512 ? N : 512)>, dim3(N / RW, count),
                          dim3(3 * (N / 8) * RW), 0, s, ea, eb, 0, M, Q, c, macc, tw, ms, om,
                          tick);
  else
    hipExtLaunchKernelGGL(k_step_row<N>, dim3(N / RW, count), dim3(N / 8 * RW), 0, s, ea, eb, 0,
                          M, Q, c, macc, tw, ms, om, tick);
}

// Kernels of one time step.  first: the state's column transforms are made
// here (otherwise the previous step's last stage left them in M); chain: the
// last stage also emits the next step's column transforms; masked: the state
// outside the update mask is known to be zero (written by an earlier step of
// the same run), so only the disc is written back.  slot: CFL maximum slot of
// this step (steps alternate, each clears the other's at its first stage).
template <int N>
static int step_seq(r2d3c_plan *p, StepConst &c, cplx *psi, cplx *vz, hipStream_t s,
                    double *dt_out, bool first, bool chain, bool masked, int slot) {
  const int nc = p->g.nc, ncp = (nc + 1) & ~1;
  cplx *M = p->mstep, *Q = p->mstep + 6 * p->fs;
  constexpr int RW = RowStep<N>::ROWS;
  const cplx *tw = p->tw;
  unsigned long long *mb = p->maxbits + slot, *mother = p->maxbits + (slot ^ 1);
  hipEvent_t ea, eb;
  if (first) {
    ev_pair(p, 0, &ea, &eb);
    hipExtLaunchKernelGGL((N >= 512 && p->col_cs) ? k_step_colinv<N, true> : k_step_colinv<N, false>,
                          dim3(round16(ncp + nc)), dim3(N / 8), 0, s, ea, eb, 0,
                          (const cplx *)psi, (const cplx *)vz, M, c, mb, tw);
  }
  for (int o = 4; o >= 1; --o) {
    unsigned long long *macc = (o == 4 && c.cfl_mode) ? mb : nullptr;
    ev_pair(p, 1, &ea, &eb);
    if constexpr (RowV2<N, 2>::THREADS >= 64) {
      if (p->row_bpc == 0)
        launch_row_np<N>(p, M, Q, c, macc, tw, 0, nullptr, 1, s, ea, eb);
      else if (p->row_rw == 1 && RowV2<N, 1>::THREADS >= 64)
        hipExtLaunchKernelGGL(k_step_row2<N, 1>, dim3(row2_grid<N, 1>(p->cus, p->row_bpc)),
                              dim3(RowV2<N, 1>::THREADS), 0, s, ea, eb, 0, (const cplx *)M, Q, c,
                              macc, tw, (size_t)0, (const double *)nullptr, (int *)nullptr);
      else
        hipExtLaunchKernelGGL(k_step_row2<N, 2>, dim3(row2_grid<N, 2>(p->cus, p->row_bpc)),
                              dim3(RowV2<N, 2>::THREADS), 0, s, ea, eb, 0, (const cplx *)M, Q, c,
                              macc, tw, (size_t)0, (const double *)nullptr, (int *)nullptr);
    } else {
      launch_row_np<N>(p, M, Q, c, macc, tw, 0, nullptr, 1, s, ea, eb);
    }
    int mode = FM_EMIT;
    if (o == 1) mode = FM_STATE | (masked ? FM_DISK : FM_ZERO) | (chain ? FM_EMIT : 0);
    const int nblk = round16(2 * ncp + ((mode & FM_ZERO) ? (N / 2 + 1 - nc) : 0));
    unsigned long long *mres = o == 4 ? mother : nullptr;
    ev_pair(p, o == 1 ? 3 : 2, &ea, &eb);
    const int nsl = col_slices(p, (long long)nblk * (N / 8));
    const int nz = (mode & FM_STATE) ? 1 : nsl;
    // latency-bound (split) launches load the update's inputs early
    constexpr bool ESM = N <= R2D_COL_EARLY_MAXN;
    const bool cs = N >= 512 && p->col_cs, early = ESM && nsl > 1;
    hipExtLaunchKernelGGL(early ? (cs ? k_step_colfused<N, true, ESM> : k_step_colfused<N, false, ESM>)
                                : (cs ? k_step_colfused<N, true> : k_step_colfused<N, false>),
                          dim3(nblk, 1, nz), dim3(N / 8), 0, s, ea, eb, 0,
                          (const cplx *)Q, psi, vz, M, c, o, mode,
                          (const unsigned long long *)mb, mres, dt_out, tw);
  }
  LAUNCHCHK("step");
  return 0;
}

extern "C" float r2d3c_randu(int *idum);

// randu values of the band modes of const_inj for an nprocs-rank run: rank r
// owns the kx columns [lo, hi) of its slab (the split of `range`,
// src/fftp.f90:17-47) and restarts randu at seed r on every call, drawing in
// its loop order (src/pseudospec2D_Rot2D3C.f90:1032-1057).  Uploaded once per
// (kdn, kup, nprocs).
static int ci_table(r2d3c_plan *p, double kdn, double kup, int nprocs, hipStream_t s) {
  nprocs = nprocs < 1 ? 1 : nprocs;
  const int n = p->n, ncol = n / 2 + 1;
  if (nprocs > ncol) return set_err(R2D3C_EINVAL, "const_inj: more ranks than kx columns");
  if (p->ci_key[0] == kdn && p->ci_key[1] == kup && p->ci_key[2] == nprocs) return 0;
  p->ci_host.assign((size_t)ncol * n, 0.0f);
  const int w1 = ncol / nprocs, w2 = ncol % nprocs;
  for (int r = 0, lo = 0; r < nprocs; ++r) {
    const int hi = lo + w1 + (r < w2 ? 1 : 0);
    int seed = r;
    for (int i = lo; i < hi; ++i) {
      const double kx = i < n / 2 ? i : i - n;
      const int j0 = i == 0 ? 1 : 0, j1 = i == 0 ? n / 2 - 1 : n - 1;
      for (int j = j0; j <= j1; ++j) {
        const double ky = j < n / 2 ? j : j - n, ka2 = kx * kx + ky * ky;
        if (ka2 <= kup * kup && ka2 >= kdn * kdn) p->ci_host[(size_t)i * n + j] = r2d3c_randu(&seed);
      }
    }
    lo = hi;
  }
  HIPCHK(hipMemcpyAsync(p->ci_rtab, p->ci_host.data(), sizeof(float) * p->ci_host.size(),
                        hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));  // ci_host may be refilled by the next key
  p->ci_key[0] = kdn;
  p->ci_key[1] = kup;
  p->ci_key[2] = nprocs;
  return 0;
}

static int const_inj_dev(r2d3c_plan *p, const cplx *ps, double kdn, double kup, double fp0, int kin,
                         cplx *fpre, cplx *fp, hipStream_t s);

template <int N>
static int run_step_ci(r2d3c_plan *p, const r2d3c_params *q, cplx *psi, cplx *vz,
                       const r2d3c_forcing *fo, int dt_mode, double dt_or_cfl, double *d_dt,
                       hipStream_t s) {
  StepConst c;
  make_step_const(p, q, dt_mode, dt_or_cfl, &c);
  const size_t sz = (size_t)(N / 2 + 1) * N;
  cplx *C1 = p->ci, *C2 = C1 + sz, *C3 = C2 + sz, *C4 = C3 + sz, *C5 = C4 + sz, *C6 = C5 + sz;
  cplx *fpre = C6 + sz, *fp = fpre + sz, *fz = fp + sz;
  double *dtp = d_dt ? d_dt : p->dtslot;
  int rc;
  if (dt_mode == R2D3C_DT_CFL) {
    rc = r2d3c_cfl(p, q, dt_or_cfl, (const double *)psi, dtp, s);
  } else {
    k_set_f64<<<1, 1, 0, s>>>(dtp, dt_or_cfl);
    rc = 0;
  }
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(C1, psi, sizeof(cplx) * sz, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(C2, vz, sizeof(cplx) * sz, hipMemcpyDeviceToDevice, s));
  const int g = launch_grid(sz);
  for (int o = 4; o >= 1; --o) {
    if ((rc = const_inj_dev(p, C1, fo->ci_kdn, fo->ci_kup, fo->ci_fp0, 1, fpre, fp, s))) return rc;
    if ((rc = const_inj_dev(p, C2, fo->ci_kdn, fo->ci_kup, fo->ci_fz0, 0, fpre, fz, s))) return rc;
    k_spec_op<<<g, 256, 0, s>>>(C1, C4, OP_LAP, N, p->g.kmax, p->g.tiny);
    if ((rc = run_poisson<N>(p, C1, C4, C5, s))) return rc;
    if ((rc = run_poisson<N>(p, C1, C2, C6, s))) return rc;
    k_spec_op<<<g, 256, 0, s>>>(C1, C3, OP_DY, N, p->g.kmax, p->g.tiny);
    k_spec_op<<<g, 256, 0, s>>>(C2, C4, OP_DY, N, p->g.kmax, p->g.tiny);
    k_update<<<g, 256, 0, s>>>(psi, vz, C1, C2, C3, C4, C5, C6, fp, fz, c, dtp, o);
  }
  k_spec_op<<<g, 256, 0, s>>>(C1, psi, OP_MASK, N, p->g.kmax, p->g.tiny);
  k_spec_op<<<g, 256, 0, s>>>(C2, vz, OP_MASK, N, p->g.kmax, p->g.tiny);
  LAUNCHCHK("iflow=2 step");
  return 0;
}

template <int N>
static int run_steps(r2d3c_plan *p, const r2d3c_params *q, cplx *psi, cplx *vz,
                     const r2d3c_forcing *fo, int nsteps, int dt_mode, double dt_or_cfl,
                     double *d_dt, hipStream_t s) {
  if (fo) {  // one forcing kind per call: the iflow = 2 table is built from fo[0]
    for (int st = 1; st < nsteps; ++st)
      if (fo[st].ci_enabled != fo[0].ci_enabled ||
          (fo[0].ci_enabled && (fo[st].ci_kdn != fo[0].ci_kdn || fo[st].ci_kup != fo[0].ci_kup ||
                                fo[st].ci_nprocs != fo[0].ci_nprocs)))
        return set_err(R2D3C_EINVAL,
                       "every step of one call must have the same ci_enabled and, with "
                       "iflow = 2, the same ci_kdn, ci_kup and ci_nprocs");
  }
  if (fo && fo[0].ci_enabled) {  // iflow = 2 runs stage by stage
    if ((fo[0].ci_kdn > fo[0].ci_kup) || fo[0].ci_kup < 0.0)
      return set_err(R2D3C_EINVAL, "const_inj band must satisfy 0 <= kdn <= kup");
    int rc = ci_table(p, fo[0].ci_kdn, fo[0].ci_kup, fo[0].ci_nprocs, s);
    if (rc) return rc;
    for (int st = 0; st < nsteps; ++st) {
      rc = run_step_ci<N>(p, q, psi, vz, &fo[st], dt_mode, dt_or_cfl, d_dt ? d_dt + st : nullptr, s);
      if (rc) return rc;
    }
    return 0;
  }
  StepConst c;
  make_step_const(p, q, dt_mode, dt_or_cfl, &c);
  for (int st = 0; st < nsteps; ++st) {
    std::memset(&c.fp, 0, sizeof(c.fp));
    std::memset(&c.fz, 0, sizeof(c.fz));
    c.fpd = c.fzd = nullptr;
    if (fo) {
      int rc = make_force_set(&fo[st].psi, N, &c.fp);
      if (rc) return rc;
      rc = make_force_set(&fo[st].vz, N, &c.fz);
      if (rc) return rc;
      c.fpd = (const cplx *)fo[st].dense_psi;
      c.fzd = (const cplx *)fo[st].dense_vz;
    }
    double *dt_out = d_dt ? d_dt + st : p->dtslot;
    int rc = step_seq<N>(p, c, psi, vz, s, dt_out, st == 0, st + 1 < nsteps, st > 0, st & 1);
    if (rc) return rc;
  }
  return 0;
}


// n = 16384 (definitions at the end of this file)
constexpr int kBigN = 16384, kBigH = 8192, kBigT = 1024;
static int big_c2r(r2d3c_plan *p, const cplx *in, int op, double *out, hipStream_t s,
                   int ncols = kBigN / 2 + 1, cplx *mix = nullptr);
static int big_r2c(r2d3c_plan *p, const double *in, cplx *out, hipStream_t s);
static int big_poisson(r2d3c_plan *p, const cplx *a, const cplx *b, cplx *c, hipStream_t s);
static int big_cfl(r2d3c_plan *p, const r2d3c_params *q, double cfl, const cplx *psi, double *d_dt,
                   hipStream_t s);
static int big_step(r2d3c_plan *p, const r2d3c_params *q, cplx *psi, cplx *vz,
                    const r2d3c_forcing *fo, int dt_mode, double dt_or_cfl, double *d_dt,
                    hipStream_t s);

// Geometry of an n x n plan: dealiasing ka2 <= kmax = (n/3)^2
// (src/Rot2D3C.para.f90:213-214), the fused step's kept columns kx <= K and
// the field pitches.  Nonzero if the cutoff is not n/3 (never for a power of
// two; the row loaders rely on it).
static int plan_geometry(r2d3c_plan *p, int n) {
  const double kmax = ((double)n / 3.0) * ((double)n / 3.0);
  int K = 0;
  while ((double)(K + 1) * (K + 1) <= kmax) ++K;
  p->g.n = n;
  p->g.nc = K + 1;
  if (p->g.nc != n / 3 + 1) return 1;
  p->g.kp = (K + 2) & ~1;
  p->g.ncf = n / 2 + 1;
  p->g.kpf = n / 2 + 2;
  p->g.kmax = kmax;
  p->g.tiny = 0.000001;
  p->fs = (size_t)n * p->g.kp;
  p->fsf = (size_t)n * p->g.kpf;
  return 0;
}

// Twiddles exp(-2 pi i t/n), t < n, from the extended-precision circle,
// rounded once to double.
static void twiddles_host(int n, cplx *tw) {
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < n; ++t) {
    const long double a = twopi * (long double)t / (long double)n;
    tw[t] = make_double2((double)cosl(a), (double)(-sinl(a)));
  }
}

// ---- r2d3c_run_noalloc: the fused step on a static device scratch array ----
//
// Layout of `scratch` (every offset a multiple of 256 bytes): the twiddle
// tables of n = 16, 32, ..., kNaMaxN back to back; the CFL maximum slots and
// the dt slot (512 bytes); the fused step's 6 mixed + 2 bracket fields of the
// largest grid.  Each device holds its own copy for the life of the process.
constexpr int kNaMaxN = R2D3C_NOALLOC_MAX_N;
constexpr size_t kNaTwBytes = ((size_t)(2 * kNaMaxN - 16) * sizeof(cplx) + 255) & ~(size_t)255;
constexpr size_t kNaFlagBytes = 512;
constexpr size_t kNaMaxK = kNaMaxN / 3;  // K = n/3 for every power of two
constexpr size_t kNaMstepBytes = 8 * sizeof(cplx) * kNaMaxN * ((kNaMaxK + 2) & ~(size_t)1);
constexpr size_t kNaScratchBytes = kNaTwBytes + kNaFlagBytes + kNaMstepBytes;
static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

constexpr int kNaMaxDevices = 64;
struct NaDevice {
  unsigned char *ptr = nullptr;  // device address of this device's `scratch`
  int cus = 0;
  int last_n = 0;                // grid of the previous call (0: none yet)
};
static std::mutex g_na_mu;
static NaDevice g_na[kNaMaxDevices];

static size_t na_tw_offset(int n) {  // entries before n's table: 16 + ... + n/2
  return (size_t)(n - 16);
}

__global__ void k_zero_cplx(cplx *a, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = mk(0.0, 0.0);
}

// Set-up once per device, with that device current: the device address of
// `scratch` (the host-side name is not a device address) and the twiddle
// tables, copied synchronously.
static int na_device(int dev, NaDevice **out) {
  std::lock_guard<std::mutex> lock(g_na_mu);
  NaDevice &d = g_na[dev];
  if (!d.ptr) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    HIPCHK(hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, dev));
    std::vector<cplx> tw(2 * kNaMaxN - 16);
    for (int n = 16; n <= kNaMaxN; n *= 2) twiddles_host(n, tw.data() + na_tw_offset(n));
    HIPCHK(hipMemcpy(ptr, tw.data(), sizeof(cplx) * tw.size(), hipMemcpyHostToDevice));
    d.ptr = (unsigned char *)ptr;
  }
  *out = &d;
  return 0;
}

extern "C" {

const char *r2d3c_last_error(void) { return g_err.c_str(); }

int r2d3c_plan_create(r2d3c_plan **out, int n) {
  if (!out) return set_err(R2D3C_EINVAL, "plan pointer is NULL");
  *out = nullptr;
  if (n < 16 || (n > 4096 && n != 16384) || (n & (n - 1)))
    return set_err(R2D3C_ENOTSUP, "n must be a power of two in [16, 4096], or 16384");
  r2d3c_plan *p = new r2d3c_plan();
  p->n = n;
  HIPCHK(hipGetDevice(&p->device));
  HIPCHK(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device));
  if (plan_geometry(p, n)) { delete p; return set_err(R2D3C_EINVAL, "dealias cutoff"); }
  std::vector<cplx> tw(n);
  twiddles_host(n, tw.data());
  // n = 16384 steps stage by stage: no fused-step workspace, a third real field
  const bool big = n == 16384;
  size_t b_tw = sizeof(cplx) * n, b_ms = big ? 256 : sizeof(cplx) * 8 * p->fs,
         b_mf = sizeof(cplx) * 4 * p->fsf;
  size_t b_re = sizeof(double) * (size_t)n * n;
  hipError_t e;
  if ((e = hipMalloc(&p->tw, b_tw)) != hipSuccess || (e = hipMalloc(&p->mstep, b_ms)) != hipSuccess ||
      (e = hipMalloc(&p->mfull, b_mf)) != hipSuccess || (e = hipMalloc(&p->real, b_re)) != hipSuccess ||
      (e = hipMalloc(&p->maxbits, 256)) != hipSuccess || (e = hipMalloc(&p->dtslot, 256)) != hipSuccess ||
      (e = hipMalloc(&p->partial, sizeof(double) * kReduceBlocks)) != hipSuccess ||
      (e = hipMalloc(&p->real2, b_re)) != hipSuccess ||
      (e = hipMalloc(&p->spec3, 3 * sizeof(cplx) * (size_t)(n / 2 + 1) * n)) != hipSuccess ||
      (e = hipMalloc(&p->dpart, sizeof(double) * kCondSums * kReduceBlocks)) != hipSuccess ||
      (e = hipMalloc(&p->ci, 9 * sizeof(cplx) * (size_t)(n / 2 + 1) * n)) != hipSuccess ||
      (e = hipMalloc(&p->ci_rtab, sizeof(float) * (size_t)(n / 2 + 1) * n)) != hipSuccess ||
      (e = hipMalloc(&p->ci_efp, 256)) != hipSuccess ||
      (big && (e = hipMalloc(&p->real3, b_re)) != hipSuccess) ||
      (e = hipMemcpy(p->tw, tw.data(), b_tw, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemset(p->mstep, 0, b_ms)) != hipSuccess || (e = hipMemset(p->mfull, 0, b_mf)) != hipSuccess) {
    int rc = hip_err(e, "plan allocation");
    r2d3c_plan_destroy(p);
    return rc;
  }
  p->bytes = (long long)(b_tw + b_ms + b_mf + (big ? 3 : 2) * b_re + 512 + sizeof(double) * kReduceBlocks +
                         3 * sizeof(cplx) * (size_t)(n / 2 + 1) * n +
                         sizeof(double) * kCondSums * kReduceBlocks +
                         9 * sizeof(cplx) * (size_t)(n / 2 + 1) * n +
                         sizeof(float) * (size_t)(n / 2 + 1) * n + 256);
  *out = p;
  return 0;
}

int r2d3c_plan_destroy(r2d3c_plan *p) {
  if (!p) return 0;
  for (hipEvent_t e : p->evpool) (void)hipEventDestroy(e);
  (void)hipFree(p->tw);
  (void)hipFree(p->mstep);
  (void)hipFree(p->mfull);
  (void)hipFree(p->real);
  if (p->real3) (void)hipFree(p->real3);
  (void)hipFree(p->maxbits);
  (void)hipFree(p->dtslot);
  (void)hipFree(p->partial);
  (void)hipFree(p->real2);
  (void)hipFree(p->spec3);
  (void)hipFree(p->dpart);
  (void)hipFree(p->ci);
  (void)hipFree(p->ci_rtab);
  (void)hipFree(p->ci_efp);
  delete p;
  return 0;
}

long long r2d3c_plan_workspace_bytes(const r2d3c_plan *p) { return p ? p->bytes : 0; }

int r2d3c_timing_enable(r2d3c_plan *p, int enable) {
  if (!p) return set_err(R2D3C_EINVAL, "NULL plan");
  p->timing = enable != 0;
  p->evused = 0;
  p->recs.clear();
  return 0;
}

int r2d3c_timing_read(r2d3c_plan *p, double *ms, long long *launches) {
  if (!p || !ms || !launches) return set_err(R2D3C_EINVAL, "NULL argument");
  for (int k = 0; k < R2D3C_NKERNELS; ++k) { ms[k] = 0.0; launches[k] = 0; }
  for (const auto &r : p->recs) {
    HIPCHK(hipEventSynchronize(r.b));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
    ms[r.kind] += t;
    launches[r.kind] += 1;
  }
  p->evused = 0;
  p->recs.clear();
  return 0;
}
int r2d3c_plan_size(const r2d3c_plan *p) { return p ? p->n : 0; }

// Streaming copy that measures the achievable HBM bandwidth: 16-byte loads
// and stores, four loads in flight per thread before their stores,
// grid-stride over 16 workgroups per CU.
__global__ __launch_bounds__(256) void k_copy_probe(const cplx *__restrict__ src,
                                                    cplx *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * 1024;
  for (size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x; i < n; i += stride) {
    cplx v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = i + 256 * u < n ? src[i + 256 * u] : mk(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i + 256 * u < n) dst[i + 256 * u] = v[u];
  }
}

int r2d3c_copy_probe(void *dst, const void *src, long long bytes, void *stream) {
  if (!dst || !src) return set_err(R2D3C_EINVAL, "NULL argument");
  if (bytes < 0 || bytes % 16) return set_err(R2D3C_EINVAL, "bytes: a non-negative multiple of 16");
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return set_err(R2D3C_EINVAL, "pointers: 16-byte aligned");
  const size_t n = (size_t)bytes / 16;
  if (!n) return 0;
  int dev = 0, cus = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const size_t blocks = std::min((n + 1023) / 1024, (size_t)cus * 16);
  k_copy_probe<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>((const cplx *)src, (cplx *)dst, n);
  LAUNCHCHK("copy probe");
  return 0;
}

#define NEED(x) \
  if (!(x)) return set_err(R2D3C_EINVAL, "NULL argument: " #x)

int r2d3c_c2r(r2d3c_plan *p, const double *in, double *out, void *stream) {
  NEED(p); NEED(in); NEED(out);
  if (p->n == kBigN) return big_c2r(p, (const cplx *)in, OP_RAW, out, (hipStream_t)stream);
  R2D_DISPATCH(p->n, run_c2r, p, (const cplx *)in, out, (hipStream_t)stream);
}

int r2d3c_r2c(r2d3c_plan *p, const double *in, double *out, void *stream) {
  NEED(p); NEED(in); NEED(out);
  if (p->n == kBigN) return big_r2c(p, in, (cplx *)out, (hipStream_t)stream);
  R2D_DISPATCH(p->n, run_r2c, p, in, (cplx *)out, (hipStream_t)stream);
}

static int spec_op(r2d3c_plan *p, const double *a, double *b, int op, void *stream) {
  NEED(p); NEED(a); NEED(b);
  const size_t total = (size_t)(p->n / 2 + 1) * p->n;
  k_spec_op<<<launch_grid(total), 256, 0, (hipStream_t)stream>>>((const cplx *)a, (cplx *)b, op, p->n,
                                                                   p->g.kmax, p->g.tiny);
  LAUNCHCHK("spectral operator");
  return 0;
}

int r2d3c_derivk2(r2d3c_plan *p, const double *a, double *b, int dir, void *stream) {
  return spec_op(p, a, b, dir == 1 ? OP_DX : (dir == 2 ? OP_DY : OP_MASK), stream);
}
int r2d3c_laplak2(r2d3c_plan *p, const double *a, double *b, void *stream) {
  return spec_op(p, a, b, OP_LAP, stream);
}
int r2d3c_inv_laplak2(r2d3c_plan *p, const double *a, double *b, void *stream) {
  return spec_op(p, a, b, OP_INVLAP, stream);
}

int r2d3c_poisson(r2d3c_plan *p, const double *a, const double *b, double *c, void *stream) {
  NEED(p); NEED(a); NEED(b); NEED(c);
  if (p->n == kBigN)
    return big_poisson(p, (const cplx *)a, (const cplx *)b, (cplx *)c, (hipStream_t)stream);
  R2D_DISPATCH(p->n, run_poisson, p, (const cplx *)a, (const cplx *)b, (cplx *)c,
               (hipStream_t)stream);
}

static int reduce(r2d3c_plan *p, const double *a, const double *b, int mode, int kin, double *out,
                  void *stream) {
  hipStream_t s = (hipStream_t)stream;
  k_reduce_partial<<<kReduceBlocks, 256, 0, s>>>((const cplx *)a, (const cplx *)b, mode, kin, p->n,
                                                  p->g.kmax, p->g.tiny, p->partial);
  k_reduce_finish<<<1, 256, 0, s>>>(p->partial, kReduceBlocks, out);
  LAUNCHCHK("reduction");
  return 0;
}

int r2d3c_energy(r2d3c_plan *p, const double *a, int kin, double *d_out, void *stream) {
  NEED(p); NEED(a); NEED(d_out);
  return reduce(p, a, a, 0, kin, d_out, stream);
}

int r2d3c_inerprod(r2d3c_plan *p, const double *a, const double *b, int kin, double *d_out,
                   void *stream) {
  NEED(p); NEED(a); NEED(b); NEED(d_out);
  return reduce(p, a, b, 1, kin, d_out, stream);
}

static int check_params(const r2d3c_params *q) {
  NEED(q);
  if (q->nn < 0 || q->mm < 0 || q->nnv < 0 || q->mmv < 0)
    return set_err(R2D3C_EINVAL, "dissipation orders must be >= 0");
  return 0;
}

int r2d3c_cfl(r2d3c_plan *p, const r2d3c_params *q, double cfl, const double *psi, double *d_dt,
              void *stream) {
  NEED(p); NEED(psi); NEED(d_dt);
  int rc = check_params(q);
  if (rc) return rc;
  if (p->n == kBigN) return big_cfl(p, q, cfl, (const cplx *)psi, d_dt, (hipStream_t)stream);
  R2D_DISPATCH(p->n, run_cfl, p, q, cfl, (const cplx *)psi, d_dt, (hipStream_t)stream);
}

int r2d3c_step(r2d3c_plan *p, const r2d3c_params *q, double *psi, double *vz,
               const r2d3c_forcing *forcing, int dt_mode, double dt_or_cfl, double *d_dt,
               void *stream) {
  NEED(p); NEED(psi); NEED(vz);
  int rc = check_params(q);
  if (rc) return rc;
  if (dt_mode != R2D3C_DT_CFL && dt_mode != R2D3C_DT_FIXED)
    return set_err(R2D3C_EINVAL, "dt_mode must be R2D3C_DT_CFL or R2D3C_DT_FIXED");
  if (psi == vz) return set_err(R2D3C_EINVAL, "psi and vz must not alias");
  if (p->n == kBigN) return r2d3c_run(p, q, psi, vz, forcing, 1, dt_mode, dt_or_cfl, d_dt, stream);
  R2D_DISPATCH(p->n, run_steps, p, q, (cplx *)psi, (cplx *)vz, forcing, 1, dt_mode, dt_or_cfl,
               d_dt, (hipStream_t)stream);
}

int r2d3c_run(r2d3c_plan *p, const r2d3c_params *q, double *psi, double *vz,
              const r2d3c_forcing *forcing, int n_steps, int dt_mode, double dt_or_cfl,
              double *d_dt, void *stream) {
  NEED(p); NEED(psi); NEED(vz);
  int rc = check_params(q);
  if (rc) return rc;
  if (dt_mode != R2D3C_DT_CFL && dt_mode != R2D3C_DT_FIXED)
    return set_err(R2D3C_EINVAL, "dt_mode must be R2D3C_DT_CFL or R2D3C_DT_FIXED");
  if (psi == vz) return set_err(R2D3C_EINVAL, "psi and vz must not alias");
  if (n_steps < 0) return set_err(R2D3C_EINVAL, "n_steps must be >= 0");
  if (n_steps == 0) return 0;
  if (p->n == kBigN) {
    for (int st = 0; st < n_steps; ++st) {
      const r2d3c_forcing *f = forcing ? forcing + st : nullptr;
      if (f && f->ci_enabled) return set_err(R2D3C_ENOTSUP, "iflow = 2 is not supported at n = 16384");
      rc = big_step(p, q, (cplx *)psi, (cplx *)vz, f, dt_mode, dt_or_cfl, d_dt ? d_dt + st : nullptr,
                    (hipStream_t)stream);
      if (rc) return rc;
    }
    return 0;
  }
  R2D_DISPATCH(p->n, run_steps, p, q, (cplx *)psi, (cplx *)vz, forcing, n_steps, dt_mode,
               dt_or_cfl, d_dt, (hipStream_t)stream);
}

int r2d3c_run_noalloc(int n, const r2d3c_params *q, double *psi, double *vz,
                      const r2d3c_forcing *forcing, int n_steps, int dt_mode, double dt_or_cfl,
                      double *d_dt, void *stream) {
  NEED(psi); NEED(vz);
  if (n < 16 || n > kNaMaxN || (n & (n - 1)))
    return set_err(R2D3C_EINVAL, "r2d3c_run_noalloc: n must be a power of two in [16, " +
                                     std::to_string(kNaMaxN) + "]");
  int rc = check_params(q);
  if (rc) return rc;
  if (dt_mode != R2D3C_DT_CFL && dt_mode != R2D3C_DT_FIXED)
    return set_err(R2D3C_EINVAL, "dt_mode must be R2D3C_DT_CFL or R2D3C_DT_FIXED");
  if (psi == vz) return set_err(R2D3C_EINVAL, "psi and vz must not alias");
  if (n_steps < 0) return set_err(R2D3C_EINVAL, "n_steps must be >= 0");
  if (forcing)
    for (int st = 0; st < n_steps; ++st)
      if (forcing[st].ci_enabled)
        return set_err(R2D3C_ENOTSUP, "r2d3c_run_noalloc: iflow = 2 needs a plan (r2d3c_run)");
  if (n_steps == 0) return 0;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) return set_err(R2D3C_EINVAL, "device index out of range");
  NaDevice *d = nullptr;
  if ((rc = na_device(dev, &d))) return rc;
  // a plan over the scratch array: the same kernels and launch shapes as a
  // plan of r2d3c_plan_create with the default schedules
  r2d3c_plan pl;
  pl.n = n;
  pl.device = dev;
  pl.cus = d->cus;
  if (plan_geometry(&pl, n)) return set_err(R2D3C_EINVAL, "dealias cutoff");
  pl.tw = (cplx *)d->ptr + na_tw_offset(n);
  pl.maxbits = (unsigned long long *)(d->ptr + kNaTwBytes);
  pl.dtslot = (double *)(d->ptr + kNaTwBytes + 256);
  pl.mstep = (cplx *)(d->ptr + kNaTwBytes + kNaFlagBytes);
  hipStream_t s = (hipStream_t)stream;
  {
    // a new plan's fields start zeroed; after a call at another n the
    // array holds that grid's layout, so clear the fields this grid uses
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (d->last_n != n) {
      k_zero_cplx<<<1024, 256, 0, s>>>(pl.mstep, 8 * pl.fs);
      LAUNCHCHK("r2d3c_run_noalloc clear");
      d->last_n = n;
    }
  }
  switch (n) {
    case 16: return run_steps<16>(&pl, q, (cplx *)psi, (cplx *)vz, forcing, n_steps, dt_mode, dt_or_cfl, d_dt, s);
    case 32: return run_steps<32>(&pl, q, (cplx *)psi, (cplx *)vz, forcing, n_steps, dt_mode, dt_or_cfl, d_dt, s);
    case 64: return run_steps<64>(&pl, q, (cplx *)psi, (cplx *)vz, forcing, n_steps, dt_mode, dt_or_cfl, d_dt, s);
    case 128: return run_steps<128>(&pl, q, (cplx *)psi, (cplx *)vz, forcing, n_steps, dt_mode, dt_or_cfl, d_dt, s);
    case 256: return run_steps<256>(&pl, q, (cplx *)psi, (cplx *)vz, forcing, n_steps, dt_mode, dt_or_cfl, d_dt, s);
    case 512: return run_steps<512>(&pl, q, (cplx *)psi, (cplx *)vz, forcing, n_steps, dt_mode, dt_or_cfl, d_dt, s);
    case 1024: return run_steps<1024>(&pl, q, (cplx *)psi, (cplx *)vz, forcing, n_steps, dt_mode, dt_or_cfl, d_dt, s);
    default: return set_err(R2D3C_ENOTSUP, "unsupported grid size");
  }
}

int r2d3c_forcing_fill(r2d3c_plan *p, const r2d3c_forcing *fo, double dt, double *fp, double *fz,
                       void *stream) {
  NEED(p); NEED(fo);
  if (!(dt > 0.0)) return set_err(R2D3C_EINVAL, "dt must be positive");
  const size_t total = (size_t)(p->n / 2 + 1) * p->n;
  ForceSet a, b;
  int rc = make_force_set(&fo->psi, p->n, &a);
  if (rc) return rc;
  rc = make_force_set(&fo->vz, p->n, &b);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (fp) k_force_fill<<<launch_grid(total), 256, 0, s>>>((cplx *)fp, a, dt, p->n);
  if (fz) k_force_fill<<<launch_grid(total), 256, 0, s>>>((cplx *)fz, b, dt, p->n);
  LAUNCHCHK("forcing fill");
  return 0;
}

int r2d3c_cond_check(r2d3c_plan *p, const r2d3c_params *q, double kup, const double *psi,
                     const double *vz, const double *fp, const double *fz, int nprocs,
                     double *d_out, void *stream) {
  NEED(p); NEED(psi); NEED(vz); NEED(fp); NEED(fz); NEED(d_out);
  int rc = check_params(q);
  if (rc) return rc;
  nprocs = nprocs < 1 ? 1 : nprocs;
  if (nprocs > p->n) return set_err(R2D3C_EINVAL, "cond_check: more ranks than rows");
  if (nprocs > kReduceBlocks) return set_err(R2D3C_ENOTSUP, "cond_check: at most 256 ranks");
  hipStream_t s = (hipStream_t)stream;
  const int n = p->n;
  CondConst c;
  c.n = n; c.nn = q->nn; c.mm = q->mm; c.nnv = q->nnv; c.mmv = q->mmv;
  c.kmax = p->g.kmax; c.tiny = p->g.tiny;
  c.band_lo = kup * kup;
  c.band_hi = (double)2.01f * kup * kup;  // `(2.01)*kup*kup`: a default-REAL literal (:559)
  k_cond_partial<<<kReduceBlocks, 256, 0, s>>>((const cplx *)psi, (const cplx *)vz, (const cplx *)fp,
                                               (const cplx *)fz, c, p->dpart);
  k_cond_finish<<<1, 32, 0, s>>>(p->dpart, kReduceBlocks, d_out);
  LAUNCHCHK("cond_check sums");
  // extrema of c2r(laplak2(ps)) (-omega*n^2) and c2r(derivk2(vz,0)) (:573-586)
  const size_t sz = (size_t)(n / 2 + 1) * n, nn2 = (size_t)n * n;
  cplx *t = p->spec3;
  k_spec_op<<<launch_grid(sz), 256, 0, s>>>((const cplx *)psi, t, OP_LAP, n, p->g.kmax, p->g.tiny);
  rc = r2d3c_c2r(p, (const double *)t, p->real2, stream);
  if (rc) return rc;
  auto extrema = [&](double *out) {
    if (nprocs == 1) {
      k_minmax_partial<<<kReduceBlocks, 256, 0, s>>>(p->real2, nn2, p->dpart);
      k_minmax_finish<<<1, 1, 0, s>>>(p->dpart, kReduceBlocks, out);
    } else {
      k_minmax_slabs<<<nprocs, 256, 0, s>>>(p->real2, n, p->dpart);
      k_minmax_finish_ranks<<<1, 1, 0, s>>>(p->dpart, nprocs, out);
    }
  };
  extrema(d_out + kCondSums);
  k_spec_op<<<launch_grid(sz), 256, 0, s>>>((const cplx *)vz, t, OP_MASK, n, p->g.kmax, p->g.tiny);
  rc = r2d3c_c2r(p, (const double *)t, p->real2, stream);
  if (rc) return rc;
  extrema(d_out + kCondSums + 2);
  LAUNCHCHK("cond_check extrema");
  return 0;
}

}  // extern "C"

static int const_inj_dev(r2d3c_plan *p, const cplx *ps, double kdn, double kup, double fp0, int kin,
                         cplx *fpre, cplx *fp, hipStream_t s) {
  const int n = p->n;
  const size_t sz = (size_t)(n / 2 + 1) * n;
  const int g = launch_grid(sz);
  k_ci_pre<<<g, 256, 0, s>>>(ps, fpre, kdn * kdn, kup * kup, n);
  int rc = r2d3c_inerprod(p, (const double *)ps, (const double *)fpre, kin, p->ci_efp, s);
  if (rc) return rc;
  k_ci_post<<<g, 256, 0, s>>>(ps, fpre, p->ci_rtab, p->ci_efp, fp0, kdn * kdn, kup * kup, fp, n);
  LAUNCHCHK("const_inj");
  return 0;
}

extern "C" {

int r2d3c_const_inj(r2d3c_plan *p, const double *ps, double kdn, double kup, double fp0, int kin,
                    int nprocs, double *fp, void *stream) {
  NEED(p); NEED(ps); NEED(fp);
  if (kdn > kup || kup < 0.0) return set_err(R2D3C_EINVAL, "const_inj band must satisfy 0 <= kdn <= kup");
  hipStream_t s = (hipStream_t)stream;
  int rc = ci_table(p, kdn, kup, nprocs, s);
  if (rc) return rc;
  const size_t sz = (size_t)(p->n / 2 + 1) * p->n;
  return const_inj_dev(p, (const cplx *)ps, kdn, kup, fp0, kin, p->ci + 6 * sz, (cplx *)fp, s);
}

int r2d3c_const_inj_last(r2d3c_plan *p, double *fp, double *fz, void *stream) {
  NEED(p);
  const size_t sz = (size_t)(p->n / 2 + 1) * p->n;
  hipStream_t s = (hipStream_t)stream;
  if (fp) HIPCHK(hipMemcpyAsync(fp, p->ci + 7 * sz, sizeof(cplx) * sz, hipMemcpyDeviceToDevice, s));
  if (fz) HIPCHK(hipMemcpyAsync(fz, p->ci + 8 * sz, sizeof(cplx) * sz, hipMemcpyDeviceToDevice, s));
  return 0;
}

int r2d3c_spectrum(r2d3c_plan *p, const double *psi, const double *vz, double *d_out, void *stream) {
  NEED(p); NEED(psi); NEED(vz); NEED(d_out);
  k_shells<<<p->n / 2 + 1, 256, 0, (hipStream_t)stream>>>((const cplx *)psi, (const cplx *)vz, nullptr,
                                                          nullptr, 0, p->n, d_out);
  LAUNCHCHK("spectrum");
  return 0;
}

int r2d3c_transfers(r2d3c_plan *p, const double *psi, const double *vz, double *d_out,
                    void *stream) {
  NEED(p); NEED(psi); NEED(vz); NEED(d_out);
  hipStream_t s = (hipStream_t)stream;
  const int n = p->n;
  const size_t sz = (size_t)(n / 2 + 1) * n;
  cplx *c1 = p->spec3, *c2 = c1 + sz, *w = c2 + sz;
  // c1 = laplak2(ps); c2 = poisson(ps, c1); c1 = poisson(ps, vz)  (:759-761)
  k_spec_op<<<launch_grid(sz), 256, 0, s>>>((const cplx *)psi, w, OP_LAP, n, p->g.kmax, p->g.tiny);
  int rc = r2d3c_poisson(p, psi, (const double *)w, (double *)c2, stream);
  if (rc) return rc;
  rc = r2d3c_poisson(p, psi, vz, (double *)c1, stream);
  if (rc) return rc;
  k_shells<<<n / 2 + 1, 256, 0, s>>>((const cplx *)psi, (const cplx *)vz, c1, c2, 1, n, d_out);
  LAUNCHCHK("transfers");
  return 0;
}

// randu: src/pseudospec2D_mod.f90:52-71 (REAL results, am = 1./im in REAL)
float r2d3c_randu(int *idum) {
  const int iq = 127773, ir = 2836, mask = 123459876, ia = 16807, im = 2147483647;
  const float am = 1.0f / (float)im;
  int v = *idum ^ mask;
  const int k = v / iq;
  v = ia * (v - k * iq) - ir * k;
  if (v < 0) v += im;
  volatile float r = am * (float)v;  // no contraction with the next line
  r = (r - 0.5f) * 2.0f;
  *idum = v ^ mask;
  return r;
}

void r2d3c_rand_force_draw(double kup, int *seed, double *kx, double *ky, double *phase) {
  const double pi = 3.1415926535897932384;  // src/pseudospec2D_mod.f90:35
  const double theta = (double)r2d3c_randu(seed) * pi;
  int tmp = *seed + 1;
  const double ph = (double)r2d3c_randu(&tmp) * pi;
  *kx = std::floor(kup * std::cos(theta) + 0.5);
  *ky = std::floor(kup * std::sin(theta) + 0.5);
  *phase = ph;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// schedules and their autotuner
// ---------------------------------------------------------------------------
static bool schedule_valid(int n, int rw, int bpc) {
  // bpc = 0: the non-persistent row kernel, one block per 2 rows
  if (bpc == 0) return rw == 2;
  if (bpc < 1 || bpc > 4 || (rw != 1 && rw != 2)) return false;
  // one row per block needs n/8 >= 64 threads (whole waves)
  return rw == 2 || n >= 512;
}

template <int N>
static int tune_steps(r2d3c_plan *p, cplx *psi, cplx *vz, int reps, hipStream_t s, float *ms) {
  r2d3c_params q{1e-3, 0.1, 1e-3, 0.1, 1.0, 1, 2, 1, 2};
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  // best of three trials of `reps` chained steps, after two warm-up steps
  int rc = run_steps<N>(p, &q, psi, vz, nullptr, 2, R2D3C_DT_FIXED, 1e-6, nullptr, s);
  for (int trial = 0; trial < 3 && !rc; ++trial) {
    HIPCHK(hipEventRecord(a, s));
    rc = run_steps<N>(p, &q, psi, vz, nullptr, reps, R2D3C_DT_FIXED, 1e-6, nullptr, s);
    HIPCHK(hipEventRecord(b, s));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    if (trial == 0 || t < *ms) *ms = t;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

extern "C" {

int r2d3c_plan_set_schedule(r2d3c_plan *p, int row_rows, int row_blocks_per_cu) {
  NEED(p);
  if (!schedule_valid(p->n, row_rows, row_blocks_per_cu))
    return set_err(R2D3C_EINVAL, "schedule: rows per block 1 (n >= 512) or 2, 1..4 blocks per CU");
  p->row_rw = row_rows;
  p->row_bpc = row_blocks_per_cu;
  return 0;
}

int r2d3c_plan_set_column_split(r2d3c_plan *p, int mode) {
  NEED(p);
  if (mode < -1 || mode > 1) return set_err(R2D3C_EINVAL, "column split: -1, 0 or 1");
  p->col_split = mode;
  return 0;
}

int r2d3c_plan_set_row_split(r2d3c_plan *p, int mode) {
  NEED(p);
  if (mode < -1 || mode > 1) return set_err(R2D3C_EINVAL, "row split: -1, 0 or 1");
  p->row_split = mode;
  return 0;
}

int r2d3c_plan_get_row_split(const r2d3c_plan *p, int *mode) {
  NEED(p); NEED(mode);
  *mode = p->row_split;
  return 0;
}

int r2d3c_plan_set_bracket_pass(r2d3c_plan *p, int fused) {
  NEED(p);
  if (fused != 0 && fused != 1) return set_err(R2D3C_EINVAL, "bracket pass: 0 or 1");
  p->big_fused = fused;
  return 0;
}

int r2d3c_plan_get_bracket_pass(const r2d3c_plan *p, int *fused) {
  NEED(p); NEED(fused);
  *fused = p->big_fused;
  return 0;
}

int r2d3c_plan_get_column_split(const r2d3c_plan *p, int *mode) {
  NEED(p); NEED(mode);
  *mode = p->col_split;
  return 0;
}

int r2d3c_plan_set_column_radix(r2d3c_plan *p, int form) {
  NEED(p);
  if (form != 0 && form != 1) return set_err(R2D3C_EINVAL, "column radix form: 0 or 1");
  p->col_cs = form;
  return 0;
}

int r2d3c_plan_get_column_radix(const r2d3c_plan *p, int *form) {
  NEED(p); NEED(form);
  *form = p->col_cs;
  return 0;
}

int r2d3c_plan_get_schedule(const r2d3c_plan *p, int *row_rows, int *row_blocks_per_cu) {
  NEED(p); NEED(row_rows); NEED(row_blocks_per_cu);
  *row_rows = p->row_rw;
  *row_blocks_per_cu = p->row_bpc;
  return 0;
}

int r2d3c_autotune(r2d3c_plan *p, int reps, double *ms_per_step, void *stream) {
  NEED(p);
  if (reps < 1) return set_err(R2D3C_EINVAL, "reps must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  const int n = p->n;
  const size_t sz = (size_t)(n / 2 + 1) * n * sizeof(cplx);
  cplx *psi = nullptr, *vz = nullptr;
  HIPCHK(hipMalloc(&psi, sz));
  hipError_t e = hipMalloc(&vz, sz);
  if (e != hipSuccess) { (void)hipFree(psi); return hip_err(e, "hipMalloc"); }
  // the timings do not depend on the values: a zero state is enough
  int rc = 0;
  if ((e = hipMemsetAsync(psi, 0, sz, s)) != hipSuccess || (e = hipMemsetAsync(vz, 0, sz, s)) != hipSuccess)
    rc = hip_err(e, "hipMemsetAsync");
  // candidates: (rows, blocks per CU, column radix form); the column form
  // matters only from n = 512, the row schedule only from n = 256
  const int cand[10][3] = {{2, 0, 0}, {2, 1, 0}, {2, 2, 0}, {1, 1, 0}, {1, 2, 0},
                           {2, 0, 1}, {2, 1, 1}, {2, 2, 1}, {1, 1, 1}, {1, 2, 1}};
  int best = -1;
  float best_ms = 0.f;
  const int rw0 = p->row_rw, bpc0 = p->row_bpc, cs0 = p->col_cs;
  for (int k = 0; k < 10 && !rc; ++k) {
    if (!schedule_valid(n, cand[k][0], cand[k][1])) continue;
    if (n < 256 && k > 0) break;  // no persistent row kernel below 256: nothing to tune
    if (n < 512 && cand[k][2]) break;  // one column radix schedule below 512
    p->row_rw = cand[k][0];
    p->row_bpc = cand[k][1];
    p->col_cs = cand[k][2];
    float ms = 0.f;
    auto tune = [&](auto tag) -> int {
      constexpr int NN = decltype(tag)::value;
      return tune_steps<NN>(p, psi, vz, reps, s, &ms);
    };
    switch (n) {
      case 16: rc = tune(std::integral_constant<int, 16>()); break;
      case 32: rc = tune(std::integral_constant<int, 32>()); break;
      case 64: rc = tune(std::integral_constant<int, 64>()); break;
      case 128: rc = tune(std::integral_constant<int, 128>()); break;
      case 256: rc = tune(std::integral_constant<int, 256>()); break;
      case 512: rc = tune(std::integral_constant<int, 512>()); break;
      case 1024: rc = tune(std::integral_constant<int, 1024>()); break;
      case 2048: rc = tune(std::integral_constant<int, 2048>()); break;
      case 4096: rc = tune(std::integral_constant<int, 4096>()); break;
      default: rc = set_err(R2D3C_ENOTSUP, "unsupported grid size");
    }
    if (!rc && (best < 0 || ms < best_ms)) { best = k; best_ms = ms; }
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(psi);
  (void)hipFree(vz);
  if (rc || best < 0) {
    p->row_rw = rw0;
    p->row_bpc = bpc0;
    p->col_cs = cs0;
    return rc ? rc : set_err(R2D3C_EINVAL, "no schedule could be timed");
  }
  p->row_rw = cand[best][0];
  p->row_bpc = cand[best][1];
  p->col_cs = cand[best][2];
  if (ms_per_step) *ms_per_step = best_ms / reps;
  return 0;
}

}  // extern "C"

// the network must hold at least one n x n image for the fused kernels
static int check_net(const r2d3c_plan *p, const r2d3c_cnn *net) {
  int b = 0, h = 0, w = 0;
  int rc = r2d3c_cnn_shape(net, &b, &h, &w);
  if (rc) return rc;
  if (b < 1 || h != p->n || w != p->n)
    return set_err(R2D3C_EINVAL, "closure: the network must take n x n images of the plan's grid");
  return 0;
}

// ---------------------------------------------------------------------------
// ensembles of independent runs
// ---------------------------------------------------------------------------
// A run whose step starts with the learned closure: inputs -> network ->
// dense forcing (fp, fz), which the fused step adds like iflow = 1 forcing.
struct EnsClosure {
  int member = 0;                   // index within its class
  r2d3c_cnn *net = nullptr;
  double sin[3] = {1, 1, 1}, sout[2] = {1, 1}, kmin = 0.0;
  double kclip = 0.0;               // backscatter clipping on |k| >= kclip (0: none)
  const double *psi = nullptr, *vz = nullptr;
  cplx *fp = nullptr, *fz = nullptr;
};

struct EnsClass {
  int n = 0, first = 0, count = 0;  // runs [first, first + count) of the ensemble, all of size n
  std::vector<EnsClosure> closures;
  r2d3c_plan *plan = nullptr;       // geometry, twiddles and schedule of size n
  EnsMember *d_mem = nullptr;
  cplx *work = nullptr;             // 8 mixed fields per run, run i at 8 fs i
  double *d_om = nullptr;           // 2 Omega n^2 per run (row kernel)
  StepConst c0{};                   // geometry shared by the class (row kernel)
  hipStream_t st = nullptr;         // branch stream of this class in the step graph
  hipEvent_t done = nullptr;
  // own streams (concurrent 2): the forcing-table step of this class and its
  // graphs of one step and of kblock steps, replayed on st
  int *d_step = nullptr;
  hipGraph_t g1 = nullptr, gk = nullptr;
  hipGraphExec_t x1 = nullptr, xk = nullptr;
};

struct r2d3c_ensemble {
  int count = 0;
  std::vector<EnsClass> cls;
  std::vector<int> order;           // run index in class order -> caller's index
  std::vector<r2d3c_member> mem;    // caller's runs
  hipStream_t es = nullptr;         // private stream: launches and graph capture
  hipEvent_t ev_in = nullptr, ev_out = nullptr, fork = nullptr;
  ForceSet *d_ftab = nullptr;
  size_t ftab_cap = 0;              // ForceSet entries allocated
  std::vector<ForceSet> h_ftab;
  int *d_step = nullptr;
  int period = 1;                   // closures refreshed every period-th step of a run call
  hipGraph_t graph = nullptr;       // one step with the closures
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph2 = nullptr;      // one step holding the closures' last forcing
  hipGraphExec_t exec2 = nullptr;
  hipGraph_t graphk = nullptr;      // a block of kblock steps (a multiple of period),
  hipGraphExec_t execk = nullptr;   // closures on its steps 0, period, 2 period, ...
  int kblock = 0;
  const ForceSet *graph_ftab = nullptr;
  bool graph_valid = false;
  // size classes of a step: 0 one after another on the ensemble's stream, 1
  // as concurrent branches of the step graph, 2 (default) each class on its
  // own stream for a whole run call (no closures; else as 0)
  int concurrent = 2;
  // stage launch form: 0 per-size kernels; 1 one persistent launch per RK4
  // stage for every class (k_ens_stage); 2 one launch per phase for every
  // class (k_ens_rows, k_ens_cols).  The class table, the grid barrier, its
  // time-out flag and the persistent grid size.
  int persistent = 0;
  PsClass *d_ps = nullptr;
  unsigned *d_bar = nullptr;
  int *d_err = nullptr;
  int ps_grid = 0, ps_rows = 0, ps_cols = 0;
};

// The size-class mode in effect: own streams (2) only with the per-size
// kernels and for ensembles without closures (their graphs alternate steps
// with and without the network); else one class after another.
static int ens_mode(const r2d3c_ensemble *e) {
  if (e->concurrent != 2) return e->concurrent;
  if (e->persistent) return 0;
  for (const auto &k : e->cls)
    if (!k.closures.empty()) return 0;
  return 2;
}

// tick: the step counter this class's first row launch advances (the first
// class of a step whose classes follow each other), or NULL.
template <int N>
static int run_closure_inputs_mixed(r2d3c_plan *p, const cplx *M, const double *sc,
                                    unsigned short *in_bf16, hipStream_t s);
template <int N>
static int run_closure_forcing_rows(r2d3c_plan *p, r2d3c_cnn *net, const double *sc,
                                    hipStream_t s);
void r2d3c_internal_cnn_forcing_gain(const r2d3c_cnn *c, const double **gain, int *n);
int r2d3c_internal_cnn_forward_mixed(r2d3c_cnn *c, const double *M, size_t fs, int kp, int nc,
                                     const double *tw, double s0, double s1, double s2,
                                     void *stream);

template <int N>
static int ens_class_step(r2d3c_ensemble *e, EnsClass &k, const ForceSet *ftab, hipStream_t s,
                          bool first, bool masked, bool chain, bool closures, int *tick) {
  const r2d3c_plan *p = k.plan;
  const int nc = p->g.nc, ncp = (nc + 1) & ~1;
  const cplx *tw = p->tw;
  EnsArgs a{k.d_mem, ftab, ens_mode(e) == 2 ? k.d_step : e->d_step, e->count, k.first};
  const bool cs = N >= 512 && p->col_cs;  // the columns' radix schedule
  if (first) {
    if (cs) k_ens_colinv<N, true><<<dim3(round16(ncp + nc), k.count), N / 8, 0, s>>>(a, tw);
    else k_ens_colinv<N, false><<<dim3(round16(ncp + nc), k.count), N / 8, 0, s>>>(a, tw);
  }
  // The closure forcing of this step, from the step's state (else the
  // forcing of the last evaluation is held).  A steady step's state is zero
  // outside the update mask and its column transforms are in the workspace
  // (emitted by the last stage before), so the network inputs come from
  // there.  (The chain on a side branch of the step graph beside the first
  // row launch measured slower: the graph ran the branch's launches one after
  // the other anyway, with a queue switch between.)
  // A lone closure run whose first column launch is latency-bound (early
  // loads, canonical order) also folds the forcing's column pass into that
  // launch (CfFuse): one launch fewer per evaluation.  Its forcing arrays
  // hold zeros past the kept columns from the call's first step (unfolded).
  const int nsl1 = col_slices(p, (long long)round16(2 * ncp) * k.count * (N / 8));
  const bool fold = R2D_CL_FOLD && closures && k.closures.size() == 1 &&
                    k.closures[0].kclip <= 0.0 && !first && masked && nsl1 > 1 &&
                    N <= R2D_COL_EARLY_MAXN && (N < 512 || cs);
  // Where the first row launch is the parallel-pair kernel, the forcing's
  // row pass rides along in extra blocks of that launch (k_step_row3_cf).
  bool row_plain = true;
  if constexpr (RowV2<N, 2>::THREADS >= 64) row_plain = p->row_bpc == 0;
  const bool fold_row = fold && row_plain && row3_takes_cf<N>(p);
  EnsArgs a4 = a;
  CfRowArgs cfr{};
  for (size_t ci = 0; closures && ci < k.closures.size(); ++ci) {
    EnsClosure &cl = k.closures[ci];
    int rc;
    if (fold) {
      const cplx *Mw = k.work + (size_t)cl.member * 8 * p->fs;
      if (N == 256 && R2D_CL_FOLD0) {
        // the input pass inside the network's first layer (k_conv0_mixed)
        const double inv = 1.0 / ((double)N * N);
        rc = r2d3c_internal_cnn_forward_mixed(cl.net, (const double *)Mw, p->fs, p->g.kp,
                                              p->g.nc, (const double *)p->tw, cl.sin[0] * inv,
                                              cl.sin[1] * inv, cl.sin[2] * inv, s);
      } else {
        void *in = nullptr;
        rc = r2d3c_cnn_buffers(cl.net, &in, nullptr);
        if (!rc) rc = run_closure_inputs_mixed<N>(k.plan, Mw, cl.sin, (unsigned short *)in, s);
        if (!rc) rc = r2d3c_cnn_forward_buffers(cl.net, 1, s);
      }
      if (!rc && !fold_row) rc = run_closure_forcing_rows<N>(k.plan, cl.net, cl.sout, s);
      if (!rc && fold_row) {
        const float *z = nullptr;
        rc = r2d3c_cnn_buffers(cl.net, nullptr, &z);
        cfr = CfRowArgs{z, p->mfull, cl.sout[0], cl.sout[1], p->g.ncf, p->g.kpf, cl.member, 0};
      }
      if (rc) return rc;
      const double *gain = nullptr;
      int ngain = 0;
      r2d3c_internal_cnn_forcing_gain(cl.net, &gain, &ngain);
      a4.cf = CfFuse{p->mfull, gain, cl.fp, cl.fz, cl.kmin * cl.kmin, p->g.kpf, ngain, cl.member};
      continue;
    }
    if (!first && masked) {
      void *in = nullptr;
      rc = r2d3c_cnn_buffers(cl.net, &in, nullptr);
      if (!rc)
        rc = run_closure_inputs_mixed<N>(k.plan, k.work + (size_t)cl.member * 8 * p->fs, cl.sin,
                                         (unsigned short *)in, s);
    } else {
      rc = r2d3c_closure_net_inputs(k.plan, cl.psi, cl.vz, cl.sin, cl.net, s);
    }
    if (!rc) rc = r2d3c_cnn_forward_buffers(cl.net, 1, s);
    if (!rc)
      rc = r2d3c_closure_net_forcing(k.plan, cl.net, cl.sout, cl.kmin, (double *)cl.fp,
                                     (double *)cl.fz, s);
    if (!rc && cl.kclip > 0.0)
      rc = r2d3c_closure_clip(k.plan, cl.psi, cl.vz, cl.kclip, (double *)cl.fp, (double *)cl.fz, s);
    if (rc) return rc;
  }
  for (int o = 4; o >= 1; --o) {
    // one row launch covers the class: the runs' workspaces are 8 fs apart
    const size_t ms = 8 * p->fs;
    const cplx *M0 = k.work;
    cplx *Q0 = k.work + 6 * p->fs;
    int *tk = o == 4 ? tick : nullptr;
    if constexpr (RowV2<N, 2>::THREADS >= 64) {
      if (p->row_bpc == 0)
        launch_row_np<N>(p, M0, Q0, k.c0, nullptr, tw, ms, k.d_om, k.count, s, nullptr, nullptr,
                         tk, (o == 4 && fold_row) ? &cfr : nullptr);
      else if (p->row_rw == 1 && RowV2<N, 1>::THREADS >= 64)
        k_step_row2<N, 1><<<dim3(row2_grid<N, 1>(p->cus, p->row_bpc), k.count),
                            RowV2<N, 1>::THREADS, 0, s>>>(M0, Q0, k.c0, nullptr, tw, ms, k.d_om,
                                                          tk);
      else
        k_step_row2<N, 2><<<dim3(row2_grid<N, 2>(p->cus, p->row_bpc), k.count),
                            RowV2<N, 2>::THREADS, 0, s>>>(M0, Q0, k.c0, nullptr, tw, ms, k.d_om,
                                                          tk);
    } else {
      launch_row_np<N>(p, M0, Q0, k.c0, nullptr, tw, ms, k.d_om, k.count, s, nullptr, nullptr, tk,
                       (o == 4 && fold_row) ? &cfr : nullptr);
    }
    int mode = FM_EMIT;
    if (o == 1) mode = FM_STATE | (masked ? FM_DISK : FM_ZERO) | (chain ? FM_EMIT : 0);
    const int nblk = round16(2 * ncp + ((mode & FM_ZERO) ? (N / 2 + 1 - nc) : 0));
    const int nsl = col_slices(p, (long long)nblk * k.count * (N / 8));
    // latency-bound columns: emission slices as thread groups of one block
    // (the last stage too), or as separate blocks (not the last stage)
    if constexpr (4 * (N / 8) <= 1024 && N <= 1024) {
      if (nsl > 1 && R2D_ENS_GSLICES >= ((mode & FM_STATE) ? 1 : 2)) {
        if (cs) k_ens_colfused_s<N, true, 4><<<dim3(nblk, k.count), 4 * (N / 8), 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
        else k_ens_colfused_s<N, false, 4><<<dim3(nblk, k.count), 4 * (N / 8), 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
        continue;
      }
    }
    const int nz = (mode & FM_STATE) ? 1 : nsl;
    // latency-bound (split) launches load the update's inputs early
    constexpr bool ESM = N <= R2D_COL_EARLY_MAXN;
    const bool early = ESM && nsl > 1;
    // columns per block: fill a wavefront below n = 512 (GC divides ncp and nblk)
    const int gc = (N / 8 >= 64 || !R2D_ENS_COL_GROUPS) ? 1 : (64 / (N / 8) >= 4 && ncp % 4 == 0 ? 4 : 2);
    if constexpr (N >= 32 && 4 * (N / 8) <= 64) {
      if (gc == 4) {
        if (early) k_ens_colfused_g<N, false, 4, ESM><<<dim3(nblk / 4, k.count, nz), 4 * (N / 8), 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
        else k_ens_colfused_g<N, false, 4><<<dim3(nblk / 4, k.count, nz), 4 * (N / 8), 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
        continue;
      }
    }
    if constexpr (N >= 32 && 2 * (N / 8) <= 64) {
      if (gc == 2) {
        if (early) k_ens_colfused_g<N, false, 2, ESM><<<dim3(nblk / 2, k.count, nz), 2 * (N / 8), 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
        else k_ens_colfused_g<N, false, 2><<<dim3(nblk / 2, k.count, nz), 2 * (N / 8), 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
        continue;
      }
    }
    if (early) {
      if (cs) k_ens_colfused<N, true, ESM><<<dim3(nblk, k.count, nz), N / 8, 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
      else k_ens_colfused<N, false, ESM><<<dim3(nblk, k.count, nz), N / 8, 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
    } else {
      if (cs) k_ens_colfused<N, true><<<dim3(nblk, k.count, nz), N / 8, 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
      else k_ens_colfused<N, false><<<dim3(nblk, k.count, nz), N / 8, 0, s>>>(o == 4 ? a4 : a, o, mode, tw);
    }
  }
  return 0;
}

static int ens_dispatch(r2d3c_ensemble *e, EnsClass &k, const ForceSet *ftab, hipStream_t s,
                        bool first, bool masked, bool chain, bool closures = true,
                        int *tick = nullptr) {
  int rc;
  switch (k.n) {
    case 16: rc = ens_class_step<16>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    case 32: rc = ens_class_step<32>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    case 64: rc = ens_class_step<64>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    case 128: rc = ens_class_step<128>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    case 256: rc = ens_class_step<256>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    case 512: rc = ens_class_step<512>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    case 1024: rc = ens_class_step<1024>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    case 2048: rc = ens_class_step<2048>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    case 4096: rc = ens_class_step<4096>(e, k, ftab, s, first, masked, chain, closures, tick); break;
    default: return set_err(R2D3C_ENOTSUP, "unsupported grid size");
  }
  if (rc) return rc;
  LAUNCHCHK("ensemble step");
  return 0;
}

// Whether a step's first row launch advances the forcing-table step (the
// per-size classes one after another) rather than a tick kernel.
static bool ens_row_ticks(const r2d3c_ensemble *e) {
  return !e->persistent && ens_mode(e) != 1;
}

// Mode 2: nsteps steady steps of class k alone, captured on its own stream;
// its first row launch of each step advances its own forcing-table step.
static int ens_capture_cls(r2d3c_ensemble *e, EnsClass &k, const ForceSet *ftab, int nsteps,
                           hipGraph_t *graph, hipGraphExec_t *exec) {
  if (*exec) { (void)hipGraphExecDestroy(*exec); *exec = nullptr; }
  if (*graph) { (void)hipGraphDestroy(*graph); *graph = nullptr; }
  HIPCHK(hipStreamBeginCapture(k.st, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int j = 0; j < nsteps && !rc; ++j)
    rc = ens_dispatch(e, k, ftab, k.st, false, true, true, false, k.d_step);
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(k.st, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (ee != hipSuccess) return hip_err(ee, "hipStreamEndCapture");
  *graph = g;
  HIPCHK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return 0;
}

// Captures one steady time step of every class: the classes one after another
// or as concurrent branches, or the all-size stage launches; the forcing-table
// step advances at the step's start (ens_row_ticks).
// nsteps steps, closures evaluated on steps j with j % every == 0 (every 0:
// none), in one graph: consecutive replays of a one-step graph leave a gap
// between graphs on the device that the steps of one graph do not.
static int ens_capture(r2d3c_ensemble *e, const ForceSet *ftab, int nsteps, int every,
                       hipGraph_t *graph, hipGraphExec_t *exec) {
  HIPCHK(hipStreamSynchronize(e->es));  // no launch of the old graph may still run
  if (*exec) { (void)hipGraphExecDestroy(*exec); *exec = nullptr; }
  if (*graph) { (void)hipGraphDestroy(*graph); *graph = nullptr; }
  e->graph_valid = false;
  int nk = 0;
  if (e->persistent) {  // the class table, largest size first
    std::vector<PsClass> h;
    for (auto &k : e->cls) {
      PsClass c{};
      c.a = EnsArgs{k.d_mem, ftab, e->d_step, e->count, k.first};
      c.tw = k.plan->tw;
      c.om = k.d_om;
      c.n = k.n;
      c.count = k.count;
      c.ucol = (k.plan->g.nc + 2048 / k.n - 1) / (2048 / k.n);
      h.push_back(c);
    }
    std::stable_sort(h.begin(), h.end(), [](const PsClass &a, const PsClass &b) { return a.n > b.n; });
    int re = 0, ce = 0;
    for (auto &c : h) {
      re += c.count * (c.n * c.n / 2048);
      ce += c.count * 2 * c.ucol;
      c.row_end = re;
      c.col_end = ce;
    }
    nk = (int)h.size();
    e->ps_rows = re;
    e->ps_cols = ce;
    HIPCHK(hipMemcpy(e->d_ps, h.data(), sizeof(PsClass) * h.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(hipStreamBeginCapture(e->es, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  hipError_t he = hipSuccess;
  for (int j = 0; j < nsteps && he == hipSuccess && !rc; ++j) {
    const bool closures = every > 0 && j % every == 0;
    // the forcing-table step: advanced by the first row launch when the
    // per-size classes follow each other, else by a tick at the step's start
    const bool row_tick = ens_row_ticks(e);
    if (!row_tick) k_ens_tick<<<1, 1, 0, e->es>>>(e->d_step);
    he = hipGetLastError();
    if (he == hipSuccess) he = hipEventRecord(e->fork, e->es);
    if (e->persistent) {
      for (int o = 4; o >= 1 && he == hipSuccess; --o) {
        const int mode = o == 1 ? (FM_STATE | FM_DISK | FM_EMIT) : FM_EMIT;
        if (e->persistent == 1) {
          k_ens_stage<<<e->ps_grid, kPsT, 0, e->es>>>(e->d_ps, nk, e->d_bar, e->d_err, kPsSpin, o,
                                                      mode);
        } else {
          k_ens_rows<<<round16(e->ps_rows), kPsT, 0, e->es>>>(e->d_ps, nk);
          k_ens_cols<<<round16(e->ps_cols), kPsT, 0, e->es>>>(e->d_ps, nk, o, mode);
        }
        he = hipGetLastError();
      }
    }
    // largest sizes first: with concurrent branches theirs is the longest path
    std::vector<EnsClass *> by_size;
    for (auto &k : e->cls) by_size.push_back(&k);
    std::stable_sort(by_size.begin(), by_size.end(),
                     [](const EnsClass *a, const EnsClass *b) { return a->n > b->n; });
    for (EnsClass *kp : by_size) {
      EnsClass &k = *kp;
      if (he != hipSuccess || rc || e->persistent) break;
      if (ens_mode(e) != 1) {  // one class after another, each with the whole GPU
        rc = ens_dispatch(e, k, ftab, e->es, false, true, true, closures,
                          row_tick && kp == by_size[0] ? e->d_step : nullptr);
        continue;
      }
      he = hipStreamWaitEvent(k.st, e->fork, 0);
      if (he == hipSuccess) rc = ens_dispatch(e, k, ftab, k.st, false, true, true, closures);
      if (he == hipSuccess && !rc) he = hipEventRecord(k.done, k.st);
      if (he == hipSuccess && !rc) he = hipStreamWaitEvent(e->es, k.done, 0);
    }
  }  // steps of the graph
  if (he == hipSuccess && !rc) he = hipGetLastError();
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(e->es, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (he != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_err(he, "ensemble capture"); }
  if (ee != hipSuccess) return hip_err(ee, "hipStreamEndCapture");
  *graph = g;
  HIPCHK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return 0;
}

// (Re)captures the step graphs a run call needs: with the closures, and, when
// closures are refreshed less often than every step, without them.
static int ens_capture_all(r2d3c_ensemble *e, const ForceSet *ftab) {
  bool any = false;
  for (auto &k : e->cls) any |= !k.closures.empty();
  if (ens_mode(e) == 2) {
    HIPCHK(hipStreamSynchronize(e->es));  // and so every class stream's last replay
    e->graph_valid = false;
    e->kblock = std::max(1, R2D_ENS_BLOCK);
    for (auto &k : e->cls) {
      int rc = ens_capture_cls(e, k, ftab, 1, &k.g1, &k.x1);
      if (!rc && e->kblock > 1) rc = ens_capture_cls(e, k, ftab, e->kblock, &k.gk, &k.xk);
      if (rc) return rc;
    }
    e->graph_ftab = ftab;
    e->graph_valid = true;
    return 0;
  }
  int rc = ens_capture(e, ftab, 1, 1, &e->graph, &e->exec);
  if (!rc && any && e->period > 1) {
    rc = ens_capture(e, ftab, 1, 0, &e->graph2, &e->exec2);
  } else {  // no held-forcing steps: drop a graph captured for an earlier set-up
    if (e->exec2) { (void)hipGraphExecDestroy(e->exec2); e->exec2 = nullptr; }
    if (e->graph2) { (void)hipGraphDestroy(e->graph2); e->graph2 = nullptr; }
  }
  // the block graph: about R2D_ENS_BLOCK steps, a multiple of the period
  e->kblock = e->period * std::max(1, R2D_ENS_BLOCK / e->period);
  if (!rc && R2D_ENS_BLOCK > 1)
    rc = ens_capture(e, ftab, e->kblock, any ? e->period : 0, &e->graphk, &e->execk);
  if (rc) return rc;
  e->graph_ftab = ftab;
  e->graph_valid = true;
  return 0;
}

static void ens_free(r2d3c_ensemble *e) {
  if (!e) return;
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->exec2) (void)hipGraphExecDestroy(e->exec2);
  if (e->graph2) (void)hipGraphDestroy(e->graph2);
  if (e->execk) (void)hipGraphExecDestroy(e->execk);
  if (e->graphk) (void)hipGraphDestroy(e->graphk);
  for (auto &k : e->cls) {
    for (auto &cl : k.closures) {
      (void)hipFree(cl.fp);
      (void)hipFree(cl.fz);
    }
    if (k.d_mem) (void)hipFree(k.d_mem);
    if (k.work) (void)hipFree(k.work);
    if (k.d_om) (void)hipFree(k.d_om);
    if (k.x1) (void)hipGraphExecDestroy(k.x1);
    if (k.g1) (void)hipGraphDestroy(k.g1);
    if (k.xk) (void)hipGraphExecDestroy(k.xk);
    if (k.gk) (void)hipGraphDestroy(k.gk);
    if (k.d_step) (void)hipFree(k.d_step);
    if (k.st) (void)hipStreamDestroy(k.st);
    if (k.done) (void)hipEventDestroy(k.done);
    if (k.plan) r2d3c_plan_destroy(k.plan);
  }
  if (e->d_ftab) (void)hipFree(e->d_ftab);
  if (e->d_step) (void)hipFree(e->d_step);
  if (e->d_ps) (void)hipFree(e->d_ps);
  if (e->d_bar) (void)hipFree(e->d_bar);
  if (e->d_err) (void)hipFree(e->d_err);
  if (e->ev_in) (void)hipEventDestroy(e->ev_in);
  if (e->ev_out) (void)hipEventDestroy(e->ev_out);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->es) (void)hipStreamDestroy(e->es);
  delete e;
}

extern "C" {

int r2d3c_ensemble_create(r2d3c_ensemble **out, const r2d3c_member *members, int count) {
  if (!out) return set_err(R2D3C_EINVAL, "ensemble pointer is NULL");
  *out = nullptr;
  NEED(members);
  if (count < 1) return set_err(R2D3C_EINVAL, "an ensemble needs at least one run");
  for (int m = 0; m < count; ++m) {
    const r2d3c_member &r = members[m];
    if (r.n < 16 || r.n > 4096 || (r.n & (r.n - 1)))
      return set_err(R2D3C_ENOTSUP, "run sizes must be powers of two in [16, 4096]");
    if (!r.psi || !r.vz || r.psi == r.vz) return set_err(R2D3C_EINVAL, "run state pointers");
    if (!(r.dt > 0.0)) return set_err(R2D3C_EINVAL, "ensembles step with a fixed dt > 0");
    int rc = check_params(&r.params);
    if (rc) return rc;
  }
  r2d3c_ensemble *e = new r2d3c_ensemble();
  e->count = count;
  e->mem.assign(members, members + count);
  int rc = 0;
  hipError_t he = hipSuccess;
  // classes by size, in the order sizes first appear; runs keep their order
  std::vector<int> sizes;
  for (int m = 0; m < count; ++m) {
    bool seen = false;
    for (int z : sizes) seen |= z == members[m].n;
    if (!seen) sizes.push_back(members[m].n);
  }
  for (int z : sizes) {
    EnsClass k;
    k.n = z;
    k.first = (int)e->order.size();
    for (int m = 0; m < count; ++m)
      if (members[m].n == z) e->order.push_back(m);
    k.count = (int)e->order.size() - k.first;
    e->cls.push_back(k);
  }
  for (auto &k : e->cls) {
    if ((rc = r2d3c_plan_create(&k.plan, k.n))) break;
    const size_t fs = k.plan->fs;
    if ((he = hipMalloc(&k.work, sizeof(cplx) * 8 * fs * k.count)) != hipSuccess) break;
    if ((he = hipMalloc(&k.d_mem, sizeof(EnsMember) * k.count)) != hipSuccess) break;
    std::vector<EnsMember> hm(k.count);
    for (int i = 0; i < k.count; ++i) {
      const r2d3c_member &r = members[e->order[k.first + i]];
      make_step_const(k.plan, &r.params, R2D3C_DT_FIXED, r.dt, &hm[i].c);
      hm[i].psi = (cplx *)r.psi;
      hm[i].vz = (cplx *)r.vz;
      hm[i].M = k.work + (size_t)i * 8 * fs;
    }
    if ((he = hipMemcpy(k.d_mem, hm.data(), sizeof(EnsMember) * k.count, hipMemcpyHostToDevice)) != hipSuccess) break;
    std::vector<double> om(k.count);
    for (int i = 0; i < k.count; ++i) om[i] = hm[i].c.two_omega_n2;
    k.c0 = hm[0].c;
    if ((he = hipMalloc(&k.d_om, sizeof(double) * k.count)) != hipSuccess) break;
    if ((he = hipMemcpy(k.d_om, om.data(), sizeof(double) * k.count, hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((he = hipStreamCreateWithFlags(&k.st, hipStreamNonBlocking)) != hipSuccess) break;
    if ((he = hipEventCreateWithFlags(&k.done, hipEventDisableTiming)) != hipSuccess) break;
    if ((he = hipMalloc(&k.d_step, sizeof(int))) != hipSuccess) break;
  }
  if (!rc && he == hipSuccess) he = hipStreamCreateWithFlags(&e->es, hipStreamNonBlocking);
  if (!rc && he == hipSuccess) he = hipEventCreateWithFlags(&e->ev_in, hipEventDisableTiming);
  if (!rc && he == hipSuccess) he = hipEventCreateWithFlags(&e->ev_out, hipEventDisableTiming);
  if (!rc && he == hipSuccess) he = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
  if (!rc && he == hipSuccess) he = hipMalloc(&e->d_step, sizeof(int));
  if (!rc && he != hipSuccess) rc = hip_err(he, "ensemble set-up");
  if (rc) { ens_free(e); return rc; }
  *out = e;
  return 0;
}

int r2d3c_ensemble_destroy(r2d3c_ensemble *e) {
  if (e && e->es) (void)hipStreamSynchronize(e->es);
  ens_free(e);
  return 0;
}

int r2d3c_ensemble_set_schedule(r2d3c_ensemble *e, int n, int row_rows, int row_blocks_per_cu) {
  NEED(e);
  for (auto &k : e->cls)
    if (k.n == n) {
      int rc = r2d3c_plan_set_schedule(k.plan, row_rows, row_blocks_per_cu);
      if (!rc) e->graph_valid = false;
      return rc;
    }
  return set_err(R2D3C_EINVAL, "no run of that size in the ensemble");
}

int r2d3c_ensemble_set_stage_launch(r2d3c_ensemble *e, int form) {
  NEED(e);
  if (form < 0 || form > 2) return set_err(R2D3C_EINVAL, "stage launch form: 0, 1 or 2");
  const int on = form;
  if (on) {
    if (e->cls.size() > (size_t)kPsMaxClasses)
      return set_err(R2D3C_ENOTSUP, "all-size stage launch: at most 5 size classes");
    for (auto &k : e->cls) {
      if (k.n < 64 || k.n > 1024)
        return set_err(R2D3C_ENOTSUP, "all-size stage launch: run sizes 64 to 1024 only");
      if (!k.closures.empty())
        return set_err(R2D3C_ENOTSUP, "all-size stage launch: runs without closures only");
    }
    HIPCHK(hipStreamSynchronize(e->es));
    if (!e->d_ps) {
      HIPCHK(hipMalloc(&e->d_ps, sizeof(PsClass) * kPsMaxClasses));
      HIPCHK(hipMalloc(&e->d_bar, 2 * sizeof(unsigned)));
      HIPCHK(hipMalloc(&e->d_err, sizeof(int)));
    }
    HIPCHK(hipMemset(e->d_bar, 0, 2 * sizeof(unsigned)));
    HIPCHK(hipMemset(e->d_err, 0, sizeof(int)));
    // every block resident: the occupancy limit, a multiple of 16 blocks
    int per_cu = 0, cus = e->cls[0].plan->cus;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_ens_stage, kPsT, 0));
    e->ps_grid = (per_cu * cus) & ~15;
    if (e->ps_grid < 16) return set_err(R2D3C_ENOTSUP, "persistent stage: occupancy too low");
    // the per-size kernels of a call's first step use the all-size
    // launches' arithmetic: the sequential row kernel, Stockham columns
    for (auto &k : e->cls) {
      k.plan->row_rw = 2;
      k.plan->row_bpc = 0;
      k.plan->row_split = 0;
      k.plan->col_cs = 1;
    }
  }
  if (e->persistent != on) e->graph_valid = false;
  e->persistent = on;
  return 0;
}

int r2d3c_ensemble_get_stage_launch(const r2d3c_ensemble *e, int *form, int *grid) {
  NEED(e);
  if (form) *form = e->persistent;
  if (grid) *grid = e->persistent == 1 ? e->ps_grid : 0;
  return 0;
}

int r2d3c_ensemble_sync(r2d3c_ensemble *e) {
  NEED(e);
  HIPCHK(hipStreamSynchronize(e->es));
  if (e->d_err) {
    int err = 0;
    HIPCHK(hipMemcpy(&err, e->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
      HIPCHK(hipMemset(e->d_bar, 0, 2 * sizeof(unsigned)));
      HIPCHK(hipMemset(e->d_err, 0, sizeof(int)));
      return set_err(R2D3C_EFAIL, "persistent stage: a grid barrier wait timed out (not every "
                                  "block resident); the state of that run call is invalid");
    }
  }
  return 0;
}

int r2d3c_ensemble_set_concurrent(r2d3c_ensemble *e, int concurrent) {
  NEED(e);
  if (concurrent < 0 || concurrent > 2) return set_err(R2D3C_EINVAL, "concurrent: 0, 1 or 2");
  if (e->concurrent != concurrent) e->graph_valid = false;
  e->concurrent = concurrent;
  return 0;
}

int r2d3c_ensemble_set_row_split(r2d3c_ensemble *e, int n, int mode) {
  NEED(e);
  for (auto &k : e->cls)
    if (k.n == n) {
      int rc = r2d3c_plan_set_row_split(k.plan, mode);
      if (!rc) e->graph_valid = false;
      return rc;
    }
  return set_err(R2D3C_EINVAL, "no run of that size in the ensemble");
}

int r2d3c_ensemble_set_column_split(r2d3c_ensemble *e, int n, int mode) {
  NEED(e);
  for (auto &k : e->cls)
    if (k.n == n) {
      int rc = r2d3c_plan_set_column_split(k.plan, mode);
      if (!rc) e->graph_valid = false;
      return rc;
    }
  return set_err(R2D3C_EINVAL, "no run of that size in the ensemble");
}

int r2d3c_ensemble_set_column_radix(r2d3c_ensemble *e, int n, int form) {
  NEED(e);
  for (auto &k : e->cls)
    if (k.n == n) {
      int rc = r2d3c_plan_set_column_radix(k.plan, form);
      if (!rc) e->graph_valid = false;
      return rc;
    }
  return set_err(R2D3C_EINVAL, "no run of that size in the ensemble");
}

int r2d3c_ensemble_set_closure(r2d3c_ensemble *e, int run, r2d3c_cnn *net, const double *scale_in,
                               const double *scale_out, double kmin, double kclip) {
  NEED(e);
  if (run < 0 || run >= e->count) return set_err(R2D3C_EINVAL, "no such run");
  if (net && e->persistent)
    return set_err(R2D3C_ENOTSUP, "closures need the per-size stage launches "
                                  "(r2d3c_ensemble_set_stage_launch(ens, 0))");
  int j = 0;
  while (e->order[j] != run) ++j;
  EnsClass *kp = nullptr;
  for (auto &k : e->cls)
    if (j >= k.first && j < k.first + k.count) kp = &k;
  EnsClass &k = *kp;
  const int i = j - k.first, n = k.n;
  HIPCHK(hipStreamSynchronize(e->es));
  // drop an existing closure of this run
  for (size_t c = 0; c < k.closures.size(); ++c)
    if (k.closures[c].member == i) {
      EnsClosure &cl = k.closures[c];
      (void)hipFree(cl.fp);
      (void)hipFree(cl.fz);
      k.closures.erase(k.closures.begin() + c);
      break;
    }
  EnsMember m;
  HIPCHK(hipMemcpy(&m, k.d_mem + i, sizeof(EnsMember), hipMemcpyDeviceToHost));
  m.c.fpd = m.c.fzd = nullptr;
  if (net) {
    NEED(scale_in); NEED(scale_out);
    int rc = check_net(k.plan, net);
    if (rc) return rc;
    EnsClosure cl;
    cl.member = i;
    cl.net = net;
    for (int c = 0; c < 3; ++c) cl.sin[c] = scale_in[c];
    for (int c = 0; c < 2; ++c) cl.sout[c] = scale_out[c];
    cl.kmin = kmin;
    cl.kclip = kclip > 0.0 ? kclip : 0.0;
    cl.psi = (const double *)m.psi;
    cl.vz = (const double *)m.vz;
    const size_t sz = (size_t)(n / 2 + 1) * n;
    hipError_t he = hipMalloc(&cl.fp, sizeof(cplx) * sz);
    if (he == hipSuccess) he = hipMalloc(&cl.fz, sizeof(cplx) * sz);
    if (he != hipSuccess) {
      (void)hipFree(cl.fp);
      (void)hipFree(cl.fz);
      return hip_err(he, "closure buffers");
    }
    m.c.fpd = cl.fp;
    m.c.fzd = cl.fz;
    k.closures.push_back(cl);
  }
  HIPCHK(hipMemcpy(k.d_mem + i, &m, sizeof(EnsMember), hipMemcpyHostToDevice));
  e->graph_valid = false;
  return 0;
}

int r2d3c_ensemble_set_closure_period(r2d3c_ensemble *e, int period) {
  NEED(e);
  if (period < 1) return set_err(R2D3C_EINVAL, "closure period must be >= 1");
  HIPCHK(hipStreamSynchronize(e->es));
  e->period = period;
  e->graph_valid = false;
  return 0;
}

int r2d3c_ensemble_run(r2d3c_ensemble *e, const r2d3c_forcing *forcing, int n_steps,
                       void *stream) {
  NEED(e);
  if (n_steps < 0) return set_err(R2D3C_EINVAL, "n_steps must be >= 0");
  if (n_steps == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const ForceSet *ftab = nullptr;
  if (forcing) {
    const size_t need = (size_t)n_steps * e->count * 2;
    e->h_ftab.assign(need, ForceSet{});
    for (int st = 0; st < n_steps; ++st)
      for (int j = 0; j < e->count; ++j) {
        const int m = e->order[j];
        const r2d3c_forcing &f = forcing[(size_t)st * e->count + m];
        if (f.dense_psi || f.dense_vz || f.ci_enabled)
          return set_err(R2D3C_ENOTSUP, "ensembles take rand_force draws only");
        ForceSet *dst = &e->h_ftab[((size_t)st * e->count + j) * 2];
        int rc = make_force_set(&f.psi, e->mem[m].n, dst);
        if (!rc) rc = make_force_set(&f.vz, e->mem[m].n, dst + 1);
        if (rc) return rc;
      }
    HIPCHK(hipStreamSynchronize(e->es));  // the previous run may still read the table
    if (need > e->ftab_cap) {
      if (e->d_ftab) HIPCHK(hipFree(e->d_ftab));
      e->d_ftab = nullptr;
      e->ftab_cap = 0;
      HIPCHK(hipMalloc(&e->d_ftab, sizeof(ForceSet) * need));
      e->ftab_cap = need;
    }
    HIPCHK(hipMemcpyAsync(e->d_ftab, e->h_ftab.data(), sizeof(ForceSet) * need,
                          hipMemcpyHostToDevice, e->es));
    ftab = e->d_ftab;
  }
  HIPCHK(hipEventRecord(e->ev_in, s));
  HIPCHK(hipStreamWaitEvent(e->es, e->ev_in, 0));
  // step 0 during the first step: with row ticks its first row launch
  // advances the counter from -1
  const bool row_tick = ens_row_ticks(e);
  if (ens_mode(e) == 2) {
    // every class on its own stream, from its own graphs, for the whole call
    for (auto &k : e->cls) k_set_i32<<<1, 1, 0, e->es>>>(k.d_step, -1);
    for (auto &k : e->cls) {
      int rc = ens_dispatch(e, k, ftab, e->es, true, false, n_steps > 1, true, k.d_step);
      if (rc) return rc;
    }
    if (n_steps > 1) {
      LAUNCHCHK("ensemble first step");
      if (!e->graph_valid || e->graph_ftab != ftab) {
        int rc = ens_capture_all(e, ftab);
        if (rc) return rc;
      }
      // each class's graph launches in one go, smallest size first (on the
      // 24/12/4 mix 0.94-0.96 ms per step against 1.04 largest first and
      // 1.03 with the classes' launches interleaved)
      std::vector<EnsClass *> by_size;
      for (auto &k : e->cls) by_size.push_back(&k);
      std::stable_sort(by_size.begin(), by_size.end(),
                       [](const EnsClass *a, const EnsClass *b) { return a->n < b->n; });
      HIPCHK(hipEventRecord(e->fork, e->es));
      for (EnsClass *kp : by_size) {
        EnsClass &k = *kp;
        HIPCHK(hipStreamWaitEvent(k.st, e->fork, 0));
        int st = 1;
        for (; k.xk && n_steps - st >= e->kblock; st += e->kblock)
          HIPCHK(hipGraphLaunch(k.xk, k.st));
        for (; st < n_steps; ++st) HIPCHK(hipGraphLaunch(k.x1, k.st));
        HIPCHK(hipEventRecord(k.done, k.st));
        HIPCHK(hipStreamWaitEvent(e->es, k.done, 0));
      }
    }
    HIPCHK(hipEventRecord(e->ev_out, e->es));
    HIPCHK(hipStreamWaitEvent(s, e->ev_out, 0));
    return 0;
  }
  k_set_i32<<<1, 1, 0, e->es>>>(e->d_step, row_tick ? -1 : 0);
  // the first step reads the state's own columns and zeroes what lies
  // outside the update mask, as the first step of r2d3c_run does
  for (auto &k : e->cls) {
    int rc = ens_dispatch(e, k, ftab, e->es, true, false, n_steps > 1, true,
                          row_tick && &k == &e->cls[0] ? e->d_step : nullptr);
    if (rc) return rc;
  }
  if (n_steps > 1) {
    LAUNCHCHK("ensemble first step");
    if (!e->graph_valid || e->graph_ftab != ftab) {
      int rc = ens_capture_all(e, ftab);
      if (rc) return rc;
    }
    // single steps up to a multiple of the period, then blocks of kblock
    // steps, then single steps again (the same kernels either way)
    auto single = [&](int st) {
      return hipGraphLaunch(st % e->period && e->exec2 ? e->exec2 : e->exec, e->es);
    };
    int st = 1;
    for (; st < n_steps && st % e->period; ++st) HIPCHK(single(st));
    for (; e->execk && n_steps - st >= e->kblock; st += e->kblock)
      HIPCHK(hipGraphLaunch(e->execk, e->es));
    for (; st < n_steps; ++st) HIPCHK(single(st));
  }
  HIPCHK(hipEventRecord(e->ev_out, e->es));
  HIPCHK(hipStreamWaitEvent(s, e->ev_out, 0));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned closure: coarse-graining and the closure's inputs and forcing
// ---------------------------------------------------------------------------
// Box filter of width h (transfer function sinc(kx h/2) sinc(ky h/2), 1 when
// h = 0) and truncation of a fine spectral field to the coarse grid's dealiased
// modes (k^2 <= (nc/3)^2); coefficients rescaled from nf^2 to nc^2 (the
// unnormalised convention).
__device__ __forceinline__ double box_tf(double k, double h) {
  const double a = 0.5 * k * h;
  return a == 0.0 ? 1.0 : sin(a) / a;
}
__global__ void k_filter(const cplx *__restrict__ fine, int nf, cplx *__restrict__ coarse, int nc,
                         double kmax, double h, double scale) {
  const size_t total = (size_t)(nc / 2 + 1) * nc;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / nc), j = (int)(e % nc);
    const int kx = kxof(i, nc), ky = kyof(j, nc);
    const double ka2 = (double)kx * kx + (double)ky * ky;
    cplx v = mk(0.0, 0.0);
    if (ka2 <= kmax && kx >= 0) {
      const int jf = ky >= 0 ? ky : ky + nf;
      const cplx a = fine[(size_t)kx * nf + jf];
      const double g = scale * box_tf((double)kx, h) * box_tf((double)ky, h);
      v = mk(a.x * g, a.y * g);
    }
    coarse[e] = v;
  }
}

// closure spectra: mode 0 out = -(a - b)/k^2, mode 1 out = a - b, both only
// inside the update mask tiny < k^2 <= kmax and for k^2 >= kmin2 (b may be NULL)
__global__ void k_closure_spec(const cplx *__restrict__ a, const cplx *__restrict__ b,
                               cplx *__restrict__ out, int n, double kmax, double tiny, int mode,
                               double kmin2) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    const int kx = kxof(i, n), ky = kyof(j, n);
    const double ka2 = (double)kx * kx + (double)ky * ky;
    cplx v = mk(0.0, 0.0);
    if (ka2 <= kmax && ka2 > tiny && ka2 >= kmin2) {
      v = a[e];
      if (b) v = mk(v.x - b[e].x, v.y - b[e].y);
      if (mode == 0) v = mk(-v.x / ka2, -v.y / ka2);
    }
    out[e] = v;
  }
}

// channel ch of a [n*n][nch] fp32 array <-> an n*n real field, times scale
__global__ void k_chan_pack(const double *__restrict__ r, size_t nn, double scale, int ch, int nch,
                            float *__restrict__ out) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x)
    out[e * nch + ch] = (float)(r[e] * scale);
}
__global__ void k_chan_unpack(const float *__restrict__ in, size_t nn, double scale, int ch, int nch,
                              double *__restrict__ r) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x)
    r[e] = (double)in[e * nch + ch] * scale;
}

extern "C" {

int r2d3c_filter(r2d3c_plan *fine, r2d3c_plan *coarse, const double *in, double *out, double width,
                 void *stream) {
  NEED(fine); NEED(coarse); NEED(in); NEED(out);
  const int nf = fine->n, nc = coarse->n;
  if (nc > nf) return set_err(R2D3C_EINVAL, "filter: the coarse grid must not be finer");
  if (width < 0.0) return set_err(R2D3C_EINVAL, "filter: width must be >= 0");
  const double sc = ((double)nc * nc) / ((double)nf * nf);
  const size_t total = (size_t)(nc / 2 + 1) * nc;
  k_filter<<<launch_grid(total), 256, 0, (hipStream_t)stream>>>((const cplx *)in, nf, (cplx *)out,
                                                                  nc, coarse->g.kmax, width, sc);
  LAUNCHCHK("filter");
  return 0;
}

int r2d3c_closure_inputs(r2d3c_plan *p, const double *psi, const double *vz, const double *scale,
                         float *out, void *stream) {
  NEED(p); NEED(psi); NEED(vz); NEED(scale); NEED(out);
  hipStream_t s = (hipStream_t)stream;
  const int n = p->n;
  const size_t sz = (size_t)(n / 2 + 1) * n, nn = (size_t)n * n;
  const double inv = 1.0 / ((double)n * n);
  cplx *t = p->spec3;
  // channel 0: laplacian of psi, 1: vz, 2: psi (real space, physical units)
  k_spec_op<<<launch_grid(sz), 256, 0, s>>>((const cplx *)psi, t, OP_LAP, n, p->g.kmax, p->g.tiny);
  const double *src[3] = {(const double *)t, vz, psi};
  for (int ch = 0; ch < 3; ++ch) {
    int rc = r2d3c_c2r(p, src[ch], p->real, stream);
    if (rc) return rc;
    k_chan_pack<<<launch_grid(nn), 256, 0, s>>>(p->real, nn, scale[ch] * inv, ch, 3, out);
  }
  LAUNCHCHK("closure inputs");
  return 0;
}

int r2d3c_closure_targets(r2d3c_plan *p, const double *f5, const double *c5, const double *f6,
                          const double *c6, const double *scale, double kmin, float *out,
                          void *stream) {
  NEED(p); NEED(f5); NEED(c5); NEED(f6); NEED(c6); NEED(scale); NEED(out);
  hipStream_t s = (hipStream_t)stream;
  const int n = p->n;
  const size_t sz = (size_t)(n / 2 + 1) * n, nn = (size_t)n * n;
  const double inv = 1.0 / ((double)n * n);
  cplx *t = p->spec3;
  const double *a[2] = {f5, f6}, *b[2] = {c5, c6};
  for (int ch = 0; ch < 2; ++ch) {
    k_closure_spec<<<launch_grid(sz), 256, 0, s>>>((const cplx *)a[ch], (const cplx *)b[ch], t, n,
                                                    p->g.kmax, p->g.tiny, 1, kmin * kmin);
    int rc = r2d3c_c2r(p, (const double *)t, p->real, stream);
    if (rc) return rc;
    k_chan_pack<<<launch_grid(nn), 256, 0, s>>>(p->real, nn, scale[ch] * inv, ch, 2, out);
  }
  LAUNCHCHK("closure targets");
  return 0;
}

int r2d3c_closure_forcing(r2d3c_plan *p, const float *pred, const double *scale, double kmin,
                          double *fp, double *fz, void *stream) {
  NEED(p); NEED(pred); NEED(scale); NEED(fp); NEED(fz);
  hipStream_t s = (hipStream_t)stream;
  const int n = p->n;
  const size_t sz = (size_t)(n / 2 + 1) * n, nn = (size_t)n * n;
  cplx *t = p->spec3;
  double *dst[2] = {fp, fz};
  for (int ch = 0; ch < 2; ++ch) {
    k_chan_unpack<<<launch_grid(nn), 256, 0, s>>>(pred, nn, scale[ch], ch, 2, p->real);
    int rc = r2d3c_r2c(p, p->real, (double *)t, stream);
    if (rc) return rc;
    k_closure_spec<<<launch_grid(sz), 256, 0, s>>>(t, nullptr, (cplx *)dst[ch], n, p->g.kmax,
                                                    p->g.tiny, ch == 0 ? 0 : 1, kmin * kmin);
  }
  LAUNCHCHK("closure forcing");
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// fused closure input and output (the closure inside an ensemble step)
// ---------------------------------------------------------------------------
// float -> bfloat16, round to nearest even (the network's input format)
__device__ __forceinline__ unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// column inverses of the three input fields: y = 0 laplak2(psi), 1 vz, 2 psi
template <int N>
__global__ __launch_bounds__(N / 8) void k_ci_colinv(const cplx *__restrict__ psi,
                                                      const cplx *__restrict__ vz,
                                                      cplx *__restrict__ mix, Geo g,
                                                      const cplx *__restrict__ tw) {
  constexpr int T = N / 8;
  __shared__ cplx lds[Shape<N>::LDSN];
  const int i = blockIdx.x, f = blockIdx.y, b = threadIdx.x;
  const int kx = kxof(i, N), op = f == 0 ? OP_LAP : OP_RAW;
  const cplx *src = f == 1 ? vz : psi;
  cplx v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int j = b + T * q;
    v[q] = apply_op(op, src[(size_t)i * N + j], kx, kyof(j, N), g.kmax, g.tiny);
  }
  fft_canon<N, +1, 1>(v, lds, b, tw);
  cplx *dst = mix + (size_t)f * ((size_t)N * g.kpf);
#pragma unroll
  for (int q = 0; q < 8; ++q) dst[tix(b + T * q, i, g.kpf)] = v[q];
}

// Rows y, y+1 of the three fields to the network input bf16 [y][x][16]:
// channel c = real-space field c times s[c] (s includes 1/n^2), 3..15 zero.
// The fields' column transforms f0, f1, f2 are tiled mixed fields of pitch kp
// holding columns kx < ncol (zero beyond): the full transforms of a plan
// (ncol = n/2 + 1), or a masked state's transforms in a step workspace
// (ncol = nc: its columns past n/3 are zero).
// Three packed transforms: (f0, f1) of row y, (f0, f1) of row y+1, f2 of both.
// When 3 n/8 threads fit a block (kCiPar), thread group g runs transform g in
// LDS region g, in parallel (the launch is latency-bound), and the third
// group's result is shared through its region; otherwise one group runs the
// three in turn.  The arithmetic is the same either way.
template <int N> struct CiRow {
  static constexpr int G = 3 * (N / 8) <= 1024 ? 3 : 1;  // thread groups
};
template <int N>
__global__ __launch_bounds__(CiRow<N>::G * (N / 8)) void k_ci_row(
    const cplx *__restrict__ f0, const cplx *__restrict__ f1, const cplx *__restrict__ f2,
    int kp, int ncol, const cplx *__restrict__ tw, double s0, double s1, double s2,
    unsigned short *__restrict__ out) {
  constexpr int T = N / 8, G = CiRow<N>::G, LN = Shape<N>::LDSN;
  __shared__ cplx lds[G * LN];
  const int grp = threadIdx.x / T, b = threadIdx.x - grp * T, y = 2 * blockIdx.x;
  auto load = [&](cplx v[8], const cplx *fa, int ya, const cplx *fb, int yb) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = b + T * q;
      const bool mir = k > N / 2;
      const int kk = mir ? N - k : k;
      const bool in = kk < ncol;
      v[q] = packed_z<N>(k, in ? fa[tix(ya, kk, kp)] : mk(0.0, 0.0),
                         in ? fb[tix(yb, kk, kp)] : mk(0.0, 0.0), mir);
    }
  };
  auto emit = [&](int r, const cplx z[8], const cplx v[8]) {  // row y + r
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int x = b + T * q;
      const unsigned h0 = to_bf16((float)(z[q].x * s0)), h1 = to_bf16((float)(z[q].y * s1));
      const unsigned h2 = to_bf16((float)((r == 0 ? v[q].x : v[q].y) * s2));
      uint4 *px = (uint4 *)(out + ((size_t)(y + r) * N + x) * 16);
      px[0] = make_uint4(h0 | (h1 << 16), h2, 0u, 0u);
      px[1] = make_uint4(0u, 0u, 0u, 0u);
    }
  };
  if constexpr (G == 3) {
    cplx a[8];
    if (grp < 2) load(a, f0, y + grp, f1, y + grp);
    else load(a, f2, y, f2, y + 1);
    cplx *P = lds + grp * LN;
    fft_canon<N, +1, 1>(a, P, b, tw);
    if (grp == 2) {
#pragma unroll
      for (int q = 0; q < 8; ++q) P[pad<N>(b + T * q)] = a[q];
    }
    __syncthreads();
    if (grp < 2) {
      cplx v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = lds[2 * LN + pad<N>(b + T * q)];
      emit(grp, a, v);
    }
  } else {
    cplx a[8], c[8], v[8];
    load(a, f0, y, f1, y);
    fft_canon<N, +1, 1>(a, lds, b, tw);
    load(c, f0, y + 1, f1, y + 1);
    fft_canon<N, +1, 1>(c, lds, b, tw);
    load(v, f2, y, f2, y + 1);
    fft_canon<N, +1, 1>(v, lds, b, tw);
    emit(0, a, v);
    emit(1, c, v);
  }
}

// Rows y, y+1 of the prediction (channels 0, 1 of fp32 [y][x][16], times
// s0, s1) forward transformed as packed real pairs into mixed fields 0, 1:
// thread group ch (two groups in parallel when 2 n/8 threads fit a block,
// else one group for both in turn) handles channel ch in LDS region ch.
template <int N> struct CfRow {
  static constexpr int G = 2 * (N / 8) <= 1024 ? 2 : 1;
};
// Channel ch of rows y, y+1 (one thread group of n/8 threads, LDS area P;
// every thread of the block calls it together: block barriers).  store:
// false for a group that only keeps the block's barriers in step.
template <int N>
__device__ __forceinline__ void cf_rows_group(const float *__restrict__ z, int ch, int y, int b,
                                              cplx *P, const cplx *__restrict__ tw, double sc,
                                              int ncf, int kpf, cplx *__restrict__ mix,
                                              bool store) {
  constexpr int T = N / 8;
  const size_t fs = (size_t)N * kpf;
  cplx v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = b + T * q;
    v[q] = mk((double)z[((size_t)y * N + x) * 16 + ch] * sc,
              (double)z[((size_t)(y + 1) * N + x) * 16 + ch] * sc);
  }
  fft_canon<N, -1, 1>(v, P, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) P[pad<N>(b + T * q)] = v[q];
  __syncthreads();
  if (store) unpack_store<N, 1, N / 2 + 1>(P, b, ncf, kpf, mix + ch * fs, y, mix + ch * fs, y + 1);
  __syncthreads();
}
template <int N>
__global__ __launch_bounds__(CfRow<N>::G * (N / 8)) void k_cf_row(
    const float *__restrict__ z, Geo g, const cplx *__restrict__ tw, double s0, double s1,
    cplx *__restrict__ mix) {
  constexpr int T = N / 8, G = CfRow<N>::G, LN = Shape<N>::LDSN;
  __shared__ cplx lds[G * LN];
  const int grp = threadIdx.x / T, b = threadIdx.x - grp * T, y = 2 * blockIdx.x;
  cplx *P = lds + grp * LN;
  for (int ch = grp; ch < 2; ch += G)
    cf_rows_group<N>(z, ch, y, b, P, tw, ch == 0 ? s0 : s1, g.ncf, g.kpf, mix, true);
}

// k_step_row3 plus the closure forcing's row pass of run cf.member: blocks
// past cf.nrb run it, 3 row pairs per block (one group of n/8 threads per
// channel); a launch fewer per evaluation for a lone closure run.
template <int N>
__global__ __launch_bounds__(3 * (N / 8) * RowStep<N>::ROWS) void k_step_row3_cf(
    const cplx *__restrict__ M, cplx *__restrict__ Q, StepConst c, unsigned long long *maxbits,
    const cplx *__restrict__ tw, size_t mstride, const double *__restrict__ om_tab, int *tick,
    CfRowArgs cf) {
  constexpr int T = N / 8, RW = RowStep<N>::ROWS, LN = Shape<N>::LDSN, NG = 3 * RW;
  static_assert(NG % 2 == 0, "channel pairs per block");
  __shared__ cplx lds[3 * RW * LN];
  row_tick(tick);
  const int hb = blockIdx.x;
  if (hb < cf.nrb) {
    row3_body<N>(M, Q, c, maxbits, tw, mstride, om_tab, xcd_group_n<RowStep<N>::G>(hb, cf.nrb),
                 lds);
    return;
  }
  if ((int)blockIdx.y != cf.member) return;  // uniform over the block
  const int grp = threadIdx.x / T, b = threadIdx.x - grp * T, ch = grp & 1;
  const int task = (hb - cf.nrb) * (NG / 2) + grp / 2;  // row pair
  const bool ok = task < N / 2;
  cf_rows_group<N>(cf.z, ch, 2 * (ok ? task : 0), b, lds + grp * LN, tw, ch == 0 ? cf.s0 : cf.s1,
                   cf.ncf, cf.kpf, cf.mix, ok);
}

template <int N>
static void launch_row_np(const r2d3c_plan *p, const cplx *M, cplx *Q, const StepConst &c,
                          unsigned long long *macc, const cplx *tw, size_t ms, const double *om,
                          int count, hipStream_t s, hipEvent_t ea, hipEvent_t eb, int *tick,
                          const CfRowArgs *cfr) {
  if (!cfr || !row3_takes_cf<N>(p)) {
    launch_row_np_impl<N>(p, M, Q, c, macc, tw, ms, om, count, s, ea, eb, tick);
    return;
  }
  constexpr int NN = N <= 512 ? N : 512, RW = RowStep<NN>::ROWS, PB = 3 * RW / 2;
  CfRowArgs a = *cfr;
  a.nrb = NN / RW;
  hipExtLaunchKernelGGL(k_step_row3_cf<NN>, dim3(a.nrb + (NN / 2 + PB - 1) / PB, count),
                        dim3(3 * (NN / 8) * RW), 0, s, ea, eb, 0, M, Q, c, macc, tw, ms, om, tick,
                        a);
}

// column forwards of the two predicted fields with the closure epilogue:
// y = 0: fp = -tau5/k^2, y = 1: fz = tau6, both on the update mask and
// k^2 >= kmin2 only, times gain[int(|k| + 0.5)] for shells below ngain
template <int N>
__global__ __launch_bounds__(N / 8) void k_cf_col(const cplx *__restrict__ mix, Geo g,
                                                   const cplx *__restrict__ tw, double kmin2,
                                                   const double *__restrict__ gain, int ngain,
                                                   cplx *__restrict__ fp, cplx *__restrict__ fz) {
  constexpr int T = N / 8;
  __shared__ cplx lds[Shape<N>::LDSN];
  const int i = blockIdx.x, f = blockIdx.y, b = threadIdx.x;
  const cplx *src = mix + (size_t)f * ((size_t)N * g.kpf);
  cplx v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = src[tix(b + T * q, i, g.kpf)];
  fft_canon<N, -1, 1>(v, lds, b, tw);
  cplx *dst = f == 0 ? fp : fz;
  const int kx = kxof(i, N);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int j = b + T * q, ky = kyof(j, N);
    const double ka2 = (double)kx * kx + (double)ky * ky;
    cplx o = mk(0.0, 0.0);
    if (ka2 <= g.kmax && ka2 > g.tiny && ka2 >= kmin2) {
      o = f == 0 ? mk(-v[q].x / ka2, -v[q].y / ka2) : v[q];
      const int shell = (int)(sqrt(ka2) + 0.5);
      if (shell < ngain) o = mk(o.x * gain[shell], o.y * gain[shell]);
    }
    dst[(size_t)i * N + j] = o;
  }
}

// Backscatter clipping on the modes |k| >= kclip: where the closure forcing
// injects energy (Re(conj(psi) fp) > 0 for psi, whose energy is k^2 |psi|^2;
// Re(conj(vz) fz) > 0 for vz) the injecting component along the mode's own
// coefficient is removed, fp -= Re(conj(psi) fp) / |psi|^2 psi (same for
// fz), leaving the energy-neutral part.
__global__ void k_closure_clip(const cplx *__restrict__ psi, const cplx *__restrict__ vz,
                               cplx *__restrict__ fp, cplx *__restrict__ fz, int n,
                               double kclip2) {
  const size_t total = (size_t)(n / 2 + 1) * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double kx = kxof(i, n), ky = kyof(j, n);
    if (kx * kx + ky * ky < kclip2) continue;
    const cplx a = psi[e], f = fp[e];
    const double ta = a.x * f.x + a.y * f.y, na = a.x * a.x + a.y * a.y;
    if (ta > 0.0 && na > 0.0) fp[e] = mk(f.x - ta / na * a.x, f.y - ta / na * a.y);
    const cplx b = vz[e], g = fz[e];
    const double tb = b.x * g.x + b.y * g.y, nb = b.x * b.x + b.y * b.y;
    if (tb > 0.0 && nb > 0.0) fz[e] = mk(g.x - tb / nb * b.x, g.y - tb / nb * b.y);
  }
}

void r2d3c_internal_cnn_forcing_gain(const r2d3c_cnn *c, const double **gain, int *n);

template <int N>
static int run_closure_io(r2d3c_plan *p, int dir, const cplx *psi, const cplx *vz, const double *sc,
                          double kmin, unsigned short *in_bf16, const float *z, cplx *fp, cplx *fz,
                          hipStream_t s, const r2d3c_cnn *net = nullptr) {
  if (dir == 0) {
    k_ci_colinv<N><<<dim3(N / 2 + 1, 3), N / 8, 0, s>>>(psi, vz, p->mfull, p->g, p->tw);
    const double inv = 1.0 / ((double)N * N);
    const size_t fs = (size_t)N * p->g.kpf;
    k_ci_row<N><<<N / 2, CiRow<N>::G * (N / 8), 0, s>>>(
        p->mfull, p->mfull + fs, p->mfull + 2 * fs, p->g.kpf, N / 2 + 1, p->tw, sc[0] * inv,
        sc[1] * inv, sc[2] * inv, in_bf16);
  } else {
    k_cf_row<N><<<N / 2, CfRow<N>::G * (N / 8), 0, s>>>(z, p->g, p->tw, sc[0], sc[1], p->mfull);
    const double *gain = nullptr;
    int ngain = 0;
    r2d3c_internal_cnn_forcing_gain(net, &gain, &ngain);
    k_cf_col<N><<<dim3(N / 2 + 1, 2), N / 8, 0, s>>>(p->mfull, p->g, p->tw, kmin * kmin, gain,
                                                      ngain, fp, fz);
  }
  LAUNCHCHK("closure io");
  return 0;
}

// The network inputs of a masked state from its column transforms in a step
// workspace (fields 0, 2, 4 of M: psi, -k^2 psi = omega, vz, as the step's
// last stage emits them), instead of transforming the state again: the same
// values as r2d3c_closure_net_inputs for a state that is zero outside the
// update mask.
template <int N>
static int run_closure_inputs_mixed(r2d3c_plan *p, const cplx *M, const double *sc,
                                    unsigned short *in_bf16, hipStream_t s) {
  const double inv = 1.0 / ((double)N * N);
  const size_t fs = p->fs;
  k_ci_row<N><<<N / 2, CiRow<N>::G * (N / 8), 0, s>>>(M + 2 * fs, M + 4 * fs, M, p->g.kp,
                                                       p->g.nc, p->tw, sc[0] * inv, sc[1] * inv,
                                                       sc[2] * inv, in_bf16);
  LAUNCHCHK("closure inputs (mixed)");
  return 0;
}

// The forcing's row pass alone (k_cf_row into the plan's mixed fields); the
// column pass then runs inside the first stage's column launch (CfFuse).
template <int N>
static int run_closure_forcing_rows(r2d3c_plan *p, r2d3c_cnn *net, const double *sc,
                                    hipStream_t s) {
  const float *z = nullptr;
  const int rc = r2d3c_cnn_buffers(net, nullptr, &z);
  if (rc) return rc;
  k_cf_row<N><<<N / 2, CfRow<N>::G * (N / 8), 0, s>>>(z, p->g, p->tw, sc[0], sc[1], p->mfull);
  LAUNCHCHK("closure forcing rows");
  return 0;
}

extern "C" {

int r2d3c_closure_net_inputs(r2d3c_plan *p, const double *psi, const double *vz,
                             const double *scale, r2d3c_cnn *net, void *stream) {
  NEED(p); NEED(psi); NEED(vz); NEED(scale); NEED(net);
  void *in = nullptr;
  int rc = check_net(p, net);
  if (!rc) rc = r2d3c_cnn_buffers(net, &in, nullptr);
  if (rc) return rc;
  R2D_DISPATCH(p->n, run_closure_io, p, 0, (const cplx *)psi, (const cplx *)vz, scale, 0.0,
               (unsigned short *)in, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int r2d3c_closure_net_forcing(r2d3c_plan *p, r2d3c_cnn *net, const double *scale, double kmin,
                              double *fp, double *fz, void *stream) {
  NEED(p); NEED(net); NEED(scale); NEED(fp); NEED(fz);
  const float *z = nullptr;
  int rc = check_net(p, net);
  if (!rc) rc = r2d3c_cnn_buffers(net, nullptr, &z);
  if (rc) return rc;
  R2D_DISPATCH(p->n, run_closure_io, p, 1, nullptr, nullptr, scale, kmin, nullptr, z, (cplx *)fp,
               (cplx *)fz, (hipStream_t)stream, net);
}

int r2d3c_closure_clip(r2d3c_plan *p, const double *psi, const double *vz, double kclip,
                       double *fp, double *fz, void *stream) {
  NEED(p); NEED(psi); NEED(vz); NEED(fp); NEED(fz);
  if (!(kclip > 0.0)) return set_err(R2D3C_EINVAL, "kclip must be > 0");
  const size_t total = (size_t)(p->n / 2 + 1) * p->n;
  k_closure_clip<<<launch_grid(total), 256, 0, (hipStream_t)stream>>>(
      (const cplx *)psi, (const cplx *)vz, (cplx *)fp, (cplx *)fz, p->n, kclip * kclip);
  LAUNCHCHK("closure clip");
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// n = 16384.  A length-16384 transform exceeds one block's LDS and threads, so
// it is done as two 8192-point halves (the even and the odd samples) in one
// block of 1024 threads, 16 points per thread, then a radix-2 combine in
// registers.  The 2048 step's row kernel (all six derivative rows transformed
// in one kernel, three products held in registers) would need more registers
// than 1024 threads have, so four derivative fields go through HBM as real
// fields and the bracket pass (k_big_row_jac2) transforms psi_x + i psi_y in
// registers and reads them.
// ---------------------------------------------------------------------------

// In: ev[q] = x[2(b + 1024 q)], od[q] = x[2(b + 1024 q) + 1].
// Out: ev[q] = X[b + 1024 q], od[q] = X[b + 1024 q + 8192] (sign S).
// With R2D_BIG_TWLAST the thread loads its twiddles once: w^(2b) for both
// halves' last passes (tw_last) and w^b, w^(b + 1024) for the combine, whose
// other factors w^(b + 1024 q) are these times exp(S i pi u / 4), q = 2u or
// 2u + 1 (16 table loads fewer per transform).
#ifndef R2D_BIG_TWLAST
#define R2D_BIG_TWLAST 1
#endif
template <int S>
__device__ __forceinline__ void fft16k(cplx ev[8], cplx od[8], cplx *lds, int b,
                                       const cplx *__restrict__ tw) {
#if R2D_BIG_TWLAST
  const TwLast tls = tw_last<kBigH, 2>(tw, b);
  const TwLast *tl = &tls;
#else
  const TwLast *tl = nullptr;
#endif
  // opaque copies of the thread index keep each half's index math from being
  // hoisted over the other half (at 1024 threads it would spill)
  int b1 = b, b2 = b;
  asm volatile("" : "+v"(b1));
  fft_canon<kBigH, S, 1, 2>(ev, lds, b1, tw, tl);
  asm volatile("" : "+v"(b2));
  fft_canon<kBigH, S, 1, 2>(od, lds, b2, tw, tl);
#if R2D_BIG_TWLAST
  const cplx c0 = twid<S>(tw, b), c1 = twid<S>(tw, b + kBigT);
#endif
#pragma unroll
  for (int q = 0; q < 8; ++q) {
#if R2D_BIG_TWLAST
    const cplx w = rot8<S>((q & 1) ? c1 : c0, q >> 1);
#else
    const cplx w = twid<S>(tw, b + kBigT * q);
#endif
    const cplx t = cmul(od[q], w), e = ev[q];
    ev[q] = cadd(e, t);
    od[q] = csub(e, t);
  }
}

// column inverse of S[i][:] (with op, times scale) -> M(i, y)
__global__ __launch_bounds__(kBigT) void k_big_col_inv(const cplx *__restrict__ spec,
                                                        cplx *__restrict__ mix, int op,
                                                        double scale, Geo g,
                                                        const cplx *__restrict__ tw) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  const int i = blockIdx.x, b = threadIdx.x, kx = kxof(i, N);
  cplx ev[8], od[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int j = 2 * (b + kBigT * q);
    ev[q] = cscale(apply_op(op, spec[(size_t)i * N + j], kx, kyof(j, N), g.kmax, g.tiny), scale);
    od[q] = cscale(apply_op(op, spec[(size_t)i * N + j + 1], kx, kyof(j + 1, N), g.kmax, g.tiny),
                   scale);
  }
  fft16k<+1>(ev, od, lds, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int y = b + kBigT * q;
    mix[tix(y, i, g.kpf)] = ev[q];
    mix[tix(y + kBigH, i, g.kpf)] = od[q];
  }
}

// Column range of packed slot (q, parity) of k_big_row_c2r over its
// threads b = 0..kBigT-1: k = 2 (b + kBigT q) + parity, read at kk = k (k <=
// n/2) or n - k.
__host__ __device__ constexpr int big_kk_min(int q, int par) {
  return q < 4 ? 2 * kBigT * q + par : kBigN - (2 * kBigT * q + 2 * kBigT - 2 + par);
}
__host__ __device__ constexpr int big_kk_max(int q, int par) {
  return q < 4 ? 2 * kBigT * q + 2 * kBigT - 2 + par
               : (q == 4 && par == 0 ? kBigN / 2 : kBigN - (2 * kBigT * q + par));
}

// Packed inputs of one 16384-point row transform in fft16k's input order
// (ev[q]: k = 2(b + 1024 q), od[q]: k + 1): Z[k] = A[k] + i B[k] for k <= n/2
// and conj(A[n-k]) + i conj(B[n-k]) above (packed_z), with A from row ya of
// the mixed field FA and B from row yb of FB (tiled layout, pitch kp).
// Columns kx >= KC are taken as zero (not read): KC is the number of columns
// read (n/3 + 1 for the step's dealiased fields, n/2 + 1 for full
// transforms); slots wholly past it are neither loaded nor tested, slots
// wholly below it are not tested.  Two q slots (four points) at a time: a
// group's loads are issued before its arithmetic, without branches (points
// past KC read column 0 and are discarded), so they are in flight together.
template <int KC>
__device__ __forceinline__ void big_load_packed(cplx ev[8], cplx od[8], const cplx *__restrict__ FA,
                                                int ya, const cplx *__restrict__ FB, int yb, int b,
                                                int kp) {
  constexpr int N = kBigN, ncols = KC;
#pragma unroll
  for (int q0 = 0; q0 < 8; q0 += 2) {
    cplx ra[4], rb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + (u >> 1), par = u & 1;
      if (big_kk_min(q, par) >= ncols) continue;  // no thread reads this slot
      const int k = 2 * (b + kBigT * q) + par;
      const int kk = k > N / 2 ? N - k : k;
      const int kl = (big_kk_max(q, par) < ncols || kk < ncols) ? kk : 0;
      ra[u] = FA[tix(ya, kl, kp)];
      rb[u] = FB[tix(yb, kl, kp)];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + (u >> 1), par = u & 1;
      const int k = 2 * (b + kBigT * q) + par;
      const bool mir = k > N / 2;
      const int kk = mir ? N - k : k;
      const bool in = big_kk_min(q, par) < ncols && (big_kk_max(q, par) < ncols || kk < ncols);
      const cplx zz = in ? packed_z<N>(k, ra[u], rb[u], mir) : mk(0.0, 0.0);
      if (u & 1) od[q0 + (u >> 1)] = zz;
      else ev[q0 + (u >> 1)] = zz;
    }
    __builtin_amdgcn_sched_barrier(0);  // the next group's loads stay below
  }
}

// rows y, y+1 of M (half spectra) -> R[y][:], R[y+1][:] as one packed
// transform; columns kx >= KC are taken as zero (not read).
template <int KC>
__global__ __launch_bounds__(kBigT) void k_big_row_c2r(const cplx *__restrict__ mix,
                                                        double *__restrict__ out, Geo g,
                                                        const cplx *__restrict__ tw) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  // Consecutive workgroups go to the 8 XCDs in turn: the two row pairs of one
  // 4-row tile (which share every 64-byte chunk of the tiled layout) are
  // given to consecutive workgroups of one XCD, so the second finds the
  // chunks in that XCD's L2.
  const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3;
  const int b = threadIdx.x, y = 2 * ((((l >> 1) << 3) + xcd) * 2 + (l & 1));
  cplx ev[8], od[8];
  big_load_packed<KC>(ev, od, mix, y, mix, y + 1, b, g.kpf);
  fft16k<+1>(ev, od, lds, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = b + kBigT * q;
    out[(size_t)y * N + x] = ev[q].x;
    out[(size_t)(y + 1) * N + x] = ev[q].y;
    out[(size_t)y * N + x + kBigH] = od[q].x;
    out[(size_t)(y + 1) * N + x + kBigH] = od[q].y;
  }
}

// ---------------------------------------------------------------------------
// Split column inverse.  A column's 16384-point transform is X[k] = E[k] +
// w^k O[k], X[k + 8192] = E[k] - w^k O[k] (w = exp(S 2 pi i / 16384)) with E,
// O the 8192-point transforms of the even and the odd samples.  k_big_col_inv_h
// computes E or O in a block of 512 threads (16 points per thread, two
// 4096-point halves through 70 KB of LDS), so two such blocks are resident per
// CU and one block's loads and stores overlap the other's transform (one
// 1024-thread block per CU, as k_big_col_inv runs, leaves them in turn).  The
// halves go to a field of the same size in the layout hix: at each column the
// 64-byte chunk of E[4t..4t+3] is followed by that of O[4t..4t+3], so each
// block writes whole chunks and E[k], O[k] share one 128-byte line.  The row
// pass that reads it (k_big_row_c2r_h) transforms rows k and k + 8192 as its
// packed pair, forming both from E[k] and O[k] with one twiddle per row.
// ---------------------------------------------------------------------------
constexpr int kBigT2 = 512;
// slot of half p, row k < 8192, column kx in a split field of column pitch kp
__device__ __forceinline__ size_t hix(int k, int p, int kx, int kp) {
  return (size_t)(k >> 2) * (size_t)(kp * 8) + (size_t)kx * 8 + p * 4 + (k & 3);
}

// In: ev[q] = u[2(b + 512 q)], od[q] = u[2(b + 512 q) + 1] (u of length 8192).
// Out: ev[q] = U[b + 512 q], od[q] = U[b + 512 q + 4096] (sign S).  tw is the
// 16384-point table.
template <int S>
__device__ __forceinline__ void fft8k(cplx ev[8], cplx od[8], cplx *lds, int b,
                                      const cplx *__restrict__ tw) {
  constexpr int H = 4096;
  const TwLast tls = tw_last<H, 4>(tw, b);
  int b1 = b, b2 = b;
  asm volatile("" : "+v"(b1));
  fft_canon<H, S, 1, 4>(ev, lds, b1, tw, &tls);
  asm volatile("" : "+v"(b2));
  fft_canon<H, S, 1, 4>(od, lds, b2, tw, &tls);
  // combine factors exp(S 2 pi i (b + 512 q) / 8192): the loaded b and
  // b + 512 times exp(S i pi u / 4), q = 2u or 2u + 1
  const cplx c0 = twid<S>(tw, 2 * b), c1 = twid<S>(tw, 2 * (b + kBigT2));
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const cplx w = rot8<S>((q & 1) ? c1 : c0, q >> 1);
    const cplx t = cmul(od[q], w), e = ev[q];
    ev[q] = cadd(e, t);
    od[q] = csub(e, t);
  }
}

// half p (0: even samples, 1: odd) of the column inverse of S[i][:] (with op,
// times scale) -> H[hix(k, p, i)], k < 8192.  Block 8 l +
// x (x = 0..7) takes column 8 (l / 2) + x, half l % 2: both halves of a column
// go to one XCD back to back, so the second finds the first's lines (each
// 64-byte line holds samples of both) in that XCD's L2.
__global__ __launch_bounds__(kBigT2) __attribute__((amdgpu_waves_per_eu(4)))
void k_big_col_inv_h(const cplx *__restrict__ spec, cplx *__restrict__ mix, int op, double scale,
                     Geo g, const cplx *__restrict__ tw, int ncols) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<4096>::LDSN];
  const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3, p = l & 1;
  const int i = ((l >> 1) << 3) + xcd, b = threadIdx.x;
  if (i >= ncols) return;  // whole block, before any barrier
  const int kx = kxof(i, N);
  cplx ev[8], od[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int j = 4 * (b + kBigT2 * q) + p;
    ev[q] = cscale(apply_op(op, spec[(size_t)i * N + j], kx, kyof(j, N), g.kmax, g.tiny), scale);
    od[q] = cscale(apply_op(op, spec[(size_t)i * N + j + 2], kx, kyof(j + 2, N), g.kmax, g.tiny),
                   scale);
  }
  fft8k<+1>(ev, od, lds, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = b + kBigT2 * q;
    mix[hix(k, p, i, g.kpf)] = ev[q];
    mix[hix(k + 4096, p, i, g.kpf)] = od[q];
  }
}

// Packed inputs of rows y and y + 8192 from a split mixed field H (E[y] and
// O[y] at hix(y, 0 or 1, kx)): A = E + w O, B = E - w O with
// w = exp(2 pi i y / 16384), then as big_load_packed.
template <int KC>
__device__ __forceinline__ void big_load_packed_split(cplx ev[8], cplx od[8],
                                                      const cplx *__restrict__ H, int y, cplx w,
                                                      int b, int kp) {
  constexpr int N = kBigN, ncols = KC;
#pragma unroll
  for (int q0 = 0; q0 < 8; q0 += 2) {
    cplx ra[4], rb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + (u >> 1), par = u & 1;
      if (big_kk_min(q, par) >= ncols) continue;  // no thread reads this slot
      const int k = 2 * (b + kBigT * q) + par;
      const int kk = k > N / 2 ? N - k : k;
      const int kl = (big_kk_max(q, par) < ncols || kk < ncols) ? kk : 0;
      ra[u] = H[hix(y, 0, kl, kp)];
      rb[u] = H[hix(y, 1, kl, kp)];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + (u >> 1), par = u & 1;
      const int k = 2 * (b + kBigT * q) + par;
      const bool mir = k > N / 2;
      const int kk = mir ? N - k : k;
      const bool in = big_kk_min(q, par) < ncols && (big_kk_max(q, par) < ncols || kk < ncols);
      const cplx t = cmul(rb[u], w);
      const cplx zz = in ? packed_z<N>(k, cadd(ra[u], t), csub(ra[u], t), mir) : mk(0.0, 0.0);
      if (u & 1) od[q0 + (u >> 1)] = zz;
      else ev[q0 + (u >> 1)] = zz;
    }
    __builtin_amdgcn_sched_barrier(0);  // the next group's loads stay below
  }
}

// rows y, y + 8192 of M, from the split field H -> R[y][:], R[y + 8192][:] as
// one packed transform; columns kx >= KC are taken as zero (not read).  Block
// 8 l + x takes y = 4 (8 (l / 4) + x) + l % 4: the four blocks that share the
// 128-byte lines of H's rows 4t..4t+3 run back to back on one XCD.
template <int KC>
__global__ __launch_bounds__(kBigT) void k_big_row_c2r_h(const cplx *__restrict__ H,
                                                          double *__restrict__ out, Geo g,
                                                          const cplx *__restrict__ tw) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3;
  const int b = threadIdx.x, y = 4 * (((l >> 2) << 3) + xcd) + (l & 3);
  cplx ev[8], od[8];
  big_load_packed_split<KC>(ev, od, H, y, twid<+1>(tw, y), b, g.kpf);
  fft16k<+1>(ev, od, lds, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = b + kBigT * q;
    out[(size_t)y * N + x] = ev[q].x;
    out[(size_t)(y + kBigH) * N + x] = ev[q].y;
    out[(size_t)y * N + x + kBigH] = od[q].x;
    out[(size_t)(y + kBigH) * N + x + kBigH] = od[q].y;
  }
}

// rows y, y+1 of R as one packed forward transform: Z[y/2][k], k = 0..N-1
__global__ __launch_bounds__(kBigT) void k_big_row_r2c(const double *__restrict__ in,
                                                        cplx *__restrict__ zrows,
                                                        const cplx *__restrict__ tw) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  const int b = threadIdx.x, y = 2 * blockIdx.x;
  cplx ev[8], od[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = 2 * (b + kBigT * q);
    ev[q] = mk(in[(size_t)y * N + x], in[(size_t)(y + 1) * N + x]);
    od[q] = mk(in[(size_t)y * N + x + 1], in[(size_t)(y + 1) * N + x + 1]);
  }
  fft16k<-1>(ev, od, lds, b, tw);
  cplx *zr = zrows + (size_t)blockIdx.x * N;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    zr[b + kBigT * q] = ev[q];
    zr[b + kBigT * q + kBigH] = od[q];
  }
}

// rows y, y+1 of the bracket J = (a b - c d) / n4 as one packed forward
// transform: Z[y/2][k], k = 0..N-1 (the product a b is rounded first, as when
// it is stored)
__global__ __launch_bounds__(kBigT) void k_big_row_r2c_jac(const double *__restrict__ a,
                                                            const double *__restrict__ b,
                                                            const double *__restrict__ c,
                                                            const double *__restrict__ d,
                                                            double n4, cplx *__restrict__ zrows,
                                                            const cplx *__restrict__ tw) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  const int t = threadIdx.x, y = 2 * blockIdx.x;
  cplx ev[8], od[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = 2 * (t + kBigT * q);
    const size_t e0 = (size_t)y * N + x, e1 = e0 + N;
    const double2 a0 = *(const double2 *)(a + e0), a1 = *(const double2 *)(a + e1);
    const double2 b0 = *(const double2 *)(b + e0), b1 = *(const double2 *)(b + e1);
    const double2 c0 = *(const double2 *)(c + e0), c1 = *(const double2 *)(c + e1);
    const double2 d0 = *(const double2 *)(d + e0), d1 = *(const double2 *)(d + e1);
    ev[q] = mk((mul_rn(a0.x, b0.x) - c0.x * d0.x) / n4, (mul_rn(a1.x, b1.x) - c1.x * d1.x) / n4);
    od[q] = mk((mul_rn(a0.y, b0.y) - c0.y * d0.y) / n4, (mul_rn(a1.y, b1.y) - c1.y * d1.y) / n4);
  }
  fft16k<-1>(ev, od, lds, t, tw);
  cplx *zr = zrows + (size_t)blockIdx.x * N;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    zr[t + kBigT * q] = ev[q];
    zr[t + kBigT * q + kBigH] = od[q];
  }
}

// Column i of the half spectra, unpacked from the packed row transforms
// (A = (Z[k] + conj Z[-k])/2 of row 2p, B = (Z[k] - conj Z[-k])/(2i) of row
// 2p+1), forward transformed along y: S[i][:]
__global__ __launch_bounds__(kBigT) void k_big_col_fwd(const cplx *__restrict__ zrows,
                                                        cplx *__restrict__ spec,
                                                        const cplx *__restrict__ tw, int ncols) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  // The 8 columns of one 128-byte line go to consecutive workgroups of one
  // XCD (workgroups go to the XCDs in turn), so 7 of them find their rows'
  // lines in that XCD's L2.  The grid is rounded up to 64 workgroups.
  const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3;
  const int i = ((l >> 3) << 6) + (xcd << 3) + (l & 7), b = threadIdx.x;
  if (i >= ncols) return;
  cplx ev[8], od[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const cplx *zr = zrows + (size_t)(b + kBigT * q) * N;
    const cplx w = zr[i], wm = zr[(N - i) & (N - 1)];
    ev[q] = mk(0.5 * (w.x + wm.x), 0.5 * (w.y - wm.y));
    od[q] = mk(0.5 * (w.y + wm.y), -0.5 * (w.x - wm.x));
  }
  fft16k<-1>(ev, od, lds, b, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    spec[(size_t)i * N + b + kBigT * q] = ev[q];
    spec[(size_t)i * N + b + kBigT * q + kBigH] = od[q];
  }
}

// ---- the fused bracket pass of the n = 16384 step ----
constexpr int kBigNC = kBigN / 3 + 1;  // kept columns kx < nc

// DFT of length 16384, sign S, by decimation in frequency, for inputs in
// natural halves: in ev[q] = x[b + 1024 q], od[q] = x[b + 1024 q + 8192];
// out ev[q] = X[2(b + 1024 q)], od[q] = X[2(b + 1024 q) + 1].  The first
// radix-2 pass pairs x[j] with x[j + n/2] in registers; the halves are then
// two 8192-point transforms (even and odd outputs).
template <int S>
__device__ __forceinline__ void fft16k_dif(cplx ev[8], cplx od[8], cplx *lds, int b,
                                           const cplx *__restrict__ tw) {
#if R2D_BIG_TWLAST
  const cplx c0 = twid<S>(tw, b), c1 = twid<S>(tw, b + kBigT);
#endif
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const cplx e = ev[q], o = od[q];
#if R2D_BIG_TWLAST
    const cplx w = rot8<S>((q & 1) ? c1 : c0, q >> 1);
#else
    const cplx w = twid<S>(tw, b + kBigT * q);
#endif
    ev[q] = cadd(e, o);
    od[q] = cmul(csub(e, o), w);
  }
#if R2D_BIG_TWLAST
  const TwLast tls = tw_last<kBigH, 2>(tw, b);
  const TwLast *tl = &tls;
#else
  const TwLast *tl = nullptr;
#endif
  int b1 = b, b2 = b;
  asm volatile("" : "+v"(b1));
  fft_canon<kBigH, S, 1, 2>(ev, lds, b1, tw, tl);
  asm volatile("" : "+v"(b2));
  fft_canon<kBigH, S, 1, 2>(od, lds, b2, tw, tl);
}

// psi_x + i psi_y along row y from the kept columns of its two column
// transforms MX (d/dx) and MY (d/dy): ev[q] at x = b + 1024 q, od[q] at
// x + 8192 (unnormalised, as c2r).  The step's bracket pass and the n = 16384
// CFL_condition share it, so both see the same values.
__device__ __forceinline__ void big_grad_row(cplx ev[8], cplx od[8], const cplx *__restrict__ MX,
                                             const cplx *__restrict__ MY, int y, int b, int kp,
                                             cplx *lds, const cplx *__restrict__ tw) {
  big_load_packed<kBigNC>(ev, od, MX, y, MY, y, b, kp);
  fft16k<+1>(ev, od, lds, b, tw);
}
// max over the block of psi_x^2 + psi_y^2 into maxbits (bit pattern, atomicMax)
__device__ __forceinline__ void big_fold_max(const cplx ev[8], const cplx od[8],
                                             unsigned long long *maxbits) {
  double m = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    m = fmax(m, ev[q].x * ev[q].x + ev[q].y * ev[q].y);
    m = fmax(m, od[q].x * od[q].x + od[q].y * od[q].y);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
}

// The two brackets of a stage along row y, in one pass (poisson's
// arithmetic, src/pseudospec2D_Rot2D3C.f90:195-216, of
// src/Rot2D3C.para.f90:572-573): z = psi_x + i psi_y is inverse transformed
// in registers (big_grad_row); w_x, w_y (w = laplacian psi) and vz_x, vz_y
// come as real rows; in place of z the block forms
// W = J1 + i J2,  J1 = (psi_x w_y - psi_y w_x)/n^4,  J2 = (psi_x vz_y - psi_y vz_x)/n^4
// and transforms it forward (fft16k_dif): row y of Z = FFT(J1) + i FFT(J2).
// The two row spectra are split through LDS, A = (Z[k] + conj Z[n-k])/2 =
// FFT(J1)[k] and B = (Z[k] - conj Z[n-k])/(2i) = FFT(J2)[k], and stored for
// the kept k < nc in the tiled mixed layout (fa, fb; pitch kp), from which
// the column passes read whole 64-byte chunks.  With maxbits the row's
// maximum of psi_x^2 + psi_y^2 is folded in (CFL_condition's maximum).
__global__ __launch_bounds__(kBigT) void k_big_row_jac2(
    const cplx *__restrict__ MX, const cplx *__restrict__ MY, const double *__restrict__ wx,
    const double *__restrict__ wy, const double *__restrict__ vx, const double *__restrict__ vy,
    double n4, cplx *__restrict__ fa, cplx *__restrict__ fb, Geo g, const cplx *__restrict__ tw,
    unsigned long long *maxbits) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  // the 4 rows of one tile (one 64-byte chunk per kx) run on consecutive
  // workgroups of one XCD, so three of them find the chunks in its L2
  const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3;
  const int b = threadIdx.x, y = 4 * (((l >> 2) << 3) + xcd) + (l & 3);
  cplx ev[8], od[8];
  big_grad_row(ev, od, MX, MY, y, b, g.kpf, lds, tw);
  if (maxbits) big_fold_max(ev, od, maxbits);
  const size_t r = (size_t)y * N;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const size_t x0 = r + b + kBigT * q, x1 = x0 + kBigH;
    const double a0 = wx[x0], b0 = wy[x0], c0 = vx[x0], d0 = vy[x0];
    const double a1 = wx[x1], b1 = wy[x1], c1 = vx[x1], d1 = vy[x1];
    const double px0 = ev[q].x, py0 = ev[q].y, px1 = od[q].x, py1 = od[q].y;
    ev[q] = mk((mul_rn(px0, b0) - py0 * a0) / n4, (mul_rn(px0, d0) - py0 * c0) / n4);
    od[q] = mk((mul_rn(px1, b1) - py1 * a1) / n4, (mul_rn(px1, d1) - py1 * c1) / n4);
    if (q & 1) __builtin_amdgcn_sched_barrier(0);  // two slots' loads in flight at a time
  }
  fft16k_dif<-1>(ev, od, lds, b, tw);
  // even k (ev, index m = k/2), then odd k (od, m = (k-1)/2), through the
  // LDS the transform has freed; Z[n-k] sits at index (n/2 - m) mod n/2 of
  // the even half and n/2 - 1 - m of the odd half
  const int kp = g.kp;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int q = 0; q < 8; ++q) lds[b + kBigT * q] = h ? od[q] : ev[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {  // 2 (b + 1024 q) < nc only for q < 3
      const int m = b + kBigT * q, k = 2 * m + h;
      if (k < kBigNC) {
        const cplx z = h ? od[q] : ev[q], w = lds[h ? kBigH - 1 - m : (kBigH - m) & (kBigH - 1)];
        fa[tix(y, k, kp)] = mk(0.5 * (z.x + w.x), 0.5 * (z.y - w.y));
        fb[tix(y, k, kp)] = mk(0.5 * (z.y + w.y), -0.5 * (z.x - w.x));
      }
    }
    __syncthreads();
  }
}

// CFL_condition's maximum at n = 16384 from the column transforms MX, MY of
// d/dx psi and d/dy psi (the bracket pass's first transform)
__global__ __launch_bounds__(kBigT) void k_big_row_gradmax(const cplx *__restrict__ MX,
                                                            const cplx *__restrict__ MY, Geo g,
                                                            const cplx *__restrict__ tw,
                                                            unsigned long long *maxbits) {
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  cplx ev[8], od[8];
  big_grad_row(ev, od, MX, MY, blockIdx.x, threadIdx.x, g.kpf, lds, tw);
  big_fold_max(ev, od, maxbits);
}

// Column i < nc of a bracket's row spectra (tiled mixed layout, pitch kp),
// forward transformed along y: ev[q] = S[i][j], od[q] = S[i][j + 8192],
// j = b + 1024 q.  Rows 2p and 2p+1 share a 64-byte chunk: one 32-byte load.
__device__ __forceinline__ void big_col_fwd_tiled(cplx ev[8], cplx od[8],
                                                  const cplx *__restrict__ f, int i, int kp, int b,
                                                  cplx *lds, const cplx *__restrict__ tw) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const size_t t = tix(2 * (b + kBigT * q), i, kp);
    ev[q] = f[t];
    od[q] = f[t + 1];
  }
  fft16k<-1>(ev, od, lds, b, tw);
}
// Neighbouring columns (whose 64-byte chunks share 128-byte lines) go to
// consecutive workgroups of one XCD, 8 at a time (workgroups go to the XCDs
// in turn); grids are rounded up to 64 workgroups.
__device__ __forceinline__ int big_col_of_block() {
  const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3;
  return ((l >> 3) << 6) + (xcd << 3) + (l & 7);
}


// a = a * b;  a = (a - b * c) / n4;  max of a^2 + b^2 (bit pattern, atomicMax)
__global__ void k_real_mul(double *__restrict__ a, const double *__restrict__ b, size_t nn) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x)
    a[e] = a[e] * b[e];
}
__global__ void k_real_bracket(double *__restrict__ a, const double *__restrict__ b,
                               const double *__restrict__ c, double n4, size_t nn) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x)
    a[e] = (a[e] - b[e] * c[e]) / n4;
}
__global__ __launch_bounds__(256) void k_real_gradmax(const double *__restrict__ a,
                                                       const double *__restrict__ b, size_t nn,
                                                       unsigned long long *maxbits) {
  double m = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x)
    m = fmax(m, a[e] * a[e] + b[e] * b[e]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
}

// The stage update of src/Rot2D3C.para.f90:578-599 at element e = (i, j) with
// the step's forcing: rand_force draws (amplitude from the device dt) and/or
// dense arrays.  c1, c2 hold the stage's psi and vz; the rotation terms' d/dy
// of them (C3, C4 of the reference, :575-576) are formed here.  C6 comes from
// c6[e] or, with REG, is t6r.  The result goes to o1, o2, which may be c1, c2
// or ps, vz themselves (the element is read before it is written).
template <bool REG>
__device__ __forceinline__ void big_update_at(size_t e, int i, int j, const cplx *ps,
                                              const cplx *vz, const cplx *c1, const cplx *c2,
                                              cplx *o1, cplx *o2, const cplx *__restrict__ c5,
                                              const cplx *__restrict__ c6, cplx t6r,
                                              const StepConst &c, double dt, int o) {
  const int n = kBigN;
  const double kx = kxof(i, n), ky = kyof(j, n), ka2 = kx * kx + ky * ky;
  cplx a = mk(0.0, 0.0), b = mk(0.0, 0.0);
  if (ka2 <= c.g.kmax && ka2 > c.g.tiny) {
    cplx f = mk(0.0, 0.0), g = mk(0.0, 0.0);
    for (int q = 0; q < c.fp.nent; ++q)
      if (c.fp.e[q].i == i && c.fp.e[q].j == j) f = force_value(c.fp, q, dt, c.n2);
    for (int q = 0; q < c.fz.nent; ++q)
      if (c.fz.e[q].i == i && c.fz.e[q].j == j) g = force_value(c.fz, q, dt, c.n2);
    if (c.fpd) f = cadd(f, c.fpd[e]);
    if (c.fzd) g = cadd(g, c.fzd[e]);
    const cplx s1 = c1[e], s2 = c2[e];
    const cplx t3 = mk(-ky * s1.y, ky * s1.x), t4 = mk(-ky * s2.y, ky * s2.x);
    const double tmp1 = dt / (double)o, tmp2 = 1.0 / ka2;
    double tmp3 = 1.0 + (c.nu * ipow(ka2, c.nn) + c.hnu * ipow(tmp2, c.mm)) * tmp1;
    const cplx t5 = c5[e];
    cplx t = mk((-t5.x + c.two_omega * t4.x) * tmp2 + f.x, (-t5.y + c.two_omega * t4.y) * tmp2 + f.y);
    t = cadd(ps[e], cscale(t, tmp1));
    a = mk(t.x / tmp3, t.y / tmp3);
    tmp3 = 1.0 + (c.nuv * ipow(ka2, c.nnv) + c.hnuv * ipow(tmp2, c.mmv)) * tmp1;
    const cplx t6 = REG ? t6r : c6[e];
    t = mk(t6.x + c.two_omega * t3.x + g.x, t6.y + c.two_omega * t3.y + g.y);
    t = cadd(vz[e], cscale(t, tmp1));
    b = mk(t.x / tmp3, t.y / tmp3);
  }
  o1[e] = a;
  o2[e] = b;
}

// The stage update over the whole half plane (stage-by-stage form)
__global__ void k_big_update(const cplx *ps, const cplx *vz, const cplx *c1, const cplx *c2,
                             cplx *o1, cplx *o2, const cplx *__restrict__ c5,
                             const cplx *__restrict__ c6, StepConst c,
                             const double *__restrict__ dtp, int o) {
  const int n = c.g.n;
  const size_t total = (size_t)(n / 2 + 1) * n;
  const double dt = dtp ? *dtp : c.dt_fixed;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x)
    big_update_at<false>(e, (int)(e / n), (int)(e % n), ps, vz, c1, c2, o1, o2, c5, c6,
                         mk(0.0, 0.0), c, dt, o);
}


// The two halves of big_update_at for one element: psi's update from C5
// (reads c2 for C4 = d/dy c2) and vz's from C6 (reads c1 for C3 = d/dy c1),
// the same arithmetic.
__device__ __forceinline__ void big_update_psi_at(size_t e, int i, int j, const cplx *ps,
                                                  const cplx *c2, cplx *o1, cplx t5,
                                                  const StepConst &c, double dt, int o) {
  const int n = kBigN;
  const double kx = kxof(i, n), ky = kyof(j, n), ka2 = kx * kx + ky * ky;
  cplx a = mk(0.0, 0.0);
  if (ka2 <= c.g.kmax && ka2 > c.g.tiny) {
    cplx f = mk(0.0, 0.0);
    for (int q = 0; q < c.fp.nent; ++q)
      if (c.fp.e[q].i == i && c.fp.e[q].j == j) f = force_value(c.fp, q, dt, c.n2);
    if (c.fpd) f = cadd(f, c.fpd[e]);
    const cplx s2 = c2[e];
    const cplx t4 = mk(-ky * s2.y, ky * s2.x);
    const double tmp1 = dt / (double)o, tmp2 = 1.0 / ka2;
    const double tmp3 = 1.0 + (c.nu * ipow(ka2, c.nn) + c.hnu * ipow(tmp2, c.mm)) * tmp1;
    cplx t = mk((-t5.x + c.two_omega * t4.x) * tmp2 + f.x, (-t5.y + c.two_omega * t4.y) * tmp2 + f.y);
    t = cadd(ps[e], cscale(t, tmp1));
    a = mk(t.x / tmp3, t.y / tmp3);
  }
  o1[e] = a;
}
__device__ __forceinline__ void big_update_vz_at(size_t e, int i, int j, const cplx *vz,
                                                 const cplx *c1, cplx *o2, cplx t6,
                                                 const StepConst &c, double dt, int o) {
  const int n = kBigN;
  const double kx = kxof(i, n), ky = kyof(j, n), ka2 = kx * kx + ky * ky;
  cplx b = mk(0.0, 0.0);
  if (ka2 <= c.g.kmax && ka2 > c.g.tiny) {
    cplx g = mk(0.0, 0.0);
    for (int q = 0; q < c.fz.nent; ++q)
      if (c.fz.e[q].i == i && c.fz.e[q].j == j) g = force_value(c.fz, q, dt, c.n2);
    if (c.fzd) g = cadd(g, c.fzd[e]);
    const cplx s1 = c1[e];
    const cplx t3 = mk(-ky * s1.y, ky * s1.x);
    const double tmp1 = dt / (double)o, tmp2 = 1.0 / ka2;
    const double tmp3 = 1.0 + (c.nuv * ipow(ka2, c.nnv) + c.hnuv * ipow(tmp2, c.mmv)) * tmp1;
    cplx t = mk(t6.x + c.two_omega * t3.x + g.x, t6.y + c.two_omega * t3.y + g.y);
    t = cadd(vz[e], cscale(t, tmp1));
    b = mk(t.x / tmp3, t.y / tmp3);
  }
  o2[e] = b;
}

// Both brackets' column forwards with the stage update (the fused form's
// last kernel of a stage): column i of FFT(J1) (fa) carries psi's update,
// then column i of FFT(J2) (fb) vz's, so neither spectrum goes through HBM.
// psi's update writes o1 before vz's reads c1, so o1 must not alias c1 (the
// stages alternate between two pairs of stage fields).  Each transform's
// output goes half by half to LDS slots only its thread reads back.  With
// zero_tail the grid also covers the columns nc..n/2, which it sets to zero
// in o1, o2 (the last stage, which writes the state: the update mask).
__global__ __launch_bounds__(kBigT) void k_big_col_update2(
    const cplx *__restrict__ fa, const cplx *__restrict__ fb, int kp, const cplx *ps,
    const cplx *vz, const cplx *c1, const cplx *c2, cplx *o1, cplx *o2, StepConst c,
    const double *__restrict__ dtp, int o, int zero_tail, const cplx *__restrict__ tw) {
  constexpr int N = kBigN;
  __shared__ cplx lds[Shape<kBigH>::LDSN];
  const int i = big_col_of_block(), b = threadIdx.x;
  if (i >= kBigNC) {
    if (zero_tail && i <= N / 2) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const size_t e = (size_t)i * N + b + kBigT * q;
        o1[e] = mk(0.0, 0.0);
        o2[e] = mk(0.0, 0.0);
      }
    }
    return;
  }
  const double dt = dtp ? *dtp : c.dt_fixed;
#pragma unroll 1
  for (int f = 0; f < 2; ++f) {
    cplx ev[8], od[8];
    if (f) __syncthreads();  // every thread's psi slots read before the transform reuses them
    int bb = b;  // opaque copy: no index math hoisted across the two transforms
    asm volatile("" : "+v"(bb));
    big_col_fwd_tiled(ev, od, f ? fb : fa, i, kp, bb, lds, tw);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int q = 0; q < 8; ++q) lds[b + kBigT * q] = h ? od[q] : ev[q];
#pragma unroll 1
      for (int q = 0; q < 8; ++q) {
        const int j = b + kBigT * q + h * kBigH;
        const size_t e = (size_t)i * N + j;
        if (f == 0) big_update_psi_at(e, i, j, ps, c2, o1, lds[b + kBigT * q], c, dt, o);
        else big_update_vz_at(e, i, j, vz, c1, o2, lds[b + kBigT * q], c, dt, o);
      }
    }
  }
}

// ncols < n/2 + 1: the input is zero in the columns kx >= ncols (dealiased)
// mix: the mixed field between the two passes (default: the plan's first)
static int big_c2r(r2d3c_plan *p, const cplx *in, int op, double *out, hipStream_t s, int ncols,
                   cplx *mix) {
  if (ncols != kBigN / 3 + 1 && ncols != kBigN / 2 + 1)
    return set_err(R2D3C_EINVAL, "16384 c2r: column count n/3 + 1 or n/2 + 1");
  if (!mix) mix = p->mfull;
  k_big_col_inv<<<ncols, kBigT, 0, s>>>(in, mix, op, 1.0, p->g, p->tw);
  if (ncols == kBigN / 3 + 1)
    k_big_row_c2r<kBigN / 3 + 1><<<kBigN / 2, kBigT, 0, s>>>(mix, out, p->g, p->tw);
  else
    k_big_row_c2r<kBigN / 2 + 1><<<kBigN / 2, kBigT, 0, s>>>(mix, out, p->g, p->tw);
  LAUNCHCHK("c2r 16384");
  return 0;
}
// The step's dealiased c2r through the split column inverse: the column
// halves into mix, rows y and y + 8192 combined in the row pass.
#ifndef R2D_BIG_SPLIT
#define R2D_BIG_SPLIT 1
#endif
static int big_c2r_split(r2d3c_plan *p, const cplx *in, int op, double *out, hipStream_t s,
                         cplx *mix) {
  k_big_col_inv_h<<<16 * ((kBigNC + 7) / 8), kBigT2, 0, s>>>(in, mix, op, 1.0, p->g, p->tw,
                                                             kBigNC);
  k_big_row_c2r_h<kBigNC><<<kBigN / 2, kBigT, 0, s>>>(mix, out, p->g, p->tw);
  LAUNCHCHK("c2r 16384 (split)");
  return 0;
}
// columns kx < ncols of r2c((a b - c d) / n^4)
static int big_r2c_jac(r2d3c_plan *p, const double *a, const double *b, const double *c,
                       const double *d, cplx *out, int ncols, hipStream_t s) {
  const double n4 = (double)kBigN * kBigN * (double)kBigN * kBigN;
  k_big_row_r2c_jac<<<kBigN / 2, kBigT, 0, s>>>(a, b, c, d, n4, p->mfull, p->tw);
  k_big_col_fwd<<<(ncols + 63) / 64 * 64, kBigT, 0, s>>>(p->mfull, out, p->tw, ncols);
  LAUNCHCHK("r2c bracket 16384");
  return 0;
}
static int big_r2c(r2d3c_plan *p, const double *in, cplx *out, hipStream_t s) {
  k_big_row_r2c<<<kBigN / 2, kBigT, 0, s>>>(in, p->mfull, p->tw);
  k_big_col_fwd<<<(kBigN / 2 + 1 + 63) / 64 * 64, kBigT, 0, s>>>(p->mfull, out, p->tw,
                                                                  kBigN / 2 + 1);
  LAUNCHCHK("r2c 16384");
  return 0;
}
// c = r2c{(c2r(d_x a) c2r(d_y b) - c2r(d_y a) c2r(d_x b))/n^4} (poisson, :169-226)
static int big_poisson(r2d3c_plan *p, const cplx *a, const cplx *b, cplx *c, hipStream_t s) {
  const size_t nn = (size_t)kBigN * kBigN;
  const double n4 = (double)kBigN * kBigN * (double)kBigN * kBigN;
  int rc;
  if ((rc = big_c2r(p, a, OP_DX, p->real, s)) || (rc = big_c2r(p, b, OP_DY, p->real2, s))) return rc;
  k_real_mul<<<launch_grid(nn), 256, 0, s>>>(p->real, p->real2, nn);
  if ((rc = big_c2r(p, a, OP_DY, p->real2, s)) || (rc = big_c2r(p, b, OP_DX, p->real3, s))) return rc;
  k_real_bracket<<<launch_grid(nn), 256, 0, s>>>(p->real, p->real2, p->real3, n4, nn);
  LAUNCHCHK("poisson 16384");
  return big_r2c(p, p->real, c, s);
}
static int big_cfl(r2d3c_plan *p, const r2d3c_params *q, double cfl, const cplx *psi, double *d_dt,
                   hipStream_t s) {
  const size_t nn = (size_t)kBigN * kBigN;
  k_set_u64<<<1, 1, 0, s>>>(p->maxbits, 0ull);
  if (p->big_fused) {  // the fused step's own maximum (its first row transform)
    cplx *MX = p->mfull, *MY = MX + p->fsf;
    k_big_col_inv<<<kBigNC, kBigT, 0, s>>>(psi, MX, OP_DX, 1.0, p->g, p->tw);
    k_big_col_inv<<<kBigNC, kBigT, 0, s>>>(psi, MY, OP_DY, 1.0, p->g, p->tw);
    k_big_row_gradmax<<<kBigN, kBigT, 0, s>>>(MX, MY, p->g, p->tw, p->maxbits);
  } else {
    int rc;
    if ((rc = big_c2r(p, psi, OP_DX, p->real, s, kBigNC)) ||
        (rc = big_c2r(p, psi, OP_DY, p->real2, s, kBigNC)))
      return rc;
    k_real_gradmax<<<kReduceBlocks, 256, 0, s>>>(p->real, p->real2, nn, p->maxbits);
  }
  k_dt_from_max<<<1, 1, 0, s>>>(p->maxbits, make_dt_const(kBigN, q, cfl), d_dt);
  LAUNCHCHK("cfl 16384");
  return 0;
}
// One time step at n = 16384 (src/Rot2D3C.para.f90:399,545-611) on the
// dealiased columns, in one of two forms.  Fused bracket pass (default): per
// stage four c2r (d/dx, d/dy of laplacian(psi) and of vz) into real fields,
// the column transforms of d/dx, d/dy psi, k_big_row_jac2 (psi's derivatives
// in registers, both brackets and their row transforms), and two column
// forwards, the second carrying the update.  Stage by stage: six c2r (the two
// brackets share psi's derivatives), the two brackets formed inside their r2c
// row pass, and the update over the half plane.  The update forms d/dy of psi
// and vz itself.  Stage 4 reads psi, vz directly and stage 1 writes them, so
// the step copies no state; its output is already the masked state.  In CFL
// mode the step's dt comes from stage 4's d/dx, d/dy of psi, which are the
// fields CFL_condition transforms.
static int big_step(r2d3c_plan *p, const r2d3c_params *q, cplx *psi, cplx *vz,
                    const r2d3c_forcing *fo, int dt_mode, double dt_or_cfl, double *d_dt,
                    hipStream_t s) {
  StepConst c;
  make_step_const(p, q, dt_mode, dt_or_cfl, &c);
  if (fo) {
    int rc = make_force_set(&fo->psi, kBigN, &c.fp);
    if (!rc) rc = make_force_set(&fo->vz, kBigN, &c.fz);
    if (rc) return rc;
    c.fpd = (const cplx *)fo->dense_psi;
    c.fzd = (const cplx *)fo->dense_vz;
  }
  const size_t sz = (size_t)(kBigN / 2 + 1) * kBigN, nn = (size_t)kBigN * kBigN;
  cplx *C1 = p->ci, *C2 = C1 + sz, *C3 = C2 + sz, *C4 = C3 + sz, *C5 = C4 + sz, *C6 = C5 + sz;
  // six real fields: three real scratch fields, and C3, C4 and the first
  // transfers scratch field (each holds n^2 doubles)
  double *R1 = p->real, *R2 = p->real2, *R3 = p->real3, *R4 = (double *)C3, *R5 = (double *)C4,
         *R6 = (double *)p->spec3;
  const int nc = p->g.nc;
  double *dtp = d_dt ? d_dt : p->dtslot;
  int rc;
  if (dt_mode != R2D3C_DT_CFL) k_set_f64<<<1, 1, 0, s>>>(dtp, dt_or_cfl);
  const int g = launch_grid(sz);
  if (p->big_fused) {
    // Fused form: d/dx, d/dy of psi stay column transforms (MX, MY); the four
    // other derivative fields go through HBM as real fields; k_big_row_jac2
    // forms both brackets and their row transforms in one pass, and the
    // second bracket's column transform carries the update.
    cplx *MX = p->mfull, *MY = MX + p->fsf, *MS = MY + p->fsf, *FA = MS, *FB = FA + p->fs;
    const double n4 = (double)kBigN * kBigN * (double)kBigN * kBigN;
    const bool cfl = dt_mode == R2D3C_DT_CFL;
    // stage fields: C1, C2 after stages 4 and 2, C5, C6 after stage 3 (the
    // merged update must not write the fields it reads)
    for (int o = 4; o >= 1; --o) {
      const cplx *a = o == 4 ? psi : (o == 2 ? C5 : C1), *v = o == 4 ? vz : (o == 2 ? C6 : C2);
      if (R2D_BIG_SPLIT) {
        if ((rc = big_c2r_split(p, a, OP_LDX, R1, s, MS)) ||
            (rc = big_c2r_split(p, a, OP_LDY, R2, s, MS)) ||
            (rc = big_c2r_split(p, v, OP_DX, R3, s, MS)) ||
            (rc = big_c2r_split(p, v, OP_DY, R4, s, MS)))
          return rc;
      } else if ((rc = big_c2r(p, a, OP_LDX, R1, s, nc, MS)) ||
                 (rc = big_c2r(p, a, OP_LDY, R2, s, nc, MS)) ||
                 (rc = big_c2r(p, v, OP_DX, R3, s, nc, MS)) ||
                 (rc = big_c2r(p, v, OP_DY, R4, s, nc, MS))) {
        return rc;
      }
      k_big_col_inv<<<kBigNC, kBigT, 0, s>>>(a, MX, OP_DX, 1.0, p->g, p->tw);
      k_big_col_inv<<<kBigNC, kBigT, 0, s>>>(a, MY, OP_DY, 1.0, p->g, p->tw);
      if (o == 4 && cfl) k_set_u64<<<1, 1, 0, s>>>(p->maxbits, 0ull);
      // FA, FB overlay MS (two fields of n kp entries in 2 fsf): the four
      // c2r above are done with it
      k_big_row_jac2<<<kBigN, kBigT, 0, s>>>(MX, MY, R1, R2, R3, R4, n4, FA, FB, p->g, p->tw,
                                             o == 4 && cfl ? p->maxbits : nullptr);
      if (o == 4 && cfl)
        k_dt_from_max<<<1, 1, 0, s>>>(p->maxbits, make_dt_const(kBigN, q, dt_or_cfl), dtp);
      const int ncol = o == 1 ? kBigN / 2 + 1 : kBigNC;
      cplx *o1 = o == 1 ? psi : (o == 3 ? C5 : C1), *o2 = o == 1 ? vz : (o == 3 ? C6 : C2);
      k_big_col_update2<<<(ncol + 63) / 64 * 64, kBigT, 0, s>>>(
          FA, FB, p->g.kp, psi, vz, a, v, o1, o2, c, dtp, o, o == 1, p->tw);
    }
    LAUNCHCHK("step 16384");
    return 0;
  }
  for (int o = 4; o >= 1; --o) {
    const cplx *a = o == 4 ? psi : C1, *v = o == 4 ? vz : C2;
    if ((rc = big_c2r(p, a, OP_DX, R1, s, nc)) || (rc = big_c2r(p, a, OP_DY, R2, s, nc))) return rc;
    if (o == 4 && dt_mode == R2D3C_DT_CFL) {
      k_set_u64<<<1, 1, 0, s>>>(p->maxbits, 0ull);
      k_real_gradmax<<<kReduceBlocks, 256, 0, s>>>(R1, R2, nn, p->maxbits);
      k_dt_from_max<<<1, 1, 0, s>>>(p->maxbits, make_dt_const(kBigN, q, dt_or_cfl), dtp);
    }
    if ((rc = big_c2r(p, a, OP_LDX, R3, s, nc)) || (rc = big_c2r(p, a, OP_LDY, R4, s, nc)) ||
        (rc = big_c2r(p, v, OP_DX, R5, s, nc)) || (rc = big_c2r(p, v, OP_DY, R6, s, nc)))
      return rc;
    // poisson(psi, laplacian psi) and poisson(psi, vz) (:572-573)
    if ((rc = big_r2c_jac(p, R1, R4, R2, R3, C5, nc, s)) ||
        (rc = big_r2c_jac(p, R1, R6, R2, R5, C6, nc, s)))
      return rc;
    k_big_update<<<g, 256, 0, s>>>(psi, vz, a, v, o == 1 ? psi : C1, o == 1 ? vz : C2, C5, C6, c,
                                   dtp, o);
  }
  LAUNCHCHK("step 16384");
  return 0;
}
