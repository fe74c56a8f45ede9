// fft_core.h -- FP64 radix-8 Stockham FFT building blocks for gfx950.
//
// One length-N transform is done by T = N/8 threads.  Between passes the data
// goes through LDS; before the first and after the last pass it lives in
// registers in the "canonical" distribution: thread b holds elements
// b + T*q, q = 0..7, in v[q].  The first pass of a forward transform and the
// last pass of an inverse one (and vice versa) therefore see the same
// distribution, so a kernel can feed the output of one transform into the next
// without touching LDS or global memory.
//
// Pass schedule: radix 8, then radix 8 while at least 16 points remain, then
// one radix 2/4/8 pass.  2048 = 8*8*8*4, 256 = 8*8*4, 1024 = 8*8*8*2.
// Twiddles come from a table tw[t] = exp(-2 pi i t / N) built on the host in
// extended precision; the inverse uses its conjugate.  A radix-8 butterfly
// loads w, w^2, w^4 and forms w^3, w^5, w^6, w^7 by products (at most two
// extra roundings), which cuts the twiddle loads from 7 to 3.
#pragma once
#include <hip/hip_runtime.h>


namespace r2d {

// Sizes (bit mask of log2 N) whose radix-8 transforms run their middle
// passes in place (see fft_canon).
#ifndef R2D_INPLACE
#define R2D_INPLACE ((1 << 11) | (1 << 13))
#endif
// Grids from this size up pad their LDS arrays with one slot per 8 entries,
// smaller ones with one per 16 (see pad below).  Off by default: at 2048 the
// row kernel measured 1.5 us slower with it and the column kernels gained
// less than the run-to-run spread.
#ifndef R2D_PAD8_MIN_N
#define R2D_PAD8_MIN_N (1 << 30)
#endif

typedef double2 cplx;

__device__ __forceinline__ cplx mk(double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
// a product rounded on its own: never fused into the add that uses it
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cplx cconj(cplx a) { return mk(a.x, -a.y); }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return mk(a.x * s, a.y * s); }
// a * (S i)
template <int S> __device__ __forceinline__ cplx mul_si(cplx a) {
  return S > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

template <int N> struct Shape {
  static_assert(N >= 16 && (N & (N - 1)) == 0, "N must be a power of two >= 16");
  static constexpr int T = N / 8;                          // threads per transform
  static constexpr int LOG = ilog2(N);
  static constexpr int NMID = (LOG - 4) / 3;               // middle radix-8 passes
  static constexpr int RL = 1 << (LOG - 3 - 3 * NMID);     // last radix: 2, 4 or 8
  static constexpr int PS = N >= R2D_PAD8_MIN_N ? 3 : 4;  // one pad slot per 2^PS
  // in-place passes (fft_canon): sizes in the bit mask R2D_INPLACE of log2 N
  static constexpr bool IP = ((R2D_INPLACE) >> LOG) & 1;
  static constexpr int SK2 = RL == 2 ? 2 : 0;  // in-place slot skew per N/16 positions
  static constexpr int LDSN = N + (N >> PS) + (IP ? RL + 15 * SK2 : 0);  // padded LDS slots
  static constexpr int WR = 512 + (512 >> PS);              // one wave's four-step region
};

// LDS slot of element i of a length-N transform: one pad slot every 2^PS
// entries.  With one every 8, every access pattern of the passes below is
// free of bank conflicts: 8 lanes writing 8 consecutive elements each (the
// first pass) land on slots 9b + m, distinct modulo 8, and the strided reads
// stay consecutive within aligned lane groups.  One every 16 leaves the first
// pass's writes 2-way conflicted (38% slower 16-byte LDS traffic in an access
// test) but takes 1/16 less LDS, which keeps more blocks of the LDS-bound
// small-grid kernels resident, and the transforms here are not bound by LDS
// bandwidth, so it is the default.
template <int N> __device__ __forceinline__ int pad(int i) { return i + (i >> Shape<N>::PS); }

// In-place form of the radix-8 passes (Shape<N>::IP): the same butterflies
// with the same twiddles as the Stockham passes below, so the same results
// bit for bit, but each middle pass writes its outputs back to the LDS slots
// it read, so no barrier is needed between its reads and its writes.  The
// first pass stores thread b's outputs m at position m + 8 rho(b), where
// rho(b) reverses the digits of b (the input index b + T m has digits
// m_1 = m of weight T, then m_2 .. m_{NMID+1} of weights N/8^s and a last
// digit of radix RL and weight 1); middle pass s (span L) works on positions
// k 8L + j + mL (thread b: k = b / L, j = b % L) and the last pass on
// j + m N/RL.  Position P lies in slot pad(P) + P / (N/RL), plus 2 P/(N/16)
// when RL = 2: the skew makes the first pass's writes (whose positions are
// multiples of N/16 apart across lanes) and every other pass's accesses free
// of bank conflicts.
template <int N> __device__ __forceinline__ int ip_rho(int b) {
  using F = Shape<N>;
  int rho = 0, f = 1;
#pragma unroll
  for (int s = 2; s <= F::NMID + 1; ++s) {
    rho += ((b >> (F::LOG - 3 * s)) & 7) * f;
    f *= 8;
  }
  return rho + (b & (F::RL - 1)) * f;
}
template <int N> __device__ __forceinline__ int ip_slot(int P) {
  return pad<N>(P) + P / (N / Shape<N>::RL) + Shape<N>::SK2 * (P / (N / 16));
}
// slot(P0 + m L) - slot(P0) for a middle pass of span L (P0 = k 8L + j)
template <int N> constexpr int ip_step(int L, int m) {
  constexpr int PS = Shape<N>::PS;
  return (L >= (1 << PS) ? m * (L + (L >> PS)) : 8 * m + ((8 * m) >> PS)) +
         (L == N / 16 ? Shape<N>::SK2 * m : 0);
}

template <int S> __device__ __forceinline__ cplx twid(const cplx *__restrict__ tw, int t) {
  cplx w = tw[t];
  if (S > 0) w.y = -w.y;
  return w;
}
// a raw table value exp(-2 pi i t / N) as the sign-S twiddle exp(S 2 pi i t / N)
template <int S> __device__ __forceinline__ cplx tw_s(cplx w) { return S > 0 ? mk(w.x, -w.y) : w; }
// a * exp(S i pi u / 4), u = 0..3 (u = 2 exact; u = 1, 3 one rounding per part)
template <int S> __device__ __forceinline__ cplx rot8(cplx a, int u) {
  const double c = 0.70710678118654752440;
  if (u == 1) return mk(c * (a.x - S * a.y), c * (a.y + S * a.x));
  if (u == 2) return S > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
  if (u == 3) return mk(-c * (a.x + S * a.y), c * (S * a.x - a.y));
  return a;
}

// Twiddles of one thread that every transform of a kernel shares: the last
// pass's table values at j = b (exp(-2 pi i t/N) for t = b, 2b, 4b: the
// butterflies j = b + T u take them times exp(S i pi u/4), exp(S i pi u/2),
// exp(S i pi u), and a mirrored pass's j = 2T - b takes S i conj(w_b) and
// -conj(w_2b)).  Loaded once, they save the last pass's table loads in every
// transform that takes them.
struct TwLast {
  cplx l1, l2, l4;
};
template <int N, int TS = 1>
__device__ __forceinline__ TwLast tw_last(const cplx *__restrict__ tw, int b) {
  constexpr int RL = Shape<N>::RL;
  TwLast t;
  t.l1 = tw[TS * b];
  t.l2 = RL >= 4 ? tw[TS * 2 * b] : t.l1;
  t.l4 = RL == 8 ? tw[TS * 4 * b] : t.l1;
  return t;
}

template <int S> __device__ __forceinline__ void dft2(cplx &a, cplx &b) {
  cplx t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

// Z: bit i set when input i is known to be zero; its additions are left
// out (the values agree with the full butterfly up to the sign of zero).
template <int S, int Z = 0>
__device__ __forceinline__ void dft4(cplx &x0, cplx &x1, cplx &x2, cplx &x3) {
  cplx a0, a1, a2, a3;
  if constexpr ((Z & 5) == 5) { a0 = mk(0.0, 0.0); a1 = a0; }
  else if constexpr (Z & 4) { a0 = x0; a1 = x0; }
  else if constexpr (Z & 1) { a0 = x2; a1 = mk(-x2.x, -x2.y); }
  else { a0 = cadd(x0, x2); a1 = csub(x0, x2); }
  if constexpr ((Z & 10) == 10) { a2 = mk(0.0, 0.0); a3 = a2; }
  else if constexpr (Z & 8) { a2 = x1; a3 = mul_si<S>(x1); }
  else if constexpr (Z & 2) { a2 = x3; a3 = mul_si<S>(mk(-x3.x, -x3.y)); }
  else { a2 = cadd(x1, x3); a3 = mul_si<S>(csub(x1, x3)); }
  x0 = cadd(a0, a2);
  x2 = csub(a0, a2);
  x1 = cadd(a1, a3);
  x3 = csub(a1, a3);
}

// y_k = sum_m x_m exp(S 2 pi i m k / 8); Z: bit m set when x_m is known zero
template <int S, int Z = 0> __device__ __forceinline__ void dft8(cplx *x) {
  constexpr int ZE = (Z & 1) | ((Z >> 1) & 2) | ((Z >> 2) & 4) | ((Z >> 3) & 8);
  constexpr int ZO = ((Z >> 1) & 1) | ((Z >> 2) & 2) | ((Z >> 3) & 4) | ((Z >> 4) & 8);
  cplx e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
  cplx o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
  dft4<S, ZE>(e0, e1, e2, e3);
  dft4<S, ZO>(o0, o1, o2, o3);
  // exp(S i pi/4) = c (1 + S i) and exp(3 S i pi/4) = c (-1 + S i): the factor
  // c is folded into the final additions as fused multiply-adds
  const double c = 0.70710678118654752440;
  const cplx t1 = mk(o1.x - S * o1.y, o1.y + S * o1.x);
  const cplx t3 = mk(-o3.x - S * o3.y, S * o3.x - o3.y);
  o2 = mul_si<S>(o2);  // * exp(S i pi/2)
  x[0] = cadd(e0, o0); x[4] = csub(e0, o0);
  x[1] = mk(fma(c, t1.x, e1.x), fma(c, t1.y, e1.y));
  x[5] = mk(fma(-c, t1.x, e1.x), fma(-c, t1.y, e1.y));
  x[2] = cadd(e2, o2); x[6] = csub(e2, o2);
  x[3] = mk(fma(c, t3.x, e3.x), fma(c, t3.y, e3.y));
  x[7] = mk(fma(-c, t3.x, e3.x), fma(-c, t3.y, e3.y));
}

template <int R, int S> __device__ __forceinline__ void dftR(cplx *x) {
  if constexpr (R == 2) dft2<S>(x[0], x[1]);
  else if constexpr (R == 4) dft4<S>(x[0], x[1], x[2], x[3]);
  else dft8<S>(x);
}

// Orders this wave's LDS writes before its later LDS reads (compiler and
// memory model); LDS instructions of one wave execute in order.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// The transforms' LDS hand-offs: a block barrier, or (WAVE) only this
// wave's ordering, for a transform whose threads and LDS area lie in one
// wavefront (no s_barrier: other waves of the block go on).
template <bool WAVE> __device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) wave_lds_sync();
  else __syncthreads();
}

// Unnormalised DFT of length N, sign S (-1 forward, +1 inverse), of the
// canonical register set v of thread b (0 <= b < N/8).  ZQ: bit q set when
// v[q] is zero on entry for every thread (its additions in the first pass are
// left out).  MIR (radix-4 last pass only): the thread's second last-pass
// butterfly is the mirror j' = 2T - b of its first, j = b (j' = T for b = 0),
// so on return v[u + 2m] = X[j_u + 2T m] and every output's mirror X[N - k] is
// in the same thread.  CLOSE: end with the closing barrier.  tl: the thread's
// last-pass twiddles (tw_last), or null to load them from the table.  lds points at this
// transform's LDS area; consecutive slots are IL entries apart, so IL
// transforms can share one interleaved LDS array.  tw holds TS*N entries
// (exp(-2 pi i t / (TS N))): a longer transform's table serves its halves.  All threads of the block
// must call this together (it contains block barriers).  The last pass reads
// the LDS area and a closing barrier follows it, so the area is free for any
// thread to write again on return.  WAVE: the transform's threads and LDS
// area lie in one wavefront; wave-level ordering replaces the barriers.
template <int N, int S, int IL, int TS = 1, int ZQ = 0, bool MIR = false, bool CLOSE = true,
          bool WAVE = false>
__device__ __forceinline__ void fft_canon(cplx v[8], cplx *lds, int b,
                                          const cplx *__restrict__ tw,
                                          const TwLast *tl = nullptr) {
  using F = Shape<N>;
  constexpr int T = F::T;
  // LDS slots as one base per pass plus compile-time offsets: with Q = 2^PS
  // (the pad period) dividing T, pad(i + m*K) = pad(i) + m*(K + K/Q) for K = T,
  // N/RL and spans L >= Q, and the span-8 writes of Q = 16 have offsets
  // 8m + m/2 (their bases are 0..7 modulo 64).  Used where every offset fits
  // the 16-bit immediate of an LDS instruction; otherwise (small or very long
  // transforms) pad() per slot.
  constexpr int RL0 = F::RL, PS = F::PS, PQ = 1 << PS;
  constexpr bool LIN = T % PQ == 0 && (N / RL0) % PQ == 0 &&
                       16 * IL * 7 * (T + (T >> PS)) <= 65535 &&
                       16 * IL * (RL0 - 1) * (N / RL0 + ((N / RL0) >> PS)) <= 65535;
  constexpr bool IP = F::IP;
  // pass 1: radix 8, span 1; inputs b + m*T are exactly v[m]
  dft8<S, ZQ>(v);
  {
    // Stockham: pad(8b + m) = pad(8b) + m; in place: slot(8 rho + m) = slot(8 rho) + m
    cplx *w = lds + (IP ? ip_slot<N>(8 * ip_rho<N>(b)) : pad<N>(8 * b)) * IL;
#pragma unroll
    for (int m = 0; m < 8; ++m) w[m * IL] = v[m];
  }
  lds_sync<WAVE>();
  int L = 8;
#pragma unroll
  for (int p = 0; p < F::NMID; ++p) {
    const int j = b & (L - 1);
    const int st = (N / (8 * L)) * j;
    // the pass's table twiddles are loaded first: they stay in flight over
    // the LDS reads and the barrier
    const cplx w1 = twid<S>(tw, TS * (st)), w2 = twid<S>(tw, TS * (2 * st)), w4 = twid<S>(tw, TS * (4 * st));
    cplx x[8];
    // in place: the butterfly's positions k 8L + j + mL, read and written back
    cplx *const ipw = lds + ip_slot<N>((b - j) * 8 + j) * IL;
    if constexpr (IP) {
#pragma unroll
      for (int m = 0; m < 8; ++m) x[m] = ipw[ip_step<N>(L, m) * IL];
    } else if constexpr (LIN) {
      const cplx *r = lds + pad<N>(b) * IL;
#pragma unroll
      for (int m = 0; m < 8; ++m) x[m] = r[m * (T + (T >> PS)) * IL];
    } else {
#pragma unroll
      for (int m = 0; m < 8; ++m) x[m] = lds[pad<N>(b + m * T) * IL];
    }
    if constexpr (!IP) lds_sync<WAVE>();
    {  // three table loads; the other powers by products (error <= 2 ulp more)
      const cplx w3 = cmul(w1, w2);
      x[1] = cmul(x[1], w1); x[2] = cmul(x[2], w2); x[3] = cmul(x[3], w3);
      x[4] = cmul(x[4], w4); x[5] = cmul(x[5], cmul(w1, w4));
      x[6] = cmul(x[6], cmul(w2, w4)); x[7] = cmul(x[7], cmul(w3, w4));
    }
    dft8<S>(x);
    const int base = (b - j) * 8 + j;
    if constexpr (IP) {
#pragma unroll
      for (int m = 0; m < 8; ++m) ipw[ip_step<N>(L, m) * IL] = x[m];
    } else if constexpr (LIN) {
      cplx *w = lds + pad<N>(base) * IL;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        w[(L >= PQ ? m * (L + (L >> PS)) : 8 * m + ((8 * m) >> PS)) * IL] = x[m];
    } else {
#pragma unroll
      for (int m = 0; m < 8; ++m) lds[pad<N>(base + m * L) * IL] = x[m];
    }
    lds_sync<WAVE>();
    L *= 8;
  }
  // last pass: radix RL, span N/RL; outputs j + m*N/RL land in canonical order
  constexpr int RL = F::RL, U = 8 / RL;
  static_assert(!MIR || RL == 4, "mirrored last pass: radix 4 only");
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = MIR && u == 1 ? (b ? 2 * T - b : T) : b + T * u;
    // table twiddles ahead of the LDS reads, or the thread's shared ones
    cplx w1, w2, w4;
    if (tl) {
      const cplx a1 = tw_s<S>(tl->l1), a2 = tw_s<S>(tl->l2), a4 = tw_s<S>(tl->l4);
      if (MIR && u == 1) {  // j = 2T - b, or T for b = 0
        const double c = 0.70710678118654752440;
        w1 = b ? mul_si<S>(cconj(a1)) : mk(c, S * c);
        w2 = b ? mk(-a2.x, a2.y) : mk(0.0, (double)S);
        w4 = w1;
      } else {  // j = b + T u; u <= 1 wherever w2 or w4 is used (RL >= 4)
        w1 = rot8<S>(a1, u);
        w2 = u ? mul_si<S>(a2) : a2;
        w4 = a4;
      }
    } else {
      w1 = twid<S>(tw, TS * (j));
      w2 = RL >= 4 ? twid<S>(tw, TS * (2 * j)) : w1;
      w4 = RL == 8 ? twid<S>(tw, TS * (4 * j)) : w1;
    }
    cplx x[RL];
    if constexpr (IP) {  // slot(j + m N/RL) = slot(j) + m (N/RL + N/RL/Q + 1 + 8 SK2)
      const cplx *r = lds + ip_slot<N>(j) * IL;
#pragma unroll
      for (int m = 0; m < RL; ++m)
        x[m] = r[m * (N / RL + ((N / RL) >> PS) + 1 + 8 * F::SK2) * IL];
    } else if constexpr (LIN) {
      const cplx *r = lds + pad<N>(j) * IL;
#pragma unroll
      for (int m = 0; m < RL; ++m) x[m] = r[m * (N / RL + ((N / RL) >> PS)) * IL];
    } else {
#pragma unroll
      for (int m = 0; m < RL; ++m) x[m] = lds[pad<N>(j + m * (N / RL)) * IL];
    }
    if constexpr (RL == 2) {
      x[1] = cmul(x[1], w1);
    } else if constexpr (RL == 4) {
      x[1] = cmul(x[1], w1); x[2] = cmul(x[2], w2); x[3] = cmul(x[3], cmul(w1, w2));
    } else {
      const cplx w3 = cmul(w1, w2);
      x[1] = cmul(x[1], w1); x[2] = cmul(x[2], w2); x[3] = cmul(x[3], w3);
      x[4] = cmul(x[4], w4); x[5] = cmul(x[5], cmul(w1, w4));
      x[6] = cmul(x[6], cmul(w2, w4)); x[7] = cmul(x[7], cmul(w3, w4));
    }
    dftR<RL, S>(x);
#pragma unroll
    for (int m = 0; m < RL; ++m) v[u + m * U] = x[m];
  }
  if constexpr (CLOSE) lds_sync<WAVE>();
}

// ---------------------------------------------------------------------------
// Four-step transforms for N >= 512: N = 512 * W.  Each wavefront (64 lanes x
// 8 points) does a 512-point sub-transform whose two exchanges stay inside
// the wave (ordered LDS traffic of one wave needs no s_barrier), and one
// radix-W exchange crosses the W waves of the transform.  Two block barriers
// per transform instead of six, so waves drift apart and overlap their VALU
// and LDS work.
//   DIT: input  v[q] of thread b = 64 w + l holds x[w + W (l + 64 q)]
//        ("interleaved"), output canonical (X[b + (N/8) q]).
//   DIF: input canonical, output interleaved.
// A forward DIF followed by an inverse DIT (or the reverse) therefore needs
// no redistribution between the two.  LDS: W regions of Shape<N>::WR slots (512 + pad)
// = N + N/16 slots, contiguous per transform.
// ---------------------------------------------------------------------------


// The three table twiddles w^st, w^2st, w^4st of a radix-8 pass; loaded
// ahead of the pass's LDS reads so that they are in flight meanwhile.
struct Tw3 {
  cplx w1, w2, w4;
};
template <int S>
__device__ __forceinline__ Tw3 tw3_load(const cplx *__restrict__ tw, int st) {
  return Tw3{twid<S>(tw, st), twid<S>(tw, 2 * st), twid<S>(tw, 4 * st)};
}
__device__ __forceinline__ void twiddle8(cplx *x, const Tw3 &t) {
  const cplx w1 = t.w1, w2 = t.w2, w4 = t.w4;
  const cplx w3 = cmul(w1, w2);
  x[1] = cmul(x[1], w1); x[2] = cmul(x[2], w2); x[3] = cmul(x[3], w3);
  x[4] = cmul(x[4], w4); x[5] = cmul(x[5], cmul(w1, w4));
  x[6] = cmul(x[6], cmul(w2, w4)); x[7] = cmul(x[7], cmul(w3, w4));
}

// 512-point transform of one wavefront: lane l holds elements l + 64 q in
// v[q] on entry and exit.  L: this wave's Shape<512 TS>::WR-slot region.  Twiddles
// W_512^t = tw[t * TS] (tw is the table of length 512 * TS).
template <int S, int TS>
__device__ __forceinline__ void fft_wave512(cplx v[8], cplx *L, int l,
                                            const cplx *__restrict__ tw) {
  dft8<S>(v);  // span 1
#pragma unroll
  for (int m = 0; m < 8; ++m) L[pad<512 * TS>(l * 8 + m)] = v[m];
  wave_lds_sync();
  {  // span 8
    const int j = l & 7;
    const Tw3 t = tw3_load<S>(tw, 8 * j * TS);
    cplx x[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) x[m] = L[pad<512 * TS>(l + 64 * m)];
    wave_lds_sync();
    twiddle8(x, t);
    dft8<S>(x);
    const int base = (l - j) * 8 + j;
#pragma unroll
    for (int m = 0; m < 8; ++m) L[pad<512 * TS>(base + 8 * m)] = x[m];
    wave_lds_sync();
  }
  {  // span 64, last
    const Tw3 t = tw3_load<S>(tw, l * TS);
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = L[pad<512 * TS>(l + 64 * m)];
    twiddle8(v, t);
    dft8<S>(v);
    wave_lds_sync();
  }
}

template <int N, int S>
__device__ __forceinline__ void fft_dit4(cplx v[8], cplx *P, int b, const cplx *__restrict__ tw) {
  constexpr int W = N / 512, T = N / 8, U = 8 / W;
  const int w = b >> 6, l = b & 63;
  cplx *L = P + w * Shape<N>::WR;
  fft_wave512<S, W>(v, L, l, tw);
  if constexpr (W > 1) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k2 = l + 64 * q;
      L[pad<N>(k2)] = w ? cmul(v[q], twid<S>(tw, w * k2)) : v[q];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k2 = b + T * u;
      cplx x[W];
#pragma unroll
      for (int w2 = 0; w2 < W; ++w2) x[w2] = P[w2 * Shape<N>::WR + pad<N>(k2)];
      dftR<W, S>(x);
#pragma unroll
      for (int k1 = 0; k1 < W; ++k1) v[u + U * k1] = x[k1];
    }
  }
  __syncthreads();
}

template <int N, int S>
__device__ __forceinline__ void fft_dif4(cplx v[8], cplx *P, int b, const cplx *__restrict__ tw) {
  constexpr int W = N / 512, T = N / 8, U = 8 / W;
  const int w = b >> 6, l = b & 63;
  if constexpr (W > 1) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n2 = b + T * u;
      cplx x[W];
#pragma unroll
      for (int n1 = 0; n1 < W; ++n1) x[n1] = v[u + U * n1];
      dftR<W, S>(x);
#pragma unroll
      for (int k1 = 0; k1 < W; ++k1)
        P[k1 * Shape<N>::WR + pad<N>(n2)] = k1 ? cmul(x[k1], twid<S>(tw, n2 * k1)) : x[k1];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = P[w * Shape<N>::WR + pad<N>(l + 64 * q)];
  }
  fft_wave512<S, W>(v, P + w * Shape<N>::WR, l, tw);
  __syncthreads();
}

// Element index held in v[q] by thread b in the interleaved distribution
// (the canonical one for N < 512, or with CS, where fft_canon is used both
// ways).
template <int N, bool CS = false> __device__ __forceinline__ int ilv(int b, int q) {
  if constexpr (N >= 512 && !CS) {
    return (b >> 6) + (N / 512) * ((b & 63) + 64 * q);
  } else {
    return b + (N / 8) * q;
  }
}

// DIF / DIT dispatch, two radix schedules: the four-step form for N >= 512
// (512-point wave-local sub-transforms and one radix N/512 exchange), or, below
// 512 or with CS, fft_canon (radix-8 passes and a last radix 2, 4 or 8,
// canonical order both ways).  P: the transform's LDS area with slot stride
// IL (IL = 1 required for the four-step form).
// tl: the thread's last-pass twiddles for the radix-8 Stockham form (unused by
// the four-step form).
template <int N, int S, int IL, bool CS = false>
__device__ __forceinline__ void fft_dif(cplx v[8], cplx *P, int b, const cplx *__restrict__ tw,
                                        const TwLast *tl = nullptr) {
  if constexpr (N >= 512 && !CS) fft_dif4<N, S>(v, P, b, tw);
  else fft_canon<N, S, IL>(v, P, b, tw, tl);
}
template <int N, int S, int IL, bool CS = false>
__device__ __forceinline__ void fft_dit(cplx v[8], cplx *P, int b, const cplx *__restrict__ tw,
                                        const TwLast *tl = nullptr) {
  if constexpr (N >= 512 && !CS) fft_dit4<N, S>(v, P, b, tw);
  else fft_canon<N, S, IL>(v, P, b, tw, tl);
}

// Packed half-spectrum loader for a row transform.  Two real rows a, b with
// half spectra A, B (kx = 0..n/2) are transformed as one complex row
// z = a + i b, whose full spectrum is Z[k] = A[k] + i B[k] (k <= n/2) and
// Z[n-k] = conj(A[k]) + i conj(B[k]).  Imaginary parts at k = 0 and k = n/2
// are dropped first, as FFTW's c2r does.
template <int N>
__device__ __forceinline__ cplx packed_z(int k, cplx A, cplx B, bool mirrored) {
  if (!mirrored) {
    if (k == 0 || k == N / 2) { A.y = 0.0; B.y = 0.0; }
    return mk(A.x - B.y, A.y + B.x);
  }
  return mk(A.x + B.y, B.x - A.y);
}
// Slot of (y, kx) in a tiled mixed field of column pitch kp: M[y/4][kx][y%4].
__device__ __forceinline__ size_t tix(int y, int kx, int kp) {
  return (size_t)(y >> 2) * (size_t)(kp * 4) + (size_t)kx * 4 + (y & 3);
}

}  // namespace r2d
