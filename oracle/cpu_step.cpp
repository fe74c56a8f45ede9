// cpu_step.cpp -- compiled CPU restatement of the Rot2D3C time step (test
// infrastructure: the benchmark's CPU baseline and a cross-check of the
// numpy oracle; nothing in the product loads it).
//
// It runs the reference's step as the reference does it, on the host cores
// with OpenMP: every stage makes two Poisson brackets from full 2D transforms
// (4 complex-to-real + 1 real-to-complex each, 40 transforms per step;
// src/Rot2D3C.para.f90:571-576, src/pseudospec2D_Rot2D3C.f90:169-226), then
// the semi-implicit update (src/Rot2D3C.para.f90:578-599) and the final
// masked copy (:610-611).  The 2D transforms follow src/fftp2D.f90:133-290:
// c2r is complex along ky then real along x (ignoring the imaginary parts at
// kx = 0 and n/2), r2c is real along x then complex along ky, both
// unnormalised.  The 1D transforms are this file's own radix-4 Stockham
// FFT (autosorting, out of place; a last radix-2 pass when log2 n is odd)
// with twiddles from the extended-precision circle; two real rows travel as
// one complex row.  Layouts are the reference's one-rank arrays: spectral
// [kx][ky] (ky fastest, kx = 0..n/2), real [y][x].
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

using cd = std::complex<double>;

struct Fft {
  int n = 0, logn = 0;
  std::vector<cd> tw;  // exp(-2 pi i t / n), t < n

  explicit Fft(int n_) : n(n_) {
    while ((1 << logn) < n) ++logn;
    tw.resize(n);
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; ++t) {
      const long double a = twopi * (long double)t / (long double)n;
      tw[t] = cd((double)cosl(a), (double)(-sinl(a)));
    }
  }

  // In-place unnormalised transform of x (length n) with work buffer w:
  // sign -1 forward (e^{-i}), +1 backward (e^{+i}).  Stockham DIF: a pass of
  // sub-length L and stride s reads x[q + s (p + k L/4)] and writes
  // y[q + s (4 p + k)] times w_L^{k p}, k = 0..3.
  void run(cd *x, cd *w, int sign) const {
    cd *a = x, *b = w;
    int L = n, s = 1;
    while (L >= 4) {
      const int L4 = L / 4, step = n / L;
      for (int p = 0; p < L4; ++p) {
        cd w1 = tw[(size_t)p * step], w2 = tw[(size_t)2 * p * step], w3 = tw[(size_t)3 * p * step];
        if (sign > 0) { w1 = std::conj(w1); w2 = std::conj(w2); w3 = std::conj(w3); }
        const cd *x0 = a + (size_t)s * p, *x1 = a + (size_t)s * (p + L4),
                 *x2 = a + (size_t)s * (p + 2 * L4), *x3 = a + (size_t)s * (p + 3 * L4);
        cd *y = b + (size_t)s * 4 * p;
        for (int q = 0; q < s; ++q) {
          const cd u0 = x0[q], u1 = x1[q], u2 = x2[q], u3 = x3[q];
          const cd s02 = u0 + u2, d02 = u0 - u2, s13 = u1 + u3, d13 = u1 - u3;
          // -i d13 for the forward sign, +i d13 for the backward one
          const cd r13 = sign < 0 ? cd(d13.imag(), -d13.real()) : cd(-d13.imag(), d13.real());
          y[q] = s02 + s13;
          y[q + s] = w1 * (d02 + r13);
          y[q + 2 * s] = w2 * (s02 - s13);
          y[q + 3 * s] = w3 * (d02 - r13);
        }
      }
      std::swap(a, b);
      L /= 4;
      s *= 4;
    }
    if (L == 2) {
      for (int q = 0; q < s; ++q) {
        const cd u0 = a[q], u1 = a[q + s];
        b[q] = u0 + u1;
        b[q + s] = u0 - u1;
      }
      std::swap(a, b);
    }
    if (a != x) std::memcpy(x, a, sizeof(cd) * n);
  }
};

struct Step {
  int n, nh;  // nh = n/2 + 1 columns
  Fft fft;
  double kmax, tiny;
  std::vector<cd> mix;  // [kx][y] after the column pass, nh x n

  explicit Step(int n_) : n(n_), nh(n_ / 2 + 1), fft(n_), mix((size_t)nh * n_) {
    kmax = ((double)n / 3.0) * ((double)n / 3.0);  // src/Rot2D3C.para.f90:213
    tiny = 0.000001;
  }
  double ka(int i) const { return (double)(i < n / 2 ? i : i - n); }  // :220-223
  double kx(int i) const { return i == n / 2 ? -(double)(n / 2) : (double)i; }

  // c2r: spectral [kx][ky] -> real [y][x] (src/fftp2D.f90:214-290)
  void c2r(const cd *spec, double *real) {
#pragma omp parallel
    {
      std::vector<cd> w(n), z(n), z2(n);
#pragma omp for schedule(static)
      for (int i = 0; i < nh; ++i) {
        cd *col = mix.data() + (size_t)i * n;
        std::memcpy(col, spec + (size_t)i * n, sizeof(cd) * n);
        fft.run(col, w.data(), +1);
      }
#pragma omp for schedule(static)
      for (int y = 0; y < n; y += 4) {  // four rows: two packed transforms
        for (int pr = 0; pr < 2; ++pr) {
          cd *zz = pr ? z2.data() : z.data();
          const int ya = y + 2 * pr, yb = ya + 1;
          for (int i = 0; i < nh; ++i) {
            cd A = mix[(size_t)i * n + ya], B = mix[(size_t)i * n + yb];
            if (i == 0 || i == n / 2) { A = cd(A.real(), 0.0); B = cd(B.real(), 0.0); }
            zz[i] = A + cd(-B.imag(), B.real());  // A + i B
            if (i > 0 && i < n / 2) {
              const cd Ac = std::conj(A), Bc = std::conj(B);
              zz[n - i] = Ac + cd(-Bc.imag(), Bc.real());
            }
          }
          fft.run(zz, w.data(), +1);
          double *ra = real + (size_t)ya * n, *rb = real + (size_t)yb * n;
          for (int x = 0; x < n; ++x) { ra[x] = zz[x].real(); rb[x] = zz[x].imag(); }
        }
      }
    }
  }

  // r2c: real [y][x] -> spectral [kx][ky] (src/fftp2D.f90:133-211)
  void r2c(const double *real, cd *spec) {
#pragma omp parallel
    {
      std::vector<cd> w(n), z(n);
#pragma omp for schedule(static)
      for (int y = 0; y < n; y += 2) {
        const double *ra = real + (size_t)y * n, *rb = ra + n;
        for (int x = 0; x < n; ++x) z[x] = cd(ra[x], rb[x]);
        fft.run(z.data(), w.data(), -1);
        for (int i = 0; i < nh; ++i) {
          const cd Z = z[i], Zc = std::conj(z[(n - i) % n]);
          const cd A = 0.5 * (Z + Zc), D = 0.5 * (Z - Zc);
          mix[(size_t)i * n + y] = A;
          mix[(size_t)i * n + y + 1] = cd(D.imag(), -D.real());  // D / i
        }
      }
#pragma omp for schedule(static)
      for (int i = 0; i < nh; ++i) {
        cd *col = spec + (size_t)i * n;
        std::memcpy(col, mix.data() + (size_t)i * n, sizeof(cd) * n);
        fft.run(col, w.data(), -1);
      }
    }
  }

  // derivk2 (src/pseudospec2D_Rot2D3C.f90:6-68): dir 1 i kx a, dir 2 i ky a,
  // zero where ka2 > kmax
  void derivk2(const cd *a, cd *b, int dir) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nh; ++i)
      for (int j = 0; j < n; ++j) {
        const size_t e = (size_t)i * n + j;
        const double k1 = kx(i), k2 = ka(j), ka2 = k1 * k1 + k2 * k2;
        if (ka2 > kmax) { b[e] = 0.0; continue; }
        const double k = dir == 1 ? k1 : k2;
        b[e] = cd(-k * a[e].imag(), k * a[e].real());
      }
  }

  // poisson (src/pseudospec2D_Rot2D3C.f90:169-226)
  void poisson(const cd *a, const cd *b, cd *c, cd *c1, cd *c2, double *r1, double *r2,
               double *r3) {
    const size_t nn = (size_t)n * n;
    const double n4 = (double)n * n * (double)n * n;
    derivk2(a, c1, 1);
    derivk2(b, c2, 2);
    c2r(c1, r1);
    c2r(c2, r2);
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < nn; ++e) r3[e] = r1[e] * r2[e];
    derivk2(a, c1, 2);
    derivk2(b, c2, 1);
    c2r(c1, r1);
    c2r(c2, r2);
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < nn; ++e) r3[e] = (r3[e] - r1[e] * r2[e]) / n4;
    r2c(r3, c);
  }
};

struct Draw {
  double kx, ky, phase, f0, kupk;
  int enabled;
};

// The dense array of one rand_force draw (src/pseudospec2D_Rot2D3C.f90:
// 1097-1122): amplitude sqrt(2 f0)/(sqrt(dt) kup^kin) n^2 e^{i phase} at the
// half-plane index whose ka equals +-ky and the full index whose ka equals
// kx; only the kx = 0 column gets a conjugate partner.
void force_fill(const Step &st, const Draw &d, double dt, cd *f) {
  std::memset((void *)f, 0, sizeof(cd) * (size_t)st.nh * st.n);
  if (!d.enabled) return;
  const int n = st.n;
  const double amp = std::sqrt(2.0 * d.f0) / (std::sqrt(dt) * d.kupk) * (double)n * (double)n;
  const cd val(amp * std::cos(d.phase), amp * std::sin(d.phase));
  for (int i = 1; i <= n / 2; ++i)
    if (st.kx(i) == d.ky || st.kx(i) == -d.ky)
      for (int j = 0; j < n; ++j)
        if (st.ka(j) == d.kx) f[(size_t)i * n + j] = val;
  if (st.kx(0) == d.ky)
    for (int j = 1; j < n / 2; ++j)
      if (st.ka(j) == d.kx) {
        f[j] = val;
        f[n - j] = std::conj(val);
      }
}

}  // namespace

extern "C" {

int r2d3c_cpu_threads(int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  return omp_get_max_threads();
}

int r2d3c_cpu_c2r(int n, const double *spec, double *real) {
  Step st(n);
  st.c2r((const cd *)spec, real);
  return 0;
}

int r2d3c_cpu_r2c(int n, const double *real, double *spec) {
  Step st(n);
  st.r2c(real, (cd *)spec);
  return 0;
}

// nsteps fixed-dt steps in place.  prm: nu, hnu, nuv, hnuv, omega, nn, mm,
// nnv, mmv (parameter.inp lines 12-20); draws: 2 per step (psi, vz), each
// kx, ky, phase, f0, kupk, enabled (NULL: no forcing).
int r2d3c_cpu_run(int n, const double *prm, double *psi, double *vz, int nsteps, double dt,
                  const double *draws) {
  if (n < 4 || (n & (n - 1))) return -1;
  Step st(n);
  const size_t S = (size_t)st.nh * n, R = (size_t)n * n;
  std::vector<cd> C1(S), C2(S), C3(S), C4(S), C5(S), C6(S), t1(S), t2(S), fp(S), fz(S);
  std::vector<double> r1(R), r2(R), r3(R);
  const double nu = prm[0], hnu = prm[1], nuv = prm[2], hnuv = prm[3], omega = prm[4];
  const int nn = (int)prm[5], mm = (int)prm[6], nnv = (int)prm[7], mmv = (int)prm[8];
  cd *ps = (cd *)psi, *v = (cd *)vz;
  for (int step = 0; step < nsteps; ++step) {
    Draw dp{}, dz{};
    if (draws) {
      const double *q = draws + (size_t)step * 12;
      dp = Draw{q[0], q[1], q[2], q[3], q[4], (int)q[5]};
      dz = Draw{q[6], q[7], q[8], q[9], q[10], (int)q[11]};
    }
    force_fill(st, dp, dt, fp.data());
    force_fill(st, dz, dt, fz.data());
    std::memcpy((void *)C1.data(), ps, sizeof(cd) * S);  // :545-550
    std::memcpy((void *)C2.data(), v, sizeof(cd) * S);
    for (int o = 4; o >= 1; --o) {
      // C4 = laplak2(C1) (no mask, src/pseudospec2D_Rot2D3C.f90:71-95)
#pragma omp parallel for schedule(static)
      for (int i = 0; i < st.nh; ++i)
        for (int j = 0; j < n; ++j) {
          const double k1 = st.kx(i), k2 = st.ka(j);
          C4[(size_t)i * n + j] = -(k1 * k1 + k2 * k2) * C1[(size_t)i * n + j];
        }
      st.poisson(C1.data(), C4.data(), C5.data(), t1.data(), t2.data(), r1.data(), r2.data(),
                 r3.data());
      st.poisson(C1.data(), C2.data(), C6.data(), t1.data(), t2.data(), r1.data(), r2.data(),
                 r3.data());
      st.derivk2(C1.data(), C3.data(), 2);
      st.derivk2(C2.data(), C4.data(), 2);
      // the update of src/Rot2D3C.para.f90:578-599, in its evaluation order
#pragma omp parallel for schedule(static)
      for (int i = 0; i < st.nh; ++i)
        for (int j = 0; j < n; ++j) {
          const size_t e = (size_t)i * n + j;
          const double k1 = st.kx(i), k2 = st.ka(j), ka2 = k1 * k1 + k2 * k2;
          if (ka2 <= st.kmax && ka2 > st.tiny) {
            const double tmp1 = dt / (double)o, tmp2 = 1.0 / ka2;
            double tmp3 = 1.0 + (nu * std::pow(ka2, nn) + hnu * std::pow(tmp2, mm)) * tmp1;
            C1[e] = (ps[e] + ((-C5[e] + 2.0 * omega * C4[e]) * tmp2 + fp[e]) * tmp1) / tmp3;
            tmp3 = 1.0 + (nuv * std::pow(ka2, nnv) + hnuv * std::pow(tmp2, mmv)) * tmp1;
            C2[e] = (v[e] + (C6[e] + 2.0 * omega * C3[e] + fz[e]) * tmp1) / tmp3;
          } else {
            C1[e] = 0.0;
            C2[e] = 0.0;
          }
        }
    }
    // ps = derivk2(C1, 0), vz = derivk2(C2, 0): masked copies (:610-611)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < st.nh; ++i)
      for (int j = 0; j < n; ++j) {
        const size_t e = (size_t)i * n + j;
        const double k1 = st.kx(i), k2 = st.ka(j);
        const bool keep = k1 * k1 + k2 * k2 <= st.kmax;
        ps[e] = keep ? C1[e] : 0.0;
        v[e] = keep ? C2[e] : 0.0;
      }
  }
  return 0;
}

}  // extern "C"
