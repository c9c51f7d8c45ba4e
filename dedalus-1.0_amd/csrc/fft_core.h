// Mixed-radix Stockham FFT building blocks for gfx950 (CDNA4).
//
// Everything here is header-only and compiles both as HIP device code and as
// plain host C++ (g++), so the exact butterfly and Stockham indexing used by
// the kernels can be exercised on the CPU against numpy.fft.
//
// Conventions
//   cpx<T>          complex value {x = re, y = im}
//   DIR = -1        forward DFT   X[k] = sum x[n] exp(-2 pi i k n / N)
//   DIR = +1        inverse DFT   x[n] = sum X[k] exp(+2 pi i k n / N)  (unnormalised)
//   twiddle tables  tw[j] = exp(-2 pi i j / N), j in [0, N), FP64 or FP32;
//                   the inverse direction uses the conjugate.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DDX_HD __host__ __device__ inline
#else
#define DDX_HD inline
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define DDX_BARRIER() __syncthreads()
#else
#define DDX_BARRIER() ((void)0)
#endif

namespace ddx {

template <typename T>
struct cpx {
  T x, y;
};

template <typename T> DDX_HD cpx<T> mk(T a, T b) { cpx<T> r; r.x = a; r.y = b; return r; }
template <typename T> DDX_HD cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> DDX_HD cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> DDX_HD cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> DDX_HD cpx<T> scale(cpx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T> DDX_HD cpx<T> conj(cpx<T> a) { return mk(a.x, -a.y); }
// multiply by D*i, D = +1 or -1
template <int D, typename T> DDX_HD cpx<T> mul_di(cpx<T> a) {
  return D > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}
// multiply by i * s (s real)
template <typename T> DDX_HD cpx<T> mul_is(cpx<T> a, T s) { return mk(-a.y * s, a.x * s); }

// ---------------------------------------------------------------------------
// small DFTs in registers, natural order in and out
// ---------------------------------------------------------------------------

template <int D, typename T> DDX_HD void dft2(cpx<T>* v) {
  cpx<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <int D, typename T> DDX_HD void dft3(cpx<T>* v) {
  const T h = T(0.86602540378443864676);  // sin(2 pi / 3)
  cpx<T> t = v[1] + v[2];
  cpx<T> m = v[0] - scale(t, T(0.5));
  cpx<T> s = mul_di<D>(scale(v[1] - v[2], h));
  v[0] = v[0] + t;
  v[1] = m + s;
  v[2] = m - s;
}

template <int D, typename T> DDX_HD void dft4(cpx<T>* v) {
  cpx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
  cpx<T> t2 = v[1] + v[3], t3 = mul_di<D>(v[1] - v[3]);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}

template <int D, typename T> DDX_HD void dft5(cpx<T>* v) {
  const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
  const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
  cpx<T> a = v[0];
  cpx<T> t1 = v[1] + v[4], t2 = v[2] + v[3];
  cpx<T> t3 = v[1] - v[4], t4 = v[2] - v[3];
  cpx<T> m1 = a + scale(t1, c1) + scale(t2, c2);
  cpx<T> m2 = a + scale(t1, c2) + scale(t2, c1);
  cpx<T> n1 = mul_di<D>(scale(t3, s1) + scale(t4, s2));
  cpx<T> n2 = mul_di<D>(scale(t3, s2) - scale(t4, s1));
  v[0] = a + t1 + t2;
  v[1] = m1 + n1;
  v[4] = m1 - n1;
  v[2] = m2 + n2;
  v[3] = m2 - n2;
}

template <int D, typename T> DDX_HD void dft8(cpx<T>* v) {
  const T r = T(0.70710678118654752440);
  cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<D>(e);
  dft4<D>(o);
  // W8^k = exp(D 2 pi i k / 8)
  o[1] = o[1] * mk(r, T(D) * r);
  o[2] = mul_di<D>(o[2]);
  o[3] = o[3] * mk(-r, T(D) * r);
  for (int k = 0; k < 4; ++k) {
    v[k] = e[k] + o[k];
    v[k + 4] = e[k] - o[k];
  }
}

template <int D, typename T> DDX_HD void dft16(cpx<T>* v) {
  const T c = T(0.92387953251128675613), s = T(0.38268343236508977173);
  const T r = T(0.70710678118654752440);
  // four-step: n = 4 n1 + n2, k = k1 + 4 k2
  cpx<T> a[4][4];
  for (int n2 = 0; n2 < 4; ++n2) {
    cpx<T> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
    dft4<D>(t);
    for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
  }
  // twiddles W16^(n2 k1)
  a[1][1] = a[1][1] * mk(c, T(D) * s);
  a[1][2] = a[1][2] * mk(r, T(D) * r);
  a[1][3] = a[1][3] * mk(s, T(D) * c);
  a[2][1] = a[2][1] * mk(r, T(D) * r);
  a[2][2] = mul_di<D>(a[2][2]);
  a[2][3] = a[2][3] * mk(-r, T(D) * r);
  a[3][1] = a[3][1] * mk(s, T(D) * c);
  a[3][2] = a[3][2] * mk(-r, T(D) * r);
  a[3][3] = a[3][3] * mk(-c, -T(D) * s);
  for (int k1 = 0; k1 < 4; ++k1) {
    cpx<T> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4<D>(t);
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

template <int D, int R, typename T> DDX_HD void dft(cpx<T>* v) {
  if (R == 2) dft2<D>(v);
  else if (R == 3) dft3<D>(v);
  else if (R == 4) dft4<D>(v);
  else if (R == 5) dft5<D>(v);
  else if (R == 8) dft8<D>(v);
  else if (R == 16) dft16<D>(v);
}

// ---------------------------------------------------------------------------
// radix plan
// ---------------------------------------------------------------------------

#define DDX_MAX_PASS 12

struct FftPlan {
  int n;                       // transform length
  int npass;                   // number of Stockham passes
  int radix[DDX_MAX_PASS];     // radices in execution order
};

// Factor n into passes of radix 16/8/4/2 (balanced), then 3s and 5s.
// Returns 0 on success, -1 if n has a prime factor other than 2, 3, 5.
inline int make_fft_plan(int n, FftPlan* p) {
  p->n = n;
  p->npass = 0;
  if (n < 1) return -1;
  int m = n, e2 = 0, e3 = 0, e5 = 0;
  while (m % 2 == 0) { m /= 2; ++e2; }
  while (m % 3 == 0) { m /= 3; ++e3; }
  while (m % 5 == 0) { m /= 5; ++e5; }
  if (m != 1) return -1;
  int np2 = (e2 + 3) / 4;
  if (np2 + e3 + e5 > DDX_MAX_PASS) return -1;
  for (int i = 0; i < np2; ++i) {
    int bits = e2 / np2 + (i < e2 % np2 ? 1 : 0);
    p->radix[p->npass++] = 1 << bits;
  }
  for (int i = 0; i < e5; ++i) p->radix[p->npass++] = 5;
  for (int i = 0; i < e3; ++i) p->radix[p->npass++] = 3;
  return 0;
}

// ---------------------------------------------------------------------------
// one Stockham butterfly: reads R elements i + r*(n/R) of sequence s from
// src, twiddles, DFT_R, writes to j + r*p of dst (j = (i - k) R + k, k = i mod p)
// ---------------------------------------------------------------------------

template <typename T, typename TW, int D, int R>
DDX_HD void stockham_bfly(const cpx<T>* src, cpx<T>* dst, int i, int nb, int p, int es,
                          int twstride, const cpx<TW>* tw) {
  cpx<T> v[R];
  for (int r = 0; r < R; ++r) v[r] = src[(i + r * nb) * es];
  int k = i % p;
  if (k) {
    for (int r = 1; r < R; ++r) {
      cpx<TW> w = tw[r * k * twstride];
      cpx<T> wt = mk(T(w.x), T(D < 0 ? w.y : -w.y));
      v[r] = v[r] * wt;
    }
  }
  dft<D, R>(v);
  int j = (i - k) * R + k;
  for (int r = 0; r < R; ++r) dst[(j + r * p) * es] = v[r];
}

// Batched in-LDS FFT over `nseq` sequences of length plan.n.  Element e of
// sequence s lives at buf[s * ss + e * es].  Ping-pongs between A and B;
// returns the buffer holding the result.  Work is spread over threads
// tid = 0..nthr-1 (host callers pass tid = 0, nthr = 1).  When seq_fast is
// set, consecutive threads take consecutive sequences (used when sequences
// are adjacent in memory, i.e. ss == 1), otherwise consecutive butterflies.
template <typename T, typename TW, int D>
DDX_HD cpx<T>* lds_fft(cpx<T>* A, cpx<T>* B, const FftPlan& plan, int nseq, int es, int ss,
                       bool seq_fast, const cpx<TW>* tw, int tid, int nthr) {
  const int n = plan.n;
  int p = 1;
  for (int pass = 0; pass < plan.npass; ++pass) {
    const int R = plan.radix[pass];
    const int nb = n / R;
    const int twstride = n / (p * R);
    const int total = nseq * nb;
    for (int b = tid; b < total; b += nthr) {
      int s, i;
      if (seq_fast) { s = b % nseq; i = b / nseq; }
      else { s = b / nb; i = b - s * nb; }
      const cpx<T>* src = A + s * ss;
      cpx<T>* dst = B + s * ss;
      switch (R) {
        case 2: stockham_bfly<T, TW, D, 2>(src, dst, i, nb, p, es, twstride, tw); break;
        case 3: stockham_bfly<T, TW, D, 3>(src, dst, i, nb, p, es, twstride, tw); break;
        case 4: stockham_bfly<T, TW, D, 4>(src, dst, i, nb, p, es, twstride, tw); break;
        case 5: stockham_bfly<T, TW, D, 5>(src, dst, i, nb, p, es, twstride, tw); break;
        case 8: stockham_bfly<T, TW, D, 8>(src, dst, i, nb, p, es, twstride, tw); break;
        default: stockham_bfly<T, TW, D, 16>(src, dst, i, nb, p, es, twstride, tw); break;
      }
    }
    DDX_BARRIER();
    cpx<T>* t = A; A = B; B = t;
    p *= R;
  }
  return A;
}

}  // namespace ddx
