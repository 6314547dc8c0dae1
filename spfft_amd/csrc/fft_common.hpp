// Device-side complex arithmetic and in-register DFT butterflies shared by
// the generic and the fixed-length kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace spfft {
namespace dev {

template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ Cx<T> cscale(Cx<T> a, T s) {
  return {a.re * s, a.im * s};
}
template <typename T>
__device__ __forceinline__ Cx<T> cconj(Cx<T> a) {
  return {a.re, -a.im};
}
template <typename T>
__device__ __forceinline__ bool cnonzero(Cx<T> a) {
  return a.re != T(0) || a.im != T(0);
}
// Multiply by -i (forward) or +i (backward).
template <bool BWD, typename T>
__device__ __forceinline__ Cx<T> rot(Cx<T> a) {
  return BWD ? Cx<T>{-a.im, a.re} : Cx<T>{a.im, -a.re};
}
// ---- twiddles (table entries Tw<T>: the root, or a compensated FP32 pair)

template <typename T>
__device__ __forceinline__ TwPair<T> cconj(TwPair<T> a) {
  return {cconj(a.hi), cconj(a.lo)};
}
template <bool BWD, typename W>
__device__ __forceinline__ W twiddle(const W* __restrict__ w, int t) {
  const W v = w[t];
  return BWD ? cconj(v) : v;
}
// a (hi + lo): the small lo product is formed first and enters the hi
// product's fused multiply-adds.
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, TwPair<T> b) {
  const T lr = a.re * b.lo.re - a.im * b.lo.im;
  const T li = a.re * b.lo.im + a.im * b.lo.re;
  return {fma(a.re, b.hi.re, fma(-a.im, b.hi.im, lr)), fma(a.re, b.hi.im, fma(a.im, b.hi.re, li))};
}
// The root part of a table entry (exact for roots whose components are 0 or
// +-1, where lo is zero).
template <typename T>
__device__ __forceinline__ Cx<T> tw_root(Cx<T> w) {
  return w;
}
template <typename T>
__device__ __forceinline__ Cx<T> tw_root(TwPair<T> w) {
  return w.hi;
}
template <typename T>
__device__ __forceinline__ Tw<T> tw_one() {
  if constexpr (sizeof(T) == 4) {
    return Tw<T>{Cx<T>{T(1), T(0)}, Cx<T>{T(0), T(0)}};
  } else {
    return Tw<T>{T(1), T(0)};
  }
}

// x k and x (c + i s) for constants written in double, rounded once to T
// (the butterflies' fixed roots; the twiddle tables carry the compensated
// roots).
template <typename T>
__device__ __forceinline__ T kmul(T x, double k) {
  return x * static_cast<T>(k);
}
template <typename T>
__device__ __forceinline__ Cx<T> ckmul(Cx<T> x, double c, double s) {
  const T cc = static_cast<T>(c), ss = static_cast<T>(s);
  return {x.re * cc - x.im * ss, x.re * ss + x.im * cc};
}
template <typename T>
__device__ __forceinline__ Cx<T> kscale(Cx<T> a, double k) {
  return {kmul(a.re, k), kmul(a.im, k)};
}

// ---- butterflies: in-register DFT of length R with sign -1 (fwd) / +1 (bwd)

// acc + a c, c the real (IM = false) or imaginary (IM = true) part of a table
// entry, negated when NEG; a compensated FP32 entry adds its lo part first.
template <bool IM, bool NEG, typename T>
__device__ __forceinline__ Cx<T> cfma_part(Cx<T> acc, Cx<T> a, Cx<T> t) {
  const T c = NEG ? -(IM ? t.im : t.re) : (IM ? t.im : t.re);
  return {fma(a.re, c, acc.re), fma(a.im, c, acc.im)};
}
template <bool IM, bool NEG, typename T>
__device__ __forceinline__ Cx<T> cfma_part(Cx<T> acc, Cx<T> a, TwPair<T> t) {
  const T hi = NEG ? -(IM ? t.hi.im : t.hi.re) : (IM ? t.hi.im : t.hi.re);
  const T lo = NEG ? -(IM ? t.lo.im : t.lo.re) : (IM ? t.lo.im : t.lo.re);
  return {fma(a.re, hi, fma(a.re, lo, acc.re)), fma(a.im, hi, fma(a.im, lo, acc.im))};
}

// Term r of output q of the odd-radix butterfly below (compile-time indices).
template <typename T, int R, int Q, int I>
struct OddTerms {
  static constexpr int H = (R - 1) / 2;
  static constexpr int r = I + 1;
  static constexpr int m = (r * Q) % R;  // W_R^m
  static constexpr bool UP = m > H;      // W_R^m = conj W_R^(R - m)
  static constexpr int mm = UP ? R - m : m;
  __device__ __forceinline__ static void run(Cx<T>& x, Cx<T>& y, const Cx<T>* a, const Cx<T>* b,
                                             const Tw<T>* t) {
    if constexpr (m == 0) {  // composite R: W_R^0 = 1
      x = cadd(x, a[r]);
    } else {
      x = cfma_part<false, false>(x, a[r], t[mm]);
      y = cfma_part<true, UP>(y, b[r], t[mm]);
    }
    if constexpr (I + 1 < H) OddTerms<T, R, Q, I + 1>::run(x, y, a, b, t);
  }
};
template <typename T, int R, int Q>
struct OddOutputs {
  __device__ __forceinline__ static void run(Cx<T>* v, const Cx<T>* a, const Cx<T>* b,
                                             const Tw<T>* t, Cx<T> v0) {
    Cx<T> x = v0, y = {T(0), T(0)};
    OddTerms<T, R, Q, 0>::run(x, y, a, b, t);
    v[Q] = Cx<T>{x.re - y.im, x.im + y.re};          // x + i y
    v[R - Q] = Cx<T>{x.re + y.im, x.im - y.re};      // x - i y
    if constexpr (Q + 1 <= (R - 1) / 2) OddOutputs<T, R, Q + 1>::run(v, a, b, t, v0);
  }
};

template <typename T, int R, bool BWD>
struct Butterfly {
  // DFT of length R from the length-n table: W_R^m = W_n^(m n / R).
  __device__ __forceinline__ static void run(Cx<T>* v, const Tw<T>* __restrict__ w, int n) {
    const int step = n / R;
    if constexpr (R % 2 == 1) {
      // Odd R: W_R^(R - m) = conj W_R^m, so with a_r = v_r + v_(R-r) and
      // b_r = v_r - v_(R-r) (r = 1 .. H = (R - 1) / 2), output q and R - q
      // are x_q +- i y_q, x_q = v_0 + sum_r a_r Re W^(rq), y_q = sum_r b_r
      // Im W^(rq): H roots and 4 H^2 real products instead of (R - 1)^2
      // complex ones.
      constexpr int H = (R - 1) / 2;
      Cx<T> a[H + 1], b[H + 1];
      Tw<T> t[H + 1];
      const Cx<T> v0 = v[0];
      Cx<T> s = v0;
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        a[r] = cadd(v[r], v[R - r]);
        b[r] = csub(v[r], v[R - r]);
        t[r] = twiddle<BWD>(w, r * step);
        s = cadd(s, a[r]);
      }
      OddOutputs<T, R, 1>::run(v, a, b, t, v0);
      v[0] = s;
    } else {
      // direct DFT
      Cx<T> out[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        Cx<T> acc = v[0];
#pragma unroll
        for (int r = 1; r < R; ++r) {
          acc = cadd(acc, cmul(v[r], twiddle<BWD>(w, ((r * q) % R) * step)));
        }
        out[q] = acc;
      }
#pragma unroll
      for (int q = 0; q < R; ++q) v[q] = out[q];
    }
  }
};

template <typename T, bool BWD>
struct Butterfly<T, 2, BWD> {
  __device__ __forceinline__ static void run(Cx<T>* v, const Tw<T>*, int) {
    const Cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T, bool BWD>
struct Butterfly<T, 4, BWD> {
  __device__ __forceinline__ static void run(Cx<T>* v, const Tw<T>*, int) {
    const Cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const Cx<T> t2 = cadd(v[1], v[3]), t3 = rot<BWD>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <typename T, bool BWD>
struct Butterfly<T, 8, BWD> {
  __device__ __forceinline__ static void run(Cx<T>* v, const Tw<T>* w, int n) {
    Cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    Cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Butterfly<T, 4, BWD>::run(e, w, n);
    Butterfly<T, 4, BWD>::run(o, w, n);
    constexpr double h = 0.70710678118654752440;
    // o[k] *= exp(-/+ 2 pi i k / 8)
    const Cx<T> o1 = BWD ? Cx<T>{kmul(o[1].re - o[1].im, h), kmul(o[1].re + o[1].im, h)}
                         : Cx<T>{kmul(o[1].re + o[1].im, h), kmul(o[1].im - o[1].re, h)};
    const Cx<T> o2 = rot<BWD>(o[2]);
    const Cx<T> o3 = BWD ? Cx<T>{kmul(o[3].re + o[3].im, -h), kmul(o[3].re - o[3].im, h)}
                         : Cx<T>{kmul(o[3].im - o[3].re, h), kmul(o[3].re + o[3].im, -h)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// 16 = 2 x 8: X[k] = E[k] + W16^k O[k], X[k + 8] = E[k] - W16^k O[k] with E, O
// the length-8 DFTs of the even and odd inputs.
template <typename T, bool BWD>
struct Butterfly<T, 16, BWD> {
  __device__ __forceinline__ static void run(Cx<T>* v, const Tw<T>* w, int n) {
    Cx<T> e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      e[k] = v[2 * k];
      o[k] = v[2 * k + 1];
    }
    Butterfly<T, 8, BWD>::run(e, w, n);
    Butterfly<T, 8, BWD>::run(o, w, n);
    // cos and sin of 2 pi k / 16, k = 0..7
    constexpr double c[8] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                             0.38268343236508977173, 0.0, -0.38268343236508977173,
                             -0.70710678118654752440, -0.92387953251128675613};
    constexpr double s[8] = {0.0, 0.38268343236508977173, 0.70710678118654752440,
                             0.92387953251128675613, 1.0, 0.92387953251128675613,
                             0.70710678118654752440, 0.38268343236508977173};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      Cx<T> t = o[k];
      if (k == 4) {
        t = rot<BWD>(t);
      } else if (k != 0) {
        t = ckmul(t, c[k], BWD ? s[k] : -s[k]);
      }
      v[k] = cadd(e[k], t);
      v[k + 8] = csub(e[k], t);
    }
  }
};

// 24 = 8 x 3: A[n2] = the length-8 DFT of x[3 n1 + n2], then
// X[k1 + 8 k2] = sum_n2 W3^(n2 k2) W24^(n2 k1) A[n2][k1].
template <typename T, bool BWD>
struct Butterfly<T, 24, BWD> {
  __device__ __forceinline__ static void run(Cx<T>* v, const Tw<T>* w, int n);
};

template <typename T, bool BWD>
struct Butterfly<T, 3, BWD> {
  __device__ __forceinline__ static void run(Cx<T>* v, const Tw<T>*, int) {
    constexpr double s3 = 0.86602540378443864676;  // sin(2 pi / 3)
    const Cx<T> t1 = cadd(v[1], v[2]);
    const Cx<T> t2 = csub(v[1], v[2]);
    const Cx<T> m = csub(v[0], cscale(t1, T(0.5)));
    const Cx<T> r = rot<BWD>(kscale(t2, s3));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <typename T, bool BWD>
struct Butterfly<T, 5, BWD> {
  __device__ __forceinline__ static void run(Cx<T>* v, const Tw<T>*, int) {
    constexpr double c1 = 0.30901699437494742410;   // cos(2 pi / 5)
    constexpr double c2 = -0.80901699437494742410;  // cos(4 pi / 5)
    constexpr double s1 = 0.95105651629515357212;   // sin(2 pi / 5)
    constexpr double s2 = 0.58778525229247312917;   // sin(4 pi / 5)
    const Cx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const Cx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const Cx<T> b1 = cadd(v[0], cadd(kscale(t1, c1), kscale(t2, c2)));
    const Cx<T> b2 = cadd(v[0], cadd(kscale(t1, c2), kscale(t2, c1)));
    const Cx<T> d1 = rot<BWD>(cadd(kscale(t3, s1), kscale(t4, s2)));
    const Cx<T> d2 = rot<BWD>(csub(kscale(t3, s2), kscale(t4, s1)));
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(b1, d1);
    v[4] = csub(b1, d1);
    v[2] = cadd(b2, d2);
    v[3] = csub(b2, d2);
  }
};

template <typename T, bool BWD>
__device__ __forceinline__ void Butterfly<T, 24, BWD>::run(Cx<T>* v, const Tw<T>* w, int n) {
  Cx<T> a[3][8];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) a[n2][n1] = v[3 * n1 + n2];
    Butterfly<T, 8, BWD>::run(a[n2], w, n);
  }
  // cos and sin of 2 pi m / 24, m = 0..14
  constexpr double c15 = 0.96592582628906828675, s15 = 0.25881904510252076235;
  constexpr double c30 = 0.86602540378443864676, c45 = 0.70710678118654752440;
  constexpr double c[15] = {1.0,  c15,  c30,  c45,  0.5,  s15,  0.0,  -s15,
                            -0.5, -c45, -c30, -c15, -1.0, -c15, -c30};
  constexpr double s[15] = {0.0, s15, 0.5, c45, c30, c15, 1.0, c15,
                            c30, c45, 0.5, s15, 0.0, -s15, -0.5};
#pragma unroll
  for (int n2 = 1; n2 < 3; ++n2) {
#pragma unroll
    for (int k1 = 1; k1 < 8; ++k1) {
      const int m = n2 * k1;
      a[n2][k1] = ckmul(a[n2][k1], c[m], BWD ? s[m] : -s[m]);
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) {
    Cx<T> t[3] = {a[0][k1], a[1][k1], a[2][k1]};
    Butterfly<T, 3, BWD>::run(t, w, n);
    v[k1] = t[0];
    v[k1 + 8] = t[1];
    v[k1 + 16] = t[2];
  }
}

}  // namespace dev
}  // namespace spfft
