// CPU FFT of the host execution path (SPFFT_PU_HOST): unnormalised 1D DFT
// of any length, out[k] = sum_j in[j] exp(sign 2 pi i j k / n).
//
// Mixed-radix Stockham, decimation in frequency, self-sorting: with stride s
// (product of the radices already applied) and n' = n / s, a radix-r stage
// reads x[q + s (p + j m)] (m = n' / r, p < m, q < s, j < r), forms the
// r-point DFT t_k = sum_j x_j w_r^(j k), and writes
// y[q + s (r p + k)] = t_k w_n'^(p k). Radices 4, 2 and 3 have dedicated
// butterflies; any other prime factor is a direct DFT against the length-n
// table. Twiddles exp(-2 pi i t / n) are evaluated in long double after
// reduction of t to (-n/2, n/2] and rounded once to T.
#pragma once

#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.hpp"

namespace spfft {

template <typename T>
class HostFft {
public:
  using C = Cx<T>;

  HostFft() = default;
  explicit HostFft(int n) : n_(n) {
    int m = n;
    while (m % 4 == 0) { radix_.push_back(4); m /= 4; }
    while (m % 2 == 0) { radix_.push_back(2); m /= 2; }
    for (int f = 3; m > 1; f += 2) {
      if (static_cast<long long>(f) * f > m) f = m;
      while (m % f == 0) { radix_.push_back(f); m /= f; }
    }
    w_.resize(n > 0 ? n : 1);
    const long double twoPi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; ++t) {
      const long long tt = 2LL * t > n ? static_cast<long long>(t) - n : t;
      const long double a = -twoPi * static_cast<long double>(tt) / static_cast<long double>(n);
      w_[t] = C{static_cast<T>(std::cos(a)), static_cast<T>(std::sin(a))};
    }
  }

  int size() const { return n_; }

  // Elements of scratch that run() needs.
  size_t work_size() const { return 2 * static_cast<size_t>(n_ > 0 ? n_ : 1); }

  // `in` and `out` may be the same array; `work` holds work_size() elements.
  void run(const C* in, C* out, int sign, C* work) const {
    const int n = n_;
    if (n <= 1) {
      if (n == 1) out[0] = in[0];
      return;
    }
    const C* x = in;
    C* bufs[2] = {work, work + n};
    int s = 1;
    for (size_t i = 0; i < radix_.size(); ++i) {
      C* y = bufs[i % 2];
      stage(radix_[i], s, x, y, sign);
      x = y;
      s *= radix_[i];
    }
    std::memcpy(static_cast<void*>(out), static_cast<const void*>(x), sizeof(C) * n);
  }

private:
  C tw(long long t, int sign) const {
    const C v = w_[t];
    return sign > 0 ? C{v.re, -v.im} : v;
  }
  static C add(C a, C b) { return {a.re + b.re, a.im + b.im}; }
  static C sub(C a, C b) { return {a.re - b.re, a.im - b.im}; }
  static C mul(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
  // multiply by sign * i
  static C rot(C a, int sign) { return sign > 0 ? C{-a.im, a.re} : C{a.im, -a.re}; }

  void stage(int r, int s, const C* x, C* y, int sign) const {
    const int n = n_;
    const int m = n / (s * r);
    C a[64];
    std::vector<C> big;
    C* t = a;
    if (r > 32) {
      big.resize(2 * static_cast<size_t>(r));
      t = big.data();
    }
    C* u = t + r;
    for (int p = 0; p < m; ++p) {
      for (int q = 0; q < s; ++q) {
        for (int j = 0; j < r; ++j) u[j] = x[q + static_cast<size_t>(s) * (p + j * m)];
        butterfly(r, u, t, sign);
        for (int k = 0; k < r; ++k) {
          const long long e = static_cast<long long>(p) * k * s;
          const C v = k == 0 || p == 0 ? t[k] : mul(t[k], tw(e, sign));
          y[q + static_cast<size_t>(s) * (r * p + k)] = v;
        }
      }
    }
  }

  // t[k] = sum_j u[j] w_r^(j k), w_r = exp(sign 2 pi i / r)
  void butterfly(int r, const C* u, C* t, int sign) const {
    switch (r) {
      case 2:
        t[0] = add(u[0], u[1]);
        t[1] = sub(u[0], u[1]);
        return;
      case 4: {
        const C b0 = add(u[0], u[2]), b1 = sub(u[0], u[2]);
        const C b2 = add(u[1], u[3]), b3 = rot(sub(u[1], u[3]), sign);
        t[0] = add(b0, b2);
        t[2] = sub(b0, b2);
        t[1] = add(b1, b3);
        t[3] = sub(b1, b3);
        return;
      }
      case 3: {
        // w_3 = -1/2 + sign i sqrt(3)/2
        const T h = static_cast<T>(0.866025403784438646763723170752936183L);
        const C sm = add(u[1], u[2]), d = sub(u[1], u[2]);
        const C base{u[0].re - T(0.5) * sm.re, u[0].im - T(0.5) * sm.im};
        const C rd = rot(C{h * d.re, h * d.im}, sign);
        t[0] = add(u[0], sm);
        t[1] = add(base, rd);
        t[2] = sub(base, rd);
        return;
      }
      default: {
        // direct DFT: w_r^(j k) = table entry ((j k) mod r) * n / r
        const long long step = n_ / r;
        for (int k = 0; k < r; ++k) {
          C acc = u[0];
          for (int j = 1; j < r; ++j) {
            acc = add(acc, mul(u[j], tw(((static_cast<long long>(j) * k) % r) * step, sign)));
          }
          t[k] = acc;
        }
        return;
      }
    }
  }

  int n_ = 0;
  std::vector<int> radix_;
  std::vector<C> w_;  // exp(-2 pi i t / n)
};

}  // namespace spfft
