// HIP kernels of the sparse 3D FFT for CDNA4 (gfx950), generic-length path.
//
// Every stage kernel stages a batch of 1D lines in LDS, runs a mixed-radix
// Stockham FFT over them (radix 2/3/4/5/7/8/11/13 butterflies in registers,
// any other prime as a direct DFT from a twiddle table) and writes the result
// with the next stage's layout, so the transpose between stages happens at
// the LDS -> HBM boundary. Twiddles come from one table per length holding
// exp(-2 pi i t / n) generated on the GPU in FP64; the backward direction uses
// the conjugate, which is exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <unordered_map>

#include "hip_check.hpp"
#include "kernels.hpp"
#include "profiling.hpp"
#include "fft_common.hpp"

namespace spfft {

extern __shared__ __attribute__((aligned(16))) unsigned char spfft_lds[];

namespace {

using namespace dev;

constexpr int kBlock = 256;
constexpr size_t kLdsDefault = 64 * 1024;
constexpr size_t kLdsMax = 160 * 1024;

// ---- one Stockham pass over `batch` contiguous lines of length n in LDS.
// Ns = product of the radices already applied. Butterfly j reads
// src[j + r n/R], twiddles by W^(k r n/(Ns R)) with k = j mod Ns and writes
// dst[(j / Ns) Ns R + k + r Ns]; after the last pass the order is natural.
template <typename T, int R, bool BWD>
__device__ __forceinline__ void stockham_pass(const Cx<T>* __restrict__ src,
                                              Cx<T>* __restrict__ dst, int n, int Ns, int batch,
                                              const Tw<T>* __restrict__ w) {
  const int m = n / R;
  const int twStep = n / (Ns * R);
  const int total = batch * m;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int b = t / m;
    const int j = t - b * m;
    const Cx<T>* s = src + b * n;
    Cx<T>* d = dst + b * n;
    const int k = j % Ns;
    Cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * m];
    if (k != 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<BWD>(w, k * r * twStep));
    }
    Butterfly<T, R, BWD>::run(v, w, n);
    const int o = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[o + r * Ns] = v[r];
  }
}

// Any radix (primes above 7): one thread per output, direct DFT.
template <typename T, bool BWD>
__device__ __forceinline__ void stockham_pass_any(const Cx<T>* __restrict__ src,
                                                  Cx<T>* __restrict__ dst, int n, int R, int Ns,
                                                  int batch, const Tw<T>* __restrict__ w) {
  const int m = n / R;
  const int twStep = n / (Ns * R);
  const int total = batch * n;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int b = t / n;
    const int rem = t - b * n;
    const int q = rem / m;
    const int j = rem - q * m;
    const Cx<T>* s = src + b * n;
    const int k = j % Ns;
    Cx<T> acc = {T(0), T(0)};
    for (int r = 0; r < R; ++r) {
      int e = k * r * twStep + ((r * q) % R) * m;
      if (e >= n) e -= n;
      acc = cadd(acc, cmul(s[j + r * m], twiddle<BWD>(w, e)));
    }
    dst[b * n + (j / Ns) * Ns * R + k + q * Ns] = acc;
  }
}

// Full FFT of `batch` lines staged in a (scratch b); returns the buffer that
// holds the result. Caller synchronises before the call.
template <typename T, bool BWD>
__device__ Cx<T>* lds_fft(Cx<T>* a, Cx<T>* b, const FftDesc& f, int batch,
                          const Tw<T>* __restrict__ w) {
  int Ns = 1;
  for (int p = 0; p < f.npass; ++p) {
    const int R = f.radix[p];
    switch (R) {
      case 2: stockham_pass<T, 2, BWD>(a, b, f.n, Ns, batch, w); break;
      case 3: stockham_pass<T, 3, BWD>(a, b, f.n, Ns, batch, w); break;
      case 4: stockham_pass<T, 4, BWD>(a, b, f.n, Ns, batch, w); break;
      case 5: stockham_pass<T, 5, BWD>(a, b, f.n, Ns, batch, w); break;
      case 7: stockham_pass<T, 7, BWD>(a, b, f.n, Ns, batch, w); break;
      case 8: stockham_pass<T, 8, BWD>(a, b, f.n, Ns, batch, w); break;
      // 11 and 13: the DFT of the radix from the length-n table, unrolled
      case 11: stockham_pass<T, 11, BWD>(a, b, f.n, Ns, batch, w); break;
      case 13: stockham_pass<T, 13, BWD>(a, b, f.n, Ns, batch, w); break;
      default: stockham_pass_any<T, BWD>(a, b, f.n, R, Ns, batch, w); break;
    }
    __syncthreads();
    Cx<T>* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// Hermitian fill of one line held in LDS, with the reference's rule
// (src/symmetry/symmetry_host.hpp:47-90): for each pair (k, n-k), a non-zero
// lower entry overwrites the upper one with its conjugate, otherwise a
// non-zero upper entry fills the lower one; a non-zero self-paired entry
// (k = n/2) is conjugated. Exact compares, as in the reference.
template <typename T>
__device__ __forceinline__ void hermitian_fill_line(Cx<T>* line, int n) {
  for (int k = 1 + threadIdx.x; 2 * k <= n; k += blockDim.x) {
    const int kk = n - k;
    if (kk == k) {
      if (cnonzero(line[k])) line[k] = cconj(line[k]);
    } else if (cnonzero(line[k])) {
      line[kk] = cconj(line[k]);
    } else if (cnonzero(line[kk])) {
      line[k] = cconj(line[kk]);
    }
  }
}

template <typename T>
__device__ __forceinline__ void zero_lds(Cx<T>* a, int count) {
  for (int t = threadIdx.x; t < count; t += blockDim.x) a[t] = Cx<T>{T(0), T(0)};
}

// Thread-strided loop over [begin, end) with the global loads of kUnroll
// iterations in flight together: load(t) issues iteration t's loads and
// returns their results, use(t, x) consumes them (LDS or global stores).
constexpr int kUnroll = 8;
template <typename Load, typename Use>
__device__ __forceinline__ void batched_loop(int begin, int end, Load load, Use use) {
  for (int t0 = begin + static_cast<int>(threadIdx.x); t0 < end;
       t0 += kUnroll * static_cast<int>(blockDim.x)) {
    decltype(load(t0)) x[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int t = t0 + u * static_cast<int>(blockDim.x);
      if (t < end) x[u] = load(t);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int t = t0 + u * static_cast<int>(blockDim.x);
      if (t < end) use(t, x[u]);
    }
  }
}

// ---------------------------------------------------------------- backward

template <typename T>
__global__ void __launch_bounds__(kBlock) k_z_backward(StageArgs<T> a) {
  const int nz = a.nz;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(spfft_lds);
  Cx<T>* B = A + a.batch * nz;
  const int s0 = blockIdx.x * a.batch;
  const int ns = min(a.batch, a.nst - s0);
  zero_lds(A, ns * nz);
  __syncthreads();
  const int e0 = a.stickBegin[s0], e1 = a.stickBegin[s0 + ns];
  const int base = s0 * nz;
  const Cx<T>* in = reinterpret_cast<const Cx<T>*>(a.in);
  // entries of kUnroll iterations, then their values, all in flight together;
  // a repeated triplet's earlier copies are skipped (the last one wins)
  for (int f0 = e0 + static_cast<int>(threadIdx.x); f0 < e1;
       f0 += kUnroll * static_cast<int>(blockDim.x)) {
    int slot[kUnroll], val[kUnroll];
    bool keep[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int e = f0 + u * static_cast<int>(blockDim.x);
      keep[u] = false;
      if (e < e1) {
        slot[u] = a.entrySlot[e];
        val[u] = a.entryValue[e];
        keep[u] = !(e + 1 < e1 && a.entrySlot[e + 1] == slot[u]);
      }
    }
    Cx<T> x[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (keep[u]) x[u] = in[val[u]];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (keep[u]) A[slot[u] - base] = x[u];
    }
  }
  __syncthreads();
  if (a.r2c && s0 == 0 && a.zeroZeroStick == 0) {
    hermitian_fill_line(A, nz);
    __syncthreads();
  }
  const Cx<T>* r = lds_fft<T, true>(A, B, a.fft, ns, a.tw);
  for (int t = threadIdx.x; t < ns * nz; t += blockDim.x) {
    const int sl = t % ns, z = t / ns;
    a.zbuf[static_cast<size_t>(z) * a.nst + s0 + sl] = r[sl * nz + z];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_y_backward(StageArgs<T> a) {
  const int ny = a.ny;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(spfft_lds);
  Cx<T>* B = A + a.batch * ny;
  const int z = blockIdx.y;
  const int u0 = blockIdx.x * a.batch;
  const int nc = min(a.batch, a.ncol - u0);
  zero_lds(A, nc * ny);
  __syncthreads();
  const Cx<T>* row = a.zbuf + static_cast<size_t>(z) * a.nst;
  struct Gathered {
    int at;
    Cx<T> x;
  };
  batched_loop(
      a.colBegin[u0], a.colBegin[u0 + nc],
      [&](int s) { return Gathered{(a.stickCol[s] - u0) * ny + a.stickY[s], row[s]}; },
      [&](int, const Gathered& g) { A[g.at] = g.x; });
  __syncthreads();
  if (a.r2c && u0 == 0 && a.colX[0] == 0) {
    hermitian_fill_line(A, ny);
    __syncthreads();
  }
  const Cx<T>* r = lds_fft<T, true>(A, B, a.fft, nc, a.tw);
  for (int t = threadIdx.x; t < nc * ny; t += blockDim.x) {
    const int ul = t % nc, y = t / nc;
    a.ybuf[(static_cast<size_t>(z) * ny + y) * a.ncol + u0 + ul] = r[ul * ny + y];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_x_backward(StageArgs<T> a) {
  const int nx = a.nx, ny = a.ny, ncol = a.ncol;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(spfft_lds);
  Cx<T>* B = A + a.batch * nx;
  const int z = blockIdx.y;
  const int y0 = blockIdx.x * a.batch;
  const int nr = min(a.batch, ny - y0);
  zero_lds(A, nr * nx);
  __syncthreads();
  const Cx<T>* src = a.ybuf + (static_cast<size_t>(z) * ny + y0) * ncol;
  struct Gathered {
    int at;
    Cx<T> x;
  };
  batched_loop(
      0, nr * ncol,
      [&](int t) {
        const int yl = t / ncol, u = t - yl * ncol;
        return Gathered{yl * nx + a.colX[u], src[t]};
      },
      [&](int, const Gathered& g) { A[g.at] = g.x; });
  __syncthreads();
  if (a.r2c) {
    // Complete the hermitian half along x; the imaginary parts of the DC and
    // (even length) Nyquist entries do not contribute to a real signal.
    for (int t = threadIdx.x; t < nr * nx; t += blockDim.x) {
      const int yl = t / nx, k = t - yl * nx;
      Cx<T>* line = A + yl * nx;
      if (k >= a.nxf) {
        line[k] = cconj(line[nx - k]);
      } else if (k == 0 || 2 * k == nx) {
        line[k].im = T(0);
      }
    }
    __syncthreads();
  }
  const Cx<T>* r = lds_fft<T, true>(A, B, a.fft, nr, a.tw);
  const size_t outBase = (static_cast<size_t>(z) * ny + y0) * nx;
  if (a.r2c) {
    for (int t = threadIdx.x; t < nr * nx; t += blockDim.x) a.out[outBase + t] = r[t].re;
  } else {
    Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out);
    for (int t = threadIdx.x; t < nr * nx; t += blockDim.x) out[outBase + t] = r[t];
  }
}

// ----------------------------------------------------------------- forward

template <typename T>
__global__ void __launch_bounds__(kBlock) k_x_forward(StageArgs<T> a) {
  const int nx = a.nx, ny = a.ny, ncol = a.ncol;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(spfft_lds);
  Cx<T>* B = A + a.batch * nx;
  const int z = blockIdx.y;
  const int y0 = blockIdx.x * a.batch;
  const int nr = min(a.batch, ny - y0);
  const size_t inBase = (static_cast<size_t>(z) * ny + y0) * nx;
  if (a.r2c) {
    batched_loop(
        0, nr * nx, [&](int t) { return a.in[inBase + t]; },
        [&](int t, T x) { A[t] = Cx<T>{x, T(0)}; });
  } else {
    const Cx<T>* in = reinterpret_cast<const Cx<T>*>(a.in);
    batched_loop(
        0, nr * nx, [&](int t) { return in[inBase + t]; }, [&](int t, Cx<T> x) { A[t] = x; });
  }
  __syncthreads();
  const Cx<T>* r = lds_fft<T, false>(A, B, a.fft, nr, a.tw);
  Cx<T>* dst = a.ybuf + (static_cast<size_t>(z) * ny + y0) * ncol;
  batched_loop(
      0, nr * ncol,
      [&](int t) {
        const int yl = t / ncol, u = t - yl * ncol;
        return yl * nx + a.colX[u];
      },
      [&](int t, int at) { dst[t] = r[at]; });
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_y_forward(StageArgs<T> a) {
  const int ny = a.ny;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(spfft_lds);
  Cx<T>* B = A + a.batch * ny;
  const int z = blockIdx.y;
  const int u0 = blockIdx.x * a.batch;
  const int nc = min(a.batch, a.ncol - u0);
  batched_loop(
      0, nc * ny,
      [&](int t) {
        const int ul = t % nc, y = t / nc;
        return a.ybuf[(static_cast<size_t>(z) * ny + y) * a.ncol + u0 + ul];
      },
      [&](int t, Cx<T> x) { A[(t % nc) * ny + t / nc] = x; });
  __syncthreads();
  const Cx<T>* r = lds_fft<T, false>(A, B, a.fft, nc, a.tw);
  Cx<T>* row = a.zbuf + static_cast<size_t>(z) * a.nst;
  batched_loop(
      a.colBegin[u0], a.colBegin[u0 + nc],
      [&](int s) { return (a.stickCol[s] - u0) * ny + a.stickY[s]; },
      [&](int s, int at) { row[s] = r[at]; });
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_z_forward(StageArgs<T> a) {
  const int nz = a.nz;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(spfft_lds);
  Cx<T>* B = A + a.batch * nz;
  const int s0 = blockIdx.x * a.batch;
  const int ns = min(a.batch, a.nst - s0);
  batched_loop(
      0, ns * nz,
      [&](int t) { return a.zbuf[static_cast<size_t>(t / ns) * a.nst + s0 + t % ns]; },
      [&](int t, Cx<T> x) { A[(t % ns) * nz + t / ns] = x; });
  __syncthreads();
  const Cx<T>* r = lds_fft<T, false>(A, B, a.fft, ns, a.tw);
  const int base = s0 * nz;
  Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out);
  batched_loop(
      a.stickBegin[s0], a.stickBegin[s0 + ns],
      [&](int e) { return make_int2(a.entrySlot[e] - base, a.entryValue[e]); },
      [&](int, int2 sv) {
        Cx<T> c = r[sv.x];
        if (a.useScale) c = cscale(c, a.scale);
        out[sv.y] = c;
      });
}

template <typename T>
__global__ void k_twiddles(Tw<T>* tw, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  // Angle reduced to (-pi, pi] so that the argument of sincospi is exact up
  // to one rounding of t / n.
  const int tt = (2 * t > n) ? t - n : t;
  double s, c;
  sincospi(-2.0 * static_cast<double>(tt) / static_cast<double>(n), &s, &c);
  if constexpr (sizeof(T) == 4) {
    // compensated pair: the FP64 root is hi + lo to about 2^-48
    const float ch = static_cast<float>(c), sh = static_cast<float>(s);
    tw[t] = Tw<T>{Cx<float>{ch, sh}, Cx<float>{static_cast<float>(c - static_cast<double>(ch)),
                                               static_cast<float>(s - static_cast<double>(sh))}};
  } else {
    tw[t] = Tw<T>{static_cast<T>(c), static_cast<T>(s)};
  }
}

// Raise the dynamic LDS limit of a kernel once per process.
template <typename Kernel>
void launch_stage(Kernel kernel, dim3 grid, size_t lds, hipStream_t stream,
                  const void* argsPtr) {
  (void)grid;
  (void)stream;
  (void)argsPtr;
  if (lds <= kLdsDefault) return;
  static std::mutex mutex;
  static std::unordered_map<const void*, bool> raised;
  const void* key = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lock(mutex);
  if (raised.count(key)) return;
  check_hip(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(kLdsMax)));
  raised[key] = true;
}

}  // namespace

FftDesc make_fft_desc(int n) {
  FftDesc d;
  d.n = n;
  int m = n;
  auto push = [&](int r) {
    if (d.npass >= kMaxPasses) throw GPUFFTError();
    d.radix[d.npass++] = r;
    m /= r;
  };
  while (m % 8 == 0) push(8);
  while (m % 4 == 0) push(4);
  while (m % 2 == 0) push(2);
  for (int r : {3, 5, 7}) {
    while (m % r == 0) push(r);
  }
  for (int p = 11; m > 1; p += 2) {
    if (static_cast<long long>(p) * p > m) p = m;
    while (m % p == 0) push(p);
  }
  return d;
}

template <typename T>
int stage_batch(int n, int count, size_t* ldsBytes) {
  const size_t perLine = 2 * static_cast<size_t>(n) * sizeof(Cx<T>);
  if (perLine > kLdsMax) throw GPUFFTError();
  int batch = static_cast<int>(std::max<size_t>(1, kLdsDefault / perLine));
  batch = std::min(batch, 64);
  batch = std::max(1, std::min(batch, count));
  *ldsBytes = perLine * batch;
  return batch;
}

template <typename T>
void generate_twiddles(Tw<T>* tw, int n, hipStream_t stream) {
  const int blocks = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_twiddles<T>, dim3(blocks), dim3(kBlock), 0, stream, tw, n);
  check_launch();
}

#define SPFFT_STAGE_LAUNCHER(NAME, KERNEL, LEN, COUNT, GRID)                      \
  template <typename T>                                                          \
  void NAME(const StageArgs<T>& a, hipStream_t stream) {                         \
    size_t lds = 0;                                                              \
    (void)stage_batch<T>(a.LEN, COUNT, &lds);                                    \
    lds = 2 * static_cast<size_t>(a.LEN) * sizeof(Cx<T>) * a.batch;              \
    launch_stage(KERNEL<T>, GRID, lds, stream, &a);                              \
    prof::KernelSpan span(stream, #KERNEL);                                      \
    hipLaunchKernelGGL(KERNEL<T>, GRID, dim3(kBlock), lds, stream, a);           \
    check_launch();                                                              \
  }

SPFFT_STAGE_LAUNCHER(launch_z_backward, k_z_backward, nz, a.nst,
                     dim3((a.nst + a.batch - 1) / a.batch))
SPFFT_STAGE_LAUNCHER(launch_y_backward, k_y_backward, ny, a.ncol,
                     dim3((a.ncol + a.batch - 1) / a.batch, a.nz))
SPFFT_STAGE_LAUNCHER(launch_x_backward, k_x_backward, nx, a.ny,
                     dim3((a.ny + a.batch - 1) / a.batch, a.nz))
SPFFT_STAGE_LAUNCHER(launch_x_forward, k_x_forward, nx, a.ny,
                     dim3((a.ny + a.batch - 1) / a.batch, a.nz))
SPFFT_STAGE_LAUNCHER(launch_y_forward, k_y_forward, ny, a.ncol,
                     dim3((a.ncol + a.batch - 1) / a.batch, a.nz))
SPFFT_STAGE_LAUNCHER(launch_z_forward, k_z_forward, nz, a.nst,
                     dim3((a.nst + a.batch - 1) / a.batch))

#undef SPFFT_STAGE_LAUNCHER

#define SPFFT_INSTANTIATE(T)                                                      \
  template int stage_batch<T>(int, int, size_t*);                                \
  template void generate_twiddles<T>(Tw<T>*, int, hipStream_t);                  \
  template void launch_z_backward<T>(const StageArgs<T>&, hipStream_t);          \
  template void launch_y_backward<T>(const StageArgs<T>&, hipStream_t);          \
  template void launch_x_backward<T>(const StageArgs<T>&, hipStream_t);          \
  template void launch_x_forward<T>(const StageArgs<T>&, hipStream_t);           \
  template void launch_y_forward<T>(const StageArgs<T>&, hipStream_t);           \
  template void launch_z_forward<T>(const StageArgs<T>&, hipStream_t);

SPFFT_INSTANTIATE(double)
SPFFT_INSTANTIATE(float)

}  // namespace spfft
