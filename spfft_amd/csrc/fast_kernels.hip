// Fixed-length kernels of the C2C sparse 3D FFT for gfx950 (see
// fast_kernels.hpp for the decomposition).
//
// FFT engine (wave level): a line of N points is owned by P = N / E lanes of
// one wavefront, lane q holding elements q + P e (e < E) in registers. Each
// Stockham pass applies radix-R butterflies to lane-local elements (R divides
// E, so every butterfly's inputs sit in one lane); all passes but the last
// exchange results through the line's own LDS region, whose index map
// i -> i + i / 8 (FP64; i / 16 for FP32) keeps the stride-R writes free of
// bank conflicts. The last pass leaves the line in natural order in the
// lanes' registers. Exchanges are wave-local (no workgroup barrier), so the
// waves of a workgroup run their FFTs independently.
#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_map>

#include "fast_kernels.hpp"
#include "fft_common.hpp"
#include "hip_check.hpp"
#include "profiling.hpp"

namespace spfft {

extern __shared__ __attribute__((aligned(16))) unsigned char spfft_lds_fast[];

namespace {

using namespace dev;

constexpr size_t kLdsDefault = 64 * 1024;
constexpr size_t kLdsMax = 160 * 1024;
constexpr int kWave = 64;

// ---- line shapes and radix schedules (compile time)

// Elements per lane: 8 for powers of two, 24 for lengths with a factor 3.
constexpr int elems_per_lane(int n) { return n % 3 == 0 ? 24 : 8; }

template <int N>
struct Line {
  static constexpr int E = elems_per_lane(N);
  static constexpr int P = N / E;      // lanes per line
  static constexpr int PER_WAVE = kWave / P;
  static_assert(N % E == 0 && kWave % P == 0, "line shape");
};

constexpr int wave_radix(int m, int e) {
  return (m % 8 == 0 && e % 8 == 0)   ? 8
         : (m % 4 == 0 && e % 4 == 0) ? 4
         : (m % 2 == 0 && e % 2 == 0) ? 2
         : (m % 3 == 0 && e % 3 == 0) ? 3
         : (m % 5 == 0 && e % 5 == 0) ? 5
                                      : -1;
}

// Padded index inside an exchange region: one spare element every 128 bytes.
template <typename T>
constexpr int pad_shift() {
  return sizeof(T) == 8 ? 3 : 4;
}
template <typename T>
__device__ __forceinline__ int phys(int i) {
  return i + (i >> pad_shift<T>());
}
template <typename T>
constexpr int padded_len(int n) {
  return n + ((n - 1) >> pad_shift<T>());
}
constexpr int make_odd(int n) { return n | 1; }
constexpr int round_up(int n, int m) { return (n + m - 1) / m * m; }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Stockham pass (and, recursively, the remaining ones) of the wave-level
// FFT. `line` is this line's exchange region: element i at line[phys(i) * es]
// (es = 1) or line[i * es] (column regions, es > 1). The twiddle table holds
// W_{N * TWS}. NS = product of the radices already applied.
template <typename T, int N, bool BWD, int TWS, int NS = 1>
__device__ __forceinline__ void wave_fft(Cx<T>* v, Cx<T>* line, int es, int q,
                                         const Cx<T>* __restrict__ tw) {
  constexpr int E = Line<N>::E, PL = Line<N>::P;
  constexpr int R = wave_radix(N / NS, E);
  static_assert(R > 0, "no radix divides the lane's element count");
  constexpr int EB = E / R;
  constexpr bool LAST = NS * R == N;
  constexpr int STEP = N / (NS * R) * TWS;
#pragma unroll
  for (int b = 0; b < EB; ++b) {
    const int j = q + PL * b;
    Cx<T> u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[b + r * EB];
    if constexpr (NS > 1) {
      const int k = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) u[r] = cmul(u[r], twiddle<BWD>(tw, k * r * STEP));
    }
    Butterfly<T, R, BWD>::run(u, tw, N);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * EB] = u[r];
    } else {
      const int o = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = o + r * NS;
        line[(es == 1 ? phys<T>(i) : i * es)] = u[r];
      }
    }
  }
  if constexpr (!LAST) {
    wave_sync();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = q + PL * e;
      v[e] = line[(es == 1 ? phys<T>(i) : i * es)];
    }
    wave_sync();
    wave_fft<T, N, BWD, TWS, NS * R>(v, line, es, q, tw);
  }
}

// Workgroup id -> (plane, class). Workgroups are dealt to the 8 XCDs round
// robin, so the classes of a plane get ids with the same id mod 8 and share
// that XCD's L2 for the plane's input.
template <int R>
__device__ __forceinline__ void plane_class(int b, int& z, int& c) {
  z = (b / (8 * R)) * 8 + (b % 8);
  c = (b / 8) % R;
}

// ------------------------------------------------------------- z kernels

template <typename T, int NZ>
struct ZShape {
  static constexpr int NT = 256;
  static constexpr int S = (NT / kWave) * Line<NZ>::PER_WAVE;  // sticks per workgroup
  static constexpr int LZ = padded_len<T>(NZ);
  static size_t lds() { return static_cast<size_t>(S) * LZ * sizeof(Cx<T>); }
};

template <typename T, int NZ>
__global__ void __launch_bounds__((ZShape<T, NZ>::NT)) k_z_backward_fast(FastArgs<T> a) {
  using Z = ZShape<T, NZ>;
  using L = Line<NZ>;
  constexpr int S = Z::S, LZ = Z::LZ, NT = Z::NT;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  const int s0 = blockIdx.x * S;
  const int ns = min(S, a.nst - s0);
  for (int t = threadIdx.x; t < S * LZ; t += NT) tile[t] = Cx<T>{T(0), T(0)};
  __syncthreads();
  const int e0 = a.stickBegin[s0], e1 = a.stickBegin[s0 + ns];
  for (int e = e0 + threadIdx.x; e < e1; e += NT) {
    const int slot = a.entrySlot[e];
    if (e + 1 < e1 && a.entrySlot[e + 1] == slot) continue;  // repeated triplet: last one wins
    const int rel = slot - s0 * NZ;
    const int sl = rel / NZ;
    const size_t v = static_cast<size_t>(a.entryValue[e]);
    tile[sl * LZ + phys<T>(rel - sl * NZ)] = Cx<T>{a.in[2 * v], a.in[2 * v + 1]};
  }
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int sl = (threadIdx.x / kWave) * L::PER_WAVE + lane / L::P;
  const int q = lane % L::P;
  Cx<T>* line = tile + sl * LZ;
  Cx<T> v[L::E];
#pragma unroll
  for (int e = 0; e < L::E; ++e) v[e] = line[phys<T>(q + L::P * e)];
  wave_sync();
  wave_fft<T, NZ, true, 1>(v, line, 1, q, a.twZ);
#pragma unroll
  for (int e = 0; e < L::E; ++e) line[phys<T>(q + L::P * e)] = v[e];
  __syncthreads();
  for (int t = threadIdx.x; t < S * NZ; t += NT) {
    const int s = t % S, zz = t / S;
    if (s < ns) a.zbuf[static_cast<size_t>(zz) * a.nst + s0 + s] = tile[s * LZ + phys<T>(zz)];
  }
}

template <typename T, int NZ>
__global__ void __launch_bounds__((ZShape<T, NZ>::NT)) k_z_forward_fast(FastArgs<T> a) {
  using Z = ZShape<T, NZ>;
  using L = Line<NZ>;
  constexpr int S = Z::S, LZ = Z::LZ, NT = Z::NT;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  const int s0 = blockIdx.x * S;
  const int ns = min(S, a.nst - s0);
  for (int t = threadIdx.x; t < S * NZ; t += NT) {
    const int s = t % S, zz = t / S;
    tile[s * LZ + phys<T>(zz)] = s < ns ? a.zbuf[static_cast<size_t>(zz) * a.nst + s0 + s]
                                        : Cx<T>{T(0), T(0)};
  }
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int sl = (threadIdx.x / kWave) * L::PER_WAVE + lane / L::P;
  const int q = lane % L::P;
  Cx<T>* line = tile + sl * LZ;
  Cx<T> v[L::E];
#pragma unroll
  for (int e = 0; e < L::E; ++e) v[e] = line[phys<T>(q + L::P * e)];
  wave_sync();
  wave_fft<T, NZ, false, 1>(v, line, 1, q, a.twZ);
#pragma unroll
  for (int e = 0; e < L::E; ++e) line[phys<T>(q + L::P * e)] = v[e];
  __syncthreads();
  const int e0 = a.stickBegin[s0], e1 = a.stickBegin[s0 + ns];
  for (int e = e0 + threadIdx.x; e < e1; e += NT) {
    const int rel = a.entrySlot[e] - s0 * NZ;
    const int s = rel / NZ;
    Cx<T> c = tile[s * LZ + phys<T>(rel - s * NZ)];
    if (a.useScale) c = cscale(c, a.scale);
    const size_t v = static_cast<size_t>(a.entryValue[e]);
    a.out[2 * v] = c.re;
    a.out[2 * v + 1] = c.im;
  }
}

// ------------------------------------------------------------ xy kernels

template <typename T, int N>
struct XYShape {
  static constexpr int R = fast_classes(N);     // residue classes per plane
  static constexpr int M = N / R;               // rows (bwd) / columns (fwd) per class
  static constexpr int NT = N >= 128 && N < 192 ? 512 : 256;
  static constexpr int NW = NT / kWave;
  // backward: A = [columns][LA] (y transforms), reused as B = [rows][LB] (x transforms)
  static constexpr int LA = make_odd(padded_len<T>(M));
  static constexpr int LB = padded_len<T>(N);
  static constexpr int ROWS_ALLOC = round_up(M, Line<N>::PER_WAVE);
  static_assert(NW * Line<N>::PER_WAVE >= M, "all rows of a class in one round");
  // forward: Tb = [rows y][LT], columns transformed in place with stride LT;
  // the rows have room for the spare columns of a partial last wave
  static constexpr int LT = make_odd(padded_len<T>(round_up(M, Line<N>::PER_WAVE)));
  static size_t lds_backward(int ncol) {
    const size_t a = static_cast<size_t>(round_up(ncol, Line<M>::PER_WAVE)) * LA;
    const size_t b = static_cast<size_t>(ROWS_ALLOC) * LB;
    return (a > b ? a : b) * sizeof(Cx<T>);
  }
  static size_t lds_forward() {
    return static_cast<size_t>(round_up(N, Line<M>::PER_WAVE)) * LT * sizeof(Cx<T>);
  }
};

template <typename T, int N>
__global__ void __launch_bounds__((XYShape<T, N>::NT)) k_xy_backward_fast(FastArgs<T> a) {
  using S = XYShape<T, N>;
  using LM = Line<S::M>;
  using LX = Line<N>;
  constexpr int R = S::R, M = S::M, LA = S::LA, LB = S::LB, NW = S::NW;
  const int ncol = a.ncol;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  int z, r;
  plane_class<R>(blockIdx.x, z, r);
  if (z >= a.nz) return;
  const Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.nst;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;

  // y: for each column u, decimation in frequency to the rows y = r + R m
  // (t[j] = W^(r j) sum_l X[j + M l] W_R^(r l)), then the M-point FFT.
  {
    const int q = lane % LM::P;
    for (int u0 = wave * LM::PER_WAVE; u0 < ncol; u0 += NW * LM::PER_WAVE) {
      const int u = u0 + lane / LM::P;
      Cx<T> v[LM::E];
      if (u < ncol) {
        const int16_t* rel = a.stickRel + static_cast<size_t>(u) * N;
        const Cx<T>* col = zrow + a.colBegin[u];
        int s[LM::E][R];
#pragma unroll
        for (int e = 0; e < LM::E; ++e) {
#pragma unroll
          for (int l = 0; l < R; ++l) s[e][l] = rel[q + LM::P * e + M * l];
        }
#pragma unroll
        for (int e = 0; e < LM::E; ++e) {
          const int j = q + LM::P * e;
          Cx<T> acc = {T(0), T(0)};
#pragma unroll
          for (int l = 0; l < R; ++l) {
            if (s[e][l] >= 0) {
              Cx<T> x = col[s[e][l]];
              const int ex = ((r * l) % R) * M;
              if (ex) x = cmul(x, twiddle<true>(a.twY, ex));
              acc = cadd(acc, x);
            }
          }
          v[e] = r ? cmul(acc, twiddle<true>(a.twY, r * j)) : acc;
        }
      } else {
#pragma unroll
        for (int e = 0; e < LM::E; ++e) v[e] = Cx<T>{T(0), T(0)};
      }
      Cx<T>* line = A + u * LA;
      wave_fft<T, M, true, R>(v, line, 1, q, a.twY);
#pragma unroll
      for (int e = 0; e < LM::E; ++e) line[phys<T>(q + LM::P * e)] = v[e];
    }
  }
  __syncthreads();

  // x: row m (y = r + R m) holds column u's value at x = colX[u].
  const int m = wave * LX::PER_WAVE + lane / LX::P;
  const int q = lane % LX::P;
  Cx<T> v[LX::E];
#pragma unroll
  for (int e = 0; e < LX::E; ++e) {
    const int u = a.xCol[q + LX::P * e];
    v[e] = (m < M && u >= 0) ? A[u * LA + phys<T>(m)] : Cx<T>{T(0), T(0)};
  }
  __syncthreads();  // A is dead from here on; its space holds the row lines
  if (wave * LX::PER_WAVE < M) {
    Cx<T>* line = A + m * LB;
    wave_fft<T, N, true, 1>(v, line, 1, q, a.twX);
    if (m < M) {
      Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out) +
                   (static_cast<size_t>(z) * N + r + R * m) * N;
#pragma unroll
      for (int e = 0; e < LX::E; ++e) out[q + LX::P * e] = v[e];
    }
  }
}

template <typename T, int N>
__global__ void __launch_bounds__((XYShape<T, N>::NT)) k_xy_forward_fast(FastArgs<T> a) {
  using S = XYShape<T, N>;
  using LM = Line<S::M>;
  using LY = Line<N>;
  constexpr int R = S::R, MX = S::M, LT = S::LT, NW = S::NW;
  Cx<T>* Tb = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  int z, c;
  plane_class<R>(blockIdx.x, z, c);
  if (z >= a.nz) return;
  const Cx<T>* in = reinterpret_cast<const Cx<T>*>(a.in) + static_cast<size_t>(z) * N * N;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;

  // x: every row, decimation in frequency to the columns x = c + R m, then
  // the MX-point FFT; row y's result lands in Tb[y][m].
  {
    const int q = lane % LM::P;
    for (int y0 = wave * LM::PER_WAVE; y0 < N; y0 += NW * LM::PER_WAVE) {
      const int y = y0 + lane / LM::P;
      Cx<T> v[LM::E];
      if (y < N) {
        const Cx<T>* row = in + static_cast<size_t>(y) * N;
#pragma unroll
        for (int e = 0; e < LM::E; ++e) {
          const int j = q + LM::P * e;
          Cx<T> x[R];
#pragma unroll
          for (int l = 0; l < R; ++l) x[l] = row[j + MX * l];
          Cx<T> acc = x[0];
#pragma unroll
          for (int l = 1; l < R; ++l) {
            const int ex = ((c * l) % R) * MX;
            acc = cadd(acc, ex ? cmul(x[l], twiddle<false>(a.twX, ex)) : x[l]);
          }
          v[e] = c ? cmul(acc, twiddle<false>(a.twX, c * j)) : acc;
        }
      } else {
#pragma unroll
        for (int e = 0; e < LM::E; ++e) v[e] = Cx<T>{T(0), T(0)};
      }
      Cx<T>* line = Tb + y * LT;
      wave_fft<T, MX, false, R>(v, line, 1, q, a.twX);
#pragma unroll
      for (int e = 0; e < LM::E; ++e) line[phys<T>(q + LM::P * e)] = v[e];
    }
  }
  __syncthreads();

  // y: column m of the class, in place with element stride LT.
  {
    const int q = lane % LY::P;
    for (int m0 = wave * LY::PER_WAVE; m0 < MX; m0 += NW * LY::PER_WAVE) {
      // lanes past the last column use a spare column of the padded rows
      const int m = m0 + lane / LY::P;
      Cx<T>* col = Tb + phys<T>(m);
      Cx<T> v[LY::E];
#pragma unroll
      for (int e = 0; e < LY::E; ++e) v[e] = col[(q + LY::P * e) * LT];
      wave_sync();
      wave_fft<T, N, false, 1>(v, col, LT, q, a.twY);
      if (m < MX) {
#pragma unroll
        for (int e = 0; e < LY::E; ++e) col[(q + LY::P * e) * LT] = v[e];
      }
    }
  }
  __syncthreads();
  const int i0 = a.resBegin[c], i1 = a.resBegin[c + 1];
  Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.nst;
  for (int i = i0 + threadIdx.x; i < i1; i += S::NT) {
    const int s = a.resSticks[i];
    const int m = a.colX[a.stickCol[s]] / R;
    zrow[s] = Tb[a.stickY[s] * LT + phys<T>(m)];
  }
}

// Raise the dynamic LDS limit of a kernel once per process (the attribute
// call is far more expensive than a launch).
template <typename Kernel>
void prepare_lds(Kernel kernel, size_t lds) {
  if (lds > kLdsMax) throw GPUFFTError();
  if (lds <= kLdsDefault) return;
  static std::mutex mutex;
  static std::unordered_map<const void*, size_t> raised;
  const void* key = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lock(mutex);
  auto it = raised.find(key);
  if (it != raised.end() && it->second >= kLdsMax) return;
  check_hip(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(kLdsMax)));
  raised[key] = kLdsMax;
}

template <typename T, int NZ>
void z_backward_n(const FastArgs<T>& a, hipStream_t stream) {
  using Z = ZShape<T, NZ>;
  auto kernel = k_z_backward_fast<T, NZ>;
  prepare_lds(kernel, Z::lds());
  prof::KernelSpan span(stream, "k_z_backward_fast");
  hipLaunchKernelGGL(kernel, dim3((a.nst + Z::S - 1) / Z::S), dim3(Z::NT), Z::lds(), stream, a);
  check_launch();
}

template <typename T, int NZ>
void z_forward_n(const FastArgs<T>& a, hipStream_t stream) {
  using Z = ZShape<T, NZ>;
  auto kernel = k_z_forward_fast<T, NZ>;
  prepare_lds(kernel, Z::lds());
  prof::KernelSpan span(stream, "k_z_forward_fast");
  hipLaunchKernelGGL(kernel, dim3((a.nst + Z::S - 1) / Z::S), dim3(Z::NT), Z::lds(), stream, a);
  check_launch();
}

template <int N>
int plane_grid(int nz) {
  return (nz + 7) / 8 * 8 * fast_classes(N);
}

template <typename T, int N>
void xy_backward_n(const FastArgs<T>& a, hipStream_t stream) {
  using S = XYShape<T, N>;
  const size_t lds = S::lds_backward(a.ncol);
  auto kernel = k_xy_backward_fast<T, N>;
  prepare_lds(kernel, lds);
  prof::KernelSpan span(stream, "k_xy_backward_fast");
  hipLaunchKernelGGL(kernel, dim3(plane_grid<N>(a.nz)), dim3(S::NT), lds, stream, a);
  check_launch();
}

template <typename T, int N>
void xy_forward_n(const FastArgs<T>& a, hipStream_t stream) {
  using S = XYShape<T, N>;
  const size_t lds = S::lds_forward();
  auto kernel = k_xy_forward_fast<T, N>;
  prepare_lds(kernel, lds);
  prof::KernelSpan span(stream, "k_xy_forward_fast");
  hipLaunchKernelGGL(kernel, dim3(plane_grid<N>(a.nz)), dim3(S::NT), lds, stream, a);
  check_launch();
}

template <typename T, int N>
bool fits(int ncol) {
  return XYShape<T, N>::lds_backward(ncol) <= kLdsMax && XYShape<T, N>::lds_forward() <= kLdsMax;
}

bool fast_length(int n) { return n == 32 || n == 64 || n == 96 || n == 128 || n == 192; }

#define SPFFT_FAST_SWITCH(N, CALL) \
  switch (N) {                     \
    case 32: CALL(32); break;      \
    case 64: CALL(64); break;      \
    case 96: CALL(96); break;      \
    case 128: CALL(128); break;    \
    case 192: CALL(192); break;    \
    default: throw GPUFFTError();  \
  }

}  // namespace

template <typename T>
bool fast_path_supported(int nx, int ny, int nz, int ncol) {
  if (nx != ny || !fast_length(nx) || !fast_length(nz) || ncol < 1) return false;
  bool ok = false;
#define SPFFT_FITS(N) ok = fits<T, N>(ncol)
  SPFFT_FAST_SWITCH(nx, SPFFT_FITS)
#undef SPFFT_FITS
  return ok;
}

template <typename T>
void launch_fast_z_backward(int nz, const FastArgs<T>& a, hipStream_t stream) {
#define SPFFT_CALL(N) z_backward_n<T, N>(a, stream)
  SPFFT_FAST_SWITCH(nz, SPFFT_CALL)
#undef SPFFT_CALL
}

template <typename T>
void launch_fast_z_forward(int nz, const FastArgs<T>& a, hipStream_t stream) {
#define SPFFT_CALL(N) z_forward_n<T, N>(a, stream)
  SPFFT_FAST_SWITCH(nz, SPFFT_CALL)
#undef SPFFT_CALL
}

template <typename T>
void launch_fast_xy_backward(int nx, int ny, const FastArgs<T>& a, hipStream_t stream) {
  if (nx != ny) throw GPUFFTError();
#define SPFFT_CALL(N) xy_backward_n<T, N>(a, stream)
  SPFFT_FAST_SWITCH(nx, SPFFT_CALL)
#undef SPFFT_CALL
}

template <typename T>
void launch_fast_xy_forward(int nx, int ny, const FastArgs<T>& a, hipStream_t stream) {
  if (nx != ny) throw GPUFFTError();
#define SPFFT_CALL(N) xy_forward_n<T, N>(a, stream)
  SPFFT_FAST_SWITCH(nx, SPFFT_CALL)
#undef SPFFT_CALL
}

#define SPFFT_INSTANTIATE_FAST(T)                                                        \
  template bool fast_path_supported<T>(int, int, int, int);                              \
  template void launch_fast_z_backward<T>(int, const FastArgs<T>&, hipStream_t);         \
  template void launch_fast_z_forward<T>(int, const FastArgs<T>&, hipStream_t);          \
  template void launch_fast_xy_backward<T>(int, int, const FastArgs<T>&, hipStream_t);   \
  template void launch_fast_xy_forward<T>(int, int, const FastArgs<T>&, hipStream_t);

SPFFT_INSTANTIATE_FAST(double)
SPFFT_INSTANTIATE_FAST(float)

}  // namespace spfft
