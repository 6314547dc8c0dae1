// Fixed-length kernels of the sparse 3D FFT (C2C and R2C) for gfx950 (see
// fast_kernels.hpp for the decomposition).
//
// FFT engine (wave level): a line of N points is owned by P = N / E lanes of
// one wavefront, lane q holding elements q + P e (e < E) in registers. Each
// Stockham pass applies radix-R butterflies to lane-local elements (R divides
// E, so every butterfly's inputs sit in one lane); all passes but the last
// exchange results through the line's own LDS region, whose index map
// i -> i + i / 8 keeps most of the stride-R writes free of
// bank conflicts. The last pass leaves the line in natural order in the
// lanes' registers. Exchanges are wave-local (no workgroup barrier), so the
// waves of a workgroup run their FFTs independently.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>
#include <unordered_map>

#include "fast_kernels.hpp"
#include "fft_common.hpp"
#include "hip_check.hpp"
#include "profiling.hpp"

namespace spfft {

extern __shared__ __attribute__((aligned(16))) unsigned char spfft_lds_fast[];

// One call into the z kernels or the split plane stage of one length. Those
// kernels are compiled in four groups of lengths (four translation units per
// precision and family, so that they build in parallel); each group's
// dispatcher launches the call and returns true when it holds the length.
template <typename T>
struct ZCall {
  enum Op { Backward, Forward, Batch, Geometry } op = Backward;
  const FastArgs<T>* args = nullptr;  // the transform, or the batch's table
  const int2* map = nullptr;          // batches: (transform, block) of each block
  int blocks = 0;
  int useRuns = 0;
  bool forward = false;  // batches
  bool rc = false;       // batches
  int* geom = nullptr;   // Geometry: sticks per block, line length, pad shift
  hipStream_t stream = nullptr;
};
template <typename T>
struct SplitCall {
  bool y = false;  // the y kernel (columns) or the x kernel (rows)
  bool forward = false;
  bool batch = false;
  bool rc = false;                    // batches (single transforms: FastArgs::r2c)
  const FastArgs<T>* args = nullptr;  // the transform, or the batch's table
  int ntrans = 0, ny = 0, nz = 0, maxNcol = 0;  // batches
  hipStream_t stream = nullptr;
};
bool z_group0(int n, const ZCall<double>& c);
bool z_group0(int n, const ZCall<float>& c);
bool z_group1(int n, const ZCall<double>& c);
bool z_group1(int n, const ZCall<float>& c);
bool split_group0(int n, const SplitCall<double>& c);
bool split_group0(int n, const SplitCall<float>& c);
bool split_group1(int n, const SplitCall<double>& c);
bool split_group1(int n, const SplitCall<float>& c);
bool z_group2(int n, const ZCall<double>& c);
bool z_group2(int n, const ZCall<float>& c);
bool split_group2(int n, const SplitCall<double>& c);
bool split_group2(int n, const SplitCall<float>& c);
bool z_group3(int n, const ZCall<double>& c);
bool z_group3(int n, const ZCall<float>& c);
bool split_group3(int n, const SplitCall<double>& c);
bool split_group3(int n, const SplitCall<float>& c);

namespace {

using namespace dev;

constexpr size_t kLdsDefault = 64 * 1024;
constexpr size_t kLdsMax = 160 * 1024;
constexpr int kWave = 64;

// ---- line shapes and radix schedules (compile time)

// Elements per lane: 8 for powers of two, 24 for lengths with a factor 3.
// The FP64 z kernels use half of that (more lanes per line, more waves in
// flight).
constexpr int elems_per_lane(int n) { return n % 3 == 0 ? 24 : 8; }
constexpr int z_elems_per_lane(int n) { return n % 3 == 0 ? 12 : 4; }

template <int N, int EL = elems_per_lane(N)>
struct Line {
  static constexpr int E = EL;
  static constexpr int P = N / E;             // lanes per line
  static constexpr int PER_WAVE = kWave / P;  // whole lines per wave
  // Lanes past the last whole line of a wave (P not dividing the wave): they
  // run a dummy line in an exchange region of their own and store nothing.
  static constexpr bool SPARE = PER_WAVE * P < kWave;
  static_assert(N % E == 0 && P <= kWave, "line shape");
  // True for the spare lanes of a wave.
  __device__ static bool spare(int lane) { return SPARE && lane / P >= PER_WAVE; }
};

constexpr int wave_radix(int m, int e) {
  return (m % 8 == 0 && e % 8 == 0)     ? 8
         : (m % 4 == 0 && e % 4 == 0)   ? 4
         : (m % 2 == 0 && e % 2 == 0)   ? 2
         : (m % 3 == 0 && e % 3 == 0)   ? 3
         : (m % 5 == 0 && e % 5 == 0)   ? 5
         : (m % 7 == 0 && e % 7 == 0)   ? 7
         : (m % 11 == 0 && e % 11 == 0) ? 11
         : (m % 13 == 0 && e % 13 == 0) ? 13
                                        : -1;
}
// Radices without a constant butterfly: their DFT reads W_R from the line's
// own twiddle table (W_N).
constexpr bool table_radix(int r) { return r == 7 || r == 11 || r == 13; }
// True when every pass of the wave-level schedule of an n-point line with e
// elements per lane finds a radix dividing both the remaining length and e,
// and the n / e lanes of a line fit a wavefront.
constexpr bool dividing_schedule(int n, int e) {
  if (e <= 0 || n % e != 0 || n / e > kWave) return false;
  int m = n;
  while (m > 1) {
    const int r = wave_radix(m, e);
    if (r < 0) return false;
    m /= r;
  }
  return true;
}

// Mixed schedule, for lines whose lane element count cannot take every radix
// (an odd e = 15, 25 or 27 with factors 2 in n): the first pass
// reads the lane's own elements and the last pass leaves them in natural
// order, so their radices divide e; the passes between them read their
// butterflies' inputs from the line's LDS region anyway, so a lane there
// takes ceil((n / r) / P) butterflies of any radix r (the surplus ones of the
// last lanes are masked). Middle radices are picked to keep the fewest
// elements per lane (r * ceil((n / r) / P)), the larger radix on ties.
struct MixedSchedule {
  int radix[16] = {};
  int count = 0;
  bool ok = false;
};
constexpr int dividing_pick(int m, int e) {
  return (m % 13 == 0 && e % 13 == 0)   ? 13
         : (m % 11 == 0 && e % 11 == 0) ? 11
         : (m % 7 == 0 && e % 7 == 0)   ? 7
         : (m % 5 == 0 && e % 5 == 0)   ? 5
         : (m % 3 == 0 && e % 3 == 0)   ? 3
         : (m % 8 == 0 && e % 8 == 0)   ? 8
         : (m % 4 == 0 && e % 4 == 0)   ? 4
         : (m % 2 == 0 && e % 2 == 0)   ? 2
                                        : -1;
}
constexpr MixedSchedule mixed_schedule(int n, int e) {
  MixedSchedule s;
  if (e <= 0 || n % e != 0 || n / e > kWave) return s;
  const int p = n / e;
  const int first = dividing_pick(n, e);
  if (first < 0) return s;
  const int last = dividing_pick(n / first, e);
  if (last < 0) return s;
  s.radix[s.count++] = first;
  int m = n / (first * last);
  const int cand[8] = {8, 4, 2, 3, 5, 7, 11, 13};
  while (m > 1) {
    int best = -1, bestCost = 1 << 30;
    for (int i = 0; i < 8; ++i) {
      const int r = cand[i];
      if (m % r != 0) continue;
      const int cost = r * ((n / r + p - 1) / p);
      if (cost < bestCost || (cost == bestCost && r > best)) {
        best = r;
        bestCost = cost;
      }
    }
    if (best < 0 || s.count >= 15) return s;
    s.radix[s.count++] = best;
    m /= best;
  }
  s.radix[s.count++] = last;
  s.ok = true;
  return s;
}
// Product of the radices of the passes before `pass`.
constexpr int mixed_ns(int n, int e, int pass) {
  const MixedSchedule s = mixed_schedule(n, e);
  int ns = 1;
  for (int i = 0; i < pass; ++i) ns *= s.radix[i];
  return ns;
}
// A line of n points with e elements per lane has a wave-level schedule.
constexpr bool schedule_ok(int n, int e) {
  return dividing_schedule(n, e) || mixed_schedule(n, e).ok;
}

// Lengths served by the split plane stage only (no class-split or whole-plane
// kernels): elements per lane of their full-length lines and of their z lines
// (FP64, FP32), chosen so that each line's schedule exists (schedule_ok), as
// a dividing schedule (every radix divides the lane's elements, e.g. 80 = 8
// lanes x 10: 2 2 2 2 5) or a mixed one (mixed_schedule, e.g. 120 = 8 x 15:
// 5 8 3; 144 = 16 x 9: 3 2 2 2 2 3). Lanes per line not dividing the wave
// leave spare lanes (50 = 10 x 5, 90 = 9 x 10, ...). Radix 7, 11 and 13 take
// their roots from the line's twiddle table; lengths with a factor 11 or 13
// need lines of 22 or 26 elements. Measured on the MI355X, lines of 3 to 10
// elements run fastest (lines of 21 to 27 held up to 256 VGPRs: 81^3 FP64
// 13.2k -> 32.4k pairs/s at 3 elements), except at 120, 200 and 240, at 400
// and 432 in FP64 and at 288 and 320 in FP32, which keep longer lines (their
// shorter ones measured 3 to 30% slower); FP32 448 and 512 take longer
// lines (28; 16 and, for z, 32) so
// that a workgroup's row segments are 64 to 128 bytes (512^3: +36%). Odd
// lengths run on the same lines; an R2C plane with an odd row count leaves
// the last packed row pair with one row.
struct LaneElems {
  int full64, full32, z64, z32;
};
constexpr LaneElems lane_elems(int n) {
  switch (n) {
    case 36: return {3, 3, 3, 3};
    case 40: return {4, 4, 4, 4};
    case 42: return {6, 6, 6, 6};
    case 44: return {2, 2, 2, 2};
    case 45: return {3, 3, 3, 3};
    case 48: return {4, 4, 4, 4};
    case 50: return {5, 5, 5, 5};
    case 52: return {2, 2, 2, 2};
    case 54: return {3, 3, 3, 3};
    case 56: return {4, 4, 4, 4};
    case 60: return {6, 6, 6, 6};
    case 63: return {3, 3, 3, 3};
    case 66: return {6, 6, 6, 6};
    case 70: return {10, 10, 10, 10};
    case 72: return {4, 4, 4, 4};
    case 75: return {5, 5, 5, 5};
    case 78: return {6, 6, 6, 6};
    case 80: return {4, 4, 4, 4};
    case 81: return {3, 3, 3, 3};
    case 84: return {6, 6, 6, 6};
    case 88: return {4, 4, 4, 4};
    case 90: return {3, 3, 3, 3};
    case 98: return {7, 7, 7, 7};
    case 99: return {9, 9, 9, 9};
    case 100: return {5, 5, 5, 5};
    case 104: return {4, 4, 4, 4};
    case 105: return {15, 15, 15, 15};
    case 108: return {6, 6, 6, 6};
    case 110: return {10, 10, 10, 10};
    case 112: return {4, 4, 4, 4};
    case 117: return {9, 9, 9, 9};
    case 120: return {15, 15, 15, 15};
    case 121: return {11, 11, 11, 11};
    case 125: return {5, 5, 5, 5};
    case 126: return {6, 6, 6, 6};
    case 130: return {10, 10, 10, 10};
    case 132: return {12, 12, 12, 12};
    case 135: return {9, 9, 9, 9};
    case 140: return {10, 10, 10, 10};
    case 144: return {8, 8, 8, 8};
    case 147: return {7, 7, 7, 7};
    case 150: return {5, 5, 5, 5};
    case 154: return {14, 14, 14, 14};
    case 156: return {12, 12, 12, 12};
    case 160: return {8, 8, 8, 8};
    case 162: return {6, 6, 6, 6};
    case 165: return {15, 15, 15, 15};
    case 168: return {12, 14, 12, 14};
    case 169: return {13, 13, 13, 13};
    case 175: return {25, 25, 25, 25};
    case 176: return {16, 16, 16, 16};
    case 180: return {12, 12, 12, 12};
    case 182: return {14, 14, 14, 14};
    case 189: return {3, 9, 3, 9};
    case 195: return {15, 15, 15, 15};
    case 196: return {7, 7, 7, 7};
    case 198: return {18, 18, 18, 18};
    case 200: return {10, 10, 10, 10};
    case 208: return {16, 16, 16, 16};
    case 210: return {10, 10, 10, 10};
    case 216: return {12, 12, 12, 12};
    case 220: return {20, 20, 20, 20};
    case 224: return {8, 14, 8, 14};
    case 225: return {9, 9, 9, 9};
    case 231: return {21, 21, 21, 21};
    case 234: return {18, 18, 18, 18};
    case 240: return {15, 15, 15, 15};
    case 242: return {22, 22, 22, 22};
    case 243: return {9, 9, 9, 9};
    case 252: return {12, 12, 12, 12};
    case 256: return {8, 8, 4, 8};
    case 260: return {20, 20, 20, 20};
    case 264: return {22, 22, 22, 22};
    case 270: return {9, 9, 9, 9};
    case 273: return {21, 21, 21, 21};
    case 275: return {25, 25, 25, 25};
    case 280: return {10, 10, 10, 10};
    case 286: return {26, 26, 26, 26};
    case 288: return {9, 18, 9, 18};
    case 297: return {27, 27, 27, 27};
    case 300: return {10, 10, 10, 10};
    case 312: return {24, 24, 24, 24};
    case 320: return {10, 20, 10, 10};
    case 325: return {25, 25, 25, 25};
    case 330: return {22, 22, 22, 22};
    case 336: return {12, 21, 12, 21};
    case 338: return {26, 26, 26, 26};
    case 351: return {27, 27, 27, 27};
    case 352: return {22, 22, 22, 22};
    case 360: return {12, 12, 12, 12};
    case 384: return {12, 24, 12, 12};
    case 390: return {15, 15, 15, 15};
    case 400: return {25, 25, 25, 25};
    case 416: return {26, 26, 26, 26};
    case 420: return {15, 15, 15, 15};
    case 432: return {27, 27, 27, 27};
    case 440: return {22, 22, 22, 22};
    case 448: return {14, 28, 14, 28};
    case 450: return {15, 15, 15, 15};
    case 462: return {21, 21, 21, 21};
    case 480: return {15, 15, 15, 15};
    case 500: return {10, 10, 10, 10};
    case 512: return {8, 16, 8, 32};
    default: return {0, 0, 0, 0};
  }
}
constexpr bool split_only_length(int n) { return lane_elems(n).full64 > 0; }
template <typename T>
constexpr int split_only_elems(int n) {
  return sizeof(T) == 8 ? lane_elems(n).full64 : lane_elems(n).full32;
}
template <typename T>
constexpr int split_only_z_elems(int n) {
  return sizeof(T) == 8 ? lane_elems(n).z64 : lane_elems(n).z32;
}

// Lines of the full-tile kernels take the whole lane's elements in their
// first pass (radix 16 or 24 in registers) so that a line needs one exchange.
constexpr int full_radix(int m, int e) {
  return (m % 24 == 0 && e % 24 == 0)   ? 24
         : (m % 16 == 0 && e % 16 == 0) ? 16
                                        : wave_radix(m, e);
}

// Padded index inside an exchange region: one spare element every 8
// elements (128 bytes in FP64, 64 bytes in FP32; the denser FP32 padding
// measured slightly faster than one per 128 bytes).
template <typename T>
constexpr int pad_shift() {
  return 3;
}
template <typename T>
__device__ __forceinline__ int phys(int i) {
  return i + (i >> pad_shift<T>());
}
// Exchange slot of element i: phys(i) (PE = 0), or, for the full-tile lines
// of P lanes holding PE elements each, one spare slot per PE elements:
// slot(i) = i + i / PE. The P lanes of such a line then reach P distinct
// slots mod P in every exchange write (i = PE q + r: (PE + 1) q + r) and read
// (i = q + P e), and slot offsets stay affine in r and e (immediate offsets
// of one base address).
template <typename T, int PE>
__device__ __forceinline__ int slot(int i) {
  if constexpr (PE == 0) {
    return phys<T>(i);
  } else {
    return i + static_cast<int>(static_cast<unsigned>(i) / PE);
  }
}
template <typename T>
constexpr int padded_len(int n) {
  return n + ((n - 1) >> pad_shift<T>());
}
constexpr int make_odd(int n) { return n | 1; }
constexpr int round_up(int n, int m) { return (n + m - 1) / m * m; }
// Offset (in elements) of an LDS twiddle table placed after n elements:
// aligned to a whole table entry (16 bytes), so each entry loads with one
// instruction.
template <typename T>
__host__ __device__ constexpr int tw_offset(int n) {
  return round_up(n, static_cast<int>(sizeof(Tw<T>) / sizeof(Cx<T>)));
}

// Element `i` of an array smaller than 4 GiB, addressed with a 32-bit byte
// offset (lets the compiler use a scalar base and one offset register).
template <typename X>
__device__ __forceinline__ X at32(const X* base, int i) {
  const unsigned off = static_cast<unsigned>(i) * static_cast<unsigned>(sizeof(X));
  return *reinterpret_cast<const X*>(reinterpret_cast<const char*>(base) + off);
}

// Workgroup barrier for LDS data: waits for this wave's LDS operations only.
// __syncthreads() releases at workgroup scope for all memory, which waits for
// every global load and store in flight (loads issued ahead of the barrier
// would land before it). No kernel here passes data between waves through
// global memory.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Streaming store (no cache allocation on the way to HBM): for batch outputs
// that no later kernel of the chain reads while they could still be cached.
template <typename T>
__device__ __forceinline__ void store_stream(Cx<T>* p, Cx<T> v) {
  __builtin_nontemporal_store(v.re, &p->re);
  __builtin_nontemporal_store(v.im, &p->im);
}
// Two consecutive FP32 elements (16 bytes, 16-byte aligned) in one streaming
// store.
__device__ __forceinline__ void store_stream_pair(Cx<float>* p, Cx<float> a, Cx<float> b) {
  using F4 = __attribute__((ext_vector_type(4))) float;
  const F4 w = {a.re, a.im, b.re, b.im};
  __builtin_nontemporal_store(w, reinterpret_cast<F4*>(p));
}
template <bool STREAM, typename T>
__device__ __forceinline__ void store(Cx<T>* p, Cx<T> v) {
  if constexpr (STREAM) {
    store_stream(p, v);
  } else {
    *p = v;
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct NoHook {
  __device__ void operator()() const {}
};

// One Stockham pass (and, recursively, the remaining ones) of the wave-level
// FFT. `line` is this line's exchange region: element i at
// line[slot<PE>(i) * es] (rows: es = 1; columns of a tile: es = row length).
// The twiddle table holds W_{N * TWS}. NS = product of the radices already
// applied. `hook` runs once, after the first pass has written its outputs to
// LDS (the line's registers are free then): callers issue loads there.
// Pass PASS of the mixed schedule (see mixed_schedule). v: the lane's E
// elements (input of the first pass, output of the last); between passes the
// line lives in its LDS region.
template <typename T, int N, bool BWD, int TWS, int E, int PASS = 0>
__device__ __forceinline__ void wave_fft_mixed(Cx<T>* v, Cx<T>* line, int es, int q,
                                               const Tw<T>* __restrict__ tw) {
  constexpr MixedSchedule S = mixed_schedule(N, E);
  static_assert(S.ok, "no mixed schedule");
  constexpr int PL = N / E;
  constexpr int R = S.radix[PASS];
  static_assert(!table_radix(R) || TWS == 1, "table butterflies need the line's own table");
  constexpr int NB = N / R;  // butterflies of the pass
  constexpr bool FIRST = PASS == 0, LAST = PASS == S.count - 1;
  static_assert(!(FIRST || LAST) || E % R == 0, "first and last radices divide E");
  constexpr int BL = FIRST || LAST ? E / R : (NB + PL - 1) / PL;  // butterflies per lane
  constexpr int NS = mixed_ns(N, E, PASS);
  constexpr int STEP = N / (NS * R) * TWS;
  Cx<T> u[BL][R];
#pragma unroll
  for (int b = 0; b < BL; ++b) {
    const int j = q + PL * b;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if constexpr (FIRST) {
        u[b][r] = v[b + r * (E / R)];
      } else {
        const int jj = j < NB ? j : 0;
        u[b][r] = line[slot<T, 0>(jj + r * NB) * es];
      }
    }
  }
  if constexpr (!FIRST) wave_sync();
#pragma unroll
  for (int b = 0; b < BL; ++b) {
    const int j = q + PL * b;
    if constexpr (NS > 1) {
      const int k = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) u[b][r] = cmul(u[b][r], twiddle<BWD>(tw, k * r * STEP));
    }
    Butterfly<T, R, BWD>::run(u[b], tw, N);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * (E / R)] = u[b][r];
    } else {
      const int o = (j / NS) * NS * R + (j % NS);
      if (FIRST || j < NB) {
#pragma unroll
        for (int r = 0; r < R; ++r) line[slot<T, 0>(o + r * NS) * es] = u[b][r];
      }
    }
  }
  if constexpr (!LAST) {
    wave_sync();
    wave_fft_mixed<T, N, BWD, TWS, E, PASS + 1>(v, line, es, q, tw);
  }
}

template <typename T, int N, bool BWD, int TWS, int E = elems_per_lane(N), int NS = 1,
          int PE = 0, typename Hook = NoHook>
__device__ __forceinline__ void wave_fft(Cx<T>* v, Cx<T>* line, int es, int q,
                                         const Tw<T>* __restrict__ tw, Hook hook = Hook()) {
  if constexpr (NS == 1 && PE == 0 && !dividing_schedule(N, E)) {
    static_assert(std::is_same<Hook, NoHook>::value, "mixed schedules take no hook");
    wave_fft_mixed<T, N, BWD, TWS, E>(v, line, es, q, tw);
    return;
  } else {
  constexpr int PL = N / E;
  constexpr int R = PE > 0 ? full_radix(N / NS, E) : wave_radix(N / NS, E);
  static_assert(R > 0, "no radix divides the lane's element count");
  // radices without a constant butterfly (7, 11, 13) read W_R from the line's table
  static_assert(!table_radix(R) || TWS == 1, "table butterflies need the line's own table");
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
        line[slot<T, PE>(i) * es] = u[r];
      }
    }
  }
  if constexpr (!LAST) {
    hook();
    wave_sync();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = q + PL * e;
      v[e] = line[slot<T, PE>(i) * es];
    }
    wave_sync();
    wave_fft<T, N, BWD, TWS, E, NS * R, PE>(v, line, es, q, tw);
  }
  }
}

// One output of the decimation-in-frequency split over R inputs x[l] with
// the class's twiddles w[l] = W_R^(+-r l): sum_l w[l] x[l]. For R = 4 the
// twiddles are 1, w1, w2 = +-1, w3 = conj(w1), and the sum needs 10 real
// operations instead of 18: w1 x1 + conj(w1) x3 = (c p.re - s d.im,
// c p.im + s d.re) with p = x1 + x3, d = x1 - x3, w1 = c + i s.
template <typename T, int R>
__device__ __forceinline__ Cx<T> dif_sum(const Cx<T>* x, const Tw<T>* w) {
  if constexpr (R == 4) {
    // W_4 roots have components 0 and +-1: their table entries are exact
    const Cx<T> w1 = tw_root(w[1]);
    const T w2 = tw_root(w[2]).re;
    const T pr = x[1].re + x[3].re, pi = x[1].im + x[3].im;
    const T dr = x[1].re - x[3].re, di = x[1].im - x[3].im;
    return Cx<T>{x[0].re + w2 * x[2].re + w1.re * pr - w1.im * di,
                 x[0].im + w2 * x[2].im + w1.re * pi + w1.im * dr};
  } else {
    Cx<T> acc = x[0];
#pragma unroll
    for (int l = 1; l < R; ++l) acc = cadd(acc, cmul(x[l], w[l]));
    return acc;
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
  // FP64 and lengths with a factor 3: narrow lines (more waves in flight);
  // FP32 powers of two: full lines, so a workgroup owns 16 sticks and its
  // [z][stick] row segments are whole 128-byte lines
  using L = Line<NZ, split_only_length(NZ)
                            ? split_only_z_elems<T>(NZ)
                            : ((sizeof(T) == 4 && NZ % 3 != 0) ? elems_per_lane(NZ)
                                                               : z_elems_per_lane(NZ))>;
  static_assert(schedule_ok(NZ, L::E), "z line schedule");
  static constexpr int NT = 256;
  static constexpr int S = (NT / kWave) * L::PER_WAVE;  // sticks per workgroup
  static constexpr int LZ = padded_len<T>(NZ);
  // entries per thread when no triplet repeats (S full sticks)
  static constexpr int K = (S * NZ + NT - 1) / NT;
  // tile, one spare element that absorbs repeated triplets, the spare lanes'
  // dummy lines (one per wave), twiddle table
  static constexpr int DUMMY = S * LZ + 1;
  static constexpr int TW = DUMMY + (L::SPARE ? (NT / kWave) * LZ : 0);
  // Stick of this lane within the workgroup (S for a spare lane: none) and
  // the tile row its line exchanges through.
  __device__ static int stick(int lane) {
    return L::spare(lane) ? S : static_cast<int>(threadIdx.x / kWave) * L::PER_WAVE + lane / L::P;
  }
  template <typename C>
  __device__ static C* line(C* tile, int lane) {
    return L::spare(lane) ? tile + DUMMY + (threadIdx.x / kWave) * LZ : tile + stick(lane) * LZ;
  }
  // twiddle table after the tile
  static size_t lds() {
    return static_cast<size_t>(tw_offset<T>(TW)) * sizeof(Cx<T>) +
           static_cast<size_t>(NZ) * sizeof(Tw<T>);
  }
};

// Hermitian fill of the (0,0) stick in the tile (reference
// symmetry_host.hpp:72-89): a non-zero lower entry wins, an upper entry fills
// a zero lower one, the self-paired entry is conjugated.
template <typename T, int NZ, int NT>
__device__ __forceinline__ void fill_zero_zero_stick(Cx<T>* tile) {
  for (int k = 1 + threadIdx.x; 2 * k <= NZ; k += NT) {
    Cx<T>& lo = tile[phys<T>(k)];
    Cx<T>& hi = tile[phys<T>(NZ - k)];
    if (2 * k == NZ) {
      if (cnonzero(lo)) lo = cconj(lo);
    } else if (cnonzero(lo)) {
      hi = cconj(lo);
    } else if (cnonzero(hi)) {
      lo = cconj(hi);
    }
  }
}

// Sticks whose values are one run in user order (zInfo[s] = {first value,
// first storage z, length}: value of storage z = first + ((z - z0) mod NZ)
// for the len positions of the run): every lane loads its z values straight
// into registers, with no entry table, zero fill or scatter.
template <typename T, int NZ, bool RC, bool STREAM>
__device__ __forceinline__ void z_backward_runs(const FastArgs<T>& a, int bid) {
  using Z = ZShape<T, NZ>;
  using L = typename Z::L;
  constexpr int S = Z::S, LZ = Z::LZ, NT = Z::NT;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(tile + tw_offset<T>(Z::TW));
  const int s0 = bid * S;
  const int ns = min(S, a.nst - s0);
  const int lane = threadIdx.x % kWave;
  const int sl = Z::stick(lane);
  const int q = lane % L::P;
  const bool live = sl < S && s0 + sl < a.nst;
  const int4 run = a.zInfo[live ? s0 + sl : a.nst - 1];
  const Cx<T>* in = reinterpret_cast<const Cx<T>*>(a.in);
  Cx<T> v[L::E];
#pragma unroll
  for (int e = 0; e < L::E; ++e) {
    int pos = q + L::P * e - run.y;
    pos += pos < 0 ? NZ : 0;
    const bool present = live && pos < run.z;
    const Cx<T> x = in[present ? run.x + pos : run.x];
    v[e] = present ? x : Cx<T>{T(0), T(0)};
  }
  for (int t = threadIdx.x; t < NZ; t += NT) twl[t] = a.twZ[t];
  Cx<T>* line = Z::line(tile, lane);
  if (RC && s0 == 0 && a.zeroZeroStick == 0) {
#pragma unroll
    for (int e = 0; e < L::E; ++e) line[phys<T>(q + L::P * e)] = v[e];
    __syncthreads();
    fill_zero_zero_stick<T, NZ, NT>(tile);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < L::E; ++e) v[e] = line[phys<T>(q + L::P * e)];
    wave_sync();
  } else {
    __syncthreads();  // twiddles
  }
  // the xy kernels read element nst of every row as zero
  if (bid == 0) {
    for (int t = threadIdx.x; t < NZ; t += NT) {
      a.zbuf[static_cast<size_t>(t) * a.pitch + a.nst] = Cx<T>{T(0), T(0)};
    }
  }
  wave_fft<T, NZ, true, 1, L::E>(v, line, 1, q, twl);
#pragma unroll
  for (int e = 0; e < L::E; ++e) line[phys<T>(q + L::P * e)] = v[e];
  __syncthreads();
  for (int t = threadIdx.x; t < S * NZ; t += NT) {
    const int s = t % S, zz = t / S;
    if (s < ns) store<STREAM>(&a.zbuf[static_cast<size_t>(zz) * a.pitch + s0 + s], tile[s * LZ + phys<T>(zz)]);
  }
}

// STREAM: streaming stores of the [z][stick] rows (batches, whose
// intermediates exceed the caches anyway)
template <typename T, int NZ, bool RC, bool STREAM>
__device__ __forceinline__ void z_backward_body(const FastArgs<T>& a, int bid) {
  using Z = ZShape<T, NZ>;
  using L = typename Z::L;
  constexpr int S = Z::S, LZ = Z::LZ, NT = Z::NT, K = Z::K;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(tile + tw_offset<T>(Z::TW));
  const int s0 = bid * S;
  const int ns = min(S, a.nst - s0);
  if (a.hasRuns) {
    z_backward_runs<T, NZ, RC, STREAM>(a, bid);
    return;
  }
  const int e0 = a.stickBegin[s0], e1 = a.stickBegin[s0 + ns];
  // Entries carry their tile position (a repeated triplet's earlier copies
  // point at the spare slot after the tile, so the last one wins). All of a
  // thread's entry loads are issued together, then all its value loads.
  int ev[K], ep[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = e0 + threadIdx.x + k * NT;
    const int2 x = a.zEntries[e < e1 ? e : e0];
    ev[k] = x.x;
    ep[k] = e < e1 && x.y >= 0 ? x.y : S * LZ;
  }
  // twiddles in flight while the tile is cleared (TWK per thread)
  constexpr int TWK = (NZ + NT - 1) / NT;
  Tw<T> tw[TWK];
#pragma unroll
  for (int k = 0; k < TWK; ++k) {
    const int t = threadIdx.x + k * NT;
    tw[k] = a.twZ[t < NZ ? t : 0];
  }
  for (int t = threadIdx.x; t < S * LZ; t += NT) tile[t] = Cx<T>{T(0), T(0)};
  T vr[K], vi[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const Cx<T> x = reinterpret_cast<const Cx<T>*>(a.in)[ev[k]];
    vr[k] = x.re;
    vi[k] = x.im;
  }
#pragma unroll
  for (int k = 0; k < TWK; ++k) {
    const int t = threadIdx.x + k * NT;
    if (t < NZ) twl[t] = tw[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) tile[ep[k]] = Cx<T>{vr[k], vi[k]};
  // more entries than S full sticks only with repeated triplets
  for (int e = e0 + threadIdx.x + K * NT; e < e1; e += NT) {
    const int2 x = a.zEntries[e];
    tile[x.y < 0 ? S * LZ : x.y] = reinterpret_cast<const Cx<T>*>(a.in)[x.x];
  }
  __syncthreads();
  if (RC && s0 == 0 && a.zeroZeroStick == 0) {
    fill_zero_zero_stick<T, NZ, NT>(tile);
    __syncthreads();
  }
  // the xy kernels read element nst of every row as zero
  if (bid == 0) {
    for (int t = threadIdx.x; t < NZ; t += NT) {
      a.zbuf[static_cast<size_t>(t) * a.pitch + a.nst] = Cx<T>{T(0), T(0)};
    }
  }
  const int lane = threadIdx.x % kWave;
  const int q = lane % L::P;
  Cx<T>* line = Z::line(tile, lane);
  Cx<T> v[L::E];
#pragma unroll
  for (int e = 0; e < L::E; ++e) v[e] = line[phys<T>(q + L::P * e)];
  wave_sync();
  wave_fft<T, NZ, true, 1, L::E>(v, line, 1, q, twl);
#pragma unroll
  for (int e = 0; e < L::E; ++e) line[phys<T>(q + L::P * e)] = v[e];
  __syncthreads();
  for (int t = threadIdx.x; t < S * NZ; t += NT) {
    const int s = t % S, zz = t / S;
    if (s < ns) store<STREAM>(&a.zbuf[static_cast<size_t>(zz) * a.pitch + s0 + s], tile[s * LZ + phys<T>(zz)]);
  }
}

template <typename T, int NZ>
__device__ __forceinline__ void z_forward_body(const FastArgs<T>& a, int bid) {
  using Z = ZShape<T, NZ>;
  using L = typename Z::L;
  constexpr int S = Z::S, LZ = Z::LZ, NT = Z::NT, K = Z::K;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(tile + tw_offset<T>(Z::TW));
  const int s0 = bid * S;
  const int ns = min(S, a.nst - s0);
  if constexpr (S * NZ != L::E * NT || NT % S != 0) {
    // spare lanes: no whole-tile thread mapping
    for (int t = threadIdx.x; t < S * NZ; t += NT) {
      const int s = t % S, zz = t / S;
      tile[s * LZ + phys<T>(zz)] = s < ns ? a.zbuf[static_cast<size_t>(zz) * a.pitch + s0 + s]
                                          : Cx<T>{T(0), T(0)};
    }
  } else {
    // thread t holds stick t % S at z = t / S + P k, k < E (S NZ = E NT): all
    // loads in flight before the tile stores
    constexpr int ZS = NT / S;  // z rows per pass of the workgroup
    const int s = threadIdx.x % S, zz0 = threadIdx.x / S;
    const int sc = s < ns ? s : 0;
    Cx<T> w[L::E];
#pragma unroll
    for (int k = 0; k < L::E; ++k) {
      w[k] = a.zbuf[static_cast<size_t>(zz0 + ZS * k) * a.pitch + s0 + sc];
    }
#pragma unroll
    for (int k = 0; k < L::E; ++k) {
      tile[s * LZ + phys<T>(zz0 + ZS * k)] = s < ns ? w[k] : Cx<T>{T(0), T(0)};
    }
  }
  for (int t = threadIdx.x; t < NZ; t += NT) twl[t] = a.twZ[t];
  if (a.hasRuns) {
    // sticks stored as runs: each lane writes its outputs straight to the
    // run (see z_backward_runs)
    const int lane = threadIdx.x % kWave;
    const int sl = Z::stick(lane);
    const int q = lane % L::P;
    const bool live = sl < S && s0 + sl < a.nst;
    const int4 run = a.zInfo[live ? s0 + sl : a.nst - 1];
    __syncthreads();
    Cx<T>* line = Z::line(tile, lane);
    Cx<T> v[L::E];
#pragma unroll
    for (int e = 0; e < L::E; ++e) v[e] = line[phys<T>(q + L::P * e)];
    wave_sync();
    wave_fft<T, NZ, false, 1, L::E>(v, line, 1, q, twl);
    Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out);
#pragma unroll
    for (int e = 0; e < L::E; ++e) {
      int pos = q + L::P * e - run.y;
      pos += pos < 0 ? NZ : 0;
      if (live && pos < run.z) out[run.x + pos] = a.useScale ? cscale(v[e], a.scale) : v[e];
    }
    return;
  }
  // the gather's entries are loaded while the sticks arrive
  const int e0 = a.stickBegin[s0], e1 = a.stickBegin[s0 + ns];
  int ev[K], ep[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = e0 + threadIdx.x + k * NT;
    const int2 x = a.zEntries[e < e1 ? e : e0];
    ev[k] = e < e1 ? x.x : -1;
    ep[k] = x.y & 0x7fffffff;
  }
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int q = lane % L::P;
  Cx<T>* line = Z::line(tile, lane);
  Cx<T> v[L::E];
#pragma unroll
  for (int e = 0; e < L::E; ++e) v[e] = line[phys<T>(q + L::P * e)];
  wave_sync();
  wave_fft<T, NZ, false, 1, L::E>(v, line, 1, q, twl);
#pragma unroll
  for (int e = 0; e < L::E; ++e) line[phys<T>(q + L::P * e)] = v[e];
  __syncthreads();
  Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    Cx<T> c = tile[ep[k]];
    if (a.useScale) c = cscale(c, a.scale);
    if (ev[k] >= 0) out[ev[k]] = c;
  }
  for (int e = e0 + threadIdx.x + K * NT; e < e1; e += NT) {
    const int2 x = a.zEntries[e];
    Cx<T> c = tile[x.y & 0x7fffffff];
    if (a.useScale) c = cscale(c, a.scale);
    out[x.x] = c;
  }
}

// ------------------------------------------------------------ xy kernels

// Full-length lines of the class-split kernels (backward x, forward y): FP64
// lengths with a factor 3 take 12 elements per lane (radices 4, 4, 2, 3 at
// 96), since 24 FP64 elements need about 230 registers per lane and halve the
// waves a SIMD holds.
template <typename T, int N>
using FullLine = Line<N, split_only_length(N)
                             ? split_only_elems<T>(N)
                             : ((sizeof(T) == 8 && N % 3 == 0) ? 12 : elems_per_lane(N))>;

template <typename T, int N>
struct XYShape {
  static constexpr int R = fast_classes(N);     // residue classes per plane
  static constexpr int M = N / R;               // rows (bwd) / columns (fwd) per class
  using FL = FullLine<T, N>;
  // enough waves for all rows (columns) of a class in one round of lines
  static constexpr int NT = (N >= 128 && N < 192) || M > 4 * FL::PER_WAVE ? 512 : 256;
  static constexpr int NW = NT / kWave;
  static_assert(N <= NT, "one twiddle per thread");
  // lines of M points (forward x, backward y): half the usual elements per
  // lane, so that a lane's gathers form one batch and twice as many lanes
  // share them; the backward y lines of lengths with a factor 3 take a
  // quarter (6 elements: 12 would keep 48 gathers per lane in flight)
  using YL = Line<M, elems_per_lane(M) / 2>;
  using YB = Line<M, elems_per_lane(M) / (M % 3 == 0 ? 4 : 2)>;
  static_assert(!FL::SPARE && !YL::SPARE && !YB::SPARE, "whole lines per wave");
  // backward: A = [m][LU] (column u at offset u, y transforms in column
  // mode), reused as B = [rows][LB] (x transforms)
  static constexpr int ROWS_A = padded_len<T>(M);
  static constexpr int LB = padded_len<T>(N);
  static constexpr int ROWS_ALLOC = round_up(M, FL::PER_WAVE);
  static_assert(NW * FL::PER_WAVE >= M, "all rows of a class in one round");
  // forward: Tb = [rows y][LT], columns transformed in place with stride LT;
  // the rows have room for the spare columns of a partial last wave
  static constexpr int LT = make_odd(padded_len<T>(round_up(M, FL::PER_WAVE)));
  // Row pitch of A: whole waves of columns including the zero column ncol,
  // and odd, so that the 8 lanes of one column in an FP64 LDS cycle (rows
  // q = 0..7 of the y line) hit distinct bank groups.
  static int pitch_a(int ncol) { return round_up(ncol + 1, YB::PER_WAVE) | 1; }
  // elements of the A/B region, followed by the twiddle table (N) and, for
  // R2C, the filled x = 0 column (N)
  static size_t tile_backward(int ncol) {
    const size_t a = static_cast<size_t>(ROWS_A) * pitch_a(ncol);
    const size_t b = static_cast<size_t>(ROWS_ALLOC) * LB;
    return a > b ? a : b;
  }
  static size_t lds_backward(int ncol, bool rc) {
    return (tw_offset<T>(static_cast<int>(tile_backward(ncol))) + (rc ? N : 0)) * sizeof(Cx<T>) +
           static_cast<size_t>(N) * sizeof(Tw<T>);
  }
  // rows of the tile: the y lines' exchange (padded index map) reaches row
  // padded_len(N) - 1; the x pass may round the row count up to whole waves
  static constexpr int ROWS_FWD =
      padded_len<T>(N) > round_up(N, YL::PER_WAVE) ? padded_len<T>(N)
                                                  : round_up(N, YL::PER_WAVE);
  // tile, then the MX-point twiddle table W_M of the x lines (their FFTs then
  // load no twiddles from memory, so the next rows' loads can stay in flight)
  static constexpr int TW_FWD = tw_offset<T>(ROWS_FWD * LT);
  static size_t lds_forward() {
    return static_cast<size_t>(TW_FWD) * sizeof(Cx<T>) + static_cast<size_t>(M) * sizeof(Tw<T>);
  }
};

template <typename T, int N, bool RC>
__device__ __forceinline__ void xy_backward_body(const FastArgs<T>& a, int bid) {
  using S = XYShape<T, N>;
  using LM = typename S::YB;
  using LX = typename S::FL;
  constexpr int R = S::R, M = S::M, LB = S::LB, NW = S::NW;
  const int ncol = a.ncol;
  const int LU = a.pitchA;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  int z, r;
  plane_class<R>(bid, z, r);
  z += a.zBegin;
  if (z >= a.nz) return;
  const Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.pitch;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  // column of each x this lane handles in the x transform (loaded early);
  // for R2C the upper half of a row is the conjugate of x' = N - x
  int uu[LX::E];
#pragma unroll
  for (int e = 0; e < LX::E; ++e) {
    const int x = lane % LX::P + LX::P * e;
    uu[e] = a.xCol[(RC && 2 * x > N) ? N - x : x];
  }
  // The twiddle table W_N (x and y have the same length) is staged in LDS
  // after the tile.
  const int tileA = S::ROWS_A * LU, tileB = S::ROWS_ALLOC * LB;
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(A + tw_offset<T>(tileA > tileB ? tileA : tileB));
  if (threadIdx.x < N) twl[threadIdx.x] = a.twX[threadIdx.x];
  // R2C: hermitian fill of the plane's x = 0 column before its y transform
  // (reference symmetry_host.hpp:47-58, same rule as the (0,0) stick), into
  // LDS after the twiddles; the column's gathers read it from there.
  Cx<T>* fill = reinterpret_cast<Cx<T>*>(twl + N);
  const bool fillCol = RC && a.colX[0] == 0;
  if (fillCol) {
    for (int y = threadIdx.x; y < N; y += S::NT) {
      const int yp = y == 0 ? 0 : N - y;
      const Cx<T> xs = zrow[a.stickAt[(y % M) * R + y / M]];
      const Cx<T> xp = zrow[a.stickAt[(yp % M) * R + yp / M]];
      Cx<T> o = xs;
      if (y != 0) {
        if (yp == y) {
          o = cnonzero(xs) ? cconj(xs) : xs;
        } else if (y < yp) {
          o = cnonzero(xs) ? xs : (cnonzero(xp) ? cconj(xp) : xs);
        } else {
          o = cnonzero(xp) ? cconj(xp) : xs;
        }
      }
      fill[y] = o;
    }
  }
  __syncthreads();

  // y: for each column u, decimation in frequency to the rows y = r + R m
  // (t[j] = W^(r j) sum_l X[j + M l] W_R^(r l)), then the M-point FFT. The
  // gathers are branch free (absent sticks read the zero element), so a
  // lane's loads are all in flight together.
  {
    const int q = lane % LM::P;
    Tw<T> wl[R];
    wl[0] = tw_one<T>();
#pragma unroll
    for (int l = 1; l < R; ++l) wl[l] = twiddle<true>(twl, ((r * l) % R) * M);
    // Columns up to and including ncol, which comes out zero (its stickAt
    // row points at the zero element) and serves the x pass as the column of
    // every absent x. Lanes past it read that row too.
    for (int u0 = wave * LM::PER_WAVE; u0 <= ncol; u0 += NW * LM::PER_WAVE) {
      const int u = u0 + lane / LM::P;
      const int uc = u < ncol ? u : ncol;
      const int* rel = a.stickAt + static_cast<size_t>(uc) * N;
      const Cx<T>* col = zrow;
      Cx<T> v[LM::E];
      // elements gathered per batch: at most 16 gathers per lane in flight
      constexpr int EC = (LM::E % 4 == 0 && 4 * R <= 16) ? 4 : (LM::E % 2 == 0 ? 2 : LM::E);
      static_assert(R % 4 == 0, "16-byte index loads");
#pragma unroll
      for (int e0 = 0; e0 < LM::E; e0 += EC) {
        int sv[EC][R];
#pragma unroll
        for (int e = 0; e < EC; ++e) {
          const int j = q + LM::P * (e0 + e);
#pragma unroll
          for (int l = 0; l < R; l += 4) {
            const int4 x = at32(reinterpret_cast<const int4*>(rel), (j * R + l) / 4);
            sv[e][l] = x.x;
            sv[e][l + 1] = x.y;
            sv[e][l + 2] = x.z;
            sv[e][l + 3] = x.w;
          }
        }
        Cx<T> xv[EC][R];
#pragma unroll
        for (int e = 0; e < EC; ++e) {
#pragma unroll
          for (int l = 0; l < R; ++l) xv[e][l] = at32(col, sv[e][l]);
        }
#pragma unroll
        for (int e = 0; e < EC; ++e) {
          const int j = q + LM::P * (e0 + e);
          const Tw<T> wj = twiddle<true>(twl, r * j);
          v[e0 + e] = cmul(dif_sum<T, R>(xv[e], wl), wj);
        }
      }
      Cx<T>* tcol = A + u;
      wave_fft<T, M, true, R, LM::E>(v, tcol, LU, q, twl);
#pragma unroll
      for (int e = 0; e < LM::E; ++e) tcol[(q + LM::P * e) * LU] = v[e];
    }
    if (RC && fillCol && wave == 0) {
      // Redo column 0 (x = 0, computed above from the unfilled sticks) from
      // the filled copy. Wave 0 owns column 0; all its lane groups compute
      // the same line, so their LDS writes agree.
      Cx<T> v[LM::E];
#pragma unroll
      for (int e = 0; e < LM::E; ++e) {
        const int j = q + LM::P * e;
        Cx<T> xf[R];
#pragma unroll
        for (int l = 0; l < R; ++l) xf[l] = fill[j + M * l];
        v[e] = cmul(dif_sum<T, R>(xf, wl), twiddle<true>(twl, r * j));
      }
      wave_sync();
      wave_fft<T, M, true, R, LM::E>(v, A, LU, q, twl);
#pragma unroll
      for (int e = 0; e < LM::E; ++e) A[(q + LM::P * e) * LU] = v[e];
    }
  }
  __syncthreads();

  // x: row m (y = r + R m) holds column u's value at x = colX[u]; absent x
  // read the zero column. C2C: line k is row k. R2C: line k carries the
  // hermitian rows a = 2k and b = 2k + 1 as a + i b, and its result's real
  // and imaginary parts are the two real rows. Lines past the last (a partial
  // last wave) are computed from the last one and not stored.
  static_assert(!RC || M % 2 == 0, "row pairs");
  constexpr int MX = RC ? M / 2 : M;
  const int k = wave * LX::PER_WAVE + lane / LX::P;
  const int kc = k < MX ? k : MX - 1;
  const int ma = RC ? 2 * kc : kc;
  const int q = lane % LX::P;
  Cx<T> v[LX::E];
#pragma unroll
  for (int e = 0; e < LX::E; ++e) {
    Cx<T> x = A[ma * LU + uu[e]];
    if (RC) {
      // C2R: the imaginary parts of the DC and Nyquist inputs do not
      // contribute to a real signal; the upper half is conjugated
      Cx<T> xb = A[(ma + 1) * LU + uu[e]];
      const int xx = q + LX::P * e;
      if (xx == 0 || 2 * xx == N) {
        x.im = T(0);
        xb.im = T(0);
      }
      if (2 * xx > N) {
        x = cconj(x);
        xb = cconj(xb);
      }
      x = Cx<T>{x.re - xb.im, x.im + xb.re};
    }
    v[e] = x;
  }
  __syncthreads();  // A is dead from here on; its space holds the row lines
  if (wave * LX::PER_WAVE < MX) {
    Cx<T>* line = A + k * LB;
    wave_fft<T, N, true, 1, LX::E>(v, line, 1, q, twl);
    if (k < MX) {
      const size_t row = (static_cast<size_t>(z) * N + r + R * ma) * N;
      if (RC) {
        // row b = row a + R rows
#pragma unroll
        for (int e = 0; e < LX::E; ++e) {
          a.out[row + q + LX::P * e] = v[e].re;
          a.out[row + static_cast<size_t>(R) * N + q + LX::P * e] = v[e].im;
        }
      } else {
        Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out) + row;
#pragma unroll
        for (int e = 0; e < LX::E; ++e) out[q + LX::P * e] = v[e];
      }
    }
  }
}

template <typename T, int N, bool RC>
__device__ __forceinline__ void xy_forward_body(const FastArgs<T>& a, int bid) {
  using S = XYShape<T, N>;
  using LM = typename S::YL;  // x lines of MX points, half-width like the backward y lines
  using LY = typename S::FL;
  constexpr int R = S::R, MX = S::M, LT = S::LT, NW = S::NW;
  Cx<T>* Tb = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  int z, c;
  plane_class<R>(bid, z, c);
  z += a.zBegin;
  if (z >= a.nz) return;
  const Cx<T>* in = reinterpret_cast<const Cx<T>*>(a.in) + static_cast<size_t>(z) * N * N;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  // twiddles are read from the table in global memory (L1/L2 resident): the
  // tile leaves no room for a copy at two workgroups per CU
  const Tw<T>* twl = a.twX;

  // x: every row, decimation in frequency to the columns x = c + R m, then
  // the MX-point FFT; row y's result lands in Tb[y][m]. Class twiddles are
  // loaded once; the row loads of a lane are issued together. When a lane's
  // loads for one row block fit one batch, the next block's loads are issued
  // before the current block's FFT (whose twiddles come from LDS), so they
  // overlap it.
  {
    Tw<T>* tw32 = reinterpret_cast<Tw<T>*>(Tb + S::TW_FWD);
    for (int t = threadIdx.x; t < MX; t += S::NT) tw32[t] = twl[t * R];
    const int q = lane % LM::P;
    Tw<T> wl[R];
    wl[0] = tw_one<T>();
#pragma unroll
    for (int l = 1; l < R; ++l) wl[l] = twiddle<false>(twl, ((c * l) % R) * MX);
    Tw<T> wj[LM::E];
#pragma unroll
    for (int e = 0; e < LM::E; ++e) wj[e] = twiddle<false>(twl, c * (q + LM::P * e));
    lds_barrier();  // tw32
    // at most 16 loads per lane in flight
    constexpr int EC = (LM::E % 4 == 0 && 4 * R <= 16) ? 4 : (LM::E % 2 == 0 ? 2 : LM::E);
    constexpr int STRIDE = NW * LM::PER_WAVE;  // rows per round of the workgroup
    auto load = [&](int y0, Cx<T> (&xv)[EC][R], int e0) {
      const int y = y0 + lane / LM::P;
      const int yc = y < N ? y : N - 1;
      if (RC) {
        const T* rrow = a.in + (static_cast<size_t>(z) * N + yc) * N;
#pragma unroll
        for (int e = 0; e < EC; ++e) {
#pragma unroll
          for (int l = 0; l < R; ++l) xv[e][l] = Cx<T>{rrow[q + LM::P * (e0 + e) + MX * l], T(0)};
        }
      } else {
        const Cx<T>* row = in + static_cast<size_t>(yc) * N;
#pragma unroll
        for (int e = 0; e < EC; ++e) {
#pragma unroll
          for (int l = 0; l < R; ++l) xv[e][l] = row[q + LM::P * (e0 + e) + MX * l];
        }
      }
    };
    auto dif = [&](Cx<T> (&xv)[EC][R], Cx<T>* v, int e0) {
#pragma unroll
      for (int e = 0; e < EC; ++e) v[e0 + e] = cmul(dif_sum<T, R>(xv[e], wl), wj[e0 + e]);
    };
    auto fft_store = [&](int y0, Cx<T>* v) {
      Cx<T>* line = Tb + (y0 + lane / LM::P) * LT;
      wave_fft<T, MX, false, 1, LM::E>(v, line, 1, q, tw32);
#pragma unroll
      for (int e = 0; e < LM::E; ++e) line[phys<T>(q + LM::P * e)] = v[e];
    };
    int y0 = wave * LM::PER_WAVE;
    if constexpr (EC == LM::E) {
      // one load batch per block: software-pipelined over the wave's blocks
      if (y0 < N) {
        Cx<T> xv[EC][R], v[LM::E];
        load(y0, xv, 0);
        dif(xv, v, 0);
        for (int y1 = y0 + STRIDE; y1 < N; y1 += STRIDE) {
          load(y1, xv, 0);
          fft_store(y0, v);
          dif(xv, v, 0);
          y0 = y1;
        }
        fft_store(y0, v);
      }
    } else {
      for (; y0 < N; y0 += STRIDE) {
        Cx<T> v[LM::E];
#pragma unroll
        for (int e0 = 0; e0 < LM::E; e0 += EC) {
          Cx<T> xv[EC][R];
          load(y0, xv, e0);
          dif(xv, v, e0);
        }
        fft_store(y0, v);
      }
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
      wave_fft<T, N, false, 1, LY::E>(v, col, LT, q, twl);
      if (m < MX) {
#pragma unroll
        for (int e = 0; e < LY::E; ++e) col[(q + LY::P * e) * LT] = v[e];
      }
    }
  }
  __syncthreads();
  const int i0 = a.resBegin[c], i1 = a.resBegin[c + 1];
  Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.pitch;
  for (int i = i0 + threadIdx.x; i < i1; i += S::NT) {
    const int s = a.resSticks[i];
    const int ym = a.resYM[i];  // y << 16 | column m within the class
    zrow[s] = Tb[(ym >> 16) * LT + phys<T>(ym & 0xffff)];
  }
}

// ------------------------------------------------------ whole-plane kernels
//
// One workgroup per plane for batches of transforms: the plane tile
// T[y][column] (columns of the sparse set plus a zero column) lives in LDS,
// the y transforms run in column mode on it and the x transforms on rows, so
// every input element is read from memory once (the class-split kernels read
// a plane R times). Backward: each y line gathers its sticks straight into
// registers through the stick index table and is written to the tile after
// its transform; forward: each y line's results go straight to the sticks.
// x lines have 4 elements per lane and exchange through a per-wave scratch
// region after the tile; y lines have 8.

template <typename T, int N>
struct PlaneShape {
  // powers of two: x lines of 4 elements, y lines of 8; 96: 12 and 12
  static constexpr bool F3 = N % 3 == 0;
  static constexpr int NT = F3 ? 768 : (N >= 128 ? 1024 : (N >= 64 ? 512 : 256));
  static constexpr int NW = NT / kWave;
  using XL = Line<N, F3 ? 12 : 4>;
  using YL = Line<N, F3 ? 12 : 8>;
  static_assert(!XL::SPARE && !YL::SPARE, "whole lines per wave");
  static constexpr int ROWS = padded_len<T>(N);  // the column exchange reaches padded rows
  static constexpr int LB = padded_len<T>(N);    // x line in the scratch region
  static constexpr int XROWS = NW * XL::PER_WAVE;
  static_assert(N <= NT, "one twiddle per thread");
  static_assert(N % XROWS == 0, "whole x iterations");
  // odd pitch: the 16 lanes of an LDS cycle in column mode (16 rows of one
  // column) land on distinct bank groups
  static int pitch(int ncol) { return round_up(ncol + 1, YL::PER_WAVE) | 1; }
  static size_t lds(int ncol) {
    return static_cast<size_t>(tw_offset<T>(ROWS * pitch(ncol) + XROWS * LB)) * sizeof(Cx<T>) +
           static_cast<size_t>(N) * sizeof(Tw<T>);
  }
};

// Stick indices of the y line of column u held by lane q (y = q + P e): the
// table is class-interleaved for R = 4, M = N / 4 (index of y = j + M l at
// (u M + j) R + l), so lane q reads M / P 16-byte groups g, each holding
// e = g + (M / P) l for l < 4.
template <int N, typename YL>
__device__ __forceinline__ void plane_stick_indices(const int* table, int u, int q, int* idx) {
  constexpr int M = N / 4, G = M / YL::P;
  static_assert(fast_classes(N) == 4 && M % YL::P == 0 && YL::E == 4 * G, "index layout");
  const int4* rel = reinterpret_cast<const int4*>(table) + static_cast<size_t>(u) * M;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int4 x = at32(rel, q + YL::P * g);
    idx[g] = x.x;
    idx[g + G] = x.y;
    idx[g + 2 * G] = x.z;
    idx[g + 3 * G] = x.w;
  }
}

template <typename T, int N>
__device__ __forceinline__ void plane_backward_body(const FastArgs<T>& a, int z) {
  using S = PlaneShape<T, N>;
  using XL = typename S::XL;
  using YL = typename S::YL;
  constexpr int NW = S::NW, LB = S::LB;
  const int PU = a.pitchT, ncol = a.ncol;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Cx<T>* xs = tile + S::ROWS * PU;
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(tile + tw_offset<T>(S::ROWS * PU + S::XROWS * LB));
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (threadIdx.x < N) twl[threadIdx.x] = a.twX[threadIdx.x];
  // column of each x this lane handles in the x pass (ncol = zero column)
  int uu[XL::E];
#pragma unroll
  for (int e = 0; e < XL::E; ++e) uu[e] = a.xCol[lane % XL::P + XL::P * e];
  const Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.pitch;
  // y: every column up to and including the zero column ncol, gathered
  // straight into registers through the stick index table, transformed in
  // column mode and written to the tile. The next column group's gathers
  // are issued before the current group's transform (software pipelining).
  {
    const int q = lane % YL::P;
    auto gather = [&](int g0, Cx<T>* dst) {
      const int u = g0 + lane / YL::P;
      int idx[YL::E];
      plane_stick_indices<N, YL>(a.stickAt, u < ncol ? u : ncol, q, idx);
#pragma unroll
      for (int e = 0; e < YL::E; ++e) dst[e] = at32(zrow, idx[e]);
    };
    int u0 = wave * YL::PER_WAVE;
    Cx<T> nv[YL::E];
    if (u0 <= ncol) gather(u0, nv);
    __syncthreads();  // twiddles
    for (; u0 <= ncol; u0 += NW * YL::PER_WAVE) {
      Cx<T> v[YL::E];
#pragma unroll
      for (int e = 0; e < YL::E; ++e) v[e] = nv[e];
      if (u0 + NW * YL::PER_WAVE <= ncol) gather(u0 + NW * YL::PER_WAVE, nv);
      Cx<T>* col = tile + u0 + lane / YL::P;
      wave_fft<T, N, true, 1, YL::E>(v, col, PU, q, twl);
#pragma unroll
      for (int e = 0; e < YL::E; ++e) col[(q + YL::P * e) * PU] = v[e];
    }
  }
  __syncthreads();
  // x: rows, XROWS at a time, each lane group exchanging through its own line
  const int q = lane % XL::P;
  const int lr = wave * XL::PER_WAVE + lane / XL::P;
  Cx<T>* line = xs + lr * LB;
  Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out) + static_cast<size_t>(z) * N * N;
  for (int y0 = 0; y0 < N; y0 += S::XROWS) {
    const int y = y0 + lr;
    Cx<T> v[XL::E];
#pragma unroll
    for (int e = 0; e < XL::E; ++e) v[e] = tile[y * PU + uu[e]];
    wave_fft<T, N, true, 1, XL::E>(v, line, 1, q, twl);
#pragma unroll
    for (int e = 0; e < XL::E; ++e) out[static_cast<size_t>(y) * N + q + XL::P * e] = v[e];
    wave_sync();
  }
}

template <typename T, int N>
__device__ __forceinline__ void plane_forward_body(const FastArgs<T>& a, int z) {
  using S = PlaneShape<T, N>;
  using XL = typename S::XL;
  using YL = typename S::YL;
  constexpr int NW = S::NW, LB = S::LB;
  const int PU = a.pitchT, ncol = a.ncol;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Cx<T>* xs = tile + S::ROWS * PU;
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(tile + tw_offset<T>(S::ROWS * PU + S::XROWS * LB));
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (threadIdx.x < N) twl[threadIdx.x] = a.twX[threadIdx.x];
  int uu[XL::E];
#pragma unroll
  for (int e = 0; e < XL::E; ++e) uu[e] = a.xCol[lane % XL::P + XL::P * e];
  // x: every row, read once; the sparse set's columns go into the tile. The
  // next rows' loads are issued before the current rows' transform.
  {
    const int q = lane % XL::P;
    const int lr = wave * XL::PER_WAVE + lane / XL::P;
    Cx<T>* line = xs + lr * LB;
    const Cx<T>* in = reinterpret_cast<const Cx<T>*>(a.in) + static_cast<size_t>(z) * N * N;
    Cx<T> nv[XL::E];
#pragma unroll
    for (int e = 0; e < XL::E; ++e) nv[e] = in[static_cast<size_t>(lr) * N + q + XL::P * e];
    __syncthreads();  // twiddles
    for (int y0 = 0; y0 < N; y0 += S::XROWS) {
      const int y = y0 + lr;
      Cx<T> v[XL::E];
#pragma unroll
      for (int e = 0; e < XL::E; ++e) v[e] = nv[e];
      if (y0 + S::XROWS < N) {
#pragma unroll
        for (int e = 0; e < XL::E; ++e) {
          nv[e] = in[static_cast<size_t>(y + S::XROWS) * N + q + XL::P * e];
        }
      }
      wave_fft<T, N, false, 1, XL::E>(v, line, 1, q, twl);
#pragma unroll
      for (int e = 0; e < XL::E; ++e) {
        if (uu[e] < ncol) tile[y * PU + uu[e]] = v[e];
      }
      wave_sync();
    }
  }
  __syncthreads();
  // y: the sparse set's columns in column mode
  {
    // results go straight to the sticks of the column (stick index table as
    // in the backward kernel; absent positions point at nst and are skipped)
    Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.pitch;
    const int q = lane % YL::P;
    for (int u0 = wave * YL::PER_WAVE; u0 < ncol; u0 += NW * YL::PER_WAVE) {
      const int u = u0 + lane / YL::P;
      const int uc = u < ncol ? u : ncol;
      int idx[YL::E];
      plane_stick_indices<N, YL>(a.stickAt, uc, q, idx);
      Cx<T>* col = tile + u;
      Cx<T> v[YL::E];
#pragma unroll
      for (int e = 0; e < YL::E; ++e) v[e] = col[(q + YL::P * e) * PU];
      wave_sync();
      wave_fft<T, N, false, 1, YL::E>(v, col, PU, q, twl);
#pragma unroll
      for (int e = 0; e < YL::E; ++e) {
        if (idx[e] < a.nst) zrow[idx[e]] = v[e];
      }
    }
  }
}

// ------------------------------------------------------ split plane stage
//
// Two kernels per direction around the compact column buffer
// ybuf[z][y][u] (row pitch pitchY >= ncol, u the sparse set's column):
//   backward: columns (each line gathers its column's sticks through the
//             stick index table, N-point y FFT; the workgroup's LINES
//             columns leave through an LDS tile as row segments), then rows
//             (each line places the compact row at its x positions, N-point
//             x FFT, dense row stored);
//   forward : rows (x FFT, the sparse set's columns stored compact), then
//             columns (LINES columns enter through the LDS tile as row
//             segments, y FFT, results written to the sticks).
// Lines are FullLine<T, N> (P lanes x E elements); LINES = lines per
// workgroup. No plane is read more than once. R2C: each x line carries two
// rows (a + i b). Where P does not divide the wave, its spare lanes run a
// dummy line in an exchange region of their own and store nothing.

// Waves of a split-stage workgroup: up to 4, never more lines than the plane
// has rows (columns), and the exchange regions within 64 KiB.
template <typename T, int N, typename L>
constexpr int split_waves() {
  int nw = N / L::PER_WAVE < 4 ? N / L::PER_WAVE : 4;
  constexpr int regions = L::PER_WAVE + (L::SPARE ? 1 : 0);  // exchange regions per wave
  while (nw > 1 && static_cast<size_t>(nw) * regions * padded_len<T>(N) * sizeof(Cx<T>) >
                       64 * 1024) {
    nw /= 2;
  }
  return nw;
}

template <typename T, int N>
struct SplitShape {
  using L = FullLine<T, N>;
  static_assert(schedule_ok(N, L::E), "full line schedule");
  static constexpr int NW = split_waves<T, N, L>();
  static constexpr int NT = NW * kWave;
  static constexpr int LINES = NW * L::PER_WAVE;
  static_assert(NW >= 1 && LINES <= N, "split line shape");
  static constexpr int LB = padded_len<T>(N);  // exchange region of one line
  static constexpr int TP = LINES | 1;         // odd pitch of the [y][LINES] tile
  // exchange regions: the lines', then one dummy line per wave for its spare
  // lanes (P not dividing the wave)
  static constexpr int XL = LINES + (L::SPARE ? NW : 0);
  // the exchange regions, reused as the tile between phases
  static constexpr int REGION = XL * LB > N * TP ? XL * LB : N * TP;
  // This lane's line in the workgroup (LINES for a spare lane: none) and the
  // exchange region it runs in.
  __device__ static int line_of(int wave, int lane) {
    return L::spare(lane) ? LINES : wave * L::PER_WAVE + lane / L::P;
  }
  template <typename C>
  __device__ static C* xregion(C* region, int wave, int lane) {
    return region + (L::spare(lane) ? LINES + wave : wave * L::PER_WAVE + lane / L::P) * LB;
  }
  // rows per workgroup of the row kernels (rows of length N): R2C packs two
  // real rows into each complex line
  __host__ __device__ static constexpr int row_block(bool rc) { return rc ? 2 * LINES : LINES; }
  // workgroups per plane of the row kernels (ny rows)
  __host__ __device__ static int row_groups(int ny, bool rc) {
    return (ny + row_block(rc) - 1) / row_block(rc);
  }
  // region, twiddles, the filled x = 0 column (R2C backward)
  static constexpr size_t lds() {
    return static_cast<size_t>(tw_offset<T>(REGION)) * sizeof(Cx<T>) + N * sizeof(Tw<T>) +
           N * sizeof(Cx<T>);
  }
  __host__ __device__ static int col_groups(int ncol) { return (ncol + LINES - 1) / LINES; }
};

// Stick index of (column u, y) in the class-interleaved table of the xy
// kernels (y = j + M l at (u M + j) R + l).
template <int N>
__device__ __forceinline__ int stick_at(const int* table, int u, int y) {
  constexpr int R = fast_classes(N), M = N / R;
  return table[(u * M + y % M) * R + y / M];
}

template <typename T, int N, bool RC>
__device__ __forceinline__ void ysplit_backward_body(const FastArgs<T>& a, int bid) {
  using S = SplitShape<T, N>;
  using L = typename S::L;
  constexpr int E = L::E, P = L::P;
  const int groups = S::col_groups(a.ncol);
  const int z = bid / groups + a.zBegin;
  const int u0 = (bid % groups) * S::LINES;
  if (z >= a.nz) return;
  Cx<T>* region = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(region + tw_offset<T>(S::REGION));
  Cx<T>* fill = reinterpret_cast<Cx<T>*>(twl + N);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int q = lane % P;
  const int line = S::line_of(wave, lane);
  const int u = u0 + line;
  // row ncol of the table: the zero element (also for spare lanes)
  const int uc = line < S::LINES && u < a.ncol ? u : a.ncol;
  const Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.pitch;
  Cx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = at32(zrow, stick_at<N>(a.stickAt, uc, q + P * e));
  for (int t = threadIdx.x; t < N; t += S::NT) twl[t] = a.twY[t];
  if (RC && u0 == 0 && a.colX[0] == 0) {
    // hermitian fill of the x = 0 column (reference symmetry_host.hpp:47-58)
    if (line == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) fill[q + P * e] = v[e];
    }
    __syncthreads();
    for (int k = 1 + threadIdx.x; 2 * k <= N; k += S::NT) {
      Cx<T>& lo = fill[k];
      Cx<T>& hi = fill[N - k];
      if (2 * k == N) {
        if (cnonzero(lo)) lo = cconj(lo);
      } else if (cnonzero(lo)) {
        hi = cconj(lo);
      } else if (cnonzero(hi)) {
        lo = cconj(hi);
      }
    }
    __syncthreads();
    if (line == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = fill[q + P * e];
    }
  } else {
    __syncthreads();  // twiddles
  }
  Cx<T>* lr = S::xregion(region, wave, lane);
  wave_fft<T, N, true, 1, E>(v, lr, 1, q, twl);
  __syncthreads();  // the exchange regions become the tile
  if (line < S::LINES) {
#pragma unroll
    for (int e = 0; e < E; ++e) region[(q + P * e) * S::TP + line] = v[e];
  }
  __syncthreads();
  Cx<T>* yrows = a.ybuf + static_cast<size_t>(z) * N * a.pitchY + u0;
  const int nc = min(S::LINES, a.ncol - u0);
  for (int t = threadIdx.x; t < N * S::LINES; t += S::NT) {
    const int y = t / S::LINES, c = t - y * S::LINES;
    if (c < nc) yrows[static_cast<size_t>(y) * a.pitchY + c] = region[y * S::TP + c];
  }
}

template <typename T, int N, bool RC>
__device__ __forceinline__ void xsplit_backward_body(const FastArgs<T>& a, int bid) {
  using S = SplitShape<T, N>;
  using L = typename S::L;
  constexpr int E = L::E, P = L::P;
  const int ny = a.ny;  // rows of the plane (N is the row length)
  const int groups = S::row_groups(ny, RC);
  const int z = bid / groups + a.zBegin;
  const int y0 = (bid % groups) * S::row_block(RC);
  if (z >= a.nz) return;
  Cx<T>* region = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(region + tw_offset<T>(S::REGION));
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int q = lane % P;
  const int line = S::line_of(wave, lane);
  // C2C: row y0 + line. R2C: rows y0 + 2 line (real part of the line's
  // result) and y0 + 2 line + 1 (imaginary part), which an odd ny leaves
  // out of the last pair (its input is zero, its output not stored).
  const int y = line < S::LINES ? y0 + (RC ? 2 * line : line) : ny;  // spare lanes: no row
  const int yc = y < ny ? y : ny - 1 - (RC ? 1 : 0);
  const bool hasB = yc + 1 < ny;
  const int ncol = a.ncol;
  const Cx<T>* yrow = a.ybuf + (static_cast<size_t>(z) * ny + yc) * a.pitchY;
  const size_t offB = RC && hasB ? a.pitchY : 0;
  // compact row at its x positions; absent x read column 0 and are zeroed;
  // R2C: the upper half is the conjugate of x' = N - x, and the hermitian rows
  // a and b enter as a + i b (two real outputs from one complex transform)
  Cx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = q + P * e;
    const int u = a.xCol[(RC && 2 * x > N) ? N - x : x];
    const int uc = u < ncol ? u : 0;
    Cx<T> w = yrow[uc];
    if (u >= ncol) w = Cx<T>{T(0), T(0)};
    if (RC) {
      Cx<T> wb = yrow[offB + uc];
      if (u >= ncol || !hasB) wb = Cx<T>{T(0), T(0)};
      if (x == 0 || 2 * x == N) {
        w.im = T(0);
        wb.im = T(0);
      }
      if (2 * x > N) {
        w = cconj(w);
        wb = cconj(wb);
      }
      w = Cx<T>{w.re - wb.im, w.im + wb.re};
    }
    v[e] = w;
  }
  for (int t = threadIdx.x; t < N; t += S::NT) twl[t] = a.twX[t];
  __syncthreads();
  wave_fft<T, N, true, 1, E>(v, S::xregion(region, wave, lane), 1, q, twl);
  if (y < ny) {
    const size_t row = (static_cast<size_t>(z) * ny + y) * N;
    if (RC) {
#pragma unroll
      for (int e = 0; e < E; ++e) a.out[row + q + P * e] = v[e].re;
      if (hasB) {
#pragma unroll
        for (int e = 0; e < E; ++e) a.out[row + N + q + P * e] = v[e].im;
      }
    } else {
      Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out) + row;
#pragma unroll
      for (int e = 0; e < E; ++e) out[q + P * e] = v[e];
    }
  }
}

template <typename T, int N, bool RC>
__device__ __forceinline__ void xsplit_forward_body(const FastArgs<T>& a, int bid) {
  using S = SplitShape<T, N>;
  using L = typename S::L;
  constexpr int E = L::E, P = L::P;
  const int ny = a.ny;  // rows of the plane (N is the row length)
  const int groups = S::row_groups(ny, RC);
  const int z = bid / groups + a.zBegin;
  const int y0 = (bid % groups) * S::row_block(RC);
  if (z >= a.nz) return;
  Cx<T>* region = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(region + tw_offset<T>(S::REGION));
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int q = lane % P;
  const int line = S::line_of(wave, lane);
  // C2C: row y0 + line. R2C: the real rows a = y0 + 2 line and b = a + 1
  // enter as a + i b; an odd ny leaves b out of the last pair (zero input,
  // its spectrum not stored).
  const int y = line < S::LINES ? y0 + (RC ? 2 * line : line) : ny;  // spare lanes: no row
  const int yc = y < ny ? y : ny - 1 - (RC ? 1 : 0);
  const bool hasB = yc + 1 < ny;
  const size_t row = (static_cast<size_t>(z) * ny + yc) * N;
  Cx<T> v[E];
  if (RC) {
    const size_t rowB = hasB ? row + N : row;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const T b = a.in[rowB + q + P * e];
      v[e] = Cx<T>{a.in[row + q + P * e], hasB ? b : T(0)};
    }
  } else {
    const Cx<T>* in = reinterpret_cast<const Cx<T>*>(a.in) + row;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = in[q + P * e];
  }
  int uu[E];
#pragma unroll
  for (int e = 0; e < E; ++e) uu[e] = a.xCol[q + P * e];
  for (int t = threadIdx.x; t < N; t += S::NT) twl[t] = a.twX[t];
  __syncthreads();
  Cx<T>* lr = S::xregion(region, wave, lane);
  wave_fft<T, N, false, 1, E>(v, lr, 1, q, twl);
  if (RC) {
    // Z = A + i B with A, B the spectra of the real rows a and b:
    // A[x] = (Z[x] + conj Z[N - x]) / 2, B[x] = -i (Z[x] - conj Z[N - x]) / 2.
    // The partner of x is held by lane (P - q) % P; it comes through the
    // line's exchange region (wave-local).
    wave_sync();
#pragma unroll
    for (int e = 0; e < E; ++e) lr[phys<T>(q + P * e)] = v[e];
    wave_sync();
    if (y < ny) {
      Cx<T>* ya = a.ybuf + (static_cast<size_t>(z) * ny + y) * a.pitchY;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int x = q + P * e;
        if (uu[e] < a.ncol) {
          const Cx<T> zp = lr[phys<T>(x == 0 ? 0 : N - x)];
          const T h = T(0.5);
          ya[uu[e]] = Cx<T>{h * (v[e].re + zp.re), h * (v[e].im - zp.im)};
          if (hasB) ya[a.pitchY + uu[e]] = Cx<T>{h * (v[e].im + zp.im), h * (zp.re - v[e].re)};
        }
      }
    }
  } else if (y < ny) {
    Cx<T>* yrow = a.ybuf + (static_cast<size_t>(z) * ny + y) * a.pitchY;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (uu[e] < a.ncol) yrow[uu[e]] = v[e];
    }
  }
}

template <typename T, int N>
__device__ __forceinline__ void ysplit_forward_body(const FastArgs<T>& a, int bid) {
  using S = SplitShape<T, N>;
  using L = typename S::L;
  constexpr int E = L::E, P = L::P;
  const int groups = S::col_groups(a.ncol);
  const int z = bid / groups + a.zBegin;
  const int u0 = (bid % groups) * S::LINES;
  if (z >= a.nz) return;
  Cx<T>* region = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(region + tw_offset<T>(S::REGION));
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int q = lane % P;
  const int line = S::line_of(wave, lane);
  const int u = line < S::LINES ? u0 + line : a.ncol;  // spare lanes: no column
  const int nc = min(S::LINES, a.ncol - u0);
  // LINES columns of every row through the tile (row segments); columns past
  // ncol are zero. Thread t holds column t % LINES of rows t / LINES + P k,
  // k < E (N LINES = E NT), all loads in flight before the tile stores.
  const Cx<T>* yrows = a.ybuf + static_cast<size_t>(z) * N * a.pitchY + u0;
  if constexpr (N * S::LINES != E * S::NT || S::NT % S::LINES != 0) {
    // spare lanes: no whole-tile thread mapping
    for (int t = threadIdx.x; t < N * S::LINES; t += S::NT) {
      const int y = t / S::LINES, c = t - y * S::LINES;
      region[y * S::TP + c] =
          c < nc ? yrows[static_cast<size_t>(y) * a.pitchY + c] : Cx<T>{T(0), T(0)};
    }
  } else {
    constexpr int RS = S::NT / S::LINES;  // rows per pass of the workgroup
    const int c = threadIdx.x % S::LINES, y0 = threadIdx.x / S::LINES;
    const int cc = c < nc ? c : 0;
    Cx<T> w[E];
#pragma unroll
    for (int k = 0; k < E; ++k) w[k] = yrows[static_cast<size_t>(y0 + RS * k) * a.pitchY + cc];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      region[(y0 + RS * k) * S::TP + c] = c < nc ? w[k] : Cx<T>{T(0), T(0)};
    }
  }
  for (int t = threadIdx.x; t < N; t += S::NT) twl[t] = a.twY[t];
  __syncthreads();
  Cx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = region[(q + P * e) * S::TP + (line < S::LINES ? line : 0)];
  __syncthreads();  // the tile becomes the exchange regions
  wave_fft<T, N, false, 1, E>(v, S::xregion(region, wave, lane), 1, q, twl);
  if (u < a.ncol) {
    Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.pitch;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int s = stick_at<N>(a.stickAt, u, q + P * e);
      if (s < a.nst) zrow[s] = v[e];
    }
  }
}

// ------------------------------------------------- full-tile plane kernels
//
// One workgroup per plane with the whole plane in LDS: tile row y holds the
// plane's columns at their compact positions u < ncol, and, with row pitch
// PU > ROWS, room for the x line of row y, which runs in place (no exchange
// scratch). Lines have P lanes and E = N / P elements per lane (64: 8 x 8,
// 96: 4 x 24, 128: 8 x 16), so every y and x line is one Stockham exchange
// after an in-register first pass, and all lines of a phase run at once (one
// line per lane group). Exchange slots have one spare per E elements
// (slot<E>, ROWS slots per line), and the lines sharing an LDS cycle are 8
// (or 4) columns or rows apart, so they use disjoint banks. Element ROWS of
// each row is zero: absent x positions read it.
//
// Each workgroup takes planes blockIdx.x, blockIdx.x + gridDim.x, ... The
// FP32 launches are persistent (as many workgroups as fit on the GPU at once;
// one per CU at 96 and 128), and the loads of a workgroup's next
// plane (backward: its y lines' gathers, forward: its rows) are issued during
// the current plane's second phase. The FP64 launches have one workgroup per
// plane. Barriers are LDS-only, so loads in flight cross them.

template <typename T, int N>
struct FullShape {
  // lanes x elements per line: 64: 8 x 8, 96: 4 x 24, 128: 8 x 16
  static constexpr int E = N == 96 ? 24 : N / 8;
  using L = Line<N, E>;
  static constexpr int P = L::P;
  static_assert(P == 8 || P == 4, "lines of 4 or 8 lanes");
  static constexpr int PE = E;                  // one spare slot per E elements
  static constexpr int ROWS = N + (N - 1) / E;  // exchange slots of a line
  static constexpr int PU = (ROWS + 1) | 1;     // odd row pitch, zero element at ROWS
  static constexpr int NW = N / L::PER_WAVE;    // one line per row (column)
  static constexpr int NT = NW * kWave;
  // persistent grid (FP32), or one plane per workgroup (FP64, where staggered
  // workgroup starts measured faster than the prefetching plane loop)
  static constexpr bool PERSISTENT = sizeof(T) == 4;
  // tile, twiddle table, x -> compact column map of the current plane
  static constexpr size_t lds() {
    return static_cast<size_t>(tw_offset<T>(ROWS * PU)) * sizeof(Cx<T>) + N * sizeof(Tw<T>) +
           N * sizeof(int);
  }
};

// The tile offset of column `line`, opaque to the compiler: column addresses
// reach past the 64 KiB immediate offset of LDS instructions, and without
// this the loop-invariant part of every one of them would be hoisted out of
// the plane loop into a register of its own.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Row (column) of lane group g of wave w. An LDS cycle serves 16 lanes: with
// 8 lanes per line its two lines, g = 2k and 2k + 1, are 8 apart; with 4
// lanes per line its four lines, g = 4k .. 4k + 3, are 4 apart.
template <int P>
__device__ __forceinline__ int full_line(int w, int g) {
  if constexpr (P == 8) {
    return 16 * (w >> 1) + 4 * (w & 1) + (g >> 1) + 8 * (g & 1);
  } else {
    return 16 * w + 4 * (g & 3) + (g >> 2);
  }
}

// Values of y line `line` of plane z (a column of the sparse set), gathered
// from the [z][stick] buffer through its stick indices.
template <typename T, typename L>
__device__ __forceinline__ void full_gather(const FastArgs<T>& a, int z, const int* idx,
                                            Cx<T>* v) {
  const Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.pitch;
#pragma unroll
  for (int e = 0; e < L::E; ++e) v[e] = at32(zrow, idx[e]);
}

// Values of x line `line` (a row) of plane z of the input grid.
template <typename T, int N, typename L>
__device__ __forceinline__ void full_rows(const FastArgs<T>& a, int z, int line, int q,
                                          Cx<T>* v) {
  const Cx<T>* in =
      reinterpret_cast<const Cx<T>*>(a.in) + (static_cast<size_t>(z) * N + line) * N;
#pragma unroll
  for (int e = 0; e < L::E; ++e) v[e] = in[q + L::P * e];
}

// Arguments read from a device table hold generic pointers, so the compiler
// would access them with flat instructions (which share the LDS counter and
// serialise against LDS traffic). Rebasing each pointer on a kernel argument
// (a global pointer) keeps the address and lets the accesses be global ones.
// A null pointer stays null (no offset is formed from it).
template <typename P>
__device__ __forceinline__ P* rebase(const char* anchor, P* p) {
  if (p == nullptr) return nullptr;
  return (P*)(anchor + (reinterpret_cast<const char*>(p) - anchor));
}
template <typename T>
__device__ __forceinline__ FastArgs<T> on_global(FastArgs<T> a, const void* table) {
  const char* g = static_cast<const char*>(table);
  a.zbuf = rebase(g, a.zbuf);
  a.in = rebase(g, a.in);
  a.out = rebase(g, a.out);
  a.entryValue = rebase(g, a.entryValue);
  a.entrySlot = rebase(g, a.entrySlot);
  a.stickBegin = rebase(g, a.stickBegin);
  a.stickY = rebase(g, a.stickY);
  a.stickCol = rebase(g, a.stickCol);
  a.colX = rebase(g, a.colX);
  a.colBegin = rebase(g, a.colBegin);
  a.stickAt = rebase(g, a.stickAt);
  a.zEntries = rebase(g, a.zEntries);
  a.zInfo = rebase(g, a.zInfo);
  a.xCol = rebase(g, a.xCol);
  a.resSticks = rebase(g, a.resSticks);
  a.resBegin = rebase(g, a.resBegin);
  a.resYM = rebase(g, a.resYM);
  a.ybuf = rebase(g, a.ybuf);
  a.twX = rebase(g, a.twX);
  a.twY = rebase(g, a.twY);
  a.twZ = rebase(g, a.twZ);
  return a;
}

// ---- entry points: one transform (arguments by value) or a batch of
// transforms of one shape (argument table in device memory). A batched z
// launch maps each workgroup to (transform, block) through `map`; a batched
// xy launch gives every transform the same number of workgroups.

template <typename T, int NZ, bool RC>
__global__ void __launch_bounds__((ZShape<T, NZ>::NT)) k_z_backward_fast(FastArgs<T> a) {
  z_backward_body<T, NZ, RC, false>(a, blockIdx.x);
}
template <typename T, int NZ>
__global__ void __launch_bounds__((ZShape<T, NZ>::NT)) k_z_forward_fast(FastArgs<T> a) {
  z_forward_body<T, NZ>(a, blockIdx.x);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((XYShape<T, N>::NT)) k_xy_backward_fast(FastArgs<T> a) {
  xy_backward_body<T, N, RC>(a, blockIdx.x);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((XYShape<T, N>::NT)) k_xy_forward_fast(FastArgs<T> a) {
  xy_forward_body<T, N, RC>(a, blockIdx.x);
}
// useRuns = 0: the entry-table addressing even for sets made of stick runs
template <typename T, int NZ, bool RC>
__global__ void __launch_bounds__((ZShape<T, NZ>::NT))
    k_z_backward_batch(const FastArgs<T>* __restrict__ batch, const int2* __restrict__ map,
                       int useRuns) {
  const int2 m = map[blockIdx.x];
  FastArgs<T> a = on_global(batch[m.x], batch);
  a.hasRuns = a.hasRuns && useRuns;
  z_backward_body<T, NZ, RC, true>(a, m.y);
}
template <typename T, int NZ>
__global__ void __launch_bounds__((ZShape<T, NZ>::NT))
    k_z_forward_batch(const FastArgs<T>* __restrict__ batch, const int2* __restrict__ map,
                      int useRuns) {
  const int2 m = map[blockIdx.x];
  FastArgs<T> a = on_global(batch[m.x], batch);
  a.hasRuns = a.hasRuns && useRuns;
  z_forward_body<T, NZ>(a, m.y);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((XYShape<T, N>::NT))
    k_xy_backward_batch(const FastArgs<T>* __restrict__ batch, int perTransform) {
  const int t = blockIdx.x / perTransform;
  xy_backward_body<T, N, RC>(on_global(batch[t], batch), blockIdx.x - t * perTransform);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((XYShape<T, N>::NT))
    k_xy_forward_batch(const FastArgs<T>* __restrict__ batch, int perTransform) {
  const int t = blockIdx.x / perTransform;
  xy_forward_body<T, N, RC>(on_global(batch[t], batch), blockIdx.x - t * perTransform);
}

template <typename T, int N, bool RC>
__global__ void __launch_bounds__((SplitShape<T, N>::NT)) k_ysplit_backward(FastArgs<T> a) {
  ysplit_backward_body<T, N, RC>(a, blockIdx.x);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((SplitShape<T, N>::NT)) k_xsplit_backward(FastArgs<T> a) {
  xsplit_backward_body<T, N, RC>(a, blockIdx.x);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((SplitShape<T, N>::NT)) k_xsplit_forward(FastArgs<T> a) {
  xsplit_forward_body<T, N, RC>(a, blockIdx.x);
}
template <typename T, int N>
__global__ void __launch_bounds__((SplitShape<T, N>::NT)) k_ysplit_forward(FastArgs<T> a) {
  ysplit_forward_body<T, N>(a, blockIdx.x);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((SplitShape<T, N>::NT))
    k_ysplit_backward_batch(const FastArgs<T>* __restrict__ batch, int perTransform) {
  const int t = blockIdx.x / perTransform;
  ysplit_backward_body<T, N, RC>(on_global(batch[t], batch), blockIdx.x - t * perTransform);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((SplitShape<T, N>::NT))
    k_xsplit_backward_batch(const FastArgs<T>* __restrict__ batch, int perTransform) {
  const int t = blockIdx.x / perTransform;
  xsplit_backward_body<T, N, RC>(on_global(batch[t], batch), blockIdx.x - t * perTransform);
}
template <typename T, int N, bool RC>
__global__ void __launch_bounds__((SplitShape<T, N>::NT))
    k_xsplit_forward_batch(const FastArgs<T>* __restrict__ batch, int perTransform) {
  const int t = blockIdx.x / perTransform;
  xsplit_forward_body<T, N, RC>(on_global(batch[t], batch), blockIdx.x - t * perTransform);
}
template <typename T, int N>
__global__ void __launch_bounds__((SplitShape<T, N>::NT))
    k_ysplit_forward_batch(const FastArgs<T>* __restrict__ batch, int perTransform) {
  const int t = blockIdx.x / perTransform;
  ysplit_forward_body<T, N>(on_global(batch[t], batch), blockIdx.x - t * perTransform);
}

template <typename T, int N>
__global__ void __launch_bounds__((PlaneShape<T, N>::NT))
    k_plane_backward_batch(const FastArgs<T>* __restrict__ batch, int nz) {
  const int t = blockIdx.x / nz;
  plane_backward_body<T, N>(on_global(batch[t], batch), blockIdx.x - t * nz);
}
template <typename T, int N>
__global__ void __launch_bounds__((PlaneShape<T, N>::NT))
    k_plane_forward_batch(const FastArgs<T>* __restrict__ batch, int nz) {
  const int t = blockIdx.x / nz;
  plane_forward_body<T, N>(on_global(batch[t], batch), blockIdx.x - t * nz);
}
// Entry t of an argument table, read through the constant address space:
// the table is not written while a kernel runs, and loads inside a loop with
// global stores would otherwise be vector loads (whose wait also waits for
// every load issued before them).
template <typename T>
__device__ __forceinline__ FastArgs<T> table_args(const FastArgs<T>* batch, int t) {
  static_assert(sizeof(FastArgs<T>) % 4 == 0, "whole dwords");
  using ConstWords = const __attribute__((address_space(4))) unsigned*;
  // t is uniform; saying so keeps the loads scalar where the compiler cannot
  // prove it (inside the FFT hooks)
  const ConstWords src = (ConstWords)(batch + __builtin_amdgcn_readfirstlane(t));
  struct Words {
    unsigned w[sizeof(FastArgs<T>) / 4];
  } words;
#pragma unroll
  for (int i = 0; i < static_cast<int>(sizeof(FastArgs<T>) / 4); ++i) words.w[i] = src[i];
  return on_global(__builtin_bit_cast(FastArgs<T>, words), batch);
}

// A uniform value, loaded into scalar registers here: a table field used
// only under divergent control flow would otherwise be loaded there, with a
// vector load whose wait also waits for every load issued before it.
template <typename X>
__device__ __forceinline__ X pin(X x) {
  asm volatile("" : "+s"(x));
  return x;
}
// The same for a pointer into global memory: rebased on a kernel argument so
// that the accesses stay global (not flat) ones.
template <typename P>
__device__ __forceinline__ P* pin(const void* anchor, P* p) {
  return rebase(static_cast<const char*>(anchor), pin(p));
}

// Full-tile kernels: `total` = planes of the batch (transforms x nz), the
// grid at most that. Every transform of a batch has the same N, so the
// twiddle table is loaded once.
template <typename T, int N>
__global__ void __launch_bounds__((FullShape<T, N>::NT))
    k_full_backward_batch(const FastArgs<T>* __restrict__ batch, int nz, int total) {
  using S = FullShape<T, N>;
  using L = typename S::L;
  constexpr int E = L::E, P = L::P, PU = S::PU, ROWS = S::ROWS;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(tile + tw_offset<T>(ROWS * PU));
  int* xmap = reinterpret_cast<int*>(twl + N);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int q = lane % P;
  const int line = full_line<P>(wave, lane / P);
  const int line0 = full_line<P>(wave, 0);  // the wave's lowest line
  const int xi = static_cast<int>(threadIdx.x % N);  // x of this thread's map entry
  const int G = gridDim.x;
  // In flight at the top of the loop: plane p's y line values (v) and map
  // entry (xc), plane p + G's stick indices (idx). Every lane loads: lines
  // past ncol read the zero column of the index table (all entries nst, the
  // zero element), so no load sits under a divergent branch.
  int p = blockIdx.x;
  int t = p / nz;
  FastArgs<T> a = table_args(batch, t);
  if (threadIdx.x < N) {
    twl[threadIdx.x] = a.twX[threadIdx.x];
    tile[threadIdx.x * PU + ROWS] = Cx<T>{T(0), T(0)};
  }
  Cx<T> v[E];
  int idx[E];
  plane_stick_indices<N, L>(a.stickAt, min(line, a.ncol), q, idx);
  full_gather<T, L>(a, p - t * nz, idx, v);
  int xc = a.xCol[xi];
  int tn = p + G < total ? (p + G) / nz : t;
  FastArgs<T> an = table_args(batch, tn);
  plane_stick_indices<N, L>(an.stickAt, min(line, an.ncol), q, idx);
  for (;;) {
    const int z = p - t * nz;
    const int ncol = a.ncol;
    lds_barrier();  // tile and map free (first plane: twiddles, zero elements)
    // y: column `line`, from the registers gathered earlier (waves whose
    // lines are all past ncol skip it; the others transform their idle lines
    // too, into tile columns nobody reads)
    if (line0 < ncol) {
      Cx<T>* col = tile + opaque(line);
      wave_fft<T, N, true, 1, E, 1, S::PE>(v, col, PU, q, twl);
#pragma unroll
      for (int e = 0; e < E; ++e) col[(q + P * e) * PU] = v[e];
    }
    if (threadIdx.x < N) xmap[threadIdx.x] = xc < ncol ? xc : ROWS;
    lds_barrier();
    // x: row `line`, in place. Once its first pass is in LDS, plane p + G's
    // gathers (indices already in registers) and map entry, and plane
    // p + 2G's indices go out.
    Cx<T>* row = tile + line * PU;
    Cx<T> x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = row[xmap[q + P * e]];
    Cx<T>* out = reinterpret_cast<Cx<T>*>(a.out) + (static_cast<size_t>(z) * N + line) * N;
    const int pn = p + G;
    wave_fft<T, N, true, 1, E, 1, S::PE>(x, row, 1, q, twl, [&] {
      if (pn < total) {
        full_gather<T, L>(an, pn - tn * nz, idx, v);
        xc = pin(batch, an.xCol)[xi];
        const int pnn = pn + G;
        const FastArgs<T> ann = table_args(batch, pnn < total ? pnn / nz : tn);
        plane_stick_indices<N, L>(pin(batch, ann.stickAt), min(line, pin(ann.ncol)), q, idx);
      }
    });
    if constexpr (sizeof(T) == 4 && P == 8) {
      // 128-byte row segments per store instruction: the row goes back to
      // LDS in natural order and each lane stores two consecutive elements
#pragma unroll
      for (int e = 0; e < E; ++e) row[q + P * e] = x[e];
      wave_sync();
#pragma unroll
      for (int k = 0; k < E / 2; ++k) {
        const int i = 2 * (q + P * k);
        store_stream_pair(out + i, row[i], row[i + 1]);
      }
    } else if constexpr (sizeof(T) == 4 && P == 4) {
      // 4-lane lines: the wave's 16 rows (16 w .. 16 w + 15) go back to LDS
      // in natural order, then 8 lanes per row store two consecutive
      // elements each: 8 rows x 128 bytes per store instruction
#pragma unroll
      for (int e = 0; e < E; ++e) row[q + P * e] = x[e];
      wave_sync();
      Cx<T>* plane = reinterpret_cast<Cx<T>*>(a.out) + static_cast<size_t>(z) * N * N;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int r = 16 * wave + lane / 8 + 8 * j;
        const Cx<T>* src = tile + r * PU;
#pragma unroll
        for (int k = 0; k < N / 16; ++k) {
          const int i = 2 * (lane % 8 + 8 * k);
          store_stream_pair(plane + static_cast<size_t>(r) * N + i, src[i], src[i + 1]);
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) store_stream(&out[q + P * e], x[e]);
    }
    if (pn >= total) break;
    p = pn;
    t = tn;
    a = an;
    tn = pn + G < total ? (pn + G) / nz : t;
    an = table_args(batch, tn);
  }
}

template <typename T, int N>
__global__ void __launch_bounds__((FullShape<T, N>::NT))
    k_full_forward_batch(const FastArgs<T>* __restrict__ batch, int nz, int total) {
  using S = FullShape<T, N>;
  using L = typename S::L;
  constexpr int E = L::E, P = L::P, PU = S::PU;
  Cx<T>* tile = reinterpret_cast<Cx<T>*>(spfft_lds_fast);
  Tw<T>* twl = reinterpret_cast<Tw<T>*>(tile + tw_offset<T>(S::ROWS * PU));
  int* xmap = reinterpret_cast<int*>(twl + N);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int q = lane % P;
  const int line = full_line<P>(wave, lane / P);
  const int line0 = full_line<P>(wave, 0);
  const int xi = static_cast<int>(threadIdx.x % N);
  const int G = gridDim.x;
  // In flight at the top of the loop: plane p's rows (v).
  int p = blockIdx.x;
  int t = p / nz;
  FastArgs<T> a = table_args(batch, t);
  Cx<T> v[E];
  full_rows<T, N, L>(a, p - t * nz, line, q, v);
  if (threadIdx.x < N) {
    twl[threadIdx.x] = a.twX[threadIdx.x];
    xmap[threadIdx.x] = a.xCol[threadIdx.x];
  }
  for (;;) {
    const int z = p - t * nz;
    const int ncol = a.ncol;
    lds_barrier();  // tile and map free (first plane: twiddles)
    // x: row `line`, in place; the sparse set's columns stay in the row
    Cx<T>* row = tile + line * PU;
    wave_fft<T, N, false, 1, E, 1, S::PE>(v, row, 1, q, twl);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int u = xmap[q + P * e];
      if (u < ncol) row[u] = v[e];
    }
    lds_barrier();
    // y: column `line`, results straight to its sticks. Waves whose lines
    // are all past ncol skip the transform; the others transform their idle
    // lines too and write them, like absent positions, to element nst (the
    // backward z pass rewrites it with zero). Once the first pass is in LDS,
    // the next plane's rows and map entry and this line's stick indices go
    // out.
    const int pn = p + G;
    const int tn = pn < total ? pn / nz : t;
    const FastArgs<T> an = table_args(batch, tn);
    int xc = 0;
    int idx[E];
    auto prefetch = [&] {
      if (pn < total) {
        xc = pin(batch, an.xCol)[xi];
        full_rows<T, N, L>(an, pn - tn * nz, line, q, v);
      }
    };
    if (line0 < ncol) {
      const int* stickAt = pin(batch, a.stickAt);
      Cx<T>* col = tile + opaque(line);
      Cx<T> w[E];
#pragma unroll
      for (int e = 0; e < E; ++e) w[e] = col[(q + P * e) * PU];
      wave_sync();
      wave_fft<T, N, false, 1, E, 1, S::PE>(w, col, PU, q, twl, [&] {
        prefetch();
        plane_stick_indices<N, L>(stickAt, min(line, ncol), q, idx);
      });
      Cx<T>* zrow = a.zbuf + static_cast<size_t>(z) * a.pitch;
#pragma unroll
      for (int e = 0; e < E; ++e) zrow[idx[e]] = w[e];
    } else {
      prefetch();
    }
    if (pn >= total) break;
    if (threadIdx.x < N) xmap[threadIdx.x] = xc;
    p = pn;
    t = tn;
    a = an;
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
  auto kernel = a.r2c ? k_z_backward_fast<T, NZ, true> : k_z_backward_fast<T, NZ, false>;
  prepare_lds(kernel, Z::lds());
  prof::KernelSpan span(stream, a.r2c ? "k_z_backward_r2c" : "k_z_backward_fast", true);
  hipExtLaunchKernelGGL(kernel, dim3((a.nst + Z::S - 1) / Z::S), dim3(Z::NT), Z::lds(), stream,
                        span.start(), span.stop(), 0, a);
  check_launch();
}

template <typename T, int NZ>
void z_forward_n(const FastArgs<T>& a, hipStream_t stream) {
  using Z = ZShape<T, NZ>;
  auto kernel = k_z_forward_fast<T, NZ>;
  prepare_lds(kernel, Z::lds());
  prof::KernelSpan span(stream, "k_z_forward_fast", true);
  hipExtLaunchKernelGGL(kernel, dim3((a.nst + Z::S - 1) / Z::S), dim3(Z::NT), Z::lds(), stream,
                        span.start(), span.stop(), 0, a);
  check_launch();
}

template <int N>
int plane_grid(int nz) {
  return (nz + 7) / 8 * 8 * fast_classes(N);
}

template <typename T, int N>
void xy_backward_n(FastArgs<T> a, hipStream_t stream) {
  using S = XYShape<T, N>;
  const size_t lds = S::lds_backward(a.ncol, a.r2c != 0);
  a.pitchA = S::pitch_a(a.ncol);
  auto kernel = a.r2c ? k_xy_backward_fast<T, N, true> : k_xy_backward_fast<T, N, false>;
  prepare_lds(kernel, lds);
  prof::KernelSpan span(stream, a.r2c ? "k_xy_backward_r2c" : "k_xy_backward_fast", true);
  hipExtLaunchKernelGGL(kernel, dim3(plane_grid<N>(a.nz - a.zBegin)), dim3(S::NT), lds, stream,
                        span.start(), span.stop(), 0, a);
  check_launch();
}

template <typename T, int N>
void xy_forward_n(const FastArgs<T>& a, hipStream_t stream) {
  using S = XYShape<T, N>;
  const size_t lds = S::lds_forward();
  auto kernel = a.r2c ? k_xy_forward_fast<T, N, true> : k_xy_forward_fast<T, N, false>;
  prepare_lds(kernel, lds);
  prof::KernelSpan span(stream, a.r2c ? "k_xy_forward_r2c" : "k_xy_forward_fast", true);
  hipExtLaunchKernelGGL(kernel, dim3(plane_grid<N>(a.nz - a.zBegin)), dim3(S::NT), lds, stream,
                        span.start(), span.stop(), 0, a);
  check_launch();
}

template <typename T, int N>
bool fits(int ncol, bool rc) {
  return XYShape<T, N>::lds_backward(ncol, rc) <= kLdsMax &&
         XYShape<T, N>::lds_forward() <= kLdsMax;
}

// Lengths with class-split (and, for some, whole-plane) kernels, and all
// lengths with z and split-stage kernels.
bool square_length(int n) { return n == 32 || n == 64 || n == 96 || n == 128 || n == 192; }
bool fast_length(int n) { return square_length(n) || split_only_length(n); }

#define SPFFT_SQUARE_SWITCH(N, CALL) \
  switch (N) {                       \
    case 32: CALL(32); break;        \
    case 64: CALL(64); break;        \
    case 96: CALL(96); break;        \
    case 128: CALL(128); break;      \
    case 192: CALL(192); break;      \
    default: throw GPUFFTError();    \
  }

// The fixed lengths, in the four groups of the z kernels and the split plane
// stage (group 0 holds the class-split lengths; group 2 the even lengths
// with a factor 11 or 13; group 3 odd multiples of 11 and 13, 11^2 and 13^2
// and their doubles).
#define SPFFT_LENGTHS_0(X)                                                              \
  X(32) X(40) X(48) X(50) X(54) X(56) X(60) X(64) X(72) X(80) X(84) X(90) X(96) X(98)  \
  X(100) X(108) X(112) X(120) X(126) X(128) X(144) X(150) X(160) X(162) X(168) X(180) \
  X(192) X(196) X(200) X(216) X(224) X(240) X(252) X(256)
#define SPFFT_LENGTHS_1(X)                                                               \
  X(36) X(42) X(45) X(63) X(70) X(75) X(81) X(105) X(125) X(135) X(140) X(147) X(175)   \
  X(189) X(210) X(225) X(243) X(270) X(280) X(288) X(300) X(320) X(336) X(360) X(384)  \
  X(400) X(420) X(432) X(448) X(450) X(480) X(500) X(512)
#define SPFFT_LENGTHS_2(X)                                                                \
  X(44) X(52) X(66) X(78) X(88) X(104) X(110) X(130) X(132) X(154) X(156) X(176) X(182)  \
  X(198) X(208) X(220) X(234) X(260) X(264) X(312) X(330) X(352) X(390) X(416) X(440)   \
  X(462)
#define SPFFT_LENGTHS_3(X)                                                               \
  X(99) X(117) X(121) X(165) X(169) X(195) X(231) X(242) X(273) X(275) X(286) X(297)   \
  X(325) X(338) X(351)

}  // namespace

template <typename T, int N>
void xy_backward_batch_n(bool rc, const FastArgs<T>* batch, int ntrans, int nz, int maxNcol,
                         hipStream_t stream) {
  using S = XYShape<T, N>;
  const size_t lds = S::lds_backward(maxNcol, rc);
  auto kernel = rc ? k_xy_backward_batch<T, N, true> : k_xy_backward_batch<T, N, false>;
  prepare_lds(kernel, lds);
  prof::KernelSpan span(stream, rc ? "k_xy_backward_batch_r2c" : "k_xy_backward_batch", true);
  const int per = plane_grid<N>(nz);
  hipExtLaunchKernelGGL(kernel, dim3(per * ntrans), dim3(S::NT), lds, stream, span.start(),
                        span.stop(), 0, batch, per);
  check_launch();
}

template <typename T, int N>
void xy_forward_batch_n(bool rc, const FastArgs<T>* batch, int ntrans, int nz,
                        hipStream_t stream) {
  using S = XYShape<T, N>;
  const size_t lds = S::lds_forward();
  auto kernel = rc ? k_xy_forward_batch<T, N, true> : k_xy_forward_batch<T, N, false>;
  prepare_lds(kernel, lds);
  prof::KernelSpan span(stream, rc ? "k_xy_forward_batch_r2c" : "k_xy_forward_batch", true);
  const int per = plane_grid<N>(nz);
  hipExtLaunchKernelGGL(kernel, dim3(per * ntrans), dim3(S::NT), lds, stream, span.start(),
                        span.stop(), 0, batch, per);
  check_launch();
}

template <typename T, int NZ>
void z_batch_n(bool forward, bool rc, const FastArgs<T>* batch, const int2* map, int blocks,
               int useRuns, hipStream_t stream) {
  using Z = ZShape<T, NZ>;
  if (forward) {
    auto kernel = k_z_forward_batch<T, NZ>;
    prepare_lds(kernel, Z::lds());
    prof::KernelSpan span(stream, "k_z_forward_batch", true);
    hipExtLaunchKernelGGL(kernel, dim3(blocks), dim3(Z::NT), Z::lds(), stream,
                          span.start(), span.stop(), 0, batch, map, useRuns);
  } else {
    auto kernel = rc ? k_z_backward_batch<T, NZ, true> : k_z_backward_batch<T, NZ, false>;
    prepare_lds(kernel, Z::lds());
    prof::KernelSpan span(stream, rc ? "k_z_backward_batch_r2c" : "k_z_backward_batch", true);
    hipExtLaunchKernelGGL(kernel, dim3(blocks), dim3(Z::NT), Z::lds(), stream,
                          span.start(), span.stop(), 0, batch, map, useRuns);
  }
  check_launch();
}

// One call into the z kernels of length NZ.
template <typename T, int NZ>
void z_call(const ZCall<T>& c) {
  switch (c.op) {
    case ZCall<T>::Backward: z_backward_n<T, NZ>(*c.args, c.stream); break;
    case ZCall<T>::Forward: z_forward_n<T, NZ>(*c.args, c.stream); break;
    case ZCall<T>::Batch:
      z_batch_n<T, NZ>(c.forward, c.rc, c.args, c.map, c.blocks, c.useRuns, c.stream);
      break;
    case ZCall<T>::Geometry:
      c.geom[0] = ZShape<T, NZ>::S;
      c.geom[1] = ZShape<T, NZ>::LZ;
      c.geom[2] = pad_shift<T>();
      break;
  }
}

template <typename T>
void z_dispatch(int nz, const ZCall<T>& c) {
  if (!z_group0(nz, c) && !z_group1(nz, c) && !z_group2(nz, c) && !z_group3(nz, c)) {
    throw GPUFFTError();
  }
}

template <typename T>
void fast_z_geometry(int nz, int* sticksPerBlock, int* lineLength, int* padShift) {
  int g[3] = {};
  ZCall<T> c;
  c.op = ZCall<T>::Geometry;
  c.geom = g;
  z_dispatch(nz, c);
  *sticksPerBlock = g[0];
  *lineLength = g[1];
  *padShift = g[2];
}

template <typename T>
bool fast_path_supported(int nx, int ny, int nz, int ncol, bool r2c) {
  (void)r2c;
  // the z kernels and the split plane stage serve every fast length
  return fast_length(nx) && fast_length(ny) && fast_length(nz) && ncol >= 1;
}

template <typename T>
bool class_split_supported(int nx, int ny, int ncol, bool r2c) {
  if (nx != ny || !square_length(nx) || ncol < 1) return false;
  bool ok = false;
#define SPFFT_FITS(N) ok = fits<T, N>(ncol, r2c)
  SPFFT_SQUARE_SWITCH(nx, SPFFT_FITS)
#undef SPFFT_FITS
  return ok;
}

template <typename T>
void launch_fast_z_backward(int nz, const FastArgs<T>& a, hipStream_t stream) {
  ZCall<T> c;
  c.op = ZCall<T>::Backward;
  c.args = &a;
  c.stream = stream;
  z_dispatch(nz, c);
}

template <typename T>
void launch_fast_z_forward(int nz, const FastArgs<T>& a, hipStream_t stream) {
  ZCall<T> c;
  c.op = ZCall<T>::Forward;
  c.args = &a;
  c.stream = stream;
  z_dispatch(nz, c);
}

template <typename T>
void launch_fast_xy_backward(int nx, int ny, const FastArgs<T>& a, hipStream_t stream) {
  if (nx != ny) throw GPUFFTError();
#define SPFFT_CALL(N) xy_backward_n<T, N>(a, stream)
  SPFFT_SQUARE_SWITCH(nx, SPFFT_CALL)
#undef SPFFT_CALL
}

template <typename T>
void launch_fast_xy_forward(int nx, int ny, const FastArgs<T>& a, hipStream_t stream) {
  if (nx != ny) throw GPUFFTError();
#define SPFFT_CALL(N) xy_forward_n<T, N>(a, stream)
  SPFFT_SQUARE_SWITCH(nx, SPFFT_CALL)
#undef SPFFT_CALL
}

template <typename T>
int fast_xy_pitch_a(int n, int ncol) {
  int p = 0;
#define SPFFT_PITCH(N) p = XYShape<T, N>::pitch_a(ncol)
  SPFFT_SQUARE_SWITCH(n, SPFFT_PITCH)
#undef SPFFT_PITCH
  return p;
}

template <typename T>
void launch_fast_z_batch(bool forward, bool r2c, int nz, const FastArgs<T>* batch,
                         const int2* map, int blocks, bool useRuns, hipStream_t stream) {
  ZCall<T> c;
  c.op = ZCall<T>::Batch;
  c.args = batch;
  c.map = map;
  c.blocks = blocks;
  c.useRuns = useRuns ? 1 : 0;
  c.forward = forward;
  c.rc = r2c;
  c.stream = stream;
  z_dispatch(nz, c);
}

template <typename T>
void launch_fast_xy_batch(bool forward, bool r2c, int n, const FastArgs<T>* batch, int ntrans,
                          int nz, int maxNcol, hipStream_t stream) {
  if (forward) {
#define SPFFT_CALL(N) xy_forward_batch_n<T, N>(r2c, batch, ntrans, nz, stream)
    SPFFT_SQUARE_SWITCH(n, SPFFT_CALL)
#undef SPFFT_CALL
  } else {
#define SPFFT_CALL(N) xy_backward_batch_n<T, N>(r2c, batch, ntrans, nz, maxNcol, stream)
    SPFFT_SQUARE_SWITCH(n, SPFFT_CALL)
#undef SPFFT_CALL
  }
}

// The full-tile kernels serve N = 64 and 128 when the tile fits, and 96 in
// FP32 (FP64 lines of 24 elements do not fit in the register file).
template <typename T, int N>
constexpr bool full_tile_fits() {
  if constexpr (N == 64 || N == 128 || (N == 96 && sizeof(T) == 4)) {
    return FullShape<T, N>::lds() <= kLdsMax;
  } else {
    return false;
  }
}

// Resident workgroups per CU of the persistent plane kernel (0: not
// persistent).
template <typename T, int N>
int plane_slots_n(bool forward) {
  if constexpr (full_tile_fits<T, N>()) {
    using F = FullShape<T, N>;
    if (!F::PERSISTENT) return 0;
    auto kernel = forward ? k_full_forward_batch<T, N> : k_full_backward_batch<T, N>;
    prepare_lds(kernel, F::lds());
    int perCu = 0;
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, F::NT, F::lds()));
    return std::max(1, perCu);
  } else {
    return 0;
  }
}

template <typename T, int N>
void plane_batch_n(bool forward, const FastArgs<T>* batch, int ntrans, int nz, int maxNcol,
                   int slots, hipStream_t stream) {
  if constexpr (full_tile_fits<T, N>()) {
    using F = FullShape<T, N>;
    auto kernel = forward ? k_full_forward_batch<T, N> : k_full_backward_batch<T, N>;
    prepare_lds(kernel, F::lds());
    // persistent grid: the workgroups that are resident at once
    const int total = nz * ntrans;
    int grid = total;
    if (F::PERSISTENT) {
      int device = 0, cus = 0, perCu = 0;
      check_hip(hipGetDevice(&device));
      check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
      check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, F::NT, F::lds()));
      perCu = std::max(1, perCu);
      if (slots > 0) perCu = std::min(perCu, slots);
      grid = std::max(1, std::min(total, perCu * cus));
    }
    prof::KernelSpan span(stream, forward ? "k_full_forward_batch" : "k_full_backward_batch",
                          true);
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(F::NT), F::lds(), stream, span.start(),
                          span.stop(), 0, batch, nz, total);
    check_launch();
  } else {
    using S = PlaneShape<T, N>;
    const size_t lds = S::lds(maxNcol);
    auto kernel = forward ? k_plane_forward_batch<T, N> : k_plane_backward_batch<T, N>;
    prepare_lds(kernel, lds);
    prof::KernelSpan span(stream, forward ? "k_plane_forward_batch" : "k_plane_backward_batch",
                          true);
    hipExtLaunchKernelGGL(kernel, dim3(nz * ntrans), dim3(S::NT), lds, stream, span.start(),
                          span.stop(), 0, batch, nz);
    check_launch();
  }
}

#define SPFFT_PLANE_SWITCH(N, CALL) \
  switch (N) {                      \
    case 32: CALL(32); break;       \
    case 64: CALL(64); break;       \
    case 96: CALL(96); break;       \
    case 128: CALL(128); break;     \
    default: throw GPUFFTError();   \
  }

template <typename T>
bool plane_path_supported(int n, int ncol, bool r2c) {
  if (r2c || !(n == 32 || n == 64 || n == 96 || n == 128)) return false;
  size_t lds = 0;
#define SPFFT_PLANE_LDS(N) lds = full_tile_fits<T, N>() ? 0 : PlaneShape<T, N>::lds(ncol)
  SPFFT_PLANE_SWITCH(n, SPFFT_PLANE_LDS)
#undef SPFFT_PLANE_LDS
  return lds <= kLdsMax;
}

template <typename T>
int plane_pitch(int n, int ncol) {
  int p = 0;
#define SPFFT_PLANE_PITCH(N) p = PlaneShape<T, N>::pitch(ncol)
  SPFFT_PLANE_SWITCH(n, SPFFT_PLANE_PITCH)
#undef SPFFT_PLANE_PITCH
  return p;
}

template <typename T>
void launch_plane_xy_batch(bool forward, int n, const FastArgs<T>* batch, int ntrans, int nz,
                           int maxNcol, int slots, hipStream_t stream) {
#define SPFFT_CALL(N) plane_batch_n<T, N>(forward, batch, ntrans, nz, maxNcol, slots, stream)
  SPFFT_PLANE_SWITCH(n, SPFFT_CALL)
#undef SPFFT_CALL
}

template <typename T>
int plane_xy_slots(bool forward, int n) {
  int s = 0;
#define SPFFT_CALL(N) s = plane_slots_n<T, N>(forward)
  SPFFT_PLANE_SWITCH(n, SPFFT_CALL)
#undef SPFFT_CALL
  return s;
}

// ---- split plane stage launches

template <typename T, int N, typename Kernel, typename... Args>
void split_launch(Kernel kernel, const char* name, int blocks, hipStream_t stream, Args... args) {
  using S = SplitShape<T, N>;
  prepare_lds(kernel, S::lds());
  prof::KernelSpan span(stream, name, true);
  hipExtLaunchKernelGGL(kernel, dim3(blocks), dim3(S::NT), S::lds(), stream, span.start(),
                        span.stop(), 0, args...);
  check_launch();
}

// One stage of the split plane stage: the y kernel (columns of length NY)
// or the x kernel (rows of length NX, ny rows).
template <typename T, int NY>
void ysplit_n(bool forward, const FastArgs<T>& a, hipStream_t stream) {
  using S = SplitShape<T, NY>;
  const int blocks = S::col_groups(a.ncol) * (a.nz - a.zBegin);
  if (forward) {
    split_launch<T, NY>(k_ysplit_forward<T, NY>, "k_ysplit_forward", blocks, stream, a);
  } else {
    split_launch<T, NY>(a.r2c ? k_ysplit_backward<T, NY, true> : k_ysplit_backward<T, NY, false>,
                        "k_ysplit_backward", blocks, stream, a);
  }
}
template <typename T, int NX>
void xsplit_n(bool forward, const FastArgs<T>& a, hipStream_t stream) {
  using S = SplitShape<T, NX>;
  const int blocks = S::row_groups(a.ny, a.r2c != 0) * (a.nz - a.zBegin);
  if (forward) {
    split_launch<T, NX>(a.r2c ? k_xsplit_forward<T, NX, true> : k_xsplit_forward<T, NX, false>,
                        "k_xsplit_forward", blocks, stream, a);
  } else {
    split_launch<T, NX>(a.r2c ? k_xsplit_backward<T, NX, true> : k_xsplit_backward<T, NX, false>,
                        "k_xsplit_backward", blocks, stream, a);
  }
}
template <typename T, int NY>
void ysplit_batch_n(bool forward, bool rc, const FastArgs<T>* batch, int ntrans, int nz,
                    int maxNcol, hipStream_t stream) {
  using S = SplitShape<T, NY>;
  const int per = S::col_groups(maxNcol) * nz;
  if (forward) {
    split_launch<T, NY>(k_ysplit_forward_batch<T, NY>, "k_ysplit_forward_batch", per * ntrans,
                        stream, batch, per);
  } else {
    split_launch<T, NY>(rc ? k_ysplit_backward_batch<T, NY, true>
                           : k_ysplit_backward_batch<T, NY, false>,
                        "k_ysplit_backward_batch", per * ntrans, stream, batch, per);
  }
}
template <typename T, int NX>
void xsplit_batch_n(bool forward, bool rc, const FastArgs<T>* batch, int ntrans, int ny, int nz,
                    hipStream_t stream) {
  using S = SplitShape<T, NX>;
  const int per = S::row_groups(ny, rc) * nz;
  if (forward) {
    split_launch<T, NX>(rc ? k_xsplit_forward_batch<T, NX, true>
                           : k_xsplit_forward_batch<T, NX, false>,
                        "k_xsplit_forward_batch", per * ntrans, stream, batch, per);
  } else {
    split_launch<T, NX>(rc ? k_xsplit_backward_batch<T, NX, true>
                           : k_xsplit_backward_batch<T, NX, false>,
                        "k_xsplit_backward_batch", per * ntrans, stream, batch, per);
  }
}

template <typename T>
int split_pitch(int ncol, int nx) {
  // 128-byte row segments, but never wider than a row of the grid: the
  // buffer lives in the grid's work buffer of nx * ny * nz elements (ncol <=
  // nx, since columns are distinct x)
  const int q = static_cast<int>(128 / sizeof(Cx<T>));
  const int p = (ncol + q - 1) / q * q;
  return p <= nx ? p : nx;
}

// One call into the split plane stage of length N (the y kernel of columns
// of length N, or the x kernel of rows of length N).
template <typename T, int N>
void split_call(const SplitCall<T>& c) {
  if (!c.batch) {
    if (c.y) {
      ysplit_n<T, N>(c.forward, *c.args, c.stream);
    } else {
      xsplit_n<T, N>(c.forward, *c.args, c.stream);
    }
  } else if (c.y) {
    ysplit_batch_n<T, N>(c.forward, c.rc, c.args, c.ntrans, c.nz, c.maxNcol, c.stream);
  } else {
    xsplit_batch_n<T, N>(c.forward, c.rc, c.args, c.ntrans, c.ny, c.nz, c.stream);
  }
}

template <typename T>
void split_dispatch(int n, const SplitCall<T>& c) {
  if (!split_group0(n, c) && !split_group1(n, c) && !split_group2(n, c) &&
      !split_group3(n, c)) {
    throw GPUFFTError();
  }
}

template <typename T>
void launch_split_xy(bool forward, int nx, int ny, const FastArgs<T>& a, hipStream_t stream) {
  // backward: y then x; forward: x then y
  for (int k = 0; k < 2; ++k) {
    SplitCall<T> c;
    c.y = (k == 0) != forward;
    c.forward = forward;
    c.args = &a;
    c.stream = stream;
    split_dispatch(c.y ? ny : nx, c);
  }
}

template <typename T>
void launch_split_xy_batch(bool forward, bool r2c, int nx, int ny, const FastArgs<T>* batch,
                           int ntrans, int nz, int maxNcol, hipStream_t stream) {
  for (int k = 0; k < 2; ++k) {
    SplitCall<T> c;
    c.y = (k == 0) != forward;
    c.forward = forward;
    c.batch = true;
    c.rc = r2c;
    c.args = batch;
    c.ntrans = ntrans;
    c.ny = ny;
    c.nz = nz;
    c.maxNcol = maxNcol;
    c.stream = stream;
    split_dispatch(c.y ? ny : nx, c);
  }
}

// Dispatchers of one group of lengths.
#define SPFFT_Z_CASE(N) \
  case N:               \
    z_call<T, N>(c);    \
    return true;
#define SPFFT_SPLIT_CASE(N) \
  case N:                   \
    split_call<T, N>(c);    \
    return true;
#define SPFFT_DEFINE_GROUP(NAME, CALL, TYPE, CASE, LENGTHS) \
  bool NAME(int n, const CALL<TYPE>& c) {                   \
    using T = TYPE;                                         \
    switch (n) {                                            \
      LENGTHS(CASE)                                         \
      default: return false;                                \
    }                                                       \
  }

// z kernels
#define SPFFT_INSTANTIATE_FAST_Z(T)                                                      \
  template void fast_z_geometry<T>(int, int*, int*, int*);                               \
  template void launch_fast_z_backward<T>(int, const FastArgs<T>&, hipStream_t);         \
  template void launch_fast_z_forward<T>(int, const FastArgs<T>&, hipStream_t);          \
  template void launch_fast_z_batch<T>(bool, bool, int, const FastArgs<T>*, const int2*, \
                                       int, bool, hipStream_t);
// split plane stage
#define SPFFT_INSTANTIATE_FAST_SPLIT(T)                                                  \
  template int split_pitch<T>(int, int);                                                 \
  template void launch_split_xy<T>(bool, int, int, const FastArgs<T>&, hipStream_t);     \
  template void launch_split_xy_batch<T>(bool, bool, int, int, const FastArgs<T>*, int,  \
                                         int, int, hipStream_t);
// class-split and whole-plane kernels, path queries
#define SPFFT_INSTANTIATE_FAST_PLANE(T)                                                  \
  template bool plane_path_supported<T>(int, int, bool);                                 \
  template int plane_pitch<T>(int, int);                                                 \
  template void launch_plane_xy_batch<T>(bool, int, const FastArgs<T>*, int, int, int,   \
                                         int, hipStream_t);                              \
  template int plane_xy_slots<T>(bool, int);                                             \
  template int fast_xy_pitch_a<T>(int, int);                                             \
  template void launch_fast_xy_batch<T>(bool, bool, int, const FastArgs<T>*, int, int,   \
                                        int, hipStream_t);                               \
  template bool fast_path_supported<T>(int, int, int, int, bool);                        \
  template bool class_split_supported<T>(int, int, int, bool);                           \
  template void launch_fast_xy_backward<T>(int, int, const FastArgs<T>&, hipStream_t);   \
  template void launch_fast_xy_forward<T>(int, int, const FastArgs<T>&, hipStream_t);

// One precision and one kernel family (and, for the z and split kernels, one
// group of lengths) per translation unit, so that they build in parallel:
// the fast_kernels_*.hip files include this file with SPFFT_FAST_UNIT set (1
// to 17); compiled alone it is unit 0.
#ifndef SPFFT_FAST_UNIT
#define SPFFT_FAST_UNIT 0
#endif
#if SPFFT_FAST_UNIT == 0
SPFFT_INSTANTIATE_FAST_PLANE(double)
#elif SPFFT_FAST_UNIT == 1
SPFFT_DEFINE_GROUP(z_group0, ZCall, double, SPFFT_Z_CASE, SPFFT_LENGTHS_0)
SPFFT_INSTANTIATE_FAST_Z(double)
#elif SPFFT_FAST_UNIT == 2
SPFFT_DEFINE_GROUP(split_group0, SplitCall, double, SPFFT_SPLIT_CASE, SPFFT_LENGTHS_0)
SPFFT_INSTANTIATE_FAST_SPLIT(double)
#elif SPFFT_FAST_UNIT == 3
SPFFT_INSTANTIATE_FAST_PLANE(float)
#elif SPFFT_FAST_UNIT == 4
SPFFT_DEFINE_GROUP(z_group0, ZCall, float, SPFFT_Z_CASE, SPFFT_LENGTHS_0)
SPFFT_INSTANTIATE_FAST_Z(float)
#elif SPFFT_FAST_UNIT == 5
SPFFT_DEFINE_GROUP(split_group0, SplitCall, float, SPFFT_SPLIT_CASE, SPFFT_LENGTHS_0)
SPFFT_INSTANTIATE_FAST_SPLIT(float)
#elif SPFFT_FAST_UNIT == 6
SPFFT_DEFINE_GROUP(z_group1, ZCall, double, SPFFT_Z_CASE, SPFFT_LENGTHS_1)
#elif SPFFT_FAST_UNIT == 7
SPFFT_DEFINE_GROUP(split_group1, SplitCall, double, SPFFT_SPLIT_CASE, SPFFT_LENGTHS_1)
#elif SPFFT_FAST_UNIT == 8
SPFFT_DEFINE_GROUP(z_group1, ZCall, float, SPFFT_Z_CASE, SPFFT_LENGTHS_1)
#elif SPFFT_FAST_UNIT == 9
SPFFT_DEFINE_GROUP(split_group1, SplitCall, float, SPFFT_SPLIT_CASE, SPFFT_LENGTHS_1)
#elif SPFFT_FAST_UNIT == 10
SPFFT_DEFINE_GROUP(z_group2, ZCall, double, SPFFT_Z_CASE, SPFFT_LENGTHS_2)
#elif SPFFT_FAST_UNIT == 11
SPFFT_DEFINE_GROUP(split_group2, SplitCall, double, SPFFT_SPLIT_CASE, SPFFT_LENGTHS_2)
#elif SPFFT_FAST_UNIT == 12
SPFFT_DEFINE_GROUP(z_group2, ZCall, float, SPFFT_Z_CASE, SPFFT_LENGTHS_2)
#elif SPFFT_FAST_UNIT == 13
SPFFT_DEFINE_GROUP(split_group2, SplitCall, float, SPFFT_SPLIT_CASE, SPFFT_LENGTHS_2)
#elif SPFFT_FAST_UNIT == 14
SPFFT_DEFINE_GROUP(z_group3, ZCall, double, SPFFT_Z_CASE, SPFFT_LENGTHS_3)
#elif SPFFT_FAST_UNIT == 15
SPFFT_DEFINE_GROUP(split_group3, SplitCall, double, SPFFT_SPLIT_CASE, SPFFT_LENGTHS_3)
#elif SPFFT_FAST_UNIT == 16
SPFFT_DEFINE_GROUP(z_group3, ZCall, float, SPFFT_Z_CASE, SPFFT_LENGTHS_3)
#else
SPFFT_DEFINE_GROUP(split_group3, SplitCall, float, SPFFT_SPLIT_CASE, SPFFT_LENGTHS_3)
#endif

}  // namespace spfft
