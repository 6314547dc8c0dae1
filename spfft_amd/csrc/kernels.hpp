// Launch interface of the HIP kernels of the sparse 3D FFT (generic path).
//
// Data flow of one backward transform (forward is the exact mirror):
//   z stage : sparse values -> LDS sticks (zero filled) -> z FFT -> zbuf[z][stick]
//   y stage : zbuf row of plane z -> LDS columns (distinct x only) -> y FFT
//             -> ybuf[z][y][column]
//   x stage : ybuf rows -> LDS rows of length dimX (zero filled) -> x FFT
//             (or hermitian completion + FFT for C2R) -> dense space domain
// The scatter into sticks, the stick<->plane transpose and the zero fill of
// the planes are all folded into the LDS staging of these kernels, so no
// memset and no separate transpose kernel touches HBM.
#pragma once

#include <hip/hip_runtime.h>

namespace spfft {

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
  T re, im;
};

// Twiddle table entry exp(-2 pi i t / n), generated in FP64 on the GPU. FP64
// transforms keep the root rounded once; FP32 transforms keep it as a
// compensated pair hi + lo with hi = fl32(w) and lo = fl32(w - hi), so a
// product with it carries the root to about 2^-48 and only the product's
// own rounding remains.
template <typename T>
struct alignas(2 * sizeof(Cx<T>)) TwPair {
  Cx<T> hi, lo;
};
template <typename T>
struct TwSelect {
  using type = Cx<T>;
};
template <>
struct TwSelect<float> {
  using type = TwPair<float>;
};
template <typename T>
using Tw = typename TwSelect<T>::type;

constexpr int kMaxPasses = 24;

// Mixed-radix Stockham schedule for one FFT length.
struct FftDesc {
  int n = 1;
  int npass = 0;
  int radix[kMaxPasses] = {};
};

FftDesc make_fft_desc(int n);

template <typename T>
struct StageArgs {
  int nx = 0, ny = 0, nz = 0, nxf = 0;  // dims, nxf = dimX/2+1 for R2C
  int nst = 0;                          // number of z-sticks
  int ncol = 0;                         // number of distinct x among sticks
  int batch = 1;                        // 1D transforms per workgroup
  int r2c = 0;
  int zeroZeroStick = 0;
  FftDesc fft;                 // schedule of the stage's FFT length
  const Tw<T>* tw = nullptr;   // exp(-2 pi i t / n), t in [0, n)
  const int* entryValue = nullptr;
  const int* entrySlot = nullptr;
  const int* stickBegin = nullptr;
  const int* stickY = nullptr;
  const int* stickCol = nullptr;
  const int* colX = nullptr;
  const int* colBegin = nullptr;
  const T* in = nullptr;
  T* out = nullptr;
  Cx<T>* zbuf = nullptr;
  Cx<T>* ybuf = nullptr;
  T scale = T(1);
  int useScale = 0;
};

// Per-workgroup LDS need of a stage for FFT length n and the batch chosen.
template <typename T>
int stage_batch(int n, int count, size_t* ldsBytes);

template <typename T>
void generate_twiddles(Tw<T>* tw, int n, hipStream_t stream);

template <typename T>
void launch_z_backward(const StageArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_y_backward(const StageArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_x_backward(const StageArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_x_forward(const StageArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_y_forward(const StageArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_z_forward(const StageArgs<T>& a, hipStream_t stream);

}  // namespace spfft
