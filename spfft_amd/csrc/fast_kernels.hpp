// Fixed-length fast path of the sparse 3D FFT (lengths with compile-time
// radix schedules), C2C and R2C. Two kernels per direction:
//
//   z kernel  : sparse values -> LDS sticks -> z FFT -> zbuf[z][stick]
//               (forward: zbuf -> z FFT -> gather/scale into user order)
//   xy kernel : one workgroup per (plane z, residue class), no intermediate
//               between the y and x transforms.
//     backward, class r of R: rows y = r + R m. Decimation in frequency over
//       y: t_u[j] = W_NY^(r j) sum_l X_u[j + M l] W_R^(r l), then an M-point
//       FFT per column gives the rows' y-transformed values; the rows are
//       x-transformed in LDS and streamed to the dense plane.
//     forward, class c of R: columns x = c + R m. Decimation in frequency
//       over x on every row, M-point x FFTs, then full-length y FFTs on the
//       class's columns; results are gathered at the stick positions.
//   Each class reads the whole plane input (from L2 after the first class);
//   the classes of a plane are mapped to the same XCD. Class-split kernels
//   exist for nx = ny in {32, 64, 96, 128, 192}.
//   split plane stage (every fixed length, any nx and ny): a y kernel and an
//   x kernel per direction around a compact column buffer [z][y][column]
//   held in the grid's work buffer; every plane element is read once.
//   R2C (a compile-time variant): the backward z kernel fills the (0,0)
//   stick hermitian, the backward plane kernels fill the x = 0 column of
//   each plane and complete every row from its half spectrum; two hermitian
//   rows a, b enter one complex x FFT as a + i b, whose real and imaginary
//   parts are the two real rows (the split stage's forward x kernel packs two
//   real rows the same way and separates their spectra); the forward
//   class-split kernel reads real rows.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace spfft {

template <typename T>
struct FastArgs {
  int nz = 0, nst = 0, ncol = 0;
  int ny = 0;      // rows of a plane (the split stage's x kernels)
  int zBegin = 0;  // single-transform xy launches: planes [zBegin, nz) (chunked host I/O)
  int pitch = 0;   // row pitch of zbuf in elements (nst + 1 rounded up to 128 bytes);
                   // element nst of every row is zero
  int pitchA = 0;  // row pitch of the backward xy kernel's column tile (set at launch)
  int r2c = 0;            // real space domain, hermitian half in x
  int zeroZeroStick = 0;  // 0 when the (x=0, y=0) stick exists
  Cx<T>* zbuf = nullptr;
  const T* in = nullptr;
  T* out = nullptr;
  const int* entryValue = nullptr;
  const int* entrySlot = nullptr;
  const int* stickBegin = nullptr;
  const int* stickY = nullptr;
  const int* stickCol = nullptr;
  const int* colX = nullptr;
  const int* colBegin = nullptr;
  const int* stickAt = nullptr;       // [ncol + 1][M][R]: stick index at (column, y = j + M l),
                                      // or nst (the zero element); row ncol is all nst
  const int2* zEntries = nullptr;     // per value (sorted by slot): {value index,
                                      //  LDS tile position | repeated << 31}
  const int4* zInfo = nullptr;        // per stick {first value, first storage z, length} when
                                      // every stick's values are one run in user order
  int hasRuns = 0;                    // zInfo is valid (set on the host; kernels select the
                                      // stick-run path on this flag, never on the pointer)
  const int* xCol = nullptr;          // [NX]: column id of x, or ncol (a zero column)
  const int* resSticks = nullptr;     // sticks grouped by colX mod R
  const int* resBegin = nullptr;      // R + 1 offsets into resSticks
  const int* resYM = nullptr;         // per resSticks entry: y << 16 | (x / R)
  int pitchT = 0;                     // whole-plane kernels: row pitch of the plane tile
  Cx<T>* ybuf = nullptr;              // split plane kernels: compact columns [z][y][pitchY]
  int pitchY = 0;                     //   (column u of row y at u; pitchY >= ncol)
  const Tw<T>* twX = nullptr;
  const Tw<T>* twY = nullptr;
  const Tw<T>* twZ = nullptr;
  T scale = T(1);
  int useScale = 0;
};

// Residue classes per plane used by the xy kernels of length n.
// (A divisor of n: the stick index table is class-interleaved with it.)
constexpr int fast_classes(int n) {
  return n >= 192 && n % 8 == 0 ? 8 : (n % 4 == 0 ? 4 : (n % 2 == 0 ? 2 : 1));
}

// Geometry of the z kernels' LDS tile: sticks per workgroup, padded line
// length and the padding shift of the index map i -> i + (i >> shift).
template <typename T>
void fast_z_geometry(int nz, int* sticksPerBlock, int* lineLength, int* padShift);

// Row pitch of the [z][stick] buffer: room for nst sticks and one zero
// element, rounded up to a 128-byte multiple.
template <typename T>
inline int fast_pitch(int nst) {
  const int q = static_cast<int>(128 / (2 * sizeof(T)));
  return (nst + 1 + q - 1) / q * q;
}

// True when (nx, ny, nz) have fast kernels: z kernels and the split plane
// stage for every length in {32, 36, 40, 42, 44, 45, 48, 50, 52, 54, 56, 60,
// 63, 64, 66, 70, 72, 75, 78, 80, 81, 84, 88, 90, 96, 98, 99, 100, 104, 105,
// 108, 110, 112, 117, 120, 121, 125, 126, 128, 130, 132, 135, 140, 144, 147,
// 150, 154, 156, 160, 162, 165, 168, 169, 175, 176, 180, 182, 189, 192, 195,
// 196, 198, 200, 208, 210, 216, 220, 224, 225, 231, 234, 240, 242, 243, 252,
// 256, 260, 264, 270, 273, 275, 280, 286, 288, 297, 300, 312, 320, 325, 330,
// 336, 338, 351, 352, 360, 384, 390, 400, 416, 420, 432, 440, 448, 450, 462,
// 480, 500, 512} (108 lengths).
template <typename T>
bool fast_path_supported(int nx, int ny, int nz, int ncol, bool r2c);
// True when the class-split plane kernels serve the plane too: nx = ny in
// {32, 64, 96, 128, 192} and their LDS fits ncol columns.
template <typename T>
bool class_split_supported(int nx, int ny, int ncol, bool r2c);

// Row pitch of the backward xy kernel's column tile for ncol columns.
template <typename T>
int fast_xy_pitch_a(int n, int ncol);

// Batched launches over transforms of one shape. `batch` is a device table of
// arguments; `map` gives (transform, block) for every z workgroup.
// All transforms of a batch have the same type (r2c).
// useRuns = false: the per-value entry tables even for sets of stick runs.
template <typename T>
void launch_fast_z_batch(bool forward, bool r2c, int nz, const FastArgs<T>* batch,
                         const int2* map, int blocks, bool useRuns, hipStream_t stream);
template <typename T>
void launch_fast_xy_batch(bool forward, bool r2c, int n, const FastArgs<T>* batch, int ntrans,
                          int nz, int maxNcol, hipStream_t stream);

// Whole-plane kernels for batches (one workgroup per plane; no class split,
// so the plane's input is read once): power-of-two lengths up to 128, C2C,
// when the plane tile for `ncol` columns fits in LDS.
template <typename T>
bool plane_path_supported(int n, int ncol, bool r2c);
template <typename T>
int plane_pitch(int n, int ncol);
// slots > 0 caps the resident workgroups per CU of the persistent (FP32
// full-tile) kernels; 0 keeps all that fit.
template <typename T>
void launch_plane_xy_batch(bool forward, int n, const FastArgs<T>* batch, int ntrans, int nz,
                           int maxNcol, int slots, hipStream_t stream);
// Resident workgroups per CU of the persistent plane kernel for length n (0
// when that kernel is not persistent).
template <typename T>
int plane_xy_slots(bool forward, int n);

// Split plane stage: two kernels per direction with the compact column
// buffer ybuf between them (y stage on the sparse set's columns, x stage on
// rows). Every plane element is read once from memory, at the cost of the
// intermediate; it serves large lengths better than the class-split kernels,
// which read each plane once per residue class. Single transforms (planes
// [zBegin, nz)) and batches.
// Row pitch of the compact column buffer [z][y][column]: ncol rounded up to
// 128 bytes, at most nx (the buffer must fit nx * ny * nz elements).
template <typename T>
int split_pitch(int ncol, int nx);
template <typename T>
void launch_split_xy(bool forward, int nx, int ny, const FastArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_split_xy_batch(bool forward, bool r2c, int nx, int ny, const FastArgs<T>* batch,
                           int ntrans, int nz, int maxNcol, hipStream_t stream);

template <typename T>
void launch_fast_z_backward(int nz, const FastArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_fast_z_forward(int nz, const FastArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_fast_xy_backward(int nx, int ny, const FastArgs<T>& a, hipStream_t stream);
template <typename T>
void launch_fast_xy_forward(int nx, int ny, const FastArgs<T>& a, hipStream_t stream);

}  // namespace spfft
