#include "execution_host.hpp"

#include <algorithm>
#include <cstring>
#include <thread>

namespace spfft {

namespace {

// Runs f(begin, end, worker) over [0, n) split into `threads` contiguous
// chunks, the first on the calling thread.
template <typename F>
void parallel_for(int threads, long long n, F&& f) {
  const int w = static_cast<int>(std::max<long long>(1, std::min<long long>(threads, n)));
  if (w <= 1) {
    if (n > 0) f(0LL, n, 0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(w - 1);
  auto chunk = [&](int k) { return n * k / w; };
  for (int k = 1; k < w; ++k) pool.emplace_back([&, k] { f(chunk(k), chunk(k + 1), k); });
  f(chunk(0), chunk(1), 0);
  for (std::thread& t : pool) t.join();
}

template <typename T>
bool nonzero(const Cx<T>& v) {
  return v.re != T(0) || v.im != T(0);
}

// Hermitian fill of a line of n elements with stride `stride` (reference
// symmetry_host.hpp:47-58 and :72-89): k = 1 .. n/2 in order, then
// k = n/2 + 1 .. n - 1, each non-zero entry writes its conjugate to n - k.
// On non-hermitian input a non-zero lower-half entry therefore wins, an upper
// entry fills a zero lower one, and a self-paired entry is conjugated.
template <typename T>
void hermitian_fill(Cx<T>* v, int n, size_t stride) {
  for (int k = 1; k < n; ++k) {
    const Cx<T> a = v[k * stride];
    if (nonzero(a)) v[(n - k) * stride] = Cx<T>{a.re, -a.im};
  }
}

}  // namespace

template <typename T>
ExecutionHost<T>::ExecutionHost(const std::shared_ptr<IndexMap>& map,
                                const std::shared_ptr<GridInternal<T>>& grid, int numThreads)
    : map_(map),
      grid_(grid),
      threads_(std::max(1, numThreads)),
      scale_(static_cast<T>(1.0 / static_cast<double>(static_cast<unsigned long long>(map->dimX) *
                                                      static_cast<unsigned long long>(map->dimY) *
                                                      static_cast<unsigned long long>(map->dimZ)))),
      fftX_(map->dimX),
      fftY_(map->dimY),
      fftZ_(map->dimZ) {
  const int nst = map_->num_sticks();
  stickX_.resize(nst);
  stickY_.resize(nst);
  for (int s = 0; s < nst; ++s) {
    stickX_[s] = map_->stickKey[s] / map_->dimY;
    stickY_[s] = map_->stickKey[s] % map_->dimY;
  }
  uniqueX_ = stickX_;
  std::sort(uniqueX_.begin(), uniqueX_.end());
  uniqueX_.erase(std::unique(uniqueX_.begin(), uniqueX_.end()), uniqueX_.end());
  lineLen_ = static_cast<size_t>(std::max({map_->dimX, map_->dimY, map_->dimZ, 1}));
}

template <typename T>
void ExecutionHost<T>::backward(const T* input, T* output) {
  const int nx = map_->dimX, ny = map_->dimY, nz = map_->dimZ, nxf = map_->dimXFreq;
  const int nst = map_->num_sticks();
  const bool r2c = map_->type == SPFFT_TRANS_R2C;
  C* sticks = reinterpret_cast<C*>(grid_->host_array1());
  C* planes = reinterpret_cast<C*>(grid_->host_array2());
  const C* in = reinterpret_cast<const C*>(input);
  const size_t work = std::max(fftX_.work_size(), std::max(fftY_.work_size(), fftZ_.work_size()));

  // sticks: zero, scatter (serial, so the last copy of a repeated triplet
  // wins), hermitian fill of the (0,0) stick
  std::fill(sticks, sticks + static_cast<size_t>(nst) * nz, C{T(0), T(0)});
  for (int i = 0; i < map_->num_values(); ++i) sticks[map_->valueSlot[i]] = in[i];
  if (r2c && map_->zeroZeroStick < nst) {
    hermitian_fill(sticks + static_cast<size_t>(map_->zeroZeroStick) * nz, nz, 1);
  }
  // z transform of every stick, transposed into zeroed planes[z][x][y]
  std::fill(planes, planes + static_cast<size_t>(nz) * nxf * ny, C{T(0), T(0)});
  parallel_for(threads_, nst, [&](long long b, long long e, int) {
    std::vector<C> buf(work);
    for (long long s = b; s < e; ++s) {
      C* stick = sticks + static_cast<size_t>(s) * nz;
      fftZ_.run(stick, stick, +1, buf.data());
      C* dst = planes + static_cast<size_t>(stickX_[s]) * ny + stickY_[s];
      for (int z = 0; z < nz; ++z) dst[static_cast<size_t>(z) * nxf * ny] = stick[z];
    }
  });
  // R2C: hermitian fill of the x = 0 line of every plane (all zero when no
  // stick has x = 0); y transform on the distinct x only
  const long long nu = static_cast<long long>(uniqueX_.size());
  parallel_for(threads_, nz * nu, [&](long long b, long long e, int) {
    std::vector<C> buf(work);
    for (long long zu = b; zu < e; ++zu) {
      const int z = static_cast<int>(zu / nu), x = uniqueX_[zu % nu];
      C* col = planes + (static_cast<size_t>(z) * nxf + x) * ny;
      if (r2c && x == 0) hermitian_fill(col, ny, 1);
      fftY_.run(col, col, +1, buf.data());
    }
  });
  // x transform of every row into the space domain
  parallel_for(threads_, static_cast<long long>(nz) * ny, [&](long long b, long long e, int) {
    std::vector<C> buf(work), line(nx);
    for (long long zy = b; zy < e; ++zy) {
      const int z = static_cast<int>(zy / ny), y = static_cast<int>(zy % ny);
      const C* src = planes + static_cast<size_t>(z) * nxf * ny + y;
      for (int k = 0; k < nxf; ++k) line[k] = src[static_cast<size_t>(k) * ny];
      if (!r2c) {
        C* out = reinterpret_cast<C*>(output) + static_cast<size_t>(zy) * nx;
        fftX_.run(line.data(), out, +1, buf.data());
      } else {
        // C2R: the imaginary parts of the DC and Nyquist inputs do not
        // contribute to a real signal; the upper half is the conjugate of
        // the lower one
        line[0].im = T(0);
        if (nx % 2 == 0) line[nx / 2].im = T(0);
        for (int k = nxf; k < nx; ++k) line[k] = C{line[nx - k].re, -line[nx - k].im};
        fftX_.run(line.data(), line.data(), +1, buf.data());
        T* out = output + static_cast<size_t>(zy) * nx;
        for (int x = 0; x < nx; ++x) out[x] = line[x].re;
      }
    }
  });
}

template <typename T>
void ExecutionHost<T>::forward(const T* input, T* output, SpfftScalingType scaling) {
  const int nx = map_->dimX, ny = map_->dimY, nz = map_->dimZ, nxf = map_->dimXFreq;
  const int nst = map_->num_sticks();
  const bool r2c = map_->type == SPFFT_TRANS_R2C;
  C* sticks = reinterpret_cast<C*>(grid_->host_array1());
  C* planes = reinterpret_cast<C*>(grid_->host_array2());
  const size_t work = std::max(fftX_.work_size(), std::max(fftY_.work_size(), fftZ_.work_size()));

  // x transform of every row into planes[z][x][y] (x < xFreq)
  parallel_for(threads_, static_cast<long long>(nz) * ny, [&](long long b, long long e, int) {
    std::vector<C> buf(work), line(nx);
    for (long long zy = b; zy < e; ++zy) {
      const int z = static_cast<int>(zy / ny), y = static_cast<int>(zy % ny);
      if (!r2c) {
        fftX_.run(reinterpret_cast<const C*>(input) + static_cast<size_t>(zy) * nx, line.data(),
                  -1, buf.data());
      } else {
        const T* row = input + static_cast<size_t>(zy) * nx;
        for (int x = 0; x < nx; ++x) line[x] = C{row[x], T(0)};
        fftX_.run(line.data(), line.data(), -1, buf.data());
      }
      C* dst = planes + static_cast<size_t>(z) * nxf * ny + y;
      for (int k = 0; k < nxf; ++k) dst[static_cast<size_t>(k) * ny] = line[k];
    }
  });
  // y transform on the distinct x only
  const long long nu = static_cast<long long>(uniqueX_.size());
  parallel_for(threads_, nz * nu, [&](long long b, long long e, int) {
    std::vector<C> buf(work);
    for (long long zu = b; zu < e; ++zu) {
      const int z = static_cast<int>(zu / nu), x = uniqueX_[zu % nu];
      C* col = planes + (static_cast<size_t>(z) * nxf + x) * ny;
      fftY_.run(col, col, -1, buf.data());
    }
  });
  // gather every stick from the planes, z transform
  parallel_for(threads_, nst, [&](long long b, long long e, int) {
    std::vector<C> buf(work);
    for (long long s = b; s < e; ++s) {
      C* stick = sticks + static_cast<size_t>(s) * nz;
      const C* src = planes + static_cast<size_t>(stickX_[s]) * ny + stickY_[s];
      for (int z = 0; z < nz; ++z) stick[z] = src[static_cast<size_t>(z) * nxf * ny];
      fftZ_.run(stick, stick, -1, buf.data());
    }
  });
  // gather the values in user order, scaled by one multiply per component
  C* out = reinterpret_cast<C*>(output);
  const bool scaled = scaling == SPFFT_FULL_SCALING;
  parallel_for(threads_, map_->num_values(), [&](long long b, long long e, int) {
    for (long long i = b; i < e; ++i) {
      C v = sticks[map_->valueSlot[i]];
      if (scaled) v = C{v.re * scale_, v.im * scale_};
      out[i] = v;
    }
  });
}

template class ExecutionHost<double>;
template class ExecutionHost<float>;

}  // namespace spfft
