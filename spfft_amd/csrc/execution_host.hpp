// CPU execution of one transform (SPFFT_PU_HOST), the reference's host path
// (src/execution/execution_host.cpp:50-120, :247-352) restated on the
// library's own FFT (host_fft.hpp) and std::thread workers.
//
// Buffers (the grid's two host arrays, X*Y*Z complex each):
//   array 1: sticks [stick][z] during the z stage, and the space domain
//            (z, y, x) (complex for C2C, real for R2C) -- they alias, as in the
//            reference: the space domain returned by space_domain() is array 1
//   array 2: planes [z][xFreq][y], y contiguous
// Backward: zero sticks, scatter values (last copy of a repeated triplet
// wins), R2C hermitian fill of the (0,0) stick, z FFT per stick, transpose
// into zeroed planes, R2C hermitian fill of every plane's x = 0 line, y FFT
// on the distinct x only, x FFT per row (C2R: hermitian completion, real
// part). Forward mirrors it, ending in gather and optional scaling by
// (T)(1.0 / (double)(Nx Ny Nz)).
#pragma once

#include <memory>
#include <vector>

#include "grid_internal.hpp"
#include "host_fft.hpp"
#include "index_map.hpp"
#include "spfft/types.h"

namespace spfft {

template <typename T>
class ExecutionHost {
public:
  ExecutionHost(const std::shared_ptr<IndexMap>& map, const std::shared_ptr<GridInternal<T>>& grid,
                int numThreads);

  // input: 2 * numValues T (user order); output: the space domain.
  void backward(const T* input, T* output);
  // input: the space domain; output: 2 * numValues T.
  void forward(const T* input, T* output, SpfftScalingType scaling);

  T* space_domain() const { return grid_->host_array1(); }

private:
  using C = Cx<T>;

  std::shared_ptr<IndexMap> map_;
  std::shared_ptr<GridInternal<T>> grid_;
  int threads_ = 1;
  T scale_;
  HostFft<T> fftX_, fftY_, fftZ_;
  std::vector<int> stickX_, stickY_;  // storage x and y of every stick
  std::vector<int> uniqueX_;          // distinct storage x among the sticks
  size_t lineLen_ = 1;                // longest line
};

}  // namespace spfft
