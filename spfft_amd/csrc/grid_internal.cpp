#include "grid_internal.hpp"

#include <cstdlib>
#include <thread>

#include "hip_check.hpp"
#include "spfft/exceptions.hpp"

namespace spfft {

namespace {
int default_num_threads() {
  // The reference takes omp_get_max_threads() (grid_internal.cpp:69-73),
  // i.e. OMP_NUM_THREADS when set, otherwise the hardware thread count.
  if (const char* env = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(env);
    if (v > 0) return v;
  }
  const unsigned hw = std::thread::hardware_concurrency();
  return hw > 0 ? static_cast<int>(hw) : 1;
}
}  // namespace

template <typename T>
GridInternal<T>::GridInternal(int maxDimX, int maxDimY, int maxDimZ, int maxNumLocalZColumns,
                              SpfftProcessingUnitType processingUnit, int numThreads)
    : maxDimX_(maxDimX),
      maxDimY_(maxDimY),
      maxDimZ_(maxDimZ),
      maxNumLocalZColumns_(maxNumLocalZColumns),
      processingUnit_(processingUnit),
      numThreads_(numThreads < 1 ? default_num_threads() : numThreads) {
  if (maxDimX <= 0 || maxDimY <= 0 || maxDimZ <= 0 || maxNumLocalZColumns < 0) {
    throw InvalidParameterError();
  }
  if (!(processingUnit & (SPFFT_PU_HOST | SPFFT_PU_GPU))) throw InvalidParameterError();
  allocate();
}

template <typename T>
GridInternal<T>::GridInternal(const GridInternal& other)
    : maxDimX_(other.maxDimX_),
      maxDimY_(other.maxDimY_),
      maxDimZ_(other.maxDimZ_),
      maxNumLocalZColumns_(other.maxNumLocalZColumns_),
      processingUnit_(other.processingUnit_),
      numThreads_(other.numThreads_),
      deviceId_(other.deviceId_) {
  if (processingUnit_ & SPFFT_PU_GPU) {
    DeviceGuard guard(deviceId_);
    allocate();
  } else {
    allocate();
  }
}

template <typename T>
void GridInternal<T>::allocate() {
  const size_t hostSize = 2 * static_cast<size_t>(maxDimX_) * maxDimY_ * maxDimZ_;
  if (!(processingUnit_ & SPFFT_PU_GPU)) {
    // host-only grid: no HIP call at all
    host1_.reset(new HostBuffer<T>(hostSize, false));
    host2_.reset(new HostBuffer<T>(hostSize, false));
    return;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    throw GPUNoDeviceError();
  }
  check_hip(hipGetDevice(&deviceId_));
  const size_t gridSize = static_cast<size_t>(maxDimX_) * maxDimY_ * maxDimZ_;
  space_ = DeviceBuffer<T>(2 * gridSize);
  work_ = DeviceBuffer<T>(2 * gridSize);
  // rows of the [z][stick] buffer hold one element past the last stick (kept
  // zero by the fixed-length kernels) and are padded to 128 bytes (16 FP32
  // sticks)
  const size_t pitch = (static_cast<size_t>(maxNumLocalZColumns_) + 1 + 15) / 16 * 16;
  sticks_ = DeviceBuffer<T>(2 * static_cast<size_t>(maxDimZ_) * pitch);
  if (processingUnit_ & SPFFT_PU_HOST) {
    host1_.reset(new HostBuffer<T>(hostSize, true));
    host2_.reset(new HostBuffer<T>(hostSize, true));
  }
}

template <typename T>
T* GridInternal<T>::space_host() {
  if (host1_) return host1_->data();
  std::lock_guard<std::mutex> lock(hostMutex_);
  if (!host_) {
    const size_t gridSize = static_cast<size_t>(maxDimX_) * maxDimY_ * maxDimZ_;
    host_.reset(new PinnedBuffer<T>(2 * gridSize));
  }
  return host_->data();
}

template <typename T>
void* GridInternal<T>::host_staging(size_t bytes) {
  std::lock_guard<std::mutex> lock(hostMutex_);
  if (!staging_ || staging_->size() < bytes) {
    staging_.reset();
    staging_.reset(new HostBuffer<char>(bytes, true));
  }
  return staging_->data();
}

template class GridInternal<double>;
template class GridInternal<float>;

}  // namespace spfft
