#include "transform_internal.hpp"

#include "hip_check.hpp"
#include "no_alloc.hpp"
#include "spfft/exceptions.hpp"

namespace spfft {

template <typename T>
TransformInternal<T>::TransformInternal(SpfftProcessingUnitType processingUnit,
                                        std::shared_ptr<GridInternal<T>> grid,
                                        std::shared_ptr<IndexMap> map)
    : processingUnit_(processingUnit), map_(std::move(map)), grid_(std::move(grid)) {
  // Checks in the order of the reference (transform_internal.cpp:52-77).
  if (!grid_) throw InvalidParameterError();
  if (map_->dimZ > grid_->max_local_z_length()) throw InvalidParameterError();
  if (map_->num_sticks() > grid_->max_num_local_z_columns()) throw InvalidParameterError();
  if (map_->dimX > grid_->max_dim_x() || map_->dimY > grid_->max_dim_y() ||
      map_->dimZ > grid_->max_dim_z()) {
    throw InvalidParameterError();
  }
  if (!(processingUnit & grid_->processing_unit())) throw InvalidParameterError();
  if (processingUnit != SPFFT_PU_HOST && processingUnit != SPFFT_PU_GPU) {
    throw InvalidParameterError();
  }
  if (map_->dimX <= 0 || map_->dimY <= 0 || map_->dimZ <= 0) throw InvalidParameterError();
  if (processingUnit == SPFFT_PU_HOST) {
    // CPU execution on the grid's host arrays (reference
    // transform_internal.cpp:90-97); no HIP call is made for it
    execHost_.reset(new ExecutionHost<T>(map_, grid_, grid_->num_threads()));
    return;
  }
  DeviceGuard guard(grid_->device_id());
  exec_.reset(new ExecutionGPU<T>(map_, grid_));
}

template <typename T>
TransformInternal<T> TransformInternal<T>::clone() const {
  std::shared_ptr<GridInternal<T>> newGrid(new GridInternal<T>(*grid_));
  TransformInternal<T> t(processingUnit_, std::move(newGrid), map_);
  t.execMode_ = execMode_;
  return t;
}

template <typename T>
T* TransformInternal<T>::space_domain_data(SpfftProcessingUnitType location) {
  if (execHost_) {
    // a host transform has its space domain in host memory only
    // (reference transform_internal.cpp:341-356)
    if (location != SPFFT_PU_HOST) throw InvalidParameterError();
    return execHost_->space_domain();
  }
  return location == SPFFT_PU_GPU ? exec_->space_domain_gpu() : exec_->space_domain_host();
}

template <typename T>
void TransformInternal<T>::forward(SpfftProcessingUnitType inputLocation, T* output,
                                   SpfftScalingType scaling) {
  forward(space_domain_data(inputLocation), output, scaling);
}

template <typename T>
void TransformInternal<T>::forward(const T* input, T* output, SpfftScalingType scaling) {
  if (execHost_) {
    execHost_->forward(input, output, scaling);
    return;
  }
  check_preceding();
  DeviceGuard guard(grid_->device_id());
  exec_->forward(input, output, scaling);
  exec_->synchronize(execMode_);
}

template <typename T>
void TransformInternal<T>::backward(const T* input, SpfftProcessingUnitType outputLocation) {
  backward(input, space_domain_data(outputLocation));
}

template <typename T>
void TransformInternal<T>::backward(const T* input, T* output) {
  if (execHost_) {
    execHost_->backward(input, output);
    return;
  }
  check_preceding();
  DeviceGuard guard(grid_->device_id());
  exec_->backward(input, output);
  exec_->synchronize(execMode_);
}

template <typename T>
void TransformInternal<T>::forward_enqueue(const T* input, T* output, SpfftScalingType scaling) {
  if (execHost_) {
    execHost_->forward(input, output, scaling);
    return;
  }
  check_preceding();
  DeviceGuard guard(grid_->device_id());
  exec_->forward(input, output, scaling, true);
}

template <typename T>
void TransformInternal<T>::backward_enqueue(const T* input, T* output) {
  if (execHost_) {
    execHost_->backward(input, output);
    return;
  }
  check_preceding();
  DeviceGuard guard(grid_->device_id());
  exec_->backward(input, output, true);
}

template <typename T>
void TransformInternal<T>::synchronize() {
  if (execHost_) return;  // host execution is synchronous
  DeviceGuard guard(grid_->device_id());
  exec_->synchronize(execMode_, true);
}

namespace {
template <typename T>
bool run_batched(const std::vector<TransformInternal<T>*>& ts,
                 const std::vector<std::shared_ptr<ExecutionGPU<T>>>& execs,
                 const std::vector<const T*>& in, const std::vector<T*>& out, bool forward,
                 const std::vector<SpfftScalingType>& scaling, int device) {
  std::vector<ExecutionGPU<T>*> ex;
  for (const auto& e : execs) ex.push_back(e.get());
  DeviceGuard guard(device);
  if (!ExecutionGPU<T>::run_batch(ex, in, out, forward, scaling)) return false;
  bool sync = false;
  for (const TransformInternal<T>* t : ts) sync |= t->execution_mode() == SPFFT_EXEC_SYNCHRONOUS;
  if (sync) check_hip(hipStreamSynchronize(nullptr));
  return true;
}
}  // namespace

template <typename T>
bool TransformInternal<T>::batch_backward(const std::vector<TransformInternal*>& ts,
                                          const std::vector<const T*>& in,
                                          const std::vector<T*>& out) {
  if (ts.empty()) return false;
  for (TransformInternal* t : ts) {
    if (!t->exec_) return false;  // host transforms run one by one
  }
  check_preceding();
  std::vector<std::shared_ptr<ExecutionGPU<T>>> execs;
  for (TransformInternal* t : ts) execs.push_back(t->exec_);
  return run_batched<T>(ts, execs, in, out, false, {}, ts[0]->device_id());
}

template <typename T>
bool TransformInternal<T>::batch_forward(const std::vector<TransformInternal*>& ts,
                                         const std::vector<const T*>& in,
                                         const std::vector<T*>& out,
                                         const std::vector<SpfftScalingType>& scaling) {
  if (ts.empty()) return false;
  for (TransformInternal* t : ts) {
    if (!t->exec_) return false;  // host transforms run one by one
  }
  check_preceding();
  std::vector<std::shared_ptr<ExecutionGPU<T>>> execs;
  for (TransformInternal* t : ts) execs.push_back(t->exec_);
  return run_batched<T>(ts, execs, in, out, true, scaling, ts[0]->device_id());
}

template <typename T>
void TransformInternal<T>::batch_noalloc(TransformInternal* const* ts, size_t n,
                                         const T* const* in, T* const* out, bool forward,
                                         const SpfftScalingType* scaling) {
  if (n < 1 || n > static_cast<size_t>(kNoAllocMaxTransforms) || !ts || !in || !out) {
    throw InvalidParameterError();
  }
  ExecutionGPU<T>* ex[kNoAllocMaxTransforms];
  for (size_t t = 0; t < n; ++t) {
    if (!ts[t] || !ts[t]->exec_ || ts[t]->device_id() != ts[0]->device_id()) {
      throw InvalidParameterError();
    }
    ex[t] = ts[t]->exec_.get();
  }
  check_preceding();
  DeviceGuard guard(ts[0]->device_id());
  const NoAllocScratch s = noalloc_scratch(ts[0]->device_id());
  {
    std::lock_guard<std::mutex> lock(*s.stageLock);
    ExecutionGPU<T>::run_batch_noalloc(ex, n, in, out, forward, scaling, s.device,
                                       kNoAllocScratchBytes, s.stage, kNoAllocStageBytes);
  }
  bool sync = false;
  for (size_t t = 0; t < n; ++t) sync |= ts[t]->execution_mode() == SPFFT_EXEC_SYNCHRONOUS;
  if (sync) check_hip(hipStreamSynchronize(nullptr));
}

template class TransformInternal<double>;
template class TransformInternal<float>;

}  // namespace spfft
