// Transform core: argument checks, device selection, stage sequencing and
// the space-domain accessors (reference: src/spfft/transform_internal.hpp:55-149,
// src/spfft/transform_internal.cpp:45-363).
#pragma once

#include <memory>
#include <vector>

#include "execution.hpp"
#include "execution_host.hpp"
#include "grid_internal.hpp"
#include "index_map.hpp"
#include "spfft/types.h"

namespace spfft {

template <typename T>
class TransformInternal {
public:
  TransformInternal(SpfftProcessingUnitType processingUnit,
                    std::shared_ptr<GridInternal<T>> grid, std::shared_ptr<IndexMap> map);

  // Same parameters on a deep copy of the grid (reference transform_internal.cpp:176-179).
  TransformInternal clone() const;

  void forward(SpfftProcessingUnitType inputLocation, T* output, SpfftScalingType scaling);
  void forward(const T* input, T* output, SpfftScalingType scaling);
  void backward(const T* input, SpfftProcessingUnitType outputLocation);
  void backward(const T* input, T* output);

  // Enqueue without the final synchronisation (batched execution).
  void forward_enqueue(const T* input, T* output, SpfftScalingType scaling);
  void backward_enqueue(const T* input, T* output);
  void synchronize();

  // One fused launch chain for a batch of transforms of one shape (see
  // ExecutionGPU::run_batch); false when the batch does not qualify.
  static bool batch_backward(const std::vector<TransformInternal*>& ts,
                             const std::vector<const T*>& in, const std::vector<T*>& out);
  static bool batch_forward(const std::vector<TransformInternal*>& ts,
                            const std::vector<const T*>& in, const std::vector<T*>& out,
                            const std::vector<SpfftScalingType>& scaling);

  // Allocation-free batched call (ExecutionGPU::run_batch_noalloc) of n
  // GPU transforms of one device; the tables go to that device's fixed
  // scratch. Synchronises when any transform is in synchronous mode.
  static void batch_noalloc(TransformInternal* const* ts, size_t n, const T* const* in,
                            T* const* out, bool forward, const SpfftScalingType* scaling);

  T* space_domain_data(SpfftProcessingUnitType location);

  SpfftTransformType type() const { return map_->type; }
  int dim_x() const { return map_->dimX; }
  int dim_y() const { return map_->dimY; }
  int dim_z() const { return map_->dimZ; }
  int num_local_xy_planes() const { return map_->dimZ; }
  int local_xy_plane_offset() const { return 0; }
  SpfftProcessingUnitType processing_unit() const { return processingUnit_; }
  int device_id() const { return grid_->device_id(); }
  int num_threads() const { return grid_->num_threads(); }
  int num_local_elements() const { return map_->num_values(); }
  long long num_global_elements() const { return map_->num_values(); }
  long long global_size() const {
    return static_cast<long long>(map_->dimX) * map_->dimY * map_->dimZ;
  }
  SpfftExecType execution_mode() const { return execMode_; }
  void set_execution_mode(SpfftExecType mode) { execMode_ = mode; }
  bool shared_grid(const TransformInternal& other) const { return other.grid_ == grid_; }
  const IndexMap& index_map() const { return *map_; }

private:
  SpfftProcessingUnitType processingUnit_;
  SpfftExecType execMode_ = SPFFT_EXEC_SYNCHRONOUS;
  std::shared_ptr<IndexMap> map_;
  std::shared_ptr<GridInternal<T>> grid_;
  std::shared_ptr<ExecutionGPU<T>> exec_;       // SPFFT_PU_GPU transforms
  std::shared_ptr<ExecutionHost<T>> execHost_;  // SPFFT_PU_HOST transforms
};

}  // namespace spfft
