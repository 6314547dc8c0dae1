// Public C++ classes, batched entry points and the C ABI, for double and
// float. Behaviour of each entry point follows the reference:
//   C++ Transform / TransformFloat : src/spfft/transform.cpp:45-176,
//                                    src/spfft/transform_float.cpp
//   C++ Grid / GridFloat           : src/spfft/grid.cpp:40-86
//   multi_transform_*              : src/spfft/multi_transform_internal.hpp:46-175
//   C functions                    : src/spfft/transform.cpp:178-562,
//                                    src/spfft/grid.cpp:88-290,
//                                    src/spfft/multi_transform.cpp:84-147
// Every C function returns SPFFT_INVALID_HANDLE_ERROR for a null handle
// (create functions excepted), the error_code() of a thrown GenericError,
// and SPFFT_UNKNOWN_ERROR for anything else.
#include <memory>
#include <vector>

#include "grid_internal.hpp"
#include "index_map.hpp"
#include "no_alloc.hpp"
#include "spfft/exceptions.hpp"
#include "spfft_amd/no_alloc.h"
#include "spfft/spfft.h"
#include "spfft/spfft.hpp"
#include "transform_internal.hpp"

namespace spfft {

namespace {

template <typename T>
std::shared_ptr<TransformInternal<T>> make_transform(
    const std::shared_ptr<GridInternal<T>>& grid, SpfftProcessingUnitType processingUnit,
    SpfftTransformType transformType, int dimX, int dimY, int dimZ, int localZLength,
    int numLocalElements, SpfftIndexFormatType indexFormat, const int* indices) {
  if (dimX < 0 || dimY < 0 || dimZ < 0 || localZLength < 0 || numLocalElements < 0 ||
      (!indices && numLocalElements > 0)) {
    throw InvalidParameterError();
  }
  if (indexFormat != SPFFT_INDEX_TRIPLETS) throw InternalError();
  std::shared_ptr<IndexMap> map(new IndexMap(
      build_index_map(transformType, dimX, dimY, dimZ, numLocalElements, indices)));
  return std::make_shared<TransformInternal<T>>(processingUnit, grid, std::move(map));
}

template <typename T>
std::shared_ptr<TransformInternal<T>> make_independent(int maxNumThreads,
                                                       SpfftProcessingUnitType processingUnit,
                                                       SpfftTransformType transformType, int dimX,
                                                       int dimY, int dimZ, int numLocalElements,
                                                       SpfftIndexFormatType indexFormat,
                                                       const int* indices) {
  if (dimX < 0 || dimY < 0 || dimZ < 0 || numLocalElements < 0 ||
      (!indices && numLocalElements > 0)) {
    throw InvalidParameterError();
  }
  if (indexFormat != SPFFT_INDEX_TRIPLETS) throw InternalError();
  std::shared_ptr<IndexMap> map(new IndexMap(
      build_index_map(transformType, dimX, dimY, dimZ, numLocalElements, indices)));
  std::shared_ptr<GridInternal<T>> grid(new GridInternal<T>(
      dimX, dimY, dimZ, map->num_sticks(), processingUnit, maxNumThreads));
  return std::make_shared<TransformInternal<T>>(processingUnit, std::move(grid), std::move(map));
}

}  // namespace

// Batched execution (reference multi_transform_internal.hpp:46-175).
// Transforms of one shape on the fixed-length path run as one launch chain
// (one z launch and one xy launch for the whole batch); otherwise every
// transform is enqueued on its own stream before any is synchronised, so
// independent transforms still overlap on the GPU.
template <typename TransformType>
class MultiTransformInternal {
public:
  using ValueType = typename TransformType::ValueType;

  static void forward(int n, TransformType* transforms, const ValueType* const* inputs,
                      ValueType* const* outputs, const SpfftScalingType* scalings) {
    check(n, transforms);
    {
      std::vector<TransformInternal<ValueType>*> ts;
      for (int t = 0; t < n; ++t) ts.push_back(transforms[t].transform_.get());
      if (TransformInternal<ValueType>::batch_forward(
              ts, std::vector<const ValueType*>(inputs, inputs + n),
              std::vector<ValueType*>(outputs, outputs + n),
              std::vector<SpfftScalingType>(scalings, scalings + n))) {
        return;
      }
    }
    for (int t = 0; t < n; ++t) {
      transforms[t].transform_->forward_enqueue(inputs[t], outputs[t], scalings[t]);
    }
    for (int t = 0; t < n; ++t) transforms[t].transform_->synchronize();
  }

  static void forward(int n, TransformType* transforms, const SpfftProcessingUnitType* locations,
                      ValueType* const* outputs, const SpfftScalingType* scalings) {
    std::vector<const ValueType*> inputs(n > 0 ? n : 0);
    for (int t = 0; t < n; ++t) {
      inputs[t] = transforms[t].transform_->space_domain_data(locations[t]);
    }
    forward(n, transforms, inputs.data(), outputs, scalings);
  }

  static void backward(int n, TransformType* transforms, const ValueType* const* inputs,
                       ValueType* const* outputs) {
    check(n, transforms);
    {
      std::vector<TransformInternal<ValueType>*> ts;
      for (int t = 0; t < n; ++t) ts.push_back(transforms[t].transform_.get());
      if (TransformInternal<ValueType>::batch_backward(
              ts, std::vector<const ValueType*>(inputs, inputs + n),
              std::vector<ValueType*>(outputs, outputs + n))) {
        return;
      }
    }
    for (int t = 0; t < n; ++t) transforms[t].transform_->backward_enqueue(inputs[t], outputs[t]);
    for (int t = 0; t < n; ++t) transforms[t].transform_->synchronize();
  }

  static void backward(int n, TransformType* transforms, const ValueType* const* inputs,
                       const SpfftProcessingUnitType* locations) {
    std::vector<ValueType*> outputs(n > 0 ? n : 0);
    for (int t = 0; t < n; ++t) {
      outputs[t] = transforms[t].transform_->space_domain_data(locations[t]);
    }
    backward(n, transforms, inputs, outputs.data());
  }

  // Allocation-free batched call (no counterpart in the reference): the
  // C handles are used in place, no C++ objects are built.
  static void noalloc(int n, void* const* handles, const ValueType* const* inputs,
                      ValueType* const* outputs, bool forward, const SpfftScalingType* scalings) {
    if (n < 1 || n > kNoAllocMaxTransforms || !handles) throw InvalidParameterError();
    TransformInternal<ValueType>* ts[kNoAllocMaxTransforms];
    for (int t = 0; t < n; ++t) {
      if (!handles[t]) throw InvalidParameterError();
      ts[t] = reinterpret_cast<TransformType*>(handles[t])->transform_.get();
    }
    for (int a = 0; a < n; ++a) {
      for (int b = a + 1; b < n; ++b) {
        if (ts[a]->shared_grid(*ts[b])) throw InvalidParameterError();
      }
    }
    TransformInternal<ValueType>::batch_noalloc(ts, static_cast<size_t>(n), inputs, outputs,
                                                forward, scalings);
  }

private:
  static void check(int n, TransformType* transforms) {
    // Transforms of one grid share buffers and cannot run together
    // (reference multi_transform_internal.hpp:69-76).
    for (int a = 0; a < n; ++a) {
      for (int b = a + 1; b < n; ++b) {
        if (transforms[a].transform_->shared_grid(*transforms[b].transform_)) {
          throw InvalidParameterError();
        }
      }
    }
  }
};

// ------------------------------------------------------------ C++ classes

#define SPFFT_DEFINE_CLASSES(TRANSFORM, GRID, T)                                                  \
  TRANSFORM::TRANSFORM(int maxNumThreads, SpfftProcessingUnitType processingUnit,                 \
                       SpfftTransformType transformType, int dimX, int dimY, int dimZ,            \
                       int numLocalElements, SpfftIndexFormatType indexFormat,                    \
                       const int* indices)                                                        \
      : transform_(make_independent<T>(maxNumThreads, processingUnit, transformType, dimX, dimY, \
                                       dimZ, numLocalElements, indexFormat, indices)) {}          \
  TRANSFORM::TRANSFORM(const std::shared_ptr<GridInternal<T>>& grid,                              \
                       SpfftProcessingUnitType processingUnit, SpfftTransformType transformType, \
                       int dimX, int dimY, int dimZ, int localZLength, int numLocalElements,      \
                       SpfftIndexFormatType indexFormat, const int* indices)                      \
      : transform_(make_transform<T>(grid, processingUnit, transformType, dimX, dimY, dimZ,       \
                                     localZLength, numLocalElements, indexFormat, indices)) {}    \
  TRANSFORM::TRANSFORM(std::shared_ptr<TransformInternal<T>> transform)                           \
      : transform_(std::move(transform)) {}                                                       \
  TRANSFORM TRANSFORM::clone() const {                                                            \
    return TRANSFORM(std::make_shared<TransformInternal<T>>(transform_->clone()));                \
  }                                                                                               \
  SpfftTransformType TRANSFORM::type() const { return transform_->type(); }                       \
  int TRANSFORM::dim_x() const { return transform_->dim_x(); }                                    \
  int TRANSFORM::dim_y() const { return transform_->dim_y(); }                                    \
  int TRANSFORM::dim_z() const { return transform_->dim_z(); }                                    \
  int TRANSFORM::local_z_length() const { return transform_->num_local_xy_planes(); }             \
  int TRANSFORM::local_z_offset() const { return transform_->local_xy_plane_offset(); }           \
  int TRANSFORM::local_slice_size() const { return dim_x() * dim_y() * local_z_length(); }        \
  long long int TRANSFORM::global_size() const { return transform_->global_size(); }              \
  int TRANSFORM::num_local_elements() const { return transform_->num_local_elements(); }          \
  long long int TRANSFORM::num_global_elements() const {                                          \
    return transform_->num_global_elements();                                                     \
  }                                                                                               \
  SpfftProcessingUnitType TRANSFORM::processing_unit() const {                                    \
    return transform_->processing_unit();                                                         \
  }                                                                                               \
  int TRANSFORM::device_id() const { return transform_->device_id(); }                            \
  int TRANSFORM::num_threads() const { return transform_->num_threads(); }                        \
  SpfftExecType TRANSFORM::execution_mode() const { return transform_->execution_mode(); }        \
  void TRANSFORM::set_execution_mode(SpfftExecType mode) { transform_->set_execution_mode(mode); } \
  T* TRANSFORM::space_domain_data(SpfftProcessingUnitType dataLocation) {                         \
    return transform_->space_domain_data(dataLocation);                                           \
  }                                                                                               \
  void TRANSFORM::forward(SpfftProcessingUnitType inputLocation, T* output,                       \
                          SpfftScalingType scaling) {                                             \
    transform_->forward(inputLocation, output, scaling);                                          \
  }                                                                                               \
  void TRANSFORM::forward(const T* input, T* output, SpfftScalingType scaling) {                  \
    transform_->forward(input, output, scaling);                                                  \
  }                                                                                               \
  void TRANSFORM::backward(const T* input, SpfftProcessingUnitType outputLocation) {              \
    transform_->backward(input, outputLocation);                                                  \
  }                                                                                               \
  void TRANSFORM::backward(const T* input, T* output) { transform_->backward(input, output); }   \
                                                                                                  \
  GRID::GRID(int maxDimX, int maxDimY, int maxDimZ, int maxNumLocalZColumns,                      \
             SpfftProcessingUnitType processingUnit, int maxNumThreads)                           \
      : grid_(new GridInternal<T>(maxDimX, maxDimY, maxDimZ, maxNumLocalZColumns, processingUnit, \
                                  maxNumThreads)) {}                                              \
  GRID::GRID(const GRID& other) : grid_(new GridInternal<T>(*other.grid_)) {}                     \
  GRID& GRID::operator=(const GRID& other) {                                                      \
    if (this != &other) grid_.reset(new GridInternal<T>(*other.grid_));                           \
    return *this;                                                                                 \
  }                                                                                               \
  TRANSFORM GRID::create_transform(SpfftProcessingUnitType processingUnit,                        \
                                   SpfftTransformType transformType, int dimX, int dimY,          \
                                   int dimZ, int localZLength, int numLocalElements,              \
                                   SpfftIndexFormatType indexFormat, const int* indices) const {  \
    return TRANSFORM(grid_, processingUnit, transformType, dimX, dimY, dimZ, localZLength,        \
                     numLocalElements, indexFormat, indices);                                     \
  }                                                                                               \
  int GRID::max_dim_x() const { return grid_->max_dim_x(); }                                      \
  int GRID::max_dim_y() const { return grid_->max_dim_y(); }                                      \
  int GRID::max_dim_z() const { return grid_->max_dim_z(); }                                      \
  int GRID::max_num_local_z_columns() const { return grid_->max_num_local_z_columns(); }          \
  int GRID::max_local_z_length() const { return grid_->max_local_z_length(); }                    \
  SpfftProcessingUnitType GRID::processing_unit() const { return grid_->processing_unit(); }      \
  int GRID::device_id() const { return grid_->device_id(); }                                      \
  int GRID::num_threads() const { return grid_->num_threads(); }

SPFFT_DEFINE_CLASSES(Transform, Grid, double)
SPFFT_DEFINE_CLASSES(TransformFloat, GridFloat, float)
#undef SPFFT_DEFINE_CLASSES

#define SPFFT_DEFINE_MULTI(TRANSFORM, T)                                                         \
  void multi_transform_forward(int numTransforms, TRANSFORM* transforms,                         \
                               const SpfftProcessingUnitType* inputLocations,                    \
                               T* const* outputPointers, const SpfftScalingType* scalingTypes) { \
    MultiTransformInternal<TRANSFORM>::forward(numTransforms, transforms, inputLocations,        \
                                               outputPointers, scalingTypes);                    \
  }                                                                                              \
  void multi_transform_forward(int numTransforms, TRANSFORM* transforms,                         \
                               const T* const* inputPointers, T* const* outputPointers,          \
                               const SpfftScalingType* scalingTypes) {                           \
    MultiTransformInternal<TRANSFORM>::forward(numTransforms, transforms, inputPointers,         \
                                               outputPointers, scalingTypes);                    \
  }                                                                                              \
  void multi_transform_backward(int numTransforms, TRANSFORM* transforms,                        \
                                const T* const* inputPointers,                                   \
                                const SpfftProcessingUnitType* outputLocations) {                \
    MultiTransformInternal<TRANSFORM>::backward(numTransforms, transforms, inputPointers,        \
                                                outputLocations);                                \
  }                                                                                              \
  void multi_transform_backward(int numTransforms, TRANSFORM* transforms,                        \
                                const T* const* inputPointers, T* const* outputPointers) {       \
    MultiTransformInternal<TRANSFORM>::backward(numTransforms, transforms, inputPointers,        \
                                                outputPointers);                                 \
  }

SPFFT_DEFINE_MULTI(Transform, double)
SPFFT_DEFINE_MULTI(TransformFloat, float)
#undef SPFFT_DEFINE_MULTI

}  // namespace spfft

// ------------------------------------------------------------------ C ABI

namespace {

template <typename F>
SpfftError guarded(F&& f) {
  try {
    f();
  } catch (const spfft::GenericError& e) {
    return e.error_code();
  } catch (...) {
    return SPFFT_UNKNOWN_ERROR;
  }
  return SPFFT_SUCCESS;
}

template <typename H, typename F>
SpfftError with_handle(H handle, F&& f) {
  if (!handle) return SPFFT_INVALID_HANDLE_ERROR;
  return guarded(std::forward<F>(f));
}

// Copy the C handle array into C++ objects that share the same transforms.
template <typename TransformType, typename Handle>
std::vector<TransformType> gather(int n, Handle* handles) {
  std::vector<TransformType> out;
  out.reserve(n > 0 ? n : 0);
  for (int t = 0; t < n; ++t) {
    if (!handles[t]) throw spfft::InvalidParameterError();
    out.push_back(*reinterpret_cast<TransformType*>(handles[t]));
  }
  return out;
}

}  // namespace

#define SPFFT_DEFINE_C_API(PREFIX, GRIDH, TRANSFORMH, GRID, TRANSFORM, T)                          \
  SpfftError PREFIX##grid_create(GRIDH* grid, int maxDimX, int maxDimY, int maxDimZ,              \
                                 int maxNumLocalZColumns, SpfftProcessingUnitType processingUnit, \
                                 int maxNumThreads) {                                             \
    return guarded([&] {                                                                          \
      *grid = new spfft::GRID(maxDimX, maxDimY, maxDimZ, maxNumLocalZColumns, processingUnit,     \
                              maxNumThreads);                                                     \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##grid_destroy(GRIDH grid) {                                                   \
    return with_handle(grid, [&] { delete reinterpret_cast<spfft::GRID*>(grid); });               \
  }                                                                                               \
  SpfftError PREFIX##grid_max_dim_x(GRIDH grid, int* v) {                                         \
    return with_handle(grid, [&] { *v = reinterpret_cast<spfft::GRID*>(grid)->max_dim_x(); });    \
  }                                                                                               \
  SpfftError PREFIX##grid_max_dim_y(GRIDH grid, int* v) {                                         \
    return with_handle(grid, [&] { *v = reinterpret_cast<spfft::GRID*>(grid)->max_dim_y(); });    \
  }                                                                                               \
  SpfftError PREFIX##grid_max_dim_z(GRIDH grid, int* v) {                                         \
    return with_handle(grid, [&] { *v = reinterpret_cast<spfft::GRID*>(grid)->max_dim_z(); });    \
  }                                                                                               \
  SpfftError PREFIX##grid_max_num_local_z_columns(GRIDH grid, int* v) {                           \
    return with_handle(                                                                           \
        grid, [&] { *v = reinterpret_cast<spfft::GRID*>(grid)->max_num_local_z_columns(); });     \
  }                                                                                               \
  SpfftError PREFIX##grid_max_local_z_length(GRIDH grid, int* v) {                                \
    return with_handle(grid,                                                                      \
                       [&] { *v = reinterpret_cast<spfft::GRID*>(grid)->max_local_z_length(); }); \
  }                                                                                               \
  SpfftError PREFIX##grid_processing_unit(GRIDH grid, SpfftProcessingUnitType* v) {               \
    return with_handle(grid,                                                                      \
                       [&] { *v = reinterpret_cast<spfft::GRID*>(grid)->processing_unit(); });    \
  }                                                                                               \
  SpfftError PREFIX##grid_device_id(GRIDH grid, int* v) {                                         \
    return with_handle(grid, [&] { *v = reinterpret_cast<spfft::GRID*>(grid)->device_id(); });    \
  }                                                                                               \
  SpfftError PREFIX##grid_num_threads(GRIDH grid, int* v) {                                       \
    return with_handle(grid, [&] { *v = reinterpret_cast<spfft::GRID*>(grid)->num_threads(); });  \
  }                                                                                               \
  SpfftError PREFIX##transform_create(                                                            \
      TRANSFORMH* transform, GRIDH grid, SpfftProcessingUnitType processingUnit,                  \
      SpfftTransformType transformType, int dimX, int dimY, int dimZ, int localZLength,           \
      int numLocalElements, SpfftIndexFormatType indexFormat, const int* indices) {               \
    return with_handle(grid, [&] {                                                                \
      *transform = new spfft::TRANSFORM(reinterpret_cast<spfft::GRID*>(grid)->create_transform(  \
          processingUnit, transformType, dimX, dimY, dimZ, localZLength, numLocalElements,        \
          indexFormat, indices));                                                                 \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_create_independent(                                                \
      TRANSFORMH* transform, int maxNumThreads, SpfftProcessingUnitType processingUnit,           \
      SpfftTransformType transformType, int dimX, int dimY, int dimZ, int numLocalElements,       \
      SpfftIndexFormatType indexFormat, const int* indices) {                                     \
    return guarded([&] {                                                                          \
      *transform = new spfft::TRANSFORM(maxNumThreads, processingUnit, transformType, dimX, dimY, \
                                        dimZ, numLocalElements, indexFormat, indices);            \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_destroy(TRANSFORMH transform) {                                    \
    return with_handle(transform, [&] { delete reinterpret_cast<spfft::TRANSFORM*>(transform); }); \
  }                                                                                               \
  SpfftError PREFIX##transform_clone(TRANSFORMH transform, TRANSFORMH* newTransform) {            \
    return with_handle(transform, [&] {                                                           \
      *newTransform =                                                                             \
          new spfft::TRANSFORM(reinterpret_cast<spfft::TRANSFORM*>(transform)->clone());          \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_forward(TRANSFORMH transform,                                      \
                                       SpfftProcessingUnitType inputLocation, T* output,          \
                                       SpfftScalingType scaling) {                                \
    return with_handle(transform, [&] {                                                           \
      reinterpret_cast<spfft::TRANSFORM*>(transform)->forward(inputLocation, output, scaling);    \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_forward_ptr(TRANSFORMH transform, const T* input, T* output,       \
                                           SpfftScalingType scaling) {                            \
    return with_handle(transform, [&] {                                                           \
      reinterpret_cast<spfft::TRANSFORM*>(transform)->forward(input, output, scaling);            \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_backward(TRANSFORMH transform, const T* input,                     \
                                        SpfftProcessingUnitType outputLocation) {                 \
    return with_handle(transform, [&] {                                                           \
      reinterpret_cast<spfft::TRANSFORM*>(transform)->backward(input, outputLocation);            \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_backward_ptr(TRANSFORMH transform, const T* input, T* output) {    \
    return with_handle(transform, [&] {                                                           \
      reinterpret_cast<spfft::TRANSFORM*>(transform)->backward(input, output);                    \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_get_space_domain(TRANSFORMH transform,                             \
                                                SpfftProcessingUnitType dataLocation, T** data) { \
    return with_handle(transform, [&] {                                                           \
      *data = reinterpret_cast<spfft::TRANSFORM*>(transform)->space_domain_data(dataLocation);    \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_dim_x(TRANSFORMH transform, int* v) {                              \
    return with_handle(transform,                                                                 \
                       [&] { *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->dim_x(); });    \
  }                                                                                               \
  SpfftError PREFIX##transform_dim_y(TRANSFORMH transform, int* v) {                              \
    return with_handle(transform,                                                                 \
                       [&] { *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->dim_y(); });    \
  }                                                                                               \
  SpfftError PREFIX##transform_dim_z(TRANSFORMH transform, int* v) {                              \
    return with_handle(transform,                                                                 \
                       [&] { *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->dim_z(); });    \
  }                                                                                               \
  SpfftError PREFIX##transform_local_z_length(TRANSFORMH transform, int* v) {                     \
    return with_handle(transform, [&] {                                                           \
      *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->local_z_length();                      \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_local_slice_size(TRANSFORMH transform, int* v) {                   \
    return with_handle(transform, [&] {                                                           \
      *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->local_slice_size();                    \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_local_z_offset(TRANSFORMH transform, int* v) {                     \
    return with_handle(transform, [&] {                                                           \
      *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->local_z_offset();                      \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_global_size(TRANSFORMH transform, long long int* v) {              \
    return with_handle(                                                                           \
        transform, [&] { *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->global_size(); });  \
  }                                                                                               \
  /* Reports the local z length, as the reference does (transform.cpp:433-445). */               \
  SpfftError PREFIX##transform_num_local_elements(TRANSFORMH transform, int* v) {                 \
    return with_handle(transform, [&] {                                                           \
      *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->local_z_length();                      \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_num_global_elements(TRANSFORMH transform, long long int* v) {      \
    return with_handle(transform, [&] {                                                           \
      *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->num_global_elements();                 \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_device_id(TRANSFORMH transform, int* v) {                          \
    return with_handle(transform,                                                                 \
                       [&] { *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->device_id(); }); \
  }                                                                                               \
  SpfftError PREFIX##transform_num_threads(TRANSFORMH transform, int* v) {                        \
    return with_handle(transform, [&] {                                                           \
      *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->num_threads();                         \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_execution_mode(TRANSFORMH transform, SpfftExecType* v) {           \
    return with_handle(transform, [&] {                                                           \
      *v = reinterpret_cast<spfft::TRANSFORM*>(transform)->execution_mode();                      \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##transform_set_execution_mode(TRANSFORMH transform, SpfftExecType mode) {     \
    return with_handle(transform, [&] {                                                           \
      reinterpret_cast<spfft::TRANSFORM*>(transform)->set_execution_mode(mode);                   \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##multi_transform_forward(int numTransforms, TRANSFORMH* transforms,           \
                                             const SpfftProcessingUnitType* inputLocations,       \
                                             T* const* outputPointers,                            \
                                             const SpfftScalingType* scalingTypes) {              \
    return guarded([&] {                                                                          \
      auto ts = gather<spfft::TRANSFORM>(numTransforms, transforms);                              \
      spfft::multi_transform_forward(numTransforms, ts.data(), inputLocations, outputPointers,    \
                                     scalingTypes);                                               \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##multi_transform_forward_ptr(int numTransforms, TRANSFORMH* transforms,       \
                                                 const T* const* inputPointers,                   \
                                                 T* const* outputPointers,                        \
                                                 const SpfftScalingType* scalingTypes) {          \
    return guarded([&] {                                                                          \
      auto ts = gather<spfft::TRANSFORM>(numTransforms, transforms);                              \
      spfft::multi_transform_forward(numTransforms, ts.data(), inputPointers, outputPointers,     \
                                     scalingTypes);                                               \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##multi_transform_backward(int numTransforms, TRANSFORMH* transforms,          \
                                              const T* const* inputPointers,                      \
                                              const SpfftProcessingUnitType* outputLocations) {   \
    return guarded([&] {                                                                          \
      auto ts = gather<spfft::TRANSFORM>(numTransforms, transforms);                              \
      spfft::multi_transform_backward(numTransforms, ts.data(), inputPointers, outputLocations);  \
    });                                                                                           \
  }                                                                                               \
  SpfftError PREFIX##multi_transform_backward_ptr(int numTransforms, TRANSFORMH* transforms,      \
                                                  const T* const* inputPointers,                  \
                                                  T* const* outputPointers) {                     \
    return guarded([&] {                                                                          \
      auto ts = gather<spfft::TRANSFORM>(numTransforms, transforms);                              \
      spfft::multi_transform_backward(numTransforms, ts.data(), inputPointers, outputPointers);   \
    });                                                                                           \
  }

extern "C" {
SPFFT_DEFINE_C_API(spfft_, SpfftGrid, SpfftTransform, Grid, Transform, double)

static_assert(SPFFT_AMD_NOALLOC_SCRATCH_BYTES == spfft::kNoAllocScratchBytes &&
                  SPFFT_AMD_NOALLOC_MAX_TRANSFORMS == spfft::kNoAllocMaxTransforms,
              "public limits of the allocation-free entry point");

SpfftError spfft_amd_multi_transform_noalloc(int numTransforms, SpfftTransform* transforms,
                                             const double* const* inputPointers,
                                             double* const* outputPointers, int forward,
                                             const SpfftScalingType* scalingTypes) {
  return guarded([&] {
    spfft::MultiTransformInternal<spfft::Transform>::noalloc(
        numTransforms, transforms, inputPointers, outputPointers, forward != 0, scalingTypes);
  });
}

SPFFT_DEFINE_C_API(spfft_float_, SpfftFloatGrid, SpfftFloatTransform, GridFloat, TransformFloat,
                   float)
}

#undef SPFFT_DEFINE_C_API
