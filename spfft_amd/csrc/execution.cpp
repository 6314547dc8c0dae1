#include "execution.hpp"

#include <algorithm>
#include <cstdlib>
#include <utility>
#include <vector>

#include "hip_check.hpp"
#include "spfft/exceptions.hpp"

namespace spfft {

namespace {

// Host or device view of a user pointer (reference:
// src/gpu_util/gpu_pointer_translation.hpp:41-71). Device and managed
// allocations are used in place; anything else goes through a copy.
template <typename P>
std::pair<P*, P*> translate_pointer(P* ptr) {
  hipPointerAttribute_t attr{};
  const hipError_t status = hipPointerGetAttributes(&attr, static_cast<const void*>(ptr));
  if (status != hipSuccess) {
    (void)hipGetLastError();
    if (status != hipErrorInvalidValue) check_hip(status);
    return {ptr, nullptr};
  }
  if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) {
    return {nullptr, static_cast<P*>(attr.devicePointer)};
  }
  return {attr.hostPointer ? static_cast<P*>(attr.hostPointer) : ptr, nullptr};
}

// SPFFT_AMD_GENERIC=1 forces the generic stage kernels (used to compare the
// two paths).
bool disable_fast_path() {
  const char* v = std::getenv("SPFFT_AMD_GENERIC");
  return v && v[0] == '1';
}

}  // namespace

template <typename T>
ExecutionGPU<T>::ExecutionGPU(const std::shared_ptr<IndexMap>& map,
                              const std::shared_ptr<GridInternal<T>>& grid)
    : map_(map),
      grid_(grid),
      scale_(static_cast<T>(1.0 / static_cast<double>(static_cast<unsigned long long>(map->dimX) *
                                                      static_cast<unsigned long long>(map->dimY) *
                                                      static_cast<unsigned long long>(map->dimZ)))) {
  check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  check_hip(hipEventCreateWithFlags(&startEvent_, hipEventDisableTiming));
  check_hip(hipEventCreateWithFlags(&endEvent_, hipEventDisableTiming));

  const StickTables t = build_stick_tables(*map_);
  numCols_ = static_cast<int>(t.colX.size());
  entryValue_ = DeviceBuffer<int>::from_host(t.entryValue);
  entrySlot_ = DeviceBuffer<int>::from_host(t.entrySlot);
  stickBegin_ = DeviceBuffer<int>::from_host(t.stickBegin);
  stickY_ = DeviceBuffer<int>::from_host(t.stickY);
  stickCol_ = DeviceBuffer<int>::from_host(t.stickCol);
  colX_ = DeviceBuffer<int>::from_host(t.colX);
  colBegin_ = DeviceBuffer<int>::from_host(t.colBegin);

  fftX_ = make_fft_desc(map_->dimX);
  fftY_ = make_fft_desc(map_->dimY);
  fftZ_ = make_fft_desc(map_->dimZ);
  size_t lds = 0;
  zBatch_ = stage_batch<T>(map_->dimZ, std::max(1, map_->num_sticks()), &lds);
  yBatch_ = stage_batch<T>(map_->dimY, std::max(1, numCols_), &lds);
  xBatch_ = stage_batch<T>(map_->dimX, map_->dimY, &lds);

  // Fixed-length kernels when the lengths have compile-time schedules.
  fast_ = !disable_fast_path() &&
          fast_path_supported<T>(map_->dimX, map_->dimY, map_->dimZ, numCols_,
                                 map_->type == SPFFT_TRANS_R2C);
  if (fast_) {
    const int ny = map_->dimY;
    std::vector<int> at(static_cast<size_t>(numCols_) * ny, -1);
    for (int s = 0; s < map_->num_sticks(); ++s) {
      at[static_cast<size_t>(t.stickCol[s]) * ny + t.stickY[s]] = s;
    }
    stickAt_ = DeviceBuffer<int>::from_host(at);
    // Entry list for the z kernels: value index and position in the LDS tile
    // of the workgroup that owns the stick; earlier copies of a repeated
    // triplet are flagged so that the last one wins on scatter.
    int S = 0, LZ = 0, shift = 0;
    fast_z_geometry<T>(map_->dimZ, &S, &LZ, &shift);
    std::vector<int2> ent(map_->num_values());
    for (int e = 0; e < map_->num_values(); ++e) {
      const int slot = t.entrySlot[e];
      const int st = slot / map_->dimZ, z = slot % map_->dimZ;
      int pos = (st % S) * LZ + z + (z >> shift);
      if (e + 1 < map_->num_values() && t.entrySlot[e + 1] == slot) pos |= int(0x80000000u);
      ent[e] = make_int2(t.entryValue[e], pos);
    }
    zEntries_ = DeviceBuffer<int2>::from_host(ent);
    const int classes = fast_classes(map_->dimX);
    std::vector<int> resBegin(classes + 1, 0), resSticks;
    resSticks.reserve(map_->num_sticks());
    for (int c = 0; c < classes; ++c) {
      resBegin[c] = static_cast<int>(resSticks.size());
      for (int s = 0; s < map_->num_sticks(); ++s) {
        if (t.colX[t.stickCol[s]] % classes == c) resSticks.push_back(s);
      }
    }
    resBegin[classes] = static_cast<int>(resSticks.size());
    std::vector<int> resYM(resSticks.size());
    for (size_t i = 0; i < resSticks.size(); ++i) {
      const int st = resSticks[i];
      resYM[i] = (t.stickY[st] << 16) | (t.colX[t.stickCol[st]] / classes);
    }
    resYM_ = DeviceBuffer<int>::from_host(resYM);
    std::vector<int> xCol(map_->dimX, -1);
    for (int u = 0; u < numCols_; ++u) xCol[t.colX[u]] = u;
    xCol_ = DeviceBuffer<int>::from_host(xCol);
    resSticks_ = DeviceBuffer<int>::from_host(resSticks);
    resBegin_ = DeviceBuffer<int>::from_host(resBegin);
  }

  twX_ = DeviceBuffer<Cx<T>>(map_->dimX);
  twY_ = DeviceBuffer<Cx<T>>(map_->dimY);
  twZ_ = DeviceBuffer<Cx<T>>(map_->dimZ);
  generate_twiddles<T>(twX_.data(), map_->dimX, stream_);
  generate_twiddles<T>(twY_.data(), map_->dimY, stream_);
  generate_twiddles<T>(twZ_.data(), map_->dimZ, stream_);
  check_hip(hipStreamSynchronize(stream_));
}

template <typename T>
ExecutionGPU<T>::~ExecutionGPU() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (startEvent_) (void)hipEventDestroy(startEvent_);
  if (endEvent_) (void)hipEventDestroy(endEvent_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

template <typename T>
size_t ExecutionGPU<T>::space_domain_size() const {
  const size_t n = static_cast<size_t>(map_->dimX) * map_->dimY * map_->dimZ;
  return map_->type == SPFFT_TRANS_R2C ? n : 2 * n;
}

template <typename T>
hipStream_t ExecutionGPU<T>::begin(bool concurrent) {
  // An error left by earlier GPU work is reported, not absorbed
  // (reference execution_gpu.cpp:256-258).
  if (hipGetLastError() != hipSuccess) throw GPUPrecedingError();
  if (!concurrent) return nullptr;
  check_hip(hipEventRecord(startEvent_, nullptr));
  check_hip(hipStreamWaitEvent(stream_, startEvent_, 0));
  return stream_;
}

template <typename T>
StageArgs<T> ExecutionGPU<T>::base_args() const {
  StageArgs<T> a;
  a.nx = map_->dimX;
  a.ny = map_->dimY;
  a.nz = map_->dimZ;
  a.nxf = map_->dimXFreq;
  a.nst = map_->num_sticks();
  a.ncol = numCols_;
  a.r2c = map_->type == SPFFT_TRANS_R2C ? 1 : 0;
  a.zeroZeroStick = map_->zeroZeroStick;
  a.entryValue = entryValue_.data();
  a.entrySlot = entrySlot_.data();
  a.stickBegin = stickBegin_.data();
  a.stickY = stickY_.data();
  a.stickCol = stickCol_.data();
  a.colX = colX_.data();
  a.colBegin = colBegin_.data();
  a.zbuf = reinterpret_cast<Cx<T>*>(grid_->sticks_gpu());
  a.ybuf = reinterpret_cast<Cx<T>*>(grid_->work_gpu());
  a.scale = scale_;
  return a;
}

template <typename T>
FastArgs<T> ExecutionGPU<T>::fast_args() const {
  FastArgs<T> a;
  a.nz = map_->dimZ;
  a.nst = map_->num_sticks();
  a.ncol = numCols_;
  a.zbuf = reinterpret_cast<Cx<T>*>(grid_->sticks_gpu());
  a.entryValue = entryValue_.data();
  a.entrySlot = entrySlot_.data();
  a.stickBegin = stickBegin_.data();
  a.stickY = stickY_.data();
  a.stickCol = stickCol_.data();
  a.colX = colX_.data();
  a.colBegin = colBegin_.data();
  a.stickAt = stickAt_.data();
  a.zEntries = zEntries_.data();
  a.pitch = fast_pitch<T>(map_->num_sticks());
  a.r2c = map_->type == SPFFT_TRANS_R2C ? 1 : 0;
  a.zeroZeroStick = map_->zeroZeroStick;
  a.resSticks = resSticks_.data();
  a.resBegin = resBegin_.data();
  a.xCol = xCol_.data();
  a.resYM = resYM_.data();
  a.twX = twX_.data();
  a.twY = twY_.data();
  a.twZ = twZ_.data();
  a.scale = scale_;
  return a;
}

template <typename T>
void ExecutionGPU<T>::backward(const T* input, T* output, bool concurrent) {
  hipStream_t stream = begin(concurrent);
  const size_t numValueReals = 2 * static_cast<size_t>(map_->num_values());
  auto in = translate_pointer(input);
  auto out = translate_pointer(output);
  const T* inGPU = in.second;
  if (numValueReals > 0 && !inGPU) {
    // Host input: stage in the work buffer, which is free until the y stage.
    check_hip(hipMemcpyAsync(grid_->work_gpu(), in.first, numValueReals * sizeof(T),
                             hipMemcpyHostToDevice, stream));
    inGPU = grid_->work_gpu();
  }
  T* outGPU = out.second ? out.second : grid_->space_gpu();

  if (fast_) {
    FastArgs<T> f = fast_args();
    f.in = inGPU;
    f.out = outGPU;
    launch_fast_z_backward<T>(map_->dimZ, f, stream);
    launch_fast_xy_backward<T>(map_->dimX, map_->dimY, f, stream);
    if (out.first) {
      check_hip(hipMemcpyAsync(out.first, outGPU, space_domain_size() * sizeof(T),
                               hipMemcpyDeviceToHost, stream));
    }
    return;
  }

  StageArgs<T> a = base_args();
  a.in = inGPU;
  a.out = outGPU;
  if (a.nst > 0) {
    a.fft = fftZ_;
    a.tw = twZ_.data();
    a.batch = zBatch_;
    launch_z_backward<T>(a, stream);
    a.fft = fftY_;
    a.tw = twY_.data();
    a.batch = yBatch_;
    launch_y_backward<T>(a, stream);
  }
  a.fft = fftX_;
  a.tw = twX_.data();
  a.batch = xBatch_;
  launch_x_backward<T>(a, stream);

  if (out.first) {
    check_hip(hipMemcpyAsync(out.first, outGPU, space_domain_size() * sizeof(T),
                             hipMemcpyDeviceToHost, stream));
  }
}

template <typename T>
void ExecutionGPU<T>::forward(const T* input, T* output, SpfftScalingType scaling,
                              bool concurrent) {
  hipStream_t stream = begin(concurrent);
  const size_t numValueReals = 2 * static_cast<size_t>(map_->num_values());
  auto in = translate_pointer(input);
  auto out = translate_pointer(output);
  const T* inGPU = in.second;
  if (!inGPU) {
    check_hip(hipMemcpyAsync(grid_->space_gpu(), in.first, space_domain_size() * sizeof(T),
                             hipMemcpyHostToDevice, stream));
    inGPU = grid_->space_gpu();
  }
  if (numValueReals == 0) return;
  // Host output: the z stage writes into the work buffer, which is free again
  // once the y stage has consumed it.
  T* outGPU = out.second ? out.second : grid_->work_gpu();

  if (fast_) {
    FastArgs<T> f = fast_args();
    f.in = inGPU;
    f.out = outGPU;
    f.useScale = scaling == SPFFT_FULL_SCALING ? 1 : 0;
    launch_fast_xy_forward<T>(map_->dimX, map_->dimY, f, stream);
    launch_fast_z_forward<T>(map_->dimZ, f, stream);
    if (out.first) {
      check_hip(hipMemcpyAsync(out.first, outGPU, numValueReals * sizeof(T),
                               hipMemcpyDeviceToHost, stream));
    }
    return;
  }

  StageArgs<T> a = base_args();
  a.in = inGPU;
  a.out = outGPU;
  a.useScale = scaling == SPFFT_FULL_SCALING ? 1 : 0;
  a.fft = fftX_;
  a.tw = twX_.data();
  a.batch = xBatch_;
  launch_x_forward<T>(a, stream);
  a.fft = fftY_;
  a.tw = twY_.data();
  a.batch = yBatch_;
  launch_y_forward<T>(a, stream);
  a.fft = fftZ_;
  a.tw = twZ_.data();
  a.batch = zBatch_;
  launch_z_forward<T>(a, stream);

  if (out.first) {
    check_hip(hipMemcpyAsync(out.first, outGPU, numValueReals * sizeof(T), hipMemcpyDeviceToHost,
                             stream));
  }
}

template <typename T>
void ExecutionGPU<T>::synchronize(SpfftExecType mode, bool concurrent) {
  if (!concurrent) {
    if (mode != SPFFT_EXEC_ASYNCHRONOUS) check_hip(hipStreamSynchronize(nullptr));
    return;
  }
  if (mode == SPFFT_EXEC_ASYNCHRONOUS) {
    check_hip(hipEventRecord(endEvent_, stream_));
    check_hip(hipStreamWaitEvent(nullptr, endEvent_, 0));
  } else {
    check_hip(hipStreamSynchronize(stream_));
  }
}

template <typename T>
bool ExecutionGPU<T>::run_batch(const std::vector<ExecutionGPU<T>*>& ex,
                                const std::vector<const T*>& in, const std::vector<T*>& out,
                                bool forward, const std::vector<SpfftScalingType>& scaling) {
  const size_t n = ex.size();
  if (n < 2) return false;
  const IndexMap& m0 = *ex[0]->map_;
  for (const ExecutionGPU<T>* e : ex) {
    const IndexMap& m = *e->map_;
    if (!e->fast_ || m.dimX != m0.dimX || m.dimY != m0.dimY || m.dimZ != m0.dimZ ||
        m.type != m0.type || m.num_values() == 0 || e->grid_->device_id() != ex[0]->grid_->device_id()) {
      return false;
    }
  }
  int S = 0, LZ = 0, shift = 0;
  fast_z_geometry<T>(m0.dimZ, &S, &LZ, &shift);
  std::vector<FastArgs<T>> args(n);
  std::vector<int2> zmap;
  int maxNcol = 0;
  for (size_t t = 0; t < n; ++t) {
    auto i = translate_pointer(in[t]);
    auto o = translate_pointer(out[t]);
    if (!i.second || !o.second) return false;
    FastArgs<T> f = ex[t]->fast_args();
    f.in = i.second;
    f.out = o.second;
    f.useScale = forward && scaling[t] == SPFFT_FULL_SCALING ? 1 : 0;
    f.pitchA = fast_xy_pitch_a<T>(m0.dimX, f.ncol);
    maxNcol = std::max(maxNcol, f.ncol);
    const int blocks = (f.nst + S - 1) / S;
    for (int b = 0; b < blocks; ++b) zmap.push_back(make_int2(static_cast<int>(t), b));
    args[t] = f;
  }
  if (hipGetLastError() != hipSuccess) throw GPUPrecedingError();
  // Argument tables live in stream-ordered allocations on the default stream.
  FastArgs<T>* dArgs = nullptr;
  int2* dMap = nullptr;
  check_hip(hipMallocAsync(reinterpret_cast<void**>(&dArgs), n * sizeof(FastArgs<T>), nullptr));
  check_hip(hipMallocAsync(reinterpret_cast<void**>(&dMap), zmap.size() * sizeof(int2), nullptr));
  check_hip(hipMemcpyAsync(dArgs, args.data(), n * sizeof(FastArgs<T>), hipMemcpyHostToDevice,
                           nullptr));
  check_hip(hipMemcpyAsync(dMap, zmap.data(), zmap.size() * sizeof(int2), hipMemcpyHostToDevice,
                           nullptr));
  const int blocks = static_cast<int>(zmap.size());
  const bool r2c = m0.type == SPFFT_TRANS_R2C;
  const int nt = static_cast<int>(n);
  if (forward) {
    launch_fast_xy_batch<T>(true, r2c, m0.dimX, dArgs, nt, m0.dimZ, maxNcol, nullptr);
    launch_fast_z_batch<T>(true, r2c, m0.dimZ, dArgs, dMap, blocks, nullptr);
  } else {
    launch_fast_z_batch<T>(false, r2c, m0.dimZ, dArgs, dMap, blocks, nullptr);
    launch_fast_xy_batch<T>(false, r2c, m0.dimX, dArgs, nt, m0.dimZ, maxNcol, nullptr);
  }
  check_hip(hipFreeAsync(dArgs, nullptr));
  check_hip(hipFreeAsync(dMap, nullptr));
  return true;
}

template class ExecutionGPU<double>;
template class ExecutionGPU<float>;

}  // namespace spfft
