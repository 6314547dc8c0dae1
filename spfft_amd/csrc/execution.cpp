#include "execution.hpp"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <list>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "hip_check.hpp"
#include "profiling.hpp"
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

// Choice of plane kernels for batches: SPFFT_AMD_PLANE=0 keeps the
// class-split kernels, =1 forces the whole-plane kernels where they apply;
// unset, the autotuner decides. Returns -1 (tune), 0 or 1.
int plane_kernels_forced() {
  const char* v = std::getenv("SPFFT_AMD_PLANE");
  if (v && v[0] == '0') return 0;
  if (v && v[0] == '1') return 1;
  return -1;
}

// Autotuner decisions between the class-split and the whole-plane kernels,
// per (value bytes, length, direction): timed once, on the first qualifying
// batch, on that batch's own buffers. SPFFT_AMD_TUNE_LOG=1 prints them.
struct TuneTable {
  std::mutex mutex;
  std::map<std::tuple<int, int, int>, int> choice;
};
TuneTable& tune_table() {
  static TuneTable* t = new TuneTable();
  return *t;
}
bool tune_log() {
  const char* v = std::getenv("SPFFT_AMD_TUNE_LOG");
  return v && v[0] == '1';
}

// SPFFT_AMD_GRAPHS=0 replays batched calls with direct launches instead of
// captured HIP graphs.
bool graphs_enabled() {
  const char* v = std::getenv("SPFFT_AMD_GRAPHS");
  return !(v && v[0] == '0');
}

std::atomic<std::uint64_t> next_serial{1};

// A batched call's device argument table, block map and captured launch
// chain, keyed by everything they depend on: direction, device, and per
// transform its serial number (never reused, so a destroyed transform's plan
// can never match again), its device input/output pointers and scaling.
struct BatchPlan {
  std::vector<std::uint64_t> key;
  void* args = nullptr;
  void* map = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  bool plane = false;  // plane kernels the graph was captured with
};

// Most recently used first. Heap-allocated and never destroyed: plans hold
// device memory that must not be released after the runtime shuts down.
struct BatchCache {
  std::mutex mutex;
  std::list<BatchPlan> plans;
};
constexpr size_t kMaxBatchPlans = 8;

BatchCache& batch_cache() {
  static BatchCache* cache = new BatchCache();
  return *cache;
}

void release_plan(BatchPlan& p) {
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  if (p.args) (void)hipFree(p.args);
  if (p.map) (void)hipFree(p.map);
  p = BatchPlan{};
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
  serial_ = next_serial.fetch_add(1);
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
    // Stick index at (column, y); absent positions, and the extra row
    // `numCols_` used by the lanes past the last column, point at the zero
    // element that every row of the [z][stick] buffer holds at index nst.
    // Stored class-interleaved: for column u and j < M = Ny / R the R entries
    // of y = j + M l (l < R) are adjacent, so a lane of the backward plane
    // kernel reads all the indices of one decimation-in-frequency output with
    // R / 4 16-byte loads.
    const int ny = map_->dimY;
    const int R = fast_classes(map_->dimX), M = ny / R;
    std::vector<int> at(static_cast<size_t>(numCols_ + 1) * ny, map_->num_sticks());
    for (int s = 0; s < map_->num_sticks(); ++s) {
      const int y = t.stickY[s];
      at[(static_cast<size_t>(t.stickCol[s]) * M + y % M) * R + y / M] = s;
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
    // Sticks whose values form one run in user order: storage z0, z0 + 1, ...
    // (cyclically, e.g. centred z ascending) at consecutive value indices.
    // When every stick is such a run the z kernels address values directly.
    {
      const int nz = map_->dimZ;
      std::vector<int4> runs(map_->num_sticks());
      bool all = map_->num_sticks() > 0;
      for (int st = 0; st < map_->num_sticks() && all; ++st) {
        const int b = t.stickBegin[st], e = t.stickBegin[st + 1];
        const int len = e - b;
        // storage z of the entries ascend (sorted by slot); one run has at
        // most one gap, and starts after it
        int start = 0;
        for (int k = b + 1; k < e; ++k) {
          const int za = t.entrySlot[k - 1] % nz, zb = t.entrySlot[k] % nz;
          if (zb == za) { all = false; break; }  // repeated triplet
          if (zb != za + 1) {
            if (start != 0) { all = false; break; }
            start = k - b;
          }
        }
        if (!all) break;
        const int z0 = t.entrySlot[b + start] % nz;
        const int v0 = t.entryValue[b + start];
        for (int k = 0; k < len && all; ++k) {
          const int kk = b + (start + k) % len;
          const int z = t.entrySlot[kk] % nz;
          if ((z - z0 + nz) % nz != k || t.entryValue[kk] != v0 + k) all = false;
        }
        runs[st] = make_int4(v0, z0, len, 0);
      }
      if (all) zInfo_ = DeviceBuffer<int4>::from_host(runs);
    }
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
    // column of x, or the all-zero column numCols_ of the xy tile
    std::vector<int> xCol(map_->dimX, numCols_);
    for (int u = 0; u < numCols_; ++u) xCol[t.colX[u]] = u;
    xCol_ = DeviceBuffer<int>::from_host(xCol);
    resSticks_ = DeviceBuffer<int>::from_host(resSticks);
    resBegin_ = DeviceBuffer<int>::from_host(resBegin);
    // tile pitch for the whole-plane batch kernels (0: not applicable)
    if (plane_path_supported<T>(map_->dimX, numCols_, map_->type == SPFFT_TRANS_R2C)) {
      pitchT_ = plane_pitch<T>(map_->dimX, numCols_);
    }
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
  a.zInfo = zInfo_.data();
  a.pitchT = pitchT_;
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
  std::vector<const T*> din(n);
  std::vector<T*> dout(n);
  for (size_t t = 0; t < n; ++t) {
    auto i = translate_pointer(in[t]);
    auto o = translate_pointer(out[t]);
    if (!i.second || !o.second) return false;
    din[t] = i.second;
    dout[t] = o.second;
  }
  if (hipGetLastError() != hipSuccess) throw GPUPrecedingError();

  std::vector<std::uint64_t> key;
  key.reserve(3 + 4 * n);
  key.push_back(forward ? 1 : 0);
  key.push_back(static_cast<std::uint64_t>(ex[0]->grid_->device_id()));
  key.push_back(n);
  for (size_t t = 0; t < n; ++t) {
    key.push_back(ex[t]->serial_);
    key.push_back(reinterpret_cast<std::uintptr_t>(din[t]));
    key.push_back(reinterpret_cast<std::uintptr_t>(dout[t]));
    key.push_back(forward && scaling[t] == SPFFT_FULL_SCALING ? 1 : 0);
  }

  int S = 0, LZ = 0, shift = 0;
  fast_z_geometry<T>(m0.dimZ, &S, &LZ, &shift);
  int blocks = 0, maxNcol = 0;
  for (size_t t = 0; t < n; ++t) {
    blocks += (ex[t]->map_->num_sticks() + S - 1) / S;
    maxNcol = std::max(maxNcol, ex[t]->numCols_);
  }
  const bool r2c = m0.type == SPFFT_TRANS_R2C;
  const int nt = static_cast<int>(n);
  // whole-plane kernels are a candidate when every transform has its tile
  // positions and the widest tile fits
  bool planeOk = plane_kernels_forced() != 0 && plane_path_supported<T>(m0.dimX, maxNcol, r2c);
  for (size_t t = 0; t < n && planeOk; ++t) planeOk = ex[t]->pitchT_ > 0;
  bool plane = false;
  auto launch_xy_variant = [&](bool usePlane, const FastArgs<T>* dArgs, hipStream_t s) {
    if (usePlane) {
      launch_plane_xy_batch<T>(forward, m0.dimX, dArgs, nt, m0.dimZ, maxNcol, s);
    } else {
      launch_fast_xy_batch<T>(forward, r2c, m0.dimX, dArgs, nt, m0.dimZ, maxNcol, s);
    }
  };
  auto launch_xy = [&](const FastArgs<T>* dArgs, hipStream_t s) {
    launch_xy_variant(plane, dArgs, s);
  };
  auto launch_chain = [&](const FastArgs<T>* dArgs, const int2* dMap, hipStream_t s) {
    if (forward) {
      launch_xy(dArgs, s);
      launch_fast_z_batch<T>(true, r2c, m0.dimZ, dArgs, dMap, blocks, s);
    } else {
      launch_fast_z_batch<T>(false, r2c, m0.dimZ, dArgs, dMap, blocks, s);
      launch_xy(dArgs, s);
    }
  };

  BatchCache& cache = batch_cache();
  std::lock_guard<std::mutex> lock(cache.mutex);
  auto hit = std::find_if(cache.plans.begin(), cache.plans.end(),
                          [&](const BatchPlan& p) { return p.key == key; });
  if (hit == cache.plans.end()) {
    if (cache.plans.size() >= kMaxBatchPlans) {
      // the evicted plan's tables may still be read by queued work
      check_hip(hipStreamSynchronize(nullptr));
      release_plan(cache.plans.back());
      cache.plans.pop_back();
    }
    std::vector<FastArgs<T>> args(n);
    std::vector<int2> zmap;
    zmap.reserve(blocks);
    for (size_t t = 0; t < n; ++t) {
      FastArgs<T> f = ex[t]->fast_args();
      f.in = din[t];
      f.out = dout[t];
      f.useScale = static_cast<int>(key[3 + 4 * t + 3]);
      f.pitchA = fast_xy_pitch_a<T>(m0.dimX, f.ncol);
      const int b = (f.nst + S - 1) / S;
      for (int k = 0; k < b; ++k) zmap.push_back(make_int2(static_cast<int>(t), k));
      args[t] = f;
    }
    BatchPlan plan;
    plan.key = std::move(key);
    check_hip(hipMalloc(&plan.args, n * sizeof(FastArgs<T>)));
    check_hip(hipMalloc(&plan.map, zmap.size() * sizeof(int2)));
    check_hip(hipMemcpyAsync(plan.args, args.data(), n * sizeof(FastArgs<T>),
                             hipMemcpyHostToDevice, nullptr));
    check_hip(hipMemcpyAsync(plan.map, zmap.data(), zmap.size() * sizeof(int2),
                             hipMemcpyHostToDevice, nullptr));
    cache.plans.push_front(std::move(plan));
    hit = cache.plans.begin();
  } else if (hit != cache.plans.begin()) {
    cache.plans.splice(cache.plans.begin(), cache.plans, hit);
    hit = cache.plans.begin();
  }
  BatchPlan& plan = *hit;
  const auto* dArgs = static_cast<const FastArgs<T>*>(plan.args);
  const auto* dMap = static_cast<const int2*>(plan.map);

  if (planeOk) {
    const int forced = plane_kernels_forced();
    const auto tkey = std::make_tuple(static_cast<int>(sizeof(T)), m0.dimX, forward ? 1 : 0);
    TuneTable& tt = tune_table();
    std::lock_guard<std::mutex> tlock(tt.mutex);
    auto it = tt.choice.find(tkey);
    if (forced == 1) {
      plane = true;
    } else if (it != tt.choice.end()) {
      plane = it->second != 0;
    } else {
      // Time both plane kernels on this batch. The xy launch is idempotent
      // (backward: [z][stick] -> outputs, after the z launch has filled the
      // buffer; forward: inputs -> [z][stick]), and the call's real launch
      // chain follows.
      if (!forward) launch_fast_z_batch<T>(false, r2c, m0.dimZ, dArgs, dMap, blocks, nullptr);
      hipEvent_t e0 = nullptr, e1 = nullptr;
      check_hip(hipEventCreate(&e0));
      check_hip(hipEventCreate(&e1));
      float ms[2] = {0.f, 0.f};
      for (int c = 0; c < 2; ++c) {
        launch_xy_variant(c == 1, dArgs, nullptr);
        check_hip(hipEventRecord(e0, nullptr));
        for (int k = 0; k < 3; ++k) launch_xy_variant(c == 1, dArgs, nullptr);
        check_hip(hipEventRecord(e1, nullptr));
        check_hip(hipEventSynchronize(e1));
        check_hip(hipEventElapsedTime(&ms[c], e0, e1));
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      plane = ms[1] < ms[0];
      tt.choice[tkey] = plane ? 1 : 0;
      if (tune_log()) {
        std::fprintf(stderr, "spfft_amd tune: %s N=%d %s, %zu transforms: class-split %.3f ms, "
                     "whole-plane %.3f ms -> %s\n", sizeof(T) == 8 ? "double" : "float",
                     m0.dimX, forward ? "forward" : "backward", n, ms[0] / 3, ms[1] / 3,
                     plane ? "whole-plane" : "class-split");
      }
    }
  }
  if (plan.exec && plan.plane != plane) {
    (void)hipGraphExecDestroy(plan.exec);
    (void)hipGraphDestroy(plan.graph);
    plan.exec = nullptr;
    plan.graph = nullptr;
  }
  if (!graphs_enabled() || prof::enabled()) {
    launch_chain(dArgs, dMap, nullptr);
    return true;
  }
  if (!plan.exec) {
    plan.plane = plane;
    // Capture the two launches once on a private stream; later calls with
    // the same key replay the graph on the default stream.
    hipStream_t cs = nullptr;
    check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    try {
      launch_chain(dArgs, dMap, cs);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(cs, &g);
      if (g) (void)hipGraphDestroy(g);
      (void)hipStreamDestroy(cs);
      throw;
    }
    check_hip(hipStreamEndCapture(cs, &plan.graph));
    check_hip(hipStreamDestroy(cs));
    check_hip(hipGraphInstantiate(&plan.exec, plan.graph, nullptr, nullptr, 0));
  }
  check_hip(hipGraphLaunch(plan.exec, nullptr));
  return true;
}

template class ExecutionGPU<double>;
template class ExecutionGPU<float>;

}  // namespace spfft
