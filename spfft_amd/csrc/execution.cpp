#include "execution.hpp"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <list>
#include <unordered_map>
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

// A user pointer classified for the copy path: device memory is used in
// place, pinned host memory is copied by DMA directly, pageable host memory
// goes through the grid's pinned ring.
enum class MemKind { Device, Pinned, Pageable };
template <typename P>
MemKind classify(P* ptr, P** view) {
  hipPointerAttribute_t attr{};
  const hipError_t status = hipPointerGetAttributes(&attr, static_cast<const void*>(ptr));
  if (status != hipSuccess) {
    (void)hipGetLastError();
    if (status != hipErrorInvalidValue) check_hip(status);
    *view = ptr;
    return MemKind::Pageable;
  }
  if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) {
    *view = static_cast<P*>(attr.devicePointer);
    return MemKind::Device;
  }
  *view = attr.hostPointer ? static_cast<P*>(attr.hostPointer) : ptr;
  return attr.type == hipMemoryTypeHost ? MemKind::Pinned : MemKind::Pageable;
}

// Host transfers are split into chunks of about this size (whole planes for
// the space domain), so copies overlap the kernels and, for pageable memory,
// the host-side copies overlap the DMA.
constexpr size_t kIoChunkBytes = size_t(4) << 20;

// SPFFT_AMD_GENERIC=1 forces the generic stage kernels (used to compare the
// two paths).
bool disable_fast_path() {
  const char* v = std::getenv("SPFFT_AMD_GENERIC");
  return v && v[0] == '1';
}

// Choice of plane kernels: SPFFT_AMD_PLANE=0 keeps the class-split kernels,
// =1 forces the whole-plane kernels where they apply (batches), =2 the split
// plane stage; unset, the autotuner decides. Returns -1 (tune), 0, 1 or 2. A
// forced family skips the tuner (defaults for the other choices, nothing
// recorded).
int plane_kernels_forced() {
  const char* v = std::getenv("SPFFT_AMD_PLANE");
  if (v && v[0] == '0') return 0;
  if (v && v[0] == '1') return 1;
  if (v && v[0] == '2') return 2;
  return -1;
}

// Mean time of `reps` launches after one warm-up launch, in milliseconds.
template <typename F>
float time_launches(F&& launch, hipEvent_t e0, hipEvent_t e1, int reps = 3) {
  launch();
  check_hip(hipEventRecord(e0, nullptr));
  for (int k = 0; k < reps; ++k) launch();
  check_hip(hipEventRecord(e1, nullptr));
  check_hip(hipEventSynchronize(e1));
  float ms = 0.f;
  check_hip(hipEventElapsedTime(&ms, e0, e1));
  return ms / static_cast<float>(reps);
}

// Two timing events for one tuning session.
struct TuneEvents {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TuneEvents() {
    check_hip(hipEventCreate(&e0));
    check_hip(hipEventCreate(&e1));
  }
  ~TuneEvents() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  TuneEvents(const TuneEvents&) = delete;
  TuneEvents& operator=(const TuneEvents&) = delete;
};

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
  TuneChoice choice;  // kernel choices the graph was captured with
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

// Device tables of one index set: built once per (device, precision, set,
// path) and shared by every transform with an identical set, so bands over
// one set hold, and a batch reads, one copy.
template <typename T>
std::shared_ptr<TransformTables<T>> build_tables(const std::shared_ptr<IndexMap>& map, bool fast,
                                                 int device, hipStream_t stream) {
  auto tabPtr = std::make_shared<TransformTables<T>>();
  TransformTables<T>& tb = *tabPtr;
  tb.map = map;
  tb.device = device;
  tb.allowFast = fast;
  const StickTables t = build_stick_tables(*map);
  tb.numCols = static_cast<int>(t.colX.size());
  // fixed-length kernels when the lengths have compile-time schedules
  tb.fast = fast && fast_path_supported<T>(map->dimX, map->dimY, map->dimZ, tb.numCols,
                                           map->type == SPFFT_TRANS_R2C);
  tb.classSplit = tb.fast && class_split_supported<T>(map->dimX, map->dimY, tb.numCols,
                                                      map->type == SPFFT_TRANS_R2C);
  const bool useFast = tb.fast;
  tb.entryValue = DeviceBuffer<int>::from_host(t.entryValue);
  tb.entrySlot = DeviceBuffer<int>::from_host(t.entrySlot);
  tb.stickBegin = DeviceBuffer<int>::from_host(t.stickBegin);
  tb.stickY = DeviceBuffer<int>::from_host(t.stickY);
  tb.stickCol = DeviceBuffer<int>::from_host(t.stickCol);
  tb.colX = DeviceBuffer<int>::from_host(t.colX);
  tb.colBegin = DeviceBuffer<int>::from_host(t.colBegin);

  if (useFast) {
    // Stick index at (column, y); absent positions, and the extra row
    // `tb.numCols` used by the lanes past the last column, point at the zero
    // element that every row of the [z][stick] buffer holds at index nst.
    // Stored class-interleaved: for column u and j < M = Ny / R the R entries
    // of y = j + M l (l < R) are adjacent, so a lane of the backward plane
    // kernel reads all the indices of one decimation-in-frequency output with
    // R / 4 16-byte loads.
    const int ny = map->dimY;
    const int R = fast_classes(ny), M = ny / R;
    std::vector<int> at(static_cast<size_t>(tb.numCols + 1) * ny, map->num_sticks());
    for (int s = 0; s < map->num_sticks(); ++s) {
      const int y = t.stickY[s];
      at[(static_cast<size_t>(t.stickCol[s]) * M + y % M) * R + y / M] = s;
    }
    tb.stickAt = DeviceBuffer<int>::from_host(at);
    // Entry list for the z kernels: value index and position in the LDS tile
    // of the workgroup that owns the stick; earlier copies of a repeated
    // triplet are flagged so that the last one wins on scatter.
    int S = 0, LZ = 0, shift = 0;
    fast_z_geometry<T>(map->dimZ, &S, &LZ, &shift);
    std::vector<int2> ent(map->num_values());
    for (int e = 0; e < map->num_values(); ++e) {
      const int slot = t.entrySlot[e];
      const int st = slot / map->dimZ, z = slot % map->dimZ;
      int pos = (st % S) * LZ + z + (z >> shift);
      if (e + 1 < map->num_values() && t.entrySlot[e + 1] == slot) pos |= int(0x80000000u);
      ent[e] = make_int2(t.entryValue[e], pos);
    }
    tb.zEntries = DeviceBuffer<int2>::from_host(ent);
    // Sticks whose values form one run in user order: storage z0, z0 + 1, ...
    // (cyclically, e.g. centred z ascending) at consecutive value indices.
    // When every stick is such a run the z kernels address values directly.
    {
      const int nz = map->dimZ;
      std::vector<int4> runs(map->num_sticks());
      bool all = map->num_sticks() > 0;
      for (int st = 0; st < map->num_sticks() && all; ++st) {
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
      if (all) tb.zInfo = DeviceBuffer<int4>::from_host(runs);
    }
    const int classes = fast_classes(map->dimX);
    std::vector<int> resBegin(classes + 1, 0), resSticks;
    resSticks.reserve(map->num_sticks());
    for (int c = 0; c < classes; ++c) {
      resBegin[c] = static_cast<int>(resSticks.size());
      for (int s = 0; s < map->num_sticks(); ++s) {
        if (t.colX[t.stickCol[s]] % classes == c) resSticks.push_back(s);
      }
    }
    resBegin[classes] = static_cast<int>(resSticks.size());
    std::vector<int> resYM(resSticks.size());
    for (size_t i = 0; i < resSticks.size(); ++i) {
      const int st = resSticks[i];
      resYM[i] = (t.stickY[st] << 16) | (t.colX[t.stickCol[st]] / classes);
    }
    tb.resYM = DeviceBuffer<int>::from_host(resYM);
    // column of x, or the all-zero column tb.numCols of the xy tile
    std::vector<int> xCol(map->dimX, tb.numCols);
    for (int u = 0; u < tb.numCols; ++u) xCol[t.colX[u]] = u;
    tb.xCol = DeviceBuffer<int>::from_host(xCol);
    tb.resSticks = DeviceBuffer<int>::from_host(resSticks);
    tb.resBegin = DeviceBuffer<int>::from_host(resBegin);
    // tile pitch for the whole-plane batch kernels (0: not applicable)
    if (map->dimX == map->dimY &&
        plane_path_supported<T>(map->dimX, tb.numCols, map->type == SPFFT_TRANS_R2C)) {
      tb.pitchT = plane_pitch<T>(map->dimX, tb.numCols);
    }
  }

  tb.twX = DeviceBuffer<Tw<T>>(map->dimX);
  tb.twY = DeviceBuffer<Tw<T>>(map->dimY);
  tb.twZ = DeviceBuffer<Tw<T>>(map->dimZ);
  generate_twiddles<T>(tb.twX.data(), map->dimX, stream);
  generate_twiddles<T>(tb.twY.data(), map->dimY, stream);
  generate_twiddles<T>(tb.twZ.data(), map->dimZ, stream);
  check_hip(hipStreamSynchronize(stream));
  return tabPtr;
}


// Hash of everything the tables depend on.
std::uint64_t table_key(const IndexMap& m, bool fast, int device, size_t realBytes) {
  std::uint64_t h = 1469598103934665603ull;
  auto mix = [&h](std::uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  };
  mix(static_cast<std::uint64_t>(m.dimX));
  mix(static_cast<std::uint64_t>(m.dimY));
  mix(static_cast<std::uint64_t>(m.dimZ));
  mix(static_cast<std::uint64_t>(m.type));
  mix(fast ? 1 : 0);
  mix(static_cast<std::uint64_t>(device));
  mix(realBytes);
  for (int v : m.valueSlot) mix(static_cast<std::uint32_t>(v));
  for (int v : m.stickKey) mix(static_cast<std::uint32_t>(v));
  return h;
}

bool same_set(const IndexMap& a, const IndexMap& b) {
  return a.dimX == b.dimX && a.dimY == b.dimY && a.dimZ == b.dimZ && a.type == b.type &&
         a.valueSlot == b.valueSlot && a.stickKey == b.stickKey;
}

template <typename T>
std::shared_ptr<TransformTables<T>> acquire_tables(const std::shared_ptr<IndexMap>& map,
                                                   bool fast, int device, hipStream_t stream) {
  static std::mutex mutex;
  static auto* registry =
      new std::unordered_multimap<std::uint64_t, std::weak_ptr<TransformTables<T>>>();
  const std::uint64_t key = table_key(*map, fast, device, sizeof(T));
  std::lock_guard<std::mutex> lock(mutex);
  auto range = registry->equal_range(key);
  for (auto it = range.first; it != range.second;) {
    std::shared_ptr<TransformTables<T>> t = it->second.lock();
    if (!t) {
      it = registry->erase(it);
      continue;
    }
    if (t->device == device && t->allowFast == fast && same_set(*t->map, *map)) return t;
    ++it;
  }
  std::shared_ptr<TransformTables<T>> t = build_tables<T>(map, fast, device, stream);
  registry->emplace(key, t);
  return t;
}

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

  const bool fast = !disable_fast_path();
  tab_ = acquire_tables<T>(map_, fast, grid_->device_id(), stream_);
  numCols_ = tab_->numCols;
  fast_ = tab_->fast;
  classSplit_ = tab_->classSplit;
  pitchT_ = tab_->pitchT;

  fftX_ = make_fft_desc(map_->dimX);
  fftY_ = make_fft_desc(map_->dimY);
  fftZ_ = make_fft_desc(map_->dimZ);
  size_t lds = 0;
  zBatch_ = stage_batch<T>(map_->dimZ, std::max(1, map_->num_sticks()), &lds);
  yBatch_ = stage_batch<T>(map_->dimY, std::max(1, numCols_), &lds);
  xBatch_ = stage_batch<T>(map_->dimX, map_->dimY, &lds);
}

template <typename T>
ExecutionGPU<T>::~ExecutionGPU() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (hipEvent_t e : ioEvents_) {
    if (e) (void)hipEventDestroy(e);
  }
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
  a.entryValue = tab_->entryValue.data();
  a.entrySlot = tab_->entrySlot.data();
  a.stickBegin = tab_->stickBegin.data();
  a.stickY = tab_->stickY.data();
  a.stickCol = tab_->stickCol.data();
  a.colX = tab_->colX.data();
  a.colBegin = tab_->colBegin.data();
  a.zbuf = reinterpret_cast<Cx<T>*>(grid_->sticks_gpu());
  a.ybuf = reinterpret_cast<Cx<T>*>(grid_->work_gpu());
  a.scale = scale_;
  return a;
}

template <typename T>
FastArgs<T> ExecutionGPU<T>::fast_args() const {
  FastArgs<T> a;
  a.nz = map_->dimZ;
  a.ny = map_->dimY;
  a.nst = map_->num_sticks();
  a.ncol = numCols_;
  a.zbuf = reinterpret_cast<Cx<T>*>(grid_->sticks_gpu());
  a.entryValue = tab_->entryValue.data();
  a.entrySlot = tab_->entrySlot.data();
  a.stickBegin = tab_->stickBegin.data();
  a.stickY = tab_->stickY.data();
  a.stickCol = tab_->stickCol.data();
  a.colX = tab_->colX.data();
  a.colBegin = tab_->colBegin.data();
  a.stickAt = tab_->stickAt.data();
  a.zEntries = tab_->zEntries.data();
  a.pitch = fast_pitch<T>(map_->num_sticks());
  a.r2c = map_->type == SPFFT_TRANS_R2C ? 1 : 0;
  a.zeroZeroStick = map_->zeroZeroStick;
  a.resSticks = tab_->resSticks.data();
  a.resBegin = tab_->resBegin.data();
  a.xCol = tab_->xCol.data();
  a.zInfo = tab_->zInfo.data();
  a.hasRuns = tab_->zInfo.data() != nullptr ? 1 : 0;
  a.pitchT = pitchT_;
  a.resYM = tab_->resYM.data();
  a.ybuf = reinterpret_cast<Cx<T>*>(grid_->work_gpu());
  a.pitchY = split_pitch<T>(numCols_, map_->dimX);
  // the split stage's column buffer must fit the grid's work buffer
  if (static_cast<size_t>(map_->dimZ) * map_->dimY * a.pitchY >
      static_cast<size_t>(grid_->max_dim_x()) * grid_->max_dim_y() * grid_->max_dim_z()) {
    throw InternalError();
  }
  a.twX = tab_->twX.data();
  a.twY = tab_->twY.data();
  a.twZ = tab_->twZ.data();
  a.scale = scale_;
  return a;
}

template <typename T>
TuneChoice ExecutionGPU<T>::single_choice(bool forward, const FastArgs<T>& f, hipStream_t stream,
                                          bool staged) {
  TuneChoice c;
  c.zRuns = f.hasRuns;
  const bool classSplit = classSplit_;
  const int forced = plane_kernels_forced();
  if (forced >= 0 || !classSplit) {
    // single transforms have no whole-plane kernels; x != y planes and
    // lengths without class-split kernels only the split stage
    c.plane = forced == 2 || !classSplit ? 2 : 0;
    if (forced >= 0) return c;
  }
  // batched fallbacks on the transform's own stream keep the defaults
  if (stream != nullptr) return c;
  const TuneKey key = make_tune_key(grid_->device_id(), static_cast<int>(sizeof(T)), map_->dimX,
                                    map_->dimY, map_->dimZ, map_->type == SPFFT_TRANS_R2C,
                                    forward, numCols_, 1);
  if (tune_lookup(key, &c)) {
    c.zRuns = c.zRuns && f.hasRuns;
    if (c.plane == 1) c.plane = 0;
    if (!classSplit) c.plane = 2;
    return c;
  }
  // Host data staged in the work buffer, which the split stage's candidate
  // launches would overwrite: no timing on this call (defaults).
  if (staged) {
    c = TuneChoice();
    c.zRuns = f.hasRuns;
    c.plane = classSplit ? 0 : 2;
    return c;
  }
  // Every launch below is idempotent: backward z reads the values and writes
  // the [z][stick] buffer, the plane stage reads that buffer and writes the
  // output; forward the plane stage reads the input and writes the buffer,
  // z reads it and writes the values. The call's real launches follow.
  TuneEvents ev;
  FastArgs<T> a = f;
  a.zBegin = 0;
  a.nz = map_->dimZ;
  auto z = [&](int runs) {
    a.hasRuns = runs;
    if (forward) {
      launch_fast_z_forward<T>(map_->dimZ, a, nullptr);
    } else {
      launch_fast_z_backward<T>(map_->dimZ, a, nullptr);
    }
  };
  auto xy_fused = [&] {
    if (forward) {
      launch_fast_xy_forward<T>(map_->dimX, map_->dimY, a, nullptr);
    } else {
      launch_fast_xy_backward<T>(map_->dimX, map_->dimY, a, nullptr);
    }
  };
  auto xy_two = [&] { launch_split_xy<T>(forward, map_->dimX, map_->dimY, a, nullptr); };
  float zr = 0.f, ze = 0.f, fused = 0.f, two = 0.f;
  auto time_z = [&] {
    if (f.hasRuns) {
      zr = time_launches([&] { z(1); }, ev.e0, ev.e1);
      ze = time_launches([&] { z(0); }, ev.e0, ev.e1);
      c.zRuns = zr <= ze ? 1 : 0;
    } else {
      z(0);
    }
  };
  auto time_xy = [&] {
    two = time_launches(xy_two, ev.e0, ev.e1);
    if (classSplit) {
      fused = time_launches(xy_fused, ev.e0, ev.e1);
      c.plane = two < fused ? 2 : 0;
    } else {
      c.plane = 2;
    }
  };
  if (forward) {
    time_xy();
    time_z();
  } else {
    time_z();
    time_xy();
  }
  tune_record(key, c);
  if (tune_log_enabled()) {
    std::fprintf(stderr, "spfft_amd tune: %s N=%d Nz=%d %s %s, single transform, %d columns: z "
                 "runs %.4f / entries %.4f ms; class-split %.4f, split stage %.4f ms -> plane=%d "
                 "zruns=%d\n", sizeof(T) == 8 ? "double" : "float", map_->dimX, map_->dimZ,
                 map_->type == SPFFT_TRANS_R2C ? "R2C" : "C2C", forward ? "forward" : "backward",
                 numCols_, zr, ze, fused, two, c.plane, c.zRuns);
  }
  return c;
}

template <typename T>
bool ExecutionGPU<T>::device_accessible(const void* ptr) {
  const void* view = nullptr;
  return classify(ptr, &view) == MemKind::Device;
}

template <typename T>
void ExecutionGPU<T>::io_prepare(size_t slotBytes) {
  if (!ioEvents_[0]) {
    for (hipEvent_t& e : ioEvents_) check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  ring_ = static_cast<char*>(grid_->host_staging(2 * slotBytes));
  slotBytes_ = slotBytes;
}

template <typename T>
int ExecutionGPU<T>::planes_per_chunk(size_t planeBytes) const {
  const int nz = map_->dimZ;
  int ppc = static_cast<int>(std::max<size_t>(1, kIoChunkBytes / std::max<size_t>(1, planeBytes)));
  ppc = std::max(ppc, (nz + kIoMaxChunks - 1) / kIoMaxChunks);
  return std::min(ppc, std::max(1, nz));
}

template <typename T>
void ExecutionGPU<T>::upload_piece(char* dst, const char* src, size_t bytes, bool pinned, int k,
                                   hipStream_t copy) {
  hipEvent_t* done = &ioEvents_[kIoMaxChunks];
  if (pinned) {
    check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, copy));
  } else {
    // the slot is free once the DMA of the piece two earlier has read it
    char* r = ring_ + (k & 1) * slotBytes_;
    if (k >= 2) check_hip(hipEventSynchronize(done[(k - 2) % kIoMaxChunks]));
    std::memcpy(r, src, bytes);
    check_hip(hipMemcpyAsync(dst, r, bytes, hipMemcpyHostToDevice, copy));
  }
  check_hip(hipEventRecord(done[k % kIoMaxChunks], copy));
}

template <typename T>
void ExecutionGPU<T>::upload(char* dst, const char* src, size_t bytes, bool pinned,
                             hipStream_t compute, hipStream_t copy) {
  hipEvent_t* done = &ioEvents_[kIoMaxChunks];
  const int pieces = static_cast<int>((bytes + slotBytes_ - 1) / slotBytes_);
  for (int k = 0; k < pieces; ++k) {
    const size_t off = static_cast<size_t>(k) * slotBytes_;
    upload_piece(dst + off, src + off, std::min(slotBytes_, bytes - off), pinned, k, copy);
  }
  if (pieces > 0 && copy != compute) {
    check_hip(hipStreamWaitEvent(compute, done[(pieces - 1) % kIoMaxChunks], 0));
  }
  // the ring is free again when the call returns (a later call refills it)
  if (pieces > 0 && !pinned) check_hip(hipEventSynchronize(done[(pieces - 1) % kIoMaxChunks]));
}

template <typename T>
void ExecutionGPU<T>::download(char* dst, const char* src, size_t bytes, size_t pieceBytes,
                               bool pinned, const hipEvent_t* ready, int readyCount,
                               hipStream_t compute, hipStream_t copy) {
  hipEvent_t* done = &ioEvents_[kIoMaxChunks];
  const int pieces = static_cast<int>((bytes + pieceBytes - 1) / pieceBytes);
  auto range = [&](int k, size_t* off, size_t* len) {
    *off = static_cast<size_t>(k) * pieceBytes;
    *len = std::min(pieceBytes, bytes - *off);
  };
  auto drain = [&](int k) {
    size_t off = 0, len = 0;
    range(k, &off, &len);
    check_hip(hipEventSynchronize(done[k % kIoMaxChunks]));
    std::memcpy(dst + off, ring_ + (k & 1) * slotBytes_, len);
  };
  for (int k = 0; k < pieces; ++k) {
    size_t off = 0, len = 0;
    range(k, &off, &len);
    if (copy != compute) {
      check_hip(hipStreamWaitEvent(copy, ready[std::min(k, readyCount - 1)], 0));
    }
    if (pinned) {
      check_hip(hipMemcpyAsync(dst + off, src + off, len, hipMemcpyDeviceToHost, copy));
    } else {
      if (k >= 2) drain(k - 2);  // frees the slot
      check_hip(hipMemcpyAsync(ring_ + (k & 1) * slotBytes_, src + off, len,
                               hipMemcpyDeviceToHost, copy));
      check_hip(hipEventRecord(done[k % kIoMaxChunks], copy));
    }
  }
  if (pinned) {
    if (copy != compute && pieces > 0) {
      check_hip(hipEventRecord(done[0], copy));
      check_hip(hipStreamWaitEvent(compute, done[0], 0));
    }
  } else {
    for (int k = std::max(0, pieces - 2); k < pieces; ++k) drain(k);
  }
}

template <typename T>
void ExecutionGPU<T>::backward(const T* input, T* output, bool concurrent) {
  hipStream_t stream = begin(concurrent);
  // copies go to the transform's own stream, beside the kernels on the
  // default stream (batched fallbacks run both on the transform's stream)
  hipStream_t copy = concurrent ? stream : stream_;
  const size_t valueBytes = 2 * static_cast<size_t>(map_->num_values()) * sizeof(T);
  const T* inView = nullptr;
  T* outView = nullptr;
  const MemKind inKind = classify(input, &inView);
  const MemKind outKind = classify(output, &outView);
  const bool hostOut = outKind != MemKind::Device;
  const size_t planeBytes = space_domain_size() / map_->dimZ * sizeof(T);
  const int ppc = hostOut ? planes_per_chunk(planeBytes) : map_->dimZ;
  const int chunks = (map_->dimZ + ppc - 1) / ppc;
  if (inKind != MemKind::Device || hostOut) {
    io_prepare(std::max(kIoChunkBytes, static_cast<size_t>(ppc) * planeBytes));
    if (copy != stream) {
      // the copy stream must not touch the buffers before earlier work ends
      check_hip(hipEventRecord(startEvent_, stream));
      check_hip(hipStreamWaitEvent(copy, startEvent_, 0));
    }
  }
  const T* inGPU = inView;
  if (valueBytes > 0 && inKind != MemKind::Device) {
    // host values: staged into the work buffer, free until the plane stage
    upload(reinterpret_cast<char*>(grid_->work_gpu()), reinterpret_cast<const char*>(inView),
           valueBytes, inKind == MemKind::Pinned, stream, copy);
    inGPU = grid_->work_gpu();
  }
  T* outGPU = hostOut ? grid_->space_gpu() : outView;
  hipEvent_t* ready = &ioEvents_[0];

  if (fast_) {
    FastArgs<T> f = fast_args();
    f.in = inGPU;
    f.out = outGPU;
    const TuneChoice choice =
        single_choice(false, f, stream, inKind != MemKind::Device || hostOut);
    f.hasRuns = choice.zRuns;
    launch_fast_z_backward<T>(map_->dimZ, f, stream);
    // host output: the plane stage runs in chunks of planes, each copied out
    // while the next is computed
    for (int c = 0; c < chunks; ++c) {
      f.zBegin = c * ppc;
      f.nz = std::min(map_->dimZ, (c + 1) * ppc);
      if (choice.plane == 2) {
        launch_split_xy<T>(false, map_->dimX, map_->dimY, f, stream);
      } else {
        launch_fast_xy_backward<T>(map_->dimX, map_->dimY, f, stream);
      }
      if (hostOut) check_hip(hipEventRecord(ready[c], stream));
    }
  } else {
    StageArgs<T> a = base_args();
    a.in = inGPU;
    a.out = outGPU;
    if (a.nst > 0) {
      a.fft = fftZ_;
      a.tw = tab_->twZ.data();
      a.batch = zBatch_;
      launch_z_backward<T>(a, stream);
      a.fft = fftY_;
      a.tw = tab_->twY.data();
      a.batch = yBatch_;
      launch_y_backward<T>(a, stream);
    }
    a.fft = fftX_;
    a.tw = tab_->twX.data();
    a.batch = xBatch_;
    launch_x_backward<T>(a, stream);
    if (hostOut) {
      for (int c = 0; c < chunks; ++c) check_hip(hipEventRecord(ready[c], stream));
    }
  }
  if (hostOut) {
    download(reinterpret_cast<char*>(outView), reinterpret_cast<const char*>(outGPU),
             space_domain_size() * sizeof(T), static_cast<size_t>(ppc) * planeBytes,
             outKind == MemKind::Pinned, ready, chunks, stream, copy);
  }
}

template <typename T>
void ExecutionGPU<T>::forward(const T* input, T* output, SpfftScalingType scaling,
                              bool concurrent) {
  hipStream_t stream = begin(concurrent);
  hipStream_t copy = concurrent ? stream : stream_;
  const size_t valueBytes = 2 * static_cast<size_t>(map_->num_values()) * sizeof(T);
  const T* inView = nullptr;
  T* outView = nullptr;
  const MemKind inKind = classify(input, &inView);
  const MemKind outKind = classify(output, &outView);
  if (valueBytes == 0) return;
  const bool hostIn = inKind != MemKind::Device;
  const size_t planeBytes = space_domain_size() / map_->dimZ * sizeof(T);
  const int ppc = hostIn ? planes_per_chunk(planeBytes) : map_->dimZ;
  const int chunks = (map_->dimZ + ppc - 1) / ppc;
  const size_t chunkBytes = static_cast<size_t>(ppc) * planeBytes;
  const size_t spaceBytes = space_domain_size() * sizeof(T);
  if (hostIn || outKind != MemKind::Device) {
    io_prepare(std::max(kIoChunkBytes, chunkBytes));
    if (copy != stream) {
      check_hip(hipEventRecord(startEvent_, stream));
      check_hip(hipStreamWaitEvent(copy, startEvent_, 0));
    }
  }
  const T* inGPU = hostIn ? grid_->space_gpu() : inView;
  // Host output: the z stage writes into the work buffer, which is free again
  // once the plane stage has consumed its input.
  T* outGPU = outKind == MemKind::Device ? outView : grid_->work_gpu();
  hipEvent_t* done = &ioEvents_[kIoMaxChunks];
  auto upload_chunk = [&](int c) {
    const size_t off = static_cast<size_t>(c) * chunkBytes;
    upload_piece(reinterpret_cast<char*>(grid_->space_gpu()) + off,
                 reinterpret_cast<const char*>(inView) + off, std::min(chunkBytes, spaceBytes - off),
                 inKind == MemKind::Pinned, c, copy);
    if (copy != stream) check_hip(hipStreamWaitEvent(stream, done[c % kIoMaxChunks], 0));
  };

  if (fast_) {
    FastArgs<T> f = fast_args();
    f.in = inGPU;
    f.out = outGPU;
    f.useScale = scaling == SPFFT_FULL_SCALING ? 1 : 0;
    const TuneChoice choice =
        single_choice(true, f, stream, hostIn || outKind != MemKind::Device);
    // host input: each chunk of planes is transformed as soon as it arrives
    for (int c = 0; c < chunks; ++c) {
      if (hostIn) upload_chunk(c);
      f.zBegin = c * ppc;
      f.nz = std::min(map_->dimZ, (c + 1) * ppc);
      if (choice.plane == 2) {
        launch_split_xy<T>(true, map_->dimX, map_->dimY, f, stream);
      } else {
        launch_fast_xy_forward<T>(map_->dimX, map_->dimY, f, stream);
      }
    }
    f.zBegin = 0;
    f.nz = map_->dimZ;
    f.hasRuns = choice.zRuns;
    launch_fast_z_forward<T>(map_->dimZ, f, stream);
  } else {
    if (hostIn) {
      for (int c = 0; c < chunks; ++c) upload_chunk(c);
    }
  }
  if (hostIn && inKind == MemKind::Pageable) {
    // the ring is free again when the call returns (a later call refills it)
    check_hip(hipEventSynchronize(done[(chunks - 1) % kIoMaxChunks]));
  }
  if (!fast_) {
    StageArgs<T> a = base_args();
    a.in = inGPU;
    a.out = outGPU;
    a.useScale = scaling == SPFFT_FULL_SCALING ? 1 : 0;
    a.fft = fftX_;
    a.tw = tab_->twX.data();
    a.batch = xBatch_;
    launch_x_forward<T>(a, stream);
    a.fft = fftY_;
    a.tw = tab_->twY.data();
    a.batch = yBatch_;
    launch_y_forward<T>(a, stream);
    a.fft = fftZ_;
    a.tw = tab_->twZ.data();
    a.batch = zBatch_;
    launch_z_forward<T>(a, stream);
  }
  if (outKind != MemKind::Device) {
    hipEvent_t* ready = &ioEvents_[0];
    check_hip(hipEventRecord(ready[0], stream));
    download(reinterpret_cast<char*>(outView), reinterpret_cast<const char*>(outGPU), valueBytes,
             slotBytes_, outKind == MemKind::Pinned, ready, 1, stream, copy);
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
bool ExecutionGPU<T>::batch_qualifies(ExecutionGPU<T>* const* ex, size_t n, const T* const* in,
                                      T* const* out, const T** din, T** dout,
                                      BatchShape* shape) {
  if (n < 1) return false;
  const IndexMap& m0 = *ex[0]->map_;
  for (size_t t = 0; t < n; ++t) {
    const IndexMap& m = *ex[t]->map_;
    if (!ex[t]->fast_ || m.dimX != m0.dimX || m.dimY != m0.dimY || m.dimZ != m0.dimZ ||
        m.type != m0.type || m.num_values() == 0 ||
        ex[t]->grid_->device_id() != ex[0]->grid_->device_id()) {
      return false;
    }
  }
  for (size_t t = 0; t < n; ++t) {
    auto i = translate_pointer(in[t]);
    auto o = translate_pointer(out[t]);
    if (!i.second || !o.second) return false;
    din[t] = i.second;
    dout[t] = o.second;
  }
  int S = 0, LZ = 0, shift = 0;
  fast_z_geometry<T>(m0.dimZ, &S, &LZ, &shift);
  shape->blocks = 0;
  shape->maxNcol = 0;
  shape->runs = true;
  shape->classSplit = true;
  for (size_t t = 0; t < n; ++t) {
    shape->classSplit = shape->classSplit && ex[t]->classSplit_;
    shape->blocks += (ex[t]->map_->num_sticks() + S - 1) / S;
    shape->maxNcol = std::max(shape->maxNcol, ex[t]->numCols_);
    shape->runs = shape->runs && ex[t]->tab_->zInfo.data() != nullptr;
  }
  shape->r2c = m0.type == SPFFT_TRANS_R2C;
  shape->dimX = m0.dimX;
  shape->dimY = m0.dimY;
  shape->dimZ = m0.dimZ;
  shape->device = ex[0]->grid_->device_id();
  return true;
}

template <typename T>
bool ExecutionGPU<T>::plane_candidate(ExecutionGPU<T>* const* ex, size_t n, const T* const* din,
                                      T* const* dout, bool forward, const BatchShape& shape) {
  // whole-plane kernels are a candidate when every transform has its tile
  // positions and the widest tile fits
  bool ok = plane_kernels_forced() != 0 && shape.dimX == shape.dimY &&
            plane_path_supported<T>(shape.dimX, shape.maxNcol, shape.r2c);
  for (size_t t = 0; t < n && ok; ++t) ok = ex[t]->pitchT_ > 0;
  // The FP32 whole-plane kernels move two elements per 16-byte access, so
  // they need 16-byte aligned space-domain buffers; a complex64 view at an odd
  // element offset (8-byte aligned) takes the class-split kernels instead.
  for (size_t t = 0; t < n && ok; ++t) {
    const auto space = reinterpret_cast<std::uintptr_t>(forward ? din[t] : dout[t]);
    ok = space % 16 == 0;
  }
  return ok;
}

template <typename T>
void ExecutionGPU<T>::fill_batch_tables(ExecutionGPU<T>* const* ex, size_t n, const T* const* din,
                                        T* const* dout, bool forward,
                                        const SpfftScalingType* scaling, FastArgs<T>* args,
                                        int2* zmap) {
  int S = 0, LZ = 0, shift = 0;
  fast_z_geometry<T>(ex[0]->map_->dimZ, &S, &LZ, &shift);
  size_t k = 0;
  for (size_t t = 0; t < n; ++t) {
    FastArgs<T> f = ex[t]->fast_args();
    f.in = din[t];
    f.out = dout[t];
    f.useScale = forward && scaling[t] == SPFFT_FULL_SCALING ? 1 : 0;
    f.pitchA = ex[t]->classSplit_ ? fast_xy_pitch_a<T>(ex[0]->map_->dimX, f.ncol) : 0;
    const int b = (f.nst + S - 1) / S;
    for (int j = 0; j < b; ++j) zmap[k++] = make_int2(static_cast<int>(t), j);
    args[t] = f;
  }
}

template <typename T>
void ExecutionGPU<T>::launch_batch(bool forward, const BatchShape& shape, int n,
                                   const TuneChoice& choice, const FastArgs<T>* dArgs,
                                   const int2* dMap, hipStream_t stream) {
  auto launch_xy = [&] {
    if (choice.plane == 2) {
      launch_split_xy_batch<T>(forward, shape.r2c, shape.dimX, shape.dimY, dArgs, n, shape.dimZ,
                               shape.maxNcol, stream);
    } else if (choice.plane == 1) {
      launch_plane_xy_batch<T>(forward, shape.dimX, dArgs, n, shape.dimZ, shape.maxNcol,
                               choice.slots, stream);
    } else {
      launch_fast_xy_batch<T>(forward, shape.r2c, shape.dimX, dArgs, n, shape.dimZ,
                              shape.maxNcol, stream);
    }
  };
  const bool runs = choice.zRuns != 0;
  if (forward) {
    launch_xy();
    launch_fast_z_batch<T>(true, shape.r2c, shape.dimZ, dArgs, dMap, shape.blocks, runs, stream);
  } else {
    launch_fast_z_batch<T>(false, shape.r2c, shape.dimZ, dArgs, dMap, shape.blocks, runs, stream);
    launch_xy();
  }
}

template <typename T>
TuneChoice ExecutionGPU<T>::tune_batch(bool forward, const BatchShape& shape, int n, bool planeOk,
                                       const FastArgs<T>* dArgs, const int2* dMap) {
  TuneChoice c;
  TuneEvents ev;
  auto z = [&](bool runs) {
    launch_fast_z_batch<T>(forward, shape.r2c, shape.dimZ, dArgs, dMap, shape.blocks, runs,
                           nullptr);
  };
  auto xy_split = [&] {
    launch_fast_xy_batch<T>(forward, shape.r2c, shape.dimX, dArgs, n, shape.dimZ, shape.maxNcol,
                            nullptr);
  };
  auto xy_plane = [&](int slots) {
    launch_plane_xy_batch<T>(forward, shape.dimX, dArgs, n, shape.dimZ, shape.maxNcol, slots,
                             nullptr);
  };
  auto xy_two = [&] {
    launch_split_xy_batch<T>(forward, shape.r2c, shape.dimX, shape.dimY, dArgs, n, shape.dimZ,
                             shape.maxNcol, nullptr);
  };
  // The forward z kernel reads the [z][stick] buffer that the xy kernel
  // writes, the backward xy kernel the one that the z kernel writes.
  if (forward) {
    if (shape.classSplit) {
      xy_split();
    } else {
      xy_two();
    }
  }
  float zr = 0.f, ze = 0.f;
  if (shape.runs) {
    zr = time_launches([&] { z(true); }, ev.e0, ev.e1);
    ze = time_launches([&] { z(false); }, ev.e0, ev.e1);
    c.zRuns = zr <= ze ? 1 : 0;
  } else if (!forward) {
    z(false);
  }
  float split = 0.f, plane0 = 0.f, plane1 = 0.f, two = 0.f;
  two = time_launches(xy_two, ev.e0, ev.e1);
  float best = two;
  c.plane = 2;
  if (shape.classSplit) {
    // x != y planes and lengths without class-split kernels have the split
    // stage only
    split = time_launches(xy_split, ev.e0, ev.e1);
    if (split <= best) {
      best = split;
      c.plane = 0;
    }
  }
  if (planeOk) {
    plane0 = time_launches([&] { xy_plane(0); }, ev.e0, ev.e1);
    if (plane0 < best) {
      best = plane0;
      c.plane = 1;
    }
    // persistent kernels with more than one resident workgroup per CU: also
    // one per CU
    if (plane_xy_slots<T>(forward, shape.dimX) > 1) {
      plane1 = time_launches([&] { xy_plane(1); }, ev.e0, ev.e1);
      if (plane1 < best) {
        c.plane = 1;
        c.slots = 1;
      }
    }
  }
  if (tune_log_enabled()) {
    std::fprintf(stderr,
                 "spfft_amd tune: %s N=%d Nz=%d %s %s, %d transforms, %d columns: z runs %.3f / "
                 "entries %.3f ms; class-split %.3f, split stage %.3f, whole-plane %.3f, one per "
                 "CU %.3f ms -> plane=%d slots=%d zruns=%d\n",
                 sizeof(T) == 8 ? "double" : "float", shape.dimX, shape.dimZ,
                 shape.r2c ? "R2C" : "C2C", forward ? "forward" : "backward", n, shape.maxNcol,
                 zr, ze, split, two, plane0, plane1, c.plane, c.slots, c.zRuns);
  }
  return c;
}

template <typename T>
bool ExecutionGPU<T>::run_batch(const std::vector<ExecutionGPU<T>*>& ex,
                                const std::vector<const T*>& in, const std::vector<T*>& out,
                                bool forward, const std::vector<SpfftScalingType>& scaling) {
  const size_t n = ex.size();
  if (n < 2) return false;
  std::vector<const T*> din(n);
  std::vector<T*> dout(n);
  BatchShape shape;
  if (!batch_qualifies(ex.data(), n, in.data(), out.data(), din.data(), dout.data(), &shape)) {
    return false;
  }
  const int nt = static_cast<int>(n);

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

  const bool planeOk = plane_candidate(ex.data(), n, din.data(), dout.data(), forward, shape);

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
    std::vector<int2> zmap(shape.blocks);
    fill_batch_tables(ex.data(), n, din.data(), dout.data(), forward,
                      forward ? scaling.data() : nullptr, args.data(), zmap.data());
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

  // kernel choices: forced plane family, recorded (this process or the cache
  // file), or timed now on this batch's buffers
  TuneChoice choice;
  const int forced = plane_kernels_forced();
  if (forced >= 0) {
    choice.plane = forced;
  } else {
    const TuneKey tkey = make_tune_key(shape.device, static_cast<int>(sizeof(T)), shape.dimX,
                                       shape.dimY, shape.dimZ, shape.r2c, forward,
                                       shape.maxNcol, nt);
    if (!tune_lookup(tkey, &choice)) {
      choice = tune_batch(forward, shape, nt, planeOk, dArgs, dMap);
      tune_record(tkey, choice);
    }
  }
  if (choice.plane == 1 && !planeOk) {
    choice.plane = 0;
    choice.slots = 0;
  }
  if (choice.plane == 0 && !shape.classSplit) choice.plane = 2;
  if (plan.exec && (plan.choice.plane != choice.plane || plan.choice.slots != choice.slots ||
                    plan.choice.zRuns != choice.zRuns)) {
    (void)hipGraphExecDestroy(plan.exec);
    (void)hipGraphDestroy(plan.graph);
    plan.exec = nullptr;
    plan.graph = nullptr;
  }
  if (!graphs_enabled() || prof::enabled()) {
    launch_batch(forward, shape, nt, choice, dArgs, dMap, nullptr);
    return true;
  }
  if (!plan.exec) {
    plan.choice = choice;
    // Capture the two launches once on a private stream; later calls with
    // the same key replay the graph on the default stream.
    hipStream_t cs = nullptr;
    check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    try {
      launch_batch(forward, shape, nt, choice, dArgs, dMap, cs);
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

template <typename T>
void ExecutionGPU<T>::run_batch_noalloc(ExecutionGPU<T>* const* ex, size_t n, const T* const* in,
                                        T* const* out, bool forward,
                                        const SpfftScalingType* scaling, void* scratch,
                                        size_t scratchBytes, void* stage, size_t stageBytes) {
  if (n < 1 || (forward && !scaling)) throw InvalidParameterError();
  // host stage: device views of the user pointers, then the image of the
  // device tables (argument table padded to 256 bytes, then the block map)
  const size_t viewBytes = 2 * n * sizeof(void*);
  if (viewBytes > stageBytes) throw InvalidParameterError();
  char* h = static_cast<char*>(stage);
  const T** din = reinterpret_cast<const T**>(h);
  T** dout = reinterpret_cast<T**>(h + n * sizeof(void*));
  BatchShape shape;
  if (!batch_qualifies(ex, n, in, out, din, dout, &shape)) throw InvalidParameterError();
  const size_t argBytes = (n * sizeof(FastArgs<T>) + 255) / 256 * 256;
  const size_t tableBytes = argBytes + static_cast<size_t>(shape.blocks) * sizeof(int2);
  if (tableBytes > scratchBytes || viewBytes + tableBytes > stageBytes) {
    throw InvalidParameterError();
  }
  char* image = h + viewBytes;
  fill_batch_tables(ex, n, din, dout, forward, scaling, reinterpret_cast<FastArgs<T>*>(image),
                    reinterpret_cast<int2*>(image + argBytes));
  // From pageable memory the copy has consumed the stage when it returns.
  check_hip(hipMemcpyAsync(scratch, image, tableBytes, hipMemcpyHostToDevice, nullptr));
  // choices recorded by earlier batched calls of this shape in this process
  // (a lookup without allocation or file access), else the defaults
  TuneChoice choice;
  const int forced = plane_kernels_forced();
  if (forced >= 0) {
    choice.plane = forced;
  } else {
    (void)tune_recorded(make_tune_key(shape.device, static_cast<int>(sizeof(T)), shape.dimX,
                                      shape.dimY, shape.dimZ, shape.r2c, forward, shape.maxNcol,
                                      static_cast<int>(n)),
                        &choice);
  }
  if (choice.plane == 1 && !plane_candidate(ex, n, din, dout, forward, shape)) {
    choice.plane = 0;
    choice.slots = 0;
  }
  if (choice.plane == 0 && !shape.classSplit) choice.plane = 2;
  const auto* dArgs = static_cast<const FastArgs<T>*>(scratch);
  const auto* dMap = reinterpret_cast<const int2*>(static_cast<const char*>(scratch) + argBytes);
  launch_batch(forward, shape, static_cast<int>(n), choice, dArgs, dMap, nullptr);
}

template class ExecutionGPU<double>;
template class ExecutionGPU<float>;

}  // namespace spfft
