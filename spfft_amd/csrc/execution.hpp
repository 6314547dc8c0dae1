// GPU execution of one transform: device index tables, twiddles, stream and
// the stage sequence (reference: src/execution/execution_gpu.hpp:59-128 and
// src/execution/execution_gpu.cpp:254-410).
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "grid_internal.hpp"
#include "index_map.hpp"
#include "fast_kernels.hpp"
#include "kernels.hpp"
#include "memory.hpp"
#include "spfft/types.h"
#include "tuner.hpp"

namespace spfft {

// Device index tables and twiddles of one index set on one device. Built once
// and shared by every transform with an identical set (same dimensions, type,
// precision and value order), so the bands of one set hold, and a batch of
// them reads, a single copy. Immutable after construction.
template <typename T>
struct TransformTables {
  std::shared_ptr<const IndexMap> map;
  int device = 0;
  bool allowFast = true;  // built with the fixed-length path allowed
  bool fast = false;      // fixed-length kernels serve this set
  bool classSplit = false;  // and the class-split plane kernels (nx = ny, LDS fits)
  int numCols = 0;
  int pitchT = 0;  // whole-plane batch kernels' tile pitch (0: not applicable)
  DeviceBuffer<Tw<T>> twX, twY, twZ;
  DeviceBuffer<int> entryValue, entrySlot, stickBegin, stickY, stickCol, colX, colBegin;
  DeviceBuffer<int> stickAt;
  DeviceBuffer<int2> zEntries;
  DeviceBuffer<int4> zInfo;  // stick runs, when every stick is one run in user order
  DeviceBuffer<int> resSticks, resBegin, xCol, resYM;
};

template <typename T>
class ExecutionGPU {
public:
  ExecutionGPU(const std::shared_ptr<IndexMap>& map, const std::shared_ptr<GridInternal<T>>& grid);
  ~ExecutionGPU();
  ExecutionGPU(const ExecutionGPU&) = delete;
  ExecutionGPU& operator=(const ExecutionGPU&) = delete;

  // Enqueue a transform. Pointers may be device or host memory.
  // Single calls run on the default (null) stream itself, which orders them
  // after earlier default-stream work and makes later default-stream work
  // wait for them -- the reference's contract (execution_gpu.cpp:333-344,
  // 403-410) without cross-stream hops. Batched calls (concurrent = true) run
  // on the transform's own stream, fenced against the default stream with
  // events, so independent transforms overlap.
  void backward(const T* input, T* output, bool concurrent = false);
  void forward(const T* input, T* output, SpfftScalingType scaling, bool concurrent = false);
  // SYNCHRONOUS: wait for the work. ASYNCHRONOUS: return; a concurrent call
  // additionally makes the default stream wait for its stream.
  void synchronize(SpfftExecType mode, bool concurrent = false);

  // Batched execution of several transforms of one shape on the fixed-length
  // path: one z launch and one xy launch for the whole batch, on the default
  // stream. The argument tables and the captured launch chain (a HIP graph)
  // are kept for the most recent batches and replayed when the same
  // transforms are called again with the same pointers. Returns false without
  // doing anything when the batch does not qualify (different shapes,
  // generic lengths, host pointers, empty sets).
  static bool run_batch(const std::vector<ExecutionGPU<T>*>& ex, const std::vector<const T*>& in,
                        const std::vector<T*>& out, bool forward,
                        const std::vector<SpfftScalingType>& scaling);

  // Pieces of the fused batch, shared by run_batch and the allocation-free
  // batched entry point (which keeps its tables in a static device array).
  struct BatchShape {
    int blocks = 0;   // z workgroups of the whole batch
    int maxNcol = 0;  // widest column set
    bool r2c = false;
    int dimX = 0, dimY = 0, dimZ = 0;
    bool runs = false;  // every transform's sticks are runs in user order
    bool classSplit = false;  // class-split plane kernels serve every transform
    int device = 0;
  };
  // True when the n transforms can run as one fused chain; then din/dout (n
  // entries each) hold the device views of the user pointers.
  static bool batch_qualifies(ExecutionGPU* const* ex, size_t n, const T* const* in,
                              T* const* out, const T** din, T** dout, BatchShape* shape);
  // True when the whole-plane kernels can serve the batch.
  static bool plane_candidate(ExecutionGPU* const* ex, size_t n, const T* const* din,
                              T* const* dout, bool forward, const BatchShape& shape);
  // Host copies of the argument table (n entries) and the z block map
  // (shape.blocks entries); `scaling` is read for forward batches only.
  static void fill_batch_tables(ExecutionGPU* const* ex, size_t n, const T* const* din,
                                T* const* dout, bool forward, const SpfftScalingType* scaling,
                                FastArgs<T>* args, int2* zmap);
  // The fused chain with the given kernel choices (choice.plane is honoured
  // only when the whole-plane kernels can serve the batch).
  static void launch_batch(bool forward, const BatchShape& shape, int n, const TuneChoice& choice,
                           const FastArgs<T>* dArgs, const int2* dMap, hipStream_t stream);
  // Times the candidate kernels on the batch's own buffers and returns the
  // fastest choice (the launches are idempotent; the call's real chain
  // follows).
  static TuneChoice tune_batch(bool forward, const BatchShape& shape, int n, bool planeOk,
                               const FastArgs<T>* dArgs, const int2* dMap);

  // run_batch for callers that cannot allocate (n >= 1 transforms). The
  // argument table and the z block map are written to `scratch`
  // (scratchBytes of device memory owned by the caller and reused by every
  // call), staged through `stage` (stageBytes of host memory); the launches
  // go to the default stream. No device or host allocation, no graph, no
  // timing: the plane-kernel family tuned by earlier batched calls is used,
  // else the class-split kernels. Throws InvalidParameterError when the batch
  // does not qualify (host pointers, mixed shapes, lengths without fixed
  // kernels, empty sets) or its tables do not fit.
  static void run_batch_noalloc(ExecutionGPU* const* ex, size_t n, const T* const* in,
                                T* const* out, bool forward, const SpfftScalingType* scaling,
                                void* scratch, size_t scratchBytes, void* stage,
                                size_t stageBytes);

  T* space_domain_gpu() const { return grid_->space_gpu(); }
  T* space_domain_host() const { return grid_->space_host(); }
  hipStream_t stream() const { return stream_; }
  // Number of T values in the space domain (2 per element for C2C).
  size_t space_domain_size() const;

  // True for device or managed memory (used in place, no staging).
  static bool device_accessible(const void* ptr);

  // Chunks of a host transfer (copies overlap kernels and host copies).
  static constexpr int kIoMaxChunks = 64;

private:
  // Checks for preceding errors and returns the stream to enqueue on.
  hipStream_t begin(bool concurrent);
  // Host I/O: events and the grid's two-slot pinned ring of slotBytes each.
  void io_prepare(size_t slotBytes);
  int planes_per_chunk(size_t planeBytes) const;
  // Host -> device piece k through the ring (pageable) or directly (pinned),
  // on `copy`; event k is recorded after it.
  void upload_piece(char* dst, const char* src, size_t bytes, bool pinned, int k,
                    hipStream_t copy);
  // A whole host range in ring-sized pieces; `compute` waits for the last.
  void upload(char* dst, const char* src, size_t bytes, bool pinned, hipStream_t compute,
              hipStream_t copy);
  // Device -> host in pieces of pieceBytes; piece k waits for ready[min(k,
  // readyCount - 1)] (recorded on `compute`). Pageable destinations are filled
  // before the call returns; pinned ones when `compute` reaches its end.
  void download(char* dst, const char* src, size_t bytes, size_t pieceBytes, bool pinned,
                const hipEvent_t* ready, int readyCount, hipStream_t compute, hipStream_t copy);
  StageArgs<T> base_args() const;
  FastArgs<T> fast_args() const;
  // Kernel choices of a single transform (plane stage: class-split or split
  // kernels; z addressing when the sticks are runs), tuned on the first call
  // of the shape with device input and output, on that call's buffers
  // (staged = host data in the work buffer: recorded choices or defaults).
  TuneChoice single_choice(bool forward, const FastArgs<T>& f, hipStream_t stream, bool staged);

  std::shared_ptr<IndexMap> map_;
  std::shared_ptr<GridInternal<T>> grid_;
  std::uint64_t serial_ = 0;  // process-unique, never reused
  hipStream_t stream_ = nullptr;
  hipEvent_t startEvent_ = nullptr, endEvent_ = nullptr;
  T scale_;
  int numCols_ = 0;
  int zBatch_ = 1, yBatch_ = 1, xBatch_ = 1;
  FftDesc fftX_, fftY_, fftZ_;
  std::shared_ptr<TransformTables<T>> tab_;  // shared with transforms of the same set
  // fixed-length path (two kernels per direction)
  bool fast_ = false;
  bool classSplit_ = false;  // class-split plane kernels too (else the split stage only)
  // whole-plane batch kernels (power-of-two lengths up to 128, C2C): tile pitch
  int pitchT_ = 0;
  // host I/O (created on the first call with a host pointer): chunk-ready
  // events [0, kIoMaxChunks), copy-done events [kIoMaxChunks, 2 kIoMaxChunks)
  std::array<hipEvent_t, 2 * kIoMaxChunks> ioEvents_{};
  char* ring_ = nullptr;
  size_t slotBytes_ = 0;
};

}  // namespace spfft
