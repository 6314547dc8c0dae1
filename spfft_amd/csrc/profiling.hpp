// Optional per-kernel timing with HIP events recorded on the launching
// stream (off by default; enabled through spfft_amd_profiling_enable).
#pragma once

#include <hip/hip_runtime_api.h>

namespace spfft {
namespace prof {

bool enabled();

// Times one kernel launch, attributed to kernel `name` (a string literal).
// Default: records a start event on construction and an end event on
// destruction, both on `stream`. attach = true: records nothing itself; the
// caller hands start() and stop() to hipExtLaunchKernelGGL, which stamps them
// when the kernel itself begins and ends (null when profiling is off).
class KernelSpan {
public:
  KernelSpan(hipStream_t stream, const char* name, bool attach = false);
  ~KernelSpan();
  KernelSpan(const KernelSpan&) = delete;
  KernelSpan& operator=(const KernelSpan&) = delete;
  hipEvent_t start() const { return start_; }
  hipEvent_t stop() const { return stop_; }

private:
  hipStream_t stream_;
  int slot_ = -1;
  bool attach_ = false;
  hipEvent_t start_ = nullptr, stop_ = nullptr;
};

}  // namespace prof
}  // namespace spfft
