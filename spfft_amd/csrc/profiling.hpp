// Optional per-kernel timing with HIP events recorded on the launching
// stream (off by default; enabled through spfft_amd_profiling_enable).
#pragma once

#include <hip/hip_runtime_api.h>

namespace spfft {
namespace prof {

bool enabled();

// Records a start event on construction and an end event on destruction,
// both on `stream`, attributed to kernel `name` (a string literal).
class KernelSpan {
public:
  KernelSpan(hipStream_t stream, const char* name);
  ~KernelSpan();
  KernelSpan(const KernelSpan&) = delete;
  KernelSpan& operator=(const KernelSpan&) = delete;

private:
  hipStream_t stream_;
  int slot_ = -1;
};

}  // namespace prof
}  // namespace spfft
