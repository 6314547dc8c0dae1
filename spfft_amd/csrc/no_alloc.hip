// Device scratch of the allocation-free batched entry point. Each GPU holds
// its own copy of `scratch` from the library's load until the process ends.
// Host code never passes the array's name to a kernel or a copy (on the host
// it only names a stand-in, not a device address): it uses the address that
// hipGetSymbolAddress returns on each device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hip_check.hpp"
#include "no_alloc.hpp"
#include "spfft/exceptions.hpp"

namespace spfft {

namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;

struct DeviceSlot {
  std::once_flag once;
  void* ptr = nullptr;
  hipError_t status = hipSuccess;
};
DeviceSlot slots[kMaxDevices];

alignas(256) unsigned char hostStage[kNoAllocStageBytes];
std::mutex stageMutex;

}  // namespace

NoAllocScratch noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw InvalidParameterError();
  DeviceSlot& s = slots[device];
  std::call_once(s.once, [&] {
    void* p = nullptr;
    s.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    s.ptr = p;
  });
  check_hip(s.status);
  if (!s.ptr) throw GPUError();
  NoAllocScratch r;
  r.device = s.ptr;
  r.stage = hostStage;
  r.stageLock = &stageMutex;
  return r;
}

}  // namespace spfft
