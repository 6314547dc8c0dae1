// HIP status -> SpFFT exception mapping (reference behaviour:
// src/gpu_util/gpu_runtime_api.hpp:113-123 maps runtime errors to the GPU*
// exception family).
#pragma once

#include <hip/hip_runtime_api.h>

#include "spfft/exceptions.hpp"

namespace spfft {

inline void check_hip(hipError_t status) {
  switch (status) {
    case hipSuccess: return;
    case hipErrorOutOfMemory: throw GPUAllocationError();
    case hipErrorNoDevice: throw GPUNoDeviceError();
    case hipErrorInvalidValue: throw GPUInvalidValueError();
    case hipErrorInvalidDevicePointer: throw GPUInvalidDevicePointerError();
    case hipErrorLaunchFailure:
    case hipErrorLaunchOutOfResources:
    case hipErrorInvalidConfiguration: throw GPULaunchError();
    default: throw GPUError();
  }
}

// An error left by earlier GPU work is reported, not absorbed (reference
// execution_gpu.cpp:256-258). Called first in every GPU entry point, before
// the call's own HIP calls (a pointer query may clear the thread's error).
inline void check_preceding() {
  if (hipGetLastError() != hipSuccess) throw GPUPrecedingError();
}

// Check the launch that was just enqueued.
inline void check_launch() {
  const hipError_t status = hipGetLastError();
  if (status != hipSuccess) throw GPULaunchError();
}

// Select the device of a grid for the lifetime of the guard
// (reference: src/gpu_util/gpu_device_guard.hpp:38-63).
class DeviceGuard {
public:
  explicit DeviceGuard(int device) {
    check_hip(hipGetDevice(&previous_));
    if (previous_ != device) {
      check_hip(hipSetDevice(device));
      switched_ = true;
    }
  }
  ~DeviceGuard() {
    if (switched_) (void)hipSetDevice(previous_);
  }
  DeviceGuard(const DeviceGuard&) = delete;
  DeviceGuard& operator=(const DeviceGuard&) = delete;

private:
  int previous_ = 0;
  bool switched_ = false;
};

}  // namespace spfft
