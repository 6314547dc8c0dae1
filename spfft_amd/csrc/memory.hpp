// Owning device and pinned-host buffers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <new>
#include <utility>
#include <vector>

#include "hip_check.hpp"

namespace spfft {

template <typename T>
class DeviceBuffer {
public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t count) : count_(count) {
    if (count_ > 0) {
      const hipError_t status = hipMalloc(reinterpret_cast<void**>(&ptr_), count_ * sizeof(T));
      if (status != hipSuccess) {
        (void)hipGetLastError();
        ptr_ = nullptr;
        throw GPUAllocationError();
      }
    }
  }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept { swap(o); }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      swap(o);
    }
    return *this;
  }

  T* data() const { return ptr_; }
  size_t size() const { return count_; }

  // Synchronous upload of a host vector into a new buffer.
  static DeviceBuffer from_host(const std::vector<T>& v) {
    DeviceBuffer b(v.size());
    if (!v.empty()) {
      check_hip(hipMemcpy(b.ptr_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    return b;
  }

private:
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    count_ = 0;
  }
  void swap(DeviceBuffer& o) noexcept {
    std::swap(ptr_, o.ptr_);
    std::swap(count_, o.count_);
  }
  T* ptr_ = nullptr;
  size_t count_ = 0;
};

template <typename T>
class PinnedBuffer {
public:
  PinnedBuffer() = default;
  explicit PinnedBuffer(size_t count) : count_(count) {
    if (count_ > 0) {
      const hipError_t status =
          hipHostMalloc(reinterpret_cast<void**>(&ptr_), count_ * sizeof(T), hipHostMallocDefault);
      if (status != hipSuccess) {
        (void)hipGetLastError();
        ptr_ = nullptr;
        throw HostAllocationError();
      }
    }
  }
  ~PinnedBuffer() {
    if (ptr_) (void)hipHostFree(ptr_);
  }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;

  T* data() const { return ptr_; }
  size_t size() const { return count_; }

private:
  T* ptr_ = nullptr;
  size_t count_ = 0;
};

// Host array: pinned (hipHostMalloc) or ordinary 64-byte aligned memory.
template <typename T>
class HostBuffer {
public:
  HostBuffer(size_t count, bool pinned) : count_(count), pinned_(pinned) {
    if (count_ == 0) return;
    if (pinned_) {
      const hipError_t status = hipHostMalloc(reinterpret_cast<void**>(&ptr_),
                                              count_ * sizeof(T), hipHostMallocDefault);
      if (status != hipSuccess) {
        (void)hipGetLastError();
        ptr_ = nullptr;
        throw HostAllocationError();
      }
    } else {
      ptr_ = static_cast<T*>(::operator new(count_ * sizeof(T), std::align_val_t(64),
                                            std::nothrow));
      if (!ptr_) throw HostAllocationError();
    }
  }
  ~HostBuffer() {
    if (!ptr_) return;
    if (pinned_) {
      (void)hipHostFree(ptr_);
    } else {
      ::operator delete(ptr_, std::align_val_t(64));
    }
  }
  HostBuffer(const HostBuffer&) = delete;
  HostBuffer& operator=(const HostBuffer&) = delete;

  T* data() const { return ptr_; }
  size_t size() const { return count_; }

private:
  T* ptr_ = nullptr;
  size_t count_ = 0;
  bool pinned_ = false;
};

}  // namespace spfft
