#include "profiling.hpp"

#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace spfft {
namespace prof {
namespace {

struct Record {
  const char* name;
  hipEvent_t start;
  hipEvent_t end;
};

std::atomic<bool> g_enabled{false};
std::mutex g_mutex;
std::vector<Record> g_records;   // recorded spans since the last report
std::vector<hipEvent_t> g_pool;  // events free for reuse

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

}  // namespace

bool enabled() { return g_enabled.load(std::memory_order_relaxed); }

KernelSpan::KernelSpan(hipStream_t stream, const char* name, bool attach)
    : stream_(stream), attach_(attach) {
  if (!enabled()) return;
  std::lock_guard<std::mutex> lock(g_mutex);
  Record r{name, take_event(), take_event()};
  if (!r.start || !r.end) return;
  if (attach_) {
    start_ = r.start;
    stop_ = r.end;
  } else {
    (void)hipEventRecord(r.start, stream_);
  }
  slot_ = static_cast<int>(g_records.size());
  g_records.push_back(r);
}

KernelSpan::~KernelSpan() {
  if (slot_ < 0 || attach_) return;
  std::lock_guard<std::mutex> lock(g_mutex);
  (void)hipEventRecord(g_records[slot_].end, stream_);
}

}  // namespace prof
}  // namespace spfft

// No counterpart in the reference ABI (declared in include/spfft_amd/profiling.h).
extern "C" __attribute__((visibility("default"))) int spfft_amd_profiling_enable(int enable) {
  spfft::prof::g_enabled.store(enable != 0);
  return 0;
}

// Waits for all recorded spans, then reports per kernel name: launches and
// total milliseconds. Clears the records. Returns the number of kernels
// written (at most maxKernels); names are copied into 64-byte slots.
extern "C" __attribute__((visibility("default"))) int spfft_amd_profiling_report(
    int maxKernels, char* names, long long* counts, double* totalMs) {
  using namespace spfft::prof;
  std::lock_guard<std::mutex> lock(g_mutex);
  std::map<std::string, std::pair<long long, double>> acc;
  for (const Record& r : g_records) {
    float ms = 0.f;
    if (hipEventSynchronize(r.end) == hipSuccess && hipEventElapsedTime(&ms, r.start, r.end) == hipSuccess) {
      auto& a = acc[r.name];
      a.first += 1;
      a.second += ms;
    } else {
      (void)hipGetLastError();
    }
    g_pool.push_back(r.start);
    g_pool.push_back(r.end);
  }
  g_records.clear();
  int k = 0;
  for (const auto& kv : acc) {
    if (k >= maxKernels) break;
    std::strncpy(names + 64 * k, kv.first.c_str(), 63);
    names[64 * k + 63] = '\0';
    counts[k] = kv.second.first;
    totalMs[k] = kv.second.second;
    ++k;
  }
  return k;
}
