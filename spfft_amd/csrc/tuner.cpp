#include "tuner.hpp"

#include <hip/hip_runtime_api.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>

namespace spfft {

namespace {

struct TuneState {
  std::mutex mutex;
  bool loaded = false;
  std::string path;  // empty: no file
  std::map<std::string, TuneChoice> choices;
  std::map<int, std::string> arch;  // device -> "gfx950:256"
  // choices this process has used, by numeric key
  std::map<std::tuple<int, int, int, int, int, int, int, int, int>, TuneChoice> used;
};

std::tuple<int, int, int, int, int, int, int, int, int> numeric(const TuneKey& k) {
  return std::make_tuple(k.device, k.valueBytes, k.n, k.ny, k.nz, k.r2c, k.forward, k.ncolBucket,
                         k.countBucket);
}

TuneState& state() {
  static TuneState* s = new TuneState();
  return *s;
}

std::string cache_path() {
  const char* v = std::getenv("SPFFT_AMD_TUNE_CACHE");
  if (v) {
    if (v[0] == '\0' || std::strcmp(v, "off") == 0) return std::string();
    return std::string(v);
  }
  std::string base;
  if (const char* x = std::getenv("XDG_CACHE_HOME"); x && x[0]) {
    base = x;
  } else if (const char* h = std::getenv("HOME"); h && h[0]) {
    base = std::string(h) + "/.cache";
  } else {
    return std::string();
  }
  return base + "/spfft_amd/tune-v1.txt";
}

// mkdir -p of the directory part of `path`
void make_parent_dirs(const std::string& path) {
  for (size_t i = 1; i < path.size(); ++i) {
    if (path[i] == '/') {
      const std::string d = path.substr(0, i);
      (void)mkdir(d.c_str(), 0755);
    }
  }
}

void load_locked(TuneState& s) {
  if (s.loaded) return;
  s.loaded = true;
  s.path = cache_path();
  if (s.path.empty()) return;
  FILE* f = std::fopen(s.path.c_str(), "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '#') continue;
    char key[256];
    TuneChoice c;
    if (std::sscanf(line, "%255s plane=%d slots=%d zruns=%d", key, &c.plane, &c.slots,
                    &c.zRuns) == 4 &&
        c.plane >= 0 && c.plane <= 2 && c.slots >= 0 && c.slots <= 64 &&
        (c.zRuns == 0 || c.zRuns == 1)) {
      s.choices[key] = c;
    }
  }
  std::fclose(f);
}

void save_locked(const TuneState& s) {
  if (s.path.empty()) return;
  make_parent_dirs(s.path);
  const std::string tmp = s.path + ".tmp" + std::to_string(static_cast<long>(getpid()));
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return;
  std::fprintf(f, "# spfft_amd kernel choices: <arch:CUs:bytes:N[xNy]:Nz:type:direction:"
                  "columns:transforms> plane= slots= zruns=\n");
  for (const auto& kv : s.choices) {
    std::fprintf(f, "%s plane=%d slots=%d zruns=%d\n", kv.first.c_str(), kv.second.plane,
                 kv.second.slots, kv.second.zRuns);
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok || std::rename(tmp.c_str(), s.path.c_str()) != 0) (void)std::remove(tmp.c_str());
}

int round_pow2(int v) {
  int p = 1;
  while (p < v) p *= 2;
  return p;
}

}  // namespace

TuneKey make_tune_key(int device, int valueBytes, int n, int ny, int nz, bool r2c, bool forward,
                      int maxNcol, int count) {
  TuneKey k;
  k.device = device;
  k.valueBytes = valueBytes;
  k.n = n;
  k.ny = ny;
  k.nz = nz;
  k.r2c = r2c ? 1 : 0;
  k.forward = forward ? 1 : 0;
  k.ncolBucket = (maxNcol + 15) / 16 * 16;
  k.countBucket = round_pow2(count < 1 ? 1 : count);
  return k;
}

std::string tune_key_text(const TuneKey& k) {
  TuneState& s = state();
  std::string arch;
  {
    std::lock_guard<std::mutex> lock(s.mutex);
    auto it = s.arch.find(k.device);
    if (it != s.arch.end()) arch = it->second;
  }
  if (arch.empty()) {
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, k.device) == hipSuccess) {
      std::string name = prop.gcnArchName;
      const size_t colon = name.find(':');
      if (colon != std::string::npos) name = name.substr(0, colon);
      arch = name + ":" + std::to_string(prop.multiProcessorCount);
    } else {
      (void)hipGetLastError();
      arch = "unknown:0";
    }
    std::lock_guard<std::mutex> lock(s.mutex);
    s.arch[k.device] = arch;
  }
  std::ostringstream o;
  o << arch << ":f" << 8 * k.valueBytes << ":N" << k.n;
  if (k.ny != k.n) o << "x" << k.ny;
  o << ":Z" << k.nz << ":"
    << (k.r2c ? "r2c" : "c2c") << ":" << (k.forward ? "fwd" : "bwd") << ":col" << k.ncolBucket
    << ":x" << k.countBucket;
  return o.str();
}

bool tune_lookup(const TuneKey& key, TuneChoice* choice) {
  const std::string text = tune_key_text(key);
  TuneState& s = state();
  std::lock_guard<std::mutex> lock(s.mutex);
  load_locked(s);
  auto it = s.choices.find(text);
  if (it == s.choices.end()) return false;
  *choice = it->second;
  s.used[numeric(key)] = it->second;
  return true;
}

bool tune_recorded(const TuneKey& key, TuneChoice* choice) {
  TuneState& s = state();
  std::lock_guard<std::mutex> lock(s.mutex);
  auto it = s.used.find(numeric(key));
  if (it == s.used.end()) return false;
  *choice = it->second;
  return true;
}

void tune_record(const TuneKey& key, const TuneChoice& choice) {
  const std::string text = tune_key_text(key);
  TuneState& s = state();
  std::lock_guard<std::mutex> lock(s.mutex);
  load_locked(s);
  s.choices[text] = choice;
  s.used[numeric(key)] = choice;
  save_locked(s);
}

bool tune_log_enabled() {
  const char* v = std::getenv("SPFFT_AMD_TUNE_LOG");
  return v && v[0] == '1';
}

}  // namespace spfft
