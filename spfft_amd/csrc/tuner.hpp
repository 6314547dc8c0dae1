// Autotuner of the fixed-length kernels. Choices are timed on the first call
// of a shape (on that call's own buffers), kept per process and persisted in a
// per-shape cache file, so later processes start from the measured choices.
//
// Key: device architecture and CU count, value bytes, Nx, Ny, Nz, C2C or R2C, direction, widest column count (rounded up to 16) and transform count
// (rounded up to a power of two; 1 for single transforms).
// Choices: the plane-kernel family of batches (class-split kernels with R
// residue classes per plane and M-point decimation-in-frequency lines, or
// whole-plane kernels with one workgroup per plane and full-length lines),
// the number of resident workgroups per CU of the persistent FP32 plane
// kernels, and the z kernels' addressing (stick runs computed from the first
// value, or the per-value entry table).
//
// Cache file: SPFFT_AMD_TUNE_CACHE names it ("off" or empty: no file);
// unset, $XDG_CACHE_HOME/spfft_amd/tune-v1.txt or ~/.cache/spfft_amd/
// tune-v1.txt. One decision per line; unreadable lines are ignored; the file
// is replaced atomically (write, then rename) when a decision is added.
#pragma once

#include <string>

namespace spfft {

struct TuneKey {
  int device = 0;
  int valueBytes = 0;
  int n = 0, ny = 0, nz = 0;
  int r2c = 0, forward = 0;
  int ncolBucket = 0, countBucket = 0;
};

struct TuneChoice {
  int plane = 0;  // plane stage: 0 class-split kernels, 1 whole-plane kernels (batches),
                  // 2 split stage (y and x kernels around a compact column buffer)
  int slots = 0;  // persistent plane kernels: resident workgroups per CU (0: all that fit)
  int zRuns = 1;  // z kernels address stick runs (when every stick of the set is a run)
};

TuneKey make_tune_key(int device, int valueBytes, int n, int ny, int nz, bool r2c, bool forward,
                      int maxNcol, int count);
// The recorded choice for `key` (from this process or the cache file).
bool tune_lookup(const TuneKey& key, TuneChoice* choice);
// Records a timed choice and rewrites the cache file.
void tune_record(const TuneKey& key, const TuneChoice& choice);
// The choice for `key` if this process has already looked it up or recorded
// it; no allocation and no file access (for the allocation-free entry point).
bool tune_recorded(const TuneKey& key, TuneChoice* choice);
// SPFFT_AMD_TUNE_LOG=1: decisions and their timings go to stderr.
bool tune_log_enabled();
std::string tune_key_text(const TuneKey& key);

}  // namespace spfft
