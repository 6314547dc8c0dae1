/* Per-kernel timing of the MI355X build. No counterpart in the reference ABI;
 * used by benchmarks to time each kernel with HIP events on the stream it
 * runs on.
 *
 * spfft_amd_profiling_enable(1) makes every kernel launch record a start and
 * an end event on its stream. spfft_amd_profiling_report() waits for them and
 * writes, per kernel name, the launch count and total milliseconds; names
 * go into consecutive 64-byte slots of `names`. It returns the number of
 * kernels written and clears the records. */
#ifndef SPFFT_AMD_PROFILING_H
#define SPFFT_AMD_PROFILING_H

#include "spfft/config.h"

#ifdef __cplusplus
extern "C" {
#endif

SPFFT_EXPORT int spfft_amd_profiling_enable(int enable);
SPFFT_EXPORT int spfft_amd_profiling_report(int maxKernels, char* names, long long* counts,
                                            double* totalMs);

#ifdef __cplusplus
}
#endif

#endif
