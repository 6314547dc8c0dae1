// Single-precision class-split and whole-plane kernels of the fixed-length
// path: the source of fast_kernels.hip compiled as a translation unit of its
// own, so that the units build in parallel.
#define SPFFT_FAST_UNIT 3
#include "fast_kernels.hip"
