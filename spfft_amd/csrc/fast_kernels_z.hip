// Double-precision z kernels of the fixed-length path: the source of
// fast_kernels.hip compiled as a translation unit of its own, so that the units
// build in parallel.
#define SPFFT_FAST_UNIT 1
#include "fast_kernels.hip"
