// Single-precision instantiations of the fixed-length kernels: the source of
// fast_kernels.hip compiled as a second translation unit, so that the two
// precisions build in parallel.
#define SPFFT_FAST_FLOAT_UNIT
#include "fast_kernels.hip"
