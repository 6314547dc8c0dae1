// Single-precision z kernels of the fixed-length path, fourth group of
// lengths (odd multiples of 11 and 13, 99 to 351): the source of
// fast_kernels.hip compiled as a translation unit of its own, so that the
// units build in parallel.
#define SPFFT_FAST_UNIT 16
#include "fast_kernels.hip"
