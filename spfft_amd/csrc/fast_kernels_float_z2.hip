// Single-precision z kernels of the fixed-length path, second group of
// lengths (36 to 512): the source of fast_kernels.hip compiled as a
// translation unit of its own, so that the units build in parallel.
#define SPFFT_FAST_UNIT 8
#include "fast_kernels.hip"
