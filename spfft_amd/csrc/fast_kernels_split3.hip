// Double-precision split plane stage of the fixed-length path, third group of
// lengths (factors 11 and 13, 44 to 462): the source of fast_kernels.hip
// compiled as a translation unit of its own, so that the units build in
// parallel.
#define SPFFT_FAST_UNIT 11
#include "fast_kernels.hip"
