/* C usage of the library (counterpart of the reference's examples/example.c):
 * a 4x4x4 C2C transform of 8 centred frequency values on the GPU, with host
 * buffers (the library stages them). Backward, then forward with full
 * scaling must return the input. Exit status 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "spfft/spfft.h"

#define CHECK(call)                                                      \
  do {                                                                   \
    SpfftError e_ = (call);                                              \
    if (e_ != SPFFT_SUCCESS) {                                           \
      fprintf(stderr, "%s failed with error %d\n", #call, (int)e_);      \
      return 1;                                                          \
    }                                                                    \
  } while (0)

int main(void) {
  const int dimX = 4, dimY = 4, dimZ = 4, n = 8;
  int indices[3 * 8];
  double values[2 * 8], result[2 * 8];
  for (int i = 0; i < n; ++i) {
    indices[3 * i] = (i & 1) ? -1 : 1;
    indices[3 * i + 1] = (i >> 1) & 1;
    indices[3 * i + 2] = (i >> 2) ? 2 : -1;
    values[2 * i] = 0.5 + i;
    values[2 * i + 1] = -0.25 * i;
  }
  SpfftGrid grid = NULL;
  SpfftTransform transform = NULL;
  CHECK(spfft_grid_create(&grid, dimX, dimY, dimZ, dimX * dimY, SPFFT_PU_GPU, -1));
  CHECK(spfft_transform_create(&transform, grid, SPFFT_PU_GPU, SPFFT_TRANS_C2C, dimX, dimY, dimZ,
                               dimZ, n, SPFFT_INDEX_TRIPLETS, indices));
  /* the grid may go before its transforms */
  CHECK(spfft_grid_destroy(grid));

  double* space = (double*)malloc(sizeof(double) * 2 * dimX * dimY * dimZ);
  CHECK(spfft_transform_backward_ptr(transform, values, space));
  CHECK(spfft_transform_forward_ptr(transform, space, result, SPFFT_FULL_SCALING));
  double err = 0.0;
  for (int i = 0; i < 2 * n; ++i) err = fmax(err, fabs(result[i] - values[i]));
  printf("max round-trip error: %.3e\n", err);
  free(space);
  CHECK(spfft_transform_destroy(transform));
  return err < 1e-12 ? 0 : 2;
}
