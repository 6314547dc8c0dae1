// C++ usage of the library (counterpart of the reference's examples/example.cpp):
// an R2C transform of a real field on the GPU through the internal space
// buffer, with exceptions for errors. Exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <vector>

#include "spfft/spfft.hpp"

int main() {
  const int dimX = 6, dimY = 5, dimZ = 4;
  // all x >= 0 half-space triplets
  std::vector<int> indices;
  for (int x = 0; x <= dimX / 2; ++x)
    for (int y = 0; y < dimY; ++y)
      for (int z = 0; z < dimZ; ++z) {
        indices.push_back(x);
        indices.push_back(y);
        indices.push_back(z);
      }
  const int n = static_cast<int>(indices.size() / 3);
  try {
    spfft::Grid grid(dimX, dimY, dimZ, (dimX / 2 + 1) * dimY, SPFFT_PU_GPU, -1);
    spfft::Transform transform = grid.create_transform(SPFFT_PU_GPU, SPFFT_TRANS_R2C, dimX, dimY,
                                                       dimZ, dimZ, n, SPFFT_INDEX_TRIPLETS,
                                                       indices.data());
    // fill the host mirror of the space domain with a real field
    double* host = transform.space_domain_data(SPFFT_PU_HOST);
    std::vector<double> field(dimX * dimY * dimZ);
    for (size_t i = 0; i < field.size(); ++i) field[i] = host[i] = std::sin(0.37 * i) + 0.1 * i;
    std::vector<double> freq(2 * n);
    transform.forward(SPFFT_PU_HOST, freq.data(), SPFFT_FULL_SCALING);
    transform.backward(freq.data(), SPFFT_PU_HOST);
    double err = 0.0;
    for (size_t i = 0; i < field.size(); ++i) err = std::fmax(err, std::fabs(host[i] - field[i]));
    std::printf("max R2C round-trip error: %.3e\n", err);
    // shared-grid transforms are rejected by multi_transform
    spfft::Transform other = transform;
    try {
      double* outs[2] = {freq.data(), freq.data()};
      SpfftProcessingUnitType locs[2] = {SPFFT_PU_HOST, SPFFT_PU_HOST};
      SpfftScalingType sc[2] = {SPFFT_NO_SCALING, SPFFT_NO_SCALING};
      spfft::Transform ts[2] = {transform, other};
      spfft::multi_transform_forward(2, ts, locs, outs, sc);
      std::printf("shared grid was not rejected\n");
      return 3;
    } catch (const spfft::InvalidParameterError&) {
    }
    return err < 1e-12 ? 0 : 2;
  } catch (const spfft::GenericError& e) {
    std::printf("error: %s (%d)\n", e.what(), static_cast<int>(e.error_code()));
    return 1;
  }
}
