// Deep copies of spfft::Grid (copy constructor and copy assignment): the
// copy has the same maxima and its own buffers, so transforms created from
// the copy and from the original do not share their space domain. Host grids,
// so this runs without a GPU. Exit code 0 on success.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "spfft/spfft.hpp"

namespace {

int fail(const char* what) {
  std::printf("FAIL: %s\n", what);
  return 1;
}

}  // namespace

int main() {
  const int n = 8;
  spfft::Grid grid(n, n, n, n * n, SPFFT_PU_HOST, 1);
  spfft::Grid copy(grid);
  spfft::Grid assigned(4, 4, 4, 16, SPFFT_PU_HOST, 1);
  assigned = grid;
  for (const spfft::Grid* g : {&copy, &assigned}) {
    if (g->max_dim_x() != n || g->max_dim_y() != n || g->max_dim_z() != n ||
        g->max_num_local_z_columns() != n * n || g->processing_unit() != SPFFT_PU_HOST ||
        g->num_threads() != 1) {
      return fail("copied grid parameters");
    }
  }
  const std::vector<int> idx = {0, 0, 0, 1, 0, 0, 0, 1, 2, -1, 3, -2};
  const int nv = static_cast<int>(idx.size() / 3);
  spfft::Transform a = grid.create_transform(SPFFT_PU_HOST, SPFFT_TRANS_C2C, n, n, n, n, nv,
                                             SPFFT_INDEX_TRIPLETS, idx.data());
  spfft::Transform b = copy.create_transform(SPFFT_PU_HOST, SPFFT_TRANS_C2C, n, n, n, n, nv,
                                             SPFFT_INDEX_TRIPLETS, idx.data());
  spfft::Transform c = assigned.create_transform(SPFFT_PU_HOST, SPFFT_TRANS_C2C, n, n, n, n, nv,
                                                 SPFFT_INDEX_TRIPLETS, idx.data());
  double* sa = a.space_domain_data(SPFFT_PU_HOST);
  double* sb = b.space_domain_data(SPFFT_PU_HOST);
  double* sc = c.space_domain_data(SPFFT_PU_HOST);
  if (sa == sb || sa == sc || sb == sc) return fail("copies share the space domain");

  // a delta at (0,0,0) gives a constant field 1; a later transform of the
  // copy with other values must leave it untouched
  std::vector<std::complex<double>> delta(nv, 0.0), other(nv, 0.0);
  delta[0] = 1.0;
  other[1] = std::complex<double>(0.5, -2.0);
  a.backward(reinterpret_cast<const double*>(delta.data()), SPFFT_PU_HOST);
  b.backward(reinterpret_cast<const double*>(other.data()), SPFFT_PU_HOST);
  c.backward(reinterpret_cast<const double*>(other.data()), SPFFT_PU_HOST);
  for (int i = 0; i < n * n * n; ++i) {
    if (sa[2 * i] != 1.0 || sa[2 * i + 1] != 0.0) return fail("original's space domain changed");
  }
  // x = 1 plane wave of amplitude (0.5, -2): |value| = sqrt(4.25) everywhere
  for (int i = 0; i < n * n * n; ++i) {
    const double m = std::hypot(sb[2 * i], sb[2 * i + 1]);
    if (std::fabs(m - std::sqrt(4.25)) > 1e-12) return fail("copy's backward");
  }
  std::printf("grid copies: independent buffers, results correct\n");
  return 0;
}
