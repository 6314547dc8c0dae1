#include "index_map.hpp"

#include <algorithm>
#include <numeric>

#include "spfft/exceptions.hpp"

namespace spfft {

IndexMap build_index_map(SpfftTransformType type, int dimX, int dimY, int dimZ, int numValues,
                         const int* triplets) {
  const bool r2c = type == SPFFT_TRANS_R2C;
  using u64 = unsigned long long;
  if (static_cast<u64>(numValues) >
      static_cast<u64>(dimX) * static_cast<u64>(dimY) * static_cast<u64>(dimZ)) {
    throw InvalidParameterError();
  }

  // One negative index anywhere switches all three dimensions to the centred
  // range (indices.hpp:129-135).
  bool centered = false;
  for (long long i = 0; i < 3LL * numValues; ++i) {
    if (triplets[i] < 0) {
      centered = true;
      break;
    }
  }
  const int maxX = (r2c || centered ? dimX / 2 + 1 : dimX) - 1;
  const int maxY = (centered ? dimY / 2 + 1 : dimY) - 1;
  const int maxZ = (centered ? dimZ / 2 + 1 : dimZ) - 1;
  const int minX = r2c ? 0 : maxX - dimX + 1;
  const int minY = maxY - dimY + 1;
  const int minZ = maxZ - dimZ + 1;
  for (int i = 0; i < numValues; ++i) {
    const int x = triplets[3 * i], y = triplets[3 * i + 1], z = triplets[3 * i + 2];
    if (x < minX || x > maxX || y < minY || y > maxY || z < minZ || z > maxZ) {
      throw InvalidIndicesError();
    }
  }

  IndexMap m;
  m.dimX = dimX;
  m.dimY = dimY;
  m.dimZ = dimZ;
  m.dimXFreq = r2c ? dimX / 2 + 1 : dimX;
  m.type = type;

  std::vector<int> keys(numValues);
  for (int i = 0; i < numValues; ++i) {
    keys[i] = to_storage_index(dimX, triplets[3 * i]) * dimY +
              to_storage_index(dimY, triplets[3 * i + 1]);
  }
  m.stickKey = keys;
  std::sort(m.stickKey.begin(), m.stickKey.end());
  m.stickKey.erase(std::unique(m.stickKey.begin(), m.stickKey.end()), m.stickKey.end());

  m.valueSlot.resize(numValues);
  for (int i = 0; i < numValues; ++i) {
    const int stick = static_cast<int>(
        std::lower_bound(m.stickKey.begin(), m.stickKey.end(), keys[i]) - m.stickKey.begin());
    m.valueSlot[i] = stick * dimZ + to_storage_index(dimZ, triplets[3 * i + 2]);
  }
  // Key 0 sorts first, so the (0,0) stick is either stick 0 or absent
  // (parameters.cpp:173-179).
  m.zeroZeroStick =
      (!m.stickKey.empty() && m.stickKey.front() == 0) ? 0 : static_cast<int>(m.stickKey.size());
  return m;
}

StickTables build_stick_tables(const IndexMap& m) {
  StickTables t;
  const int nv = m.num_values();
  const int nst = m.num_sticks();
  std::vector<int> order(nv);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return m.valueSlot[a] < m.valueSlot[b]; });
  t.entryValue.resize(nv);
  t.entrySlot.resize(nv);
  for (int e = 0; e < nv; ++e) {
    t.entryValue[e] = order[e];
    t.entrySlot[e] = m.valueSlot[order[e]];
  }
  t.stickBegin.assign(nst + 1, 0);
  for (int e = 0; e < nv; ++e) t.stickBegin[t.entrySlot[e] / m.dimZ + 1]++;
  for (int s = 0; s < nst; ++s) t.stickBegin[s + 1] += t.stickBegin[s];

  t.stickY.resize(nst);
  t.stickCol.resize(nst);
  t.colBegin.clear();
  for (int s = 0; s < nst; ++s) {
    const int x = m.stickKey[s] / m.dimY;
    t.stickY[s] = m.stickKey[s] % m.dimY;
    if (t.colX.empty() || t.colX.back() != x) {
      t.colX.push_back(x);
      t.colBegin.push_back(s);
    }
    t.stickCol[s] = static_cast<int>(t.colX.size()) - 1;
  }
  t.colBegin.push_back(nst);
  return t;
}

}  // namespace spfft

// Test hook: the library's own index conversion, callable without a GPU.
// No counterpart in the reference ABI (declared in include/spfft_amd/index_map.h).
extern "C" __attribute__((visibility("default"))) int spfft_amd_convert_index_triplets(
    int r2c, int dimX, int dimY, int dimZ, int numValues, const int* triplets, int* valueSlot,
    int* stickKey, int* numSticks, int* zeroZeroStick) {
  try {
    const spfft::IndexMap m = spfft::build_index_map(r2c ? SPFFT_TRANS_R2C : SPFFT_TRANS_C2C, dimX,
                                                     dimY, dimZ, numValues, triplets);
    for (int i = 0; i < m.num_values(); ++i) valueSlot[i] = m.valueSlot[i];
    for (int s = 0; s < m.num_sticks(); ++s) stickKey[s] = m.stickKey[s];
    *numSticks = m.num_sticks();
    *zeroZeroStick = m.zeroZeroStick;
  } catch (const spfft::GenericError& e) {
    return e.error_code();
  } catch (...) {
    return SPFFT_UNKNOWN_ERROR;
  }
  return SPFFT_SUCCESS;
}
