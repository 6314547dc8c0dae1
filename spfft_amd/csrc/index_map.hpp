// Sparse frequency index maps of one transform (host side, built once at
// creation). Semantics follow src/compression/indices.hpp:120-186 and
// src/parameters/parameters.cpp:143-180 of the reference, restated with a
// sort + binary search instead of an ordered map; the resulting integers are
// identical.
#pragma once

#include <cstdint>
#include <vector>

#include "spfft/types.h"

namespace spfft {

struct IndexMap {
  int dimX = 0, dimY = 0, dimZ = 0;
  int dimXFreq = 0;  // dimX / 2 + 1 for R2C, dimX otherwise
  SpfftTransformType type = SPFFT_TRANS_C2C;
  // Per value i (creation order): stick * dimZ + zStorage.
  std::vector<int> valueSlot;
  // Distinct column keys xStorage * dimY + yStorage in ascending order; the
  // position of a key is its stick (z-column) id.
  std::vector<int> stickKey;
  // 0 when the (x=0, y=0) column exists, number of sticks otherwise.
  int zeroZeroStick = 0;

  int num_sticks() const { return static_cast<int>(stickKey.size()); }
  int num_values() const { return static_cast<int>(valueSlot.size()); }
};

// Map an index in [-N, N) to storage order [0, N)
// (src/compression/indices.hpp:49-55).
inline int to_storage_index(int dim, int index) { return index < 0 ? dim + index : index; }

// Validate and convert interleaved (x,y,z) triplets. Throws
// InvalidParameterError when numValues exceeds the grid, InvalidIndicesError
// for an index outside the allowed range.
IndexMap build_index_map(SpfftTransformType type, int dimX, int dimY, int dimZ, int numValues,
                         const int* triplets);

// Device-facing tables derived from an IndexMap (all int32).
struct StickTables {
  // Values sorted by slot: entry e = (value index, slot). For repeated
  // triplets the entries keep creation order, so "last one wins" on scatter.
  std::vector<int> entryValue, entrySlot;
  std::vector<int> stickBegin;  // entry range per stick, size numSticks + 1
  std::vector<int> stickY;      // y storage index per stick
  std::vector<int> stickCol;    // column (distinct x) id per stick
  std::vector<int> colX;        // x storage index per column, ascending
  std::vector<int> colBegin;    // stick range per column, size numCols + 1
};

StickTables build_stick_tables(const IndexMap& map);

}  // namespace spfft
