// das_plan.h — the geometry plan object behind the C-ABI's opaque
// ustbx_plan (shared by das_capi.hip and das_batch.hip).
#pragma once
#include <stdint.h>

#include <vector>

#include "das_launch.h"

struct ustbx_plan {
  int device;
  ustbx::GeomLaunch G;
  char* dev;
  size_t dev_bytes;
  float4* tx_table;      // resident [W][P] transmit table
  size_t tx_table_bytes;
  float* unified_tx;     // [W][P] transmit delays of unified-model waves (else null)
  size_t unified_bytes;
  bool dasc_order;
  // variant 2 (receive table): 0 not built, 1 built, -1 not applicable
  int rxtab_state;
  int max_nw;            // largest active-wave count of a tile
  char* rxtab;           // one allocation: receive table, tile data, reference phase
  size_t rxtab_bytes;
  size_t coef_bytes;     // variant 3's coefficient buffer (G.coef)
  const int* tile_order; // band-major order (G.tile_order when selected)
  const int* tile_natural;  // tiles in index order, the same partition
  int n_bad;             // tiles with a non-finite nonzero weight or delay (last in both orders)
  std::vector<int> h_order;     // host copies for batches: band-major tile order,
  std::vector<int64_t> h_cost;  // terms per tile (waves x channels),
  std::vector<char> h_bad;      // and whether a tile takes the SAFE body
};
