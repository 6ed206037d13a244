// Cell-list construction on the device (the build's correct counterpart of
// NLCCalcs::createNLC / getCell, NLCStep.cpp:125-180,245-253).  Molecules
// are binned by their first primary site into fixed-capacity slots; the
// slots of every cell are then put in molecule order, so the candidate order
// (and the summation order) does not depend on atomic arrival order.
#pragma once

#include <climits>

#include "device_common.h"
#include "wave.h"

namespace mcgpu {

__device__ __forceinline__ int cells_total(const CellGrid& g) {
  return g.n[0] * g.n[1] * g.n[2];
}

// Molecule m into the cell of its first primary (mirror entry
// mol_pidx_start[m]); a full cell marks the lists incomplete.
__device__ __forceinline__ void cells_assign_one(const BoxView& bv, int m) {
  const CellGrid& g = bv.cells;
  const int p = bv.mol_pidx_start[m];
  const int c = cell_index(g, cell_coord(bv.px[p], g.inv_w[0], g.n[0]),
                           cell_coord(bv.py[p], g.inv_w[1], g.n[1]),
                           cell_coord(bv.pz[p], g.inv_w[2], g.n[2]));
  const int slot = atomicAdd(&g.count[c], 1);
  g.mol_cell[m] = c;
  if (slot < g.cap) {
    g.slots[c * g.cap + slot] = m;
  } else {
    atomicOr(g.overflow, 1);
  }
}

// One wave orders cell c's slots by molecule id: each lane takes one entry
// and its rank is the number of smaller ids in the cell (ids are distinct),
// counted against the entries read across lanes.  Cells of more than 64
// entries are ordered by one lane (insertion sort).
__device__ __forceinline__ void cells_sort_wave(const CellGrid& g, int c) {
  const int lane = wlane();
  int n = g.count[c];
  n = n < g.cap ? n : g.cap;
  int32_t* s = g.slots + (size_t)c * g.cap;
  if (n <= 64) {
    const int v = lane < n ? s[lane] : INT_MAX;
    int rank = 0;
    for (int k = 0; k < n; k++) rank += readlane_i32(v, k) < v ? 1 : 0;
    if (lane < n) {
      s[rank] = v;
      g.mol_slot[v] = rank;
    }
    return;
  }
  if (lane == 0) {
    for (int i = 1; i < n; i++) {
      const int v = s[i];
      int k = i - 1;
      while (k >= 0 && s[k] > v) {
        s[k + 1] = s[k];
        k--;
      }
      s[k + 1] = v;
    }
    for (int i = 0; i < n; i++) g.mol_slot[s[i]] = i;
  }
}

}  // namespace mcgpu
