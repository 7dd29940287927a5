// World-grid lookups shared by the LM and trajectory kernels (the grids are
// built by pk_world_grid_f32 in pk_grid.hip).
#pragma once

#include "pk_common.h"

namespace pk {

struct GridLevel {
  const int* cells;    // (n_cells, stride): count, then up to cap indices
  float lo_x, lo_y, lo_z, inv_cell;
  int nx, ny, nz, cap, stride;
  float r_max;         // largest robot sphere radius this level serves
};

// Up to two grids of increasing reach; a robot sphere uses the first whose
// r_max covers its radius, or every world sphere when none does (n == 0: no
// grid at all).
// (Two named members rather than an array: indexing an array member of a
// kernel argument makes the compiler copy the whole argument struct to
// scratch memory.)
struct GridArgs {
  GridLevel lv0, lv1;
  int n;
};
static_assert(kMaxGrids == 2, "GridArgs holds two levels");

// Host side: validate the caller's grid descriptors and fill `out`.  A grid
// built with pad serves robot spheres of radius <= pad - margin.
inline int fill_grid_args(const pk_world_grid* grids, int n_grids, float margin, GridArgs& out) {
  PK_REQUIRE(n_grids >= 0 && n_grids <= PK_MAX_GRIDS, "n_grids=%d outside [0, %d]", n_grids,
             PK_MAX_GRIDS);
  PK_REQUIRE(n_grids == 0 || grids != nullptr, "null grids");
  out.n = 0;
  for (int i = 0; i < n_grids; ++i) {
    const pk_world_grid& g = grids[i];
    PK_REQUIRE(g.cells && g.cell > 0.f && g.dims[0] > 0 && g.dims[1] > 0 && g.dims[2] > 0 &&
                   g.cap > 0 && g.stride >= g.cap + 1 && g.stride % 4 == 0 && g.pad >= 0.f,
               "invalid world grid %d", i);
    PK_REQUIRE(i == 0 || g.pad >= grids[i - 1].pad, "grids must be ordered by increasing pad");
    const GridLevel lv{g.cells, g.lo[0], g.lo[1], g.lo[2], 1.f / g.cell,
                       g.dims[0], g.dims[1], g.dims[2], g.cap, g.stride, g.pad - margin};
    if (i == 0) out.lv0 = lv; else out.lv1 = lv;
  }
  out.n = n_grids;
  return PK_OK;
}

// Candidate list of the cell holding c in one level: n entries, read with
// grid_item.  The record's first 16 bytes carry the count and the first three
// indices, so a typical lookup is a single load.  Returns false when the cell
// overflowed (the caller then tests every world sphere); a centre outside the
// grid lists nothing (no world sphere is within reach).
__device__ __forceinline__ bool level_lookup(const GridLevel& lv, float cx, float cy, float cz, int4& head,
                                             const int*& rec, int& n) {
  n = 0;
  rec = nullptr;
  head = make_int4(0, 0, 0, 0);
  const int ix = static_cast<int>(floorf((cx - lv.lo_x) * lv.inv_cell));
  const int iy = static_cast<int>(floorf((cy - lv.lo_y) * lv.inv_cell));
  const int iz = static_cast<int>(floorf((cz - lv.lo_z) * lv.inv_cell));
  if (ix < 0 || iy < 0 || iz < 0 || ix >= lv.nx || iy >= lv.ny || iz >= lv.nz) return true;
  const long long cell = (static_cast<long long>(iz) * lv.ny + iy) * lv.nx + ix;
  rec = lv.cells + cell * lv.stride;
  head = *reinterpret_cast<const int4*>(rec);
  if (head.x > lv.cap) return false;
  n = head.x;
  return true;
}

// Entry i of a looked-up list (indices in increasing order).
__device__ __forceinline__ int grid_item(const int4& head, const int* rec, int i) {
  return i == 0 ? head.y : (i == 1 ? head.z : (i == 2 ? head.w : rec[1 + i]));
}

// Lookup for a robot sphere of radius r (uniform across the wave) in the
// first level that serves r; false = test every world sphere.
__device__ __forceinline__ bool grid_lookup(const GridArgs& g, float cx, float cy, float cz, float r,
                                            int4& head, const int*& rec, int& n) {
  if (g.n > 0 && r <= g.lv0.r_max) return level_lookup(g.lv0, cx, cy, cz, head, rec, n);
  if (g.n > 1 && r <= g.lv1.r_max) return level_lookup(g.lv1, cx, cy, cz, head, rec, n);
  n = 0;
  rec = nullptr;
  head = make_int4(0, 0, 0, 0);
  return false;
}

}  // namespace pk
