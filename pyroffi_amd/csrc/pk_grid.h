// World-grid lookups shared by the LM and trajectory kernels (the grids are
// built by pk_world_grid_f32 in pk_grid.hip).
#pragma once

#include "pk_common.h"

namespace pk {

struct GridLevel {
  const int2* cells;   // (n_cells): start, count in items
  const int* items;    // world sphere indices
  float lo_x, lo_y, lo_z, inv_cell;
  int nx, ny, nz;
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
    PK_REQUIRE(g.cells && g.items && g.cell > 0.f && g.dims[0] > 0 && g.dims[1] > 0 && g.dims[2] > 0 &&
                   g.pad >= 0.f,
               "invalid world grid %d", i);
    PK_REQUIRE(reinterpret_cast<uintptr_t>(g.cells) % 8 == 0, "grid cells must be 8-byte aligned");
    PK_REQUIRE(i == 0 || g.pad >= grids[i - 1].pad, "grids must be ordered by increasing pad");
    const GridLevel lv{reinterpret_cast<const int2*>(g.cells), g.items, g.lo[0], g.lo[1], g.lo[2],
                       1.f / g.cell, g.dims[0], g.dims[1], g.dims[2], g.pad - margin};
    if (i == 0) out.lv0 = lv; else out.lv1 = lv;
  }
  out.n = n_grids;
  return PK_OK;
}

// Cell of c in one level, -1 outside the grid (no world sphere is within
// reach there).
__device__ __forceinline__ long long level_cell(const GridLevel& lv, float cx, float cy, float cz) {
  const int ix = static_cast<int>(floorf((cx - lv.lo_x) * lv.inv_cell));
  const int iy = static_cast<int>(floorf((cy - lv.lo_y) * lv.inv_cell));
  const int iz = static_cast<int>(floorf((cz - lv.lo_z) * lv.inv_cell));
  if (ix < 0 || iy < 0 || iz < 0 || ix >= lv.nx || iy >= lv.ny || iz >= lv.nz) return -1;
  return (static_cast<long long>(iz) * lv.ny + iy) * lv.nx + ix;
}

// Level serving a robot sphere of radius r: 0 or 1, or -1 when none does (the
// caller then tests every world sphere).
__device__ __forceinline__ int grid_level(const GridArgs& g, float r) {
  if (g.n > 0 && r <= g.lv0.r_max) return 0;
  if (g.n > 1 && r <= g.lv1.r_max) return 1;
  return -1;
}

// Candidate list (items pointer, count) of the cell holding c at a level;
// count 0 outside the grid.
// (The level is copied by value: a reference chosen at run time between two
// members of a kernel argument would make the compiler copy the whole argument
// to scratch memory.)
__device__ __forceinline__ GridLevel grid_level_at(const GridArgs& g, int level) {
  GridLevel lv = g.lv0;
  if (level != 0) lv = g.lv1;
  return lv;
}

__device__ __forceinline__ void grid_cell_list(const GridArgs& g, int level, float cx, float cy, float cz,
                                               const int*& items, int& count) {
  const GridLevel lv = grid_level_at(g, level);
  const long long c = level_cell(lv, cx, cy, cz);
  if (c < 0) {
    items = lv.items;
    count = 0;
    return;
  }
  const int2 h = lv.cells[c];
  items = lv.items + h.x;
  count = h.y;
}

}  // namespace pk
