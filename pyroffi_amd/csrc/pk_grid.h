// World-grid lookups shared by the LM and trajectory kernels (the grids are
// built by pk_world_grid_f32 in pk_grid.hip).
#pragma once

#include "pk_common.h"

namespace pk {

struct GridLevel {
  const int* counts;   // (n_cells)
  const int* items;    // (n_cells, cap)
  float lo_x, lo_y, lo_z, inv_cell;
  int nx, ny, nz, cap;
  float r_max;         // largest robot sphere radius this level serves
};

// Up to two grids of increasing reach; a robot sphere uses the first whose
// r_max covers its radius, or every world sphere when none does (n == 0: no
// grid at all).
struct GridArgs {
  GridLevel lv[kMaxGrids];
  int n;
};

// Host side: validate the caller's grid descriptors and fill `out`.  A grid
// built with pad serves robot spheres of radius <= pad - margin.
inline int fill_grid_args(const pk_world_grid* grids, int n_grids, float margin, GridArgs& out) {
  PK_REQUIRE(n_grids >= 0 && n_grids <= PK_MAX_GRIDS, "n_grids=%d outside [0, %d]", n_grids,
             PK_MAX_GRIDS);
  PK_REQUIRE(n_grids == 0 || grids != nullptr, "null grids");
  out.n = 0;
  for (int i = 0; i < n_grids; ++i) {
    const pk_world_grid& g = grids[i];
    PK_REQUIRE(g.counts && g.items && g.cell > 0.f && g.dims[0] > 0 && g.dims[1] > 0 &&
                   g.dims[2] > 0 && g.cap > 0 && g.pad >= 0.f,
               "invalid world grid %d", i);
    PK_REQUIRE(i == 0 || g.pad >= grids[i - 1].pad, "grids must be ordered by increasing pad");
    out.lv[i] = GridLevel{g.counts, g.items, g.lo[0], g.lo[1], g.lo[2], 1.f / g.cell,
                          g.dims[0], g.dims[1], g.dims[2], g.cap, g.pad - margin};
  }
  out.n = n_grids;
  return PK_OK;
}

// Candidate world spheres for a robot sphere of radius r (uniform across the
// wave) centred at c: the list of the cell holding c in the first level that
// serves r.  Returns false when every world sphere must be tested (no level
// serves r, or the cell overflowed); a centre outside the grid yields an
// empty list (no world sphere is within reach).
__device__ __forceinline__ bool grid_candidates(const GridArgs& g, float cx, float cy, float cz,
                                                float r, const int*& items, int& count) {
#pragma unroll
  for (int i = 0; i < kMaxGrids; ++i) {
    if (i < g.n && r <= g.lv[i].r_max) {
      const GridLevel& lv = g.lv[i];
      const int ix = static_cast<int>(floorf((cx - lv.lo_x) * lv.inv_cell));
      const int iy = static_cast<int>(floorf((cy - lv.lo_y) * lv.inv_cell));
      const int iz = static_cast<int>(floorf((cz - lv.lo_z) * lv.inv_cell));
      if (ix < 0 || iy < 0 || iz < 0 || ix >= lv.nx || iy >= lv.ny || iz >= lv.nz) {
        items = nullptr;
        count = 0;
        return true;
      }
      const long long cell = (static_cast<long long>(iz) * lv.ny + iy) * lv.nx + ix;
      const int n = lv.counts[cell];
      if (n > lv.cap) return false;
      items = lv.items + cell * lv.cap;
      count = n;
      return true;
    }
  }
  return false;
}

}  // namespace pk
