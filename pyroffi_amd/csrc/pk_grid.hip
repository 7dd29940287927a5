// Uniform grid over a sphere world, for exact culling of margin-limited costs.
//
// Cell c (an axis-aligned box of side `cell`) lists every world sphere m whose
// distance to the box is below R_m + pad, where pad = (largest robot sphere
// radius) + margin.  A robot sphere with centre inside the box can only have
// d = |x - o_m| - r - R_m < margin against spheres in that list, so a cost
// that vanishes for d >= margin (the LM collision penalty, the hinge) is exact
// when evaluated over the list alone.  A centre outside the grid is farther
// than pad from every sphere's surface (the grid box is the world's bounding
// box grown by pad) and contributes nothing.
//
// One lane per cell; the count pass and the fill pass run the same test.
#include "pk_common.h"

#include <stdint.h>

namespace pk {

template <bool FILL>
__global__ __launch_bounds__(256) void world_grid_kernel(const float* __restrict__ world, int M,
                                                         float pad, float lo_x, float lo_y, float lo_z,
                                                         float cell, int nx, int ny, int nz, int* cells,
                                                         int* __restrict__ items, long long n_items) {
  const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long n_cells = static_cast<long long>(nx) * ny * nz;
  if (c >= n_cells) return;
  const int ix = static_cast<int>(c % nx);
  const int iy = static_cast<int>((c / nx) % ny);
  const int iz = static_cast<int>(c / (static_cast<long long>(nx) * ny));
  const float bx0 = lo_x + ix * cell, by0 = lo_y + iy * cell, bz0 = lo_z + iz * cell;
  const long long start = FILL ? static_cast<long long>(cells[2 * c]) : 0;
  int n = 0;
  for (int m = 0; m < M; ++m) {
    const float ox = world[4 * m], oy = world[4 * m + 1], oz = world[4 * m + 2];
    const float reach = world[4 * m + 3] + pad;
    const float dx = fmaxf(fmaxf(bx0 - ox, ox - (bx0 + cell)), 0.f);
    const float dy = fmaxf(fmaxf(by0 - oy, oy - (by0 + cell)), 0.f);
    const float dz = fmaxf(fmaxf(bz0 - oz, oz - (bz0 + cell)), 0.f);
    // A small slack keeps the test conservative under float rounding.
    if (dx * dx + dy * dy + dz * dz <= (reach + 1e-4f) * (reach + 1e-4f)) {
      if (FILL && start >= 0 && start + n < n_items) items[start + n] = m;
      ++n;
    }
  }
  if (!FILL) cells[2 * c + 1] = n;
}

}  // namespace pk

namespace {

int check_grid_args(int32_t n_world_spheres, const float* world_spheres, float pad, float cell, int32_t nx,
                    int32_t ny, int32_t nz, const int32_t* cells, long long& n_cells) {
  PK_REQUIRE(n_world_spheres >= 0 && pad >= 0.f && cell > 0.f, "invalid world grid arguments");
  PK_REQUIRE(nx > 0 && ny > 0 && nz > 0, "grid dimensions must be positive");
  n_cells = static_cast<long long>(nx) * ny * nz;
  PK_REQUIRE(n_cells <= (1LL << 22), "grid of %lld cells is too large (limit 4M)", n_cells);
  PK_REQUIRE(cells && (n_world_spheres == 0 || world_spheres), "null pointer");
  return PK_OK;
}

}  // namespace

extern "C" int pk_world_grid_count_f32(void* stream, const float* world_spheres, int32_t n_world_spheres,
                                       float pad, float lo_x, float lo_y, float lo_z, float cell, int32_t nx,
                                       int32_t ny, int32_t nz, int32_t* cells) {
  long long n_cells = 0;
  const int st = check_grid_args(n_world_spheres, world_spheres, pad, cell, nx, ny, nz, cells, n_cells);
  if (st != PK_OK) return st;
  const unsigned blocks = static_cast<unsigned>((n_cells + 255) / 256);
  hipLaunchKernelGGL(pk::world_grid_kernel<false>, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     world_spheres, n_world_spheres, pad, lo_x, lo_y, lo_z, cell, nx, ny, nz, cells, nullptr,
                     0LL);
  return pk::check_launch("world_grid_kernel");
}

extern "C" int pk_world_grid_fill_f32(void* stream, const float* world_spheres, int32_t n_world_spheres,
                                      float pad, float lo_x, float lo_y, float lo_z, float cell, int32_t nx,
                                      int32_t ny, int32_t nz, const int32_t* cells, int32_t* items,
                                      int64_t n_items) {
  long long n_cells = 0;
  const int st = check_grid_args(n_world_spheres, world_spheres, pad, cell, nx, ny, nz, cells, n_cells);
  if (st != PK_OK) return st;
  PK_REQUIRE(n_items >= 0 && (n_items == 0 || items), "invalid items buffer");
  const unsigned blocks = static_cast<unsigned>((n_cells + 255) / 256);
  hipLaunchKernelGGL(pk::world_grid_kernel<true>, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     world_spheres, n_world_spheres, pad, lo_x, lo_y, lo_z, cell, nx, ny, nz,
                     const_cast<int*>(cells), items, static_cast<long long>(n_items));
  return pk::check_launch("world_grid_kernel");
}
