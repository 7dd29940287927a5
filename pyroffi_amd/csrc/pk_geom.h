// Signed distances from a sphere (centre p, radius r) to the reference's four
// world obstacle kinds, FP32, positive = separated.  Packed obstacle rows are
// the reference's world arrays (src/pyroffi/collision/_cuda_collision.py:98-168):
//   sphere     [4]  centre, radius
//   capsule    [7]  segment end a, segment end b, radius
//   box        [15] centre, three unit axes (columns of its rotation), half lengths
//   half-space [6]  outward unit normal, a point on the plane
// The formulas are the closed forms the reference CUDA helpers evaluate
// (src/pyroffi/cuda_kernels/_collision_cuda_helpers.cuh:14-142): distance to the
// closest point of the segment, the rotated-box signed distance function, and
// the plane offset.
#pragma once

#include <hip/hip_runtime.h>

namespace pk {

__device__ __forceinline__ float dist_sphere_sphere(float px, float py, float pz, float r, const float* o) {
  const float dx = px - o[0], dy = py - o[1], dz = pz - o[2];
  return sqrtf(dx * dx + dy * dy + dz * dz) - (r + o[3]);
}

// Closest point of segment [a, b] to p: t = clamp((p - a).(b - a) / |b - a|^2, 0, 1),
// t = 0 for a degenerate segment.
__device__ __forceinline__ float dist_sphere_capsule(float px, float py, float pz, float r, const float* o) {
  const float ux = o[3] - o[0], uy = o[4] - o[1], uz = o[5] - o[2];
  const float wx = px - o[0], wy = py - o[1], wz = pz - o[2];
  const float uu = ux * ux + uy * uy + uz * uz;
  float t = 0.f;
  if (uu > 1e-12f) t = fminf(fmaxf((wx * ux + wy * uy + wz * uz) / uu, 0.f), 1.f);
  const float dx = wx - t * ux, dy = wy - t * uy, dz = wz - t * uz;
  return sqrtf(dx * dx + dy * dy + dz * dz) - (r + o[6]);
}

// Box signed distance in the box frame: with e = |p_local| - h,
// sdf = |max(e, 0)| + min(max(e_x, e_y, e_z), 0).
__device__ __forceinline__ float dist_sphere_box(float px, float py, float pz, float r, const float* o) {
  const float dx = px - o[0], dy = py - o[1], dz = pz - o[2];
  const float e1 = fabsf(dx * o[3] + dy * o[4] + dz * o[5]) - o[12];
  const float e2 = fabsf(dx * o[6] + dy * o[7] + dz * o[8]) - o[13];
  const float e3 = fabsf(dx * o[9] + dy * o[10] + dz * o[11]) - o[14];
  const float m1 = fmaxf(e1, 0.f), m2 = fmaxf(e2, 0.f), m3 = fmaxf(e3, 0.f);
  const float outside = sqrtf(m1 * m1 + m2 * m2 + m3 * m3);
  const float inside = fminf(fmaxf(fmaxf(e1, e2), e3), 0.f);
  return outside + inside - r;
}

__device__ __forceinline__ float dist_sphere_halfspace(float px, float py, float pz, float r, const float* o) {
  return (px - o[3]) * o[0] + (py - o[4]) * o[1] + (pz - o[5]) * o[2] - r;
}

// A packed world: the four kinds back to back in the reference's column order
// (spheres, capsules, boxes, half-spaces); obstacle m of the combined index
// space is found by its kind's range.
struct WorldPrims {
  const float* data[4];
  int count[4];
  __host__ __device__ int total() const { return count[0] + count[1] + count[2] + count[3]; }
};

__device__ __forceinline__ float dist_sphere_obstacle(float px, float py, float pz, float r,
                                                      const WorldPrims& w, int m) {
  if (m < w.count[0]) return dist_sphere_sphere(px, py, pz, r, w.data[0] + 4 * m);
  m -= w.count[0];
  if (m < w.count[1]) return dist_sphere_capsule(px, py, pz, r, w.data[1] + 7 * m);
  m -= w.count[1];
  if (m < w.count[2]) return dist_sphere_box(px, py, pz, r, w.data[2] + 15 * m);
  m -= w.count[2];
  return dist_sphere_halfspace(px, py, pz, r, w.data[3] + 6 * m);
}

}  // namespace pk
