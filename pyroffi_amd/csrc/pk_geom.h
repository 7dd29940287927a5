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

// Capsule robots: a link capsule is the segment [a, b] (o[0..5]) swept by
// radius o[6] (the reference's [7] row).  Against the four obstacle kinds
// (_collision_cuda_helpers.cuh:21-142):
//   sphere     the sphere-capsule distance;
//   capsule    closest points of the two segments (clamped parameters, the
//              degenerate-segment cases handled separately);
//   box        the box signed distance at the point of the segment closest
//              to the box centre (in box coordinates), minus the radius;
//   half-space the lower of the two end-point plane offsets, minus the radius.
__device__ __forceinline__ float dist_capsule_capsule(const float* c, const float* o) {
  const float d1x = c[3] - c[0], d1y = c[4] - c[1], d1z = c[5] - c[2];
  const float d2x = o[3] - o[0], d2y = o[4] - o[1], d2z = o[5] - o[2];
  const float rx = c[0] - o[0], ry = c[1] - o[1], rz = c[2] - o[2];
  const float a = d1x * d1x + d1y * d1y + d1z * d1z;
  const float e = d2x * d2x + d2y * d2y + d2z * d2z;
  const float f = d2x * rx + d2y * ry + d2z * rz;
  constexpr float kEps = 1e-10f;
  float s = 0.f, t = 0.f;
  if (a > kEps || e > kEps) {
    if (a <= kEps) {
      t = fminf(fmaxf(f / e, 0.f), 1.f);
    } else {
      const float cc = d1x * rx + d1y * ry + d1z * rz;
      if (e <= kEps) {
        s = fminf(fmaxf(-cc / a, 0.f), 1.f);
      } else {
        const float bb = d1x * d2x + d1y * d2y + d1z * d2z;
        const float den = a * e - bb * bb;
        s = (fabsf(den) > kEps) ? fminf(fmaxf((bb * f - cc * e) / den, 0.f), 1.f) : 0.f;
        t = (bb * s + f) / e;
        if (t < 0.f) {
          t = 0.f;
          s = fminf(fmaxf(-cc / a, 0.f), 1.f);
        } else if (t > 1.f) {
          t = 1.f;
          s = fminf(fmaxf((bb - cc) / a, 0.f), 1.f);
        }
      }
    }
  }
  const float px = c[0] + s * d1x - (o[0] + t * d2x);
  const float py = c[1] + s * d1y - (o[1] + t * d2y);
  const float pz = c[2] + s * d1z - (o[2] + t * d2z);
  return sqrtf(px * px + py * py + pz * pz) - (c[6] + o[6]);
}

__device__ __forceinline__ float dist_capsule_box(const float* c, const float* o) {
  const float ax = c[0] - o[0], ay = c[1] - o[1], az = c[2] - o[2];
  const float bx = c[3] - o[0], by = c[4] - o[1], bz = c[5] - o[2];
  const float a1 = ax * o[3] + ay * o[4] + az * o[5];
  const float a2 = ax * o[6] + ay * o[7] + az * o[8];
  const float a3 = ax * o[9] + ay * o[10] + az * o[11];
  const float u1 = (bx * o[3] + by * o[4] + bz * o[5]) - a1;
  const float u2 = (bx * o[6] + by * o[7] + bz * o[8]) - a2;
  const float u3 = (bx * o[9] + by * o[10] + bz * o[11]) - a3;
  const float uu = u1 * u1 + u2 * u2 + u3 * u3;
  float t = 0.f;
  if (uu > 1e-12f) t = fminf(fmaxf(-(a1 * u1 + a2 * u2 + a3 * u3) / uu, 0.f), 1.f);
  const float e1 = fabsf(a1 + t * u1) - o[12];
  const float e2 = fabsf(a2 + t * u2) - o[13];
  const float e3 = fabsf(a3 + t * u3) - o[14];
  const float m1 = fmaxf(e1, 0.f), m2 = fmaxf(e2, 0.f), m3 = fmaxf(e3, 0.f);
  return sqrtf(m1 * m1 + m2 * m2 + m3 * m3) + fminf(fmaxf(fmaxf(e1, e2), e3), 0.f) - c[6];
}

__device__ __forceinline__ float dist_capsule_halfspace(const float* c, const float* o) {
  const float d1 = (c[0] - o[3]) * o[0] + (c[1] - o[4]) * o[1] + (c[2] - o[5]) * o[2];
  const float d2 = (c[3] - o[3]) * o[0] + (c[4] - o[4]) * o[1] + (c[5] - o[5]) * o[2];
  return fminf(d1, d2) - c[6];
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

namespace pk {

// Distance of capsule c (7 floats) to obstacle m of the packed world.
__device__ __forceinline__ float dist_capsule_obstacle(const float* c, const WorldPrims& w, int m) {
  if (m < w.count[0]) {
    const float* o = w.data[0] + 4 * m;
    const float cap[7] = {c[0], c[1], c[2], c[3], c[4], c[5], o[3]};
    return dist_sphere_capsule(o[0], o[1], o[2], c[6], cap);
  }
  m -= w.count[0];
  if (m < w.count[1]) return dist_capsule_capsule(c, w.data[1] + 7 * m);
  m -= w.count[1];
  if (m < w.count[2]) return dist_capsule_box(c, w.data[2] + 15 * m);
  m -= w.count[2];
  return dist_capsule_halfspace(c, w.data[3] + 6 * m);
}

}  // namespace pk
