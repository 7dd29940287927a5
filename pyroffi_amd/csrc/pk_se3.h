// SE(3) algebra for the gfx950 kernels, in the reference's storage convention:
// a pose is a unit Hamilton quaternion followed by a translation,
// [w, x, y, z, tx, ty, tz] ("wxyz_xyz"), and a twist is [v; omega] (linear
// part first), as in src/pyroffi/cuda_kernels/_fk_cuda_helpers.cuh:8-13.
//
// Everything here is FP32 and register resident: one lane owns one pose.
#pragma once

#include <hip/hip_runtime.h>

namespace pk {

struct Pose {
  float w, x, y, z;     // rotation
  float tx, ty, tz;     // translation
};

__device__ __forceinline__ Pose pose_identity() {
  Pose p;
  p.w = 1.f; p.x = 0.f; p.y = 0.f; p.z = 0.f;
  p.tx = 0.f; p.ty = 0.f; p.tz = 0.f;
  return p;
}

// Hamilton product a * b.
__device__ __forceinline__ void quat_mul(float aw, float ax, float ay, float az,
                                         float bw, float bx, float by, float bz,
                                         float& ow, float& ox, float& oy, float& oz) {
  ow = aw * bw - ax * bx - ay * by - az * bz;
  ox = aw * bx + ax * bw + ay * bz - az * by;
  oy = aw * by - ax * bz + ay * bw + az * bx;
  oz = aw * bz + ax * by - ay * bx + az * bw;
}

// v' = q v q*  via  u = 2 (q_v x v);  v' = v + w u + q_v x u.
__device__ __forceinline__ void quat_rotate(float w, float x, float y, float z,
                                            float vx, float vy, float vz,
                                            float& ox, float& oy, float& oz) {
  const float ux = 2.f * (y * vz - z * vy);
  const float uy = 2.f * (z * vx - x * vz);
  const float uz = 2.f * (x * vy - y * vx);
  ox = vx + w * ux + (y * uz - z * uy);
  oy = vy + w * uy + (z * ux - x * uz);
  oz = vz + w * uz + (x * uy - y * ux);
}

// a o b: rotation a.q * b.q, translation a.t + a.q (b.t).
__device__ __forceinline__ Pose compose(const Pose& a, const Pose& b) {
  Pose o;
  quat_mul(a.w, a.x, a.y, a.z, b.w, b.x, b.y, b.z, o.w, o.x, o.y, o.z);
  float rx, ry, rz;
  quat_rotate(a.w, a.x, a.y, a.z, b.tx, b.ty, b.tz, rx, ry, rz);
  o.tx = a.tx + rx;
  o.ty = a.ty + ry;
  o.tz = a.tz + rz;
  return o;
}

// Rotation matrix rows of a unit quaternion (for transforming many points).
struct Rot3 {
  float m00, m01, m02, m10, m11, m12, m20, m21, m22;
};

__device__ __forceinline__ Rot3 quat_to_rot(float w, float x, float y, float z) {
  Rot3 r;
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, xz = x * z, yz = y * z;
  const float wx = w * x, wy = w * y, wz = w * z;
  r.m00 = 1.f - 2.f * (yy + zz); r.m01 = 2.f * (xy - wz);       r.m02 = 2.f * (xz + wy);
  r.m10 = 2.f * (xy + wz);       r.m11 = 1.f - 2.f * (xx + zz); r.m12 = 2.f * (yz - wx);
  r.m20 = 2.f * (xz - wy);       r.m21 = 2.f * (yz + wx);       r.m22 = 1.f - 2.f * (xx + yy);
  return r;
}

// Exponential of the one-parameter twist (v, omega) * q, for the twist table
// of one joint.  Closed form: rotation angle theta = |omega q|,
//   R = exp([omega q]x),  t = V (v q),
//   V = I + (1 - cos th)/th^2 [w]x + (th - sin th)/th^3 [w]x^2   (w = omega q).
// Below th^2 < 1e-12 the rotation is the identity and t = v q, the same
// small-angle rule as _fk_cuda_helpers.cuh:92-101.
//
// `kind` selects a cheaper exact form for the two joint types a URDF produces:
//   1 = pure rotation (v = 0): t = 0
//   2 = pure translation (omega = 0): R = I, t = v q
//   3 = general twist
//   4 = pure rotation staged as a unit axis (ox, oy, oz) and the norm |omega|
//       in vx: theta = |omega| q, q_rot = (cos(theta/2), sin(theta/2) axis),
//       the same rotation as kind 1 without its square root and division.
//
// kNativeTrig evaluates the kind-4 half-angle sine and cosine with the
// hardware v_sin/v_cos (absolute error ~1e-6) instead of the correctly
// rounded sincosf: for solvers' internal kinematics only, never for poses
// that are returned.
template <bool kNativeTrig = false>
__device__ __forceinline__ Pose joint_exp(int kind, float q,
                                          float vx, float vy, float vz,
                                          float ox, float oy, float oz) {
  Pose p = pose_identity();
  if (kind == 4) {
    const float th = vx * q;
    if (fabsf(th) < 1e-6f) return p;  // th^2 < 1e-12: the identity, as for kind 1
    float s, c;
    if (kNativeTrig) {
      s = __sinf(0.5f * th);
      c = __cosf(0.5f * th);
    } else {
      sincosf(0.5f * th, &s, &c);
    }
    p.w = c; p.x = s * ox; p.y = s * oy; p.z = s * oz;
    return p;
  }
  if (kind == 2) {
    p.tx = vx * q; p.ty = vy * q; p.tz = vz * q;
    return p;
  }
  const float wx = ox * q, wy = oy * q, wz = oz * q;
  const float th2 = wx * wx + wy * wy + wz * wz;
  if (th2 < 1e-12f) {
    p.tx = vx * q; p.ty = vy * q; p.tz = vz * q;
    return p;
  }
  const float th = sqrtf(th2);
  float s_half, c_half;
  sincosf(0.5f * th, &s_half, &c_half);
  const float k = s_half / th;
  p.w = c_half; p.x = k * wx; p.y = k * wy; p.z = k * wz;
  if (kind == 1) return p;
  float s, c;
  sincosf(th, &s, &c);
  const float a = s / th;
  const float b = (1.f - c) / th2;
  const float cc = (th - s) / (th2 * th);
  const float ux = vx * q, uy = vy * q, uz = vz * q;
  const float cx = wy * uz - wz * uy, cy = wz * ux - wx * uz, cz = wx * uy - wy * ux;
  const float wu = wx * ux + wy * uy + wz * uz;
  p.tx = a * ux + b * cx + cc * (wx * wu - th2 * ux);
  p.ty = a * uy + b * cy + cc * (wy * wu - th2 * uy);
  p.tz = a * uz + b * cz + cc * (wz * wu - th2 * uz);
  return p;
}

// Order-preserving float <-> uint32 key, so LDS unsigned atomic min can
// reduce signed distances.
__device__ __forceinline__ unsigned int float_to_ordered(float f) {
  const unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(unsigned int k) {
  const unsigned int u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

}  // namespace pk
