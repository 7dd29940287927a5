// Kinematic-tree staging shared by the FK, sweep and IK kernels.
//
// A workgroup copies the reference's joint tables (twists, parent_tf,
// parent/actuated/mimic indices, topo order) from HBM into one LDS record per
// joint, ordered by topological position ("slot").  Every lane then walks the
// slots in order for its own configuration: all lanes read the same record at
// the same time (an LDS broadcast), so the walk costs no divergence.
//
// The per-joint rule is the reference's (_fk_cuda_helpers.cuh:171-207 and
// src/pyroffi/_robot.py:206-254):
//   q_j   = (src >= 0 ? cfg[src] : 0) * mimic_mul + mimic_off,
//           src = mimic_act_idx if set, else act_idx
//   T_pc  = parent_tf                         if src < 0 and mimic_off == 0
//         = parent_tf o exp(twist * q_j)      otherwise
//   T_w   = T_w[parent] o T_pc  (T_pc for roots).
#pragma once

#include "pk_common.h"
#include "pk_se3.h"

namespace pk {

struct JointRec {
  float ptf[7];      // parent -> joint transform
  float lin[3];      // twist, linear part
  float ang[3];      // twist, angular part
  float mul, off;    // mimic multiplier / offset
  int src;           // actuated source index, -1 when none
  int parent_slot;   // slot of the parent joint, -1 for roots
  int joint;         // original joint index
  int kind;          // 0 constant (T_pc = ptf), 2 translation, 3 general, 4 rotation
                     // (ang = unit axis, lin[0] = its original norm)
  int pad;
};
static_assert(sizeof(JointRec) == 20 * 4, "JointRec layout");

struct RobotTables {
  const float* twists;
  const float* parent_tf;
  const int* parent_idx;
  const int* act_idx;
  const float* mimic_mul;
  const float* mimic_off;
  const int* mimic_act_idx;
  const int* topo_inv;
  int n_joints;
  int n_act;
};

// Fill recs[0..J) (and slot_of[0..J), the joint -> slot map) cooperatively.
// Out-of-range indices in the tables degrade to "root"/"no source" instead of
// reading out of bounds.  Ends with a workgroup barrier.
__device__ inline void stage_joint_records(const RobotTables& t, JointRec* recs,
                                           int* slot_of) {
  const int J = t.n_joints;
  for (int i = threadIdx.x; i < J; i += blockDim.x) slot_of[i] = -1;
  __syncthreads();
  for (int i = threadIdx.x; i < J; i += blockDim.x) {
    const int j = t.topo_inv[i];
    if (j >= 0 && j < J) slot_of[j] = i;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < J; i += blockDim.x) {
    int j = t.topo_inv[i];
    if (j < 0 || j >= J) j = 0;
    JointRec r;
#pragma unroll
    for (int c = 0; c < 7; ++c) r.ptf[c] = t.parent_tf[j * 7 + c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r.lin[c] = t.twists[j * 6 + c];
      r.ang[c] = t.twists[j * 6 + 3 + c];
    }
    r.mul = t.mimic_mul[j];
    r.off = t.mimic_off[j];
    const int m = t.mimic_act_idx[j];
    int src = (m != -1) ? m : t.act_idx[j];
    if (src < 0 || src >= t.n_act) src = -1;
    r.src = src;
    const int p = t.parent_idx[j];
    int ps = (p >= 0 && p < J) ? slot_of[p] : -1;
    if (ps >= i) ps = -1;  // not topologically before us: invalid table
    r.parent_slot = ps;
    r.joint = j;
    const float l2 = r.lin[0] * r.lin[0] + r.lin[1] * r.lin[1] + r.lin[2] * r.lin[2];
    const float a2 = r.ang[0] * r.ang[0] + r.ang[1] * r.ang[1] + r.ang[2] * r.ang[2];
    int kind;
    if (src == -1 && r.off == 0.f) kind = 0;
    else if (a2 == 0.f) kind = 2;
    else if (l2 == 0.f) kind = 4;
    else kind = 3;
    if (kind == 4) {
      // Pure rotation, staged as the unit axis (in ang) and its norm (in the
      // unused lin[0]), which joint_exp kind 4 evaluates without a square
      // root or a division.
      const float na = sqrtf(a2);
#pragma unroll
      for (int c = 0; c < 3; ++c) r.ang[c] /= na;
      r.lin[0] = na;
    }
    r.kind = kind;
    r.pad = 0;
    recs[i] = r;
  }
  __syncthreads();
}

__device__ __forceinline__ Pose load_pose(const float* p) {
  Pose o;
  o.w = p[0]; o.x = p[1]; o.y = p[2]; o.z = p[3];
  o.tx = p[4]; o.ty = p[5]; o.tz = p[6];
  return o;
}

// Parent -> joint transform of one record for joint value q.
__device__ __forceinline__ Pose joint_local(const JointRec& r, float q) {
  const Pose ptf = load_pose(r.ptf);
  if (r.kind == 0) return ptf;
  const Pose d = joint_exp(r.kind, q, r.lin[0], r.lin[1], r.lin[2], r.ang[0], r.ang[1], r.ang[2]);
  return compose(ptf, d);
}

// Per-lane pose store in LDS: component c of slot i for lane l lives at
// base[(i * 7 + c) * stride + l].  stride = 65 keeps both the per-lane walk
// (consecutive lanes) and the row-major copy-out (consecutive components)
// free of bank conflicts.
constexpr int kLaneStride = kWave + 1;

__device__ __forceinline__ Pose lds_get_pose(const float* base, int slot, int lane) {
  const float* p = base + slot * 7 * kLaneStride + lane;
  Pose o;
  o.w = p[0]; o.x = p[kLaneStride]; o.y = p[2 * kLaneStride]; o.z = p[3 * kLaneStride];
  o.tx = p[4 * kLaneStride]; o.ty = p[5 * kLaneStride]; o.tz = p[6 * kLaneStride];
  return o;
}

__device__ __forceinline__ void lds_put_pose(float* base, int slot, int lane, const Pose& v) {
  float* p = base + slot * 7 * kLaneStride + lane;
  p[0] = v.w; p[kLaneStride] = v.x; p[2 * kLaneStride] = v.y; p[3 * kLaneStride] = v.z;
  p[4 * kLaneStride] = v.tx; p[5 * kLaneStride] = v.ty; p[6 * kLaneStride] = v.tz;
}

// Walk the tree for this lane's configuration; world poses of all slots go to
// `poses` (LDS, layout above).  cfg_lane(a) returns the lane's actuated value a.
template <typename CfgFn>
__device__ __forceinline__ void fk_walk(const JointRec* recs, int J, CfgFn cfg_lane,
                                        float* poses, int lane) {
  for (int i = 0; i < J; ++i) {
    const JointRec& r = recs[i];
    const float qref = (r.src >= 0) ? cfg_lane(r.src) : 0.f;
    const float q = qref * r.mul + r.off;
    const Pose tpc = joint_local(r, q);
    const Pose tw = (r.parent_slot < 0) ? tpc : compose(lds_get_pose(poses, r.parent_slot, lane), tpc);
    lds_put_pose(poses, i, lane, tw);
  }
}

// Orders one wave's LDS writes before its own later LDS reads from other
// lanes.  LDS instructions of a wave execute in order, so a wavefront-scope
// barrier only has to stop the compiler from moving accesses across it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Reductions with DPP lane moves (no LDS traffic).  Quad swaps and the
// half-row / row mirrors combine lanes inside 16-lane rows; row_bcast15 and
// row_bcast31 fold rows 0-2 into row 3.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

// Sum over the 64 lanes; the result is wave-uniform (read from lane 63).
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v = dpp_add<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_add<0x140, 0xf>(v);  // row_mirror
  v = dpp_add<0x142, 0xa>(v);  // row_bcast15 into rows 1, 3
  v = dpp_add<0x143, 0xc>(v);  // row_bcast31 into rows 2, 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Sum over each aligned group of 8 lanes, returned in every lane of the group.
__device__ __forceinline__ float group8_sum_dpp(float v) {
  v = dpp_add<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  return dpp_add<0x141, 0xf>(v);  // row_half_mirror: lane i + lane 7 - i
}

}  // namespace pk
