// Forward kinematics for gfx950: one lane = one configuration.
//
// A 64-lane workgroup owns 64 consecutive configurations.  It stages the joint
// records and the 64 x n_act configuration tile in LDS (coalesced loads), every
// lane walks the joint tree for its configuration with all poses held in an
// LDS tile, and the finished tile -- 64 x (n_out x 7) floats, contiguous in
// HBM -- is copied out with 16-byte stores.  The output is 28 B per joint (or
// link) per configuration; with one wave per SIMD (a 64-lane workgroup holds
// about 21 KB of LDS for Panda) the three phases -- staging, walk, copy-out --
// run back to back, so the kernel is latency-bound rather than at the write
// bandwidth.
//
// Reference behaviour: FkCudaImpl / fk_kernel_warp
// (src/pyroffi/cuda_kernels/_fk_cuda_kernel.cu:39-137, 236-454) and the link
// gather of src/pyroffi/_robot.py:180-196.
#include "pk_model.h"

namespace pk {

typedef float v4f __attribute__((ext_vector_type(4)));

// out_mode 0: joint poses (B, J, 7); 1: link poses (B, L, 7).
__global__ __launch_bounds__(64) void fk_kernel(
    const float* __restrict__ cfg, long long batch, int n_act, RobotTables tables,
    const int* __restrict__ link_parent_joint, int n_links, int out_mode,
    float* __restrict__ out) {
  extern __shared__ float smem[];
  const int J = tables.n_joints;
  JointRec* recs = reinterpret_cast<JointRec*>(smem);
  int* slot_of = reinterpret_cast<int*>(recs + J);
  const int n_out = out_mode ? n_links : J;
  const int row = n_out * 7;
  int* src_off = slot_of + J;   // per output float of a row: LDS offset of the pose component,
                                // -1 for an identity 1, -2 for an identity 0
  float* cfg_tile = reinterpret_cast<float*>(src_off + row);      // [a][lane], stride 65
  float* poses = cfg_tile + n_act * kLaneStride;                  // [slot*7+c][lane]

  const int lane = threadIdx.x;
  const long long b0 = static_cast<long long>(blockIdx.x) * kWave;
  const int n_valid = static_cast<int>(min(static_cast<long long>(kWave), batch - b0));

  // The configuration block and the link -> joint map are loaded into
  // registers first, so their latency overlaps the table staging.
  // Configuration tile: element f = lane + 64 i of the contiguous
  // (n_valid x n_act) block, at most kMaxAct elements per lane.
  const float* cfg_blk = cfg + b0 * n_act;
  const int n_cfg = n_valid * n_act;
  float cv[kMaxAct];
#pragma unroll
  for (int i = 0; i < kMaxAct; ++i) {
    const int f = lane + i * kWave;
    cv[i] = (f < n_cfg) ? cfg_blk[f] : 0.f;
  }
  constexpr int kRowIters = (7 * (kMaxJoints + 1) + kWave - 1) / kWave;
  int pjv[kRowIters];
#pragma unroll
  for (int i = 0; i < kRowIters; ++i) {
    const int r = lane + i * kWave;
    pjv[i] = (out_mode && r < row) ? link_parent_joint[r / 7] : -1;
  }

  stage_joint_records(tables, recs, slot_of);
#pragma unroll
  for (int i = 0; i < kRowIters; ++i) {
    const int r = lane + i * kWave;
    if (r < row) {
      const int o = r / 7, c = r - 7 * o;
      const int pj = out_mode ? pjv[i] : o;
      const int s = (pj >= 0 && pj < J) ? slot_of[pj] : -1;
      src_off[r] = (s >= 0) ? (s * 7 + c) * kLaneStride : (c == 0 ? -1 : -2);
    }
  }
#pragma unroll
  for (int i = 0; i < kMaxAct; ++i) {
    const int f = lane + i * kWave;
    if (f < n_cfg) {
      const int l = f / n_act, a = f - l * n_act;
      cfg_tile[a * kLaneStride + l] = cv[i];
    }
  }
  __syncthreads();

  if (lane < n_valid) {
    fk_walk(recs, J, [&](int a) { return cfg_tile[a * kLaneStride + lane]; }, poses, lane);
  }
  __syncthreads();

  // Copy-out: the block's rows are contiguous, n_valid * n_out * 7 floats,
  // written as 16-byte non-temporal stores (the block starts 16-byte
  // aligned: 64 rows of 7 n_out floats).  Lane f writes floats 4f .. 4f + 3;
  // float rem of configuration l comes from LDS offset src_off[rem] + l.
  float* out_blk = out + b0 * row;
  const int total = n_valid * row;
  const int n4 = total / 4;
  int l = (4 * lane) / row;
  int rem = 4 * lane - l * row;
  int adv_l = (4 * kWave) / row;            // per pass: 256 floats = adv_l rows + adv_r
  const int adv_r = 4 * kWave - adv_l * row;
  for (int f4 = lane; f4 < n4; f4 += kWave) {
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int r = rem + k, lk = l;
      if (r >= row) { r -= row; ++lk; }      // row >= 7 > 3: at most one wrap
      const int so = src_off[r];
      v[k] = (so >= 0) ? poses[so + lk] : (so == -1 ? 1.f : 0.f);
    }
    const v4f w = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(w, reinterpret_cast<v4f*>(out_blk) + f4);  // streamed past L2
    rem += adv_r;
    l += adv_l;
    if (rem >= row) { rem -= row; ++l; }
  }
  // Tail of a partial last block (total not a multiple of 4).
  for (int f = 4 * n4 + lane; f < total; f += kWave) {
    const int lk = f / row, r = f - lk * row;
    const int so = src_off[r];
    out_blk[f] = (so >= 0) ? poses[so + lk] : (so == -1 ? 1.f : 0.f);
  }
}

static size_t fk_lds_bytes(int J, int n_act, int n_out) {
  return sizeof(JointRec) * J + sizeof(int) * (J + 7 * n_out) +
         sizeof(float) * (static_cast<size_t>(n_act) + 7 * J) * kLaneStride;
}

static int launch_fk(void* stream, const float* cfg, int64_t batch, int32_t n_act,
                     const RobotTables& t, const int32_t* link_parent_joint,
                     int32_t n_links, int out_mode, float* out) {
  PK_REQUIRE(batch >= 0, "batch must be >= 0 (got %lld)", (long long)batch);
  PK_REQUIRE(t.n_joints >= 0 && t.n_joints <= kMaxJoints,
             "n_joints=%d outside [0, %d]", t.n_joints, kMaxJoints);
  PK_REQUIRE(n_act >= 0 && n_act <= kMaxAct, "n_act=%d outside [0, %d]", n_act, kMaxAct);
  if (out_mode) PK_REQUIRE(n_links >= 0 && n_links <= kMaxJoints + 1, "n_links=%d too large", n_links);
  if (batch == 0) return PK_OK;
  const int n_out = out_mode ? n_links : t.n_joints;
  if (n_out == 0) return PK_OK;
  PK_REQUIRE(cfg != nullptr && out != nullptr, "null cfg/out pointer");
  if (t.n_joints > 0)
    PK_REQUIRE(t.twists && t.parent_tf && t.parent_idx && t.act_idx && t.mimic_mul &&
                   t.mimic_off && t.mimic_act_idx && t.topo_inv,
               "null robot table pointer");
  if (out_mode) PK_REQUIRE(link_parent_joint != nullptr, "null link_parent_joint");
  const long long blocks = (batch + kWave - 1) / kWave;
  PK_REQUIRE(blocks <= 0x7fffffffLL, "batch too large");
  const size_t lds = fk_lds_bytes(t.n_joints, n_act, n_out);
  hipLaunchKernelGGL(fk_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kWave), lds,
                     static_cast<hipStream_t>(stream), cfg, static_cast<long long>(batch),
                     n_act, t, link_parent_joint, n_links, out_mode, out);
  return check_launch("fk_kernel");
}

// Adjoint of the link-pose FK: cotangents (B, L, 7) on the wxyz_xyz link
// poses -> dq (B, n_act), the vector-Jacobian product that reverse-mode AD of
// src/pyroffi/_robot.py:180-263 returns.
//
// Moving joint j at world pose T_j = (R_j, t_j) with local twist (v, w) per
// unit joint value has the spatial twist  w_s = R_j w,  v_s = R_j v + t_j x w_s,
// and every pose T below it moves as  dT/dq_j = (w_s, v_s)^ T:
//   dt = w_s x t + v_s,   d quat = 1/2 (0, w_s) * quat.
// So a link's cotangent (g_q = (g_w, g_u), g_t) acts on q_j through the wrench
//   F = g_t,   tau = t x g_t + 1/2 (w g_u - g_w u + u x g_u)   (quat = (w, u)),
// and dL/dq_j = v_s . F_sub + w_s . tau_sub with the wrench summed over the
// subtree below j.  Each lane recomputes its configuration's FK (no pose
// residual in HBM), adds every link's wrench to its parent joint, then walks
// the slots in reverse topological order: a slot's wrench is complete when it
// is reached, is projected on the slot's twist (times the mimic multiplier,
// into the actuated source), and is passed to the parent slot.
__global__ __launch_bounds__(64) void fk_vjp_kernel(
    const float* __restrict__ cfg, long long batch, int n_act, RobotTables tables,
    const int* __restrict__ link_parent_joint, int n_links, const float* __restrict__ cot,
    float* __restrict__ out_dq) {
  extern __shared__ float smem[];
  const int J = tables.n_joints;
  JointRec* recs = reinterpret_cast<JointRec*>(smem);
  int* slot_of = reinterpret_cast<int*>(recs + J);
  float* cfg_tile = reinterpret_cast<float*>(slot_of + J);   // [a][lane], stride 65
  float* poses = cfg_tile + n_act * kLaneStride;             // [slot*7+c][lane]
  float* wrench = poses + 7 * J * kLaneStride;               // [slot*6+c][lane]

  const int lane = threadIdx.x;
  const long long b0 = static_cast<long long>(blockIdx.x) * kWave;
  const int n_valid = static_cast<int>(min(static_cast<long long>(kWave), batch - b0));
  const float* cfg_blk = cfg + b0 * n_act;
  const int n_cfg = n_valid * n_act;
  stage_joint_records(tables, recs, slot_of);
  for (int f = lane; f < n_cfg; f += kWave) {
    const int l = f / n_act, a = f - l * n_act;
    cfg_tile[a * kLaneStride + l] = cfg_blk[f];
  }
  for (int i = 0; i < 6 * J; ++i) wrench[i * kLaneStride + lane] = 0.f;
  __syncthreads();
  if (lane >= n_valid) return;  // no barrier below
  fk_walk(recs, J, [&](int a) { return cfg_tile[a * kLaneStride + lane]; }, poses, lane);

  const float* g = cot + (b0 + lane) * static_cast<long long>(n_links) * 7;
  for (int l = 0; l < n_links; ++l) {
    const int pj = link_parent_joint[l];
    const int s = (pj >= 0 && pj < J) ? slot_of[pj] : -1;
    if (s < 0) continue;  // a root link's pose is the constant identity
    const Pose T = lds_get_pose(poses, s, lane);
    const float gw = g[7 * l], gx = g[7 * l + 1], gy = g[7 * l + 2], gz = g[7 * l + 3];
    const float fx = g[7 * l + 4], fy = g[7 * l + 5], fz = g[7 * l + 6];
    const float tau_x = (T.ty * fz - T.tz * fy) + 0.5f * (T.w * gx - gw * T.x + (T.y * gz - T.z * gy));
    const float tau_y = (T.tz * fx - T.tx * fz) + 0.5f * (T.w * gy - gw * T.y + (T.z * gx - T.x * gz));
    const float tau_z = (T.tx * fy - T.ty * fx) + 0.5f * (T.w * gz - gw * T.z + (T.x * gy - T.y * gx));
    float* W = wrench + s * 6 * kLaneStride + lane;
    W[0] += fx; W[kLaneStride] += fy; W[2 * kLaneStride] += fz;
    W[3 * kLaneStride] += tau_x; W[4 * kLaneStride] += tau_y; W[5 * kLaneStride] += tau_z;
  }
  // dq accumulates in the (no longer needed) configuration tile column.
  for (int a = 0; a < n_act; ++a) cfg_tile[a * kLaneStride + lane] = 0.f;
  for (int i = J - 1; i >= 0; --i) {
    const JointRec& r = recs[i];
    const float* W = wrench + i * 6 * kLaneStride + lane;
    const float F[3] = {W[0], W[kLaneStride], W[2 * kLaneStride]};
    const float tau[3] = {W[3 * kLaneStride], W[4 * kLaneStride], W[5 * kLaneStride]};
    if (r.src >= 0 && r.kind != 0) {
      float v[3] = {0.f, 0.f, 0.f}, w[3] = {0.f, 0.f, 0.f};
      if (r.kind == 4) {
        w[0] = r.ang[0] * r.lin[0]; w[1] = r.ang[1] * r.lin[0]; w[2] = r.ang[2] * r.lin[0];
      } else {
        v[0] = r.lin[0]; v[1] = r.lin[1]; v[2] = r.lin[2];
        if (r.kind == 3) { w[0] = r.ang[0]; w[1] = r.ang[1]; w[2] = r.ang[2]; }
      }
      const Pose T = lds_get_pose(poses, i, lane);
      float ws[3], vs[3];
      quat_rotate(T.w, T.x, T.y, T.z, w[0], w[1], w[2], ws[0], ws[1], ws[2]);
      quat_rotate(T.w, T.x, T.y, T.z, v[0], v[1], v[2], vs[0], vs[1], vs[2]);
      vs[0] += T.ty * ws[2] - T.tz * ws[1];
      vs[1] += T.tz * ws[0] - T.tx * ws[2];
      vs[2] += T.tx * ws[1] - T.ty * ws[0];
      const float d = vs[0] * F[0] + vs[1] * F[1] + vs[2] * F[2] + ws[0] * tau[0] + ws[1] * tau[1] +
                      ws[2] * tau[2];
      cfg_tile[r.src * kLaneStride + lane] += r.mul * d;
    }
    if (r.parent_slot >= 0) {
      float* P = wrench + r.parent_slot * 6 * kLaneStride + lane;
      P[0] += F[0]; P[kLaneStride] += F[1]; P[2 * kLaneStride] += F[2];
      P[3 * kLaneStride] += tau[0]; P[4 * kLaneStride] += tau[1]; P[5 * kLaneStride] += tau[2];
    }
  }
  float* o = out_dq + (b0 + lane) * n_act;
  for (int a = 0; a < n_act; ++a) o[a] = cfg_tile[a * kLaneStride + lane];
}

static size_t fk_vjp_lds_bytes(int J, int n_act) {
  return sizeof(JointRec) * J + sizeof(int) * J +
         sizeof(float) * (static_cast<size_t>(n_act) + 13 * J) * kLaneStride;
}

}  // namespace pk

extern "C" int pk_fk_vjp_f32(void* stream, const float* cfg, int64_t batch, int32_t n_act,
                             const float* twists, const float* parent_tf,
                             const int32_t* parent_idx, const int32_t* act_idx,
                             const float* mimic_mul, const float* mimic_off,
                             const int32_t* mimic_act_idx, const int32_t* topo_inv,
                             int32_t n_joints, const int32_t* link_parent_joint,
                             int32_t n_links, const float* cot_links, float* out_dq) {
  PK_REQUIRE(batch >= 0, "batch must be >= 0 (got %lld)", (long long)batch);
  PK_REQUIRE(n_joints >= 0 && n_joints <= pk::kMaxJoints, "n_joints=%d outside [0, %d]", n_joints,
             pk::kMaxJoints);
  PK_REQUIRE(n_act >= 0 && n_act <= pk::kMaxAct, "n_act=%d outside [0, %d]", n_act, pk::kMaxAct);
  PK_REQUIRE(n_links >= 0 && n_links <= pk::kMaxJoints + 1, "n_links=%d too large", n_links);
  if (batch == 0 || n_act == 0) return PK_OK;
  PK_REQUIRE(cfg && out_dq && (n_links == 0 || (cot_links && link_parent_joint)), "null pointer");
  if (n_joints > 0)
    PK_REQUIRE(twists && parent_tf && parent_idx && act_idx && mimic_mul && mimic_off &&
                   mimic_act_idx && topo_inv,
               "null robot table pointer");
  const long long blocks = (batch + pk::kWave - 1) / pk::kWave;
  PK_REQUIRE(blocks <= 0x7fffffffLL, "batch too large");
  const size_t lds = pk::fk_vjp_lds_bytes(n_joints, n_act);
  PK_REQUIRE(lds <= 160 * 1024, "FK adjoint needs %zu bytes of LDS", lds);
  pk::RobotTables t{twists, parent_tf, parent_idx, act_idx, mimic_mul,
                    mimic_off, mimic_act_idx, topo_inv, n_joints, n_act};
  hipLaunchKernelGGL(pk::fk_vjp_kernel, dim3(static_cast<unsigned>(blocks)), dim3(pk::kWave), lds,
                     static_cast<hipStream_t>(stream), cfg, static_cast<long long>(batch), n_act, t,
                     link_parent_joint, n_links, cot_links, out_dq);
  return pk::check_launch("fk_vjp_kernel");
}

extern "C" int pk_fk_f32(void* stream, const float* cfg, int64_t batch, int32_t n_act,
                         const float* twists, const float* parent_tf,
                         const int32_t* parent_idx, const int32_t* act_idx,
                         const float* mimic_mul, const float* mimic_off,
                         const int32_t* mimic_act_idx, const int32_t* topo_inv,
                         const int32_t* fk_level_starts, const int32_t* fk_level_joints,
                         int32_t n_joints, int32_t n_levels, float* out) {
  (void)fk_level_starts;
  (void)fk_level_joints;
  PK_REQUIRE(n_levels >= 0 && n_levels <= pk::kMaxJoints, "n_levels=%d outside [0, %d]",
             n_levels, pk::kMaxJoints);
  pk::RobotTables t{twists, parent_tf, parent_idx, act_idx, mimic_mul,
                    mimic_off, mimic_act_idx, topo_inv, n_joints, n_act};
  return pk::launch_fk(stream, cfg, batch, n_act, t, nullptr, 0, 0, out);
}

extern "C" int pk_fk_links_f32(void* stream, const float* cfg, int64_t batch, int32_t n_act,
                               const float* twists, const float* parent_tf,
                               const int32_t* parent_idx, const int32_t* act_idx,
                               const float* mimic_mul, const float* mimic_off,
                               const int32_t* mimic_act_idx, const int32_t* topo_inv,
                               int32_t n_joints, const int32_t* link_parent_joint,
                               int32_t n_links, float* out_links) {
  pk::RobotTables t{twists, parent_tf, parent_idx, act_idx, mimic_mul,
                    mimic_off, mimic_act_idx, topo_inv, n_joints, n_act};
  return pk::launch_fk(stream, cfg, batch, n_act, t, link_parent_joint, n_links, 1, out_links);
}
