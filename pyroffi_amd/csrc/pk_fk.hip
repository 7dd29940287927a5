// Forward kinematics for gfx950: one lane = one configuration.
//
// A 64-lane workgroup owns 64 consecutive configurations.  It stages the joint
// records and the 64 x n_act configuration tile in LDS (coalesced loads); every
// lane walks the joint tree for its configuration and keeps each joint's world
// pose in LDS directly at its place in the lane's output row (a link pose is
// the pose of its parent joint, and in a URDF every joint has exactly one
// child link).  The finished tile -- 64 rows of n_out x 7 floats, the exact
// image of the block's contiguous output -- is copied out with 16-byte loads
// and non-temporal 16-byte stores, with no per-element index lookups.  Output
// positions the walk does not write (root links: identity; a second link on
// the same joint: a copy) are filled after the walk.  The row stride is odd
// (the row padded by one float when 7 n_out is even), so the per-lane walk
// accesses are free of bank conflicts.
//
// Measured phases for 65,536 Panda configurations (one wave per SIMD, run back
// to back): staging ~4.9 us (kernel start plus three dependent table loads),
// walk ~6.3 us, copy-out of 24 MB ~5.7 us with the earlier per-element lookups
// costing another ~5 us on top of the store floor.
//
// Reference behaviour: FkCudaImpl / fk_kernel_warp
// (src/pyroffi/cuda_kernels/_fk_cuda_kernel.cu:39-137, 236-454) and the link
// gather of src/pyroffi/_robot.py:180-196.
#include "pk_model.h"

namespace pk {

typedef float v4f __attribute__((ext_vector_type(4)));

// LDS image of one FK workgroup (float offsets into the dynamic LDS array).
struct FkLds {
  int recs, slot_of, pos, fill, lmap, flag, cfg, tile, total;
};

// n_extra: positions past the output row for joints without a child link.
__host__ __device__ inline int fk_row_stride(int n_out, int n_extra) { return (7 * (n_out + n_extra)) | 1; }

__host__ __device__ inline FkLds fk_lds_layout(int J, int n_act, int n_out, int n_extra) {
  FkLds o;
  o.recs = 0;
  o.slot_of = o.recs + static_cast<int>(sizeof(JointRec) / 4) * J;
  o.pos = o.slot_of + J;
  o.fill = o.pos + J;
  o.lmap = o.fill + n_out;
  o.flag = o.lmap + n_out;
  o.cfg = o.flag + 1;
  o.tile = ((o.cfg + n_act * kLaneStride) + 3) & ~3;  // 16-byte aligned
  o.total = o.tile + kWave * fk_row_stride(n_out, n_extra);
  return o;
}

// out_mode 0: joint poses (B, J, 7) in joint order; 1: link poses (B, L, 7).
// In mode 1 the link map must give each joint at most one position of its
// own: joints without a child link use the n_extra spare positions, and a
// workgroup that would need more writes NaN rows (a URDF-derived link map
// never does: every joint has exactly one child link).
__global__ __launch_bounds__(64) void fk_kernel(
    const float* __restrict__ cfg, long long batch, int n_act, RobotTables tables,
    const int* __restrict__ link_parent_joint, int n_links, int out_mode, int n_extra,
    float* __restrict__ out) {
  extern __shared__ float smem[];
  const int J = tables.n_joints;
  const int n_out = out_mode ? n_links : J;
  const int row = 7 * n_out;
  const int stride = fk_row_stride(n_out, n_extra);
  const FkLds lay = fk_lds_layout(J, n_act, n_out, n_extra);
  JointRec* recs = reinterpret_cast<JointRec*>(smem + lay.recs);
  int* slot_of = reinterpret_cast<int*>(smem + lay.slot_of);
  int* pos = reinterpret_cast<int*>(smem + lay.pos);    // per slot: float offset of its pose in a row
  int* fill = reinterpret_cast<int*>(smem + lay.fill);  // per output position: -1 walk, -2 identity, else copy source
  int* lmap = reinterpret_cast<int*>(smem + lay.lmap);  // per output position: its joint slot, or -1
  int* flag = reinterpret_cast<int*>(smem + lay.flag);
  float* cfg_tile = smem + lay.cfg;                     // [a][lane], stride 65
  float* tile = smem + lay.tile;                        // [lane * stride + offset]

  const int lane = threadIdx.x;
  const long long b0 = static_cast<long long>(blockIdx.x) * kWave;
  const int n_valid = static_cast<int>(min(static_cast<long long>(kWave), batch - b0));

  // The configuration block and the link -> joint map are loaded into
  // registers first, so their latency overlaps the table staging.
  const float* cfg_blk = cfg + b0 * n_act;
  const int n_cfg = n_valid * n_act;
  float cv[kMaxAct];
#pragma unroll
  for (int i = 0; i < kMaxAct; ++i) {
    const int f = lane + i * kWave;
    cv[i] = (f < n_cfg) ? cfg_blk[f] : 0.f;
  }
  constexpr int kOutIters = (kMaxJoints + 1 + kWave - 1) / kWave;
  int pjv[kOutIters];
#pragma unroll
  for (int i = 0; i < kOutIters; ++i) {
    const int o = lane + i * kWave;
    pjv[i] = (o < n_out) ? (out_mode ? link_parent_joint[o] : o) : -1;
  }
  if (lane == 0) *flag = 0;
  for (int i = lane; i < J; i += kWave) pos[i] = -1;
  stage_joint_records(tables, recs, slot_of);  // ends with a barrier
#pragma unroll
  for (int i = 0; i < kOutIters; ++i) {
    const int o = lane + i * kWave;
    if (o < n_out) {
      const int pj = pjv[i];
      lmap[o] = (pj >= 0 && pj < J) ? slot_of[pj] : -1;
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
  // Positions: the first output position of a slot is where the walk writes
  // its pose; later positions of the same slot copy it; positions without a
  // slot read the identity.
#pragma unroll
  for (int i = 0; i < kOutIters; ++i) {
    const int o = lane + i * kWave;
    if (o < n_out) {
      const int s = lmap[o];
      int first = o;
      for (int o2 = 0; o2 < o; ++o2)
        if (s >= 0 && lmap[o2] == s) { first = o2; break; }
      fill[o] = (s < 0) ? -2 : (first == o ? -1 : 7 * first);
      if (s >= 0 && first == o) pos[s] = 7 * o;
    }
  }
  __syncthreads();
  // Slots without an output position take the spare positions past the row.
  for (int i = lane; i < J; i += kWave) {
    if (pos[i] < 0) {
      int k = 0;
      for (int i2 = 0; i2 < i; ++i2) k += (pos[i2] < 0) ? 1 : 0;
      if (k < n_extra) {
        pos[i] = 7 * (n_out + k);
      } else {
        pos[i] = 0;  // no room: the rows are overwritten with NaN below
        *flag = 1;
      }
    }
  }
  __syncthreads();

  float* trow = tile + lane * stride;
  if (lane < n_valid) {
    for (int i = 0; i < J; ++i) {
      const JointRec& r = recs[i];
      const float qref = (r.src >= 0) ? cfg_tile[r.src * kLaneStride + lane] : 0.f;
      const float q = qref * r.mul + r.off;
      const Pose tpc = joint_local(r, q);
      const Pose tw = (r.parent_slot < 0) ? tpc : compose(load_pose(trow + pos[r.parent_slot]), tpc);
      float* d = trow + pos[i];
      d[0] = tw.w; d[1] = tw.x; d[2] = tw.y; d[3] = tw.z;
      d[4] = tw.tx; d[5] = tw.ty; d[6] = tw.tz;
    }
    const bool bad = *flag != 0;
    for (int o = 0; o < n_out; ++o) {
      const int fo = fill[o];
      if (fo == -1 && !bad) continue;
      float* d = trow + 7 * o;
      if (bad) {
#pragma unroll
        for (int c = 0; c < 7; ++c) d[c] = __builtin_nanf("");
      } else if (fo == -2) {
        d[0] = 1.f;
#pragma unroll
        for (int c = 1; c < 7; ++c) d[c] = 0.f;
      } else {
        float v[7];
#pragma unroll
        for (int c = 0; c < 7; ++c) v[c] = trow[fo + c];
#pragma unroll
        for (int c = 0; c < 7; ++c) d[c] = v[c];
      }
    }
  }
  __syncthreads();

  // Copy-out of the block's n_valid * row contiguous floats.
  float* out_blk = out + b0 * row;
  const int total = n_valid * row;
  if (stride == row) {
    // The tile is the output image: 16-byte loads and non-temporal stores
    // (the block starts 16-byte aligned: 64 rows of 7 n_out floats).
    const int n4 = total / 4;
    const v4f* t4 = reinterpret_cast<const v4f*>(tile);
    for (int f4 = lane; f4 < n4; f4 += kWave) __builtin_nontemporal_store(t4[f4], reinterpret_cast<v4f*>(out_blk) + f4);
    for (int f = 4 * n4 + lane; f < total; f += kWave) out_blk[f] = tile[f];
    return;
  }
  // Padded rows: element f = (configuration l, float r) of the block.
  const int n4 = total / 4;
  int l = (4 * lane) / row;
  int rem = 4 * lane - l * row;
  const int adv_l = (4 * kWave) / row;  // per pass: 256 floats = adv_l rows + adv_r
  const int adv_r = 4 * kWave - adv_l * row;
  for (int f4 = lane; f4 < n4; f4 += kWave) {
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int r = rem + k, lk = l;
      if (r >= row) { r -= row; ++lk; }  // row >= 7 > 3: at most one wrap
      v[k] = tile[lk * stride + r];
    }
    const v4f w = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(w, reinterpret_cast<v4f*>(out_blk) + f4);
    rem += adv_r;
    l += adv_l;
    if (rem >= row) { rem -= row; ++l; }
  }
  for (int f = 4 * n4 + lane; f < total; f += kWave) {
    const int lk = f / row, r = f - lk * row;
    out_blk[f] = tile[lk * stride + r];
  }
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
  // A tree has one more link than joints; spare positions cover joints beyond that.
  const int n_extra = out_mode ? std::max(0, t.n_joints - (n_links - 1)) : 0;
  const size_t lds = sizeof(float) * static_cast<size_t>(fk_lds_layout(t.n_joints, n_act, n_out, n_extra).total);
  PK_REQUIRE(lds <= 160 * 1024, "FK tile needs %zu bytes of LDS", lds);
  hipLaunchKernelGGL(fk_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kWave), lds,
                     static_cast<hipStream_t>(stream), cfg, static_cast<long long>(batch),
                     n_act, t, link_parent_joint, n_links, out_mode, n_extra, out);
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
