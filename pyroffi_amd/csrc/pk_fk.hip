// Forward kinematics for gfx950: one lane = one configuration.
//
// A 64-lane workgroup owns 64 consecutive configurations.  It stages the joint
// records and the 64 x n_act configuration tile in LDS (coalesced loads), every
// lane walks the joint tree for its configuration with all poses held in an
// LDS tile, and the finished tile -- 64 x (n_out x 7) floats, contiguous in
// HBM -- is copied out with unit-stride stores.  The kernel is bound by the
// output write: 28 B per joint (or link) per configuration.
//
// Reference behaviour: FkCudaImpl / fk_kernel_warp
// (src/pyroffi/cuda_kernels/_fk_cuda_kernel.cu:39-137, 236-454) and the link
// gather of src/pyroffi/_robot.py:180-196.
#include "pk_model.h"

namespace pk {

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

  stage_joint_records(tables, recs, slot_of);
  for (int r = lane; r < row; r += kWave) {
    const int o = r / 7, c = r - 7 * (r / 7);
    int s;
    if (out_mode) {
      const int pj = link_parent_joint[o];
      s = (pj >= 0 && pj < J) ? slot_of[pj] : -1;
    } else {
      s = slot_of[o];
    }
    src_off[r] = (s >= 0) ? (s * 7 + c) * kLaneStride : (c == 0 ? -1 : -2);
  }
  // Configuration tile: element f of the contiguous (n_valid x n_act) block.
  const float* cfg_blk = cfg + b0 * n_act;
  for (int f = lane; f < n_valid * n_act; f += kWave) {
    const int l = f / n_act, a = f - l * n_act;
    cfg_tile[a * kLaneStride + l] = cfg_blk[f];
  }
  __syncthreads();

  if (lane < n_valid) {
    fk_walk(recs, J, [&](int a) { return cfg_tile[a * kLaneStride + lane]; }, poses, lane);
  }
  __syncthreads();

  // Copy-out: the block's rows are contiguous, n_valid * n_out * 7 floats.
  // Element f = lane + 64 i is float rem of configuration l; four elements
  // per pass keep their LDS reads in flight together.
  float* out_blk = out + b0 * row;
  const int total = n_valid * row;
  int l = lane / row;
  int rem = lane - l * row;
  constexpr int U = 4;
  for (int f0 = lane; f0 < total; f0 += U * kWave) {
    int ls[U], rs[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ls[u] = l;
      rs[u] = rem;
      rem += kWave;
      while (rem >= row) { rem -= row; ++l; }
    }
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int so = (f0 + u * kWave < total) ? src_off[rs[u]] : -2;
      v[u] = (so >= 0) ? poses[so + ls[u]] : (so == -1 ? 1.f : 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (f0 + u * kWave < total) out_blk[f0 + u * kWave] = v[u];
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

}  // namespace pk

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
