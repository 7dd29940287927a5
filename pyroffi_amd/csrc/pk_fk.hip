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
