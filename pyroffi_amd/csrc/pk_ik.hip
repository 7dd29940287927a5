// C ABI of the multi-seed LM IK (kernels in pk_ik.h), winner selection.
#include "pk_ik.h"

#include <algorithm>

namespace pk {

static int dispatch_ik(void* stream, const IkArgs& a, bool batched, int grid_waves) {
  // Padded widths: the moving-joint count (actuated + mimic + offset-only
  // joints) must fit; a robot that exceeds it gets NaN errors (see header).
  const int nee = a.n_ee <= 1 ? 1 : (a.n_ee <= 2 ? 2 : 4);
  if (a.n_act <= 8) {
    if (nee == 1) return launch_ik_8_1(stream, a, batched, grid_waves);
    if (nee == 2) return launch_ik_8_2(stream, a, batched, grid_waves);
    return launch_ik_8_4(stream, a, batched, grid_waves);
  }
  if (nee == 1) return launch_ik_16_1(stream, a, batched, grid_waves);
  if (nee == 2) return launch_ik_16_2(stream, a, batched, grid_waves);
  return launch_ik_16_4(stream, a, batched, grid_waves);
}

// Per problem: argmin over its seeds of err + w |q - prev|^2, first index on
// ties.  Seeds of problem p are [p S, (p + 1) S) or, with seed_start, the CSR
// range [seed_start[p], seed_start[p + 1]).
__global__ __launch_bounds__(64) void select_kernel(const float* __restrict__ cfgs,
                                                    const float* __restrict__ errs,
                                                    const float* __restrict__ prev,
                                                    const int* __restrict__ seed_start, int S, int A,
                                                    float cw, float* __restrict__ out_best,
                                                    int* __restrict__ out_idx) {
  const int p = blockIdx.x;
  const int lane = threadIdx.x;
  const long long s0 = seed_start ? seed_start[p] : static_cast<long long>(p) * S;
  const long long s1 = seed_start ? seed_start[p + 1] : s0 + S;
  const int n = static_cast<int>(max(0LL, s1 - s0));
  float bv = INFINITY;
  int bi = 0x7fffffff;
  for (int s = lane; s < n; s += kWave) {
    const long long g = s0 + s;
    float v = errs[g];
    if (cw != 0.f) {
      float d2 = 0.f;
      for (int i = 0; i < A; ++i) {
        const float d = cfgs[g * A + i] - prev[static_cast<long long>(p) * A + i];
        d2 += d * d;
      }
      v += cw * d2;
    }
    if (v != v) v = INFINITY;  // a NaN error never wins over a number
    if (v < bv || (v == bv && s < bi)) { bv = v; bi = s; }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float ov = __shfl_xor(bv, off, kWave);
    const int oi = __shfl_xor(bi, off, kWave);
    if (ov < bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
  }
  if (bi == 0x7fffffff) bi = 0;
  if (n > 0)
    for (int i = lane; i < A; i += kWave)
      out_best[static_cast<long long>(p) * A + i] = cfgs[(s0 + bi) * A + i];
  if (out_idx != nullptr && lane == 0) out_idx[p] = bi;
}

}  // namespace pk

namespace {

int fill_ik_args(pk::IkArgs& a, const float* seeds, int32_t n_problems, int32_t n_act,
                 const float* twists, const float* parent_tf, const int32_t* parent_idx,
                 const int32_t* act_idx, const float* mimic_mul, const float* mimic_off,
                 const int32_t* mimic_act_idx, const int32_t* topo_inv, int32_t n_joints,
                 const int32_t* target_jnts, const int32_t* ancestor_masks, int32_t n_ee,
                 const float* target_T, const float* robot_spheres_local,
                 const int32_t* robot_sphere_joint_idx, int32_t n_robot_spheres,
                 const float* world_spheres, int32_t n_world_spheres, const float* world_capsules,
                 int32_t n_world_capsules, const float* world_boxes, int32_t n_world_boxes,
                 const float* world_halfspaces, int32_t n_world_halfspaces, const float* lower,
                 const float* upper, const int32_t* fixed_mask, int32_t max_iter, float pos_weight,
                 float ori_weight, float lambda_init, float eps_pos, float eps_ori,
                 int32_t enable_collision, float collision_weight, float collision_margin,
                 float* out_cfgs, float* out_err) {
  PK_REQUIRE(n_problems >= 0 && n_act >= 0 && n_ee >= 0, "negative size");
  PK_REQUIRE(n_joints >= 0 && n_joints <= pk::kMaxJoints, "n_joints=%d outside [0, %d]", n_joints,
             pk::kMaxJoints);
  PK_REQUIRE(n_act <= pk::kMaxAct, "n_act=%d exceeds %d", n_act, pk::kMaxAct);
  PK_REQUIRE(n_ee <= pk::kMaxEE, "n_ee=%d exceeds %d", n_ee, pk::kMaxEE);
  PK_REQUIRE(max_iter >= 0, "max_iter must be >= 0");
  const bool coll = enable_collision != 0 && n_robot_spheres > 0;
  PK_REQUIRE(n_world_capsules >= 0 && n_world_boxes >= 0 && n_world_halfspaces >= 0 &&
                 7 * n_world_capsules + 15 * n_world_boxes + 6 * n_world_halfspaces <= 8192,
             "non-sphere obstacles take more than 8192 floats (%d capsules, %d boxes, %d half-spaces)",
             n_world_capsules, n_world_boxes, n_world_halfspaces);
  if (coll)
    PK_REQUIRE((n_world_capsules == 0 || world_capsules) && (n_world_boxes == 0 || world_boxes) &&
                   (n_world_halfspaces == 0 || world_halfspaces),
               "null world array");
  PK_REQUIRE(n_robot_spheres >= 0 && n_robot_spheres <= 4096, "n_robot_spheres=%d outside [0, 4096]",
             n_robot_spheres);
  PK_REQUIRE(n_world_spheres >= 0 && n_world_spheres <= 4096, "n_world_spheres=%d outside [0, 4096]",
             n_world_spheres);
  PK_REQUIRE(seeds && out_cfgs && out_err && lower && upper && fixed_mask, "null pointer");
  PK_REQUIRE(n_ee == 0 || (target_jnts && ancestor_masks && target_T), "null target pointer");
  if (coll) PK_REQUIRE(robot_spheres_local && robot_sphere_joint_idx &&
                           (n_world_spheres == 0 || world_spheres),
                       "null collision pointer");
  a.seeds = seeds;
  a.n_problems = n_problems;
  a.n_act = n_act;
  a.tables = pk::RobotTables{twists, parent_tf, parent_idx, act_idx, mimic_mul,
                             mimic_off, mimic_act_idx, topo_inv, n_joints, n_act};
  a.target_jnts = target_jnts;
  a.ancestor_masks = ancestor_masks;
  a.n_ee = n_ee;
  a.target_T = target_T;
  a.robot_spheres = robot_spheres_local;
  a.robot_sphere_joint = robot_sphere_joint_idx;
  a.n_robot_spheres = coll ? n_robot_spheres : 0;
  a.world = world_spheres;
  a.n_world = coll ? n_world_spheres : 0;
  a.world_caps = world_capsules;
  a.n_caps = coll ? n_world_capsules : 0;
  a.world_boxes = world_boxes;
  a.n_boxes = coll ? n_world_boxes : 0;
  a.world_hs = world_halfspaces;
  a.n_hs = coll ? n_world_halfspaces : 0;
  a.lower = lower;
  a.upper = upper;
  a.fixed_mask = fixed_mask;
  a.max_iter = max_iter;
  a.pos_weight = pos_weight;
  a.ori_weight = ori_weight;
  a.lambda_init = lambda_init;
  a.eps_pos = eps_pos;
  a.eps_ori = eps_ori;
  a.enable_collision = coll ? 1 : 0;
  a.collision_weight = collision_weight;
  a.collision_margin = collision_margin;
  a.out_cfgs = out_cfgs;
  a.out_err = out_err;
  a.iter_begin = 0;
  a.iter_end = max_iter;
  return PK_OK;
}

}  // namespace

extern "C" int pk_ls_ik_f32(
    void* stream, const float* seeds, int32_t n_problems, int32_t n_seeds, int32_t n_act,
    const float* twists, const float* parent_tf, const int32_t* parent_idx,
    const int32_t* act_idx, const float* mimic_mul, const float* mimic_off,
    const int32_t* mimic_act_idx, const int32_t* topo_inv, int32_t n_joints,
    const int32_t* target_jnts, const int32_t* ancestor_masks, int32_t n_ee,
    const float* target_T, const float* robot_spheres_local,
    const int32_t* robot_sphere_joint_idx, int32_t n_robot_spheres,
    const float* world_spheres, int32_t n_world_spheres, const float* world_capsules,
    int32_t n_world_capsules, const float* world_boxes, int32_t n_world_boxes,
    const float* world_halfspaces, int32_t n_world_halfspaces, const float* lower,
    const float* upper, const int32_t* fixed_mask, int32_t max_iter, float pos_weight,
    float ori_weight, float lambda_init, float eps_pos, float eps_ori,
    int32_t enable_collision, float collision_weight, float collision_margin,
    float* out_cfgs, float* out_err) {
  PK_REQUIRE(n_seeds >= 0, "negative size");
  PK_REQUIRE(n_problems <= 65535, "n_problems=%d exceeds the 65535 grid limit", n_problems);
  if (n_problems == 0 || n_seeds == 0) return PK_OK;
  pk::IkArgs a{};
  const int st = fill_ik_args(a, seeds, n_problems, n_act, twists, parent_tf, parent_idx, act_idx,
                              mimic_mul, mimic_off, mimic_act_idx, topo_inv, n_joints, target_jnts,
                              ancestor_masks, n_ee, target_T, robot_spheres_local,
                              robot_sphere_joint_idx, n_robot_spheres, world_spheres,
                              n_world_spheres, world_capsules, n_world_capsules, world_boxes,
                              n_world_boxes, world_halfspaces, n_world_halfspaces,
                              lower, upper, fixed_mask, max_iter, pos_weight, ori_weight,
                              lambda_init, eps_pos, eps_ori, enable_collision, collision_weight,
                              collision_margin, out_cfgs, out_err);
  if (st != PK_OK) return st;
  a.n_seeds = n_seeds;
  return pk::dispatch_ik(stream, a, false, 0);
}

extern "C" int pk_ls_ik_batched_f32(
    void* stream, const float* seeds, const int32_t* problem_seed_start, int32_t n_problems,
    int64_t total_seeds, int32_t n_act, const float* twists, const float* parent_tf,
    const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
    const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv,
    int32_t n_joints, const int32_t* target_jnts, const int32_t* ancestor_masks, int32_t n_ee,
    const float* target_T, const float* robot_spheres_local,
    const int32_t* robot_sphere_joint_idx, int32_t n_robot_spheres, const float* world_spheres,
    int32_t n_world_spheres, const pk_world_grid* grids, int32_t n_grids, const float* lower,
    const float* upper, const int32_t* fixed_mask, int32_t max_iter, int32_t iter_begin,
    int32_t iter_end, float* state, float pos_weight, float ori_weight, float lambda_init,
    float eps_pos, float eps_ori, int32_t enable_collision, float collision_weight,
    float collision_margin, int32_t grid_waves, float* out_cfgs, float* out_err,
    uint64_t* work_counter) {
  PK_REQUIRE(total_seeds >= 0, "negative size");
  if (n_problems == 0 || total_seeds == 0) return PK_OK;
  PK_REQUIRE(problem_seed_start != nullptr, "null problem_seed_start");
  PK_REQUIRE(0 <= iter_begin && iter_begin <= iter_end && iter_end <= max_iter,
             "iteration range [%d, %d) outside [0, %d]", iter_begin, iter_end, max_iter);
  PK_REQUIRE(state != nullptr || (iter_begin == 0 && iter_end == max_iter),
             "a partial iteration range needs a state buffer");
  pk::IkArgs a{};
  const int st = fill_ik_args(a, seeds, n_problems, n_act, twists, parent_tf, parent_idx, act_idx,
                              mimic_mul, mimic_off, mimic_act_idx, topo_inv, n_joints, target_jnts,
                              ancestor_masks, n_ee, target_T, robot_spheres_local,
                              robot_sphere_joint_idx, n_robot_spheres, world_spheres,
                              n_world_spheres, nullptr, 0, nullptr, 0, nullptr, 0, lower, upper,
                              fixed_mask, max_iter,
                              pos_weight, ori_weight, lambda_init, eps_pos, eps_ori,
                              enable_collision, collision_weight, collision_margin, out_cfgs,
                              out_err);
  if (st != PK_OK) return st;
  a.problem_seed_start = problem_seed_start;
  a.total_seeds = total_seeds;
  a.state = state;
  a.iter_begin = iter_begin;
  a.iter_end = iter_end;
  a.work = reinterpret_cast<unsigned long long*>(work_counter);
  a.grid.n = 0;
  if (a.enable_collision) {
    const int gs = pk::fill_grid_args(grids, n_grids, collision_margin, a.grid);
    if (gs != PK_OK) return gs;
  }
  return pk::dispatch_ik(stream, a, true, grid_waves);
}

extern "C" int pk_ls_ik_select_f32(void* stream, const float* cfgs, const float* errs,
                                   const float* previous, int32_t n_problems, int32_t n_seeds,
                                   int32_t n_act, float continuity_weight, float* out_best,
                                   int32_t* out_idx) {
  PK_REQUIRE(n_problems >= 0 && n_seeds >= 0 && n_act >= 0, "negative size");
  if (n_problems == 0) return PK_OK;
  PK_REQUIRE(n_seeds > 0, "n_seeds must be > 0");
  PK_REQUIRE(cfgs && errs && out_best && (continuity_weight == 0.f || previous), "null pointer");
  hipLaunchKernelGGL(pk::select_kernel, dim3(n_problems), dim3(pk::kWave), 0,
                     static_cast<hipStream_t>(stream), cfgs, errs, previous, nullptr, n_seeds, n_act,
                     continuity_weight, out_best, out_idx);
  return pk::check_launch("select_kernel");
}

extern "C" int pk_ls_ik_select_segments_f32(void* stream, const float* cfgs, const float* errs,
                                            const float* previous, const int32_t* problem_seed_start,
                                            int32_t n_problems, int32_t n_act,
                                            float continuity_weight, float* out_best,
                                            int32_t* out_idx) {
  PK_REQUIRE(n_problems >= 0 && n_act >= 0, "negative size");
  if (n_problems == 0) return PK_OK;
  PK_REQUIRE(cfgs && errs && out_best && problem_seed_start && (continuity_weight == 0.f || previous),
             "null pointer");
  hipLaunchKernelGGL(pk::select_kernel, dim3(n_problems), dim3(pk::kWave), 0,
                     static_cast<hipStream_t>(stream), cfgs, errs, previous, problem_seed_start, 0,
                     n_act, continuity_weight, out_best, out_idx);
  return pk::check_launch("select_kernel");
}
