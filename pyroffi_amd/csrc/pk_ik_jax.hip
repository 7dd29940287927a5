// C ABI of the JAX-convention IK (kernels in pk_ik_jax.h) and the HJCD-IK
// refinement-seed builder.
#include "pk_ik_jax.h"

namespace pk {

static int dispatch_jax_ik(void* stream, const JaxIkArgs& a) {
  const int nee = a.ik.n_ee <= 1 ? 1 : (a.ik.n_ee <= 2 ? 2 : 4);
  if (a.ik.n_act <= 8) {
    if (nee == 1) return launch_jax_ik_8_1(stream, a);
    if (nee == 2) return launch_jax_ik_8_2(stream, a);
    return launch_jax_ik_8_4(stream, a);
  }
  if (nee == 1) return launch_jax_ik_16_1(stream, a);
  if (nee == 2) return launch_jax_ik_16_2(stream, a);
  return launch_jax_ik_16_4(stream, a);
}

// Phase-2 seeds of hjcd_solve (_hjcd_ik.py:599-611) for one problem per
// workgroup: the top_k coarse results by error (a stable sort: ties by index,
// NaN last), tiled `repeats` times; every copy after the first top_k gets the
// perturbation noise * 0.15 * (upper - lower) (zero on fixed joints) and is
// clipped to the limits.
__global__ __launch_bounds__(256) void hjcd_refine_seeds_kernel(const float* __restrict__ cfgs,
                                                                const float* __restrict__ errs, int S, int A,
                                                                int top_k, int repeats,
                                                                const float* __restrict__ noise,
                                                                const float* __restrict__ lower,
                                                                const float* __restrict__ upper,
                                                                const int* __restrict__ fixed_mask,
                                                                float* __restrict__ out) {
  extern __shared__ float smem[];
  float* key = smem;                               // (S)
  int* sel = reinterpret_cast<int*>(key + S);      // (top_k)
  const long long p = blockIdx.x;
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const float e = errs[p * S + s];
    key[s] = (e == e) ? e : INFINITY;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const float ks = key[s];
    int rank = 0;
    for (int j = 0; j < S; ++j) {
      const float kj = key[j];
      rank += (kj < ks || (kj == ks && j < s)) ? 1 : 0;
    }
    if (rank < top_k) sel[rank] = s;
  }
  __syncthreads();
  const int R = top_k * repeats;
  for (int f = threadIdx.x; f < R * A; f += blockDim.x) {
    const int r = f / A, i = f - r * A;
    const float base = cfgs[(p * S + sel[r % top_k]) * A + i];
    float v = base;
    if (r >= top_k) {
      const float nz = fixed_mask[i] ? 0.f : noise[(p * R + r) * A + i] * 0.15f * (upper[i] - lower[i]);
      v = fmaxf(lower[i], fminf(upper[i], base + nz));
    }
    out[(p * R + r) * A + i] = v;
  }
}

static int fill_jax_args(JaxIkArgs& ja, const float* seeds, int32_t n_problems, int32_t n_seeds, int32_t n_act,
                         const float* twists, const float* parent_tf, const int32_t* parent_idx,
                         const int32_t* act_idx, const float* mimic_mul, const float* mimic_off,
                         const int32_t* mimic_act_idx, const int32_t* topo_inv, int32_t n_joints,
                         const int32_t* target_jnts, const int32_t* ancestor_masks, int32_t n_ee,
                         const float* target_T, const float* lower, const float* upper,
                         const int32_t* fixed_mask, float* out_cfgs, float* out_err) {
  PK_REQUIRE(n_problems >= 0 && n_seeds >= 0 && n_act >= 0 && n_ee >= 0, "negative size");
  PK_REQUIRE(n_problems <= 65535, "n_problems=%d exceeds 65535", n_problems);
  PK_REQUIRE(n_joints >= 0 && n_joints <= kMaxJoints, "n_joints=%d outside [0, %d]", n_joints, kMaxJoints);
  PK_REQUIRE(n_act <= kMaxAct, "n_act=%d exceeds %d", n_act, kMaxAct);
  PK_REQUIRE(n_ee >= 1 && n_ee <= kMaxEE, "n_ee=%d outside [1, %d]", n_ee, kMaxEE);
  if (n_problems == 0 || n_seeds == 0) return PK_OK;
  PK_REQUIRE(seeds && out_cfgs && out_err && lower && upper && fixed_mask && target_jnts && ancestor_masks &&
                 target_T,
             "null pointer");
  if (n_joints > 0)
    PK_REQUIRE(twists && parent_tf && parent_idx && act_idx && mimic_mul && mimic_off && mimic_act_idx &&
                   topo_inv,
               "null robot table pointer");
  IkArgs& a = ja.ik;
  a = IkArgs{};
  a.seeds = seeds;
  a.n_problems = n_problems;
  a.n_seeds = n_seeds;
  a.n_act = n_act;
  a.tables = RobotTables{twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx,
                         topo_inv, n_joints, n_act};
  a.target_jnts = target_jnts;
  a.ancestor_masks = ancestor_masks;
  a.n_ee = n_ee;
  a.target_T = target_T;
  a.lower = lower;
  a.upper = upper;
  a.fixed_mask = fixed_mask;
  a.enable_collision = 0;
  a.out_cfgs = out_cfgs;
  a.out_err = out_err;
  return PK_OK;
}

}  // namespace pk

extern "C" int pk_ik_jax_lm_f32(void* stream, const float* seeds, int32_t n_problems, int32_t n_seeds,
                                int32_t n_act, const float* twists, const float* parent_tf,
                                const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
                                const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv,
                                int32_t n_joints, const int32_t* target_jnts, const int32_t* ancestor_masks,
                                int32_t n_ee, const float* target_T, const float* lower, const float* upper,
                                const int32_t* fixed_mask, int32_t max_iter, float pos_weight, float ori_weight,
                                float lambda_init, int32_t adaptive_weights, float limit_prior_weight,
                                float kick_scale, uint64_t kick_key, float* out_cfgs, float* out_err) {
  PK_REQUIRE(max_iter >= 0, "max_iter must be >= 0");
  PK_REQUIRE(limit_prior_weight >= 0.f && kick_scale >= 0.f && lambda_init > 0.f,
             "limit_prior_weight, kick_scale must be >= 0 and lambda_init > 0");
  pk::JaxIkArgs ja{};
  const int st = pk::fill_jax_args(ja, seeds, n_problems, n_seeds, n_act, twists, parent_tf, parent_idx, act_idx,
                                   mimic_mul, mimic_off, mimic_act_idx, topo_inv, n_joints, target_jnts,
                                   ancestor_masks, n_ee, target_T, lower, upper, fixed_mask, out_cfgs, out_err);
  if (st != PK_OK || n_problems == 0 || n_seeds == 0) return st;
  ja.ik.max_iter = max_iter;
  ja.ik.pos_weight = pos_weight;
  ja.ik.ori_weight = ori_weight;
  ja.ik.lambda_init = lambda_init;
  ja.opts = pk::JlmOpts{adaptive_weights != 0, limit_prior_weight, kick_scale,
                        static_cast<unsigned long long>(kick_key)};
  ja.coarse = 0;
  return pk::dispatch_jax_ik(stream, ja);
}

extern "C" int pk_hjcd_coarse_f32(void* stream, const float* seeds, int32_t n_problems, int32_t n_seeds,
                                  int32_t n_act, const float* twists, const float* parent_tf,
                                  const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
                                  const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv,
                                  int32_t n_joints, const int32_t* target_jnts, const int32_t* ancestor_masks,
                                  int32_t n_ee, const float* target_T, const int32_t* ee_of_seed,
                                  const float* lower, const float* upper, const int32_t* fixed_mask,
                                  int32_t k_max, float epsilon, float nu, float* out_cfgs, float* out_err) {
  PK_REQUIRE(k_max >= 0, "k_max must be >= 0");
  pk::JaxIkArgs ja{};
  const int st = pk::fill_jax_args(ja, seeds, n_problems, n_seeds, n_act, twists, parent_tf, parent_idx, act_idx,
                                   mimic_mul, mimic_off, mimic_act_idx, topo_inv, n_joints, target_jnts,
                                   ancestor_masks, n_ee, target_T, lower, upper, fixed_mask, out_cfgs, out_err);
  if (st != PK_OK || n_problems == 0 || n_seeds == 0) return st;
  PK_REQUIRE(ee_of_seed != nullptr, "null ee_of_seed");
  ja.coarse = 1;
  ja.ee_of_seed = ee_of_seed;
  ja.k_max = k_max;
  ja.eps = epsilon;
  ja.nu = nu;
  return pk::dispatch_jax_ik(stream, ja);
}

extern "C" int pk_hjcd_refine_seeds_f32(void* stream, const float* coarse_cfgs, const float* coarse_err,
                                        int32_t n_problems, int32_t n_seeds, int32_t n_act, int32_t top_k,
                                        int32_t repeats, const float* noise, const float* lower,
                                        const float* upper, const int32_t* fixed_mask, float* out_seeds) {
  PK_REQUIRE(n_problems >= 0 && n_seeds >= 0 && n_act >= 0 && repeats >= 1, "invalid sizes");
  PK_REQUIRE(top_k >= 1 && top_k <= n_seeds, "top_k=%d outside [1, n_seeds=%d]", top_k, n_seeds);
  PK_REQUIRE(n_seeds <= 8192, "n_seeds=%d exceeds 8192", n_seeds);
  PK_REQUIRE(static_cast<long long>(top_k) * repeats * n_act <= 0x7fffffffLL, "too many refinement seeds");
  if (n_problems == 0 || n_act == 0) return PK_OK;
  PK_REQUIRE(coarse_cfgs && coarse_err && lower && upper && fixed_mask && out_seeds &&
                 (repeats == 1 || noise != nullptr),
             "null pointer");
  const size_t lds = sizeof(float) * n_seeds + sizeof(int) * top_k;
  hipLaunchKernelGGL(pk::hjcd_refine_seeds_kernel, dim3(n_problems), dim3(256), lds,
                     static_cast<hipStream_t>(stream), coarse_cfgs, coarse_err, n_seeds, n_act, top_k, repeats,
                     noise, lower, upper, fixed_mask, out_seeds);
  return pk::check_launch("hjcd_refine_seeds_kernel");
}
