// Pose + hinge collision cost and its hand-written adjoint, for gfx950.
//
// cost(q) = sum_c (W_c r_c)^2 + sum_{k,m} max(0, margin - d_km)^2
//   r     = [p_ee - p_t; log(q_ee q_t^-1)]  (the LM residual, _ik_cuda_helpers.cuh:471-494)
//   d_km  = |c_k - o_m| - r_k - R_m         (robot sphere k, world sphere m)
// The collision term is the CHOMP hinge^2 (src/pyroffi/optimization_engines/
// _chomp_optimization.py:140-154); the reference differentiates it with JAX AD.
//
// Adjoint (reverse accumulation over the kinematic tree, no Jacobians):
//   dcost/dc_k   = f_k = sum_m -2 h_km u_km,   u = (c - o)/|c - o|
//   per link L   : F_L = sum_{k in L} f_k,     T_L = sum_k c_k x f_k  (about the origin)
//   EE           : F += 2 W_p^2 r_pos at p_ee, T += p_ee x F + 2 W_o^2 r_ori
//                  (d|log E|^2 under a left perturbation w is 2 r_ori . w because
//                   J_l(r)^-T r = r)
//   subtree sums : (F, T)_sub(j) over the links below joint j (reverse topo order)
//   dcost/dq_a   = sum_{j : src(j) = a} mul_j (w_j . (T_sub - p_j x F_sub) + v_j . F_sub)
//                  with w_j = R_j omega_j, v_j = R_j v_j from the joint's twist
//                  and world pose (= pose of its child link).
//
// Work decomposition: a wave owns a contiguous block of configurations and
// handles one at a time; lanes sweep world obstacles (T per lane, resident in
// registers), robot spheres come from LDS as broadcasts, each link's wrench is
// a wavefront sum, and lane 0 walks the tree.
#include "pk_model.h"

#include <math.h>

#include <algorithm>

namespace pk {

struct HingeArgs {
  const float* link_poses;      // (B, L, 7)
  long long batch;
  int n_links;
  const float* spheres_local;   // (K, 4), grouped by link
  const int* link_sphere_start; // (L + 1)
  int n_spheres;
  const float* world;           // (M, 4)
  int n_world;
  float margin;
  int ee_link;                  // -1: no pose term
  const float* targets;         // (B, 7)
  float pos_weight, ori_weight;
  // tables for the adjoint
  const float* twists;
  const int* parent_idx;
  const int* act_idx;
  const int* mimic_act_idx;
  const float* mimic_mul;
  const int* topo_inv;
  const int* link_parent_joint;
  int n_joints, n_act;
  float* out_cost;              // (B)
  float* out_grad;              // (B, n_act) or null
};

struct JointAdj {
  float ang[3], lin[3];
  float mul;
  int src, parent, child_link;
};

__device__ __forceinline__ float wave_sum(float v) { return wave_sum_dpp(v); }

// Rounding allowance of the expanded reach test: 32 unit roundoffs (2^-24).
constexpr float kHingeTol = 2e-6f;

template <int T>
__global__ __launch_bounds__(64, 4) void pose_hinge_kernel(HingeArgs a) {
  extern __shared__ float4 smem4[];
  const int lane = threadIdx.x;
  const int L = a.n_links, K = a.n_spheres, M = a.n_world, J = a.n_joints;
  float4* sl = smem4;                                        // K local spheres
  float4* sc = sl + K;                                       // K world spheres
  float4* se = sc + K;                                       // K expanded entries
  float* wl = reinterpret_cast<float*>(se + K);              // L * 6 link wrenches
  float* ptile = wl + L * 6;                                 // L * 7 poses
  float* gacc = ptile + L * 7;                               // n_act
  float* wsub = gacc + a.n_act;                              // J * 6 subtree wrenches
  int* lstart = reinterpret_cast<int*>(wsub + J * 6);        // L + 1
  int* slink = lstart + (L + 1);                             // K
  int* topo = slink + K;                                     // J
  JointAdj* jt = reinterpret_cast<JointAdj*>(topo + J);      // J
  float* lbnd = reinterpret_cast<float*>(jt + J);            // L: r_link^2 + r_max^2

  if (lane == 0) {
    int prev = 0;
    for (int n = 0; n <= L; ++n) {
      const int v = (n == L) ? K : a.link_sphere_start[n];
      prev = max(prev, min(v, K));
      lstart[n] = (n == 0) ? 0 : prev;
    }
  }
  for (int k = lane; k < K; k += kWave) {
    const float* s = a.spheres_local + 4 * k;
    sl[k] = make_float4(s[0], s[1], s[2], s[3]);
  }
  for (int j = lane; j < J; j += kWave) {
    JointAdj r;
    for (int c = 0; c < 3; ++c) { r.lin[c] = a.twists[6 * j + c]; r.ang[c] = a.twists[6 * j + 3 + c]; }
    r.mul = a.mimic_mul[j];
    const int m = a.mimic_act_idx[j];
    int src = (m != -1) ? m : a.act_idx[j];
    r.src = (src >= 0 && src < a.n_act) ? src : -1;
    const int p = a.parent_idx[j];
    r.parent = (p >= 0 && p < J) ? p : -1;
    r.child_link = -1;
    for (int l = 0; l < L; ++l)
      if (a.link_parent_joint[l] == j) r.child_link = l;
    jt[j] = r;
    const int tj = a.topo_inv[j];
    topo[j] = (tj >= 0 && tj < J) ? tj : j;
  }
  __syncthreads();
  for (int n = lane; n < L; n += kWave) {
    float r2 = 0.f, rm = 0.f;
    for (int k = lstart[n]; k < lstart[n + 1]; ++k) {
      slink[k] = n;
      const float4 q = sl[k];
      r2 = fmaxf(r2, q.x * q.x + q.y * q.y + q.z * q.z);
      rm = fmaxf(rm, fabsf(q.w));
    }
    lbnd[n] = r2 * 1.001f + rm * rm;  // slack for the rounding of the rotated offsets
  }

  const int chunk = kWave * T;
  const int n_chunks = (M + chunk - 1) / chunk;
  float ox[T], oy[T], oz[T], orm[T];  // obstacle centre, radius + margin
  auto load_obstacles = [&](int c) {
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int m = c * chunk + t * kWave + lane;
      if (m < M) {
        ox[t] = a.world[4 * m]; oy[t] = a.world[4 * m + 1]; oz[t] = a.world[4 * m + 2];
        orm[t] = a.world[4 * m + 3] + a.margin;
      } else {  // contributes nothing: far away, no reach
        ox[t] = 1e6f; oy[t] = 1e6f; oz[t] = 1e6f; orm[t] = 0.f;
      }
    }
  };
  if (n_chunks > 0) load_obstacles(0);
  __syncthreads();

  const long long per = (a.batch + gridDim.x - 1) / gridDim.x;
  const long long b_begin = static_cast<long long>(blockIdx.x) * per;
  const long long b_end = min(a.batch, b_begin + per);
  const bool want_grad = a.out_grad != nullptr;
  for (long long b = b_begin; b < b_end; ++b) {
    const float* pb = a.link_poses + b * L * 7;
    for (int f = lane; f < L * 7; f += kWave) ptile[f] = pb[f];
    for (int f = lane; f < L * 6; f += kWave) wl[f] = 0.f;
    __syncthreads();
    for (int k = lane; k < K; k += kWave) {
      const float* p = ptile + 7 * slink[k];
      const float4 s = sl[k];
      float wx, wy, wz;
      quat_rotate(p[0], p[1], p[2], p[3], s.x, s.y, s.z, wx, wy, wz);
      sc[k] = make_float4(wx + p[4], wy + p[5], wz + p[6], s.w);
      se[k] = make_float4(-2.f * wx, -2.f * wy, -2.f * wz, wx * wx + wy * wy + wz * wz - s.w * s.w);
    }
    __syncthreads();

    float cost = 0.f;
    for (int c = 0; c < n_chunks; ++c) {
      if (c > 0) load_obstacles(c);
      for (int n = 0; n < L; ++n) {
        const int k0 = lstart[n], k1 = lstart[n + 1];
        if (k0 == k1) continue;
        // Reach test in the link frame: with w = R s, q = o - p and
        // R' = R_o + margin, v = |w|^2 - r^2 - 2 r R' - 2 w.q + |q|^2 - R'^2
        // = d^2 - (r + R')^2, three FMAs and one more for the radius term.
        // Its rounding is below tol = kHingeTol (r_link^2 + r_max^2 + |q|^2 +
        // R'^2), so v >= tol means the hinge is zero; the rest take the exact
        // form.
        const float px = ptile[7 * n + 4], py = ptile[7 * n + 5], pz = ptile[7 * n + 6];
        const float tb = kHingeTol * lbnd[n];
        float qx[T], qy[T], qz[T], nct[T];  // nct = tol - (|q|^2 - R'^2)
#pragma unroll
        for (int t = 0; t < T; ++t) {
          qx[t] = ox[t] - px;
          qy[t] = oy[t] - py;
          qz[t] = oz[t] - pz;
          const float qq = qx[t] * qx[t] + qy[t] * qy[t] + qz[t] * qz[t];
          nct[t] = fmaf(kHingeTol, fmaf(orm[t], orm[t], qq), tb) - fmaf(-orm[t], orm[t], qq);
        }
        float F0 = 0.f, F1 = 0.f, F2 = 0.f, T0 = 0.f, T1 = 0.f, T2 = 0.f;
        for (int k = k0; k < k1; ++k) {
          const float4 s = sc[k];
          const float4 e = se[k];
          const float m2r = -2.f * s.w;
          float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
          for (int t = 0; t < T; ++t) {
            const float v = fmaf(e.x, qx[t], fmaf(e.y, qy[t], fmaf(e.z, qz[t], fmaf(m2r, orm[t], e.w))));
            if (v < nct[t]) {
              const float dx = s.x - ox[t], dy = s.y - oy[t], dz = s.z - oz[t];
              const float d2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-20f);
              const float rs = __builtin_amdgcn_rsqf(d2);
              const float h = fmaxf((s.w + orm[t]) - d2 * rs, 0.f);  // margin - d
              cost = fmaf(h, h, cost);
              const float g = -2.f * h * rs;
              fx = fmaf(g, dx, fx);
              fy = fmaf(g, dy, fy);
              fz = fmaf(g, dz, fz);
            }
          }
          if (want_grad) {
            F0 += fx; F1 += fy; F2 += fz;
            T0 += s.y * fz - s.z * fy;
            T1 += s.z * fx - s.x * fz;
            T2 += s.x * fy - s.y * fx;
          }
        }
        if (want_grad) {
          const float w0 = wave_sum(F0), w1 = wave_sum(F1), w2 = wave_sum(F2);
          const float w3 = wave_sum(T0), w4 = wave_sum(T1), w5 = wave_sum(T2);
          if (lane == 0) {
            wl[6 * n] += w0; wl[6 * n + 1] += w1; wl[6 * n + 2] += w2;
            wl[6 * n + 3] += w3; wl[6 * n + 4] += w4; wl[6 * n + 5] += w5;
          }
        }
      }
    }
    if (n_chunks > 1) load_obstacles(0);
    cost = wave_sum(cost);

    if (lane == 0) {
      // Pose term at the end effector.
      if (a.ee_link >= 0 && a.ee_link < L) {
        const float* pe = ptile + 7 * a.ee_link;
        const float* tg = a.targets + b * 7;
        float r[6];
        r[0] = pe[4] - tg[4]; r[1] = pe[5] - tg[5]; r[2] = pe[6] - tg[6];
        float ew, ex, ey, ez;
        quat_mul(pe[0], pe[1], pe[2], pe[3], tg[0], -tg[1], -tg[2], -tg[3], ew, ex, ey, ez);
        if (ew < 0.f) { ew = -ew; ex = -ex; ey = -ey; ez = -ez; }
        const float shn = sqrtf(ex * ex + ey * ey + ez * ez);
        const float kk = (shn > 1e-6f) ? 2.f * atan2f(shn, ew) / shn : 2.f;
        r[3] = ex * kk; r[4] = ey * kk; r[5] = ez * kk;
        const float wp2 = a.pos_weight * a.pos_weight, wo2 = a.ori_weight * a.ori_weight;
        cost += wp2 * (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + wo2 * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
        if (want_grad) {
          const float Fx = 2.f * wp2 * r[0], Fy = 2.f * wp2 * r[1], Fz = 2.f * wp2 * r[2];
          float* w = wl + 6 * a.ee_link;
          w[0] += Fx; w[1] += Fy; w[2] += Fz;
          w[3] += pe[5] * Fz - pe[6] * Fy + 2.f * wo2 * r[3];
          w[4] += pe[6] * Fx - pe[4] * Fz + 2.f * wo2 * r[4];
          w[5] += pe[4] * Fy - pe[5] * Fx + 2.f * wo2 * r[5];
        }
      }
      a.out_cost[b] = cost;
      if (want_grad) {
        // Subtree wrenches in reverse topological order, then the joint gradient.
        for (int j = 0; j < J; ++j) {
          const int l = jt[j].child_link;
          for (int c = 0; c < 6; ++c) wsub[6 * j + c] = (l >= 0) ? wl[6 * l + c] : 0.f;
        }
        for (int i = J - 1; i >= 0; --i) {
          const int j = topo[i];
          const int p = jt[j].parent;
          if (p >= 0)
            for (int c = 0; c < 6; ++c) wsub[6 * p + c] += wsub[6 * j + c];
        }
        for (int q = 0; q < a.n_act; ++q) gacc[q] = 0.f;
        for (int j = 0; j < J; ++j) {
          const JointAdj& r = jt[j];
          if (r.src < 0 || r.child_link < 0) continue;
          const float* pj = ptile + 7 * r.child_link;
          float wx, wy, wz, vx, vy, vz;
          quat_rotate(pj[0], pj[1], pj[2], pj[3], r.ang[0], r.ang[1], r.ang[2], wx, wy, wz);
          quat_rotate(pj[0], pj[1], pj[2], pj[3], r.lin[0], r.lin[1], r.lin[2], vx, vy, vz);
          const float* w = wsub + 6 * j;
          const float tx = w[3] - (pj[5] * w[2] - pj[6] * w[1]);
          const float ty = w[4] - (pj[6] * w[0] - pj[4] * w[2]);
          const float tz = w[5] - (pj[4] * w[1] - pj[5] * w[0]);
          gacc[r.src] += r.mul * (wx * tx + wy * ty + wz * tz + vx * w[0] + vy * w[1] + vz * w[2]);
        }
        for (int q = 0; q < a.n_act; ++q) a.out_grad[b * a.n_act + q] = gacc[q];
      }
    }
    __syncthreads();
  }
}

static int pick_t(int M) {
  if (M <= 64) return 1;
  if (M <= 128) return 2;
  if (M <= 256) return 4;
  return 8;
}

}  // namespace pk

extern "C" int pk_pose_hinge_cost_f32(
    void* stream, const float* link_poses, int64_t batch, int32_t n_links,
    const float* spheres_local, const int32_t* link_sphere_start, int32_t n_spheres,
    const float* world_spheres, int32_t n_world_spheres, float margin, int32_t ee_link,
    const float* targets, float pos_weight, float ori_weight, const float* twists,
    const int32_t* parent_idx, const int32_t* act_idx, const int32_t* mimic_act_idx,
    const float* mimic_mul, const int32_t* topo_inv, const int32_t* link_parent_joint,
    int32_t n_joints, int32_t n_act, float* out_cost, float* out_grad) {
  PK_REQUIRE(batch >= 0 && n_links >= 0 && n_spheres >= 0 && n_world_spheres >= 0, "negative size");
  PK_REQUIRE(n_joints >= 0 && n_joints <= pk::kMaxJoints, "n_joints=%d outside [0, %d]", n_joints,
             pk::kMaxJoints);
  PK_REQUIRE(n_act >= 0 && n_act <= pk::kMaxAct, "n_act=%d outside [0, %d]", n_act, pk::kMaxAct);
  PK_REQUIRE(n_links <= pk::kMaxJoints + 1, "n_links=%d too large", n_links);
  PK_REQUIRE(ee_link < n_links, "ee_link=%d outside the %d links", ee_link, n_links);
  if (batch == 0) return PK_OK;
  PK_REQUIRE(link_poses && link_sphere_start && out_cost && (n_spheres == 0 || spheres_local) &&
                 (n_world_spheres == 0 || world_spheres) && (ee_link < 0 || targets),
             "null pointer");
  if (out_grad)
    PK_REQUIRE(twists && parent_idx && act_idx && mimic_act_idx && mimic_mul && topo_inv &&
                   link_parent_joint,
               "null adjoint table");
  pk::HingeArgs a{link_poses, batch, n_links, spheres_local, link_sphere_start, n_spheres,
                  world_spheres, n_world_spheres, margin, ee_link, targets, pos_weight,
                  ori_weight, twists, parent_idx, act_idx, mimic_act_idx, mimic_mul,
                  topo_inv, link_parent_joint, out_grad ? n_joints : 0, n_act, out_cost, out_grad};
  const int J = a.n_joints;
  const size_t lds = sizeof(float4) * 3 * n_spheres +
                     sizeof(float) * (n_links * 14 + n_act + J * 6) +
                     sizeof(int) * (n_links + 1 + n_spheres + J) + sizeof(pk::JointAdj) * J;
  PK_REQUIRE(lds <= 150 * 1024, "tables need %zu bytes of LDS", lds);
  const long long want = (batch + 3) / 4;
  const unsigned grid = static_cast<unsigned>(std::max(1LL, std::min(want, 8192LL)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (pk::pick_t(n_world_spheres)) {
    case 1: hipLaunchKernelGGL(pk::pose_hinge_kernel<1>, dim3(grid), dim3(pk::kWave), lds, s, a); break;
    case 2: hipLaunchKernelGGL(pk::pose_hinge_kernel<2>, dim3(grid), dim3(pk::kWave), lds, s, a); break;
    case 4: hipLaunchKernelGGL(pk::pose_hinge_kernel<4>, dim3(grid), dim3(pk::kWave), lds, s, a); break;
    default: hipLaunchKernelGGL(pk::pose_hinge_kernel<8>, dim3(grid), dim3(pk::kWave), lds, s, a); break;
  }
  return pk::check_launch("pose_hinge_kernel");
}
