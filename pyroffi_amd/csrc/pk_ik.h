// Multi-seed Levenberg-Marquardt IK for gfx950: one lane = one seed.
// (Templates; pk_ik_inst.hip instantiates them, pk_ik.hip holds the C ABI.)
//
// Algorithm (the reference's ls_ik_lm_kernel, _ls_ik_cuda_kernel.cu:64-394,
// with the residual/Jacobian of _ik_cuda_helpers.cuh:433-631):
//   r_pos = p_ee - p_tgt;  r_ori = log(q_ee * q_tgt^-1) (hemisphere-folded)
//   J     = geometric Jacobian ([z x (p_ee - p_j); z] revolute, [z; 0]
//           prismatic), times the mimic multiplier, ancestor-masked per EE
//   per iteration: stop if every EE has |r_pos| < eps_pos and |r_ori| < eps_ori;
//   fw = W r, Jw = W J, cur = |fw|^2 + penalty(q)
//   Jacobi column scaling s_a = |Jw[:,a]| + 1e-8; (Js^T Js + lam I) p = -Js^T fw
//   (solved in FP32 with the damping floored at 1e-6, see kDampFloor)
//   (fixed joints: zero row/col, unit diagonal); delta = p / s
//   trust region: |delta| <= R(max pos err, max ori err) in {0.38, 0.22, 0.12, 0.05}
//   5-point line search alpha in {1, .5, .25, .1, .025} on clamp(q + alpha delta),
//   all evaluated, best kept; accept if err < cur (1 - 1e-4): lam *= 0.5 (>= 1e-10),
//   else lam *= 3 (<= 1e6); track the all-time best (err, q).
//   penalty(q) = w_coll * sum over robot spheres x world spheres of
//                min(d - margin, 0)^2.
//
// gfx950 layout: the workgroup prologue folds every fixed joint into its
// nearest moving ancestor, producing a tree of "nodes" (the joints whose
// transform depends on q).  A lane holds only the node poses in registers
// (n_act of them for a robot without mimic joints), and the EE frames and
// sphere centres are constant offsets from a node.  All per-seed arrays are
// indexed by compile-time constants (MAXA = padded DOF, NEE = padded EE
// count); padded DOF are frozen joints with zero limits, padded EEs have zero
// residual, so they change nothing.  The damped normal equations are solved
// with an FP32 Cholesky in registers.
#pragma once

#include "pk_geom.h"
#include "pk_grid.h"
#include "pk_model.h"

#include <math.h>

#include <algorithm>

namespace pk {

struct NodeRec {
  float ptf[7];     // folded constant: parent node frame -> this joint (before motion)
  float lin[3], ang[3];
  float mul, off;
  int src;          // actuated index driving this node, -1 none
  int parent;       // parent node, -1 = world
  int kind;         // joint_exp kind (1..4)
  int jac_type;     // 1 revolute, 2 prismatic, 0 no Jacobian column
  float ax[3];      // unit joint axis in the joint frame (Jacobian)
  unsigned ee_mask; // bit e set when the node is an ancestor of EE e
  float qa[4];      // kind 4: ptf.q * (0, unit axis), so the local rotation is c ptf.q + s qa
  int pad[2];
};

struct IkArgs {
  const float* seeds;
  int n_problems, n_seeds, n_act;
  RobotTables tables;
  const int* target_jnts;
  const int* ancestor_masks;
  int n_ee;
  const float* target_T;
  const float* robot_spheres;
  const int* robot_sphere_joint;
  int n_robot_spheres;
  const float* world;
  int n_world;
  // Non-sphere obstacles (packed rows of pk_geom.h), tested against every
  // robot sphere (no grid).
  const float* world_caps;
  int n_caps;
  const float* world_boxes;
  int n_boxes;
  const float* world_hs;
  int n_hs;
  const float* lower;
  const float* upper;
  const int* fixed_mask;
  int max_iter;
  float pos_weight, ori_weight, lambda_init, eps_pos, eps_ori;
  int enable_collision;
  float collision_weight, collision_margin;
  float* out_cfgs;
  float* out_err;
  // Batched launch only.
  const int* problem_seed_start;  // (n_problems + 1): seeds of problem p
  long long total_seeds;
  float* state;                   // (total_seeds, 2 * n_act + 3) or null
  int iter_begin, iter_end;
  GridArgs grid;
  // When non-null, the batched kernel adds the number of sphere-pair tests
  // its collision penalty executed (per seed, summed over its iterations).
  unsigned long long* work;
};

// Shared-memory image of one problem, built by the prologue.
template <int MAXA, int NEE>
struct IkShared {
  NodeRec nodes[MAXA + kMaxJoints];   // worst case: every joint moves
  int n_nodes;
  int ee_node[NEE];
  float ee_rel[NEE][7];               // node frame -> EE frame
  float tgt[NEE][7];
  float lo[MAXA], hi[MAXA];
  int fixed[MAXA];
  int sph_start[MAXA + kMaxJoints + 2];  // spheres grouped by node; group 0 = world-fixed
  int cnt[MAXA + kMaxJoints + 2];
  int n_sph;
  int too_many_nodes;                 // moving joints exceed the padded width MAXA
  float static_pen;                   // penalty of the world-fixed spheres (group 0)
};

__device__ __forceinline__ Pose pose_from(const float* p) { return load_pose(p); }
__device__ __forceinline__ void pose_to(float* p, const Pose& v) {
  p[0] = v.w; p[1] = v.x; p[2] = v.y; p[3] = v.z; p[4] = v.tx; p[5] = v.ty; p[6] = v.tz;
}

// Single-lane tree folding (J <= 64 joints, runs once per workgroup).
template <int MAXA, int NEE>
__device__ void build_problem(const IkArgs& a, const JointRec* recs, const int* slot_of,
                              IkShared<MAXA, NEE>& sh, int* anchor, float* rel,
                              float4* sph, int* sph_node_tmp) {
  const int J = a.tables.n_joints;
  int nn = 0;
  for (int i = 0; i < J; ++i) {
    const JointRec& r = recs[i];
    int base = -1;
    Pose bt = pose_identity();
    if (r.parent_slot >= 0) {
      base = anchor[r.parent_slot];
      bt = pose_from(rel + 7 * r.parent_slot);
    }
    if (r.kind == 0) {
      anchor[i] = base;
      pose_to(rel + 7 * i, compose(bt, load_pose(r.ptf)));
    } else {
      NodeRec& n = sh.nodes[nn];
      pose_to(n.ptf, compose(bt, load_pose(r.ptf)));
      for (int c = 0; c < 3; ++c) { n.lin[c] = r.lin[c]; n.ang[c] = r.ang[c]; }
      n.mul = r.mul; n.off = r.off; n.src = r.src; n.parent = base; n.kind = r.kind;
      // kind 4 stages the unit axis in ang: the folded constant times (0, axis)
      quat_mul(n.ptf[0], n.ptf[1], n.ptf[2], n.ptf[3], 0.f, r.ang[0], r.ang[1], r.ang[2], n.qa[0], n.qa[1],
               n.qa[2], n.qa[3]);
      const float a2 = r.ang[0] * r.ang[0] + r.ang[1] * r.ang[1] + r.ang[2] * r.ang[2];
      const float l2 = r.lin[0] * r.lin[0] + r.lin[1] * r.lin[1] + r.lin[2] * r.lin[2];
      n.jac_type = 0;
      n.ax[0] = n.ax[1] = n.ax[2] = 0.f;
      if (r.src >= 0) {
        if (a2 > 1e-6f) {
          const float inv = 1.f / sqrtf(a2);
          n.jac_type = 1;
          for (int c = 0; c < 3; ++c) n.ax[c] = r.ang[c] * inv;
        } else if (l2 > 1e-6f) {
          const float inv = 1.f / sqrtf(l2);
          n.jac_type = 2;
          for (int c = 0; c < 3; ++c) n.ax[c] = r.lin[c] * inv;
        }
      }
      unsigned m = 0;
      for (int e = 0; e < a.n_ee; ++e)
        if (a.ancestor_masks[e * J + r.joint]) m |= 1u << e;
      n.ee_mask = m;
      anchor[i] = nn;
      pose_to(rel + 7 * i, pose_identity());
      ++nn;
    }
  }
  sh.n_nodes = nn;
  sh.too_many_nodes = nn > MAXA;
  for (int e = 0; e < NEE; ++e) {
    sh.ee_node[e] = -2;  // padded EE: zero residual
    if (e < a.n_ee) {
      int tj = a.target_jnts[e];
      const int s = (tj >= 0 && tj < J) ? slot_of[tj] : -1;
      if (s >= 0) {
        sh.ee_node[e] = anchor[s];
        for (int c = 0; c < 7; ++c) sh.ee_rel[e][c] = rel[7 * s + c];
      } else {
        sh.ee_node[e] = -1;
        Pose id = pose_identity();
        pose_to(sh.ee_rel[e], id);
      }
    }
  }
  // Robot spheres, regrouped by node (group g = node + 1; group 0 = world-fixed).
  const int K = a.enable_collision ? a.n_robot_spheres : 0;
  int* count = sh.cnt;
  for (int g = 0; g <= nn; ++g) count[g] = 0;
  for (int k = 0; k < K; ++k) {
    const int jidx = a.robot_sphere_joint[k];
    int g = -1;
    if (jidx >= 0 && jidx < J && slot_of[jidx] >= 0) g = anchor[slot_of[jidx]] + 1;
    sph_node_tmp[k] = g;
    if (g >= 0) ++count[g];
  }
  int acc = 0;
  for (int g = 0; g <= nn; ++g) { sh.sph_start[g] = acc; acc += count[g]; count[g] = sh.sph_start[g]; }
  sh.sph_start[nn + 1] = acc;
  sh.n_sph = acc;
  for (int k = 0; k < K; ++k) {
    const int g = sph_node_tmp[k];
    if (g < 0) continue;
    const int jidx = a.robot_sphere_joint[k];
    const int s = slot_of[jidx];
    const Pose rp = pose_from(rel + 7 * s);
    const float* sp = a.robot_spheres + 4 * k;
    float cx, cy, cz;
    quat_rotate(rp.w, rp.x, rp.y, rp.z, sp[0], sp[1], sp[2], cx, cy, cz);
    sph[count[g]++] = make_float4(cx + rp.tx, cy + rp.ty, cz + rp.tz, sp[3]);
  }
}

// Forward kinematics of the node tree for one lane.  Node parents are
// wave-uniform, so the parent lookup is a chain of scalar branches.
// kRevoluteFast: a revolute node's local transform as c ptf.q + s qa (eight
// FMAs) instead of composing ptf with the joint exponential (the same
// transform, other rounding); the JAX-convention and HJCD kernels keep the
// composed form.
template <int MAXA, int NEE, bool kRevoluteFast = true>
__device__ __forceinline__ void node_fk(const IkShared<MAXA, NEE>& sh, const float* q,
                                        Pose* P) {
#pragma unroll
  for (int n = 0; n < MAXA; ++n) {
    if (n < sh.n_nodes) {
      const NodeRec& r = sh.nodes[n];
      const int src = __builtin_amdgcn_readfirstlane(r.src);
      float qs = 0.f;
#pragma unroll
      for (int a = 0; a < MAXA; ++a)
        if (src == a) qs = q[a];
      const float qj = qs * r.mul + r.off;
      const int kind = __builtin_amdgcn_readfirstlane(r.kind);
      Pose tl;
      if (kRevoluteFast && kind == 4) {
        // Revolute: ptf o (cos(th/2), sin(th/2) axis) = (c ptf.q + s qa, ptf.t),
        // the identity rotation below |th| = 1e-6 as in joint_exp.
        const float th = r.lin[0] * qj;
        float sn = __sinf(0.5f * th), cs = __cosf(0.5f * th);
        if (fabsf(th) < 1e-6f) { sn = 0.f; cs = 1.f; }
        tl.w = cs * r.ptf[0] + sn * r.qa[0];
        tl.x = cs * r.ptf[1] + sn * r.qa[1];
        tl.y = cs * r.ptf[2] + sn * r.qa[2];
        tl.z = cs * r.ptf[3] + sn * r.qa[3];
        tl.tx = r.ptf[4]; tl.ty = r.ptf[5]; tl.tz = r.ptf[6];
      } else {
        tl = compose(load_pose(r.ptf), joint_exp<true>(kind, qj, r.lin[0], r.lin[1], r.lin[2], r.ang[0],
                                                       r.ang[1], r.ang[2]));
      }
      const int par = __builtin_amdgcn_readfirstlane(r.parent);
      Pose pp = pose_identity();
      bool has_parent = false;
#pragma unroll
      for (int m = 0; m < n; ++m)
        if (par == m) { pp = P[m]; has_parent = true; }
      P[n] = has_parent ? compose(pp, tl) : tl;
    }
  }
}

template <int MAXA, int NEE>
__device__ __forceinline__ Pose ee_pose(const IkShared<MAXA, NEE>& sh, const Pose* P, int e) {
  const int node = __builtin_amdgcn_readfirstlane(sh.ee_node[e]);
  const Pose rel = load_pose(sh.ee_rel[e]);
  Pose base = pose_identity();
  bool found = false;
#pragma unroll
  for (int n = 0; n < MAXA; ++n)
    if (node == n) { base = P[n]; found = true; }
  return found ? compose(base, rel) : rel;
}

// Residual of one EE: [p - p_t; log(q q_t^-1)] (_ik_cuda_helpers.cuh:471-494).
__device__ __forceinline__ void ee_residual(const Pose& T, const float* tgt, float* r) {
  r[0] = T.tx - tgt[4];
  r[1] = T.ty - tgt[5];
  r[2] = T.tz - tgt[6];
  float ew, ex, ey, ez;
  quat_mul(T.w, T.x, T.y, T.z, tgt[0], -tgt[1], -tgt[2], -tgt[3], ew, ex, ey, ez);
  if (ew < 0.f) { ew = -ew; ex = -ex; ey = -ey; ez = -ez; }
  const float sh = sqrtf(ex * ex + ey * ey + ez * ez);
  if (sh > 1e-6f) {
    const float k = 2.f * atan2f(sh, ew) / sh;
    r[3] = ex * k; r[4] = ey * k; r[5] = ez * k;
  } else {
    r[3] = 2.f * ex; r[4] = 2.f * ey; r[5] = 2.f * ez;
  }
}

template <int MAXA, int NEE>
__device__ __forceinline__ void residuals(const IkShared<MAXA, NEE>& sh, const Pose* P,
                                          const float (&tg)[NEE][7], float* r, Pose* Tee) {
#pragma unroll
  for (int e = 0; e < NEE; ++e) {
    if (sh.ee_node[e] == -2) {
#pragma unroll
      for (int c = 0; c < 6; ++c) r[6 * e + c] = 0.f;
      Tee[e] = pose_identity();
    } else {
      Tee[e] = ee_pose(sh, P, e);
      ee_residual(Tee[e], tg[e], r + 6 * e);
    }
  }
}

// sum over world spheres m of min(|c - o_m| - r - R'_m, 0)^2, R' = R + margin.
__device__ __forceinline__ float world_penalty_all(float cx, float cy, float cz, float r,
                                                   const float4* ws, int M) {
  float p0 = 0.f, p1 = 0.f;
  int m = 0;
  for (; m + 1 < M; m += 2) {
    const float4 o0 = ws[m], o1 = ws[m + 1];
    const float ax = cx - o0.x, ay = cy - o0.y, az = cz - o0.z;
    const float bx = cx - o1.x, by = cy - o1.y, bz = cz - o1.z;
    const float e0 = fminf(__builtin_amdgcn_sqrtf(ax * ax + ay * ay + az * az) - o0.w - r, 0.f);
    const float e1 = fminf(__builtin_amdgcn_sqrtf(bx * bx + by * by + bz * bz) - o1.w - r, 0.f);
    p0 = fmaf(e0, e0, p0);
    p1 = fmaf(e1, e1, p1);
  }
  if (m < M) {
    const float4 o0 = ws[m];
    const float ax = cx - o0.x, ay = cy - o0.y, az = cz - o0.z;
    const float e0 = fminf(__builtin_amdgcn_sqrtf(ax * ax + ay * ay + az * az) - o0.w - r, 0.f);
    p0 = fmaf(e0, e0, p0);
  }
  return p0 + p1;
}

// Same sum over the candidates of the grid cell holding c: exact, since every
// sphere that can be within the margin of a robot sphere of radius <= r_max
// (= grid pad - margin) is listed (pk_grid.hip); a radius no level serves
// falls back to the whole world, and a centre outside the grid has no such
// sphere.  The radius r is the same in every lane (one
// robot sphere at a time), so the level choice is a uniform branch.
// U candidates per pass: their index loads are in flight together (latency a
// single wave per SIMD does not hide); lanes past their count read ws[0]
// (M >= 1 here) and add nothing.
template <int U>
__device__ __forceinline__ float world_penalty_grid(float cx, float cy, float cz, float r,
                                                    const float4* ws, int M, const GridArgs& g,
                                                    unsigned& work) {
  const int level = grid_level(g, r);
  if (level < 0) {
    work += M;
    return world_penalty_all(cx, cy, cz, r, ws, M);
  }
  const int* items;
  int n;
  grid_cell_list(g, level, cx, cy, cz, items, n);
  work += n;
  float pen = 0.f;
  for (int i = 0; i < n; i += U) {
    int m[U];
    bool v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      m[u] = (i + u < n) ? items[i + u] : -1;
      v[u] = m[u] >= 0 && m[u] < M;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float4 o = ws[v[u] ? m[u] : 0];
      const float ax = cx - o.x, ay = cy - o.y, az = cz - o.z;
      const float e = fminf(__builtin_amdgcn_sqrtf(ax * ax + ay * ay + az * az) - o.w - r, 0.f);
      if (v[u]) pen = fmaf(e, e, pen);
    }
  }
  return pen;
}

// Capsules, boxes and half-spaces of the world, staged in LDS: every robot
// sphere tests all of them, sum of min(d - margin, 0)^2.
struct OtherWorld {
  const float* caps;   // (nc, 7)
  const float* boxes;  // (nb, 15)
  const float* hs;     // (nh, 6)
  int nc, nb, nh;
  float margin;
};

__device__ __forceinline__ float other_penalty(float cx, float cy, float cz, float r, const OtherWorld& ow) {
  float pen = 0.f;
  for (int m = 0; m < ow.nc; ++m) {
    const float e = fminf(dist_sphere_capsule(cx, cy, cz, r, ow.caps + 7 * m) - ow.margin, 0.f);
    pen = fmaf(e, e, pen);
  }
  for (int m = 0; m < ow.nb; ++m) {
    const float e = fminf(dist_sphere_box(cx, cy, cz, r, ow.boxes + 15 * m) - ow.margin, 0.f);
    pen = fmaf(e, e, pen);
  }
  for (int m = 0; m < ow.nh; ++m) {
    const float e = fminf(dist_sphere_halfspace(cx, cy, cz, r, ow.hs + 6 * m) - ow.margin, 0.f);
    pen = fmaf(e, e, pen);
  }
  return pen;
}

// Penalty terms of robot spheres [k0, k1) placed by the node pose T, added
// to pen sphere by sphere (world spheres, then the other shapes).
template <int MAXA>
__device__ __forceinline__ void group_penalty(float& pen, const Pose& T, int k0, int k1, const float4* sph,
                                              const float4* ws, int M, const GridArgs& grid,
                                              const OtherWorld& ow, int n_other, unsigned& work) {
  const Rot3 R = quat_to_rot(T.w, T.x, T.y, T.z);
  for (int k = k0; k < k1; ++k) {
    const float4 s = sph[k];
    const float cx = R.m00 * s.x + R.m01 * s.y + R.m02 * s.z + T.tx;
    const float cy = R.m10 * s.x + R.m11 * s.y + R.m12 * s.z + T.ty;
    const float cz = R.m20 * s.x + R.m21 * s.y + R.m22 * s.z + T.tz;
    // Four candidates per pass for up to 8 moving joints; two for the
    // wider instantiations, whose registers are already at the limit
    // (measured on the 14-DOF robot: two are 18% faster than one, four
    // 1% slower than two).
    pen += (grid.n > 0) ? world_penalty_grid<(MAXA <= 8 ? 4 : 2)>(cx, cy, cz, s.w, ws, M, grid, work)
                        : world_penalty_all(cx, cy, cz, s.w, ws, M);
    if (grid.n == 0) work += M;
    if (n_other > 0) {
      pen += other_penalty(cx, cy, cz, s.w, ow);
      work += n_other;
    }
  }
}

// Penalty of the world-fixed spheres (group 0, placed by the identity): the
// same for every seed and iteration of a problem, so the prologue evaluates
// it once per workgroup.  It is the running sum sphere_penalty used to reach
// after group 0, term for term, so starting from it changes no result.
template <int MAXA, int NEE>
__device__ __forceinline__ float static_sphere_penalty(const IkShared<MAXA, NEE>& sh, const float4* sph,
                                                       const float4* ws, int M, const GridArgs& grid,
                                                       const OtherWorld& ow) {
  float pen = 0.f;
  unsigned work = 0;
  group_penalty<MAXA>(pen, pose_identity(), sh.sph_start[0], sh.sph_start[1], sph, ws, M, grid, ow,
                      ow.nc + ow.nb + ow.nh, work);
  return pen;
}

// Collision merit penalty (unweighted) over every robot sphere: the
// world-fixed group's sum from the prologue, then the moving groups; `work`
// counts the sphere-pair tests it executes (grid candidates, or the whole
// world for a sphere no grid level serves), the world-fixed group's once per
// workgroup not included.
template <int MAXA, int NEE>
__device__ __forceinline__ float sphere_penalty(const IkShared<MAXA, NEE>& sh, const Pose* P,
                                                const float4* sph, const float4* ws, int M,
                                                const GridArgs& grid, const OtherWorld& ow, unsigned& work) {
  const int n_other = ow.nc + ow.nb + ow.nh;
  float pen = sh.static_pen;
#pragma unroll
  for (int g = 1; g <= MAXA; ++g) {
    if (g <= sh.n_nodes) {
      const int k0 = sh.sph_start[g], k1 = sh.sph_start[g + 1];
      if (k1 > k0) group_penalty<MAXA>(pen, P[g - 1], k0, k1, sph, ws, M, grid, ow, n_other, work);
    }
  }
  return pen;
}

struct IkLds {
  size_t recs, slot_of, anchor, rel, sph_tmp, sph, ws, other, total;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~static_cast<size_t>(15); }

// Byte offsets of the IK kernels' LDS regions (float4 regions 16-byte aligned).
// n_other: floats of the staged capsules, boxes and half-spaces (7, 15, 6 each).
__host__ __device__ inline IkLds ik_lds_layout(size_t shared_bytes, int J, int K, int M, int n_other = 0) {
  IkLds o;
  o.recs = align16(shared_bytes);
  o.slot_of = align16(o.recs + sizeof(JointRec) * J);
  o.anchor = o.slot_of + sizeof(int) * J;
  o.rel = o.anchor + sizeof(int) * J;
  o.sph_tmp = o.rel + sizeof(float) * 7 * J;
  o.sph = align16(o.sph_tmp + sizeof(int) * K);
  o.ws = o.sph + sizeof(float4) * K;
  o.other = o.ws + sizeof(float4) * M;
  o.total = o.other + sizeof(float) * n_other;
  return o;
}

__host__ __device__ inline int ik_other_floats(const IkArgs& a) {
  return a.enable_collision ? 7 * a.n_caps + 15 * a.n_boxes + 6 * a.n_hs : 0;
}

constexpr float kDampFloor = 1e-6f;

// Lanes below which the single-launch kernel gives every seed 8 lanes:
// 256 CUs x 4 SIMDs x 2 waves x 64 lanes.
constexpr long long kSmallBatchLanes = 256LL * 4 * 2 * 64;

__device__ __forceinline__ float ls_alpha(int i) {
  return i == 0 ? 1.f : (i == 1 ? 0.5f : (i == 2 ? 0.25f : (i == 3 ? 0.1f : 0.025f)));
}

// One step of the line-search argmin over an 8-lane group: take the partner
// lane's (error, step index) through a DPP move and keep the smaller error,
// the lower index on ties.
template <int CTRL>
__device__ __forceinline__ void ls_argmin_step(float& e, int& i) {
  const float pe = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(e), CTRL, 0xf, 0xf, false));
  const int pi = __builtin_amdgcn_mov_dpp(i, CTRL, 0xf, 0xf, false);
  if (pe < e || (pe == e && pi < i)) { e = pe; i = pi; }
}

// A pose from lane src of the wave (ds_bpermute per component).
__device__ __forceinline__ Pose shfl_pose(const Pose& p, int src) {
  Pose o;
  o.w = __shfl(p.w, src, kWave); o.x = __shfl(p.x, src, kWave);
  o.y = __shfl(p.y, src, kWave); o.z = __shfl(p.z, src, kWave);
  o.tx = __shfl(p.tx, src, kWave); o.ty = __shfl(p.ty, src, kWave); o.tz = __shfl(p.tz, src, kWave);
  return o;
}

// Per-seed solver state (registers).
template <int MAXA>
struct LmState {
  float q[MAXA], best_q[MAXA];
  float best_err, lam;
  int done;
  unsigned work;  // sphere-pair tests of the collision penalty (see IkArgs::work)
};

// Problem tables staged in LDS by the workgroup prologue.
template <int MAXA, int NEE>
struct IkView {
  const IkShared<MAXA, NEE>* sh;
  const float4* sph;
  const float4* ws;
  int M;
  bool coll;
  float wc;
  OtherWorld ow;
};

template <int MAXA, int NEE>
__device__ __forceinline__ float weighted_error(const IkView<MAXA, NEE>& v, const float (&tg)[NEE][7],
                                                const float* W, const float* q, Pose* P, float* r,
                                                Pose* Tee, const GridArgs& grid, unsigned& work) {
  node_fk(*v.sh, q, P);
  residuals(*v.sh, P, tg, r, Tee);
  float e = 0.f;
#pragma unroll
  for (int i = 0; i < 6 * NEE; ++i) { const float x = r[i] * W[i % 6]; e += x * x; }
  if (v.coll) e += v.wc * sphere_penalty(*v.sh, P, v.sph, v.ws, v.M, grid, v.ow, work);
  return e;
}

// Weighted Jacobian contribution of one node to its column for one EE:
// [z x (p_ee - p_node); z] (revolute) or [z; 0] (prismatic), times the mimic
// multiplier and the residual weights.  Explicit fmaf fixes the rounding of
// each cross-product term, so every kernel that calls this gets the same bits
// whatever registers or LDS its operands come from.
__device__ __forceinline__ void jac_column(const Pose& Pn, const float* ax, int jt, float tex, float tey,
                                           float tez, float mul, const float* W, float* out6) {
  float zx, zy, zz;
  quat_rotate(Pn.w, Pn.x, Pn.y, Pn.z, ax[0], ax[1], ax[2], zx, zy, zz);
  float g[6];
  if (jt == 1) {
    const float dx = tex - Pn.tx, dy = tey - Pn.ty, dz = tez - Pn.tz;
    g[0] = fmaf(zy, dz, -(zz * dy));
    g[1] = fmaf(zz, dx, -(zx * dz));
    g[2] = fmaf(zx, dy, -(zy * dx));
    g[3] = zx; g[4] = zy; g[5] = zz;
  } else {
    g[0] = zx; g[1] = zy; g[2] = zz; g[3] = 0.f; g[4] = 0.f; g[5] = 0.f;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) out6[k] = (mul * g[k]) * W[k];
}

// LM iterations [it0, it1) of one seed (the algorithm in the file header).
// G = 1: one lane per seed.  G = 8: the 8 lanes of an aligned group hold the
// same seed, compute the same step, and lane k evaluates line-search step k
// (k < 5); an argmin over the group gives every lane the sequential pick.
template <int MAXA, int NEE, int G = 1>
__device__ __forceinline__ void lm_iterate(const IkView<MAXA, NEE>& v, const float (&tg)[NEE][7],
                                           const float* W, const IkArgs& a, LmState<MAXA>& st,
                                           int it0, int it1, const GridArgs& grid) {
  const IkShared<MAXA, NEE>& sh = *v.sh;
  Pose P[MAXA];
  Pose Tee[NEE];
  float r[6 * NEE];
  // The error of the iterate carried into the next iteration (residuals and
  // collision penalty) is the accepted line-search point's, evaluated by the
  // same arithmetic on the same q, or the current one after a rejected step,
  // so the next iteration does not evaluate the penalty again.  With G = 8
  // (8-wide instantiations; the 16-wide ones have no registers to spare) the
  // node poses, EE poses and residuals are taken from the lane that evaluated
  // the accepted point as well, instead of being recomputed.
  constexpr bool kReuse = (G == 8) && (MAXA == 8);
  bool have = false;
  float cur_next = 0.f;
  for (int it = it0; it < it1 && !st.done; ++it) {
    if (!kReuse || !have) {
      node_fk(sh, st.q, P);
      residuals(sh, P, tg, r, Tee);
    }
    bool conv = true;
    float max_p = 0.f, max_o = 0.f;
#pragma unroll
    for (int e = 0; e < NEE; ++e) {
      const float np = sqrtf(r[6 * e] * r[6 * e] + r[6 * e + 1] * r[6 * e + 1] + r[6 * e + 2] * r[6 * e + 2]);
      const float no = sqrtf(r[6 * e + 3] * r[6 * e + 3] + r[6 * e + 4] * r[6 * e + 4] + r[6 * e + 5] * r[6 * e + 5]);
      if (e < a.n_ee && (np >= a.eps_pos || no >= a.eps_ori)) conv = false;
      max_p = fmaxf(max_p, np);
      max_o = fmaxf(max_o, no);
    }
    if (conv) { st.done = 1; break; }

    float cur = 0.f;
    float fw[6 * NEE];
#pragma unroll
    for (int i = 0; i < 6 * NEE; ++i) { fw[i] = r[i] * W[i % 6]; cur += fw[i] * fw[i]; }
    if (have) {
      cur = cur_next;  // the same sum (and penalty) the line search evaluated
    } else if (v.coll) {
      // With G lanes per seed every lane computes the same value; the work
      // count takes one lane's.
      unsigned wk = 0;
      cur += v.wc * sphere_penalty(sh, P, v.sph, v.ws, v.M, grid, v.ow, wk);
      if (G == 1 || threadIdx.x % G == 0) st.work += wk;
    }

    // Weighted geometric Jacobian, column a from the node(s) driven by a.
    float Jm[6 * NEE][MAXA];
#pragma unroll
    for (int i = 0; i < 6 * NEE; ++i)
#pragma unroll
      for (int c = 0; c < MAXA; ++c) Jm[i][c] = 0.f;
#pragma unroll
    for (int n = 0; n < MAXA; ++n) {
      if (n < sh.n_nodes) {
        const NodeRec& nr = sh.nodes[n];
        const int jt = __builtin_amdgcn_readfirstlane(nr.jac_type);
        if (jt != 0) {
          const int src = __builtin_amdgcn_readfirstlane(nr.src);
          const unsigned em = __builtin_amdgcn_readfirstlane(nr.ee_mask);
#pragma unroll
          for (int e = 0; e < NEE; ++e) {
            if (em & (1u << e)) {
              float g6[6];
              jac_column(P[n], nr.ax, jt, Tee[e].tx, Tee[e].ty, Tee[e].tz, nr.mul, W, g6);
#pragma unroll
              for (int c = 0; c < MAXA; ++c)
                if (src == c) {
#pragma unroll
                  for (int k = 0; k < 6; ++k) Jm[6 * e + k][c] += g6[k];
                }
            }
          }
        }
      }
    }

    // Column scaling and normal equations (packed lower triangle).  One
    // division per column (an IEEE FP32 division is ~10 instructions); the
    // 6 NEE entries of the column are multiplied by its reciprocal.
    float ics[MAXA];
#pragma unroll
    for (int c = 0; c < MAXA; ++c) {
      float sq = 0.f;
#pragma unroll
      for (int i = 0; i < 6 * NEE; ++i) sq += Jm[i][c] * Jm[i][c];
      ics[c] = 1.f / (sqrtf(sq) + 1e-8f);
    }
#pragma unroll
    for (int i = 0; i < 6 * NEE; ++i)
#pragma unroll
      for (int c = 0; c < MAXA; ++c) Jm[i][c] *= ics[c];
    constexpr int NT = MAXA * (MAXA + 1) / 2;
    float A[NT];
    float rhs[MAXA];
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 6 * NEE; ++k) acc += Jm[k][i] * Jm[k][j];
        A[i * (i + 1) / 2 + j] = acc;
      }
      float rb = 0.f;
#pragma unroll
      for (int k = 0; k < 6 * NEE; ++k) rb += Jm[k][i] * fw[k];
      rhs[i] = -rb;
      // Columns are unit-scaled, so the diagonal of Js^T Js is ~1 and an FP32
      // Cholesky resolves eigenvalues down to ~1e-6.  On redundant arms
      // Js^T Js is singular and lam decays to 1e-10, which FP64 (the reference)
      // survives and FP32 does not; the damping used inside the solve is
      // floored at 1e-6 while the lam schedule itself is left unchanged.
      A[i * (i + 1) / 2 + i] += fmaxf(st.lam, kDampFloor);
    }
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
      if (sh.fixed[i]) {
#pragma unroll
        for (int j = 0; j < MAXA; ++j) {
          if (j <= i) A[i * (i + 1) / 2 + j] = 0.f;
          else A[j * (j + 1) / 2 + i] = 0.f;
        }
        A[i * (i + 1) / 2 + i] = 1.f;
        rhs[i] = 0.f;
      }
    }
    // Cholesky A = L L^T in place, then two triangular solves that multiply
    // by the reciprocals of the diagonal.
    bool ok = true;
    float idg[MAXA];
#pragma unroll
    for (int k = 0; k < MAXA; ++k) {
      float sd = A[k * (k + 1) / 2 + k];
#pragma unroll
      for (int pp = 0; pp < k; ++pp) sd -= A[k * (k + 1) / 2 + pp] * A[k * (k + 1) / 2 + pp];
      if (sd <= 0.f) ok = false;
      const float lkk = sqrtf(fmaxf(sd, 1e-30f));
      A[k * (k + 1) / 2 + k] = lkk;
      const float inv = 1.f / lkk;
      idg[k] = inv;
#pragma unroll
      for (int i = k + 1; i < MAXA; ++i) {
        float t = A[i * (i + 1) / 2 + k];
#pragma unroll
        for (int pp = 0; pp < k; ++pp) t -= A[i * (i + 1) / 2 + pp] * A[k * (k + 1) / 2 + pp];
        A[i * (i + 1) / 2 + k] = t * inv;
      }
    }
    float y[MAXA];
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
      float t = rhs[i];
#pragma unroll
      for (int pp = 0; pp < i; ++pp) t -= A[i * (i + 1) / 2 + pp] * y[pp];
      y[i] = t * idg[i];
    }
    float delta[MAXA];
#pragma unroll
    for (int i = MAXA - 1; i >= 0; --i) {
      float t = y[i];
#pragma unroll
      for (int pp = i + 1; pp < MAXA; ++pp) t -= A[pp * (pp + 1) / 2 + i] * delta[pp];
      delta[i] = t * idg[i];
    }
    float dn = 0.f;
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
      delta[i] = ok ? delta[i] * ics[i] : 0.f;
      dn += delta[i] * delta[i];
    }
    // Trust region.
    float R;
    if (max_p > 1e-2f || max_o > 0.6f) R = 0.38f;
    else if (max_p > 1e-3f || max_o > 0.25f) R = 0.22f;
    else if (max_p > 2e-4f || max_o > 0.08f) R = 0.12f;
    else R = 0.05f;
    dn = sqrtf(dn);
    if (dn > R) {
      const float sc = R / (dn + 1e-18f);
#pragma unroll
      for (int i = 0; i < MAXA; ++i) delta[i] *= sc;
    }

    // Line search: evaluate all five steps, keep the best (the first on ties).
    float best_a_err = 1e30f;
    int best_a = 0;
    if (G == 1) {
#pragma unroll 1
      for (int ai = 0; ai < 5; ++ai) {
        float qt[MAXA];
#pragma unroll
        for (int i = 0; i < MAXA; ++i) qt[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], st.q[i] + ls_alpha(ai) * delta[i]));
        const float et = weighted_error(v, tg, W, qt, P, r, Tee, grid, st.work);
        if (et < best_a_err) { best_a_err = et; best_a = ai; }
      }
    } else {
      static_assert(G == 1 || G == 8, "line-search groups are 1 or 8 lanes");
      const int ai = min(static_cast<int>(threadIdx.x % G), 4);
      float qt[MAXA];
#pragma unroll
      for (int i = 0; i < MAXA; ++i) qt[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], st.q[i] + ls_alpha(ai) * delta[i]));
      unsigned wk = 0;
      Pose P2[MAXA], Tee2[NEE];
      float r2[6 * NEE];
      const float et = kReuse ? weighted_error(v, tg, W, qt, P2, r2, Tee2, grid, wk)
                              : weighted_error(v, tg, W, qt, P, r, Tee, grid, wk);
      if (threadIdx.x % G < 5) st.work += wk;  // lanes 5..7 repeat step 4
      // Sequential semantics: only errors below 1e30 replace the start value
      // (NaN never does); the first index wins ties.
      float e = (et < 1e30f) ? et : INFINITY;
      int idx = ai;
      ls_argmin_step<0xb1>(e, idx);   // quad_perm [1,0,3,2]
      ls_argmin_step<0x4e>(e, idx);   // quad_perm [2,3,0,1]
      ls_argmin_step<0x141>(e, idx);  // row_half_mirror
      if (e != INFINITY) { best_a_err = e; best_a = idx; }
      if (kReuse && best_a_err < cur * (1.f - 1e-4f)) {
        // Accepted: the next iterate's poses, EE poses and residuals from lane best_a.
        const int src = (static_cast<int>(threadIdx.x) & ~(G - 1)) + best_a;
#pragma unroll
        for (int n = 0; n < MAXA; ++n) P[n] = shfl_pose(P2[n], src);
#pragma unroll
        for (int e2 = 0; e2 < NEE; ++e2) Tee[e2] = shfl_pose(Tee2[e2], src);
#pragma unroll
        for (int i = 0; i < 6 * NEE; ++i) r[i] = __shfl(r2[i], src, kWave);
      }
    }
    cur_next = (best_a_err < cur * (1.f - 1e-4f)) ? best_a_err : cur;
    have = true;
    const float al = ls_alpha(best_a);
    float qt[MAXA];
#pragma unroll
    for (int i = 0; i < MAXA; ++i) qt[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], st.q[i] + al * delta[i]));
    if (best_a_err < cur * (1.f - 1e-4f)) {
#pragma unroll
      for (int i = 0; i < MAXA; ++i) st.q[i] = qt[i];
      st.lam = fmaxf(st.lam * 0.5f, 1e-10f);
    } else {
      st.lam = fminf(st.lam * 3.f, 1e6f);
    }
    if (best_a_err < st.best_err) {
      st.best_err = best_a_err;
#pragma unroll
      for (int i = 0; i < MAXA; ++i) st.best_q[i] = qt[i];
    }
  }
}

// Workgroup prologue shared by both kernels: stage the robot, limits and world.
template <int MAXA, int NEE>
__device__ __forceinline__ IkView<MAXA, NEE> ik_prologue(const IkArgs& a, char* base) {
  using Sh = IkShared<MAXA, NEE>;
  const int J = a.tables.n_joints;
  const int Kin = a.enable_collision ? a.n_robot_spheres : 0;
  const IkLds off = ik_lds_layout(sizeof(Sh), J, Kin, a.enable_collision ? a.n_world : 0, ik_other_floats(a));
  Sh& sh = *reinterpret_cast<Sh*>(base);
  JointRec* recs = reinterpret_cast<JointRec*>(base + off.recs);
  int* slot_of = reinterpret_cast<int*>(base + off.slot_of);
  int* anchor = reinterpret_cast<int*>(base + off.anchor);
  float* rel = reinterpret_cast<float*>(base + off.rel);
  int* sph_tmp = reinterpret_cast<int*>(base + off.sph_tmp);
  float4* sph = reinterpret_cast<float4*>(base + off.sph);
  float4* ws = reinterpret_cast<float4*>(base + off.ws);
  const int lane = threadIdx.x;
  stage_joint_records(a.tables, recs, slot_of);
  if (lane == 0) build_problem<MAXA, NEE>(a, recs, slot_of, sh, anchor, rel, sph, sph_tmp);
  for (int i = lane; i < MAXA; i += kWave) {
    const bool real = i < a.n_act;
    sh.lo[i] = real ? a.lower[i] : 0.f;
    sh.hi[i] = real ? a.upper[i] : 0.f;
    sh.fixed[i] = real ? (a.fixed_mask[i] != 0) : 1;
  }
  const int M = (a.enable_collision && Kin > 0) ? a.n_world : 0;
  for (int m = lane; m < M; m += kWave) {
    const float* o = a.world + 4 * m;
    ws[m] = make_float4(o[0], o[1], o[2], o[3] + a.collision_margin);
  }
  float* oth = reinterpret_cast<float*>(base + off.other);
  const int n_oth = (Kin > 0) ? ik_other_floats(a) : 0;
  const int nc7 = 7 * a.n_caps, nb15 = 15 * a.n_boxes;
  for (int i = lane; i < n_oth; i += kWave)
    oth[i] = (i < nc7) ? a.world_caps[i] : (i < nc7 + nb15 ? a.world_boxes[i - nc7] : a.world_hs[i - nc7 - nb15]);
  __syncthreads();
  IkView<MAXA, NEE> v;
  v.sh = &sh;
  v.sph = sph;
  v.ws = ws;
  v.M = M;
  v.ow.caps = oth;
  v.ow.boxes = oth + nc7;
  v.ow.hs = oth + nc7 + nb15;
  v.ow.nc = n_oth > 0 ? a.n_caps : 0;
  v.ow.nb = n_oth > 0 ? a.n_boxes : 0;
  v.ow.nh = n_oth > 0 ? a.n_hs : 0;
  v.ow.margin = a.collision_margin;
  v.coll = (M > 0 || n_oth > 0) && sh.n_sph > 0;
  v.wc = a.collision_weight;
  if (lane == 0) sh.static_pen = v.coll ? static_sphere_penalty(sh, sph, ws, M, a.grid, v.ow) : 0.f;
  __syncthreads();
  return v;
}

template <int MAXA, int NEE>
__device__ __forceinline__ void load_targets(const IkArgs& a, long long problem, float (&tg)[NEE][7]) {
#pragma unroll
  for (int e = 0; e < NEE; ++e)
#pragma unroll
    for (int c = 0; c < 7; ++c)
      tg[e][c] = (e < a.n_ee) ? a.target_T[(problem * a.n_ee + e) * 7 + c] : (c == 0 ? 1.f : 0.f);
}

// G lanes per seed (see lm_iterate); the group's first lane writes the result.
template <int MAXA, int NEE, int G>
__global__ __launch_bounds__(64) void ls_ik_kernel(IkArgs a) {
  extern __shared__ float4 smem4[];
  const IkView<MAXA, NEE> v = ik_prologue<MAXA, NEE>(a, reinterpret_cast<char*>(smem4));
  const int lane = threadIdx.x;
  const int p = blockIdx.y;
  const int s = blockIdx.x * (kWave / G) + lane / G;
  const bool writer = (lane % G) == 0;
  if (s >= a.n_seeds) return;
  const long long gs = static_cast<long long>(p) * a.n_seeds + s;
  if (v.sh->too_many_nodes) {  // reported to the caller as NaN errors, seeds unchanged
    if (writer) {
      for (int i = 0; i < a.n_act; ++i) a.out_cfgs[gs * a.n_act + i] = a.seeds[gs * a.n_act + i];
      a.out_err[gs] = NAN;
    }
    return;
  }
  float tg[NEE][7];
  load_targets<MAXA, NEE>(a, p, tg);
  const float W[6] = {a.pos_weight, a.pos_weight, a.pos_weight, a.ori_weight, a.ori_weight, a.ori_weight};
  LmState<MAXA> st;
#pragma unroll
  for (int i = 0; i < MAXA; ++i) {
    st.q[i] = (i < a.n_act) ? a.seeds[gs * a.n_act + i] : 0.f;
    st.best_q[i] = st.q[i];
  }
  Pose P[MAXA];
  Pose Tee[NEE];
  float r[6 * NEE];
  st.work = 0;
  st.best_err = weighted_error(v, tg, W, st.q, P, r, Tee, a.grid, st.work);
  st.lam = a.lambda_init;
  st.done = 0;
  lm_iterate<MAXA, NEE, G>(v, tg, W, a, st, 0, a.max_iter, a.grid);
  if (writer) {
#pragma unroll
    for (int i = 0; i < MAXA; ++i)
      if (i < a.n_act) a.out_cfgs[gs * a.n_act + i] = st.best_q[i];
    a.out_err[gs] = st.best_err;
  }
}

// One seed per wave for small launches of the 8-wide instantiations: the
// line search's steps run on lanes 0-7 as in the 8-lane kernel, and the step
// in between is spread over the wave through LDS -- one lane per node for the
// Jacobian contributions, one lane per column for the column sums and
// scaling, one lane per entry of the packed normal matrix and of the
// right-hand side -- instead of every lane repeating it.  The Cholesky and
// the two triangular solves (a dependency chain) run in every lane on the
// matrix read back from LDS.  Each quantity is computed by the same
// expressions, in the same order, as in lm_iterate; the compiler's choices of
// which products to fuse into FMAs differ between the two schedules, so the
// iterates agree with the one-lane kernel's to rounding (about 1e-6 after one
// iteration) rather than bit for bit, and diverge over many iterations like
// any two FP32 orderings of LM.
template <int MAXA, int NEE>
struct WaveLds {
  float P[MAXA][7];         // node poses of the current iterate
  float Tee[NEE][7];        // its EE poses
  float r[6 * NEE];         // its residuals
  float Jn[MAXA][6 * NEE];  // per node: its weighted Jacobian contribution
  float Js[MAXA][6 * NEE];  // per column: the scaled column
  float ics[MAXA];          // per column: 1 / (|column| + 1e-8)
  float A[MAXA * (MAXA + 1) / 2];
  float rhs[MAXA];
  int cnt[MAXA];            // nodes driving column c ...
  int drv[MAXA][MAXA];      // ... in node order
};

// Seeds (problems x seeds) up to which the 8-wide instantiations run one seed
// per wave: one wave per SIMD (four per CU).  With two waves sharing a SIMD
// each runs at about 0.7 of its speed alone, and the 8-lane kernel (one wave
// per 8 seeds) is as fast.
inline long long wave_seed_limit() {
  static const long long limit = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || cus <= 0)
      cus = 256;
    return 4LL * cus;
  }();
  return limit;
}

template <int MAXA, int NEE>
__device__ __forceinline__ void store_iterate(WaveLds<MAXA, NEE>& wl, const Pose* P, const Pose* Tee, const float* r) {
#pragma unroll
  for (int n = 0; n < MAXA; ++n) pose_to(wl.P[n], P[n]);
#pragma unroll
  for (int e = 0; e < NEE; ++e) pose_to(wl.Tee[e], Tee[e]);
#pragma unroll
  for (int i = 0; i < 6 * NEE; ++i) wl.r[i] = r[i];
}

template <int MAXA, int NEE>
__device__ __forceinline__ void lm_iterate_wave(const IkView<MAXA, NEE>& v, const float (&tg)[NEE][7],
                                                const float* W, const IkArgs& a, LmState<MAXA>& st,
                                                WaveLds<MAXA, NEE>& wl, float cur_next) {
  const IkShared<MAXA, NEE>& sh = *v.sh;
  const GridArgs& grid = a.grid;
  const int lane = threadIdx.x;
  constexpr int NT = MAXA * (MAXA + 1) / 2;
  for (int it = 0; it < a.max_iter && !st.done; ++it) {
    wave_sync();
    float r[6 * NEE];
#pragma unroll
    for (int i = 0; i < 6 * NEE; ++i) r[i] = wl.r[i];
    bool conv = true;
    float max_p = 0.f, max_o = 0.f;
#pragma unroll
    for (int e = 0; e < NEE; ++e) {
      const float np = sqrtf(r[6 * e] * r[6 * e] + r[6 * e + 1] * r[6 * e + 1] + r[6 * e + 2] * r[6 * e + 2]);
      const float no = sqrtf(r[6 * e + 3] * r[6 * e + 3] + r[6 * e + 4] * r[6 * e + 4] + r[6 * e + 5] * r[6 * e + 5]);
      if (e < a.n_ee && (np >= a.eps_pos || no >= a.eps_ori)) conv = false;
      max_p = fmaxf(max_p, np);
      max_o = fmaxf(max_o, no);
    }
    if (conv) { st.done = 1; break; }
    float fw[6 * NEE];
#pragma unroll
    for (int i = 0; i < 6 * NEE; ++i) fw[i] = r[i] * W[i % 6];
    const float cur = cur_next;  // the accepted point's error (residuals and penalty)

    // Jacobian contribution of node `lane` to its column.
    if (lane < MAXA) {
      float g6[6 * NEE];
#pragma unroll
      for (int i = 0; i < 6 * NEE; ++i) g6[i] = 0.f;
      if (lane < sh.n_nodes) {
        const NodeRec& nr = sh.nodes[lane];
        const int jt = nr.jac_type;
        if (jt != 0) {
          const Pose Pn = load_pose(wl.P[lane]);
          const unsigned em = nr.ee_mask;
#pragma unroll
          for (int e = 0; e < NEE; ++e)
            if (em & (1u << e)) jac_column(Pn, nr.ax, jt, wl.Tee[e][4], wl.Tee[e][5], wl.Tee[e][6], nr.mul, W, g6 + 6 * e);
        }
      }
#pragma unroll
      for (int i = 0; i < 6 * NEE; ++i) wl.Jn[lane][i] = g6[i];
    }
    wave_sync();
    // Column `lane`: its nodes' contributions summed in node order, then scaled.
    if (lane < MAXA) {
      float col[6 * NEE];
#pragma unroll
      for (int i = 0; i < 6 * NEE; ++i) col[i] = 0.f;
      const int cnt = wl.cnt[lane];
      for (int k = 0; k < cnt; ++k) {
        const int n = wl.drv[lane][k];
#pragma unroll
        for (int i = 0; i < 6 * NEE; ++i) col[i] += wl.Jn[n][i];
      }
      float sq = 0.f;
#pragma unroll
      for (int i = 0; i < 6 * NEE; ++i) sq += col[i] * col[i];
      const float ic = 1.f / (sqrtf(sq) + 1e-8f);
#pragma unroll
      for (int i = 0; i < 6 * NEE; ++i) wl.Js[lane][i] = col[i] * ic;
      wl.ics[lane] = ic;
    }
    wave_sync();
    // Packed normal matrix entry (i, j) on lane i (i + 1) / 2 + j, right-hand
    // side i on lane NT + i; damping and frozen joints as in lm_iterate.
    for (int t = lane; t < NT + MAXA; t += kWave) {
      if (t < NT) {
        int i = 0;
        while ((i + 1) * (i + 2) / 2 <= t) ++i;
        const int j = t - i * (i + 1) / 2;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 6 * NEE; ++k) acc += wl.Js[i][k] * wl.Js[j][k];
        if (i == j) acc += fmaxf(st.lam, kDampFloor);
        if (sh.fixed[i] || sh.fixed[j]) acc = (i == j) ? 1.f : 0.f;
        wl.A[t] = acc;
      } else {
        const int i = t - NT;
        float rb = 0.f;
#pragma unroll
        for (int k = 0; k < 6 * NEE; ++k) rb += wl.Js[i][k] * fw[k];
        wl.rhs[i] = sh.fixed[i] ? 0.f : -rb;
      }
    }
    wave_sync();
    float A[NT];
    float rhs[MAXA];
    float ics[MAXA];
#pragma unroll
    for (int t = 0; t < NT; ++t) A[t] = wl.A[t];
#pragma unroll
    for (int i = 0; i < MAXA; ++i) { rhs[i] = wl.rhs[i]; ics[i] = wl.ics[i]; }
    bool ok = true;
    float idg[MAXA];
#pragma unroll
    for (int k = 0; k < MAXA; ++k) {
      float sd = A[k * (k + 1) / 2 + k];
#pragma unroll
      for (int pp = 0; pp < k; ++pp) sd -= A[k * (k + 1) / 2 + pp] * A[k * (k + 1) / 2 + pp];
      if (sd <= 0.f) ok = false;
      const float lkk = sqrtf(fmaxf(sd, 1e-30f));
      A[k * (k + 1) / 2 + k] = lkk;
      const float inv = 1.f / lkk;
      idg[k] = inv;
#pragma unroll
      for (int i = k + 1; i < MAXA; ++i) {
        float t = A[i * (i + 1) / 2 + k];
#pragma unroll
        for (int pp = 0; pp < k; ++pp) t -= A[i * (i + 1) / 2 + pp] * A[k * (k + 1) / 2 + pp];
        A[i * (i + 1) / 2 + k] = t * inv;
      }
    }
    float y[MAXA];
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
      float t = rhs[i];
#pragma unroll
      for (int pp = 0; pp < i; ++pp) t -= A[i * (i + 1) / 2 + pp] * y[pp];
      y[i] = t * idg[i];
    }
    float delta[MAXA];
#pragma unroll
    for (int i = MAXA - 1; i >= 0; --i) {
      float t = y[i];
#pragma unroll
      for (int pp = i + 1; pp < MAXA; ++pp) t -= A[pp * (pp + 1) / 2 + i] * delta[pp];
      delta[i] = t * idg[i];
    }
    float dn = 0.f;
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
      delta[i] = ok ? delta[i] * ics[i] : 0.f;
      dn += delta[i] * delta[i];
    }
    float R;
    if (max_p > 1e-2f || max_o > 0.6f) R = 0.38f;
    else if (max_p > 1e-3f || max_o > 0.25f) R = 0.22f;
    else if (max_p > 2e-4f || max_o > 0.08f) R = 0.12f;
    else R = 0.05f;
    dn = sqrtf(dn);
    if (dn > R) {
      const float sc = R / (dn + 1e-18f);
#pragma unroll
      for (int i = 0; i < MAXA; ++i) delta[i] *= sc;
    }

    // Line search: lane k < 8 evaluates step min(k, 4); an argmin over lanes
    // 0-7 gives the sequential pick (first index on ties).
    Pose P2[MAXA], Tee2[NEE];
    float r2[6 * NEE];
    float e = INFINITY;
    int idx = 0;
    if (lane < 8) {
      const int ai = min(lane, 4);
      float qt[MAXA];
#pragma unroll
      for (int i = 0; i < MAXA; ++i) qt[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], st.q[i] + ls_alpha(ai) * delta[i]));
      unsigned wk = 0;
      const float et = weighted_error(v, tg, W, qt, P2, r2, Tee2, grid, wk);
      e = (et < 1e30f) ? et : INFINITY;
      idx = ai;
      ls_argmin_step<0xb1>(e, idx);   // quad_perm [1,0,3,2]
      ls_argmin_step<0x4e>(e, idx);   // quad_perm [2,3,0,1]
      ls_argmin_step<0x141>(e, idx);  // row_half_mirror
    }
    const float e0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e), 0));
    const int i0 = __builtin_amdgcn_readlane(idx, 0);
    float best_a_err = 1e30f;
    int best_a = 0;
    if (e0 != INFINITY) { best_a_err = e0; best_a = i0; }
    const bool accept = best_a_err < cur * (1.f - 1e-4f);
    if (accept && lane == best_a) store_iterate<MAXA, NEE>(wl, P2, Tee2, r2);
    cur_next = accept ? best_a_err : cur;
    const float al = ls_alpha(best_a);
    float qt[MAXA];
#pragma unroll
    for (int i = 0; i < MAXA; ++i) qt[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], st.q[i] + al * delta[i]));
    if (accept) {
#pragma unroll
      for (int i = 0; i < MAXA; ++i) st.q[i] = qt[i];
      st.lam = fmaxf(st.lam * 0.5f, 1e-10f);
    } else {
      st.lam = fminf(st.lam * 3.f, 1e6f);
    }
    if (best_a_err < st.best_err) {
      st.best_err = best_a_err;
#pragma unroll
      for (int i = 0; i < MAXA; ++i) st.best_q[i] = qt[i];
    }
  }
}

template <int MAXA, int NEE>
__global__ __launch_bounds__(64) void ls_ik_wave_kernel(IkArgs a) {
  extern __shared__ float4 smem4[];
  char* base = reinterpret_cast<char*>(smem4);
  const IkView<MAXA, NEE> v = ik_prologue<MAXA, NEE>(a, base);
  using Sh = IkShared<MAXA, NEE>;
  const int Kin = a.enable_collision ? a.n_robot_spheres : 0;
  const IkLds off = ik_lds_layout(sizeof(Sh), a.tables.n_joints, Kin, a.enable_collision ? a.n_world : 0,
                                  ik_other_floats(a));
  WaveLds<MAXA, NEE>& wl = *reinterpret_cast<WaveLds<MAXA, NEE>*>(base + align16(off.total));
  const int lane = threadIdx.x;
  const int p = blockIdx.y;
  const int s = blockIdx.x;
  const long long gs = static_cast<long long>(p) * a.n_seeds + s;
  if (v.sh->too_many_nodes) {  // reported to the caller as NaN errors, seeds unchanged
    if (lane == 0) {
      for (int i = 0; i < a.n_act; ++i) a.out_cfgs[gs * a.n_act + i] = a.seeds[gs * a.n_act + i];
      a.out_err[gs] = NAN;
    }
    return;
  }
  if (lane < MAXA) {  // the nodes driving column `lane`, in node order
    int c = 0;
    for (int n = 0; n < v.sh->n_nodes; ++n) {
      const NodeRec& nr = v.sh->nodes[n];
      if (nr.jac_type != 0 && nr.src == lane) wl.drv[lane][c++] = n;
    }
    wl.cnt[lane] = c;
  }
  float tg[NEE][7];
  load_targets<MAXA, NEE>(a, p, tg);
  const float W[6] = {a.pos_weight, a.pos_weight, a.pos_weight, a.ori_weight, a.ori_weight, a.ori_weight};
  LmState<MAXA> st;
#pragma unroll
  for (int i = 0; i < MAXA; ++i) {
    st.q[i] = (i < a.n_act) ? a.seeds[gs * a.n_act + i] : 0.f;
    st.best_q[i] = st.q[i];
  }
  Pose P[MAXA];
  Pose Tee[NEE];
  float r[6 * NEE];
  st.work = 0;
  st.best_err = weighted_error(v, tg, W, st.q, P, r, Tee, a.grid, st.work);
  if (lane == 0) store_iterate<MAXA, NEE>(wl, P, Tee, r);
  st.lam = a.lambda_init;
  st.done = 0;
  lm_iterate_wave<MAXA, NEE>(v, tg, W, a, st, wl, st.best_err);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < MAXA; ++i)
      if (i < a.n_act) a.out_cfgs[gs * a.n_act + i] = st.best_q[i];
    a.out_err[gs] = st.best_err;
  }
}

// Uneven batch: seeds of all problems in one flat list (problem p owns
// [problem_seed_start[p], problem_seed_start[p+1])).  A workgroup takes tiles
// of 64 consecutive seeds round-robin (tiles may straddle problems: the target
// is per lane).  Iterations [iter_begin, iter_end) run per launch; with a state
// buffer the solve resumes across launches (a captured graph of chunks).
template <int MAXA, int NEE>
__global__ __launch_bounds__(64) void ls_ik_batched_kernel(IkArgs a) {
  extern __shared__ float4 smem4[];
  const IkView<MAXA, NEE> v = ik_prologue<MAXA, NEE>(a, reinterpret_cast<char*>(smem4));
  const int lane = threadIdx.x;
  const long long n_tiles = (a.total_seeds + kWave - 1) / kWave;
  const int A = a.n_act;
  const int stride = 2 * A + 3;
  const float W[6] = {a.pos_weight, a.pos_weight, a.pos_weight, a.ori_weight, a.ori_weight, a.ori_weight};
  for (long long tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const long long s = tile * kWave + lane;
    if (s >= a.total_seeds) continue;
    if (v.sh->too_many_nodes) {
      if (a.iter_end >= a.max_iter) {
        for (int i = 0; i < A; ++i) a.out_cfgs[s * A + i] = a.seeds[s * A + i];
        a.out_err[s] = NAN;
      }
      continue;
    }
    // Problem of this seed: binary search in the CSR offsets.
    int lo = 0, hi = a.n_problems - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.problem_seed_start[mid] <= s) lo = mid; else hi = mid - 1;
    }
    float tg[NEE][7];
    load_targets<MAXA, NEE>(a, lo, tg);
    LmState<MAXA> st;
    if (a.iter_begin == 0 || a.state == nullptr) {
#pragma unroll
      for (int i = 0; i < MAXA; ++i) {
        st.q[i] = (i < A) ? a.seeds[s * A + i] : 0.f;
        st.best_q[i] = st.q[i];
      }
      Pose P[MAXA];
      Pose Tee[NEE];
      float r[6 * NEE];
      st.work = 0;
      st.best_err = weighted_error(v, tg, W, st.q, P, r, Tee, a.grid, st.work);
      st.lam = a.lambda_init;
      st.done = 0;
    } else {
      const float* sp = a.state + s * stride;
#pragma unroll
      for (int i = 0; i < MAXA; ++i) {
        st.q[i] = (i < A) ? sp[i] : 0.f;
        st.best_q[i] = (i < A) ? sp[A + i] : 0.f;
      }
      st.best_err = sp[2 * A];
      st.lam = sp[2 * A + 1];
      st.done = sp[2 * A + 2] != 0.f;
      st.work = 0;
    }
    lm_iterate(v, tg, W, a, st, a.iter_begin, a.iter_end, a.grid);
    if (a.work != nullptr) atomicAdd(a.work, static_cast<unsigned long long>(st.work));
    if (a.state != nullptr && a.iter_end < a.max_iter) {
      float* sp = a.state + s * stride;
#pragma unroll
      for (int i = 0; i < MAXA; ++i)
        if (i < A) { sp[i] = st.q[i]; sp[A + i] = st.best_q[i]; }
      sp[2 * A] = st.best_err;
      sp[2 * A + 1] = st.lam;
      sp[2 * A + 2] = st.done ? 1.f : 0.f;
    } else {
#pragma unroll
      for (int i = 0; i < MAXA; ++i)
        if (i < A) a.out_cfgs[s * A + i] = st.best_q[i];
      a.out_err[s] = st.best_err;
    }
  }
}

template <int MAXA, int NEE>
inline int launch_ik(void* stream, const IkArgs& a, bool batched, int grid_waves) {
  using Sh = IkShared<MAXA, NEE>;
  const int J = a.tables.n_joints;
  const int Kin = a.enable_collision ? a.n_robot_spheres : 0;
  const int M = a.enable_collision ? a.n_world : 0;
  const size_t lds = ik_lds_layout(sizeof(Sh), J, Kin, M, ik_other_floats(a)).total;
  PK_REQUIRE(lds <= 150 * 1024, "IK tables need %zu bytes of LDS (limit 150 KiB)", lds);
  if (batched) {
    const long long n_tiles = (a.total_seeds + kWave - 1) / kWave;
    long long g = grid_waves > 0 ? grid_waves : n_tiles;
    g = std::max(1LL, std::min(g, n_tiles));
    hipLaunchKernelGGL((ls_ik_batched_kernel<MAXA, NEE>), dim3(static_cast<unsigned>(g)), dim3(kWave), lds,
                       static_cast<hipStream_t>(stream), a);
    return check_launch("ls_ik_batched_kernel");
  }
  // Small batches leave most of the GPU idle and run at the latency of one
  // seed's iterations: the 8-wide instantiations give each seed a wave while
  // every SIMD holds at most one of them, and every instantiation gives
  // each seed 8 lanes (parallel line search) while the lanes still fit one
  // pass of the CUs at two waves per SIMD.
  const long long n_all = static_cast<long long>(a.n_problems) * a.n_seeds;
  if (MAXA == 8 && n_all <= wave_seed_limit()) {
    const size_t lds_w = align16(lds) + sizeof(WaveLds<MAXA, NEE>);
    PK_REQUIRE(lds_w <= 150 * 1024, "IK tables need %zu bytes of LDS (limit 150 KiB)", lds_w);
    dim3 grid(a.n_seeds, a.n_problems);
    hipLaunchKernelGGL((ls_ik_wave_kernel<MAXA, NEE>), grid, dim3(kWave), lds_w,
                       static_cast<hipStream_t>(stream), a);
    return check_launch("ls_ik_wave_kernel");
  }
  const long long lanes8 = n_all * 8;
  if (lanes8 <= kSmallBatchLanes) {
    constexpr int G = 8;
    dim3 grid((a.n_seeds + kWave / G - 1) / (kWave / G), a.n_problems);
    hipLaunchKernelGGL((ls_ik_kernel<MAXA, NEE, G>), grid, dim3(kWave), lds,
                       static_cast<hipStream_t>(stream), a);
  } else {
    dim3 grid((a.n_seeds + kWave - 1) / kWave, a.n_problems);
    hipLaunchKernelGGL((ls_ik_kernel<MAXA, NEE, 1>), grid, dim3(kWave), lds,
                       static_cast<hipStream_t>(stream), a);
  }
  return check_launch("ls_ik_kernel");
}

// One launcher per (padded DOF, padded EE count), each compiled from
// pk_ik_inst.hip in a translation unit of its own.
#define PK_IK_LAUNCHERS(X) X(8, 1) X(8, 2) X(8, 4) X(16, 1) X(16, 2) X(16, 4)
#define PK_IK_DECLARE(m, e) int launch_ik_##m##_##e(void* stream, const IkArgs& a, bool batched, int grid_waves);
PK_IK_LAUNCHERS(PK_IK_DECLARE)

}  // namespace pk
