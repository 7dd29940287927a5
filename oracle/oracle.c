/*
 * FP64 CPU oracle for the batched kinematics path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates the reference algorithms in double precision so the HIP
 * kernels can be checked against it.  It is compiled into oracle/_build/ and
 * is loaded only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py -- never by the pyroffi_amd package.
 *
 * Functions and the reference code they follow:
 *   or_fk            src/pyroffi/_robot.py:198-263 (JAX FK: fori_loop over the
 *                    topological order) and _fk_cuda_helpers.cuh:158-208; SE(3)
 *                    exp/compose per jaxlie 1.5.0 (FP64 Taylor branch below
 *                    theta^2 < 1e-10, as jaxlie does for float64)
 *   or_link_poses    src/pyroffi/_robot.py:180-196
 *   or_world_dist    CUDA semantics: _collision_cuda_kernel.cu:169-201 (padding
 *                    skipped, 1e9 when a link has no sphere); JAX semantics
 *                    (bias=1): _robot_collision.py:830-895 with the +1e-6 bias
 *                    of collision/_utils.py:45 and padding evaluated as r=-1e9
 *   or_self_dist     CUDA: _collision_cuda_kernel.cu:425-459 (full S x S);
 *                    JAX (mode 1): _robot_collision.py:796-828 (same-index
 *                    "diagonal" sphere pairs, +1e-6 bias)
 *   or_ls_ik         _ls_ik_cuda_kernel.cu:64-394 with the residual/Jacobian
 *                    of _ik_cuda_helpers.cuh:433-631 and the Cholesky of
 *                    _ik_cuda_helpers.cuh:207-243
 *   or_pose_hinge_cost  pose residual + CHOMP hinge^2 (_chomp_optimization.py:140-154)
 *   or_traj_cost     trajectory smoothness / limits / collision costs and
 *                    their analytic gradient (_sco_optimization.py:170-222,
 *                    collision/_collision.py:136-167)
 *   or_two_loop      _sco_optimization.py:119-167
 *   or_traj_lbfgs    _lbfgs_trajopt_optimization.py:76-205
 *   or_jres          the JAX residual log(T_ee^-1 T_tgt) (_ik_primitives.py:93-103,
 *                    jaxlie 1.5.0 SE3.log) and its exact Jacobian
 *   or_jlm           _ls_ik_single (_ls_ik.py:192-350) and _lm_refine_single
 *                    (_hjcd_ik.py:140-365) per seed
 *   or_hjcd_coarse   _coarse_search_single (_hjcd_ik.py:55-137)
 *   or_obstacle_dist sphere vs sphere / capsule / box / half-space
 *                    (_collision_cuda_helpers.cuh:14-142, packed rows of
 *                    collision/_cuda_collision.py:98-168)
 *   or_capsule_dist  capsule vs the same four kinds (_collision_cuda_helpers.cuh:21-142)
 *
 * Every function here is pinned by tests/test_oracle.py against checks that
 * share no code with it (known answers, scipy, finite differences).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXJ 64
#define OR_MAXA 16
#define OR_MAXEE 4

typedef struct {
  double w, x, y, z, tx, ty, tz;
} se3;

static se3 se3_load_f(const float* p) {
  se3 o = {p[0], p[1], p[2], p[3], p[4], p[5], p[6]};
  return o;
}

static void qmul(double aw, double ax, double ay, double az, double bw, double bx, double by, double bz,
                 double* o) {
  o[0] = aw * bw - ax * bx - ay * by - az * bz;
  o[1] = aw * bx + ax * bw + ay * bz - az * by;
  o[2] = aw * by - ax * bz + ay * bw + az * bx;
  o[3] = aw * bz + ax * by - ay * bx + az * bw;
}

/* Rotate v by the rotation matrix of unit quaternion q. */
static void qrot(const se3* q, const double* v, double* o) {
  const double w = q->w, x = q->x, y = q->y, z = q->z;
  const double r00 = 1 - 2 * (y * y + z * z), r01 = 2 * (x * y - w * z), r02 = 2 * (x * z + w * y);
  const double r10 = 2 * (x * y + w * z), r11 = 1 - 2 * (x * x + z * z), r12 = 2 * (y * z - w * x);
  const double r20 = 2 * (x * z - w * y), r21 = 2 * (y * z + w * x), r22 = 1 - 2 * (x * x + y * y);
  o[0] = r00 * v[0] + r01 * v[1] + r02 * v[2];
  o[1] = r10 * v[0] + r11 * v[1] + r12 * v[2];
  o[2] = r20 * v[0] + r21 * v[1] + r22 * v[2];
}

static se3 se3_mul(const se3* a, const se3* b) {
  se3 o;
  double q[4], t[3], v[3] = {b->tx, b->ty, b->tz};
  qmul(a->w, a->x, a->y, a->z, b->w, b->x, b->y, b->z, q);
  qrot(a, v, t);
  o.w = q[0]; o.x = q[1]; o.y = q[2]; o.z = q[3];
  o.tx = t[0] + a->tx; o.ty = t[1] + a->ty; o.tz = t[2] + a->tz;
  return o;
}

/* jaxlie SE3.exp of tangent [v; w]. */
static se3 se3_exp(const double* tg) {
  const double vx = tg[0], vy = tg[1], vz = tg[2], wx = tg[3], wy = tg[4], wz = tg[5];
  const double th2 = wx * wx + wy * wy + wz * wz;
  se3 o;
  const int taylor = th2 < 1e-10;
  double re, im;
  if (taylor) {
    re = 1.0 - th2 / 8.0 + th2 * th2 / 384.0;
    im = 0.5 - th2 / 48.0 + th2 * th2 / 3840.0;
  } else {
    const double th = sqrt(th2);
    re = cos(0.5 * th);
    im = sin(0.5 * th) / th;
  }
  o.w = re; o.x = im * wx; o.y = im * wy; o.z = im * wz;
  double V[9];
  if (taylor) {
    const double x = o.x, y = o.y, z = o.z, w = o.w;
    V[0] = 1 - 2 * (y * y + z * z); V[1] = 2 * (x * y - w * z); V[2] = 2 * (x * z + w * y);
    V[3] = 2 * (x * y + w * z); V[4] = 1 - 2 * (x * x + z * z); V[5] = 2 * (y * z - w * x);
    V[6] = 2 * (x * z - w * y); V[7] = 2 * (y * z + w * x); V[8] = 1 - 2 * (x * x + y * y);
  } else {
    const double th = sqrt(th2);
    const double b = (1.0 - cos(th)) / th2, c = (th - sin(th)) / (th2 * th);
    /* V = I + b [w]x + c [w]x^2 ; [w]x^2 = w w^T - th2 I */
    const double K[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
    const double W[3] = {wx, wy, wz};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        V[3 * i + j] = (i == j ? 1.0 : 0.0) + b * K[3 * i + j] + c * (W[i] * W[j] - (i == j ? th2 : 0.0));
  }
  o.tx = V[0] * vx + V[1] * vy + V[2] * vz;
  o.ty = V[3] * vx + V[4] * vy + V[5] * vz;
  o.tz = V[6] * vx + V[7] * vy + V[8] * vz;
  return o;
}

typedef struct {
  int J, A;
  const float *twists, *parent_tf, *mimic_mul, *mimic_off;
  const int32_t *parent_idx, *act_idx, *mimic_act_idx, *topo_inv;
} tables_t;

/* World poses of all joints for one configuration (joint order of the tables). */
static void fk_one(const tables_t* t, const double* q_act, se3* Tw) {
  for (int i = 0; i < t->J; ++i) {
    const int j = t->topo_inv[i];
    const int m = t->mimic_act_idx[j];
    const int src = (m != -1) ? m : t->act_idx[j];
    const double qref = (src < 0) ? 0.0 : q_act[src];
    const double q = qref * (double)t->mimic_mul[j] + (double)t->mimic_off[j];
    double tg[6];
    for (int c = 0; c < 6; ++c) tg[c] = (double)t->twists[6 * j + c] * q;
    const se3 ptf = se3_load_f(t->parent_tf + 7 * j);
    const se3 d = se3_exp(tg);
    const se3 tpc = se3_mul(&ptf, &d);
    const int p = t->parent_idx[j];
    Tw[j] = (p < 0) ? tpc : se3_mul(&Tw[p], &tpc);
  }
}

static void set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void or_fk(const double* cfg, int64_t B, int A, const float* twists, const float* parent_tf,
           const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
           const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv, int J,
           double* out, int threads) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < B; ++b) {
    se3 Tw[OR_MAXJ];
    fk_one(&t, cfg + b * A, Tw);
    for (int j = 0; j < J; ++j) {
      double* o = out + (b * J + j) * 7;
      o[0] = Tw[j].w; o[1] = Tw[j].x; o[2] = Tw[j].y; o[3] = Tw[j].z;
      o[4] = Tw[j].tx; o[5] = Tw[j].ty; o[6] = Tw[j].tz;
    }
  }
}

void or_link_poses(const double* joint_poses, int64_t B, int J, const int32_t* link_parent_joint, int L,
                   double* out) {
  for (int64_t b = 0; b < B; ++b)
    for (int l = 0; l < L; ++l) {
      const int j = link_parent_joint[l];
      double* o = out + (b * L + l) * 7;
      if (j < 0) {
        o[0] = 1; o[1] = o[2] = o[3] = o[4] = o[5] = o[6] = 0;
      } else {
        memcpy(o, joint_poses + (b * J + j) * 7, 7 * sizeof(double));
      }
    }
}

/* World-frame sphere centre from a link pose and a link-frame centre. */
static void place(const double* pose, const double* c, double* o) {
  se3 T = {pose[0], pose[1], pose[2], pose[3], pose[4], pose[5], pose[6]};
  qrot(&T, c, o);
  o[0] += T.tx; o[1] += T.ty; o[2] += T.tz;
}

static double sdist(const double* a, double ra, const double* b, double rb, int bias) {
  const double e = bias ? 1e-6 : 0.0;
  const double dx = b[0] - a[0] + e, dy = b[1] - a[1] + e, dz = b[2] - a[2] + e;
  return sqrt(dx * dx + dy * dy + dz * dz) - (ra + rb);
}

/* spheres: padded (L, S, 4) link-frame table, r < 0 = padding.
 * out (B, L, M) may be NULL; out_min (B) may be NULL. */
void or_world_dist(const double* link_poses, int64_t B, int L, const double* spheres, int S,
                   const double* world, int M, int bias, double* out, double* out_min, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < B; ++b) {
    double gmin = 1e9;
    for (int n = 0; n < L; ++n) {
      double cw[64][3];
      const int Su = S < 64 ? S : 64;
      for (int s = 0; s < Su; ++s) place(link_poses + (b * L + n) * 7, spheres + (n * S + s) * 4, cw[s]);
      for (int m = 0; m < M; ++m) {
        double best = bias ? INFINITY : 1e9;
        for (int s = 0; s < Su; ++s) {
          const double r = spheres[(n * S + s) * 4 + 3];
          if (!bias && r < 0) continue;
          const double d = sdist(cw[s], r, world + 4 * m, world[4 * m + 3], bias);
          if (d < best) best = d;
        }
        if (out) out[(b * L + n) * M + m] = best;
        if (best < gmin) gmin = best;
      }
    }
    if (out_min) out_min[b] = gmin;
  }
}

/* mode 0: CUDA semantics (all S x S pairs, padding skipped, 1e9 sentinel);
 * mode 1: JAX semantics (sphere s of link i vs sphere s of link j, +1e-6 bias). */
void or_self_dist(const double* link_poses, int64_t B, int L, const double* spheres, int S,
                  const int32_t* pair_i, const int32_t* pair_j, int P, int mode, double* out, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < B; ++b) {
    for (int p = 0; p < P; ++p) {
      const int li = pair_i[p], lj = pair_j[p];
      double best = mode ? INFINITY : 1e9;
      for (int si = 0; si < S; ++si) {
        const double ri = spheres[(li * S + si) * 4 + 3];
        if (!mode && ri < 0) continue;
        double ci[3];
        place(link_poses + (b * L + li) * 7, spheres + (li * S + si) * 4, ci);
        for (int sj = mode ? si : 0; sj < (mode ? si + 1 : S); ++sj) {
          const double rj = spheres[(lj * S + sj) * 4 + 3];
          if (!mode && rj < 0) continue;
          double cj[3];
          place(link_poses + (b * L + lj) * 7, spheres + (lj * S + sj) * 4, cj);
          const double d = sdist(ci, ri, cj, rj, mode);
          if (d < best) best = d;
        }
      }
      out[b * P + p] = best;
    }
  }
}

/* Signed distance of sphere (p, r) to one packed obstacle of kind k
 * (0 sphere [c, R], 1 capsule [a, b, R], 2 box [c, ax1, ax2, ax3, h],
 * 3 half-space [n, point]). */
static double obstacle_dist(const double* p, double r, int k, const double* o) {
  if (k == 0) {
    const double dx = p[0] - o[0], dy = p[1] - o[1], dz = p[2] - o[2];
    return sqrt(dx * dx + dy * dy + dz * dz) - (r + o[3]);
  }
  if (k == 1) {
    const double u[3] = {o[3] - o[0], o[4] - o[1], o[5] - o[2]};
    const double w[3] = {p[0] - o[0], p[1] - o[1], p[2] - o[2]};
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double t = 0.0;
    if (uu > 1e-12) {
      t = (w[0] * u[0] + w[1] * u[1] + w[2] * u[2]) / uu;
      t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    }
    const double dx = w[0] - t * u[0], dy = w[1] - t * u[1], dz = w[2] - t * u[2];
    return sqrt(dx * dx + dy * dy + dz * dz) - (r + o[6]);
  }
  if (k == 2) {
    const double d[3] = {p[0] - o[0], p[1] - o[1], p[2] - o[2]};
    double e[3], out = 0.0, in = -INFINITY;
    for (int a = 0; a < 3; ++a) {
      const double loc = d[0] * o[3 + 3 * a] + d[1] * o[4 + 3 * a] + d[2] * o[5 + 3 * a];
      e[a] = fabs(loc) - o[12 + a];
      if (e[a] > 0) out += e[a] * e[a];
      if (e[a] > in) in = e[a];
    }
    return sqrt(out) + (in < 0 ? in : 0.0) - r;
  }
  return (p[0] - o[3]) * o[0] + (p[1] - o[4]) * o[1] + (p[2] - o[5]) * o[2] - r;
}

/* Closest points of segments [a0, a1] and [b0, b1] (clamped parameters):
 * returns the distance of the points. */
static double seg_seg(const double* a0, const double* a1, const double* b0, const double* b1) {
  const double d1[3] = {a1[0] - a0[0], a1[1] - a0[1], a1[2] - a0[2]};
  const double d2[3] = {b1[0] - b0[0], b1[1] - b0[1], b1[2] - b0[2]};
  const double r[3] = {a0[0] - b0[0], a0[1] - b0[1], a0[2] - b0[2]};
  const double a = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
  const double e = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
  const double f = d2[0] * r[0] + d2[1] * r[1] + d2[2] * r[2];
  const double eps = 1e-10;
  double s = 0.0, t = 0.0;
  if (a > eps || e > eps) {
    if (a <= eps) {
      t = f / e; t = t < 0 ? 0 : (t > 1 ? 1 : t);
    } else {
      const double c = d1[0] * r[0] + d1[1] * r[1] + d1[2] * r[2];
      if (e <= eps) {
        s = -c / a; s = s < 0 ? 0 : (s > 1 ? 1 : s);
      } else {
        const double b = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
        const double den = a * e - b * b;
        if (fabs(den) > eps) { s = (b * f - c * e) / den; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
        t = (b * s + f) / e;
        if (t < 0) { t = 0; s = -c / a; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
        else if (t > 1) { t = 1; s = (b - c) / a; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
      }
    }
  }
  double d2s = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double v = a0[k] + s * d1[k] - (b0[k] + t * d2[k]);
    d2s += v * v;
  }
  return sqrt(d2s);
}

/* Signed distance of capsule c [a, b, r] to one packed obstacle of kind k
 * (_collision_cuda_helpers.cuh:21-142: the capsule-box distance takes the box
 * SDF at the segment point closest to the box centre). */
static double capsule_obstacle_dist(const double* c, int k, const double* o) {
  if (k == 0) {
    const double cap[7] = {c[0], c[1], c[2], c[3], c[4], c[5], o[3]};
    return obstacle_dist(o, c[6], 1, cap);
  }
  if (k == 1) return seg_seg(c, c + 3, o, o + 3) - (c[6] + o[6]);
  if (k == 2) {
    double al[3], u[3];
    for (int a = 0; a < 3; ++a) {
      const double* ax = o + 3 + 3 * a;
      al[a] = (c[0] - o[0]) * ax[0] + (c[1] - o[1]) * ax[1] + (c[2] - o[2]) * ax[2];
      u[a] = (c[3] - o[0]) * ax[0] + (c[4] - o[1]) * ax[1] + (c[5] - o[2]) * ax[2] - al[a];
    }
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double t = 0.0;
    if (uu > 1e-12) { t = -(al[0] * u[0] + al[1] * u[1] + al[2] * u[2]) / uu; t = t < 0 ? 0 : (t > 1 ? 1 : t); }
    const double p[3] = {al[0] + t * u[0], al[1] + t * u[1], al[2] + t * u[2]};
    /* box SDF of the box-frame point p: an axis-aligned box at the origin */
    const double box[15] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, o[12], o[13], o[14]};
    return obstacle_dist(p, c[6], 2, box);
  }
  const double d1 = (c[0] - o[3]) * o[0] + (c[1] - o[4]) * o[1] + (c[2] - o[5]) * o[2];
  const double d2 = (c[3] - o[3]) * o[0] + (c[4] - o[4]) * o[1] + (c[5] - o[5]) * o[2];
  return (d1 < d2 ? d1 : d2) - c[6];
}

/* out (N, M): capsules (N, 7) against the packed world (order as below). */
void or_capsule_dist(const double* caps, int64_t N, const double* ws, int ms, const double* wc, int mc,
                     const double* wb, int mb, const double* wh, int mh, double* out, int threads) {
  const int M = ms + mc + mb + mh;
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    for (int m = 0; m < M; ++m) {
      int k = 0, i = m;
      const double* o;
      if (i < ms) { o = ws + 4 * i; }
      else if ((i -= ms) < mc) { k = 1; o = wc + 7 * i; }
      else if ((i -= mc) < mb) { k = 2; o = wb + 15 * i; }
      else { i -= mb; k = 3; o = wh + 6 * i; }
      out[n * M + m] = capsule_obstacle_dist(caps + 7 * n, k, o);
    }
  }
}

/* out (N, M) with M = Ms + Mc + Mb + Mh obstacles in that order; sphere n is
 * centres[3n..3n+2], radius radii[n]; r < 0 is padding and reads 1e9. */
void or_obstacle_dist(const double* centres, const double* radii, int64_t N, const double* ws, int ms,
                      const double* wc, int mc, const double* wb, int mb, const double* wh, int mh, double* out,
                      int threads) {
  const int M = ms + mc + mb + mh;
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    for (int m = 0; m < M; ++m) {
      double d = 1e9;
      if (radii[n] >= 0) {
        int k = 0, i = m;
        const double* o;
        if (i < ms) { o = ws + 4 * i; }
        else if ((i -= ms) < mc) { k = 1; o = wc + 7 * i; }
        else if ((i -= mc) < mb) { k = 2; o = wb + 15 * i; }
        else { i -= mb; k = 3; o = wh + 6 * i; }
        d = obstacle_dist(centres + 3 * n, radii[n], k, o);
      }
      out[n * M + m] = d;
    }
  }
}

/* ------------------------------------------------------------------ LM IK */

typedef struct {
  const tables_t* t;
  int n_ee;
  const int32_t* target_jnts;
  const int32_t* anc;  /* (n_ee, J) */
  const double* tgt;   /* (n_ee, 7) */
  const float* rs;     /* (K, 4) joint-frame spheres */
  const int32_t* rj;
  int K;
  const float* ws;     /* (M, 4) */
  int M;
  int coll;
  double wc, margin;
  /* capsules (mc, 7), boxes (mb, 15), half-spaces (mh, 6); may be NULL/0 */
  const float *wcap, *wbox, *whs;
  int mc, mb, mh;
} ikp_t;

static void residual_ee(const se3* T, const double* tg, double* r) {
  r[0] = T->tx - tg[4]; r[1] = T->ty - tg[5]; r[2] = T->tz - tg[6];
  double e[4];
  qmul(T->w, T->x, T->y, T->z, tg[0], -tg[1], -tg[2], -tg[3], e);
  if (e[0] < 0) { e[0] = -e[0]; e[1] = -e[1]; e[2] = -e[2]; e[3] = -e[3]; }
  const double sh = sqrt(e[1] * e[1] + e[2] * e[2] + e[3] * e[3]);
  if (sh > 1e-6) {
    const double k = 2.0 * atan2(sh, e[0]) / sh;
    r[3] = e[1] * k; r[4] = e[2] * k; r[5] = e[3] * k;
  } else {
    r[3] = 2 * e[1]; r[4] = 2 * e[2]; r[5] = 2 * e[3];
  }
}

static double penalty(const ikp_t* p, const se3* Tw) {
  if (!p->coll || p->K <= 0) return 0.0;
  double pen = 0.0;
  for (int i = 0; i < p->K; ++i) {
    const int j = p->rj[i];
    if (j < 0 || j >= p->t->J) continue;
    const double c[3] = {p->rs[4 * i], p->rs[4 * i + 1], p->rs[4 * i + 2]};
    double w[3];
    qrot(&Tw[j], c, w);
    w[0] += Tw[j].tx; w[1] += Tw[j].ty; w[2] += Tw[j].tz;
    for (int m = 0; m < p->M; ++m) {
      const double o[3] = {p->ws[4 * m], p->ws[4 * m + 1], p->ws[4 * m + 2]};
      const double d = sdist(w, p->rs[4 * i + 3], o, p->ws[4 * m + 3], 0);
      if (d < p->margin) pen += (d - p->margin) * (d - p->margin);
    }
    /* the other obstacle kinds (_ls_ik_cuda_kernel.cu:195-224) */
    const int counts[3] = {p->mc, p->mb, p->mh}, width[3] = {7, 15, 6};
    const float* rows[3] = {p->wcap, p->wbox, p->whs};
    for (int k = 0; k < 3; ++k)
      for (int m = 0; m < counts[k]; ++m) {
        double o[15];
        for (int c = 0; c < width[k]; ++c) o[c] = rows[k][width[k] * m + c];
        const double d = obstacle_dist(w, p->rs[4 * i + 3], k + 1, o);
        if (d < p->margin) pen += (d - p->margin) * (d - p->margin);
      }
  }
  return p->wc * pen;
}

static double weighted_err(const ikp_t* p, const double* q, const double* W, se3* Tw, double* r) {
  fk_one(p->t, q, Tw);
  double e = 0;
  for (int k = 0; k < p->n_ee; ++k) {
    residual_ee(&Tw[p->target_jnts[k]], p->tgt + 7 * k, r + 6 * k);
    for (int c = 0; c < 6; ++c) e += (r[6 * k + c] * W[c]) * (r[6 * k + c] * W[c]);
  }
  return e + penalty(p, Tw);
}

/* Geometric Jacobian rows (6 n_ee) x A for the poses in Tw (_ik_cuda_helpers.cuh:501-566). */
static void jacobian(const ikp_t* p, const se3* Tw, double* Jm) {
  const tables_t* t = p->t;
  const int A = t->A;
  memset(Jm, 0, sizeof(double) * 6 * p->n_ee * A);
  for (int j = 0; j < t->J; ++j) {
    const int a1 = t->act_idx[j], a2 = t->mimic_act_idx[j];
    if (a1 < 0 && a2 < 0) continue;
    const float* tw = t->twists + 6 * j;
    const double ang = (double)tw[3] * tw[3] + (double)tw[4] * tw[4] + (double)tw[5] * tw[5];
    const double lin = (double)tw[0] * tw[0] + (double)tw[1] * tw[1] + (double)tw[2] * tw[2];
    int rev;
    double ax[3], z[3];
    if (ang > 1e-6) {
      rev = 1;
      for (int c = 0; c < 3; ++c) ax[c] = tw[3 + c] / sqrt(ang);
    } else if (lin > 1e-6) {
      rev = 0;
      for (int c = 0; c < 3; ++c) ax[c] = tw[c] / sqrt(lin);
    } else {
      continue;
    }
    qrot(&Tw[j], ax, z);
    for (int e = 0; e < p->n_ee; ++e) {
      if (!p->anc[e * t->J + j]) continue;
      const se3* Te = &Tw[p->target_jnts[e]];
      double g[6];
      if (rev) {
        const double arm[3] = {Te->tx - Tw[j].tx, Te->ty - Tw[j].ty, Te->tz - Tw[j].tz};
        g[0] = z[1] * arm[2] - z[2] * arm[1];
        g[1] = z[2] * arm[0] - z[0] * arm[2];
        g[2] = z[0] * arm[1] - z[1] * arm[0];
        g[3] = z[0]; g[4] = z[1]; g[5] = z[2];
      } else {
        g[0] = z[0]; g[1] = z[1]; g[2] = z[2]; g[3] = g[4] = g[5] = 0;
      }
      const double s = t->mimic_mul[j];
      for (int c = 0; c < 6; ++c) {
        if (a1 >= 0) Jm[(6 * e + c) * A + a1] += s * g[c];
        if (a2 >= 0) Jm[(6 * e + c) * A + a2] += s * g[c];
      }
    }
  }
}

/* Cholesky solve of A x = b in place (b <- x); zero step on a non-positive pivot. */
static int chol_solve(double* A, double* b, int n) {
  for (int k = 0; k < n; ++k) {
    double s = A[k * n + k];
    for (int q = 0; q < k; ++q) s -= A[k * n + q] * A[k * n + q];
    if (s <= 0.0) {
      for (int i = 0; i < n; ++i) b[i] = 0.0;
      return 0;
    }
    const double l = sqrt(s);
    A[k * n + k] = l;
    for (int i = k + 1; i < n; ++i) {
      double v = A[i * n + k];
      for (int q = 0; q < k; ++q) v -= A[i * n + q] * A[k * n + q];
      A[i * n + k] = v / l;
    }
    for (int j = k + 1; j < n; ++j) A[k * n + j] = 0.0;
  }
  double y[OR_MAXA];
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int q = 0; q < i; ++q) s -= A[i * n + q] * y[q];
    y[i] = s / A[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int q = i + 1; q < n; ++q) s -= A[q * n + i] * b[q];
    b[i] = s / A[i * n + i];
  }
  return 1;
}

static void solve_seed(const ikp_t* p, const double* seed, const double* lo, const double* hi,
                       const int32_t* fixed, int max_iter, const double* W, double lam0,
                       double eps_pos, double eps_ori, double* out_q, double* out_err) {
  const int A = p->t->A, R = 6 * p->n_ee;
  double q[OR_MAXA] = {0}, best_q[OR_MAXA] = {0}, r[6 * OR_MAXEE], Jm[6 * OR_MAXEE * OR_MAXA];
  se3 Tw[OR_MAXJ];
  for (int a = 0; a < A; ++a) q[a] = best_q[a] = seed[a];
  double best_err = weighted_err(p, q, W, Tw, r);
  double lam = lam0;
  static const double alphas[5] = {1.0, 0.5, 0.25, 0.1, 0.025};
  for (int it = 0; it < max_iter; ++it) {
    double cur = weighted_err(p, q, W, Tw, r);
    int conv = 1;
    double max_p = 0, max_o = 0;
    for (int e = 0; e < p->n_ee; ++e) {
      const double np = sqrt(r[6 * e] * r[6 * e] + r[6 * e + 1] * r[6 * e + 1] + r[6 * e + 2] * r[6 * e + 2]);
      const double no = sqrt(r[6 * e + 3] * r[6 * e + 3] + r[6 * e + 4] * r[6 * e + 4] + r[6 * e + 5] * r[6 * e + 5]);
      if (np >= eps_pos || no >= eps_ori) conv = 0;
      if (np > max_p) max_p = np;
      if (no > max_o) max_o = no;
    }
    if (conv) break;
    jacobian(p, Tw, Jm);
    double fw[6 * OR_MAXEE];
    for (int k = 0; k < R; ++k) {
      fw[k] = r[k] * W[k % 6];
      for (int a = 0; a < A; ++a) Jm[k * A + a] *= W[k % 6];
    }
    double cs[OR_MAXA];
    for (int a = 0; a < A; ++a) {
      double s2 = 0;
      for (int k = 0; k < R; ++k) s2 += Jm[k * A + a] * Jm[k * A + a];
      cs[a] = sqrt(s2) + 1e-8;
    }
    for (int k = 0; k < R; ++k)
      for (int a = 0; a < A; ++a) Jm[k * A + a] /= cs[a];
    double M[OR_MAXA * OR_MAXA], rhs[OR_MAXA];
    for (int i = 0; i < A; ++i) {
      for (int j = 0; j < A; ++j) {
        double acc = 0;
        for (int k = 0; k < R; ++k) acc += Jm[k * A + i] * Jm[k * A + j];
        M[i * A + j] = acc;
      }
      double rb = 0;
      for (int k = 0; k < R; ++k) rb += Jm[k * A + i] * fw[k];
      rhs[i] = -rb;
      M[i * A + i] += lam;
    }
    for (int a = 0; a < A; ++a) {
      if (!fixed[a]) continue;
      for (int j = 0; j < A; ++j) M[a * A + j] = M[j * A + a] = 0.0;
      M[a * A + a] = 1.0;
      rhs[a] = 0.0;
    }
    chol_solve(M, rhs, A);
    double delta[OR_MAXA], dn = 0;
    for (int a = 0; a < A; ++a) {
      delta[a] = rhs[a] / cs[a];
      dn += delta[a] * delta[a];
    }
    double Rt;
    if (max_p > 1e-2 || max_o > 0.6) Rt = 0.38;
    else if (max_p > 1e-3 || max_o > 0.25) Rt = 0.22;
    else if (max_p > 2e-4 || max_o > 0.08) Rt = 0.12;
    else Rt = 0.05;
    dn = sqrt(dn);
    if (dn > Rt) {
      const double s = Rt / (dn + 1e-18);
      for (int a = 0; a < A; ++a) delta[a] *= s;
    }
    double best_a_err = 1e30;
    int best_a = 0;
    double qt[OR_MAXA], rt[6 * OR_MAXEE];
    for (int ai = 0; ai < 5; ++ai) {
      for (int a = 0; a < A; ++a) {
        const double v = q[a] + alphas[ai] * delta[a];
        qt[a] = v < lo[a] ? lo[a] : (v > hi[a] ? hi[a] : v);
      }
      const double et = weighted_err(p, qt, W, Tw, rt);
      if (et < best_a_err) { best_a_err = et; best_a = ai; }
    }
    for (int a = 0; a < A; ++a) {
      const double v = q[a] + alphas[best_a] * delta[a];
      qt[a] = v < lo[a] ? lo[a] : (v > hi[a] ? hi[a] : v);
    }
    if (best_a_err < cur * (1.0 - 1e-4)) {
      for (int a = 0; a < A; ++a) q[a] = qt[a];
      lam = lam * 0.5 > 1e-10 ? lam * 0.5 : 1e-10;
    } else {
      lam = lam * 3.0 < 1e6 ? lam * 3.0 : 1e6;
    }
    if (best_a_err < best_err) {
      best_err = best_a_err;
      for (int a = 0; a < A; ++a) best_q[a] = qt[a];
    }
  }
  for (int a = 0; a < A; ++a) out_q[a] = best_q[a];
  *out_err = best_err;
}

/* world_other = {capsules (mc, 7), boxes (mb, 15), half-spaces (mh, 6)} as
 * float arrays (NULL when the count is 0), counts = {mc, mb, mh}. */
void or_ls_ik(const double* seeds, int P, int S, int A, const float* twists, const float* parent_tf,
              const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
              const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv, int J,
              const int32_t* target_jnts, const int32_t* ancestor_masks, int n_ee, const double* target_T,
              const float* robot_spheres, const int32_t* sphere_joint, int K, const float* world, int M,
              const double* lower, const double* upper, const int32_t* fixed_mask, int max_iter,
              double pos_weight, double ori_weight, double lambda_init, double eps_pos, double eps_ori,
              int enable_collision, double collision_weight, double collision_margin, double* out_cfgs,
              double* out_err, int threads, const float* w_caps, const float* w_boxes, const float* w_hs,
              const int32_t* other_counts) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  const double W[6] = {pos_weight, pos_weight, pos_weight, ori_weight, ori_weight, ori_weight};
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (int pi = 0; pi < P; ++pi)
    for (int si = 0; si < S; ++si) {
      ikp_t p = {&t, n_ee, target_jnts, ancestor_masks, target_T + (int64_t)pi * n_ee * 7, robot_spheres,
                 sphere_joint, K, world, M, enable_collision, collision_weight, collision_margin,
                 w_caps, w_boxes, w_hs, other_counts ? other_counts[0] : 0, other_counts ? other_counts[1] : 0,
                 other_counts ? other_counts[2] : 0};
      const int64_t g = (int64_t)pi * S + si;
      solve_seed(&p, seeds + g * A, lower, upper, fixed_mask, max_iter, W, lambda_init, eps_pos, eps_ori,
                 out_cfgs + g * A, out_err + g);
    }
}

/* Residual (6 n_ee) and weighted-free geometric Jacobian (6 n_ee x A) at q. */
void or_res_jac(const double* q, int A, const float* twists, const float* parent_tf, const int32_t* parent_idx,
                const int32_t* act_idx, const float* mimic_mul, const float* mimic_off,
                const int32_t* mimic_act_idx, const int32_t* topo_inv, int J, const int32_t* target_jnts,
                const int32_t* ancestor_masks, int n_ee, const double* target_T, double* out_r, double* out_J) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  ikp_t p = {&t, n_ee, target_jnts, ancestor_masks, target_T, NULL, NULL, 0, NULL, 0, 0, 0.0, 0.0,
             NULL, NULL, NULL, 0, 0, 0};
  se3 Tw[OR_MAXJ];
  fk_one(&t, q, Tw);
  for (int e = 0; e < n_ee; ++e) residual_ee(&Tw[target_jnts[e]], target_T + 7 * e, out_r + 6 * e);
  jacobian(&p, Tw, out_J);
}

/* Damped Gauss-Newton direction of one LM step before the trust region:
 * delta = S^-1 (Js^T Js + lam I)^-1 (-Js^T W r), Js = W J S^-1 (column scaling S). */
int or_lm_direction(const double* r, const double* Jm, int R, int A, const double* W, double lam,
                    const int32_t* fixed, double* delta) {
  double Jw[6 * OR_MAXEE * OR_MAXA], fw[6 * OR_MAXEE], cs[OR_MAXA], M[OR_MAXA * OR_MAXA], rhs[OR_MAXA];
  for (int k = 0; k < R; ++k) {
    fw[k] = r[k] * W[k % 6];
    for (int a = 0; a < A; ++a) Jw[k * A + a] = Jm[k * A + a] * W[k % 6];
  }
  for (int a = 0; a < A; ++a) {
    double s2 = 0;
    for (int k = 0; k < R; ++k) s2 += Jw[k * A + a] * Jw[k * A + a];
    cs[a] = sqrt(s2) + 1e-8;
  }
  for (int k = 0; k < R; ++k)
    for (int a = 0; a < A; ++a) Jw[k * A + a] /= cs[a];
  for (int i = 0; i < A; ++i) {
    for (int j = 0; j < A; ++j) {
      double acc = 0;
      for (int k = 0; k < R; ++k) acc += Jw[k * A + i] * Jw[k * A + j];
      M[i * A + j] = acc;
    }
    double rb = 0;
    for (int k = 0; k < R; ++k) rb += Jw[k * A + i] * fw[k];
    rhs[i] = -rb;
    M[i * A + i] += lam;
  }
  for (int a = 0; a < A; ++a) {
    if (!fixed[a]) continue;
    for (int j = 0; j < A; ++j) M[a * A + j] = M[j * A + a] = 0.0;
    M[a * A + a] = 1.0;
    rhs[a] = 0.0;
  }
  const int ok = chol_solve(M, rhs, A);
  for (int a = 0; a < A; ++a) delta[a] = rhs[a] / cs[a];
  return ok;
}

/* Pose + hinge cost per configuration (FP64), the function whose gradient the
 * adjoint kernel computes:
 *   sum (W r)^2 over the EE residual of link ee_link (-1: none)
 * + sum over real robot spheres x world spheres of max(0, margin - d)^2.
 * Hinge^2 as in src/pyroffi/optimization_engines/_chomp_optimization.py:140-154;
 * residual as in _ik_cuda_helpers.cuh:471-494. */
void or_pose_hinge_cost(const double* cfg, int64_t B, int A, const float* twists, const float* parent_tf,
                        const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
                        const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv, int J,
                        const int32_t* link_parent_joint, int L, const double* spheres, int S,
                        const double* world, int M, double margin, int ee_link, const double* targets,
                        double pos_weight, double ori_weight, double* out, int threads) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < B; ++b) {
    se3 Tw[OR_MAXJ];
    fk_one(&t, cfg + b * A, Tw);
    double cost = 0.0;
    for (int l = 0; l < L; ++l) {
      const int j = link_parent_joint[l];
      se3 T = (j < 0) ? (se3){1, 0, 0, 0, 0, 0, 0} : Tw[j];
      for (int s = 0; s < S; ++s) {
        const double* sp = spheres + (l * S + s) * 4;
        if (sp[3] < 0) continue;
        double c[3];
        qrot(&T, sp, c);
        c[0] += T.tx; c[1] += T.ty; c[2] += T.tz;
        for (int m = 0; m < M; ++m) {
          const double d = sdist(c, sp[3], world + 4 * m, world[4 * m + 3], 0);
          const double h = margin - d;
          if (h > 0) cost += h * h;
        }
      }
    }
    if (ee_link >= 0) {
      const int j = link_parent_joint[ee_link];
      se3 T = (j < 0) ? (se3){1, 0, 0, 0, 0, 0, 0} : Tw[j];
      double r[6];
      residual_ee(&T, targets + 7 * b, r);
      for (int c = 0; c < 3; ++c) cost += pos_weight * pos_weight * r[c] * r[c];
      for (int c = 3; c < 6; ++c) cost += ori_weight * ori_weight * r[c] * r[c];
    }
    out[b] = cost;
  }
}

/* ----------------------------------------------------------- trajectories */
/*
 * Trajectory cost of the L-BFGS / CHOMP trajectory optimisers:
 *   total = w_smooth * smooth + w_limits * limits + w_coll * coll
 *   smooth  _sco_optimization.py:170-189: w_vel sum|vel|^2 + w_acc sum|acc|^2 +
 *           w_jerk sum|jerk|^2, vel_t = x_{t+1} - x_t,
 *           acc_t = (-x_t + 16 x_{t+1} - 30 x_{t+2} + 16 x_{t+3} - x_{t+4}) / 12,
 *           jerk_t = acc_{t+1} - acc_t (CHOMP's variant, _chomp_optimization.py:
 *           93-111, is w_vel = 0)
 *   limits  _sco_optimization.py:192-200: sum (max(0, x-u) + max(0, l-x))^2
 *   coll    sum over waypoints of sum f(d) over _collision_distances_all
 *           (_sco_optimization.py:203-222): the JAX self distances (diagonal
 *           sphere pairs, +1e-6 bias) of every active pair and the JAX world
 *           distances d[n, m] = min over link n's spheres (+1e-6 bias, padding
 *           r = -1e9 included) of every link n and world sphere m;
 *           f = max(0, m - d)^2 (mode 0, CHOMP, _chomp_optimization.py:140-154)
 *           or -min(colldist_from_sdf(d, m), 0) (mode 1, L-BFGS,
 *           _lbfgs_trajopt_optimization.py:108-111; collision/_collision.py:136-167)
 * The gradient is analytic: per sphere centre c, dcost/dc, pulled back to the
 * joints with dc/dq_j = mul_j ((R_j w_j) x (c - p_j) + R_j v_j) for every
 * moving ancestor joint j (tests pin it with central differences).
 */
static double act_fn(double d, double m, int mode, double* dfd) {
  if (mode == 0) {
    const double h = m - d;
    if (h > 0) { *dfd = -2.0 * h; return h * h; }
    *dfd = 0.0;
    return 0.0;
  }
  const double dm = d < m ? d : m;
  if (dm < 0) { *dfd = -1.0; return 0.5 * m - dm; }
  if (d < m) { *dfd = (d - m) / (m + 1e-6); return 0.5 / (m + 1e-6) * (d - m) * (d - m); }
  *dfd = 0.0;
  return 0.0;
}

typedef struct {
  const tables_t* t;
  const int32_t* lpj; /* link parent joint (L) */
  int L;
  const double* sph;  /* (L, S, 4) link frame, r < 0 padding */
  int S;
  const int32_t *pi, *pj;
  int P;
  const double* world;
  int M;
  double margin;
  int mode;
} traj_model_t;

/* Accumulate g . dc/dq into grad for a point c on link l. */
static void pull_back(const traj_model_t* md, const se3* Tw, int l, const double* c, const double* g,
                      double* grad) {
  const tables_t* t = md->t;
  int j = md->lpj[l];
  while (j >= 0) {
    const int mm = t->mimic_act_idx[j];
    const int src = (mm != -1) ? mm : t->act_idx[j];
    if (src >= 0) {
      double w[3], v[3];
      const double ang[3] = {t->twists[6 * j + 3], t->twists[6 * j + 4], t->twists[6 * j + 5]};
      const double lin[3] = {t->twists[6 * j], t->twists[6 * j + 1], t->twists[6 * j + 2]};
      qrot(&Tw[j], ang, w);
      qrot(&Tw[j], lin, v);
      const double r[3] = {c[0] - Tw[j].tx, c[1] - Tw[j].ty, c[2] - Tw[j].tz};
      const double dc[3] = {w[1] * r[2] - w[2] * r[1] + v[0], w[2] * r[0] - w[0] * r[2] + v[1],
                            w[0] * r[1] - w[1] * r[0] + v[2]};
      grad[src] += (double)t->mimic_mul[j] * (g[0] * dc[0] + g[1] * dc[1] + g[2] * dc[2]);
    }
    j = t->parent_idx[j];
  }
}

/* Collision cost of one waypoint; grad (A) is accumulated when non-NULL. */
static double wp_coll(const traj_model_t* md, const double* q, double* grad) {
  const int L = md->L, S = md->S;
  se3 Tw[OR_MAXJ];
  fk_one(md->t, q, Tw);
  double* cw = (double*)malloc(sizeof(double) * 3 * L * S);
  for (int l = 0; l < L; ++l) {
    const int j = md->lpj[l];
    se3 T = (j < 0) ? (se3){1, 0, 0, 0, 0, 0, 0} : Tw[j];
    for (int s = 0; s < S; ++s) {
      double* o = cw + 3 * (l * S + s);
      qrot(&T, md->sph + (l * S + s) * 4, o);
      o[0] += T.tx; o[1] += T.ty; o[2] += T.tz;
    }
  }
  double cost = 0.0;
  const double e = 1e-6;
  /* world: per (link, obstacle), the link's closest sphere */
  for (int l = 0; l < L; ++l) {
    for (int m = 0; m < md->M; ++m) {
      const double* o = md->world + 4 * m;
      double best = INFINITY;
      int arg = -1;
      for (int s = 0; s < S; ++s) {
        const double d = sdist(cw + 3 * (l * S + s), md->sph[(l * S + s) * 4 + 3], o, o[3], 1);
        if (d < best) { best = d; arg = s; }
      }
      if (arg < 0) continue;
      double dfd;
      cost += act_fn(best, md->margin, md->mode, &dfd);
      if (grad && dfd != 0.0) {
        const double* c = cw + 3 * (l * S + arg);
        const double x[3] = {o[0] - c[0] + e, o[1] - c[1] + e, o[2] - c[2] + e};
        const double n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        const double g[3] = {-dfd * x[0] / n, -dfd * x[1] / n, -dfd * x[2] / n};
        pull_back(md, Tw, l, c, g, grad);
      }
    }
  }
  /* self: active pairs, same-slot sphere pairs */
  for (int p = 0; p < md->P; ++p) {
    const int li = md->pi[p], lj = md->pj[p];
    double best = INFINITY;
    int arg = -1;
    for (int s = 0; s < S; ++s) {
      const double d = sdist(cw + 3 * (li * S + s), md->sph[(li * S + s) * 4 + 3], cw + 3 * (lj * S + s),
                             md->sph[(lj * S + s) * 4 + 3], 1);
      if (d < best) { best = d; arg = s; }
    }
    if (arg < 0) continue;
    double dfd;
    cost += act_fn(best, md->margin, md->mode, &dfd);
    if (grad && dfd != 0.0) {
      const double* ci = cw + 3 * (li * S + arg);
      const double* cj = cw + 3 * (lj * S + arg);
      const double x[3] = {cj[0] - ci[0] + e, cj[1] - ci[1] + e, cj[2] - ci[2] + e};
      const double n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
      const double gj[3] = {dfd * x[0] / n, dfd * x[1] / n, dfd * x[2] / n};
      const double gi[3] = {-gj[0], -gj[1], -gj[2]};
      pull_back(md, Tw, lj, cj, gj, grad);
      pull_back(md, Tw, li, ci, gi, grad);
    }
  }
  free(cw);
  return cost;
}

/* One trajectory x (T, A): comps = {smooth, limits, coll} (unweighted by
 * w_smooth / w_limits / w_coll); grad (T, A) = gradient of the weighted total
 * when non-NULL.  w = {w_smooth, w_vel, w_acc, w_jerk, w_limits, w_coll}. */
static void traj_cost_one(const traj_model_t* md, const double* x, int T, const double* lower,
                          const double* upper, const double* w, double* comps, double* grad) {
  const int A = md->t->A;
  double sm = 0.0, lim = 0.0, col = 0.0;
  if (grad) memset(grad, 0, sizeof(double) * T * A);
  const double c5[5] = {-1.0, 16.0, -30.0, 16.0, -1.0};
  for (int a = 0; a < A; ++a) {
    for (int t = 0; t + 1 < T; ++t) {
      const double v = x[(t + 1) * A + a] - x[t * A + a];
      sm += w[1] * v * v;
      if (grad) {
        grad[(t + 1) * A + a] += w[0] * w[1] * 2 * v;
        grad[t * A + a] -= w[0] * w[1] * 2 * v;
      }
    }
    const int na = T - 4;
    double* acc = (double*)malloc(sizeof(double) * (na > 0 ? na : 1));
    for (int t = 0; t < na; ++t) {
      double s = 0;
      for (int o = 0; o < 5; ++o) s += c5[o] * x[(t + o) * A + a];
      acc[t] = s / 12.0;
      sm += w[2] * acc[t] * acc[t];
    }
    for (int t = 0; t + 1 < na; ++t) {
      const double jk = acc[t + 1] - acc[t];
      sm += w[3] * jk * jk;
    }
    if (grad) {
      for (int t = 0; t < na; ++t) {
        double ga = 2 * w[2] * acc[t];
        if (t + 1 < na) ga -= 2 * w[3] * (acc[t + 1] - acc[t]);
        if (t >= 1) ga += 2 * w[3] * (acc[t] - acc[t - 1]);
        for (int o = 0; o < 5; ++o) grad[(t + o) * A + a] += w[0] * ga * c5[o] / 12.0;
      }
    }
    free(acc);
    for (int t = 0; t < T; ++t) {
      const double xv = x[t * A + a];
      const double vu = xv > upper[a] ? xv - upper[a] : 0.0;
      const double vl = lower[a] > xv ? lower[a] - xv : 0.0;
      const double v = vu + vl;
      lim += v * v;
      if (grad) grad[t * A + a] += w[4] * 2 * v * ((xv > upper[a] ? 1.0 : 0.0) - (lower[a] > xv ? 1.0 : 0.0));
    }
  }
  double gq[OR_MAXA];
  for (int t = 0; t < T; ++t) {
    if (grad) for (int a = 0; a < A; ++a) gq[a] = 0.0;
    col += wp_coll(md, x + t * A, grad ? gq : NULL);
    if (grad) for (int a = 0; a < A; ++a) grad[t * A + a] += w[5] * gq[a];
  }
  comps[0] = sm; comps[1] = lim; comps[2] = col;
}

void or_traj_cost(const float* twists, const float* parent_tf, const int32_t* parent_idx, const int32_t* act_idx,
                  const float* mimic_mul, const float* mimic_off, const int32_t* mimic_act_idx,
                  const int32_t* topo_inv, int J, int A, const int32_t* link_parent_joint, int L,
                  const double* spheres, int S, const int32_t* pair_i, const int32_t* pair_j, int P,
                  const double* world, int M, double margin, int mode, const double* x, const int64_t* start,
                  int64_t B, const double* lower, const double* upper, const double* w, double* comps,
                  double* grad, int threads) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  traj_model_t md = {&t, link_parent_joint, L, spheres, S, pair_i, pair_j, P, world, M, margin, mode};
  set_threads(threads);
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < B; ++b) {
    const int T = (int)(start[b + 1] - start[b]);
    traj_cost_one(&md, x + start[b] * A, T, lower, upper, w, comps + 3 * b, grad ? grad + start[b] * A : NULL);
  }
}

/* Nocedal two-loop recursion with masked history (_sco_optimization.py:119-167). */
static void two_loop(const double* g, const double* s_buf, const double* y_buf, const double* rho, int m_used,
                     int newest, int m, int n, double* out) {
  double* q = (double*)malloc(sizeof(double) * n);
  double alpha[16];
  memcpy(q, g, sizeof(double) * n);
  for (int i = 0; i < m; ++i) alpha[i] = 0.0;
  for (int i = 0; i < m; ++i) {
    const int idx = (newest - i + m) % m;
    if (i >= m_used) continue;
    double sq = 0;
    for (int k = 0; k < n; ++k) sq += s_buf[idx * n + k] * q[k];
    alpha[idx] = rho[idx] * sq;
    for (int k = 0; k < n; ++k) q[k] -= alpha[idx] * y_buf[idx * n + k];
  }
  double sy = 0, yy = 0;
  for (int k = 0; k < n; ++k) {
    sy += s_buf[newest * n + k] * y_buf[newest * n + k];
    yy += y_buf[newest * n + k] * y_buf[newest * n + k];
  }
  const double gamma = sy / (yy + 1e-18);
  for (int k = 0; k < n; ++k) out[k] = gamma * q[k];
  for (int st = 0; st < m; ++st) {
    const int idx = (newest - m_used + 1 + st + 2 * m) % m;
    if (st >= m_used) continue;
    double yr = 0;
    for (int k = 0; k < n; ++k) yr += y_buf[idx * n + k] * out[k];
    const double beta = rho[idx] * yr;
    for (int k = 0; k < n; ++k) out[k] += s_buf[idx * n + k] * (alpha[idx] - beta);
  }
  for (int k = 0; k < n; ++k) out[k] = -out[k];
  free(q);
}

void or_two_loop(const double* g, const double* s_buf, const double* y_buf, const double* rho, int m_used,
                 int newest, int m, int n, double* out) {
  two_loop(g, s_buf, y_buf, rho, m_used, newest, m, n, out);
}

/* The direct L-BFGS solve of _lbfgs_trajopt_optimization.py:76-205 for each
 * trajectory.  cfg = {n_iters, m, escalation_interval} (ints) and
 * fcfg = {w_smooth, w_vel, w_acc, w_jerk, w_limits, w_collision, w_collision_max,
 * penalty_scale}; endpoints stay where x_init has them.  out_best (W, A),
 * out_best_cost (B) = the tracked best cost.  When non-NULL, out_last (W, A)
 * receives the last iterate and out_trace (B, n_iters, 3) per iteration the
 * cost of the current iterate at the current collision weight, the index of
 * the accepted step in the step list, and the accepted point's cost at the
 * maximum collision weight (the value the best tracking compares). */
void or_traj_lbfgs(const float* twists, const float* parent_tf, const int32_t* parent_idx, const int32_t* act_idx,
                   const float* mimic_mul, const float* mimic_off, const int32_t* mimic_act_idx,
                   const int32_t* topo_inv, int J, int A, const int32_t* link_parent_joint, int L,
                   const double* spheres, int S, const int32_t* pair_i, const int32_t* pair_j, int P,
                   const double* world, int M, double margin, int mode, const double* x_init,
                   const int64_t* start, int64_t B, const double* lower, const double* upper, const int32_t* cfg,
                   const double* fcfg, double* out_best, double* out_best_cost, double* out_last,
                   double* out_trace, int threads) {
  static const double alphas[5] = {1.0, 0.5, 0.25, 0.1, 0.025};
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  traj_model_t md = {&t, link_parent_joint, L, spheres, S, pair_i, pair_j, P, world, M, margin, mode};
  const int n_iters = cfg[0], m = cfg[1], esc = cfg[2];
  set_threads(threads);
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < B; ++b) {
    const int T = (int)(start[b + 1] - start[b]);
    const int n = T * A;
    double* x = (double*)malloc(sizeof(double) * n);
    double* xp = (double*)malloc(sizeof(double) * n);
    double* gp = (double*)malloc(sizeof(double) * n);
    double* g = (double*)malloc(sizeof(double) * n);
    double* dir = (double*)malloc(sizeof(double) * n);
    double* xt = (double*)malloc(sizeof(double) * n);
    double* best = (double*)malloc(sizeof(double) * n);
    double* sb = (double*)calloc((size_t)m * n, sizeof(double));
    double* yb = (double*)calloc((size_t)m * n, sizeof(double));
    double rho[16] = {0};
    memcpy(x, x_init + start[b] * A, sizeof(double) * n);
    double w[6] = {fcfg[0], fcfg[1], fcfg[2], fcfg[3], fcfg[4], fcfg[5]};
    double comps[3];
    traj_cost_one(&md, x, T, lower, upper, w, comps, g);
    for (int a = 0; a < A; ++a) { g[a] = 0; g[n - A + a] = 0; }
    double best_cost = fcfg[0] * comps[0] + fcfg[4] * comps[1] + w[5] * comps[2];
    memcpy(best, x, sizeof(double) * n);
    memcpy(xp, x, sizeof(double) * n);
    memcpy(gp, g, sizeof(double) * n);
    int m_used = 0, newest = 0;
    for (int k = 0; k < n_iters; ++k) {
      if (k > 0 && k % esc == 0) w[5] = fmin(w[5] * fcfg[7], fcfg[6]);
      traj_cost_one(&md, x, T, lower, upper, w, comps, g);
      for (int a = 0; a < A; ++a) { g[a] = 0; g[n - A + a] = 0; }
      const double cost = w[0] * comps[0] + w[4] * comps[1] + w[5] * comps[2];
      double sy = 0, yy = 0;
      for (int i = 0; i < n; ++i) {
        const double s = x[i] - xp[i], y = g[i] - gp[i];
        sy += s * y;
        yy += y * y;
      }
      if (sy > 1e-10 * yy + 1e-30 && k > 0) {
        newest = (newest + 1) % m;
        for (int i = 0; i < n; ++i) { sb[newest * n + i] = x[i] - xp[i]; yb[newest * n + i] = g[i] - gp[i]; }
        rho[newest] = 1.0 / (sy + 1e-30);
        if (m_used < m) ++m_used;
      }
      if (m_used > 0) {
        two_loop(g, sb, yb, rho, m_used, newest, m, n, dir);
      } else {
        double gn = 0;
        for (int i = 0; i < n; ++i) gn += g[i] * g[i];
        gn = sqrt(gn) + 1e-18;
        for (int i = 0; i < n; ++i) dir[i] = -g[i] / gn;
      }
      for (int a = 0; a < A; ++a) { dir[a] = 0; dir[n - A + a] = 0; }
      const double thresh = cost * (1.0 - 1e-4);
      double trial[5], tcomp[5][3];
      for (int i = 0; i < 5; ++i) {
        for (int k2 = 0; k2 < n; ++k2) xt[k2] = x[k2] + alphas[i] * dir[k2];
        traj_cost_one(&md, xt, T, lower, upper, w, tcomp[i], NULL);
        trial[i] = w[0] * tcomp[i][0] + w[4] * tcomp[i][1] + w[5] * tcomp[i][2];
      }
      int pick = -1;
      for (int i = 0; i < 5; ++i)
        if (trial[i] < thresh) { pick = i; break; }
      if (pick < 0) {
        pick = 0;
        for (int i = 1; i < 5; ++i)
          if (trial[i] < trial[pick]) pick = i;
      }
      memcpy(xp, x, sizeof(double) * n);
      memcpy(gp, g, sizeof(double) * n);
      for (int i = 0; i < n; ++i) x[i] += alphas[pick] * dir[i];
      const double eval = w[0] * tcomp[pick][0] + w[4] * tcomp[pick][1] + fcfg[6] * tcomp[pick][2];
      if (out_trace) {
        double* tr = out_trace + ((size_t)b * n_iters + k) * 3;
        tr[0] = cost;
        tr[1] = pick;
        tr[2] = eval;
      }
      if (eval < best_cost) {
        best_cost = eval;
        memcpy(best, x, sizeof(double) * n);
      }
    }
    memcpy(out_best + start[b] * A, best, sizeof(double) * n);
    if (out_last) memcpy(out_last + start[b] * A, x, sizeof(double) * n);
    out_best_cost[b] = best_cost;
    free(x); free(xp); free(gp); free(g); free(dir); free(xt); free(best); free(sb); free(yb);
  }
}
/* CHOMP's interior smoothness precision operator on one joint's timeline
 * (_chomp_optimization.py:156-166): y = 6x - 4(x[i-1] + x[i+1]) + x[i-2] + x[i+2]
 * + reg x, terms outside [0, n) dropped. */
static void chomp_metric(const double* x, int n, double reg, double* y) {
  for (int i = 0; i < n; ++i) {
    double v = (6.0 + reg) * x[i];
    if (i >= 1) v -= 4.0 * x[i - 1];
    if (i + 1 < n) v -= 4.0 * x[i + 1];
    if (i >= 2) v += x[i - 2];
    if (i + 2 < n) v += x[i + 2];
    y[i] = v;
  }
}

/* Truncated CG for M x = b (_chomp_optimization.py:169-205): zero start,
 * n_iters steps, a step taken only while the squared residual stays above
 * (tol |b| + 1e-12)^2. */
static void chomp_cg(const double* b, int n, double reg, int n_iters, double tol, double* x) {
  double* r = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* ap = (double*)malloc(sizeof(double) * n);
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; p[i] = b[i]; rr += b[i] * b[i]; }
  bb = rr;
  const double tb = tol * sqrt(bb + 1e-18) + 1e-12;
  const double tol2 = tb * tb;
  int active = 1;
  for (int it = 0; it < n_iters && active; ++it) {
    chomp_metric(p, n, reg, ap);
    double pap = 0.0;
    for (int i = 0; i < n; ++i) pap += p[i] * ap[i];
    const double alpha = rr / (pap + 1e-9);
    double rrn = 0.0;
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * ap[i];
      rrn += r[i] * r[i];
    }
    const double beta = rrn / (rr + 1e-9);
    for (int i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    rr = rrn;
    active = rrn > tol2;
  }
  free(r); free(p); free(ap);
}

/* The CHOMP trajectory optimiser of _chomp_optimization.py:208-348 for each
 * trajectory (endpoints stay where x_init has them; the model's mode is the
 * hinge^2 activation, 0).  icfg = {n_iters, use_covariant_update, n_cg_iters,
 * early_stop_patience}; fcfg = {step_size, w_smooth, w_acc, w_jerk,
 * w_collision, w_collision_max, collision_penalty_scale, w_limits,
 * smoothness_reg, cg_tol, grad_clip_norm, max_delta_per_step,
 * min_cost_improve}.  Per iteration while active: the gradient of the cost at
 * the current collision weight, clipped to norm grad_clip_norm, endpoint rows
 * zeroed; the direction -M^-1 g per joint over the interior (truncated CG) or
 * -g; the six trials x + clip(step_size alpha d, +-max_delta), alpha in
 * {1, .5, .25, .1, .05, 0}, the cheapest (first on ties) kept when it beats the
 * current cost by min_cost_improve; stall count and early stop; collision
 * weight x penalty_scale up to w_collision_max.  out_traj (W, A) = the final
 * iterate, out_cost (B) = its cost at w_collision_max; out_trace (B, n_iters,
 * 3), when non-NULL, = per iteration the current cost, the picked trial and
 * whether it was accepted (-1 in all three once stopped). */
void or_traj_chomp(const float* twists, const float* parent_tf, const int32_t* parent_idx, const int32_t* act_idx,
                   const float* mimic_mul, const float* mimic_off, const int32_t* mimic_act_idx,
                   const int32_t* topo_inv, int J, int A, const int32_t* link_parent_joint, int L,
                   const double* spheres, int S, const int32_t* pair_i, const int32_t* pair_j, int P,
                   const double* world, int M, double margin, int mode, const double* x_init,
                   const int64_t* start, int64_t B, const double* lower, const double* upper, const int32_t* icfg,
                   const double* fcfg, double* out_traj, double* out_cost, double* out_trace, int threads) {
  static const double alphas[6] = {1.0, 0.5, 0.25, 0.1, 0.05, 0.0};
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  traj_model_t md = {&t, link_parent_joint, L, spheres, S, pair_i, pair_j, P, world, M, margin, mode};
  const int n_iters = icfg[0], covariant = icfg[1], n_cg = icfg[2], patience = icfg[3];
  const double step = fcfg[0], w_smooth = fcfg[1], w_acc = fcfg[2], w_jerk = fcfg[3];
  const double w_c0 = fcfg[4], w_cmax = fcfg[5], scale = fcfg[6], w_lim = fcfg[7];
  const double reg = fcfg[8], cg_tol = fcfg[9], clip = fcfg[10], max_delta = fcfg[11], min_imp = fcfg[12];
  set_threads(threads);
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < B; ++b) {
    const int T = (int)(start[b + 1] - start[b]);
    const int n = T * A;
    double* x = (double*)malloc(sizeof(double) * n);
    double* g = (double*)malloc(sizeof(double) * n);
    double* dir = (double*)malloc(sizeof(double) * n);
    double* xt = (double*)malloc(sizeof(double) * n);
    double* xb = (double*)malloc(sizeof(double) * n);
    const int tin = T > 2 ? T - 2 : 0;
    double* col = (double*)malloc(sizeof(double) * (tin > 0 ? tin : 1));
    double* sol = (double*)malloc(sizeof(double) * (tin > 0 ? tin : 1));
    memcpy(x, x_init + start[b] * A, sizeof(double) * n);
    double w_coll = w_c0;
    int stall = 0, active = 1;
    for (int k = 0; k < n_iters; ++k) {
      double* tr = out_trace ? out_trace + ((size_t)b * n_iters + k) * 3 : NULL;
      if (!active) {
        if (tr) tr[0] = tr[1] = tr[2] = -1.0;
        continue;
      }
      double w[6] = {w_smooth, 0.0, w_acc, w_jerk, w_lim, w_coll};
      double comps[3];
      traj_cost_one(&md, x, T, lower, upper, w, comps, g);
      const double cost = w_smooth * comps[0] + w_lim * comps[1] + w_coll * comps[2];
      double gn = 0.0;
      for (int i = 0; i < n; ++i) gn += g[i] * g[i];
      gn = sqrt(gn);
      const double cs = (clip > 0.0 && gn > clip) ? clip / (gn + 1e-12) : 1.0;
      for (int i = 0; i < n; ++i) g[i] *= cs;
      for (int a = 0; a < A; ++a) { g[a] = 0.0; g[n - A + a] = 0.0; }
      if (covariant && T > 2) {
        for (int i = 0; i < n; ++i) dir[i] = 0.0;
        for (int a = 0; a < A; ++a) {
          for (int i = 0; i < tin; ++i) col[i] = g[(i + 1) * A + a];
          chomp_cg(col, tin, reg, n_cg, cg_tol, sol);
          for (int i = 0; i < tin; ++i) dir[(i + 1) * A + a] = -sol[i];
        }
      } else {
        for (int i = 0; i < n; ++i) dir[i] = -g[i];
      }
      int pick = -1;
      double best = INFINITY;
      for (int j = 0; j < 6; ++j) {
        const double al = step * alphas[j];
        for (int i = 0; i < n; ++i) {
          double d = al * dir[i];
          d = d < -max_delta ? -max_delta : (d > max_delta ? max_delta : d);
          xt[i] = x[i] + d;
        }
        double tc[3];
        traj_cost_one(&md, xt, T, lower, upper, w, tc, NULL);
        const double c = w_smooth * tc[0] + w_lim * tc[1] + w_coll * tc[2];
        if (c < best) { best = c; pick = j; memcpy(xb, xt, sizeof(double) * n); }
      }
      const int improved = best < cost - min_imp;
      if (improved) memcpy(x, xb, sizeof(double) * n);
      stall = improved ? 0 : stall + 1;
      active = stall < patience;
      w_coll = fmin(w_coll * scale, w_cmax);
      if (tr) { tr[0] = cost; tr[1] = pick; tr[2] = improved; }
    }
    double w[6] = {w_smooth, 0.0, w_acc, w_jerk, w_lim, w_cmax};
    double comps[3];
    traj_cost_one(&md, x, T, lower, upper, w, comps, NULL);
    out_cost[b] = w_smooth * comps[0] + w_lim * comps[1] + w_cmax * comps[2];
    memcpy(out_traj + start[b] * A, x, sizeof(double) * n);
    free(x); free(g); free(dir); free(xt); free(xb); free(col); free(sol);
  }
}

void or_chomp_cg(const double* b, int n, double reg, int n_iters, double tol, double* x) {
  chomp_cg(b, n, reg, n_iters, tol, x);
}

/* ----------------------------------------------------------------- SCO */

/* SCO's linearisation at one waypoint (_sco_optimization.py:224-288): per
 * collision group the smooth minimum  -tau logsumexp(-d / tau)  of its
 * distances and its gradient (the softmax-weighted distance gradients; the
 * gradient of a distance flows to the closest sphere, the first on ties, as
 * jnp.min's does).  Group 0: every active self pair (same-slot spheres);
 * group 1 + g: every (link, world sphere) pair of world group g, spheres
 * [gstart[g], gstart[g + 1]).  All with the JAX +1e-6 bias; an empty group
 * reads +inf with a zero gradient.  dk (G), Jk (G, A). */
static void sco_linearize_wp(const traj_model_t* md, const double* q, const int32_t* gstart, int n_wg, double tau,
                             double* dk, double* Jk) {
  const int L = md->L, S = md->S, A = md->t->A;
  const int G = 1 + n_wg;
  se3 Tw[OR_MAXJ];
  fk_one(md->t, q, Tw);
  double* cw = (double*)malloc(sizeof(double) * 3 * L * S);
  for (int l = 0; l < L; ++l) {
    const int j = md->lpj[l];
    se3 T = (j < 0) ? (se3){1, 0, 0, 0, 0, 0, 0} : Tw[j];
    for (int sl = 0; sl < S; ++sl) {
      double* o = cw + 3 * (l * S + sl);
      qrot(&T, md->sph + (l * S + sl) * 4, o);
      o[0] += T.tx; o[1] += T.ty; o[2] += T.tz;
    }
  }
  const double e = 1e-6;
  for (int g = 0; g < G; ++g) {
    const int n_terms = (g == 0) ? md->P : L * (gstart[g] - gstart[g - 1]);
    double* J = Jk + g * A;
    for (int a = 0; a < A; ++a) J[a] = 0.0;
    if (n_terms <= 0) { dk[g] = INFINITY; continue; }
    double* d = (double*)malloc(sizeof(double) * n_terms);
    int* arg = (int*)malloc(sizeof(int) * n_terms);
    double dmin = INFINITY;
    for (int i = 0; i < n_terms; ++i) {
      double best = INFINITY;
      int bs = 0;
      if (g == 0) {
        const int li = md->pi[i], lj = md->pj[i];
        for (int sl = 0; sl < S; ++sl) {
          const double v = sdist(cw + 3 * (li * S + sl), md->sph[(li * S + sl) * 4 + 3], cw + 3 * (lj * S + sl),
                                 md->sph[(lj * S + sl) * 4 + 3], 1);
          if (v < best) { best = v; bs = sl; }
        }
      } else {
        const int l = i / (gstart[g] - gstart[g - 1]);
        const int m = gstart[g - 1] + i % (gstart[g] - gstart[g - 1]);
        const double* o = md->world + 4 * m;
        for (int sl = 0; sl < S; ++sl) {
          const double v = sdist(cw + 3 * (l * S + sl), md->sph[(l * S + sl) * 4 + 3], o, o[3], 1);
          if (v < best) { best = v; bs = sl; }
        }
      }
      d[i] = best;
      arg[i] = bs;
      if (best < dmin) dmin = best;
    }
    double z = 0.0;
    for (int i = 0; i < n_terms; ++i) z += exp(-(d[i] - dmin) / tau);
    dk[g] = dmin - tau * log(z);
    for (int i = 0; i < n_terms; ++i) {
      const double wgt = exp(-(d[i] - dmin) / tau) / z;
      if (wgt == 0.0) continue;
      const int sl = arg[i];
      if (g == 0) {
        const int li = md->pi[i], lj = md->pj[i];
        const double* ci = cw + 3 * (li * S + sl);
        const double* cj = cw + 3 * (lj * S + sl);
        const double x[3] = {cj[0] - ci[0] + e, cj[1] - ci[1] + e, cj[2] - ci[2] + e};
        const double n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        const double gj[3] = {wgt * x[0] / n, wgt * x[1] / n, wgt * x[2] / n};
        const double gi[3] = {-gj[0], -gj[1], -gj[2]};
        pull_back(md, Tw, lj, cj, gj, J);
        pull_back(md, Tw, li, ci, gi, J);
      } else {
        const int l = i / (gstart[g] - gstart[g - 1]);
        const int m = gstart[g - 1] + i % (gstart[g] - gstart[g - 1]);
        const double* o = md->world + 4 * m;
        const double* c = cw + 3 * (l * S + sl);
        const double x[3] = {o[0] - c[0] + e, o[1] - c[1] + e, o[2] - c[2] + e};
        const double n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        const double gc[3] = {-wgt * x[0] / n, -wgt * x[1] / n, -wgt * x[2] / n};
        pull_back(md, Tw, l, c, gc, J);
      }
    }
    free(d);
    free(arg);
  }
  free(cw);
}

void or_sco_linearize(const float* twists, const float* parent_tf, const int32_t* parent_idx, const int32_t* act_idx,
                      const float* mimic_mul, const float* mimic_off, const int32_t* mimic_act_idx,
                      const int32_t* topo_inv, int J, int A, const int32_t* link_parent_joint, int L,
                      const double* spheres, int S, const int32_t* pair_i, const int32_t* pair_j, int P,
                      const double* world, int M, double margin, int mode, const double* q, int64_t W,
                      const int32_t* gstart, int n_wg, double tau, double* dk, double* Jk, int threads) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  traj_model_t md = {&t, link_parent_joint, L, spheres, S, pair_i, pair_j, P, world, M, margin, mode};
  const int G = 1 + n_wg;
  set_threads(threads);
#pragma omp parallel for schedule(dynamic)
  for (int64_t w = 0; w < W; ++w) sco_linearize_wp(&md, q + w * A, gstart, n_wg, tau, dk + w * G, Jk + w * G * A);
}

/* The inner objective of _lbfgs_inner_solve (_sco_optimization.py:291-339) for
 * one trajectory: smoothness (velocity, acceleration, jerk) + limits + w_coll
 * sum over waypoints and groups of max(0, margin - d_lin)^2 with d_lin =
 * d_k + J_k (x_t - q_k,t) + w_trust |x - q_k|^2; grad (T, A) when non-NULL. */
static double sco_inner_cost(const double* x, const double* qk, const double* dk, const double* Jk, int T, int A,
                             int G, const double* lower, const double* upper, const double* w, double w_coll,
                             double margin, double w_trust, double* grad) {
  tables_t dummy = {0, A, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL};
  traj_model_t md = {&dummy, NULL, 0, NULL, 0, NULL, NULL, 0, NULL, 0, 0.0, 0};
  double comps[3];
  const double ws[6] = {w[0], w[1], w[2], w[3], w[4], 0.0};
  traj_cost_one(&md, x, T, lower, upper, ws, comps, grad);
  double c = w[0] * comps[0] + w[4] * comps[1];
  for (int t = 0; t < T; ++t) {
    for (int g = 0; g < G; ++g) {
      const double* Jr = Jk + ((size_t)t * G + g) * A;
      double dl = dk[t * G + g];
      for (int a = 0; a < A; ++a) dl += Jr[a] * (x[t * A + a] - qk[t * A + a]);
      const double h = margin - dl;
      if (h > 0.0) {
        c += w_coll * h * h;
        if (grad)
          for (int a = 0; a < A; ++a) grad[t * A + a] -= 2.0 * w_coll * h * Jr[a];
      }
    }
    for (int a = 0; a < A; ++a) {
      const double dq = x[t * A + a] - qk[t * A + a];
      c += w_trust * dq * dq;
      if (grad) grad[t * A + a] += 2.0 * w_trust * dq;
    }
  }
  return c;
}

/* SCO per trajectory (_sco_optimization.py:291-503): n_outer rounds of
 * linearise at the current path, solve the inner problem with n_inner L-BFGS
 * steps from it (history m, the five-step line search with sufficient
 * decrease 1e-4, the best inner iterate returned), collision weight x
 * penalty_scale up to w_collision_max.  icfg = {n_outer, n_inner, m};
 * fcfg = {w_smooth, w_vel, w_acc, w_jerk, w_limits, w_collision,
 * w_collision_max, penalty_scale, margin, w_trust, tau}.  Endpoints stay where
 * x_init has them.  out_traj (W, A); out_cost (B) = the final nonlinear cost
 * at w_collision_max with the colldist activation (_eval_cost, :421-447; the
 * model's mode must be 1). */
void or_traj_sco(const float* twists, const float* parent_tf, const int32_t* parent_idx, const int32_t* act_idx,
                 const float* mimic_mul, const float* mimic_off, const int32_t* mimic_act_idx,
                 const int32_t* topo_inv, int J, int A, const int32_t* link_parent_joint, int L,
                 const double* spheres, int S, const int32_t* pair_i, const int32_t* pair_j, int P,
                 const double* world, int M, double margin_model, int mode, const double* x_init,
                 const int64_t* start, int64_t B, const double* lower, const double* upper, const int32_t* gstart,
                 int n_wg, const int32_t* icfg, const double* fcfg, double* out_traj, double* out_cost,
                 int threads) {
  static const double alphas[5] = {1.0, 0.5, 0.25, 0.1, 0.025};
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  traj_model_t md = {&t, link_parent_joint, L, spheres, S, pair_i, pair_j, P, world, M, margin_model, mode};
  const int n_outer = icfg[0], n_inner = icfg[1], m = icfg[2];
  const double w[5] = {fcfg[0], fcfg[1], fcfg[2], fcfg[3], fcfg[4]};
  const double w_c0 = fcfg[5], w_cmax = fcfg[6], scale = fcfg[7], margin = fcfg[8], w_trust = fcfg[9];
  const double tau = fcfg[10];
  const int G = 1 + n_wg;
  set_threads(threads);
#pragma omp parallel for schedule(dynamic)
  for (int64_t b = 0; b < B; ++b) {
    const int T = (int)(start[b + 1] - start[b]);
    const int n = T * A;
    double* x = (double*)malloc(sizeof(double) * n);
    double* qk = (double*)malloc(sizeof(double) * n);
    double* dk = (double*)malloc(sizeof(double) * T * G);
    double* Jk = (double*)malloc(sizeof(double) * T * G * A);
    double* xi = (double*)malloc(sizeof(double) * n);
    double* xp = (double*)malloc(sizeof(double) * n);
    double* g = (double*)malloc(sizeof(double) * n);
    double* gp = (double*)malloc(sizeof(double) * n);
    double* dir = (double*)malloc(sizeof(double) * n);
    double* xt = (double*)malloc(sizeof(double) * n);
    double* best = (double*)malloc(sizeof(double) * n);
    double* sb = (double*)calloc((size_t)m * n, sizeof(double));
    double* yb = (double*)calloc((size_t)m * n, sizeof(double));
    memcpy(x, x_init + start[b] * A, sizeof(double) * n);
    double w_coll = w_c0;
    for (int it = 0; it < n_outer; ++it) {
      memcpy(qk, x, sizeof(double) * n);
      for (int tt = 0; tt < T; ++tt) sco_linearize_wp(&md, qk + tt * A, gstart, n_wg, tau, dk + tt * G, Jk + tt * G * A);
      /* inner L-BFGS from x = q_k */
      double rho[16] = {0};
      int m_used = 0, newest = 0;
      memset(sb, 0, sizeof(double) * m * n);
      memset(yb, 0, sizeof(double) * m * n);
      memcpy(xi, qk, sizeof(double) * n);
      double c0 = sco_inner_cost(xi, qk, dk, Jk, T, A, G, lower, upper, w, w_coll, margin, w_trust, g);
      for (int a = 0; a < A; ++a) { g[a] = 0; g[n - A + a] = 0; }
      double best_cost = c0;
      memcpy(best, xi, sizeof(double) * n);
      memcpy(xp, xi, sizeof(double) * n);
      memcpy(gp, g, sizeof(double) * n);
      for (int k = 0; k < n_inner; ++k) {
        const double cost = sco_inner_cost(xi, qk, dk, Jk, T, A, G, lower, upper, w, w_coll, margin, w_trust, g);
        for (int a = 0; a < A; ++a) { g[a] = 0; g[n - A + a] = 0; }
        double sy = 0, yy = 0;
        for (int i = 0; i < n; ++i) {
          const double sv = xi[i] - xp[i], yv = g[i] - gp[i];
          sy += sv * yv;
          yy += yv * yv;
        }
        if (sy > 1e-10 * yy + 1e-30 && k > 0) {
          newest = (newest + 1) % m;
          for (int i = 0; i < n; ++i) { sb[newest * n + i] = xi[i] - xp[i]; yb[newest * n + i] = g[i] - gp[i]; }
          rho[newest] = 1.0 / (sy + 1e-30);
          if (m_used < m) ++m_used;
        }
        if (m_used > 0) {
          two_loop(g, sb, yb, rho, m_used, newest, m, n, dir);
        } else {
          double gn = 0;
          for (int i = 0; i < n; ++i) gn += g[i] * g[i];
          gn = sqrt(gn) + 1e-18;
          for (int i = 0; i < n; ++i) dir[i] = -g[i] / gn;
        }
        for (int a = 0; a < A; ++a) { dir[a] = 0; dir[n - A + a] = 0; }
        const double thresh = cost * (1.0 - 1e-4);
        double trial[5];
        for (int i = 0; i < 5; ++i) {
          for (int k2 = 0; k2 < n; ++k2) xt[k2] = xi[k2] + alphas[i] * dir[k2];
          trial[i] = sco_inner_cost(xt, qk, dk, Jk, T, A, G, lower, upper, w, w_coll, margin, w_trust, NULL);
        }
        int pick = -1;
        for (int i = 0; i < 5; ++i)
          if (trial[i] < thresh) { pick = i; break; }
        if (pick < 0) {
          pick = 0;
          for (int i = 1; i < 5; ++i)
            if (trial[i] < trial[pick]) pick = i;
        }
        memcpy(xp, xi, sizeof(double) * n);
        memcpy(gp, g, sizeof(double) * n);
        for (int i = 0; i < n; ++i) xi[i] += alphas[pick] * dir[i];
        if (trial[pick] < best_cost) {
          best_cost = trial[pick];
          memcpy(best, xi, sizeof(double) * n);
        }
      }
      memcpy(x, best, sizeof(double) * n);
      for (int a = 0; a < A; ++a) {  /* re-pin the endpoints */
        x[a] = x_init[start[b] * A + a];
        x[n - A + a] = x_init[start[b] * A + n - A + a];
      }
      w_coll = fmin(w_coll * scale, w_cmax);
    }
    double comps[3];
    const double we[6] = {w[0], w[1], w[2], w[3], w[4], w_cmax};
    traj_cost_one(&md, x, T, lower, upper, we, comps, NULL);
    out_cost[b] = w[0] * comps[0] + w[4] * comps[1] + w_cmax * comps[2];
    memcpy(out_traj + start[b] * A, x, sizeof(double) * n);
    free(x); free(qk); free(dk); free(Jk); free(xi); free(xp); free(g); free(gp); free(dir); free(xt); free(best);
    free(sb); free(yb);
  }
}

/* ------------------------------------------------- JAX-convention LM / HJCD */

/* jaxlie 1.5.0 (a third-party dependency, pinned in pyproject.toml:27, absent
 * here): SO3.log, SE3.log, inverse and compose in float64 (epsilon 1e-10),
 * as the reference's JAX residual _ik_primitives.py:93-103 evaluates them.
 * jlog_dir returns the directional derivative of log(T) under the left
 * perturbation T -> exp(d) T, branch by branch, i.e. what reverse-mode AD of
 * that code computes. */
#define JL_EPS 1e-10
#define OR_PI 3.14159265358979323846

typedef struct {
  double q[4], t[3];     /* T_rel: quaternion (as composed) and translation */
  double om[3], rho[3];  /* log = [rho; om] */
  double F, Fw, Fn2;     /* om = q_v F, dF = Fw dq_w + Fn2 d|q_v|^2 */
  double c, dc;          /* V^-1 = I - Om/2 + c Om^2, dc = dc * d(theta^2) */
} jlog_t;

static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

static void jlog_eval(const double* q, const double* t, jlog_t* L) {
  for (int i = 0; i < 4; ++i) L->q[i] = q[i];
  for (int i = 0; i < 3; ++i) L->t[i] = t[i];
  const double w = q[0], n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  if (n2 < JL_EPS) {
    L->F = 2.0 / w - 2.0 / 3.0 * n2 / (w * w * w);
    L->Fw = -2.0 / (w * w) + 2.0 * n2 / (w * w * w * w);
    L->Fn2 = -2.0 / 3.0 / (w * w * w);
  } else {
    const double n = sqrt(n2);
    if (fabs(w) < JL_EPS) {
      const double s = w > 0 ? 1.0 : -1.0;
      L->F = s * OR_PI / n;
      L->Fw = 0.0;
      L->Fn2 = -s * OR_PI / (2.0 * n2 * n);
    } else {
      const double a = atan2(w < 0 ? -n : n, fabs(w));
      L->F = 2.0 * a / n;
      L->Fw = -2.0 / (n2 + w * w);
      L->Fn2 = w / (n2 * (n2 + w * w)) - a / (n2 * n);
    }
  }
  for (int i = 0; i < 3; ++i) L->om[i] = q[1 + i] * L->F;
  const double th2 = L->om[0] * L->om[0] + L->om[1] * L->om[1] + L->om[2] * L->om[2];
  if (th2 < JL_EPS) {
    L->c = 1.0 / 12.0;
    L->dc = 0.0;
  } else {
    const double th = sqrt(th2), h = 0.5 * th;
    const double g = 1.0 - h * cos(h) / sin(h);
    const double gp = -0.5 * cos(h) / sin(h) + 0.5 * h / (sin(h) * sin(h));
    L->c = g / th2;
    L->dc = (gp / th2 - 2.0 * g / (th2 * th)) / (2.0 * th);
  }
  double ot[3], oot[3];
  cross3(L->om, t, ot);
  cross3(L->om, ot, oot);
  for (int i = 0; i < 3; ++i) L->rho[i] = t[i] - 0.5 * ot[i] + L->c * oot[i];
}

/* d log(T) for the left perturbation d = (dv, dw): out = [d rho; d om]. */
static void jlog_dir(const jlog_t* L, const double* dv, const double* dw, double* out) {
  const double* q = L->q;
  const double dqw = -0.5 * (dw[0] * q[1] + dw[1] * q[2] + dw[2] * q[3]);
  double dqv[3], wxq[3];
  cross3(dw, q + 1, wxq);
  for (int i = 0; i < 3; ++i) dqv[i] = 0.5 * (q[0] * dw[i] + wxq[i]);
  const double dn2 = 2.0 * (q[1] * dqv[0] + q[2] * dqv[1] + q[3] * dqv[2]);
  const double dF = L->Fw * dqw + L->Fn2 * dn2;
  double dom[3];
  for (int i = 0; i < 3; ++i) dom[i] = dqv[i] * L->F + q[1 + i] * dF;
  double dt[3], wxt[3];
  cross3(dw, L->t, wxt);
  for (int i = 0; i < 3; ++i) dt[i] = wxt[i] + dv[i];
  const double dth2 = 2.0 * (L->om[0] * dom[0] + L->om[1] * dom[1] + L->om[2] * dom[2]);
  double a1[3], a2[3], ot[3], oot[3], dot_[3], odt[3], oodt[3], domot[3], odomt[3];
  cross3(dom, L->t, a1);        /* dw x t */
  cross3(L->om, dt, a2);        /* w x dt */
  cross3(L->om, L->t, ot);
  cross3(L->om, ot, oot);       /* w x (w x t) */
  cross3(dom, ot, domot);       /* dw x (w x t) */
  cross3(L->om, a1, odomt);     /* w x (dw x t) */
  cross3(L->om, dt, odt);
  cross3(L->om, odt, oodt);     /* w x (w x dt) */
  (void)dot_;
  for (int i = 0; i < 3; ++i)
    out[i] = dt[i] - 0.5 * (a1[i] + a2[i]) + L->dc * dth2 * oot[i] + L->c * (domot[i] + odomt[i] + oodt[i]);
  for (int i = 0; i < 3; ++i) out[3 + i] = dom[i];
}

/* T_rel = T_act^-1 T_tgt (jaxlie inverse, then compose). */
static void jrel(const se3* A, const double* tgt, double* q, double* t) {
  qmul(A->w, -A->x, -A->y, -A->z, tgt[0], tgt[1], tgt[2], tgt[3], q);
  const double d[3] = {tgt[4] - A->tx, tgt[5] - A->ty, tgt[6] - A->tz};
  se3 inv = {A->w, -A->x, -A->y, -A->z, 0, 0, 0};
  qrot(&inv, d, t);
}

typedef struct {
  const tables_t* t;
  int n_ee;
  const int32_t* tj;   /* (n_ee) target joints */
  const double* tgt;   /* (n_ee, 7) */
} jprob_t;

/* Residuals f = log(T_ee^-1 T_tgt) of EEs [e0, e1) at q (f[6 (e - e0) ..]) and,
 * when Jm is non-NULL, their exact Jacobian (6 (e1 - e0) x A, row-major):
 * joint j of the EE's chain with spatial twist xi_s = (R_j v + t_j x w_s,
 * w_s = R_j w) moves T_ee by the body twist xi_b = Ad(T_ee^-1) xi_s, which
 * perturbs T_rel on the left by -xi_b; its column is mul_j * D log[-xi_b]
 * added to the joint's actuated source. */
static void jres_jac(const jprob_t* p, const double* q, int e0, int e1, double* f, double* Jm) {
  se3 Tw[OR_MAXJ];
  const tables_t* t = p->t;
  const int A = t->A;
  fk_one(t, q, Tw);
  if (Jm) memset(Jm, 0, sizeof(double) * 6 * (e1 - e0) * A);
  for (int e = e0; e < e1; ++e) {
    const int r0 = 6 * (e - e0);
    const se3* Te = &Tw[p->tj[e]];
    double qr[4], tr[3];
    jrel(Te, p->tgt + 7 * e, qr, tr);
    jlog_t L;
    jlog_eval(qr, tr, &L);
    for (int c = 0; c < 3; ++c) { f[r0 + c] = L.rho[c]; f[r0 + 3 + c] = L.om[c]; }
    if (!Jm) continue;
    const se3 inv = {Te->w, -Te->x, -Te->y, -Te->z, 0, 0, 0};
    for (int j = p->tj[e]; j >= 0; j = t->parent_idx[j]) {
      const int m = t->mimic_act_idx[j];
      const int src = (m != -1) ? m : t->act_idx[j];
      if (src < 0 || src >= A) continue;
      const float* tw = t->twists + 6 * j;
      const double vl[3] = {tw[0], tw[1], tw[2]}, wl[3] = {tw[3], tw[4], tw[5]};
      double ws[3], vs[3], arm[3], axw[3], vb[3], wb[3], col[6];
      qrot(&Tw[j], wl, ws);
      qrot(&Tw[j], vl, vs);
      arm[0] = Tw[j].tx - Te->tx; arm[1] = Tw[j].ty - Te->ty; arm[2] = Tw[j].tz - Te->tz;
      cross3(arm, ws, axw);
      for (int c = 0; c < 3; ++c) vs[c] += axw[c];
      qrot(&inv, vs, vb);
      qrot(&inv, ws, wb);
      jlog_dir(&L, vb, wb, col);
      const double mul = t->mimic_mul[j];
      for (int c = 0; c < 6; ++c) Jm[(r0 + c) * A + src] -= mul * col[c];
    }
  }
}

/* _adaptive_weights (_ik_primitives.py:106-116) of one 6-vector. */
static double adaptive_scale(const double* f) {
  const double pe = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) + 1e-8;
  const double oe = sqrt(f[3] * f[3] + f[4] * f[4] + f[5] * f[5]) + 1e-8;
  const double s = pe / oe;
  return s < 0.05 ? 0.05 : (s > 1.0 ? 1.0 : s);
}

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* Counter-based standard normal for the stall kicks: splitmix64 (Steele,
 * Lea and Flood) of (key, seed, iteration, joint), Box-Muller of two 24-bit
 * uniforms.  The HIP solver draws the same bits. */
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double kick_normal(uint64_t key, uint64_t gid, int it, int j) {
  const uint64_t h = mix64(key ^ mix64((gid << 20) ^ ((uint64_t)it << 6) ^ (uint64_t)j));
  const double u1 = (double)((h >> 40) + 1) * (1.0 / 16777216.0);
  const double u2 = (double)((h >> 16) & 0xFFFFFFull) * (1.0 / 16777216.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * OR_PI * u2);
}

typedef struct {
  int max_iter;
  double pos_weight, ori_weight, lambda_init;
  int adaptive;          /* 1: per-EE adaptive weights (HJCD), 0: W = (pw, pw, pw, ow, ow, ow) */
  double prior_weight;   /* soft joint-limit prior (HJCD), 0 off */
  double kick_scale;     /* stall kicks (HJCD), 0 off */
  uint64_t kick_key;
} jlm_opts_t;

static void jweights(const jlm_opts_t* o, const double* f, int n_ee, double* w) {
  for (int e = 0; e < n_ee; ++e) {
    const double s = o->adaptive ? adaptive_scale(f + 6 * e) : 0.0;
    for (int c = 0; c < 6; ++c)
      w[6 * e + c] = o->adaptive ? (c < 3 ? 1.0 : s) : (c < 3 ? o->pos_weight : o->ori_weight);
  }
}

/* One seed of _ls_ik_single (_ls_ik.py:192-350; adaptive = 0, no prior, no
 * kicks) or _lm_refine_single (_hjcd_ik.py:140-365).  Returns the all-time
 * best configuration; *out_err = the winner metric at it, sum over EEs of
 * |W f|^2 (ls_ik_solve, _ls_ik.py:503-516) or |f|^2 (hjcd_solve,
 * _hjcd_ik.py:616-630). */
static void jlm_seed(const jprob_t* p, const double* lo, const double* hi, const int32_t* fixed,
                     const jlm_opts_t* o, const double* seed, uint64_t gid, double* out_q, double* out_err) {
  static const double alphas[5] = {1.0, 0.5, 0.25, 0.1, 0.025};
  const int A = p->t->A, n_ee = p->n_ee, R6 = 6 * n_ee;
  double c[OR_MAXA], best[OR_MAXA], f[6 * OR_MAXEE], w[6 * OR_MAXEE], Jm[6 * OR_MAXEE * OR_MAXA];
  double mid[OR_MAXA], dprior[OR_MAXA];
  for (int a = 0; a < A; ++a) {
    c[a] = best[a] = seed[a];
    mid[a] = 0.5 * (lo[a] + hi[a]);
    const double half = 0.5 * (hi[a] - lo[a]) + 1e-8;
    dprior[a] = o->prior_weight / (half * half);
  }
  jres_jac(p, c, 0, n_ee, f, NULL);
  jweights(o, f, n_ee, w);
  double best_err = 0.0;
  for (int i = 0; i < R6; ++i) best_err += (w[i] * f[i]) * (w[i] * f[i]);
  double lam = o->lambda_init;
  int stall = 0;
  for (int it = 0; it < o->max_iter; ++it) {
    jres_jac(p, c, 0, n_ee, f, Jm);
    jweights(o, f, n_ee, w);
    double fw[6 * OR_MAXEE], cur = 0.0, col[OR_MAXA];
    for (int i = 0; i < R6; ++i) { fw[i] = w[i] * f[i]; cur += fw[i] * fw[i]; }
    for (int a = 0; a < A; ++a) {
      double s = 0.0;
      for (int i = 0; i < R6; ++i) { Jm[i * A + a] *= w[i]; s += Jm[i * A + a] * Jm[i * A + a]; }
      col[a] = sqrt(s) + 1e-8;
    }
    double M[OR_MAXA * OR_MAXA], rhs[OR_MAXA];
    for (int a = 0; a < A; ++a) {
      for (int b = 0; b < A; ++b) {
        double s = 0.0;
        for (int i = 0; i < R6; ++i) s += Jm[i * A + a] * Jm[i * A + b];
        M[a * A + b] = s / (col[a] * col[b]);
      }
      M[a * A + a] += lam + dprior[a] / (col[a] * col[a]);
      double s = 0.0;
      for (int i = 0; i < R6; ++i) s += Jm[i * A + a] * fw[i];
      rhs[a] = -(s / col[a] + dprior[a] * (c[a] - mid[a]) / col[a]);
    }
    chol_solve(M, rhs, A);
    double delta[OR_MAXA], dn = 0.0;
    for (int a = 0; a < A; ++a) {
      delta[a] = fixed[a] ? 0.0 : rhs[a] / col[a];
      dn += delta[a] * delta[a];
    }
    double pe = 0.0, oe = 0.0;
    for (int e = 0; e < n_ee; ++e) {
      const double* fe = f + 6 * e;
      pe = fmax(pe, sqrt(fe[0] * fe[0] + fe[1] * fe[1] + fe[2] * fe[2]));
      oe = fmax(oe, sqrt(fe[3] * fe[3] + fe[4] * fe[4] + fe[5] * fe[5]));
    }
    const double R = (pe > 1e-2 || oe > 0.6) ? 0.38 : (pe > 1e-3 || oe > 0.25) ? 0.22
                   : (pe > 2e-4 || oe > 0.08) ? 0.12 : 0.05;
    dn = sqrt(dn) + 1e-18;
    if (dn > R)
      for (int a = 0; a < A; ++a) delta[a] *= R / dn;
    double best_a = INFINITY, qa[OR_MAXA], newc[OR_MAXA];
    int pick = 0;
    for (int k = 0; k < 5; ++k) {
      for (int a = 0; a < A; ++a) qa[a] = clampd(c[a] + alphas[k] * delta[a], lo[a], hi[a]);
      double fa[6 * OR_MAXEE], ea = 0.0;
      jres_jac(p, qa, 0, n_ee, fa, NULL);
      for (int i = 0; i < R6; ++i) ea += (w[i] * fa[i]) * (w[i] * fa[i]);
      if (ea < best_a) { best_a = ea; pick = k; }
    }
    for (int a = 0; a < A; ++a) newc[a] = clampd(c[a] + alphas[pick] * delta[a], lo[a], hi[a]);
    const int improved = best_a < cur * (1.0 - 1e-4);
    lam = clampd(improved ? lam * 0.5 : lam * 3.0, 1e-10, 1e6);
    if (best_a < best_err) {
      best_err = best_a;
      for (int a = 0; a < A; ++a) best[a] = newc[a];
    }
    if (improved)
      for (int a = 0; a < A; ++a) c[a] = newc[a];
    if (o->kick_scale > 0.0) {
      stall = improved ? 0 : stall + 1;
      if (stall >= 6) {
        for (int a = 0; a < A; ++a) {
          const double k = fixed[a] ? 0.0 : kick_normal(o->kick_key, gid, it, a) * o->kick_scale;
          c[a] = clampd(c[a] + k, lo[a], hi[a]);
        }
        lam = o->lambda_init;
        stall = 0;
      }
    }
  }
  jres_jac(p, best, 0, n_ee, f, NULL);
  double e = 0.0;
  for (int i = 0; i < R6; ++i) {
    const double wi = o->adaptive ? 1.0 : ((i % 6) < 3 ? o->pos_weight : o->ori_weight);
    e += (wi * f[i]) * (wi * f[i]);
  }
  for (int a = 0; a < A; ++a) out_q[a] = best[a];
  *out_err = e;
}

/* seeds (P, S, A), target_T (P, n_ee, 7); out (P, S, A), out_err (P, S).
 * iopt = {max_iter, adaptive}; dopt = {pos_weight, ori_weight, lambda_init,
 * prior_weight, kick_scale}; seed (p, s) draws kicks with id gid0 + p S + s. */
void or_jlm(const double* seeds, int64_t P, int S, int A, const float* twists, const float* parent_tf,
            const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul, const float* mimic_off,
            const int32_t* mimic_act_idx, const int32_t* topo_inv, int J, const int32_t* target_jnts, int n_ee,
            const double* target_T, const double* lower, const double* upper, const int32_t* fixed_mask,
            const int32_t* iopt, const double* dopt, uint64_t kick_key, double* out, double* out_err,
            int threads) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  const jlm_opts_t o = {iopt[0], dopt[0], dopt[1], dopt[2], iopt[1], dopt[3], dopt[4], kick_key};
  set_threads(threads);
#pragma omp parallel for schedule(dynamic)
  for (int64_t i = 0; i < P * S; ++i) {
    const int64_t pb = i / S;
    const jprob_t pr = {&t, n_ee, target_jnts, target_T + pb * n_ee * 7};
    jlm_seed(&pr, lower, upper, fixed_mask, &o, seeds + i * A, (uint64_t)i, out + i * A, out_err + i);
  }
}

/* _coarse_search_single (_hjcd_ik.py:55-137) for every seed: seed (p, s) runs
 * k_max Hamiltonian coordinate-descent steps on EE ee_of_seed[s] alone.
 * out_err (P, S) = sum over all EEs of |f|^2 at the result (:591-597). */
void or_hjcd_coarse(const double* seeds, int64_t P, int S, int A, const float* twists, const float* parent_tf,
                    const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
                    const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv, int J,
                    const int32_t* target_jnts, int n_ee, const double* target_T, const int32_t* ee_of_seed,
                    const double* lower, const double* upper, const int32_t* fixed_mask, int k_max,
                    double epsilon, double nu, double* out, double* out_err, int threads) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  set_threads(threads);
#pragma omp parallel for schedule(dynamic)
  for (int64_t i = 0; i < P * S; ++i) {
    const int64_t pb = i / S;
    const int s = (int)(i % S);
    const jprob_t pr = {&t, n_ee, target_jnts, target_T + pb * n_ee * 7};
    const int ee = ee_of_seed[s];
    double c[OR_MAXA], mom[OR_MAXA] = {0}, f[6 * OR_MAXEE], Jm[6 * OR_MAXA];
    for (int a = 0; a < A; ++a) c[a] = seeds[i * A + a];
    for (int k = 0; k < k_max; ++k) {
      jres_jac(&pr, c, ee, ee + 1, f, Jm);
      const double pe = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
      const double oe = sqrt(f[3] * f[3] + f[4] * f[4] + f[5] * f[5]);
      if (pe < epsilon && oe < nu) continue;
      const double gate = pe < 1e-3 ? 1.0 : 0.0;
      const double so = adaptive_scale(f) * gate;
      double w[6] = {1, 1, 1, so, so, so};
      int bj = 0;
      double bs = -INFINITY, gb = 0, nb = 1;
      for (int a = 0; a < A; ++a) {
        double g = 0.0, nsq = 1e-8;
        for (int r = 0; r < 6; ++r) {
          const double jw = w[r] * Jm[r * A + a];
          g += jw * (w[r] * f[r]);
          nsq += jw * jw;
        }
        const double score = fixed_mask[a] ? -1.0 : fabs(g) / sqrt(nsq);
        if (score > bs) { bs = score; bj = a; gb = g; nb = nsq; }
      }
      const double pb2 = 0.9 * mom[bj] - 0.35 * gb / nb;
      mom[bj] = pb2;
      if (!fixed_mask[bj]) c[bj] = clampd(c[bj] + pb2, lower[bj], upper[bj]);
    }
    jres_jac(&pr, c, 0, n_ee, f, NULL);
    double e = 0.0;
    for (int r = 0; r < 6 * n_ee; ++r) e += f[r] * f[r];
    for (int a = 0; a < A; ++a) out[i * A + a] = c[a];
    out_err[i] = e;
  }
}

/* Residuals (B, n_ee, 6) of configurations cfg (B, A) against one target set
 * per configuration (B, n_ee, 7), and optionally their Jacobians (B, 6 n_ee, A). */
void or_jres(const double* cfg, int64_t B, int A, const float* twists, const float* parent_tf,
             const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul, const float* mimic_off,
             const int32_t* mimic_act_idx, const int32_t* topo_inv, int J, const int32_t* target_jnts, int n_ee,
             const double* target_T, double* out_f, double* out_J) {
  tables_t t = {J, A, twists, parent_tf, mimic_mul, mimic_off, parent_idx, act_idx, mimic_act_idx, topo_inv};
  for (int64_t b = 0; b < B; ++b) {
    const jprob_t pr = {&t, n_ee, target_jnts, target_T + b * n_ee * 7};
    jres_jac(&pr, cfg + b * A, 0, n_ee, out_f + b * 6 * n_ee, out_J ? out_J + b * 6 * n_ee * A : NULL);
  }
}
