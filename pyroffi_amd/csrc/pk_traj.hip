// Trajectory cost, its hand-written adjoint, and the direct L-BFGS trajectory
// optimiser, for gfx950.
//
// Cost of one trajectory x (T waypoints x A joints), as the reference's
// L-BFGS / CHOMP trajectory optimisers define it:
//   total  = w_smooth * smooth + w_limits * limits + w_coll * coll
//   smooth = w_vel sum|vel|^2 + w_acc sum|acc|^2 + w_jerk sum|jerk|^2
//            (_sco_optimization.py:170-189; acc is the 5-point stencil
//            (-1, 16, -30, 16, -1)/12, jerk = acc_{t+1} - acc_t)
//   limits = sum (max(0, x - u) + max(0, l - x))^2   (_sco_optimization.py:192-200)
//   coll   = sum over waypoints of sum f(d) over the JAX collision distances
//            (_sco_optimization.py:203-222): every active self pair (same-slot
//            sphere pairs of the two links) and every (link, world sphere)
//            pair (closest sphere of the link), both with the +1e-6 bias of
//            collision/_utils.py:45; f = max(0, m - d)^2 (mode 0, CHOMP,
//            _chomp_optimization.py:140-154) or -min(colldist_from_sdf(d, m), 0)
//            (mode 1, _lbfgs_trajopt_optimization.py:108-111).
//
// The optimiser is _lbfgs_direct_solve (_lbfgs_trajopt_optimization.py:76-205):
// collision-weight continuation, endpoint-masked gradients, L-BFGS history with
// the curvature test s.y > 1e-10 y.y + 1e-30, the two-loop recursion
// (_sco_optimization.py:119-167), five line-search steps {1, .5, .25, .1, .025}
// with the first one that decreases the cost by 1e-4 (relative) or else the
// cheapest, and the best iterate tracked at the maximum collision weight.
//
// gfx950 layout: one workgroup (four waves) per trajectory, so trajectories of
// different lengths run side by side in one launch.  The robot (fixed joints
// folded into moving "nodes" on the host), its spheres in node frames, the
// self-pair lists and the world live in LDS; the trajectory, its gradient and
// the search direction live in LDS as well; the L-BFGS history stays in HBM.
// A group of 8 lanes evaluates one waypoint (16 to 64 waypoints in flight per
// workgroup); every lane of the group computes the node poses in registers,
// the group writes the waypoint's world sphere centres and link bounding
// spheres to its LDS slot, and each lane then takes an eighth of the spheres
// and of the self pairs:
//   * world term: a CSR grid (pk_grid.h, small enough to stay in L2) lists,
//     per cell, the only world spheres within reach; the link's lowest-index
//     sphere within the margin of a world sphere owns that (link, world
//     sphere) term and takes the minimum over the link's spheres;
//   * self term: a bounding-sphere test per pair skips pairs farther apart
//     than the margin (exact: f vanishes there);
//   * adjoint: dcost/dc at a sphere centre c is pulled back to every ancestor
//     node a of the sphere's node, dcost/dq_a += g . (w_a x (c - p_a) + v_a)
//     with (v_a, w_a) the node's twist in world axes and p_a its origin, then
//     summed over the group and mapped to the actuated joints (mimic factors).
// The line search needs no gradient: the smoothness of x + a d is quadratic in
// a (three sums), the limits are elementwise, and the steps are evaluated in
// order until one is accepted (the same pick as evaluating all five).
#include "pk_grid.h"
#include "pk_model.h"

#include <math.h>

#include <algorithm>

namespace pk {

// A workgroup has 128, 256 or 512 threads (chosen per trajectory length on the
// host); loops and reductions read the size from blockDim.x.
constexpr int kMaxTrajThreads = 512;
constexpr int kMaxTrajWaves = kMaxTrajThreads / kWave;
constexpr int kGroup = 8;  // lanes per waypoint evaluation
static_assert(kWave % kGroup == 0, "a group must not straddle waves");
constexpr int kMaxHist = 8;
constexpr int kNumAlphas = 5;
constexpr int kNodeWords = 20;   // ptf[7] lin[3] ang[3] mul off src parent kind |ang| 1/|ang|
constexpr int kRedSlots = 16;    // values per block reduction
constexpr int kScalarWords = 16; // per-trajectory solver scalars in the state buffer

__device__ __forceinline__ float traj_alpha(int i) {
  return i == 0 ? 1.f : (i == 1 ? 0.5f : (i == 2 ? 0.25f : (i == 3 ? 0.1f : 0.025f)));
}

struct TrajModelDev {
  const float* nodes;
  int n_nodes, n_act;
  const int* link_node;
  const int* link_sph_start;
  const float* spheres;
  const float* link_bound;
  int n_links, n_spheres;
  const int* pair_links;
  const int* pair_start;
  const int* pair_items;
  int n_pairs, n_items;
  const float* world;
  int n_world;
  GridArgs grid;
  const float* lower;
  const float* upper;
  float margin;
  int mode;
  unsigned a_magic;  // quot_magic(n_act)
};

// i / d for i >= 0 and i * d < 2^32 without a division: with m = ceil(2^32 / d)
// the quotient is the high half of i * m (exact under that bound: the arrays
// indexed this way live in LDS, at most 150 KiB, so i < 40 k, and d is the
// joint count or the collision group count, both at most 64); d = 1 is the
// identity.  The trajectory loops index element
// i = t * A + a of a (T, A) array and need t and a per element.
__host__ __device__ inline unsigned quot_magic(int d) {
  return d > 1 ? static_cast<unsigned>((0x100000000ull + static_cast<unsigned>(d) - 1u) / static_cast<unsigned>(d))
               : 0u;
}
__device__ __forceinline__ int quot(int i, int d, unsigned m) {
  return d > 1 ? static_cast<int>(__umulhi(static_cast<unsigned>(i), m)) : i;
}

struct TrajWeights {
  float w_smooth, w_vel, w_acc, w_jerk, w_limits;
};

// Byte offsets of the LDS image; the same function sizes the launch on the host.
struct TrajLdsLayout {
  size_t nodes, anc, link_node, link_sph, sph_link, sph, lb, pair_links, pair_start, pair_items, world, lo,
      hi, vx, vg, vd, vt, red, gc, total;
};

__host__ __device__ inline size_t traj_align(size_t v) { return (v + 15) & ~static_cast<size_t>(15); }

__host__ __device__ inline TrajLdsLayout traj_lds_layout(int N, int L, int K, int P, int Q, int M, int A,
                                                         int max_len, int threads) {
  TrajLdsLayout o;
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t here = at;
    at = traj_align(at + bytes);
    return here;
  };
  o.nodes = take(sizeof(float) * kNodeWords * N);
  o.anc = take(sizeof(int) * N);
  o.link_node = take(sizeof(int) * L);
  o.link_sph = take(sizeof(int) * (L + 1));
  o.sph_link = take(sizeof(int) * K);
  o.sph = take(sizeof(float4) * K);
  o.lb = take(sizeof(float4) * L);
  o.pair_links = take(sizeof(int) * 2 * P);
  o.pair_start = take(sizeof(int) * (P + 1));
  o.pair_items = take(sizeof(int) * 2 * Q);
  o.world = take(sizeof(float4) * M);
  o.lo = take(sizeof(float) * A);
  o.hi = take(sizeof(float) * A);
  const size_t vec = sizeof(float) * static_cast<size_t>(max_len) * A;
  o.vx = take(vec);
  o.vg = take(vec);
  o.vd = take(vec);
  o.vt = take(vec);
  o.red = take(sizeof(float) * kMaxTrajWaves * kRedSlots);
  o.gc = take(sizeof(float4) * (K + L) * (threads / kGroup));
  o.total = at;
  return o;
}

struct TrajLds {
  const float* nodes;
  const int* anc;      // bit a set when node a is node n or one of its ancestors
  const int* link_node;
  const int* link_sph;
  const int* sph_link;  // link of each sphere
  const float4* sph;
  const float4* lb;
  const int* pair_links;
  const int* pair_start;
  const int* pair_items;
  const float4* world;
  const float* lo;
  const float* hi;
  float* vx;   // trajectory (T, A)
  float* vg;   // gradient
  float* vd;   // search direction (and scratch in the gradient pass)
  float* vt;   // accelerations of x
  float* red;  // block-reduction scratch
  float4* gc;   // per group: world sphere centres + radius (K), link bounding spheres (L)
};

// Cooperative copy of the model into LDS, with every index clamped into range.
__device__ __forceinline__ TrajLds stage_traj_model(const TrajModelDev& md, int max_len, char* base) {
  const int N = md.n_nodes, L = md.n_links, K = md.n_spheres, P = md.n_pairs, Q = md.n_items;
  const int M = md.n_world, A = md.n_act;
  const TrajLdsLayout lay = traj_lds_layout(N, L, K, P, Q, M, A, max_len, blockDim.x);
  float* nodes = reinterpret_cast<float*>(base + lay.nodes);
  int* anc = reinterpret_cast<int*>(base + lay.anc);
  int* link_node = reinterpret_cast<int*>(base + lay.link_node);
  int* link_sph = reinterpret_cast<int*>(base + lay.link_sph);
  int* sph_link = reinterpret_cast<int*>(base + lay.sph_link);
  float4* sph = reinterpret_cast<float4*>(base + lay.sph);
  float4* lb = reinterpret_cast<float4*>(base + lay.lb);
  int* pair_links = reinterpret_cast<int*>(base + lay.pair_links);
  int* pair_start = reinterpret_cast<int*>(base + lay.pair_start);
  int* pair_items = reinterpret_cast<int*>(base + lay.pair_items);
  float4* world = reinterpret_cast<float4*>(base + lay.world);
  float* lo = reinterpret_cast<float*>(base + lay.lo);
  float* hi = reinterpret_cast<float*>(base + lay.hi);
  const int tid = threadIdx.x;
  for (int i = tid; i < N * kNodeWords; i += blockDim.x) {
    float v = md.nodes[i];
    const int n = i / kNodeWords, w = i % kNodeWords;
    if (w == 15) v = (v >= 0.f && v < static_cast<float>(A)) ? v : -1.f;           // src
    if (w == 16) v = (v >= 0.f && v < static_cast<float>(n)) ? v : -1.f;           // parent
    if (w == 17) v = (v >= 1.f && v <= 3.f) ? v : 3.f;                              // kind
    if (w == 18 || w == 19) {  // axis norm and its inverse, for pure rotations (kind 1)
      const float* r = md.nodes + n * kNodeWords;
      const float na = sqrtf(r[10] * r[10] + r[11] * r[11] + r[12] * r[12]);
      v = (r[17] == 1.f && na > 0.f) ? (w == 18 ? na : 1.f / na) : 0.f;
    }
    nodes[i] = v;
  }
  for (int l = tid; l < L; l += blockDim.x) {
    const int n = md.link_node[l];
    link_node[l] = (n >= 0 && n < N) ? n : -1;
  }
  if (tid == 0) {
    int prev = 0;
    for (int l = 0; l <= L; ++l) {
      const int v = (l == L) ? K : md.link_sph_start[l];
      prev = max(prev, min(max(v, 0), K));
      link_sph[l] = (l == 0) ? 0 : prev;
    }
    prev = 0;
    for (int p = 0; p <= P; ++p) {
      const int v = (p == P) ? Q : md.pair_start[p];
      prev = max(prev, min(max(v, 0), Q));
      pair_start[p] = (p == 0) ? 0 : prev;
    }
  }
  for (int k = tid; k < K; k += blockDim.x) {
    const float* s = md.spheres + 4 * k;
    sph[k] = make_float4(s[0], s[1], s[2], s[3]);
  }
  for (int l = tid; l < L; l += blockDim.x) {
    const float* s = md.link_bound + 4 * l;
    lb[l] = make_float4(s[0], s[1], s[2], s[3]);
  }
  for (int i = tid; i < 2 * P; i += blockDim.x) {
    const int v = md.pair_links[i];
    pair_links[i] = (v >= 0 && v < L) ? v : 0;
  }
  for (int i = tid; i < 2 * Q; i += blockDim.x) {
    const int v = md.pair_items[i];
    pair_items[i] = (v >= 0 && v < K) ? v : 0;
  }
  for (int m = tid; m < M; m += blockDim.x) {
    const float* s = md.world + 4 * m;
    world[m] = make_float4(s[0], s[1], s[2], s[3]);
  }
  for (int a = tid; a < A; a += blockDim.x) {
    lo[a] = md.lower[a];
    hi[a] = md.upper[a];
  }
  __syncthreads();
  if (tid == 0) {
    for (int n = 0; n < N; ++n) {
      const int p = static_cast<int>(nodes[n * kNodeWords + 16]);
      anc[n] = (1 << n) | (p >= 0 ? anc[p] : 0);
    }
  }
  for (int l = tid; l < L; l += blockDim.x)
    for (int k = link_sph[l]; k < link_sph[l + 1]; ++k) sph_link[k] = l;
  __syncthreads();
  TrajLds o;
  o.nodes = nodes; o.anc = anc; o.link_node = link_node; o.link_sph = link_sph; o.sph_link = sph_link;
  o.sph = sph; o.lb = lb;
  o.pair_links = pair_links; o.pair_start = pair_start; o.pair_items = pair_items; o.world = world;
  o.lo = lo; o.hi = hi;
  o.vx = reinterpret_cast<float*>(base + lay.vx);
  o.vg = reinterpret_cast<float*>(base + lay.vg);
  o.vd = reinterpret_cast<float*>(base + lay.vd);
  o.vt = reinterpret_cast<float*>(base + lay.vt);
  o.red = reinterpret_cast<float*>(base + lay.red);
  o.gc = reinterpret_cast<float4*>(base + lay.gc);
  return o;
}

__device__ __forceinline__ float traj_wave_sum(float v) { return wave_sum_dpp(v); }

// Sum of NV values over the workgroup; every thread gets the totals.
template <int NV>
__device__ __forceinline__ void block_sum(float (&v)[NV], float* red) {
  static_assert(NV <= kRedSlots, "too many values");
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = traj_wave_sum(v[i]);
  const int wave = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x % kWave) == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[wave * kRedSlots + i] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < kMaxTrajWaves; ++w)
      if (w * kWave < static_cast<int>(blockDim.x)) s += red[w * kRedSlots + i];
    v[i] = s;
  }
}

__device__ __forceinline__ float block_sum1(float v, float* red) {
  float a[1] = {v};
  block_sum<1>(a, red);
  return a[0];
}

// Collision activation f(d) and df/dd.
__device__ __forceinline__ float traj_act(float d, float m, int mode, float& df) {
  if (mode == 0) {
    const float h = m - d;
    if (h > 0.f) { df = -2.f * h; return h * h; }
    df = 0.f;
    return 0.f;
  }
  const float dm = fminf(d, m);
  if (dm < 0.f) { df = -1.f; return 0.5f * m - dm; }
  if (d < m) {
    const float inv = 1.f / (m + 1e-6f);
    df = (d - m) * inv;
    return 0.5f * inv * (d - m) * (d - m);
  }
  df = 0.f;
  return 0.f;
}

template <int MAXN>
__device__ __forceinline__ Pose node_pose(const Pose (&P)[MAXN], int n) {
  Pose r = pose_identity();
#pragma unroll
  for (int i = 0; i < MAXN; ++i)
    if (n == i) r = P[i];
  return r;
}

static_assert(kGroup == 8, "group sums are 8-lane DPP reductions");
__device__ __forceinline__ float group_sum(float v) { return group8_sum_dpp(v); }

// Node kinematics of one waypoint, computed by every lane of the group:
// poses P, and (for the adjoint) each node's twist in world axes (w, v).
template <int MAXN, bool GRAD>
struct WaypointNodes {
  Pose P[MAXN];
  float w[GRAD ? MAXN : 1][3], v[GRAD ? MAXN : 1][3];
};

template <int MAXN, bool GRAD>
__device__ __forceinline__ void waypoint_fk(const TrajModelDev& md, const TrajLds& L, const float (&q)[MAXN],
                                            WaypointNodes<MAXN, GRAD>& K) {
  const int N = md.n_nodes;
#pragma unroll
  for (int n = 0; n < MAXN; ++n) {
    K.P[n] = pose_identity();
    if (n < N) {
      const float* r = L.nodes + n * kNodeWords;
      const int src = __builtin_amdgcn_readfirstlane(static_cast<int>(r[15]));
      const int par = __builtin_amdgcn_readfirstlane(static_cast<int>(r[16]));
      const int kind = __builtin_amdgcn_readfirstlane(static_cast<int>(r[17]));
      float qs = 0.f;
#pragma unroll
      for (int a = 0; a < MAXN; ++a)
        if (src == a) qs = q[a];
      const float qj = qs * r[13] + r[14];
      // A pure rotation (kind 1, non-zero axis) takes the closed form with
      // the unit axis r[10..12] / |ang| and the hardware sine and cosine.
      const float na = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r[18])));
      const Pose tl = compose(load_pose(r), (kind == 1 && na > 0.f)
          ? joint_exp<true>(4, qj, na, 0.f, 0.f, r[10] * r[19], r[11] * r[19], r[12] * r[19])
          : joint_exp(kind, qj, r[7], r[8], r[9], r[10], r[11], r[12]));
      Pose pp = pose_identity();
      bool has_parent = false;
#pragma unroll
      for (int m = 0; m < n; ++m)
        if (par == m) { pp = K.P[m]; has_parent = true; }
      K.P[n] = has_parent ? compose(pp, tl) : tl;
      if constexpr (GRAD) {
        quat_rotate(K.P[n].w, K.P[n].x, K.P[n].y, K.P[n].z, r[10], r[11], r[12], K.w[n][0], K.w[n][1], K.w[n][2]);
        quat_rotate(K.P[n].w, K.P[n].x, K.P[n].y, K.P[n].z, r[7], r[8], r[9], K.v[n][0], K.v[n][1], K.v[n][2]);
      }
    }
  }
}

// dcost/dq_node += g . dc/dq_node for a point c carried by a node whose
// ancestor set (itself included) is `mask`: dc/dq_a = w_a x (c - p_a) + v_a.
template <int MAXN, bool GRAD>
__device__ __forceinline__ void pull_point(const WaypointNodes<MAXN, GRAD>& K, int mask, float cx, float cy,
                                           float cz, float gx, float gy, float gz, float (&gn)[MAXN]) {
  if constexpr (GRAD) {
#pragma unroll
  for (int a = 0; a < MAXN; ++a) {
    if (mask & (1 << a)) {
      const float rx = cx - K.P[a].tx, ry = cy - K.P[a].ty, rz = cz - K.P[a].tz;
      const float dx = K.w[a][1] * rz - K.w[a][2] * ry + K.v[a][0];
      const float dy = K.w[a][2] * rx - K.w[a][0] * rz + K.v[a][1];
      const float dz = K.w[a][0] * ry - K.w[a][1] * rx + K.v[a][2];
      gn[a] += gx * dx + gy * dy + gz * dz;
    }
  }
  }
}

// World-frame sphere centres (radius in w) and link bounding spheres of the
// waypoint with node poses K.P, this lane's share, into the group's LDS slot
// gc = [K spheres | L link bounds].  Links are visited in table order and a
// node's rotation is rebuilt only when the node changes (the host sorts links
// by node).
template <int MAXN, bool GRAD>
__device__ __forceinline__ void waypoint_centres(const TrajModelDev& md, const TrajLds& L,
                                                 const WaypointNodes<MAXN, GRAD>& K, float4* gc) {
  const int g = threadIdx.x % kGroup;
  const int Ks = md.n_spheres;
  int prev = -2;
  Pose T = pose_identity();
  Rot3 R = quat_to_rot(1.f, 0.f, 0.f, 0.f);
  for (int l = 0; l < md.n_links; ++l) {
    const int node = __builtin_amdgcn_readfirstlane(L.link_node[l]);
    if (node != prev) {
      T = node_pose<MAXN>(K.P, node);
      R = quat_to_rot(T.w, T.x, T.y, T.z);
      prev = node;
    }
    const int k0 = L.link_sph[l], k1 = L.link_sph[l + 1];
    for (int s = k0 + g; s < k1; s += kGroup) {
      const float4 sp = L.sph[s];
      gc[s] = make_float4(R.m00 * sp.x + R.m01 * sp.y + R.m02 * sp.z + T.tx,
                          R.m10 * sp.x + R.m11 * sp.y + R.m12 * sp.z + T.ty,
                          R.m20 * sp.x + R.m21 * sp.y + R.m22 * sp.z + T.tz, sp.w);
    }
    if (g == l % kGroup) {
      const float4 b = L.lb[l];
      gc[Ks + l] = make_float4(R.m00 * b.x + R.m01 * b.y + R.m02 * b.z + T.tx,
                               R.m10 * b.x + R.m11 * b.y + R.m12 * b.z + T.ty,
                               R.m20 * b.x + R.m21 * b.y + R.m22 * b.z + T.tz, b.w);
    }
  }
}

// Collision cost of one waypoint, evaluated by the kGroup lanes of a group
// (lane g = threadIdx.x % kGroup) from the centres in gc, and with GRAD the
// group's dcost/dq_node (per node, before the mimic map).  Every lane of the
// group returns the group totals.
template <int MAXN, bool GRAD>
__device__ __forceinline__ float group_waypoint_coll(const TrajModelDev& md, const TrajLds& L, const WaypointNodes<MAXN, GRAD>& K,
                                     const float4* gc, float (&gn)[MAXN]) {
  const int g = threadIdx.x % kGroup;
  const float margin = md.margin;
  const int mode = md.mode;
  const float e = 1e-6f;
  const int Ks = md.n_spheres;
  float cost = 0.f;
#pragma unroll
  for (int a = 0; a < MAXN; ++a) gn[a] = 0.f;

  // World: spheres s = g, g + kGroup, ... of the waypoint.
  for (int s = g; s < Ks; s += kGroup) {
    const float4 c = gc[s];
    const int level = grid_level(md.grid, c.w);
    const int* items = nullptr;
    int count = 0;
    if (level >= 0) grid_cell_list(md.grid, level, c.x, c.y, c.z, items, count);
    const bool brute = level < 0;  // no level serves this radius: every world sphere
    const int total = brute ? md.n_world : count;
    if (total == 0) continue;
    const int l = L.sph_link[s];
    const int k0 = L.link_sph[l], k1 = L.link_sph[l + 1];
    const int node = L.link_node[l];
    const int mask = node >= 0 ? L.anc[node] : 0;
    for (int i = 0; i < total; ++i) {
      const int m = brute ? i : items[i];
      if (m < 0 || m >= md.n_world) continue;
      const float4 o = L.world[m];
      const float dx = o.x - c.x + e, dy = o.y - c.y + e, dz = o.z - c.z + e;
      const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
      const float d = dist - c.w - o.w;
      if (!(d < margin)) continue;
      // The (link, world sphere) term is owned by the link's lowest-index
      // sphere within the margin; the owner takes the minimum over the link.
      bool owner = true;
      for (int s2 = k0; s2 < s; ++s2) {
        const float4 c2 = gc[s2];
        const float ex = o.x - c2.x + e, ey = o.y - c2.y + e, ez = o.z - c2.z + e;
        if (sqrtf(ex * ex + ey * ey + ez * ez) - c2.w - o.w < margin) { owner = false; break; }
      }
      if (!owner) continue;
      float bd = d, bdist = dist, bx = dx, by = dy, bz = dz, bcx = c.x, bcy = c.y, bcz = c.z;
      for (int s2 = s + 1; s2 < k1; ++s2) {
        const float4 c2 = gc[s2];
        const float ex = o.x - c2.x + e, ey = o.y - c2.y + e, ez = o.z - c2.z + e;
        const float dd = sqrtf(ex * ex + ey * ey + ez * ez);
        if (dd - c2.w - o.w < bd) {
          bd = dd - c2.w - o.w; bdist = dd; bx = ex; by = ey; bz = ez; bcx = c2.x; bcy = c2.y; bcz = c2.z;
        }
      }
      float df;
      cost += traj_act(bd, margin, mode, df);
      if (GRAD && df != 0.f) {
        const float k = -df / bdist;
        pull_point<MAXN, GRAD>(K, mask, bcx, bcy, bcz, k * bx, k * by, k * bz, gn);
      }
    }
  }

  // Self: active pairs split over the group's lanes, bounding-sphere culled.
  for (int p = g; p < md.n_pairs; p += kGroup) {
    const int q0 = L.pair_start[p], q1 = L.pair_start[p + 1];
    if (q0 >= q1) continue;
    const int li = L.pair_links[2 * p], lj = L.pair_links[2 * p + 1];
    {
      const float4 bi = gc[Ks + li], bj = gc[Ks + lj];
      const float dd = sqrtf((bj.x - bi.x) * (bj.x - bi.x) + (bj.y - bi.y) * (bj.y - bi.y) +
                             (bj.z - bi.z) * (bj.z - bi.z));
      if (dd - bi.w - bj.w >= margin + 1e-4f) continue;
    }
    float best = INFINITY, bdist = 1.f, bdx = 0.f, bdy = 0.f, bdz = 0.f;
    float4 bci = make_float4(0.f, 0.f, 0.f, 0.f), bcj = bci;
    for (int it = q0; it < q1; ++it) {
      const float4 ci = gc[L.pair_items[2 * it]];
      const float4 cj = gc[L.pair_items[2 * it + 1]];
      const float dx = cj.x - ci.x + e, dy = cj.y - ci.y + e, dz = cj.z - ci.z + e;
      const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
      const float d = dist - ci.w - cj.w;
      if (d < best) { best = d; bci = ci; bcj = cj; bdx = dx; bdy = dy; bdz = dz; bdist = dist; }
    }
    if (best < margin) {
      float df;
      cost += traj_act(best, margin, mode, df);
      if (GRAD && df != 0.f) {
        const float k = df / bdist;
        const int ni = L.link_node[li], nj = L.link_node[lj];
        const int mi = ni >= 0 ? L.anc[ni] : 0, mj = nj >= 0 ? L.anc[nj] : 0;
        pull_point<MAXN, GRAD>(K, mj, bcj.x, bcj.y, bcj.z, k * bdx, k * bdy, k * bdz, gn);
        pull_point<MAXN, GRAD>(K, mi, bci.x, bci.y, bci.z, -k * bdx, -k * bdy, -k * bdz, gn);
      }
    }
  }

  cost = group_sum(cost);
  if (GRAD) {
#pragma unroll
    for (int a = 0; a < MAXN; ++a) gn[a] = group_sum(gn[a]);
  }
  return cost;
}

// dcost/dq (actuated) from the per-node sums: q_node = mul q_src + off.
template <int MAXN>
__device__ __forceinline__ void node_to_actuated(const TrajLds& L, int N, const float (&gn)[MAXN],
                                                 float (&gq)[MAXN]) {
#pragma unroll
  for (int a = 0; a < MAXN; ++a) gq[a] = 0.f;
#pragma unroll
  for (int n = 0; n < MAXN; ++n) {
    if (n < N) {
      const float* r = L.nodes + n * kNodeWords;
      const int src = __builtin_amdgcn_readfirstlane(static_cast<int>(r[15]));
      const float val = r[13] * gn[n];
#pragma unroll
      for (int a = 0; a < MAXN; ++a)
        if (src == a) gq[a] += val;
    }
  }
}

// One waypoint per group for `n_tasks` tasks; task(i, q) fills the
// configuration of task i, done(i, cost, gq) consumes the group totals (every
// lane of the group calls it).  All threads run the same number of rounds, so
// the workgroup barriers inside are uniform.
template <int MAXN, bool GRAD, typename TaskFn, typename DoneFn>
__device__ __forceinline__ void for_each_waypoint(const TrajModelDev& md, const TrajLds& L, int n_tasks,
                                                  TaskFn task, DoneFn done) {
  const int grp = threadIdx.x / kGroup;
  const int N = md.n_nodes;
  float4* gc = L.gc + grp * (md.n_spheres + md.n_links);
  const int n_groups = blockDim.x / kGroup;
  const int rounds = (n_tasks + n_groups - 1) / n_groups;
  for (int rd = 0; rd < rounds; ++rd) {
    const int i = rd * n_groups + grp;
    const bool active = i < n_tasks;
    WaypointNodes<MAXN, GRAD> K;
    if (active) {
      float q[MAXN];
      task(i, q);
      waypoint_fk<MAXN, GRAD>(md, L, q, K);
      waypoint_centres<MAXN, GRAD>(md, L, K, gc);
    }
    __syncthreads();
    if (active) {
      float gn[MAXN], gq[MAXN];
      const float c = group_waypoint_coll<MAXN, GRAD>(md, L, K, gc, gn);
      if (GRAD) node_to_actuated<MAXN>(L, N, gn, gq);
      done(i, c, gq);
    }
    __syncthreads();
  }
}

// Cost components {smooth, limits, coll} of the trajectory X (T, A) in LDS, all
// threads receiving the totals; with GRAD, L.vg = gradient of
// w_smooth smooth + w_limits limits + w_coll coll (endpoint rows zeroed when
// mask_ends).  Leaves the accelerations of x in L.vt; uses L.vd as scratch.
template <int MAXN, bool GRAD>
__device__ __forceinline__ void traj_value_grad(const TrajModelDev& md, const TrajLds& L, const float* X, int T,
                                                const TrajWeights& w, float w_coll, bool mask_ends,
                                                float (&comps)[3]) {
  const int A = md.n_act;
  const int n = T * A;
  const int tid = threadIdx.x;
  float coll = 0.f, lim = 0.f, sm = 0.f;
  const bool leader = (tid % kGroup) == 0;
  for_each_waypoint<MAXN, GRAD>(
      md, L, T,
      [&](int t, float (&q)[MAXN]) {
#pragma unroll
        for (int a = 0; a < MAXN; ++a) q[a] = (a < A) ? X[t * A + a] : 0.f;
      },
      [&](int t, float c, const float (&gq)[MAXN]) {
        if (!leader) return;
        coll += c;
        if (GRAD) {
#pragma unroll
          for (int a = 0; a < MAXN; ++a)
            if (a < A) L.vg[t * A + a] = w_coll * gq[a];
        }
      });
  const int na = T - 4;
  for (int i = tid; i < n; i += blockDim.x) {
    const int a = i - quot(i, A, md.a_magic) * A;
    const float xv = X[i];
    const float v = fmaxf(xv - L.hi[a], 0.f) + fmaxf(L.lo[a] - xv, 0.f);
    lim += v * v;
    if (GRAD)
      L.vg[i] += w.w_limits * 2.f * v * ((xv > L.hi[a] ? 1.f : 0.f) - (L.lo[a] > xv ? 1.f : 0.f));
    if (i + A < n) {
      const float dv = X[i + A] - xv;
      sm += w.w_vel * dv * dv;
    }
    if (i < na * A) {
      const float acc = (-xv + 16.f * X[i + A] - 30.f * X[i + 2 * A] + 16.f * X[i + 3 * A] -
                         X[i + 4 * A]) / 12.f;
      L.vt[i] = acc;
      sm += w.w_acc * acc * acc;
    }
  }
  __syncthreads();
  for (int i = tid; i < na * A; i += blockDim.x) {
    const int t = quot(i, A, md.a_magic);
    const float acc = L.vt[i];
    if (t + 1 < na) {
      const float jk = L.vt[i + A] - acc;
      sm += w.w_jerk * jk * jk;
    }
    if (GRAD) {
      float ga = 2.f * w.w_acc * acc;
      if (t + 1 < na) ga -= 2.f * w.w_jerk * (L.vt[i + A] - acc);
      if (t >= 1) ga += 2.f * w.w_jerk * (acc - L.vt[i - A]);
      L.vd[i] = ga;
    }
  }
  if (GRAD) {
    __syncthreads();
    const float c5[5] = {-1.f / 12.f, 16.f / 12.f, -30.f / 12.f, 16.f / 12.f, -1.f / 12.f};
    for (int i = tid; i < n; i += blockDim.x) {
      const int t = quot(i, A, md.a_magic);
      float gs = 0.f;
      if (t >= 1) gs += 2.f * w.w_vel * (X[i] - X[i - A]);
      if (t + 1 < T) gs -= 2.f * w.w_vel * (X[i + A] - X[i]);
#pragma unroll
      for (int o = 0; o < 5; ++o) {
        const int tt = t - o;
        if (tt >= 0 && tt < na) gs += L.vd[i - o * A] * c5[o];
      }
      float g = L.vg[i] + w.w_smooth * gs;
      if (mask_ends && (t == 0 || t == T - 1)) g = 0.f;
      L.vg[i] = g;
    }
  }
  float v3[3] = {sm, lim, coll};
  block_sum<3>(v3, L.red);
  comps[0] = v3[0];
  comps[1] = v3[1];
  comps[2] = v3[2];
}

template <int MAXN, bool GRAD>
__device__ __forceinline__ void traj_value_grad(const TrajModelDev& md, const TrajLds& L, int T, const TrajWeights& w,
                                                float w_coll, bool mask_ends, float (&comps)[3]) {
  traj_value_grad<MAXN, GRAD>(md, L, L.vx, T, w, w_coll, mask_ends, comps);
}

// Line search of _lbfgs_direct_solve on x + alpha_k d, alpha in {1, .5, .25,
// .1, .025}: the first step whose cost is below `thresh`, else the cheapest
// (first index on ties).  The points are evaluated in that order and the
// search stops at the first acceptable one, which picks the same step as
// evaluating all five.  Returns the step index and its cost components.
template <int MAXN>
__device__ __forceinline__ int traj_line_search(const TrajModelDev& md, const TrajLds& L, int T, const TrajWeights& w,
                                float w_coll, float thresh, float& p_sm, float& p_lim, float& p_col) {
  const int A = md.n_act;
  const int n = T * A;
  const int tid = threadIdx.x;
  const int na = T - 4;
  // Smoothness of x + a d is s0 + 2 a s1 + a^2 s2; limits per step.
  float v[3 + kNumAlphas];
#pragma unroll
  for (int k = 0; k < 3 + kNumAlphas; ++k) v[k] = 0.f;
  for (int i = tid; i < n; i += blockDim.x) {
    const int t = quot(i, A, md.a_magic);
    const int a = i - t * A;
    const float xv = L.vx[i], dv = L.vd[i];
#pragma unroll
    for (int k = 0; k < kNumAlphas; ++k) {
      const float y = xv + traj_alpha(k) * dv;
      const float u = fmaxf(y - L.hi[a], 0.f) + fmaxf(L.lo[a] - y, 0.f);
      v[3 + k] += u * u;
    }
    if (t + 1 < T) {
      const float vx = L.vx[i + A] - xv, vd = L.vd[i + A] - dv;
      v[0] += w.w_vel * vx * vx; v[1] += w.w_vel * vx * vd; v[2] += w.w_vel * vd * vd;
    }
    if (t < na) {
      const float ax = L.vt[i];
      const float ad = (-dv + 16.f * L.vd[i + A] - 30.f * L.vd[i + 2 * A] + 16.f * L.vd[i + 3 * A] -
                        L.vd[i + 4 * A]) / 12.f;
      v[0] += w.w_acc * ax * ax; v[1] += w.w_acc * ax * ad; v[2] += w.w_acc * ad * ad;
      if (t + 1 < na) {
        const float jx = L.vt[i + A] - ax;
        const float ad1 = (-L.vd[i + A] + 16.f * L.vd[i + 2 * A] - 30.f * L.vd[i + 3 * A] +
                           16.f * L.vd[i + 4 * A] - L.vd[i + 5 * A]) / 12.f;
        const float jd = ad1 - ad;
        v[0] += w.w_jerk * jx * jx; v[1] += w.w_jerk * jx * jd; v[2] += w.w_jerk * jd * jd;
      }
    }
  }
  block_sum<3 + kNumAlphas>(v, L.red);
  const bool leader = (tid % kGroup) == 0;
  int best = -1;
  float best_tc = INFINITY;
  for (int k = 0; k < kNumAlphas; ++k) {
    const float al = traj_alpha(k);
    float ck = 0.f;
    for_each_waypoint<MAXN, false>(
        md, L, T,
        [&](int t, float (&q)[MAXN]) {
#pragma unroll
          for (int a = 0; a < MAXN; ++a) q[a] = (a < A) ? L.vx[t * A + a] + al * L.vd[t * A + a] : 0.f;
        },
        [&](int, float c, const float (&)[MAXN]) {
          if (leader) ck += c;
        });
    ck = block_sum1(ck, L.red);
    const float sm = v[0] + 2.f * al * v[1] + al * al * v[2];
    float lim = 0.f;
#pragma unroll
    for (int kk = 0; kk < kNumAlphas; ++kk)
      if (kk == k) lim = v[3 + kk];
    const float tc = w.w_smooth * sm + w.w_limits * lim + w_coll * ck;
    if (best < 0 || tc < best_tc) {
      best = k; best_tc = tc; p_sm = sm; p_lim = lim; p_col = ck;
    }
    if (tc < thresh) {
      p_sm = sm; p_lim = lim; p_col = ck;
      return k;
    }
  }
  return best;
}

__device__ __forceinline__ float sel_hist(const float (&v)[kMaxHist], int i) {
  float r = 0.f;
#pragma unroll
  for (int h = 0; h < kMaxHist; ++h)
    if (h == i) r = v[h];
  return r;
}

struct CostArgs {
  TrajModelDev md;
  const float* x;
  const long long* start;
  const int* subset;  // block b works on trajectory subset[b] (all trajectories when null)
  int n_traj, max_len;
  long long n_waypoints;
  TrajWeights w;
  float w_coll;
  float* out_comps;   // (B, 3)
  float* out_grad;    // (W, A) or null
};

template <int MAXN>
__global__ __launch_bounds__(kMaxTrajThreads) void traj_cost_kernel(CostArgs a) {
  extern __shared__ float4 smem4[];
  const TrajLds L = stage_traj_model(a.md, a.max_len, reinterpret_cast<char*>(smem4));
  const int b = a.subset ? a.subset[blockIdx.x] : static_cast<int>(blockIdx.x);
  if (b < 0 || b >= a.n_traj) return;  // uniform per block
  const long long w0 = a.start[b], w1 = a.start[b + 1];
  if (w0 < 0 || w1 < w0 || w1 > a.n_waypoints || w1 - w0 > a.max_len) return;  // uniform per block
  const int T = static_cast<int>(w1 - w0);
  const int A = a.md.n_act;
  for (int i = threadIdx.x; i < T * A; i += blockDim.x) L.vx[i] = a.x[w0 * A + i];
  __syncthreads();
  float comps[3];
  if (a.out_grad != nullptr) {
    traj_value_grad<MAXN, true>(a.md, L, T, a.w, a.w_coll, false, comps);
    for (int i = threadIdx.x; i < T * A; i += blockDim.x) a.out_grad[w0 * A + i] = L.vg[i];
  } else {
    traj_value_grad<MAXN, false>(a.md, L, T, a.w, a.w_coll, false, comps);
  }
  if (threadIdx.x == 0) {
    a.out_comps[3 * b] = comps[0];
    a.out_comps[3 * b + 1] = comps[1];
    a.out_comps[3 * b + 2] = comps[2];
  }
}

struct LbfgsArgs {
  TrajModelDev md;
  const float* x_init;
  const long long* start;
  const int* subset;  // block b works on trajectory subset[b] (all trajectories when null)
  int n_traj, max_len;
  long long n_waypoints;
  TrajWeights w;
  float w_coll0, w_coll_max, penalty_scale;
  int escalation, n_iters, m_hist, it0, it1;
  float* xs;      // (W, A) current iterate
  float* xp;      // (W, A) previous iterate
  float* gp;      // (W, A) previous gradient
  float* sb;      // (m, W, A) s history
  float* yb;      // (m, W, A) y history
  float* sc;      // (B, kScalarWords) w_coll, best_cost, m_used, newest, rho[m]
  float* out_best;  // (W, A)
  float* out_cost;  // (B)
};

template <int MAXN>
__global__ __launch_bounds__(kMaxTrajThreads) void traj_lbfgs_kernel(LbfgsArgs a) {
  extern __shared__ float4 smem4[];
  const TrajLds L = stage_traj_model(a.md, a.max_len, reinterpret_cast<char*>(smem4));
  const int b = a.subset ? a.subset[blockIdx.x] : static_cast<int>(blockIdx.x);
  if (b < 0 || b >= a.n_traj) return;  // uniform per block
  const long long w0 = a.start[b], w1 = a.start[b + 1];
  if (w0 < 0 || w1 < w0 || w1 > a.n_waypoints || w1 - w0 > a.max_len) return;  // uniform per block
  const int T = static_cast<int>(w1 - w0);
  const int A = a.md.n_act;
  const int n = T * A;
  const int tid = threadIdx.x;
  const int m = a.m_hist;
  const long long off = w0 * A;
  const long long WA = a.n_waypoints * A;
  float* xs = a.xs + off;
  float* xp = a.xp + off;
  float* gp = a.gp + off;
  float* best = a.out_best + off;
  float* sb = a.sb;
  float* yb = a.yb;
  float* sc = a.sc + static_cast<long long>(b) * kScalarWords;

  float w_coll, best_cost;
  int m_used, newest;
  float rho[kMaxHist];
  if (a.it0 == 0) {
    for (int i = tid; i < n; i += blockDim.x) L.vx[i] = a.x_init[off + i];
    w_coll = a.w_coll0;
    best_cost = INFINITY;
    m_used = 0;
    newest = 0;
#pragma unroll
    for (int h = 0; h < kMaxHist; ++h) rho[h] = 0.f;
  } else {
    for (int i = tid; i < n; i += blockDim.x) L.vx[i] = xs[i];
    w_coll = sc[0];
    best_cost = sc[1];
    m_used = static_cast<int>(sc[2]);
    newest = static_cast<int>(sc[3]);
#pragma unroll
    for (int h = 0; h < kMaxHist; ++h) rho[h] = (h < m) ? sc[4 + h] : 0.f;
  }
  __syncthreads();

  for (int k = a.it0; k < a.it1; ++k) {
    if (k > 0 && k % a.escalation == 0) w_coll = fminf(w_coll * a.penalty_scale, a.w_coll_max);
    float comps[3];
    traj_value_grad<MAXN, true>(a.md, L, T, a.w, w_coll, true, comps);
    const float cost = a.w.w_smooth * comps[0] + a.w.w_limits * comps[1] + w_coll * comps[2];
    if (k == 0) {
      for (int i = tid; i < n; i += blockDim.x) {
        xp[i] = L.vx[i];
        gp[i] = L.vg[i];
        best[i] = L.vx[i];
      }
      best_cost = cost;
    }
    // Curvature pair of the last step.
    float syy[2] = {0.f, 0.f};
    for (int i = tid; i < n; i += blockDim.x) {
      const float s = L.vx[i] - xp[i], y = L.vg[i] - gp[i];
      syy[0] += s * y;
      syy[1] += y * y;
    }
    block_sum<2>(syy, L.red);
    if (syy[0] > 1e-10f * syy[1] + 1e-30f && k > 0) {
      newest = (newest + 1) % m;
      for (int i = tid; i < n; i += blockDim.x) {
        sb[newest * WA + off + i] = L.vx[i] - xp[i];
        yb[newest * WA + off + i] = L.vg[i] - gp[i];
      }
      const float r = 1.f / (syy[0] + 1e-30f);
#pragma unroll
      for (int h = 0; h < kMaxHist; ++h)
        if (h == newest) rho[h] = r;
      if (m_used < m) ++m_used;
    }
    // Search direction into L.vd.
    if (m_used > 0) {
      float alph[kMaxHist];
#pragma unroll
      for (int h = 0; h < kMaxHist; ++h) alph[h] = 0.f;
      for (int i = tid; i < n; i += blockDim.x) L.vd[i] = L.vg[i];
      for (int j = 0; j < m_used; ++j) {
        const int idx = (newest - j + m) % m;
        const float* s = sb + idx * WA + off;
        const float* y = yb + idx * WA + off;
        float sq = 0.f;
        for (int i = tid; i < n; i += blockDim.x) sq += s[i] * L.vd[i];
        const float al = sel_hist(rho, idx) * block_sum1(sq, L.red);
#pragma unroll
        for (int h = 0; h < kMaxHist; ++h)
          if (h == idx) alph[h] = al;
        for (int i = tid; i < n; i += blockDim.x) L.vd[i] -= al * y[i];
      }
      float g2[2] = {0.f, 0.f};
      {
        const float* s = sb + newest * WA + off;
        const float* y = yb + newest * WA + off;
        for (int i = tid; i < n; i += blockDim.x) {
          g2[0] += s[i] * y[i];
          g2[1] += y[i] * y[i];
        }
      }
      block_sum<2>(g2, L.red);
      const float gamma = g2[0] / (g2[1] + 1e-18f);
      for (int i = tid; i < n; i += blockDim.x) L.vd[i] *= gamma;
      for (int st = 0; st < m_used; ++st) {
        const int idx = (newest - m_used + 1 + st + 2 * m) % m;
        const float* s = sb + idx * WA + off;
        const float* y = yb + idx * WA + off;
        float yr = 0.f;
        for (int i = tid; i < n; i += blockDim.x) yr += y[i] * L.vd[i];
        const float beta = sel_hist(rho, idx) * block_sum1(yr, L.red);
        const float coef = sel_hist(alph, idx) - beta;
        for (int i = tid; i < n; i += blockDim.x) L.vd[i] += s[i] * coef;
      }
      for (int i = tid; i < n; i += blockDim.x) L.vd[i] = -L.vd[i];
    } else {
      float g2 = 0.f;
      for (int i = tid; i < n; i += blockDim.x) g2 += L.vg[i] * L.vg[i];
      const float inv = 1.f / (sqrtf(block_sum1(g2, L.red)) + 1e-18f);
      for (int i = tid; i < n; i += blockDim.x) L.vd[i] = -L.vg[i] * inv;
    }
    for (int i = tid; i < n; i += blockDim.x) {
      const int t = quot(i, A, a.md.a_magic);
      if (t == 0 || t == T - 1) L.vd[i] = 0.f;
    }
    __syncthreads();
    float p_sm = 0.f, p_lim = 0.f, p_col = 0.f;
    const int pick = traj_line_search<MAXN>(a.md, L, T, a.w, w_coll, cost * (1.f - 1e-4f), p_sm, p_lim, p_col);
    const float al = traj_alpha(pick);
    for (int i = tid; i < n; i += blockDim.x) {
      xp[i] = L.vx[i];
      gp[i] = L.vg[i];
      L.vx[i] += al * L.vd[i];
    }
    const float eval = a.w.w_smooth * p_sm + a.w.w_limits * p_lim + a.w_coll_max * p_col;
    if (eval < best_cost) {
      best_cost = eval;
      for (int i = tid; i < n; i += blockDim.x) best[i] = L.vx[i];
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += blockDim.x) xs[i] = L.vx[i];
  if (tid == 0) {
    sc[0] = w_coll;
    sc[1] = best_cost;
    sc[2] = static_cast<float>(m_used);
    sc[3] = static_cast<float>(newest);
#pragma unroll
    for (int h = 0; h < kMaxHist; ++h)
      if (h < m) sc[4 + h] = rho[h];
    if (a.it1 == a.n_iters) a.out_cost[b] = best_cost;
  }
}

// CHOMP (_chomp_optimization.py:208-330): per iteration the gradient of the
// cost at the current collision weight, clipped to a global norm, endpoint
// rows zeroed; the covariant direction -M^-1 g, M the interior second-
// difference precision operator per joint, by truncated CG (each joint's
// system on one wave, wave-level sums); the six trials
// x + clip(step alpha d, +-max_delta) evaluated in full and the cheapest (first
// on ties) kept when it beats the current cost by min_cost_improve; stall count,
// early stop and collision-weight continuation.  The alpha = 0 trial is x itself,
// whose cost is the current one.  LDS use: vx iterate, vg gradient / CG search
// vector / trial point, vd CG solution (the direction is its negation), vt CG
// residual / accelerations.
struct ChompArgs {
  TrajModelDev md;
  const float* x_init;
  const long long* start;
  int n_traj, max_len;
  long long n_waypoints;
  TrajWeights w;  // w_vel = 0: CHOMP's smoothness has no velocity term
  float step, w_coll0, w_coll_max, scale, reg, cg_tol, clip, max_delta, min_improve;
  int n_iters, covariant, n_cg, patience;
  float* out_traj;   // (W, A)
  float* out_cost;   // (B)
  float* out_trace;  // (B, n_iters, 3) or null
};

__device__ __forceinline__ float chomp_alpha(int k) {
  return k == 0 ? 1.f : (k == 1 ? 0.5f : (k == 2 ? 0.25f : (k == 3 ? 0.1f : (k == 4 ? 0.05f : 0.f))));
}

// M p at element i = t * A + a of a (T, A) array whose rows 0 and T - 1 are zero.
__device__ __forceinline__ float chomp_mp(const float* p, int i, int t, int T, int A, float reg) {
  float v = (6.f + reg) * p[i] - 4.f * (p[i - A] + p[i + A]);
  if (t >= 2) v += p[i - 2 * A];
  if (t + 2 <= T - 1) v += p[i + 2 * A];
  return v;
}

template <int MAXN>
__device__ __forceinline__ void add_at(float (&v)[MAXN], int a, float x) {
#pragma unroll
  for (int k = 0; k < MAXN; ++k)
    if (k == a) v[k] += x;
}

template <int MAXN>
__device__ __forceinline__ float get_at(const float (&v)[MAXN], int a) {
  float r = 0.f;
#pragma unroll
  for (int k = 0; k < MAXN; ++k)
    if (k == a) r = v[k];
  return r;
}

template <int MAXN>
__global__ __launch_bounds__(kMaxTrajThreads) void traj_chomp_kernel(ChompArgs a) {
  extern __shared__ float4 smem4[];
  const TrajLds L = stage_traj_model(a.md, a.max_len, reinterpret_cast<char*>(smem4));
  const int b = blockIdx.x;
  if (b >= a.n_traj) return;  // uniform per block
  const long long w0 = a.start[b], w1 = a.start[b + 1];
  if (w0 < 0 || w1 < w0 || w1 > a.n_waypoints || w1 - w0 > a.max_len) return;  // uniform per block
  const int T = static_cast<int>(w1 - w0);
  const int A = a.md.n_act;
  const int n = T * A;
  const int tid = threadIdx.x;
  const long long off = w0 * A;
  for (int i = tid; i < n; i += blockDim.x) L.vx[i] = a.x_init[off + i];
  __syncthreads();
  float w_coll = a.w_coll0;
  int stall = 0;
  bool active = true;
  for (int k = 0; k < a.n_iters; ++k) {
    float* tr = a.out_trace ? a.out_trace + (static_cast<long long>(b) * a.n_iters + k) * 3 : nullptr;
    if (!active) {
      if (tr && tid == 0) tr[0] = tr[1] = tr[2] = -1.f;
      continue;
    }
    float comps[3];
    traj_value_grad<MAXN, true>(a.md, L, L.vx, T, a.w, w_coll, false, comps);
    const float cost = a.w.w_smooth * comps[0] + a.w.w_limits * comps[1] + w_coll * comps[2];
    float g2 = 0.f;
    for (int i = tid; i < n; i += blockDim.x) g2 += L.vg[i] * L.vg[i];
    const float gn = sqrtf(block_sum1(g2, L.red));
    const float cs = (a.clip > 0.f && gn > a.clip) ? a.clip / (gn + 1e-12f) : 1.f;
    for (int i = tid; i < n; i += blockDim.x) {
      const int t = quot(i, A, a.md.a_magic);
      L.vg[i] = (t == 0 || t == T - 1) ? 0.f : L.vg[i] * cs;
    }
    __syncthreads();
    if (a.covariant && T > 2) {
      // CG on every joint's interior timeline: x = 0 (vd), r = b (vt), p = b (vg).
      // The joints' systems are independent, so each wave solves whole
      // joints (joint c on wave c mod W) with wave-level sums and wavefront
      // fences: no workgroup barrier inside the CG.
      const int n_waves = static_cast<int>(blockDim.x) / kWave;
      const int wv = tid / kWave, ln = tid % kWave;
      for (int c = wv; c < A; c += n_waves) {
        float rr = 0.f;
        for (int t = ln; t < T; t += kWave) {
          const int i = t * A + c;
          const float gv = L.vg[i];
          L.vd[i] = 0.f;
          L.vt[i] = gv;
          rr += gv * gv;
        }
        rr = traj_wave_sum(rr);
        const float tb = a.cg_tol * sqrtf(rr + 1e-18f) + 1e-12f;
        const float tol2 = tb * tb;
        wave_sync();
        for (int it = 0; it < a.n_cg; ++it) {
          float pap = 0.f;
          for (int t = 1 + ln; t < T - 1; t += kWave) {
            const int i = t * A + c;
            pap += L.vg[i] * chomp_mp(L.vg, i, t, T, A, a.reg);
          }
          const float al = rr / (traj_wave_sum(pap) + 1e-9f);
          float rn = 0.f;
          for (int t = 1 + ln; t < T - 1; t += kWave) {
            const int i = t * A + c;
            const float mp = chomp_mp(L.vg, i, t, T, A, a.reg);
            L.vd[i] += al * L.vg[i];
            const float r = L.vt[i] - al * mp;
            L.vt[i] = r;
            rn += r * r;
          }
          rn = traj_wave_sum(rn);
          wave_sync();  // every M p read of this step before the p update
          const float be = rn / (rr + 1e-9f);
          for (int t = 1 + ln; t < T - 1; t += kWave) {
            const int i = t * A + c;
            L.vg[i] = L.vt[i] + be * L.vg[i];
          }
          wave_sync();
          rr = rn;
          if (!(rn > tol2)) break;  // uniform in the wave
        }
      }
      __syncthreads();
      // The residual and search vectors of an inactive joint froze with it; the
      // solution is complete.  Direction = -x.
      for (int i = tid; i < n; i += blockDim.x) L.vd[i] = -L.vd[i];
    } else {
      for (int i = tid; i < n; i += blockDim.x) L.vd[i] = -L.vg[i];
    }
    __syncthreads();
    // Trials alpha_k, k = 0..4 (k = 5 is x itself, at the current cost).
    int pick = 5;
    float best = cost;
    for (int kk = 0; kk < 5; ++kk) {
      const float al = a.step * chomp_alpha(kk);
      for (int i = tid; i < n; i += blockDim.x)
        L.vg[i] = L.vx[i] + fminf(fmaxf(al * L.vd[i], -a.max_delta), a.max_delta);
      __syncthreads();
      float tc[3];
      traj_value_grad<MAXN, false>(a.md, L, L.vg, T, a.w, w_coll, false, tc);
      const float c = a.w.w_smooth * tc[0] + a.w.w_limits * tc[1] + w_coll * tc[2];
      if (c < best || (kk < pick && c == best)) {
        best = c;
        pick = kk;
      }
    }
    const bool improved = best < cost - a.min_improve;
    if (improved) {
      const float al = a.step * chomp_alpha(pick);
      for (int i = tid; i < n; i += blockDim.x)
        L.vx[i] += fminf(fmaxf(al * L.vd[i], -a.max_delta), a.max_delta);
    }
    stall = improved ? 0 : stall + 1;
    active = stall < a.patience;
    w_coll = fminf(w_coll * a.scale, a.w_coll_max);
    if (tr && tid == 0) {
      tr[0] = cost;
      tr[1] = static_cast<float>(pick);
      tr[2] = improved ? 1.f : 0.f;
    }
    __syncthreads();
  }
  float comps[3];
  traj_value_grad<MAXN, false>(a.md, L, L.vx, T, a.w, a.w_coll_max, false, comps);
  for (int i = tid; i < n; i += blockDim.x) a.out_traj[off + i] = L.vx[i];
  if (tid == 0) a.out_cost[b] = a.w.w_smooth * comps[0] + a.w.w_limits * comps[1] + a.w_coll_max * comps[2];
}

// ---------------------------------------------------------------- SCO
// _sco_optimization.py:224-503.  Linearisation: per waypoint and collision
// group (0 = the active self pairs, 1 + g = every (link, world sphere) pair of
// world group g) the smooth minimum dmin - tau log sum exp(-(d - dmin) / tau)
// of the JAX distances (closest same-slot pair / closest sphere of the link,
// +1e-6 bias) and its gradient: softmax weights times each distance's gradient
// at its closest sphere(s).  Links without spheres take no part (the
// reference's padded distances of about 1e9 have weight 0).  The world groups
// take one pass with a running minimum (weights rescaled when it drops), the
// self group two (the minimum, then the weights); the weights' distance
// gradients are accumulated per link as a wrench (F, c x F) and pulled back to
// the link's ancestor nodes once per link: dq_a += v_a . F + w_a . (tau - p_a x F).
constexpr int kMaxScoGroups = 8;

__device__ __forceinline__ float group8_min(float v) {
  v = fminf(v, __shfl_xor(v, 1, kWave));
  v = fminf(v, __shfl_xor(v, 2, kWave));
  return fminf(v, __shfl_xor(v, 4, kWave));
}

template <int MAXN>
__device__ __forceinline__ void pull_wrench(const WaypointNodes<MAXN, true>& K, int mask, const float (&F)[3],
                                            const float (&tq)[3], float (&gn)[MAXN]) {
#pragma unroll
  for (int a = 0; a < MAXN; ++a) {
    if (mask & (1 << a)) {
      const float px = K.P[a].tx, py = K.P[a].ty, pz = K.P[a].tz;
      const float mx = tq[0] - (py * F[2] - pz * F[1]);
      const float my = tq[1] - (pz * F[0] - px * F[2]);
      const float mz = tq[2] - (px * F[1] - py * F[0]);
      gn[a] += K.v[a][0] * F[0] + K.v[a][1] * F[1] + K.v[a][2] * F[2] + K.w[a][0] * mx + K.w[a][1] * my +
               K.w[a][2] * mz;
    }
  }
}

// |o_e - c| - r_c for robot sphere c and a world sphere centre o_e already
// shifted by e = 1e-6 per component (the world sphere's radius is subtracted
// once the closest robot sphere is known).  The hardware square root (at most
// 1 ulp) instead of the correctly rounded one: this runs for every (link
// sphere, world sphere) pair of every waypoint.
__device__ __forceinline__ float sphere_world_dist(const float4& c, const float4& oe) {
  const float dx = oe.x - c.x, dy = oe.y - c.y, dz = oe.z - c.z;
  return __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz) - c.w;  // v_sqrt_f32, <= 1 ulp
}

// Direction o_e - c from sphere c to the shifted obstacle centre, its length
// and the centre (the same expressions as sphere_world_dist, so the same bits).
__device__ __forceinline__ void closest_dir(const float4& c, const float4& oe, float& bx, float& by, float& bz,
                                            float& bn, float4& bc) {
  bx = oe.x - c.x; by = oe.y - c.y; bz = oe.z - c.z;
  bn = __builtin_amdgcn_sqrtf(bx * bx + by * by + bz * bz);
  bc = c;
}

// Closest same-slot pair of active pair p: distance, direction cj - ci + e, its length, the centres.
__device__ __forceinline__ float pair_dist(const TrajLds& L, const float4* gc, int p, float& bx, float& by, float& bz,
                                           float& bn, float4& bci, float4& bcj) {
  const float e = 1e-6f;
  float best = INFINITY;
  for (int it = L.pair_start[p]; it < L.pair_start[p + 1]; ++it) {
    const float4 ci = gc[L.pair_items[2 * it]];
    const float4 cj = gc[L.pair_items[2 * it + 1]];
    const float dx = cj.x - ci.x + e, dy = cj.y - ci.y + e, dz = cj.z - ci.z + e;
    const float n = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
    const float d = n - ci.w - cj.w;
    if (d < best) { best = d; bx = dx; by = dy; bz = dz; bn = n; bci = ci; bcj = cj; }
  }
  return best;
}

template <int MAXN>
__device__ __forceinline__ void group_smooth_min(const TrajModelDev& md, const TrajLds& L,
                                                 const WaypointNodes<MAXN, true>& K, const float4* gc,
                                                 const int* gstart, int gi, float tau, float& dk,
                                                 float (&gn)[MAXN]) {
  const int g = threadIdx.x % kGroup;
#pragma unroll
  for (int a = 0; a < MAXN; ++a) gn[a] = 0.f;
  float bx, by, bz, bn;
  float4 bc, bc2;
  float z = 0.f, dmin;
  if (gi == 0) {
    // Self pairs (a few dozen): pass 1 the group minimum, pass 2 the weights.
    dmin = INFINITY;
    for (int p = g; p < md.n_pairs; p += kGroup) dmin = fminf(dmin, pair_dist(L, gc, p, bx, by, bz, bn, bc, bc2));
    dmin = group8_min(dmin);
    if (!(dmin < INFINITY)) {  // empty group
      dk = INFINITY;
      return;
    }
    for (int p = g; p < md.n_pairs; p += kGroup) {
      const float d = pair_dist(L, gc, p, bx, by, bz, bn, bc, bc2);
      const float w = __expf(-(d - dmin) / tau);
      if (!(w > 0.f)) continue;
      z += w;
      const float k = w / bn;
      const int li = L.pair_links[2 * p], lj = L.pair_links[2 * p + 1];
      const int ni = L.link_node[li], nj = L.link_node[lj];
      pull_point<MAXN, true>(K, nj >= 0 ? L.anc[nj] : 0, bc2.x, bc2.y, bc2.z, k * bx, k * by, k * bz, gn);
      pull_point<MAXN, true>(K, ni >= 0 ? L.anc[ni] : 0, bc.x, bc.y, bc.z, -k * bx, -k * by, -k * bz, gn);
    }
  } else {
    // World group: one pass.  Each lane keeps its own running minimum mloc
    // and its weights relative to it, exp(-(d - mloc) / tau); when mloc
    // drops, the sum, the link's wrench and the node gradients gathered so
    // far are rescaled by exp(-(mloc_old - mloc_new) / tau).  The lanes'
    // sums are brought to the group minimum before they are added: the same
    // sums as weighting by the group minimum directly, in other rounding.
    const int m0 = gstart[gi - 1], m1 = gstart[gi];
    const float inv_tau = 1.f / tau;
    float mloc = INFINITY;
    for (int l = 0; l < md.n_links; ++l) {
      const int k0 = L.link_sph[l], k1 = L.link_sph[l + 1];
      if (k0 >= k1) continue;
      float F[3] = {0.f, 0.f, 0.f}, tq[3] = {0.f, 0.f, 0.f};
      // kObs obstacles per pass share each sphere centre read from LDS (and
      // give the wave independent work); their terms are then added in
      // obstacle order, the order of one obstacle at a time.  Sixteen per pass
      // measured 3% faster than eight and 27% faster than four (112 VGPRs,
      // no spills).
      constexpr int kObs = 16;
      for (int mb = m0 + g; mb < m1; mb += kGroup * kObs) {
        float4 o[kObs];
        float best[kObs];
        int bs[kObs];
#pragma unroll
        for (int u = 0; u < kObs; ++u) {
          const int m = mb + u * kGroup;
          const float4 w = L.world[m < m1 ? m : mb];
          o[u] = make_float4(w.x + 1e-6f, w.y + 1e-6f, w.z + 1e-6f, w.w);  // centre + e
          best[u] = INFINITY;
          bs[u] = -1;
        }
        for (int sp = k0; sp < k1; ++sp) {
          const float4 c = gc[sp];
#pragma unroll
          for (int u = 0; u < kObs; ++u) {
            const float d = sphere_world_dist(c, o[u]);
            if (d < best[u]) { best[u] = d; bs[u] = sp; }
          }
        }
#pragma unroll
        for (int u = 0; u < kObs; ++u) {
          if (mb + u * kGroup >= m1) break;
          const float d = best[u] - o[u].w;
          if (bs[u] >= 0) closest_dir(gc[bs[u]], o[u], bx, by, bz, bn, bc);
          if (d < mloc) {
            const float r = (mloc < INFINITY) ? __expf((d - mloc) * inv_tau) : 0.f;
            z *= r;
#pragma unroll
            for (int c = 0; c < 3; ++c) { F[c] *= r; tq[c] *= r; }
#pragma unroll
            for (int a = 0; a < MAXN; ++a) gn[a] *= r;
            mloc = d;
          }
          const float w = __expf((mloc - d) * inv_tau);
          if (w > 0.f) {
            z += w;
            const float k = -w * __builtin_amdgcn_rcpf(bn);  // d(distance)/dc = -(o - c + e) / |o - c + e|
            const float gx = k * bx, gy = k * by, gz = k * bz;
            F[0] += gx; F[1] += gy; F[2] += gz;
            tq[0] += bc.y * gz - bc.z * gy;
            tq[1] += bc.z * gx - bc.x * gz;
            tq[2] += bc.x * gy - bc.y * gx;
          }
        }
      }
      const int node = L.link_node[l];
      pull_wrench<MAXN>(K, node >= 0 ? L.anc[node] : 0, F, tq, gn);
    }
    dmin = group8_min(mloc);
    if (!(dmin < INFINITY)) {  // empty group
      dk = INFINITY;
      return;
    }
    const float r = (mloc < INFINITY) ? __expf(-(mloc - dmin) / tau) : 0.f;
    z *= r;
#pragma unroll
    for (int a = 0; a < MAXN; ++a) gn[a] *= r;
  }
  z = group_sum(z);
  dk = dmin - tau * __logf(z);
  const float inv = 1.f / z;
#pragma unroll
  for (int a = 0; a < MAXN; ++a) gn[a] = group_sum(gn[a]) * inv;
}

struct ScoLinArgs {
  TrajModelDev md;
  const float* x;
  const long long* start;
  int n_traj, max_len;
  long long n_waypoints;
  const int* gstart;  // (n_wg + 1) CSR of the world groups over the world spheres
  int n_wg;
  float tau;
  float* out_d;  // (W, G)
  float* out_J;  // (W, G, A)
};

template <int MAXN>
__global__ __launch_bounds__(kMaxTrajThreads) void traj_sco_linearize_kernel(ScoLinArgs a) {
  extern __shared__ float4 smem4[];
  const TrajLds L = stage_traj_model(a.md, a.max_len, reinterpret_cast<char*>(smem4));
  const int b = blockIdx.x;
  if (b >= a.n_traj) return;  // uniform per block
  const long long w0 = a.start[b], w1 = a.start[b + 1];
  if (w0 < 0 || w1 < w0 || w1 > a.n_waypoints || w1 - w0 > a.max_len) return;  // uniform per block
  const int T = static_cast<int>(w1 - w0);
  const int A = a.md.n_act;
  const int G = 1 + a.n_wg;
  const int grp = threadIdx.x / kGroup;
  const int g = threadIdx.x % kGroup;
  float4* gc = L.gc + grp * (a.md.n_spheres + a.md.n_links);
  const int n_groups = blockDim.x / kGroup;
  const int rounds = (T + n_groups - 1) / n_groups;
  for (int rd = 0; rd < rounds; ++rd) {
    const int t = rd * n_groups + grp;
    const bool active = t < T;
    WaypointNodes<MAXN, true> K;
    if (active) {
      float q[MAXN];
#pragma unroll
      for (int c = 0; c < MAXN; ++c) q[c] = (c < A) ? a.x[(w0 + t) * A + c] : 0.f;
      waypoint_fk<MAXN, true>(a.md, L, q, K);
      waypoint_centres<MAXN, true>(a.md, L, K, gc);
    }
    __syncthreads();
    if (active) {
      for (int gi = 0; gi < G; ++gi) {
        float dk, gn[MAXN], gq[MAXN];
        group_smooth_min<MAXN>(a.md, L, K, gc, a.gstart, gi, a.tau, dk, gn);
        if (!(dk < INFINITY)) {
#pragma unroll
          for (int c = 0; c < MAXN; ++c) gq[c] = 0.f;
        } else {
          node_to_actuated<MAXN>(L, a.md.n_nodes, gn, gq);
        }
        const long long row = (w0 + t) * G + gi;
        if (g == 0) a.out_d[row] = dk;
#pragma unroll
        for (int c = 0; c < MAXN; ++c)
          if (c < A && c % kGroup == g) a.out_J[row * A + c] = gq[c];
      }
    }
    __syncthreads();
  }
}

// The inner problem of _lbfgs_inner_solve (_sco_optimization.py:291-418) for
// one trajectory per workgroup, entirely in LDS: smoothness (velocity,
// acceleration, jerk) + limits + w_coll sum max(0, margin - d_lin)^2 with
// d_lin = d_k + J_k (x_t - q_k,t) + w_trust |x - q_k|^2; n_inner L-BFGS steps
// from x = q_k (history m, curvature test, two-loop recursion, five-step line
// search with sufficient decrease 1e-4 evaluated in order), returning the best
// iterate.
struct ScoInnerArgs {
  const float* x;       // (W, A) linearisation point = start
  const long long* start;
  int n_traj, max_len, n_act, G, m_hist, n_inner;
  long long n_waypoints;
  const float* d;       // (W, G)
  const float* J;       // (W, G, A)
  const float* lower;
  const float* upper;
  TrajWeights w;
  float w_coll, margin, w_trust;
  float* out_x;         // (W, A)
  unsigned a_magic, g_magic;  // quot_magic(n_act), quot_magic(G)
};

__host__ __device__ inline size_t sco_inner_lds_floats(int max_len, int A, int G, int m) {
  const size_t n = static_cast<size_t>(max_len) * A;
  return (9 + 2 * static_cast<size_t>(m)) * n + static_cast<size_t>(max_len) * G * (A + 2) + 2 * A +
         static_cast<size_t>(kMaxTrajWaves) * kRedSlots;
}

// Inner cost at y (LDS) and, with grad, its gradient into gout (endpoint rows
// zeroed); hk (T, G) holds each row's hinge slope -2 w_coll h for the gradient.
__device__ __forceinline__ float sco_inner_eval(const ScoInnerArgs& a, const float* y, const float* qk,
                                                const float* dk, const float* Jk, const float* lo, const float* hi,
                                                int T, float* acc, float* hk, float* gout, float* red, bool grad) {
  const int A = a.n_act, G = a.G;
  const int n = T * A;
  const int tid = threadIdx.x;
  const int na = T - 4;
  float sm = 0.f, rest = 0.f;
  for (int i = tid; i < n; i += blockDim.x) {
    const int c = i - quot(i, A, a.a_magic) * A;
    const float xv = y[i];
    const float v = fmaxf(xv - hi[c], 0.f) + fmaxf(lo[c] - xv, 0.f);
    rest += a.w.w_limits * v * v;
    const float dq = xv - qk[i];
    rest += a.w_trust * dq * dq;
    if (i + A < n) {
      const float dv = y[i + A] - xv;
      sm += a.w.w_vel * dv * dv;
    }
    if (i < na * A) {
      const float ac = (-xv + 16.f * y[i + A] - 30.f * y[i + 2 * A] + 16.f * y[i + 3 * A] - y[i + 4 * A]) / 12.f;
      acc[i] = ac;
      sm += a.w.w_acc * ac * ac;
    }
    if (grad) gout[i] = a.w.w_limits * 2.f * v * ((xv > hi[c] ? 1.f : 0.f) - (lo[c] > xv ? 1.f : 0.f)) +
                        2.f * a.w_trust * dq;
  }
  // Linearised hinge per (waypoint, group).
  for (int r = tid; r < T * G; r += blockDim.x) {
    const int t = quot(r, G, a.g_magic);
    const float* Jr = Jk + static_cast<long long>(r) * A;
    float dl = dk[r];
    for (int c = 0; c < A; ++c) dl += Jr[c] * (y[t * A + c] - qk[t * A + c]);
    const float h = a.margin - dl;
    if (h > 0.f) rest += a.w_coll * h * h;
    if (grad) hk[r] = (h > 0.f) ? -2.f * a.w_coll * h : 0.f;
  }
  __syncthreads();
  for (int i = tid; i < na * A; i += blockDim.x) {
    const int t = quot(i, A, a.a_magic);
    if (t + 1 < na) {
      const float jk = acc[i + A] - acc[i];
      sm += a.w.w_jerk * jk * jk;
    }
  }
  if (grad) {
    // hinge gradient: each element sums its waypoint's row slopes times the
    // rows' Jacobian entries
    for (int i = tid; i < n; i += blockDim.x) {
      const int t = quot(i, A, a.a_magic), c = i - t * A;
      float gh = 0.f;
      for (int gg = 0; gg < G; ++gg) {
        const int r = t * G + gg;
        gh += hk[r] * Jk[static_cast<long long>(r) * A + c];
      }
      gout[i] += gh;
    }
    __syncthreads();
    const float c5[5] = {-1.f / 12.f, 16.f / 12.f, -30.f / 12.f, 16.f / 12.f, -1.f / 12.f};
    for (int i = tid; i < n; i += blockDim.x) {
      const int t = quot(i, A, a.a_magic);
      float gs = 0.f;
      if (t >= 1) gs += 2.f * a.w.w_vel * (y[i] - y[i - A]);
      if (t + 1 < T) gs -= 2.f * a.w.w_vel * (y[i + A] - y[i]);
#pragma unroll
      for (int o = 0; o < 5; ++o) {
        const int tt = t - o;
        if (tt >= 0 && tt < na) {
          const int j = i - o * A;
          float ga = 2.f * a.w.w_acc * acc[j];
          if (tt + 1 < na) ga -= 2.f * a.w.w_jerk * (acc[j + A] - acc[j]);
          if (tt >= 1) ga += 2.f * a.w.w_jerk * (acc[j] - acc[j - A]);
          gs += ga * c5[o];
        }
      }
      float gv = gout[i] + a.w.w_smooth * gs;
      if (t == 0 || t == T - 1) gv = 0.f;
      gout[i] = gv;
    }
  }
  float v2[2] = {sm, rest};
  block_sum<2>(v2, red);
  return a.w.w_smooth * v2[0] + v2[1];
}

__global__ __launch_bounds__(512) void traj_sco_inner_kernel(ScoInnerArgs a) {
  extern __shared__ float4 smem4[];
  const int b = blockIdx.x;
  if (b >= a.n_traj) return;  // uniform per block
  const long long w0 = a.start[b], w1 = a.start[b + 1];
  if (w0 < 0 || w1 < w0 || w1 > a.n_waypoints || w1 - w0 > a.max_len) return;  // uniform per block
  const int T = static_cast<int>(w1 - w0);
  const int A = a.n_act, G = a.G, m = a.m_hist;
  const int n = T * A;
  const int cap = a.max_len * A;
  const int tid = threadIdx.x;
  float* base = reinterpret_cast<float*>(smem4);
  float* qk = base;
  float* x = qk + cap;
  float* g = x + cap;
  float* dir = g + cap;
  float* xp = dir + cap;
  float* gp = xp + cap;
  float* best = gp + cap;
  float* acc = best + cap;
  float* xt = acc + cap;
  float* sb = xt + cap;             // (m, cap)
  float* yb = sb + m * cap;         // (m, cap)
  float* dk = yb + m * cap;         // (max_len, G)
  float* Jk = dk + a.max_len * G;   // (max_len, G, A)
  float* hk = Jk + a.max_len * G * A;  // (max_len, G)
  float* lo = hk + a.max_len * G;
  float* hi = lo + A;
  float* red = hi + A;
  for (int i = tid; i < n; i += blockDim.x) {
    qk[i] = a.x[w0 * A + i];
    x[i] = qk[i];
  }
  for (int i = tid; i < T * G; i += blockDim.x) dk[i] = a.d[w0 * G + i];
  for (int i = tid; i < T * G * A; i += blockDim.x) Jk[i] = a.J[w0 * G * A + i];
  for (int c = tid; c < A; c += blockDim.x) { lo[c] = a.lower[c]; hi[c] = a.upper[c]; }
  __syncthreads();
  float rho[kMaxHist];
#pragma unroll
  for (int h = 0; h < kMaxHist; ++h) rho[h] = 0.f;
  int m_used = 0, newest = 0;
  float best_cost = sco_inner_eval(a, x, qk, dk, Jk, lo, hi, T, acc, hk, g, red, true);
  for (int i = tid; i < n; i += blockDim.x) { best[i] = x[i]; xp[i] = x[i]; gp[i] = g[i]; }
  __syncthreads();
  for (int k = 0; k < a.n_inner; ++k) {
    const float cost = sco_inner_eval(a, x, qk, dk, Jk, lo, hi, T, acc, hk, g, red, true);
    float syy[2] = {0.f, 0.f};
    for (int i = tid; i < n; i += blockDim.x) {
      const float sv = x[i] - xp[i], yv = g[i] - gp[i];
      syy[0] += sv * yv;
      syy[1] += yv * yv;
    }
    block_sum<2>(syy, red);
    if (syy[0] > 1e-10f * syy[1] + 1e-30f && k > 0) {
      newest = (newest + 1) % m;
      for (int i = tid; i < n; i += blockDim.x) {
        sb[newest * cap + i] = x[i] - xp[i];
        yb[newest * cap + i] = g[i] - gp[i];
      }
      const float r = 1.f / (syy[0] + 1e-30f);
#pragma unroll
      for (int h = 0; h < kMaxHist; ++h)
        if (h == newest) rho[h] = r;
      if (m_used < m) ++m_used;
    }
    __syncthreads();
    if (m_used > 0) {
      float alph[kMaxHist];
#pragma unroll
      for (int h = 0; h < kMaxHist; ++h) alph[h] = 0.f;
      for (int i = tid; i < n; i += blockDim.x) dir[i] = g[i];
      for (int j = 0; j < m_used; ++j) {
        const int idx = (newest - j + m) % m;
        float sq = 0.f;
        for (int i = tid; i < n; i += blockDim.x) sq += sb[idx * cap + i] * dir[i];
        const float al = sel_hist(rho, idx) * block_sum1(sq, red);
#pragma unroll
        for (int h = 0; h < kMaxHist; ++h)
          if (h == idx) alph[h] = al;
        for (int i = tid; i < n; i += blockDim.x) dir[i] -= al * yb[idx * cap + i];
      }
      float g2[2] = {0.f, 0.f};
      for (int i = tid; i < n; i += blockDim.x) {
        g2[0] += sb[newest * cap + i] * yb[newest * cap + i];
        g2[1] += yb[newest * cap + i] * yb[newest * cap + i];
      }
      block_sum<2>(g2, red);
      const float gamma = g2[0] / (g2[1] + 1e-18f);
      for (int i = tid; i < n; i += blockDim.x) dir[i] *= gamma;
      for (int st = 0; st < m_used; ++st) {
        const int idx = (newest - m_used + 1 + st + 2 * m) % m;
        float yr = 0.f;
        for (int i = tid; i < n; i += blockDim.x) yr += yb[idx * cap + i] * dir[i];
        const float beta = sel_hist(rho, idx) * block_sum1(yr, red);
        const float coef = sel_hist(alph, idx) - beta;
        for (int i = tid; i < n; i += blockDim.x) dir[i] += sb[idx * cap + i] * coef;
      }
      for (int i = tid; i < n; i += blockDim.x) dir[i] = -dir[i];
    } else {
      float gg = 0.f;
      for (int i = tid; i < n; i += blockDim.x) gg += g[i] * g[i];
      const float inv = 1.f / (sqrtf(block_sum1(gg, red)) + 1e-18f);
      for (int i = tid; i < n; i += blockDim.x) dir[i] = -g[i] * inv;
    }
    for (int i = tid; i < n; i += blockDim.x) {
      const int t = quot(i, A, a.a_magic);
      if (t == 0 || t == T - 1) dir[i] = 0.f;
    }
    __syncthreads();
    const float thresh = cost * (1.f - 1e-4f);
    int pick = -1;
    float pick_cost = INFINITY;
    int first_best = 0;
    for (int kk = 0; kk < kNumAlphas; ++kk) {
      const float al = traj_alpha(kk);
      for (int i = tid; i < n; i += blockDim.x) xt[i] = x[i] + al * dir[i];
      __syncthreads();
      const float c = sco_inner_eval(a, xt, qk, dk, Jk, lo, hi, T, acc, hk, nullptr, red, false);
      if (kk == 0 || c < pick_cost) { pick_cost = c; first_best = kk; }
      if (c < thresh) { pick = kk; pick_cost = c; break; }
    }
    if (pick < 0) pick = first_best;
    const float al = traj_alpha(pick);
    for (int i = tid; i < n; i += blockDim.x) {
      xp[i] = x[i];
      gp[i] = g[i];
      x[i] += al * dir[i];
    }
    if (pick_cost < best_cost) {
      best_cost = pick_cost;
      for (int i = tid; i < n; i += blockDim.x) best[i] = x[i];
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += blockDim.x) a.out_x[w0 * A + i] = best[i];
}

}  // namespace pk

namespace {

int fill_traj_model(const pk_traj_model* m, pk::TrajModelDev& d, int& maxn) {
  PK_REQUIRE(m != nullptr, "null model");
  PK_REQUIRE(m->n_nodes >= 0 && m->n_nodes <= 16, "n_nodes=%d outside [0, 16]", m->n_nodes);
  PK_REQUIRE(m->n_act >= 1 && m->n_act <= 16, "n_act=%d outside [1, 16]", m->n_act);
  PK_REQUIRE(m->n_links >= 0 && m->n_spheres >= 0 && m->n_pairs >= 0 && m->n_items >= 0 &&
                 m->n_world >= 0,
             "negative model size");
  PK_REQUIRE((m->n_nodes == 0 || m->nodes) && (m->n_links == 0 || (m->link_node && m->link_bound)) &&
                 m->link_sph_start && (m->n_spheres == 0 || m->spheres) &&
                 (m->n_pairs == 0 || m->pair_links) && m->pair_start &&
                 (m->n_items == 0 || m->pair_items) && (m->n_world == 0 || m->world) && m->lower &&
                 m->upper,
             "null model pointer");
  PK_REQUIRE(m->n_items == 0 || m->n_spheres > 0, "pair items without spheres");
  PK_REQUIRE(m->mode == 0 || m->mode == 1, "mode=%d (0 = hinge^2, 1 = colldist)", m->mode);
  PK_REQUIRE(m->margin >= 0.f, "negative margin");
  d.nodes = m->nodes; d.n_nodes = m->n_nodes; d.n_act = m->n_act;
  d.link_node = m->link_node; d.link_sph_start = m->link_sph_start; d.spheres = m->spheres;
  d.link_bound = m->link_bound; d.n_links = m->n_links; d.n_spheres = m->n_spheres;
  d.pair_links = m->pair_links; d.pair_start = m->pair_start; d.pair_items = m->pair_items;
  d.n_pairs = m->n_pairs; d.n_items = m->n_items; d.world = m->world; d.n_world = m->n_world;
  d.lower = m->lower; d.upper = m->upper; d.margin = m->margin; d.mode = m->mode;
  d.a_magic = pk::quot_magic(m->n_act);
  const int gs = pk::fill_grid_args(m->grids, m->n_grids, m->margin, d.grid);
  if (gs != PK_OK) return gs;
  const int width = std::max(m->n_nodes, m->n_act);
  maxn = width <= 8 ? 8 : 16;
  return PK_OK;
}

size_t traj_lds_bytes(const pk::TrajModelDev& d, int max_len, int threads) {
  return pk::traj_lds_layout(d.n_nodes, d.n_links, d.n_spheres, d.n_pairs, d.n_items, d.n_world, d.n_act,
                             max_len, threads)
      .total;
}

// Threads per trajectory workgroup: the requested 128/256/512, or (0) by the
// longest trajectory, so each of the 8-lane groups gets ~1-4 waypoints.
int pick_threads(int requested, int max_len, int& threads) {
  if (requested == 0) {
    threads = max_len <= 32 ? 128 : (max_len <= 64 ? 256 : 512);
    return PK_OK;
  }
  PK_REQUIRE(requested == 128 || requested == 256 || requested == 512,
             "block_threads=%d (0, 128, 256 or 512)", requested);
  threads = requested;
  return PK_OK;
}

}  // namespace

extern "C" int64_t pk_traj_lbfgs_state_floats(int64_t n_waypoints, int32_t n_act, int32_t n_traj,
                                              int32_t m_hist) {
  if (n_waypoints < 0 || n_act < 0 || n_traj < 0 || m_hist < 0) return -1;
  return (3 + 2 * static_cast<int64_t>(m_hist)) * n_waypoints * n_act +
         static_cast<int64_t>(pk::kScalarWords) * n_traj;
}

extern "C" int pk_traj_cost_grad_f32(void* stream, const pk_traj_model* model, const float* x,
                                     const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints,
                                     int32_t max_len, int32_t block_threads, const int32_t* traj_subset,
                                     int32_t n_subset, const float* weights, float* out_comps,
                                     float* out_grad) {
  pk::TrajModelDev d{};
  int maxn = 8;
  const int st = fill_traj_model(model, d, maxn);
  if (st != PK_OK) return st;
  PK_REQUIRE(n_traj >= 0 && n_waypoints >= 0 && max_len >= 1, "invalid trajectory sizes");
  if (n_traj == 0) return PK_OK;
  PK_REQUIRE(x && traj_start && weights && out_comps, "null pointer");
  PK_REQUIRE(traj_subset == nullptr || n_subset >= 0, "negative subset size");
  int threads = 0;
  const int ts = pick_threads(block_threads, max_len, threads);
  if (ts != PK_OK) return ts;
  const int blocks = traj_subset ? n_subset : n_traj;
  if (blocks == 0) return PK_OK;
  const size_t lds = traj_lds_bytes(d, max_len, threads);
  PK_REQUIRE(lds <= 150 * 1024, "trajectory model needs %zu bytes of LDS (limit 150 KiB)", lds);
  pk::CostArgs a{};
  a.md = d;
  a.x = x;
  a.start = reinterpret_cast<const long long*>(traj_start);
  a.subset = traj_subset;
  a.n_traj = n_traj;
  a.max_len = max_len;
  a.n_waypoints = n_waypoints;
  a.w = pk::TrajWeights{weights[0], weights[1], weights[2], weights[3], weights[4]};
  a.w_coll = weights[5];
  a.out_comps = out_comps;
  a.out_grad = out_grad;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (maxn == 8)
    hipLaunchKernelGGL(pk::traj_cost_kernel<8>, dim3(blocks), dim3(threads), lds, s, a);
  else
    hipLaunchKernelGGL(pk::traj_cost_kernel<16>, dim3(blocks), dim3(threads), lds, s, a);
  return pk::check_launch("traj_cost_kernel");
}

extern "C" int pk_traj_lbfgs_f32(void* stream, const pk_traj_model* model, const float* x_init,
                                 const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints,
                                 int32_t max_len, int32_t block_threads, const int32_t* traj_subset,
                                 int32_t n_subset, const pk_lbfgs_config* cfg, int32_t iter_begin,
                                 int32_t iter_end, float* state, float* out_best, float* out_best_cost) {
  pk::TrajModelDev d{};
  int maxn = 8;
  const int st = fill_traj_model(model, d, maxn);
  if (st != PK_OK) return st;
  PK_REQUIRE(cfg != nullptr, "null config");
  PK_REQUIRE(n_traj >= 0 && n_waypoints >= 0 && max_len >= 1, "invalid trajectory sizes");
  PK_REQUIRE(cfg->n_iters >= 0 && cfg->m_hist >= 1 && cfg->m_hist <= pk::kMaxHist &&
                 cfg->escalation_interval >= 1,
             "invalid L-BFGS configuration (m_hist in [1, %d], escalation_interval >= 1)", pk::kMaxHist);
  PK_REQUIRE(0 <= iter_begin && iter_begin <= iter_end && iter_end <= cfg->n_iters,
             "iteration range [%d, %d) outside [0, %d]", iter_begin, iter_end, cfg->n_iters);
  if (n_traj == 0 || iter_begin == iter_end) return PK_OK;
  PK_REQUIRE(x_init && traj_start && state && out_best && out_best_cost, "null pointer");
  PK_REQUIRE(traj_subset == nullptr || n_subset >= 0, "negative subset size");
  int threads = 0;
  const int ts = pick_threads(block_threads, max_len, threads);
  if (ts != PK_OK) return ts;
  const int blocks = traj_subset ? n_subset : n_traj;
  if (blocks == 0) return PK_OK;
  const size_t lds = traj_lds_bytes(d, max_len, threads);
  PK_REQUIRE(lds <= 150 * 1024, "trajectory model needs %zu bytes of LDS (limit 150 KiB)", lds);
  pk::LbfgsArgs a{};
  a.md = d;
  a.x_init = x_init;
  a.start = reinterpret_cast<const long long*>(traj_start);
  a.subset = traj_subset;
  a.n_traj = n_traj;
  a.max_len = max_len;
  a.n_waypoints = n_waypoints;
  a.w = pk::TrajWeights{cfg->w_smooth, cfg->w_vel, cfg->w_acc, cfg->w_jerk, cfg->w_limits};
  a.w_coll0 = cfg->w_collision;
  a.w_coll_max = cfg->w_collision_max;
  a.penalty_scale = cfg->penalty_scale;
  a.escalation = cfg->escalation_interval;
  a.n_iters = cfg->n_iters;
  a.m_hist = cfg->m_hist;
  a.it0 = iter_begin;
  a.it1 = iter_end;
  const long long WA = n_waypoints * d.n_act;
  a.xs = state;
  a.xp = state + WA;
  a.gp = state + 2 * WA;
  a.sb = state + 3 * WA;
  a.yb = a.sb + static_cast<long long>(cfg->m_hist) * WA;
  a.sc = a.yb + static_cast<long long>(cfg->m_hist) * WA;
  a.out_best = out_best;
  a.out_cost = out_best_cost;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (maxn == 8)
    hipLaunchKernelGGL(pk::traj_lbfgs_kernel<8>, dim3(blocks), dim3(threads), lds, s, a);
  else
    hipLaunchKernelGGL(pk::traj_lbfgs_kernel<16>, dim3(blocks), dim3(threads), lds, s, a);
  return pk::check_launch("traj_lbfgs_kernel");
}

extern "C" int pk_traj_chomp_f32(void* stream, const pk_traj_model* model, const float* x_init,
                                 const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints, int32_t max_len,
                                 int32_t block_threads, const pk_chomp_config* cfg, float* out_traj,
                                 float* out_cost, float* out_trace) {
  pk::TrajModelDev d{};
  int maxn = 8;
  const int st = fill_traj_model(model, d, maxn);
  if (st != PK_OK) return st;
  PK_REQUIRE(cfg != nullptr, "null config");
  PK_REQUIRE(n_traj >= 0 && n_waypoints >= 0 && max_len >= 1, "invalid trajectory sizes");
  PK_REQUIRE(cfg->n_iters >= 0 && cfg->n_cg_iters >= 0 && cfg->early_stop_patience >= 0,
             "invalid CHOMP configuration (negative iteration counts)");
  PK_REQUIRE(cfg->max_delta_per_step >= 0.f, "negative max_delta_per_step");
  if (n_traj == 0) return PK_OK;
  PK_REQUIRE(x_init && traj_start && out_traj && out_cost, "null pointer");
  int threads = 0;
  const int ts = pick_threads(block_threads, max_len, threads);
  if (ts != PK_OK) return ts;
  const size_t lds = traj_lds_bytes(d, max_len, threads);
  PK_REQUIRE(lds <= 150 * 1024, "trajectory model needs %zu bytes of LDS (limit 150 KiB)", lds);
  pk::ChompArgs a{};
  a.md = d;
  a.x_init = x_init;
  a.start = reinterpret_cast<const long long*>(traj_start);
  a.n_traj = n_traj;
  a.max_len = max_len;
  a.n_waypoints = n_waypoints;
  a.w = pk::TrajWeights{cfg->w_smooth, 0.f, cfg->w_acc, cfg->w_jerk, cfg->w_limits};
  a.step = cfg->step_size;
  a.w_coll0 = cfg->w_collision;
  a.w_coll_max = cfg->w_collision_max;
  a.scale = cfg->collision_penalty_scale;
  a.reg = cfg->smoothness_reg;
  a.cg_tol = cfg->cg_tol;
  a.clip = cfg->grad_clip_norm;
  a.max_delta = cfg->max_delta_per_step;
  a.min_improve = cfg->min_cost_improve;
  a.n_iters = cfg->n_iters;
  a.covariant = cfg->use_covariant_update;
  a.n_cg = cfg->n_cg_iters;
  a.patience = cfg->early_stop_patience;
  a.out_traj = out_traj;
  a.out_cost = out_cost;
  a.out_trace = out_trace;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (maxn == 8)
    hipLaunchKernelGGL(pk::traj_chomp_kernel<8>, dim3(n_traj), dim3(threads), lds, s, a);
  else
    hipLaunchKernelGGL(pk::traj_chomp_kernel<16>, dim3(n_traj), dim3(threads), lds, s, a);
  return pk::check_launch("traj_chomp_kernel");
}

extern "C" int pk_traj_sco_linearize_f32(void* stream, const pk_traj_model* model, const float* x,
                                         const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints,
                                         int32_t max_len, int32_t block_threads, const int32_t* world_group_start,
                                         int32_t n_world_groups, float temperature, float* out_d, float* out_J) {
  pk::TrajModelDev d{};
  int maxn = 8;
  const int st = fill_traj_model(model, d, maxn);
  if (st != PK_OK) return st;
  PK_REQUIRE(n_traj >= 0 && n_waypoints >= 0 && max_len >= 1, "invalid trajectory sizes");
  PK_REQUIRE(n_world_groups >= 0 && n_world_groups + 1 <= pk::kMaxScoGroups,
             "n_world_groups=%d outside [0, %d]", n_world_groups, pk::kMaxScoGroups - 1);
  PK_REQUIRE(temperature > 0.f, "temperature must be > 0");
  if (n_traj == 0) return PK_OK;
  PK_REQUIRE(x && traj_start && out_d && out_J && (n_world_groups == 0 || world_group_start), "null pointer");
  int threads = 0;
  const int ts = pick_threads(block_threads, max_len, threads);
  if (ts != PK_OK) return ts;
  const size_t lds = traj_lds_bytes(d, max_len, threads);
  PK_REQUIRE(lds <= 150 * 1024, "trajectory model needs %zu bytes of LDS (limit 150 KiB)", lds);
  pk::ScoLinArgs a{};
  a.md = d;
  a.x = x;
  a.start = reinterpret_cast<const long long*>(traj_start);
  a.n_traj = n_traj;
  a.max_len = max_len;
  a.n_waypoints = n_waypoints;
  a.gstart = world_group_start;
  a.n_wg = n_world_groups;
  a.tau = temperature;
  a.out_d = out_d;
  a.out_J = out_J;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (maxn == 8)
    hipLaunchKernelGGL(pk::traj_sco_linearize_kernel<8>, dim3(n_traj), dim3(threads), lds, s, a);
  else
    hipLaunchKernelGGL(pk::traj_sco_linearize_kernel<16>, dim3(n_traj), dim3(threads), lds, s, a);
  return pk::check_launch("traj_sco_linearize_kernel");
}

extern "C" int pk_traj_sco_inner_f32(void* stream, const float* x, const int64_t* traj_start, int32_t n_traj,
                                     int64_t n_waypoints, int32_t max_len, int32_t n_act, int32_t n_groups,
                                     const float* d, const float* J, const float* lower, const float* upper,
                                     const pk_sco_config* cfg, float w_collision, float* out_x) {
  PK_REQUIRE(cfg != nullptr, "null config");
  PK_REQUIRE(n_traj >= 0 && n_waypoints >= 0 && max_len >= 1, "invalid trajectory sizes");
  PK_REQUIRE(n_act >= 1 && n_act <= 16, "n_act=%d outside [1, 16]", n_act);
  PK_REQUIRE(n_groups >= 1 && n_groups <= pk::kMaxScoGroups, "n_groups=%d outside [1, %d]", n_groups,
             pk::kMaxScoGroups);
  PK_REQUIRE(cfg->m_lbfgs >= 1 && cfg->m_lbfgs <= pk::kMaxHist && cfg->n_inner_iters >= 0,
             "invalid SCO configuration (m_lbfgs in [1, %d])", pk::kMaxHist);
  if (n_traj == 0) return PK_OK;
  PK_REQUIRE(x && traj_start && d && J && lower && upper && out_x, "null pointer");
  const size_t lds = sizeof(float) * pk::sco_inner_lds_floats(max_len, n_act, n_groups, cfg->m_lbfgs);
  PK_REQUIRE(lds <= 150 * 1024, "SCO inner problem needs %zu bytes of LDS (limit 150 KiB): max_len too long",
             lds);
  pk::ScoInnerArgs a{};
  a.x = x;
  a.start = reinterpret_cast<const long long*>(traj_start);
  a.n_traj = n_traj;
  a.max_len = max_len;
  a.n_act = n_act;
  a.G = n_groups;
  a.a_magic = pk::quot_magic(n_act);
  a.g_magic = pk::quot_magic(n_groups);
  a.m_hist = cfg->m_lbfgs;
  a.n_inner = cfg->n_inner_iters;
  a.n_waypoints = n_waypoints;
  a.d = d;
  a.J = J;
  a.lower = lower;
  a.upper = upper;
  a.w = pk::TrajWeights{cfg->w_smooth, cfg->w_vel, cfg->w_acc, cfg->w_jerk, cfg->w_limits};
  a.w_coll = w_collision;
  a.margin = cfg->collision_margin;
  a.w_trust = cfg->w_trust;
  a.out_x = out_x;
  hipLaunchKernelGGL(pk::traj_sco_inner_kernel, dim3(n_traj), dim3(512), lds, static_cast<hipStream_t>(stream), a);
  return pk::check_launch("traj_sco_inner_kernel");
}
