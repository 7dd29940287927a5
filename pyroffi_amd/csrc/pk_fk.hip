// Forward kinematics for gfx950: one lane = one configuration.
//
// A 64-lane workgroup owns 64 consecutive configurations.  Its prologue issues
// every global load at once (the configuration block, the link -> joint map
// and one joint's table entries per lane), then builds one walk record per
// joint slot (topological position) in LDS through three barriers: the
// joint -> slot map, the records and the link map, and the output positions.
// Every lane then walks the slots for its configuration and keeps each
// joint's world pose in LDS directly at its place in the lane's output row (a
// link pose is the pose of its parent joint, and in a URDF every joint has
// exactly one child link).  The finished tile -- 64 rows of n_out x 7 floats,
// the exact image of the block's contiguous output -- is copied out with
// 16-byte loads issued in batches of eight and non-temporal 16-byte stores.
// Output positions the walk does not write are filled around it (root links:
// the identity, before; a second link on the same joint: a copy, after).  The
// row stride
// is odd (the row padded by one float when 7 n_out is even), so the per-lane
// walk accesses are free of bank conflicts.
//
// The walk is a dependency chain per lane with one wave per SIMD, where an
// instruction issues about every 8 cycles, so the work off the chain is taken
// out of it: the local transforms of the revolute and prismatic joints are
// computed first, four joints at a time (independent instruction streams the
// wave interleaves) with a branch-free sine/cosine, and the chain pass only
// composes (world = parent o local).  It avoids LDS round trips as well: each
// slot's own and parent row offsets live in the slot's lane and are read
// with readlane, the next slot's local transform is read while the current
// one is composed, and a parent that is the previous slot stays in
// registers.  A revolute joint's local rotation is c ptf.q + s (ptf.q *
// (0, axis)) for the half-angle cosine c and sine s (ptf.t unchanged), a
// prismatic joint's local translation ptf.t + q R(ptf.q) v, a constant
// joint's local transform its parent_tf; only a general screw joint takes the
// full exponential.
//
// Reference behaviour: FkCudaImpl / fk_kernel_warp
// (src/pyroffi/cuda_kernels/_fk_cuda_kernel.cu:39-137, 236-454) and the link
// gather of src/pyroffi/_robot.py:180-196.  A topological order that is not a
// permutation of the joints gives NaN rows (the reference reads whatever
// such a table indexes).
#include "pk_model.h"
#include <type_traits>

namespace pk {

typedef float v4f __attribute__((ext_vector_type(4)));

// Walk record of one joint slot, built by every FK workgroup from the tables.
struct __attribute__((aligned(16))) FkRec {
  float ptf[7];  // parent -> joint transform
  float a[4];    // kind 4: ptf.q * (0, axis); kind 2: R(ptf.q) v; kind 3: v
  float b[4];    // kind 4: b[0] = |omega|; kind 3: omega
  float mul, off;
  int src;       // actuated source, -1 for none
  int kind;      // 0 constant, 2 prismatic, 3 general twist, 4 revolute
  int pos;       // float offset of the slot's pose in the lane's row
  int pad[3];
};
static_assert(sizeof(FkRec) == 24 * 4, "FkRec layout");

// Raw table entries of one joint, staged by joint index.
constexpr int kFkRaw = 18;  // twist 6, parent_tf 7, mul, off, src, parent, spare

// LDS image of one FK workgroup (float offsets into the dynamic LDS array).
struct FkLds {
  int recs, raw, slot_of, first, fill, cfg, tile, total;
};

// n_extra: positions past the output row for joints without a child link.
__host__ __device__ inline int fk_row_stride(int n_out, int n_extra) { return (7 * (n_out + n_extra)) | 1; }

__host__ __device__ inline FkLds fk_lds_layout(int J, int n_act, int n_out, int n_extra) {
  FkLds o;
  o.recs = 0;
  o.raw = o.recs + 24 * J;
  o.slot_of = o.raw + kFkRaw * J;
  o.first = o.slot_of + J;
  o.fill = o.first + J;
  o.cfg = o.fill + n_out;
  o.tile = ((o.cfg + n_act * kLaneStride) + 3) & ~3;  // 16-byte aligned
  o.total = o.tile + kWave * fk_row_stride(n_out, n_extra);
  return o;
}

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Largest |x| the branch-free sine/cosine below is used for.
constexpr float kFkTrigMax = 1.0e4f;

// sin(x) and cos(x) without branches, for |x| <= kFkTrigMax: Cody-Waite
// reduction x = k pi/2 + r with pi/2 split into three floats (each the float
// nearest to what the previous ones leave), |r| <= pi/4, then the Taylor
// polynomials of sin r to r^9 and cos r to r^10 (truncation below 2e-9 on
// |r| <= pi/4) and the quadrant's swap and signs.  Error on every float in
// [-8, 8] and every 256th float up to 1e4: at most 1.5 ulp, 7.2e-8 absolute.
__device__ __forceinline__ void sincos_cw(float x, float& s, float& c) {
  const float k = rintf(x * 0.636619772367581343f);  // x 2/pi
  float r = fmaf(-k, 1.5707963705062866f, x);
  r = fmaf(-k, -4.371138828673793e-08f, r);
  r = fmaf(-k, -1.7151245100058819e-15f, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.7557319223985893e-06f, -1.984126984126984e-04f);  // 1/9!, -1/7!
  ps = fmaf(r2, ps, 8.333333333333333e-03f);                               // 1/5!
  ps = fmaf(r2, ps, -1.6666666666666666e-01f);                             // -1/3!
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.755731922398589e-07f, 2.48015873015873e-05f);     // -1/10!, 1/8!
  pc = fmaf(r2, pc, -1.388888888888889e-03f);                              // -1/6!
  pc = fmaf(r2, pc, 4.1666666666666664e-02f);                              // 1/4!
  pc = fmaf(r2, pc, -0.5f);
  const float cr = fmaf(r2, pc, 1.0f);
  const int qd = static_cast<int>(k) & 3;
  const float so = (qd & 1) ? cr : sr, co = (qd & 1) ? sr : cr;
  s = (qd & 2) ? -so : so;
  c = ((qd + 1) & 2) ? -co : co;
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
  const int J = tables.n_joints;  // <= 64: one slot and one joint per lane
  const int n_out = out_mode ? n_links : J;
  const int row = 7 * n_out;
  const int stride = fk_row_stride(n_out, n_extra);
  const FkLds lay = fk_lds_layout(J, n_act, n_out, n_extra);
  FkRec* recs = reinterpret_cast<FkRec*>(smem + lay.recs);
  float* raw = smem + lay.raw;                          // [joint * kFkRaw + field]
  int* slot_of = reinterpret_cast<int*>(smem + lay.slot_of);
  int* first = reinterpret_cast<int*>(smem + lay.first);  // per slot: its first output position
  int* fill = reinterpret_cast<int*>(smem + lay.fill);  // repeated positions: row offset of the pose they copy
  float* cfg_tile = smem + lay.cfg;                     // [a][lane], stride 65
  float* tile = smem + lay.tile;                        // [lane * stride + offset]

  // The workgroup is one wave: LDS hand-offs between lanes need only
  // wave_sync(), and global loads stay in flight until their first use.
  const int lane = threadIdx.x;
  const long long b0 = static_cast<long long>(blockIdx.x) * kWave;
  const int n_valid = static_cast<int>(min(static_cast<long long>(kWave), batch - b0));
  constexpr int kNoPos = 0x7fffffff;

  // A. Every global load of the prologue at once: the configuration block,
  // the link -> joint map, this lane's slot's joint and this lane's joint's
  // table entries.
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
  const bool is_slot = lane < J;
  int jt = -1;
  if (is_slot) {
    jt = tables.topo_inv[lane];
    float e[kFkRaw];
#pragma unroll
    for (int c = 0; c < 6; ++c) e[c] = tables.twists[lane * 6 + c];
#pragma unroll
    for (int c = 0; c < 7; ++c) e[6 + c] = tables.parent_tf[lane * 7 + c];
    e[13] = tables.mimic_mul[lane];
    e[14] = tables.mimic_off[lane];
    const int m = tables.mimic_act_idx[lane];
    const int ai = tables.act_idx[lane];
    int src = (m != -1) ? m : ai;
    if (src < 0 || src >= tables.n_act) src = -1;
    e[15] = __int_as_float(src);
    e[16] = __int_as_float(tables.parent_idx[lane]);
    e[17] = 0.f;
#pragma unroll
    for (int c = 0; c < kFkRaw; ++c) raw[lane * kFkRaw + c] = e[c];
    first[lane] = kNoPos;
    if (jt >= 0 && jt < J) slot_of[jt] = lane;
  }
  wave_sync();

  // B. The slot order is valid when every slot reads its own index back (J
  // distinct joints).  Records without positions and the link map (first
  // output position per slot by an LDS atomic minimum).
  const bool own = !is_slot || (jt >= 0 && jt < J && slot_of[jt] == lane);
  const bool perm = __all(own);
  int ps = -1, my_kind = -1;
  if (is_slot && perm) {
    const float* e = raw + jt * kFkRaw;
    FkRec r;
#pragma unroll
    for (int c = 0; c < 7; ++c) r.ptf[c] = e[6 + c];
    const float v[3] = {e[0], e[1], e[2]}, w[3] = {e[3], e[4], e[5]};
    r.mul = e[13];
    r.off = e[14];
    r.src = __float_as_int(e[15]);
    const int p = __float_as_int(e[16]);
    ps = (p >= 0 && p < J) ? slot_of[p] : -1;
    if (ps >= lane) ps = -1;  // not topologically before us: invalid table, treated as a root
    const float l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const float a2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    int kind;
    if (r.src == -1 && r.off == 0.f) kind = 0;
    else if (a2 == 0.f) kind = 2;
    else if (l2 == 0.f) kind = 4;
    else kind = 3;
#pragma unroll
    for (int c = 0; c < 4; ++c) { r.a[c] = 0.f; r.b[c] = 0.f; }
    if (kind == 4) {
      const float na = sqrtf(a2);
      const float inv = 1.f / na;
      quat_mul(r.ptf[0], r.ptf[1], r.ptf[2], r.ptf[3], 0.f, w[0] * inv, w[1] * inv, w[2] * inv,
               r.a[0], r.a[1], r.a[2], r.a[3]);
      r.b[0] = na;
    } else if (kind == 2) {
      quat_rotate(r.ptf[0], r.ptf[1], r.ptf[2], r.ptf[3], v[0], v[1], v[2], r.a[0], r.a[1], r.a[2]);
    } else if (kind == 3) {
#pragma unroll
      for (int c = 0; c < 3; ++c) { r.a[c] = v[c]; r.b[c] = w[c]; }
    }
    r.kind = kind;
    my_kind = kind;
    r.pos = 0;
    r.pad[0] = r.pad[1] = r.pad[2] = 0;
    recs[lane] = r;
  }
  int lm[kOutIters];
#pragma unroll
  for (int i = 0; i < kOutIters; ++i) {
    const int o = lane + i * kWave;
    const int pj = pjv[i];
    lm[i] = (perm && o < n_out && pj >= 0 && pj < J) ? slot_of[pj] : -1;
    if (lm[i] >= 0) atomicMin(&first[lm[i]], o);
  }
  wave_sync();

  // C. Row offsets: a slot's first output position, or the k-th spare
  // position past the row for the k-th slot without one; each slot's own
  // offset, its parent's and whether the parent is the previous slot stay in
  // that slot's lane (read back with readlane).  Output positions without a
  // joint (root links) get the identity now, in every row; positions that
  // repeat another position's joint are copied after the walk.
  const unsigned long long miss = __ballot(is_slot && perm && first[lane] == kNoPos);
  const bool bad = !perm || __popcll(miss) > n_extra;
  const unsigned long long screw = __ballot(my_kind == 3);
  const unsigned long long turn_slide = __ballot(my_kind == 2 || my_kind == 4);
  auto slot_pos = [&](int s) {
    const int f = first[s];
    return (f != kNoPos) ? 7 * f : 7 * (n_out + __popcll(miss & ((1ull << s) - 1ull)));
  };
  // pos | (parent pos + 1) << 12 | (parent is the previous slot) << 24 | (constant joint) << 25
  int link = 0;
  if (is_slot && !bad) {
    const int pos = slot_pos(lane);
    const int ppos = (ps >= 0) ? slot_pos(ps) : -1;
    recs[lane].pos = pos;
    link = pos | ((ppos + 1) << 12) | ((ps >= 0 && ps == lane - 1) ? (1 << 24) : 0) | ((my_kind == 0) ? (1 << 25) : 0);
  }
  unsigned long long ident[kOutIters], dup[kOutIters];
#pragma unroll
  for (int i = 0; i < kOutIters; ++i) {
    const int o = lane + i * kWave;
    const int s = lm[i];
    const bool is_dup = o < n_out && s >= 0 && first[s] != o;
    if (is_dup) fill[o] = 7 * first[s];
    ident[i] = __ballot(o < n_out && s < 0);
    dup[i] = __ballot(is_dup);
  }
  float* trow = tile + lane * stride;
  if (lane < n_valid && !bad) {
#pragma unroll
    for (int i = 0; i < kOutIters; ++i) {
      for (unsigned long long m = ident[i]; m; m &= m - 1ull) {
        float* d = trow + 7 * (i * kWave + __builtin_ctzll(m));
        d[0] = 1.f;
#pragma unroll
        for (int c = 1; c < 7; ++c) d[c] = 0.f;
      }
    }
  }
  // The configuration tile [a][lane]: element f of the block is
  // (configuration f / n_act, joint f % n_act); the quotient is a multiply
  // and shift (exact for f < 2^10, n_act <= 16).
  const unsigned div_m = (65536u + static_cast<unsigned>(n_act) - 1u) / static_cast<unsigned>(max(n_act, 1));
#pragma unroll
  for (int i = 0; i < kMaxAct; ++i) {
    const int f = lane + i * kWave;
    if (f < n_cfg) {
      const int l = static_cast<int>((static_cast<unsigned>(f) * div_m) >> 16), a = f - l * n_act;
      cfg_tile[a * kLaneStride + l] = cv[i];
    }
  }
  wave_sync();

  if (lane < n_valid) {
    if (!bad && J > 0) {
      // Local transforms of the revolute and prismatic slots, four slots at
      // a time (independent work the wave interleaves), into the slot's row
      // position; the chain pass below reads each one before it writes the
      // world pose there (a constant joint's local transform is its record's
      // parent_tf).  Both kinds share one branch-free form: th = 0 (prismatic)
      // and tq = 0 (revolute) reduce it to the other kind exactly.
      // K slots per pass, K = 4 while four remain, then 2 and 1 for the rest
      // (a padded last pass of four would compute repeated slots).
      auto local_pass = [&](auto width, unsigned long long& m) {
        constexpr int K = decltype(width)::value;
        float lc[K][7];
        int lp[K], ii[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {  // m holds at least K slots
          ii[k] = __builtin_ctzll(m);
          m &= m - 1ull;
        }
        bool big = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const FkRec& r = recs[ii[k]];
          const int src = uniform_i(r.src);
          const float qv = cfg_tile[max(src, 0) * kLaneStride + lane];
          const float q = ((src >= 0) ? qv : 0.f) * r.mul + r.off;
          const float x = 0.5f * (r.b[0] * q);
          big |= !(fabsf(x) <= kFkTrigMax);
          float s, c;
          sincos_cw(x, s, c);
          if (fabsf(x) < 0.5e-6f) { s = 0.f; c = 1.f; }  // the exponential's small-angle rule, |th| < 1e-6
          const float tq = (uniform_i(r.kind) == 2) ? q : 0.f;
          lc[k][0] = c * r.ptf[0] + s * r.a[0];
          lc[k][1] = c * r.ptf[1] + s * r.a[1];
          lc[k][2] = c * r.ptf[2] + s * r.a[2];
          lc[k][3] = c * r.ptf[3] + s * r.a[3];
          lc[k][4] = r.ptf[4] + tq * r.a[0];
          lc[k][5] = r.ptf[5] + tq * r.a[1];
          lc[k][6] = r.ptf[6] + tq * r.a[2];
          lp[k] = uniform_i(r.pos);
        }
        if (__any(big)) {
          // Joint values beyond the branch-free range: the library sine/cosine.
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const FkRec& r = recs[ii[k]];
            const int src = uniform_i(r.src);
            const float q = ((src >= 0) ? cfg_tile[src * kLaneStride + lane] : 0.f) * r.mul + r.off;
            const float x = 0.5f * (r.b[0] * q);
            float s, c;
            sincosf(x, &s, &c);
            if (fabsf(x) < 0.5e-6f) { s = 0.f; c = 1.f; }
            lc[k][0] = c * r.ptf[0] + s * r.a[0];
            lc[k][1] = c * r.ptf[1] + s * r.a[1];
            lc[k][2] = c * r.ptf[2] + s * r.a[2];
            lc[k][3] = c * r.ptf[3] + s * r.a[3];
          }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
          float* d = trow + lp[k];
#pragma unroll
          for (int c = 0; c < 7; ++c) d[c] = lc[k][c];
        }
      };
      unsigned long long m = turn_slide;
      while (__popcll(m) >= 4) local_pass(std::integral_constant<int, 4>{}, m);
      if (__popcll(m) >= 2) local_pass(std::integral_constant<int, 2>{}, m);
      if (m) local_pass(std::integral_constant<int, 1>{}, m);
      // General screw joints (rare): the full exponential.
      for (unsigned long long m = screw; m; m &= m - 1ull) {
        const FkRec& r = recs[__builtin_ctzll(m)];
        const int src = uniform_i(r.src);
        const float q = ((src >= 0) ? cfg_tile[src * kLaneStride + lane] : 0.f) * r.mul + r.off;
        const Pose ptf = load_pose(r.ptf);
        const Pose loc = compose(ptf, joint_exp(3, q, r.a[0], r.a[1], r.a[2], r.b[0], r.b[1], r.b[2]));
        float* d = trow + uniform_i(r.pos);
        d[0] = loc.w; d[1] = loc.x; d[2] = loc.y; d[3] = loc.z;
        d[4] = loc.tx; d[5] = loc.ty; d[6] = loc.tz;
      }
      // The chain: world = parent o local, in slot order.  The next slot's
      // local transform (from its row position, or a constant joint's
      // record) is read while this slot is composed.
      auto local_src = [&](int i, int lk) -> const float* {
        return ((lk >> 25) & 1) ? recs[i].ptf : trow + (lk & 0xfff);
      };
      Pose prev = pose_identity();
      Pose loc = load_pose(local_src(0, __builtin_amdgcn_readlane(link, 0)));
      // Unrolled by two so the carried poses alternate between two register
      // sets instead of being moved back at the end of every step.
#pragma unroll 2
      for (int i = 0; i < J; ++i) {
        const int lk = __builtin_amdgcn_readlane(link, i);
        const int in = min(i + 1, J - 1);
        const Pose nxt = load_pose(local_src(in, __builtin_amdgcn_readlane(link, in)));
        const int ppos = ((lk >> 12) & 0xfff) - 1;
        Pose tw = loc;
        if (ppos >= 0) {
          const Pose par = ((lk >> 24) & 1) ? prev : load_pose(trow + ppos);
          tw = compose(par, loc);
        }
        float* d = trow + (lk & 0xfff);
        d[0] = tw.w; d[1] = tw.x; d[2] = tw.y; d[3] = tw.z;
        d[4] = tw.tx; d[5] = tw.ty; d[6] = tw.tz;
        prev = tw;
        loc = nxt;
      }
#pragma unroll
      for (int i = 0; i < kOutIters; ++i) {
        for (unsigned long long m = dup[i]; m; m &= m - 1ull) {
          const int o = i * kWave + __builtin_ctzll(m);
          const int fo = uniform_i(fill[o]);
          float v[7];
#pragma unroll
          for (int c = 0; c < 7; ++c) v[c] = trow[fo + c];
          float* d = trow + 7 * o;
#pragma unroll
          for (int c = 0; c < 7; ++c) d[c] = v[c];
        }
      }
    } else if (bad) {
      for (int f = 0; f < row; ++f) trow[f] = __builtin_nanf("");
    }
  }
  wave_sync();

  // Copy-out of the block's n_valid * row contiguous floats.
  float* out_blk = out + b0 * row;
  const int total = n_valid * row;
  if (stride == row) {
    // The tile is the output image (the block starts 16-byte aligned: 64 rows
    // of 7 n_out floats).  Eight 16-byte LDS loads are issued before their
    // stores, so the wave waits for LDS once per eight stores.
    const int n4 = total / 4;
    const v4f* t4 = reinterpret_cast<const v4f*>(tile);
    v4f* o4 = reinterpret_cast<v4f*>(out_blk);
    for (int base = lane; base < n4; base += 8 * kWave) {
      v4f v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = t4[min(base + k * kWave, n4 - 1)];  // in the tile; stored only when < n4
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int f4 = base + k * kWave;
        if (f4 < n4) __builtin_nontemporal_store(v[k], o4 + f4);
      }
    }
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
