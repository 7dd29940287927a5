// Sphere-robot collision distances for gfx950.
//
// Work decomposition (the fused sweep_kernel; workgroups of kSweepWaves waves
// share the per-launch tables, each wave has its own buffers):
//   * each wave loops over its own block of configurations;
//   * for one configuration it places every robot sphere in the world frame
//     into LDS (one float4 per sphere: centre, radius, plus the expanded
//     entry of the link-frame distance test);
//   * lanes run across world obstacles -- each lane keeps T obstacles in
//     registers for the whole launch -- and every robot sphere is read once
//     from LDS as a broadcast and tested against the lane's T obstacles;
//   * per-link minima are written as coalesced rows (lanes = obstacles), the
//     per-configuration minimum is a wavefront min-reduction, and self pairs
//     are reduced in registers along each lane's run of the entry list and
//     published with LDS float atomic-min.
//
// Distance of two spheres: |c1 - c2| - r1 - r2 (positive = separated), the
// reference's sphere_sphere_dist (_collision_cuda_helpers.cuh:14-19).
#include "pk_model.h"

#include <math.h>

#include <algorithm>
#include <mutex>
#include <vector>

namespace pk {

constexpr float kBig = 1e9f;   // the reference's "no sphere" sentinel

__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_min(float v) {
  // Lanes outside ROW_MASK receive +inf, the identity of the minimum.
  return __builtin_elementwise_minimum(
      v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(INFINITY), __float_as_int(v),
                                                    CTRL, ROW_MASK, 0xf, false)));
}

// Minimum over the 64 lanes of a value that is never NaN, wave-uniform (read
// from lane 63), with the DPP steps of wave_sum_dpp.
__device__ __forceinline__ float wave_min(float v) {
  v = dpp_min<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = dpp_min<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v = dpp_min<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_min<0x140, 0xf>(v);  // row_mirror
  v = dpp_min<0x142, 0xa>(v);  // row_bcast15 into rows 1, 3
  v = dpp_min<0x143, 0xc>(v);  // row_bcast31 into rows 2, 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

struct SweepArgs {
  // source A: link poses + compact local sphere table
  const float* link_poses;   // (B, L, 7)
  const float* spheres_local;  // (K, 4)
  const int* link_sphere_start;  // (L + 1)
  // source B: world-frame sphere slots k = s * L + n
  const float* centers_soa;  // (3, B, K)
  const float* radii;        // (B, K)
  int s_per_link;            // S
  // common
  long long batch;
  int n_links;
  int n_spheres;             // K (compact for A, S * L for B)
  const float* world;        // (M, 4)
  int n_world;
  const unsigned* self_entries;
  int n_entries;
  int n_pairs;
  float* out_world;          // (B, L, M)
  float* out_min;            // (B)
  float* out_self;           // (B, P)
};

// Self entry i | j << 10 | pair << 20 with the sphere indices checked
// against K and the pair against P; entries out of range use the blank slot K
// (K < 1024) and a pair inside [0, P) (P < 2048: bit 31 is the run flag).
__device__ __forceinline__ unsigned self_entry(unsigned en, int K, int P) {
  const int i = static_cast<int>(en & 1023u), j = static_cast<int>((en >> 10) & 1023u);
  const int p = static_cast<int>(en >> 20);
  const bool ok = i < K && j < K && p < P;
  const unsigned bi = static_cast<unsigned>(ok ? i : K), bj = static_cast<unsigned>(ok ? j : K);
  return bi | (bj << 10) | (static_cast<unsigned>(max(min(p, P - 1), 0)) << 20);
}

// Expanded distance test of one link against a lane's T obstacles.
//
// With the link origin p, the rotated sphere offset w = R s and the obstacle
// offset q = o - p, |c - o|^2 = |w|^2 - 2 w.q + |q|^2: one float4 (-2w, |w|^2)
// per sphere turns a pair into three FMAs and a minimum, and |q|^2 is added
// once per group.  Rounding of that form is at most 12 u (r^2 + |q|^2), with
// u = 2^-24 and r^2 >= |w|^2 over the link; a squared distance a with
// sqrt(a) >= 0.4 (r^2 + |q|^2) + 1e-4 therefore has |sqrt(a) - |c - o|| <=
// 1e-6 m.  Groups below that bound (deep contacts) and NaN results are
// recomputed with the subtraction form, which keeps the skip-NaN minimum of
// the reference.
constexpr float kTauScale = 0.4f;
constexpr float kTauFloor = 1e-4f;

template <int T>
__device__ __forceinline__ void link_sweep(int n, int g0, int g1, const float4* sc, const float4* se,
                                           const int* gstart, const float* ptile, const float* lrad2,
                                           const float (&ox)[T], const float (&oy)[T], const float (&oz)[T],
                                           float (&acc)[T]) {
  const float px = ptile[7 * n + 4], py = ptile[7 * n + 5], pz = ptile[7 * n + 6];
  const float c0 = fmaf(kTauScale, lrad2[n], kTauFloor);
  float qx[T], qy[T], qz[T], qq[T], rt[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    qx[t] = ox[t] - px;
    qy[t] = oy[t] - py;
    qz[t] = oz[t] - pz;
    qq[t] = qx[t] * qx[t] + qy[t] * qy[t] + qz[t] * qz[t];
    rt[t] = fmaf(kTauScale, qq[t], c0);
  }
  for (int g = g0; g < g1; ++g) {
    const int k0 = gstart[g], k1 = gstart[g + 1];
    const float rg = sc[k0].w;
    float m[T];
    {
      const float4 e = se[k0];
#pragma unroll
      for (int t = 0; t < T; ++t) m[t] = fmaf(e.x, qx[t], fmaf(e.y, qy[t], fmaf(e.z, qz[t], e.w)));
    }
    int k = k0 + 1;
    for (; k + 1 < k1; k += 2) {  // two spheres per three-operand minimum
      const float4 e = se[k], f = se[k + 1];
#pragma unroll
      for (int t = 0; t < T; ++t)
        m[t] = __builtin_elementwise_minimum(
            m[t], __builtin_elementwise_minimum(fmaf(e.x, qx[t], fmaf(e.y, qy[t], fmaf(e.z, qz[t], e.w))),
                                                fmaf(f.x, qx[t], fmaf(f.y, qy[t], fmaf(f.z, qz[t], f.w)))));
    }
    if (k < k1) {
      const float4 e = se[k];
#pragma unroll
      for (int t = 0; t < T; ++t)
        m[t] = __builtin_elementwise_minimum(m[t], fmaf(e.x, qx[t], fmaf(e.y, qy[t], fmaf(e.z, qz[t], e.w))));
    }
    float v[T];
    bool bad[T];
    bool exact = false;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const float d = fast_sqrt(m[t] + qq[t]);
      bad[t] = !(d >= rt[t]);
      exact = exact || bad[t];
      v[t] = d - rg;
    }
    if (__any(exact)) {  // rare: some lane is in deep contact or saw a NaN
#pragma unroll
      for (int t = 0; t < T; ++t) {
        if (bad[t]) {
          float e2 = INFINITY;
          for (int j = k0; j < k1; ++j) {
            const float4 s = sc[j];
            const float dx = s.x - ox[t], dy = s.y - oy[t], dz = s.z - oz[t];
            e2 = fminf(e2, dx * dx + dy * dy + dz * dz);
          }
          const float x = fast_sqrt(e2) - rg;
          v[t] = (x == x) ? x : INFINITY;  // inf - inf is skipped like a NaN
        }
      }
    }
#pragma unroll
    for (int t = 0; t < T; ++t) acc[t] = __builtin_elementwise_minimum(acc[t], v[t]);
  }
}

// Waves per sweep workgroup.  They share the per-launch tables (local
// spheres, self entries, groups) in LDS and each works through its own block
// of configurations with private sphere, pose and pair-minimum buffers, so a
// CU holds several times more waves than with one table copy per wave.
constexpr int kSweepWaves = 4;

// SRC = 0: link poses + local spheres (fused path); SRC = 1: precomputed
// world-frame sphere slots (the reference FFI layout).  T = obstacles per lane.
// Six waves per SIMD (at most 80 VGPRs) for up to four obstacles per lane,
// three with eight, which need the registers.
template <int SRC, int T>
__global__ __launch_bounds__(kWave * kSweepWaves, T == 8 ? 3 : 6) void sweep_kernel(SweepArgs a) {
  extern __shared__ float smem[];
  const int tid = threadIdx.x;
  // The wave index is uniform across the wave: in a scalar register it keeps
  // the wave's buffer addresses and configuration range out of VGPRs.
  const int lane = tid % kWave, wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int L = a.n_links;
  const int K = a.n_spheres;
  const int M = a.n_world;
  const int E = a.n_entries;
  const int P = a.n_pairs;
  // LDS carve-up: shared tables, then one float4 block and one word block per wave.
  const int f4_per_wave = (K + 1) + (SRC == 0 ? K : 0);
  const int words_per_wave = (SRC == 0 ? 7 * L : 0) + P;
  float4* sl = reinterpret_cast<float4*>(smem);                          // K local spheres (SRC 0)
  float4* sc = sl + (SRC == 0 ? K : 0) + wave * f4_per_wave;            // K world spheres + 1 blank
  float4* se = sc + (K + 1);                                             // K expanded entries (SRC 0)
  const int n_it = (E + kWave - 1) / kWave;                             // self steps per lane
  unsigned* sent = reinterpret_cast<unsigned*>(sl + (SRC == 0 ? K : 0) + kSweepWaves * f4_per_wave);  // n_it * 64
  int* lstart = reinterpret_cast<int*>(sent + n_it * kWave);             // L + 1
  int* slink = lstart + (L + 1);                                         // K: link of sphere (SRC 0)
  int* gstart = slink + (SRC == 0 ? K : 0);                              // K + 1: sphere groups
  int* lgroup = gstart + (K + 1);                                        // L + 1: groups of each link
  float* lrad2 = reinterpret_cast<float*>(lgroup + (L + 1));             // L: bound of |offset|^2 (SRC 0)
  float* ptile = lrad2 + (SRC == 0 ? L : 0) + wave * words_per_wave;     // L * 7 poses (SRC 0)
  float* pmin = ptile + (SRC == 0 ? 7 * L : 0);                          // P

  // Per-launch tables, staged by the whole workgroup.
  const int nt = kWave * kSweepWaves;
  if (SRC == 0) {
    // A NaN radius makes every distance of its sphere NaN, which the minima
    // skip; a radius of -inf makes them +inf, which the minima skip as well.
    for (int k = tid; k < K; k += nt) {
      const float* s = a.spheres_local + 4 * k;
      sl[k] = make_float4(s[0], s[1], s[2], s[3] != s[3] ? -INFINITY : s[3]);
    }
    // Clamp the CSR offsets into a monotone [0, K] sequence.
    if (tid == 0) {
      int prev = 0;
      for (int n = 0; n <= L; ++n) {
        const int v = (n == L) ? K : a.link_sphere_start[n];
        prev = max(prev, min(v, K));
        lstart[n] = (n == 0) ? 0 : prev;
      }
    }
    for (int k = tid; k < K; k += nt) slink[k] = 0;
  } else {
    for (int n = tid; n <= L; n += nt) lstart[n] = n * a.s_per_link;
  }
  // The blank slot K is infinitely far from everything: distances to it are
  // +inf, which no minimum takes.
  if (lane == 0) sc[K] = make_float4(0.f, 0.f, 0.f, -INFINITY);
  // Self entries, transposed: lane l owns the contiguous run
  // [l E / 64, (l + 1) E / 64) of the entry list (grouped by pair), and step
  // `it` of all 64 runs sits at sent[it * 64 + l].  Bits: sphere i (0-9),
  // sphere j (10-19), pair (20-30), and bit 31 on the last entry of a pair
  // within a run, where the lane publishes its minimum.  Slots past the end
  // of a shorter run hold the blank sphere pair (distance +inf, no flag).
  for (int s = tid; s < n_it * kWave; s += nt) {
    const int it = s / kWave, l = s - it * kWave;
    const int e0 = static_cast<int>((static_cast<long long>(E) * l) / kWave);
    const int e1 = static_cast<int>((static_cast<long long>(E) * (l + 1)) / kWave);
    const int e = e0 + it;
    unsigned v = static_cast<unsigned>(K) | (static_cast<unsigned>(K) << 10);
    if (e < e1) {
      v = self_entry(a.self_entries[e], K, P);
      const bool last = (e + 1 == e1) || ((self_entry(a.self_entries[e + 1], K, P) >> 20) != (v >> 20));
      if (last) v |= 0x80000000u;
    }
    sent[s] = v;
  }
  __syncthreads();
  if (SRC == 0) {
    for (int n = tid; n < L; n += nt) {
      float r2 = 0.f;
      for (int k = lstart[n]; k < lstart[n + 1]; ++k) {
        slink[k] = n;
        const float4 s = sl[k];
        r2 = fmaxf(r2, s.x * s.x + s.y * s.y + s.z * s.z);
      }
      // Slack for the rounding of the rotated offsets.
      lrad2[n] = r2 * 1.001f + 1e-12f;
    }
  }
  // Sphere groups: runs of equal radius inside a link (SRC 0), or single
  // slots (SRC 1, whose radii are per configuration).  For one obstacle,
  // min_k(|c_k - o|) - r = sqrt(min_k |c_k - o|^2) - r inside a group, so a
  // group costs one square root per obstacle.
  if (tid == 0) {
    int g = 0;
    for (int n = 0; n < L; ++n) {
      lgroup[n] = g;
      for (int k = lstart[n]; k < lstart[n + 1]; ++k)
        if (SRC == 1 || k == lstart[n] || sl[k].w != sl[k - 1].w) gstart[g++] = k;
    }
    lgroup[L] = g;
    gstart[g] = K;
  }

  // This lane's obstacles.  The launcher picks T with M <= 64 T for T < 8,
  // so only T = 8 sweeps the world in several chunks and reloads them.
  const int chunk = kWave * T;
  const int n_chunks = (M + chunk - 1) / chunk;
  float ox[T], oy[T], oz[T], orad[T];
  auto load_obstacles = [&](int c) {
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int m = c * chunk + t * kWave + lane;
      if (m < M) {
        ox[t] = a.world[4 * m]; oy[t] = a.world[4 * m + 1];
        oz[t] = a.world[4 * m + 2]; orad[t] = a.world[4 * m + 3];
      } else {  // never the minimum: infinitely far
        ox[t] = 0.f; oy[t] = 0.f; oz[t] = 0.f; orad[t] = -INFINITY;
      }
    }
  };
  if (n_chunks > 0) load_obstacles(0);
  __syncthreads();

  // Each wave owns a contiguous block of configurations, so the pose rows it
  // reads and the distance rows it writes share cache lines.  From here on
  // the waves of a workgroup run independently.
  const long long n_waves = static_cast<long long>(gridDim.x) * kSweepWaves;
  const long long per = (a.batch + n_waves - 1) / n_waves;
  const long long b_begin = (static_cast<long long>(blockIdx.x) * kSweepWaves + wave) * per;
  const long long b_end = min(a.batch, b_begin + per);
  for (long long b = b_begin; b < b_end; ++b) {
    // 1. Robot spheres of configuration b in the world frame -> LDS.
    if (SRC == 0) {
      const float* pb = a.link_poses + b * L * 7;
      for (int f = lane; f < L * 7; f += kWave) ptile[f] = pb[f];
      wave_sync();
      for (int k = lane; k < K; k += kWave) {
        const float* p = ptile + 7 * slink[k];
        const float4 s = sl[k];
        float wx, wy, wz;
        quat_rotate(p[0], p[1], p[2], p[3], s.x, s.y, s.z, wx, wy, wz);
        sc[k] = make_float4(wx + p[4], wy + p[5], wz + p[6], s.w);
        se[k] = make_float4(-2.f * wx, -2.f * wy, -2.f * wz, wx * wx + wy * wy + wz * wz);
      }
    } else {
      const long long BK = a.batch * K;
      for (int k = lane; k < K; k += kWave) {
        // slot k = s * L + n is stored at LDS position n * S + s
        const int n = k % L, s = k / L;
        const long long g = b * K + k;
        const float r = a.radii[g];
        sc[n * a.s_per_link + s] =
            make_float4(a.centers_soa[g], a.centers_soa[BK + g], a.centers_soa[2 * BK + g],
                        r < 0.f ? -INFINITY : r);
      }
    }
    for (int p = lane; p < P; p += kWave) pmin[p] = kBig;
    wave_sync();

    // 2. World sweep.
    float gmin = INFINITY;
    if (a.out_world != nullptr || a.out_min != nullptr) {
      for (int c = 0; c < n_chunks; ++c) {
        if (T == 8 && c > 0) load_obstacles(c);
        if (SRC == 0 && a.out_world == nullptr) {
          // Only the configuration minimum is wanted.  Subtracting r_o is
          // monotone, so min_n fminf(kBig, acc_n - r_o) equals
          // fminf(kBig, min_n acc_n - r_o) (NaN and infinite cases included):
          // one accumulator runs through all links.
          float acc[T];
#pragma unroll
          for (int t = 0; t < T; ++t) acc[t] = INFINITY;
          for (int n = 0; n < L; ++n)
            if (lgroup[n] < lgroup[n + 1])
              link_sweep<T>(n, lgroup[n], lgroup[n + 1], sc, se, gstart, ptile, lrad2, ox, oy, oz, acc);
#pragma unroll
          for (int t = 0; t < T; ++t) gmin = __builtin_elementwise_minimum(gmin, fminf(kBig, acc[t] - orad[t]));
          continue;
        }
        for (int n = 0; n < L; ++n) {
          float acc[T];
#pragma unroll
          for (int t = 0; t < T; ++t) acc[t] = INFINITY;
          const int g1 = lgroup[n + 1];
          if (SRC == 0) {
            if (lgroup[n] < g1) link_sweep<T>(n, lgroup[n], g1, sc, se, gstart, ptile, lrad2, ox, oy, oz, acc);
          } else for (int g = lgroup[n]; g < g1; ++g) {
            const int k0 = gstart[g], k1 = gstart[g + 1];
            const float rg = sc[k0].w;
            if (SRC == 1 && rg == -INFINITY) continue;  // padding slot (wave-uniform)
            float m2[T];
#pragma unroll
            for (int t = 0; t < T; ++t) m2[t] = INFINITY;
            for (int k = k0; k < k1; ++k) {
              const float4 s = sc[k];
#pragma unroll
              for (int t = 0; t < T; ++t) {
                const float dx = s.x - ox[t], dy = s.y - oy[t], dz = s.z - oz[t];
                m2[t] = fminf(m2[t], dx * dx + dy * dy + dz * dz);
              }
            }
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t] = fminf(acc[t], fast_sqrt(m2[t]) - rg);
          }
          // The reference's per-link minimum starts at kBig and skips NaN
          // (_collision_cuda_kernel.cu:186-197): v is never NaN, so the
          // per-configuration minimum needs no NaN handling.  Padding lanes
          // (orad = -inf) give kBig.
          float v[T];
#pragma unroll
          for (int t = 0; t < T; ++t) v[t] = fminf(kBig, acc[t] - orad[t]);
          if (a.out_world != nullptr) {
#pragma unroll
            for (int t = 0; t < T; ++t) {
              const int m = c * chunk + t * kWave + lane;
              if (m < M) a.out_world[(b * L + n) * M + m] = v[t];
            }
          }
#pragma unroll
          for (int t = 0; t < T; ++t) gmin = __builtin_elementwise_minimum(gmin, v[t]);
        }
      }
      if (a.out_min != nullptr) {
        const float w = wave_min(gmin);
        if (lane == 0) a.out_min[b] = fminf(kBig, w);
      }
      if (T == 8 && n_chunks > 1) load_obstacles(0);
    }

    // 3. Self pairs.  Every lane takes n_it steps through its run of the
    // transposed entry table (a scalar loop), keeps the running minimum of
    // its current pair in a register and publishes it with an LDS float
    // atomic-min at the pair's last entry.  The running minimum skips NaN
    // distances; published values are never NaN.
    if (a.out_self != nullptr) {
      float cur = INFINITY;
      for (int it = 0; it < n_it; ++it) {
        const unsigned w = sent[it * kWave + lane];
        const float4 s1 = sc[w & 1023u];
        const float4 s2 = sc[__builtin_amdgcn_ubfe(w, 10, 10)];
        const float dx = s1.x - s2.x, dy = s1.y - s2.y, dz = s1.z - s2.z;
        cur = fminf(cur, fast_sqrt(dx * dx + dy * dy + dz * dz) - s1.w - s2.w);
        if (static_cast<int>(w) < 0) {
          atomicMin(&pmin[(w >> 20) & 2047u], cur);
          cur = INFINITY;
        }
      }
      wave_sync();
      for (int q = lane; q < P; q += kWave) a.out_self[b * P + q] = pmin[q];
    }
    wave_sync();
  }
}

// Reference-layout self distances: one lane per (configuration, pair), the
// full S x S minimum with padding skipped (_collision_cuda_kernel.cu:425-459).
__global__ __launch_bounds__(256) void self_sphere_aos_kernel(
    const float* __restrict__ centers, const float* __restrict__ radii, long long batch,
    int S, int N, const int* __restrict__ pair_i, const int* __restrict__ pair_j, int P,
    float* __restrict__ out) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= batch * P) return;
  const int p = static_cast<int>(idx % P);
  const long long b = idx / P;
  const int li = pair_i[p], lj = pair_j[p];
  float best = kBig;
  if (li >= 0 && li < N && lj >= 0 && lj < N) {
    for (int si = 0; si < S; ++si) {
      const long long ki = (b * S + si) * N + li;
      const float ri = radii[ki];
      if (ri < 0.f) continue;
      const float ax = centers[3 * ki], ay = centers[3 * ki + 1], az = centers[3 * ki + 2];
      for (int sj = 0; sj < S; ++sj) {
        const long long kj = (b * S + sj) * N + lj;
        const float rj = radii[kj];
        if (rj < 0.f) continue;
        const float dx = ax - centers[3 * kj], dy = ay - centers[3 * kj + 1],
                    dz = az - centers[3 * kj + 2];
        best = fminf(best, fast_sqrt(dx * dx + dy * dy + dz * dz) - (ri + rj));
      }
    }
  }
  out[idx] = best;
}

static int pick_t(int M) {
  if (M <= 64) return 1;
  if (M <= 128) return 2;
  if (M <= 256) return 4;
  return 8;
}

// Workgroups of `fn` resident at once on the current device (per-CU
// occupancy for `block` threads and `lds` bytes of LDS, at most `max_per_cu`,
// times the CU count), or 0 if the runtime cannot say.  Cached per (device,
// kernel, LDS size).
static long long resident_groups(const void* fn, int block, size_t lds, int max_per_cu) {
  struct Entry { int dev; const void* fn; size_t lds; long long groups; };
  static std::mutex mu;
  static std::vector<Entry> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  std::lock_guard<std::mutex> lock(mu);
  for (const Entry& e : cache)
    if (e.dev == dev && e.fn == fn && e.lds == lds) return e.groups;
  int n_cu = 0, per_cu = 0;
  long long groups = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds) == hipSuccess)
    groups = static_cast<long long>(n_cu) * std::min(per_cu, max_per_cu);
  else
    (void)hipGetLastError();
  if (cache.size() < 256) cache.push_back({dev, fn, lds, groups});
  return groups;
}

// One launch of sweep_kernel<SRC, T>: at least eight configurations per wave,
// and no more workgroups than are resident at once (each wave loops over its
// block of configurations, so a second, partial round would only run at low
// occupancy).  Residency is capped at six workgroups per CU (six waves per
// SIMD): measured on Panda, seven -- possible when the kernel needs at most
// 72 VGPRs -- made the launch 12% slower, the self phase included.
constexpr int kSweepGroupsPerCu = 6;

template <int SRC, int T>
static void launch_sweep_t(hipStream_t s, const SweepArgs& a, size_t lds) {
  const void* fn = reinterpret_cast<const void*>(&sweep_kernel<SRC, T>);
  const long long want = (a.batch + 8 * kSweepWaves - 1) / (8 * kSweepWaves);
  long long cap = resident_groups(fn, kWave * kSweepWaves, lds, kSweepGroupsPerCu);
  if (cap <= 0) cap = 8192 / kSweepWaves;
  const unsigned grid = static_cast<unsigned>(std::max(1LL, std::min(want, cap)));
  hipLaunchKernelGGL((sweep_kernel<SRC, T>), dim3(grid), dim3(kWave * kSweepWaves), lds, s, a);
}

template <int SRC>
static int launch_sweep(void* stream, const SweepArgs& a) {
  const int K = a.n_spheres, L = a.n_links;
  const size_t per_wave = sizeof(float4) * ((K + 1) + (SRC == 0 ? K : 0)) +
                          sizeof(float) * ((SRC == 0 ? 7 * L : 0) + a.n_pairs);
  const size_t n_slots = static_cast<size_t>((a.n_entries + kWave - 1) / kWave) * kWave;
  size_t lds = sizeof(float4) * (SRC == 0 ? K : 0) + sizeof(unsigned) * n_slots +
               sizeof(int) * ((L + 1) + (SRC == 0 ? K : 0) + (K + 1) + (L + 1)) +
               sizeof(float) * (SRC == 0 ? L : 0) + per_wave * kSweepWaves;
  PK_REQUIRE(a.n_entries == 0 || K < 1024, "self entries address at most 1023 spheres (got %d)", K);
  PK_REQUIRE(a.n_entries == 0 || a.n_pairs < 2048, "self entries address at most 2047 pairs (got %d)", a.n_pairs);
  PK_REQUIRE(lds <= 150 * 1024, "sphere tables need %zu bytes of LDS (limit 150 KiB)", lds);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (pick_t(a.n_world)) {
    case 1: launch_sweep_t<SRC, 1>(s, a, lds); break;
    case 2: launch_sweep_t<SRC, 2>(s, a, lds); break;
    case 4: launch_sweep_t<SRC, 4>(s, a, lds); break;
    default: launch_sweep_t<SRC, 8>(s, a, lds); break;
  }
  return check_launch("sweep_kernel");
}

int launch_world_reduced_mixed(void* stream, const float* sc, const float* sr, long long batch, int k,
                               int n_links, const float* ws, int ms, const float* wc, int mc, const float* wb,
                               int mb, const float* wh, int mh, float* out);

}  // namespace pk

extern "C" int pk_collision_world_sphere_reduced_f32(
    void* stream, const float* sphere_centers, const float* sphere_radii, int64_t batch,
    int32_t k, const float* world_spheres, int32_t n_world_spheres,
    const float* world_capsules, int32_t n_world_capsules, const float* world_boxes,
    int32_t n_world_boxes, const float* world_halfspaces, int32_t n_world_halfspaces,
    int32_t n_links, float* out) {
  PK_REQUIRE(batch >= 0 && k >= 0 && n_world_spheres >= 0 && n_world_capsules >= 0 &&
                 n_world_boxes >= 0 && n_world_halfspaces >= 0 && n_links >= 0,
             "negative size");
  const long long n_world =
      static_cast<long long>(n_world_spheres) + n_world_capsules + n_world_boxes + n_world_halfspaces;
  if (batch == 0 || n_links == 0 || n_world == 0) return PK_OK;
  PK_REQUIRE(k % n_links == 0, "k=%d is not a multiple of n_links=%d", k, n_links);
  PK_REQUIRE(sphere_centers && sphere_radii && out, "null pointer");
  if (n_world != n_world_spheres)  // capsules, boxes or half-spaces: the mixed-kind kernel
    return pk::launch_world_reduced_mixed(stream, sphere_centers, sphere_radii, batch, k, n_links,
                                          world_spheres, n_world_spheres, world_capsules,
                                          n_world_capsules, world_boxes, n_world_boxes,
                                          world_halfspaces, n_world_halfspaces, out);
  PK_REQUIRE(world_spheres != nullptr, "null pointer");
  pk::SweepArgs a{};
  a.centers_soa = sphere_centers;
  a.radii = sphere_radii;
  a.s_per_link = k / n_links;
  a.batch = batch;
  a.n_links = n_links;
  a.n_spheres = k;
  a.world = world_spheres;
  a.n_world = n_world_spheres;
  a.out_world = out;
  return pk::launch_sweep<1>(stream, a);
}

extern "C" int pk_collision_self_sphere_f32(void* stream, const float* centers,
                                            const float* radii, int64_t batch, int32_t s,
                                            int32_t n_links, const int32_t* pair_i,
                                            const int32_t* pair_j, int32_t n_pairs,
                                            float* out) {
  PK_REQUIRE(batch >= 0 && s >= 0 && n_links >= 0 && n_pairs >= 0, "negative size");
  if (batch == 0 || n_pairs == 0) return PK_OK;
  PK_REQUIRE(centers && radii && pair_i && pair_j && out, "null pointer");
  const long long total = batch * static_cast<long long>(n_pairs);
  const long long blocks = (total + 255) / 256;
  PK_REQUIRE(blocks <= 0x7fffffffLL, "batch too large");
  hipLaunchKernelGGL(pk::self_sphere_aos_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), centers, radii, static_cast<long long>(batch),
                     s, n_links, pair_i, pair_j, n_pairs, out);
  return pk::check_launch("self_sphere_aos_kernel");
}

extern "C" int pk_sphere_sweep_f32(void* stream, const float* link_poses, int64_t batch,
                                   int32_t n_links, const float* spheres_local,
                                   const int32_t* link_sphere_start, int32_t n_spheres,
                                   const float* world_spheres, int32_t n_world_spheres,
                                   const uint32_t* self_entries, int32_t n_entries,
                                   int32_t n_pairs, float* out_world, float* out_world_min,
                                   float* out_self) {
  PK_REQUIRE(batch >= 0 && n_links >= 0 && n_spheres >= 0 && n_world_spheres >= 0 &&
                 n_entries >= 0 && n_pairs >= 0,
             "negative size");
  PK_REQUIRE(n_spheres <= 1024 && n_pairs <= 4096 && (n_entries == 0 || n_pairs <= 2047),
             "self-entry encoding needs n_spheres <= 1024 and n_pairs <= 2047 (4096 without entries)");
  PK_REQUIRE(n_links <= pk::kMaxJoints + 1, "n_links=%d too large", n_links);
  if (batch == 0) return PK_OK;
  PK_REQUIRE(link_poses && link_sphere_start && (n_spheres == 0 || spheres_local),
             "null pointer");
  if (out_world || out_world_min) PK_REQUIRE(n_world_spheres == 0 || world_spheres, "null world");
  if (out_self) PK_REQUIRE(n_entries == 0 || self_entries, "null self_entries");
  // Every entry names a pair slot of out_self, so entries need at least one pair.
  if (out_self) PK_REQUIRE(n_entries == 0 || n_pairs >= 1, "n_entries=%d with n_pairs=0", n_entries);
  pk::SweepArgs a{};
  a.link_poses = link_poses;
  a.spheres_local = spheres_local;
  a.link_sphere_start = link_sphere_start;
  a.batch = batch;
  a.n_links = n_links;
  a.n_spheres = n_spheres;
  a.world = world_spheres;
  a.n_world = n_world_spheres;
  a.self_entries = out_self ? self_entries : nullptr;
  a.n_entries = out_self ? n_entries : 0;
  a.n_pairs = out_self ? n_pairs : 0;
  a.out_world = out_world;
  a.out_min = out_world_min;
  a.out_self = out_self;
  return pk::launch_sweep<0>(stream, a);
}
