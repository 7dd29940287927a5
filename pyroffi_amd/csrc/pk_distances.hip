// Output-shaped sphere distance kernels for gfx950: the reference FFI targets
// whose result is a full distance tensor ([B, K, M] per sphere, [B, N, M] per
// link over mixed obstacle kinds), and the JAX path's distance conventions.
//
// These outputs are large and each element is a few flops, so the kernels are
// bound by the output write: a flat grid-stride loop gives consecutive lanes
// consecutive output elements (coalesced 256-byte rows per wave), and the
// inputs of an element (one robot sphere, one obstacle) come through the
// caches.  The fused sweep in pk_collision.hip remains the fast path for
// sphere-only worlds and reduced outputs.
#include "pk_geom.h"
#include "pk_model.h"

namespace pk {

constexpr float kNoSphere = 1e9f;  // the reference's "no sphere" value

// out[b, k, m] = distance of robot sphere k to obstacle m; padding (r < 0)
// writes 1e9 (_collision_cuda_kernel.cu:64-161).
__global__ __launch_bounds__(256) void world_sphere_full_kernel(const float* __restrict__ sc,
                                                                const float* __restrict__ sr,
                                                                long long BK, WorldPrims w,
                                                                float* __restrict__ out) {
  const long long M = w.total();
  const long long total = BK * M;
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < total; f += stride) {
    const long long row = f / M;
    const int m = static_cast<int>(f - row * M);
    const float r = sr[row];
    float d = kNoSphere;
    if (r >= 0.f) d = dist_sphere_obstacle(sc[row], sc[BK + row], sc[2 * BK + row], r, w, m);
    __builtin_nontemporal_store(d, out + f);
  }
}

// out[b, n, m] = min over the link's sphere slots k = s * N + n (padding
// skipped, 1e9 when none is real) of the distance to obstacle m, for worlds
// with non-sphere obstacles (_collision_cuda_kernel.cu:169-407).
__global__ __launch_bounds__(256) void world_reduced_mixed_kernel(const float* __restrict__ sc,
                                                                  const float* __restrict__ sr,
                                                                  long long B, int K, int S, int N,
                                                                  WorldPrims w, float* __restrict__ out) {
  const long long M = w.total();
  const long long BK = B * K;
  const long long total = B * N * M;
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < total; f += stride) {
    const long long bn = f / M;
    const int m = static_cast<int>(f - bn * M);
    const long long b = bn / N;
    const int n = static_cast<int>(bn - b * N);
    float best = kNoSphere;
    for (int s = 0; s < S; ++s) {
      const long long k = b * K + static_cast<long long>(s) * N + n;
      const float r = sr[k];
      if (r < 0.f) continue;
      best = fminf(best, dist_sphere_obstacle(sc[k], sc[BK + k], sc[2 * BK + k], r, w, m));
    }
    __builtin_nontemporal_store(best, out + f);
  }
}

// Centre of sphere slot s of link l for configuration b: T_l applied to the
// link-frame centre of the padded (L, S, 4) table.
__device__ __forceinline__ void placed(const float* __restrict__ poses, const float* __restrict__ table,
                                       long long b, int L, int S, int l, int s, float& x, float& y, float& z,
                                       float& r) {
  const float* T = poses + (b * L + l) * 7;
  const float* c = table + (static_cast<long long>(l) * S + s) * 4;
  quat_rotate(T[0], T[1], T[2], T[3], c[0], c[1], c[2], x, y, z);
  x += T[4];
  y += T[5];
  z += T[6];
  r = c[3];
}

// The JAX distance of two spheres (src/pyroffi/collision/_geometry_pairs.py:59-68
// with the 1e-6 added to each component inside the norm, _utils.py:45):
// |(p2 - p1) + 1e-6| - (r1 + r2).  Padding radii (-1e9) take part as they do
// in JAX and make the distance about +1e9.
__device__ __forceinline__ float jax_sphere_dist(float x1, float y1, float z1, float r1, float x2, float y2,
                                                 float z2, float r2) {
  const float dx = (x2 - x1) + 1e-6f, dy = (y2 - y1) + 1e-6f, dz = (z2 - z1) + 1e-6f;
  return sqrtf(dx * dx + dy * dy + dz * dz) - (r1 + r2);
}

// out[b, l, m] = min over slots s of jax_sphere_dist(sphere (l, s), world m)
// (RobotCollisionSpherized.compute_world_collision_distance,
// src/pyroffi/collision/_robot_collision.py:830-895).
__global__ __launch_bounds__(256) void jax_world_kernel(const float* __restrict__ poses, long long B, int L,
                                                        const float* __restrict__ table, int S,
                                                        const float* __restrict__ world, int M,
                                                        float* __restrict__ out) {
  const long long total = B * L * static_cast<long long>(M);
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < total; f += stride) {
    const long long bl = f / M;
    const int m = static_cast<int>(f - bl * M);
    const long long b = bl / L;
    const int l = static_cast<int>(bl - b * L);
    const float* o = world + 4 * m;
    float best = INFINITY;
    for (int s = 0; s < S; ++s) {
      float x, y, z, r;
      placed(poses, table, b, L, S, l, s, x, y, z, r);
      best = fminf(best, jax_sphere_dist(x, y, z, r, o[0], o[1], o[2], o[3]));
    }
    __builtin_nontemporal_store(best, out + f);
  }
}

// out[b, p] = min over slots s of jax_sphere_dist(sphere (i_p, s), sphere (j_p, s)):
// the same-slot ("diagonal") pairs of
// RobotCollisionSpherized.compute_self_collision_distance
// (src/pyroffi/collision/_robot_collision.py:796-828).
__global__ __launch_bounds__(256) void jax_self_kernel(const float* __restrict__ poses, long long B, int L,
                                                       const float* __restrict__ table, int S,
                                                       const int* __restrict__ pair_i,
                                                       const int* __restrict__ pair_j, int P,
                                                       float* __restrict__ out) {
  const long long total = B * static_cast<long long>(P);
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < total; f += stride) {
    const long long b = f / P;
    const int p = static_cast<int>(f - b * P);
    const int li = pair_i[p], lj = pair_j[p];
    float best = INFINITY;
    if (li >= 0 && li < L && lj >= 0 && lj < L) {
      for (int s = 0; s < S; ++s) {
        float xi, yi, zi, ri, xj, yj, zj, rj;
        placed(poses, table, b, L, S, li, s, xi, yi, zi, ri);
        placed(poses, table, b, L, S, lj, s, xj, yj, zj, rj);
        best = fminf(best, jax_sphere_dist(xi, yi, zi, ri, xj, yj, zj, rj));
      }
    }
    __builtin_nontemporal_store(best, out + f);
  }
}

// Capsule robot vs the world: out[b, n, m] for link capsule caps[:, b, n]
// (component-major [7, B, N]) (_collision_cuda_kernel.cu:203-407).
__global__ __launch_bounds__(256) void world_capsule_kernel(const float* __restrict__ caps, long long BN,
                                                            WorldPrims w, float* __restrict__ out) {
  const long long M = w.total();
  const long long total = BN * M;
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < total; f += stride) {
    const long long bn = f / M;
    const int m = static_cast<int>(f - bn * M);
    float c[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) c[i] = caps[i * BN + bn];
    __builtin_nontemporal_store(dist_capsule_obstacle(c, w, m), out + f);
  }
}

// Capsule robot self pairs: out[b, p] = capsule-capsule distance of links
// (pair_i[p], pair_j[p]) from AoS capsules [B, N, 7] (_collision_cuda_kernel.cu:474-495).
__global__ __launch_bounds__(256) void self_capsule_kernel(const float* __restrict__ caps, long long B, int N,
                                                           const int* __restrict__ pair_i,
                                                           const int* __restrict__ pair_j, int P,
                                                           float* __restrict__ out) {
  const long long total = B * static_cast<long long>(P);
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < total; f += stride) {
    const long long b = f / P;
    const int p = static_cast<int>(f - b * P);
    const int li = pair_i[p], lj = pair_j[p];
    float d = kNoSphere;
    if (li >= 0 && li < N && lj >= 0 && lj < N) {
      float ci[7], cj[7];
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        ci[i] = caps[(b * N + li) * 7 + i];
        cj[i] = caps[(b * N + lj) * 7 + i];
      }
      d = dist_capsule_capsule(ci, cj);
    }
    __builtin_nontemporal_store(d, out + f);
  }
}

// CUDA-semantics world distances from link poses over mixed obstacle kinds:
// out[b, l, m] = min over the link's real sphere slots (r >= 0) of the
// distance to obstacle m, 1e9 for a link without spheres; the spheres are
// placed from the padded (L, S, 4) link-frame table.
__global__ __launch_bounds__(256) void world_links_mixed_kernel(const float* __restrict__ poses, long long B,
                                                                int L, const float* __restrict__ table, int S,
                                                                WorldPrims w, float* __restrict__ out) {
  const long long M = w.total();
  const long long total = B * L * M;
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < total; f += stride) {
    const long long bl = f / M;
    const int m = static_cast<int>(f - bl * M);
    const long long b = bl / L;
    const int l = static_cast<int>(bl - b * L);
    float best = kNoSphere;
    for (int s = 0; s < S; ++s) {
      float x, y, z, r;
      placed(poses, table, b, L, S, l, s, x, y, z, r);
      if (r < 0.f) continue;
      best = fminf(best, dist_sphere_obstacle(x, y, z, r, w, m));
    }
    __builtin_nontemporal_store(best, out + f);
  }
}

// Capsule robot at a configuration: each link-frame capsule [a, b, r] of the
// (N, 7) table moved by its link pose; written component-major [7, B, N]
// (the world kernel's layout) and/or AoS [B, N, 7] (the self kernel's).
__global__ __launch_bounds__(256) void capsule_place_kernel(const float* __restrict__ poses, long long B, int N,
                                                            const float* __restrict__ local,
                                                            float* __restrict__ out_soa,
                                                            float* __restrict__ out_aos) {
  const long long total = B * N;
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < total; f += stride) {
    const int n = static_cast<int>(f % N);
    const float* T = poses + f * 7;
    const float* c = local + static_cast<long long>(n) * 7;
    float v[7];
    quat_rotate(T[0], T[1], T[2], T[3], c[0], c[1], c[2], v[0], v[1], v[2]);
    quat_rotate(T[0], T[1], T[2], T[3], c[3], c[4], c[5], v[3], v[4], v[5]);
    v[0] += T[4];
    v[1] += T[5];
    v[2] += T[6];
    v[3] += T[4];
    v[4] += T[5];
    v[5] += T[6];
    v[6] = c[6];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      if (out_soa) out_soa[k * total + f] = v[k];
      if (out_aos) out_aos[f * 7 + k] = v[k];
    }
  }
}

// Per-configuration minimum over the (L, M) block of world_links_mixed_kernel's
// output: one wave per configuration, a strided scan and a wave minimum.
__global__ __launch_bounds__(64) void row_min_kernel(const float* __restrict__ dist, long long B, long long LM,
                                                     float* __restrict__ out) {
  const long long b = blockIdx.x;
  if (b >= B) return;
  float v = kNoSphere;
  for (long long i = threadIdx.x; i < LM; i += kWave) v = fminf(v, dist[b * LM + i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, kWave));
  if (threadIdx.x == 0) out[b] = v;
}

// Grid for a flat grid-stride launch of `total` elements: enough 256-thread
// workgroups to fill the chip a few times over, never more than needed.
static unsigned flat_blocks(long long total) {
  const long long need = (total + 255) / 256;
  return static_cast<unsigned>(need < 256LL * 32 ? need : 256LL * 32);
}

static int world_prims(WorldPrims& w, const float* ws, int ms, const float* wc, int mc, const float* wb,
                       int mb, const float* wh, int mh) {
  PK_REQUIRE(ms >= 0 && mc >= 0 && mb >= 0 && mh >= 0, "negative obstacle count");
  PK_REQUIRE((ms == 0 || ws) && (mc == 0 || wc) && (mb == 0 || wb) && (mh == 0 || wh),
             "null world array");
  PK_REQUIRE(static_cast<long long>(ms) + mc + mb + mh <= 0x7fffffffLL, "too many obstacles");
  w.data[0] = ws; w.data[1] = wc; w.data[2] = wb; w.data[3] = wh;
  w.count[0] = ms; w.count[1] = mc; w.count[2] = mb; w.count[3] = mh;
  return PK_OK;
}

int launch_world_reduced_mixed(void* stream, const float* sc, const float* sr, long long batch, int k,
                               int n_links, const float* ws, int ms, const float* wc, int mc, const float* wb,
                               int mb, const float* wh, int mh, float* out) {
  WorldPrims w{};
  const int st = world_prims(w, ws, ms, wc, mc, wb, mb, wh, mh);
  if (st != PK_OK) return st;
  const long long total = batch * n_links * static_cast<long long>(w.total());
  if (total == 0) return PK_OK;
  hipLaunchKernelGGL(world_reduced_mixed_kernel, dim3(flat_blocks(total)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), sc, sr, batch, k, k / n_links, n_links, w, out);
  return check_launch("world_reduced_mixed_kernel");
}

}  // namespace pk

extern "C" int pk_collision_world_sphere_f32(void* stream, const float* sphere_centers,
                                             const float* sphere_radii, int64_t batch, int32_t k,
                                             const float* world_spheres, int32_t n_world_spheres,
                                             const float* world_capsules, int32_t n_world_capsules,
                                             const float* world_boxes, int32_t n_world_boxes,
                                             const float* world_halfspaces,
                                             int32_t n_world_halfspaces, float* out) {
  PK_REQUIRE(batch >= 0 && k >= 0 && n_world_spheres >= 0 && n_world_capsules >= 0 && n_world_boxes >= 0 &&
                 n_world_halfspaces >= 0,
             "negative size");
  const long long BK = batch * static_cast<long long>(k);
  if (BK == 0 || static_cast<long long>(n_world_spheres) + n_world_capsules + n_world_boxes +
                         n_world_halfspaces == 0)
    return PK_OK;
  pk::WorldPrims w{};
  const int st = pk::world_prims(w, world_spheres, n_world_spheres, world_capsules, n_world_capsules,
                                 world_boxes, n_world_boxes, world_halfspaces, n_world_halfspaces);
  if (st != PK_OK) return st;
  const long long total = BK * w.total();
  PK_REQUIRE(sphere_centers && sphere_radii && out, "null pointer");
  hipLaunchKernelGGL(pk::world_sphere_full_kernel, dim3(pk::flat_blocks(total)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), sphere_centers, sphere_radii, BK, w, out);
  return pk::check_launch("world_sphere_full_kernel");
}

extern "C" int pk_sphere_distances_jax_f32(void* stream, const float* link_poses, int64_t batch,
                                           int32_t n_links, const float* spheres_padded,
                                           int32_t s_per_link, const float* world_spheres,
                                           int32_t n_world_spheres, const int32_t* pair_i,
                                           const int32_t* pair_j, int32_t n_pairs, float* out_world,
                                           float* out_self) {
  PK_REQUIRE(batch >= 0 && n_links >= 0 && s_per_link >= 0 && n_world_spheres >= 0 && n_pairs >= 0,
             "negative size");
  if (batch == 0 || n_links == 0) return PK_OK;
  PK_REQUIRE(link_poses && (s_per_link == 0 || spheres_padded), "null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (out_world && n_world_spheres > 0) {
    PK_REQUIRE(world_spheres != nullptr, "null world");
    const long long total = batch * n_links * static_cast<long long>(n_world_spheres);
    hipLaunchKernelGGL(pk::jax_world_kernel, dim3(pk::flat_blocks(total)), dim3(256), 0, s, link_poses,
                       static_cast<long long>(batch), n_links, spheres_padded, s_per_link, world_spheres,
                       n_world_spheres, out_world);
    const int st = pk::check_launch("jax_world_kernel");
    if (st != PK_OK) return st;
  }
  if (out_self && n_pairs > 0) {
    PK_REQUIRE(pair_i && pair_j, "null pair list");
    const long long total = batch * static_cast<long long>(n_pairs);
    hipLaunchKernelGGL(pk::jax_self_kernel, dim3(pk::flat_blocks(total)), dim3(256), 0, s, link_poses,
                       static_cast<long long>(batch), n_links, spheres_padded, s_per_link, pair_i, pair_j,
                       n_pairs, out_self);
    return pk::check_launch("jax_self_kernel");
  }
  return PK_OK;
}

extern "C" int pk_collision_world_capsule_f32(void* stream, const float* caps, int64_t batch, int32_t n_links,
                                              const float* world_spheres, int32_t n_world_spheres,
                                              const float* world_capsules, int32_t n_world_capsules,
                                              const float* world_boxes, int32_t n_world_boxes,
                                              const float* world_halfspaces, int32_t n_world_halfspaces,
                                              float* out) {
  PK_REQUIRE(batch >= 0 && n_links >= 0 && n_world_spheres >= 0 && n_world_capsules >= 0 && n_world_boxes >= 0 &&
                 n_world_halfspaces >= 0,
             "negative size");
  const long long BN = batch * static_cast<long long>(n_links);
  if (BN == 0 || static_cast<long long>(n_world_spheres) + n_world_capsules + n_world_boxes +
                         n_world_halfspaces == 0)
    return PK_OK;
  pk::WorldPrims w{};
  const int st = pk::world_prims(w, world_spheres, n_world_spheres, world_capsules, n_world_capsules,
                                 world_boxes, n_world_boxes, world_halfspaces, n_world_halfspaces);
  if (st != PK_OK) return st;
  PK_REQUIRE(caps && out, "null pointer");
  const long long total = BN * w.total();
  hipLaunchKernelGGL(pk::world_capsule_kernel, dim3(pk::flat_blocks(total)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), caps, BN, w, out);
  return pk::check_launch("world_capsule_kernel");
}

extern "C" int pk_collision_self_capsule_f32(void* stream, const float* caps, int64_t batch, int32_t n_links,
                                             const int32_t* pair_i, const int32_t* pair_j, int32_t n_pairs,
                                             float* out) {
  PK_REQUIRE(batch >= 0 && n_links >= 0 && n_pairs >= 0, "negative size");
  if (batch == 0 || n_pairs == 0) return PK_OK;
  PK_REQUIRE(caps && pair_i && pair_j && out, "null pointer");
  const long long total = batch * static_cast<long long>(n_pairs);
  hipLaunchKernelGGL(pk::self_capsule_kernel, dim3(pk::flat_blocks(total)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), caps, static_cast<long long>(batch), n_links, pair_i,
                     pair_j, n_pairs, out);
  return pk::check_launch("self_capsule_kernel");
}

extern "C" int pk_sphere_world_links_f32(void* stream, const float* link_poses, int64_t batch, int32_t n_links,
                                         const float* spheres_padded, int32_t s_per_link,
                                         const float* world_spheres, int32_t n_world_spheres,
                                         const float* world_capsules, int32_t n_world_capsules,
                                         const float* world_boxes, int32_t n_world_boxes,
                                         const float* world_halfspaces, int32_t n_world_halfspaces,
                                         float* out_world, float* out_world_min) {
  PK_REQUIRE(batch >= 0 && n_links >= 0 && s_per_link >= 0 && n_world_spheres >= 0 && n_world_capsules >= 0 &&
                 n_world_boxes >= 0 && n_world_halfspaces >= 0,
             "negative size");
  const long long M = static_cast<long long>(n_world_spheres) + n_world_capsules + n_world_boxes +
                      n_world_halfspaces;
  if (batch == 0) return PK_OK;
  PK_REQUIRE(out_world != nullptr, "out_world is required (the minimum is reduced from it)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (n_links > 0 && M > 0) {
    PK_REQUIRE(link_poses && (s_per_link == 0 || spheres_padded), "null pointer");
    pk::WorldPrims w{};
    const int ws = pk::world_prims(w, world_spheres, n_world_spheres, world_capsules, n_world_capsules,
                                   world_boxes, n_world_boxes, world_halfspaces, n_world_halfspaces);
    if (ws != PK_OK) return ws;
    const long long total = batch * n_links * M;
    hipLaunchKernelGGL(pk::world_links_mixed_kernel, dim3(pk::flat_blocks(total)), dim3(256), 0, st, link_poses,
                       static_cast<long long>(batch), n_links, spheres_padded, s_per_link, w, out_world);
    const int e = pk::check_launch("world_links_mixed_kernel");
    if (e != PK_OK) return e;
  }
  if (out_world_min) {
    PK_REQUIRE(batch <= 0x7fffffffLL, "batch too large");
    hipLaunchKernelGGL(pk::row_min_kernel, dim3(static_cast<unsigned>(batch)), dim3(pk::kWave), 0, st, out_world,
                       static_cast<long long>(batch), static_cast<long long>(n_links) * M, out_world_min);
    return pk::check_launch("row_min_kernel");
  }
  return PK_OK;
}

extern "C" int pk_capsules_place_f32(void* stream, const float* link_poses, int64_t batch, int32_t n_links,
                                     const float* capsules_local, float* out_soa, float* out_aos) {
  PK_REQUIRE(batch >= 0 && n_links >= 0, "negative size");
  if (batch == 0 || n_links == 0) return PK_OK;
  PK_REQUIRE(link_poses && capsules_local, "null pointer");
  PK_REQUIRE(out_soa || out_aos, "no output requested");
  const long long total = batch * n_links;
  hipLaunchKernelGGL(pk::capsule_place_kernel, dim3(pk::flat_blocks(total)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), link_poses, static_cast<long long>(batch), n_links,
                     capsules_local, out_soa, out_aos);
  return pk::check_launch("capsule_place_kernel");
}

extern "C" int pk_row_min_f32(void* stream, const float* dist, int64_t batch, int64_t row_len, float* out) {
  PK_REQUIRE(batch >= 0 && row_len >= 0, "negative size");
  if (batch == 0) return PK_OK;
  PK_REQUIRE(out && (row_len == 0 || dist), "null pointer");
  PK_REQUIRE(batch <= 0x7fffffffLL, "batch too large");
  hipLaunchKernelGGL(pk::row_min_kernel, dim3(static_cast<unsigned>(batch)), dim3(pk::kWave), 0,
                     static_cast<hipStream_t>(stream), dist, static_cast<long long>(batch),
                     static_cast<long long>(row_len), out);
  return pk::check_launch("row_min_kernel");
}
