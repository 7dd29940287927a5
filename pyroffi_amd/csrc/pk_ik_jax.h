// JAX-convention IK for gfx950: the reference's pure-JAX solvers, whose
// residual is the SE(3) log of the body-frame error,
//   f(q) = log(T_ee(q)^-1 T_tgt)         (_ik_primitives.py:93-103, jaxlie SE3.log)
// with the exact Jacobian that reverse-mode AD of it returns.
//
//   * jlm_seed: one seed of _ls_ik_single (_ls_ik.py:192-350) -- fixed weights
//     W = (pw, pw, pw, ow, ow, ow) -- or of _lm_refine_single (_hjcd_ik.py:140-365)
//     -- per-EE adaptive weights, a soft joint-limit prior and random kicks
//     after six non-improving steps.  Jacobi column scaling, damped normal
//     equations (Cholesky), the norm trust region, the five-step line search
//     that keeps the lowest error, the 1 - 1e-4 acceptance factor and the
//     lambda schedule, all-time best tracking.
//   * hjcd_coarse_seed: _coarse_search_single (_hjcd_ik.py:55-137), greedy
//     Hamiltonian coordinate descent on one end effector.
//
// Layout: one lane per seed, the robot folded into moving "nodes" in LDS by
// the prologue shared with the CUDA-convention solver (pk_ik.h); node poses,
// residuals, the 6 NEE x MAXA Jacobian and the packed normal matrix live in
// registers.
//
// Jacobian: a moving node n with world pose (R_n, t_n) and local twist (v, w)
// per unit joint value has the spatial twist w_s = R_n w, v_s = R_n v + t_n x w_s,
// so T_ee moves by the body twist xi_b = Ad(T_ee^-1) xi_s and T_rel = T_ee^-1 T_tgt
// by the left perturbation -xi_b; the column is -mul * D log(T_rel)[xi_b],
// the directional derivative of the log formula (jlog_dir).  FP32 evaluation:
// the two small-angle factors of the log, F (rotation) and c (V^-1), use
// their series where the closed forms cancel (|q_v|^2 / w^2 < 0.04,
// theta^2 < 0.04), so the values and derivatives stay accurate to a few ulp of
// the float64 reference.
#pragma once

#include "pk_ik.h"

namespace pk {

struct JLog {
  float qw, qx, qy, qz;  // rotation of T_rel (as composed)
  float t[3];            // translation of T_rel
  float om[3];           // rotation log: om = q_v F
  float F, Fw, Fn2;      // dF = Fw dq_w + Fn2 d|q_v|^2
  float c, dc;           // V^-1 = I - Om/2 + c Om^2, dc per unit theta^2
};

__device__ __forceinline__ void cross3f(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// log(T_act^-1 T_tgt) -> f = [rho; om] and the factors its derivative needs.
__device__ __forceinline__ void jlog_eval(const Pose& Ta, const float* tgt, JLog& L, float* f) {
  quat_mul(Ta.w, -Ta.x, -Ta.y, -Ta.z, tgt[0], tgt[1], tgt[2], tgt[3], L.qw, L.qx, L.qy, L.qz);
  quat_rotate(Ta.w, -Ta.x, -Ta.y, -Ta.z, tgt[4] - Ta.tx, tgt[5] - Ta.ty, tgt[6] - Ta.tz, L.t[0], L.t[1], L.t[2]);
  const float w = L.qw;
  const float n2 = L.qx * L.qx + L.qy * L.qy + L.qz * L.qz;
  const float w2 = w * w;
  if (n2 < 0.04f * w2) {
    // F = (2 / w) g(x), x = n2 / w^2, g = atan(sqrt x) / sqrt x (odd series);
    // the same expression covers w < 0 (the shorter rotation).
    const float x = n2 / w2;
    const float g = 1.f + x * (-1.f / 3.f + x * (1.f / 5.f + x * (-1.f / 7.f + x * (1.f / 9.f))));
    const float gp = -1.f / 3.f + x * (2.f / 5.f + x * (-3.f / 7.f + x * (4.f / 9.f)));
    const float iw = 1.f / w;
    L.F = 2.f * iw * g;
    L.Fn2 = 2.f * iw * gp / w2;
    L.Fw = -2.f * iw * iw * (g + 2.f * x * gp);
  } else {
    const float n = sqrtf(n2);
    const float a = atan2f(w < 0.f ? -n : n, fabsf(w));
    const float r2 = n2 + w2;
    L.F = 2.f * a / n;
    L.Fw = -2.f / r2;
    L.Fn2 = w / (n2 * r2) - a / (n2 * n);
  }
  L.om[0] = L.qx * L.F;
  L.om[1] = L.qy * L.F;
  L.om[2] = L.qz * L.F;
  const float th2 = L.om[0] * L.om[0] + L.om[1] * L.om[1] + L.om[2] * L.om[2];
  if (th2 < 0.04f) {
    L.c = 1.f / 12.f + th2 * (1.f / 720.f + th2 * (1.f / 30240.f + th2 * (1.f / 1209600.f)));
    L.dc = 1.f / 720.f + th2 * (1.f / 15120.f + th2 * (1.f / 403200.f));
  } else {
    const float th = sqrtf(th2), h = 0.5f * th;
    float s, co;
    sincosf(h, &s, &co);
    const float cot = co / s;
    const float g = 1.f - h * cot;
    const float gp = -0.5f * cot + 0.5f * h / (s * s);
    L.c = g / th2;
    L.dc = (gp / th2 - 2.f * g / (th2 * th)) / (2.f * th);
  }
  float ot[3], oot[3];
  cross3f(L.om, L.t, ot);
  cross3f(L.om, ot, oot);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    f[i] = L.t[i] - 0.5f * ot[i] + L.c * oot[i];
    f[3 + i] = L.om[i];
  }
}

// Directional derivative of log(T_rel) under T_rel -> exp(d) T_rel, d = (dv, dw).
__device__ __forceinline__ void jlog_dir(const JLog& L, const float* dv, const float* dw, float* out) {
  const float qv[3] = {L.qx, L.qy, L.qz};
  const float dqw = -0.5f * (dw[0] * qv[0] + dw[1] * qv[1] + dw[2] * qv[2]);
  float wxq[3], dqv[3];
  cross3f(dw, qv, wxq);
#pragma unroll
  for (int i = 0; i < 3; ++i) dqv[i] = 0.5f * (L.qw * dw[i] + wxq[i]);
  const float dn2 = 2.f * (qv[0] * dqv[0] + qv[1] * dqv[1] + qv[2] * dqv[2]);
  const float dF = L.Fw * dqw + L.Fn2 * dn2;
  float dom[3], dt[3], wxt[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) dom[i] = dqv[i] * L.F + qv[i] * dF;
  cross3f(dw, L.t, wxt);
#pragma unroll
  for (int i = 0; i < 3; ++i) dt[i] = wxt[i] + dv[i];
  const float dth2 = 2.f * (L.om[0] * dom[0] + L.om[1] * dom[1] + L.om[2] * dom[2]);
  float a1[3], a2[3], ot[3], oot[3], domot[3], odomt[3], odt[3], oodt[3];
  cross3f(dom, L.t, a1);
  cross3f(L.om, dt, a2);
  cross3f(L.om, L.t, ot);
  cross3f(L.om, ot, oot);
  cross3f(dom, ot, domot);
  cross3f(L.om, a1, odomt);
  cross3f(L.om, dt, odt);
  cross3f(L.om, odt, oodt);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    out[i] = dt[i] - 0.5f * (a1[i] + a2[i]) + L.dc * dth2 * oot[i] + L.c * (domot[i] + odomt[i] + oodt[i]);
    out[3 + i] = dom[i];
  }
}

// Residuals of every EE (f[6 e ..]; padded EEs read 0) and, with JAC, the
// Jacobian Jm[6 NEE][MAXA].
template <int MAXA, int NEE, bool JAC>
__device__ __forceinline__ void jax_res_jac(const IkShared<MAXA, NEE>& sh, const Pose* P,
                                            const float (&tg)[NEE][7], float* f, float (*Jm)[MAXA]) {
#pragma unroll
  for (int e = 0; e < NEE; ++e) {
    const int en = __builtin_amdgcn_readfirstlane(sh.ee_node[e]);
    if (en == -2) {
#pragma unroll
      for (int k = 0; k < 6; ++k) f[6 * e + k] = 0.f;
      if (JAC) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
          for (int a = 0; a < MAXA; ++a) Jm[6 * e + k][a] = 0.f;
      }
      continue;
    }
    const Pose Te = ee_pose(sh, P, e);
    JLog L;
    jlog_eval(Te, tg[e], L, f + 6 * e);
    if (!JAC) continue;
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int a = 0; a < MAXA; ++a) Jm[6 * e + k][a] = 0.f;
#pragma unroll
    for (int n = 0; n < MAXA; ++n) {
      if (n < sh.n_nodes) {
        const NodeRec& nr = sh.nodes[n];
        const unsigned em = __builtin_amdgcn_readfirstlane(nr.ee_mask);
        const int src = __builtin_amdgcn_readfirstlane(nr.src);
        const int kind = __builtin_amdgcn_readfirstlane(nr.kind);
        if (!(em & (1u << e)) || src < 0) continue;
        float wl[3] = {0.f, 0.f, 0.f}, vl[3] = {0.f, 0.f, 0.f};
        if (kind == 4) {
          wl[0] = nr.ang[0] * nr.lin[0]; wl[1] = nr.ang[1] * nr.lin[0]; wl[2] = nr.ang[2] * nr.lin[0];
        } else {
          vl[0] = nr.lin[0]; vl[1] = nr.lin[1]; vl[2] = nr.lin[2];
          if (kind == 3) { wl[0] = nr.ang[0]; wl[1] = nr.ang[1]; wl[2] = nr.ang[2]; }
        }
        float ws[3], vs[3], arm[3], axw[3], vb[3], wb[3], col[6];
        quat_rotate(P[n].w, P[n].x, P[n].y, P[n].z, wl[0], wl[1], wl[2], ws[0], ws[1], ws[2]);
        quat_rotate(P[n].w, P[n].x, P[n].y, P[n].z, vl[0], vl[1], vl[2], vs[0], vs[1], vs[2]);
        arm[0] = P[n].tx - Te.tx; arm[1] = P[n].ty - Te.ty; arm[2] = P[n].tz - Te.tz;
        cross3f(arm, ws, axw);
        vs[0] += axw[0]; vs[1] += axw[1]; vs[2] += axw[2];
        quat_rotate(Te.w, -Te.x, -Te.y, -Te.z, vs[0], vs[1], vs[2], vb[0], vb[1], vb[2]);
        quat_rotate(Te.w, -Te.x, -Te.y, -Te.z, ws[0], ws[1], ws[2], wb[0], wb[1], wb[2]);
        jlog_dir(L, vb, wb, col);
        const float mul = nr.mul;
#pragma unroll
        for (int a = 0; a < MAXA; ++a)
          if (src == a) {
#pragma unroll
            for (int k = 0; k < 6; ++k) Jm[6 * e + k][a] -= mul * col[k];
          }
      }
    }
  }
}

// _adaptive_weights (_ik_primitives.py:106-116): the orientation scale of one EE.
__device__ __forceinline__ float adaptive_scale(const float* f) {
  const float pe = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) + 1e-8f;
  const float oe = sqrtf(f[3] * f[3] + f[4] * f[4] + f[5] * f[5]) + 1e-8f;
  return fminf(fmaxf(pe / oe, 0.05f), 1.f);
}

// Counter-based standard normal for the stall kicks: splitmix64 (Steele, Lea
// and Flood) of (key, seed id, iteration, joint), Box-Muller of two 24-bit
// uniforms.  The FP64 oracle draws the same bits.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float kick_normal(unsigned long long key, unsigned long long gid, int it, int j) {
  const unsigned long long h =
      mix64(key ^ mix64((gid << 20) ^ (static_cast<unsigned long long>(it) << 6) ^ static_cast<unsigned long long>(j)));
  const float u1 = static_cast<float>((h >> 40) + 1) * (1.f / 16777216.f);
  const float u2 = static_cast<float>((h >> 16) & 0xFFFFFFull) * (1.f / 16777216.f);
  return sqrtf(-2.f * logf(u1)) * cosf(6.283185307179586f * u2);
}

struct JlmOpts {
  int adaptive;        // HJCD: per-EE adaptive weights; else W = (pw, pw, pw, ow, ow, ow)
  float prior_weight;  // soft joint-limit prior, 0 = off
  float kick_scale;    // stall kicks, 0 = off
  unsigned long long kick_key;
};

template <int NEE>
__device__ __forceinline__ void jlm_weights(const IkArgs& a, const JlmOpts& o, const float* f, float* w) {
#pragma unroll
  for (int e = 0; e < NEE; ++e) {
    const float s = o.adaptive ? adaptive_scale(f + 6 * e) : a.ori_weight;
    const float p = o.adaptive ? 1.f : a.pos_weight;
#pragma unroll
    for (int k = 0; k < 6; ++k) w[6 * e + k] = (k < 3) ? p : s;
  }
}

template <int MAXA, int NEE>
__device__ __forceinline__ float jlm_error(const IkShared<MAXA, NEE>& sh, const float (&tg)[NEE][7], const float* q,
                                           const float* w) {
  Pose P[MAXA];
  float f[6 * NEE];
  node_fk<MAXA, NEE, false>(sh, q, P);
  jax_res_jac<MAXA, NEE, false>(sh, P, tg, f, nullptr);
  float e = 0.f;
#pragma unroll
  for (int i = 0; i < 6 * NEE; ++i) { const float x = w[i] * f[i]; e += x * x; }
  return e;
}

// Solve (A) x = rhs in place for the packed lower triangle A (MAXA x MAXA);
// a non-positive pivot zeroes the solution.
template <int MAXA>
__device__ __forceinline__ void packed_cholesky_solve(float* A, float* rhs) {
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
#pragma unroll
  for (int i = MAXA - 1; i >= 0; --i) {
    float t = y[i];
#pragma unroll
    for (int pp = i + 1; pp < MAXA; ++pp) t -= A[pp * (pp + 1) / 2 + i] * rhs[pp];
    rhs[i] = ok ? t * idg[i] : 0.f;
  }
}

// One seed, max_iter iterations; returns the all-time best configuration in
// best_q and the winner metric at it (sum over EEs of |W f|^2, or |f|^2 with
// adaptive weights).
template <int MAXA, int NEE>
__device__ __forceinline__ float jlm_seed(const IkShared<MAXA, NEE>& sh, const IkArgs& a, const JlmOpts& o,
                                          const float (&tg)[NEE][7], const float* seed, unsigned long long gid,
                                          float* best_q) {
  float c[MAXA], mid[MAXA], dpr[MAXA];
#pragma unroll
  for (int i = 0; i < MAXA; ++i) {
    c[i] = (i < a.n_act) ? seed[i] : 0.f;
    best_q[i] = c[i];
    mid[i] = 0.5f * (sh.lo[i] + sh.hi[i]);
    const float half = 0.5f * (sh.hi[i] - sh.lo[i]) + 1e-8f;
    dpr[i] = (i < a.n_act) ? o.prior_weight / (half * half) : 0.f;
  }
  float w[6 * NEE], f[6 * NEE];
  {
    Pose P[MAXA];
    node_fk<MAXA, NEE, false>(sh, c, P);
    jax_res_jac<MAXA, NEE, false>(sh, P, tg, f, nullptr);
  }
  jlm_weights<NEE>(a, o, f, w);
  float best_err = 0.f;
#pragma unroll
  for (int i = 0; i < 6 * NEE; ++i) { const float x = w[i] * f[i]; best_err += x * x; }
  float lam = a.lambda_init;
  int stall = 0;
  for (int it = 0; it < a.max_iter; ++it) {
    float Jm[6 * NEE][MAXA];
    {
      Pose P[MAXA];
      node_fk<MAXA, NEE, false>(sh, c, P);
      jax_res_jac<MAXA, NEE, true>(sh, P, tg, f, Jm);
    }
    jlm_weights<NEE>(a, o, f, w);
    float fw[6 * NEE], cur = 0.f;
#pragma unroll
    for (int i = 0; i < 6 * NEE; ++i) { fw[i] = w[i] * f[i]; cur += fw[i] * fw[i]; }
    float ics[MAXA];
#pragma unroll
    for (int k = 0; k < MAXA; ++k) {
      float sq = 0.f;
#pragma unroll
      for (int i = 0; i < 6 * NEE; ++i) { Jm[i][k] *= w[i]; sq += Jm[i][k] * Jm[i][k]; }
      ics[k] = 1.f / (sqrtf(sq) + 1e-8f);
    }
#pragma unroll
    for (int i = 0; i < 6 * NEE; ++i)
#pragma unroll
      for (int k = 0; k < MAXA; ++k) Jm[i][k] *= ics[k];
    constexpr int NT = MAXA * (MAXA + 1) / 2;
    float A[NT], rhs[MAXA];
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
      rhs[i] = -(rb + dpr[i] * (c[i] - mid[i]) * ics[i]);
      // Padded columns (i >= n_act) are zero: their diagonal keeps the solve
      // definite.  The damping inside the FP32 solve is floored at 1e-6 (the
      // lambda schedule itself is the reference's), as in pk_ik.h.
      A[i * (i + 1) / 2 + i] += fmaxf(lam, kDampFloor) + dpr[i] * ics[i] * ics[i] + ((i < a.n_act) ? 0.f : 1.f);
    }
    packed_cholesky_solve<MAXA>(A, rhs);
    float delta[MAXA], dn = 0.f;
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
      delta[i] = (i < a.n_act && !sh.fixed[i]) ? rhs[i] * ics[i] : 0.f;
      dn += delta[i] * delta[i];
    }
    float max_p = 0.f, max_o = 0.f;
#pragma unroll
    for (int e = 0; e < NEE; ++e) {
      if (e < a.n_ee) {
        max_p = fmaxf(max_p, sqrtf(f[6 * e] * f[6 * e] + f[6 * e + 1] * f[6 * e + 1] + f[6 * e + 2] * f[6 * e + 2]));
        max_o = fmaxf(max_o, sqrtf(f[6 * e + 3] * f[6 * e + 3] + f[6 * e + 4] * f[6 * e + 4] +
                                   f[6 * e + 5] * f[6 * e + 5]));
      }
    }
    float R;
    if (max_p > 1e-2f || max_o > 0.6f) R = 0.38f;
    else if (max_p > 1e-3f || max_o > 0.25f) R = 0.22f;
    else if (max_p > 2e-4f || max_o > 0.08f) R = 0.12f;
    else R = 0.05f;
    dn = sqrtf(dn) + 1e-18f;
    if (dn > R) {
      const float sc = R / dn;
#pragma unroll
      for (int i = 0; i < MAXA; ++i) delta[i] *= sc;
    }
    // Line search: all five steps, the lowest error (first on ties).
    float best_a = INFINITY;
    int pick = 0;
#pragma unroll 1
    for (int k = 0; k < 5; ++k) {
      float qt[MAXA];
#pragma unroll
      for (int i = 0; i < MAXA; ++i) qt[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], c[i] + ls_alpha(k) * delta[i]));
      const float e = jlm_error(sh, tg, qt, w);
      if (e < best_a) { best_a = e; pick = k; }
    }
    float nc[MAXA];
#pragma unroll
    for (int i = 0; i < MAXA; ++i) nc[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], c[i] + ls_alpha(pick) * delta[i]));
    const bool improved = best_a < cur * (1.f - 1e-4f);
    lam = fminf(fmaxf(improved ? lam * 0.5f : lam * 3.f, 1e-10f), 1e6f);
    if (best_a < best_err) {
      best_err = best_a;
#pragma unroll
      for (int i = 0; i < MAXA; ++i) best_q[i] = nc[i];
    }
    if (improved) {
#pragma unroll
      for (int i = 0; i < MAXA; ++i) c[i] = nc[i];
    }
    if (o.kick_scale > 0.f) {
      stall = improved ? 0 : stall + 1;
      if (stall >= 6) {
#pragma unroll
        for (int i = 0; i < MAXA; ++i) {
          if (i < a.n_act) {
            const float kv = sh.fixed[i] ? 0.f : kick_normal(o.kick_key, gid, it, i) * o.kick_scale;
            c[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], c[i] + kv));
          }
        }
        lam = a.lambda_init;
        stall = 0;
      }
    }
  }
  // Winner metric at the best configuration.
  {
    Pose P[MAXA];
    node_fk<MAXA, NEE, false>(sh, best_q, P);
    jax_res_jac<MAXA, NEE, false>(sh, P, tg, f, nullptr);
  }
  float e = 0.f;
#pragma unroll
  for (int i = 0; i < 6 * NEE; ++i) {
    const float wi = o.adaptive ? 1.f : ((i % 6) < 3 ? a.pos_weight : a.ori_weight);
    const float x = wi * f[i];
    e += x * x;
  }
  return e;
}

// _coarse_search_single on EE `ee` of one seed; returns the sum over every EE
// of |f|^2 at the result.
template <int MAXA, int NEE>
__device__ __forceinline__ float hjcd_coarse_seed(const IkShared<MAXA, NEE>& sh, const IkArgs& a,
                                                  const float (&tg)[NEE][7], int ee, int k_max, float eps,
                                                  float nu, float* c) {
  float mom[MAXA];
#pragma unroll
  for (int i = 0; i < MAXA; ++i) mom[i] = 0.f;
  float f[6 * NEE];
  for (int k = 0; k < k_max; ++k) {
    float Jm[6 * NEE][MAXA];
    {
      Pose P[MAXA];
      node_fk<MAXA, NEE, false>(sh, c, P);
      jax_res_jac<MAXA, NEE, true>(sh, P, tg, f, Jm);
    }
    // Rows of this seed's EE.
    float fe[6], Je[6][MAXA];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      fe[r] = 0.f;
#pragma unroll
      for (int i = 0; i < MAXA; ++i) Je[r][i] = 0.f;
    }
#pragma unroll
    for (int e = 0; e < NEE; ++e)
      if (e == ee) {
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          fe[r] = f[6 * e + r];
#pragma unroll
          for (int i = 0; i < MAXA; ++i) Je[r][i] = Jm[6 * e + r][i];
        }
      }
    const float pe = sqrtf(fe[0] * fe[0] + fe[1] * fe[1] + fe[2] * fe[2]);
    const float oe = sqrtf(fe[3] * fe[3] + fe[4] * fe[4] + fe[5] * fe[5]);
    if (pe < eps && oe < nu) continue;
    const float so = (pe < 1e-3f) ? adaptive_scale(fe) : 0.f;
    const float w[6] = {1.f, 1.f, 1.f, so, so, so};
    float bs = -INFINITY, gb = 0.f, nb = 1.f;
    int bj = 0;
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
      if (i < a.n_act) {
        float g = 0.f, nsq = 1e-8f;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          const float jw = w[r] * Je[r][i];
          g += jw * (w[r] * fe[r]);
          nsq += jw * jw;
        }
        const float score = sh.fixed[i] ? -1.f : fabsf(g) / sqrtf(nsq);
        if (score > bs) { bs = score; bj = i; gb = g; nb = nsq; }
      }
    }
#pragma unroll
    for (int i = 0; i < MAXA; ++i) {
      if (i == bj) {
        mom[i] = 0.9f * mom[i] - 0.35f * gb / nb;
        if (!sh.fixed[i]) c[i] = fmaxf(sh.lo[i], fminf(sh.hi[i], c[i] + mom[i]));
      }
    }
  }
  {
    Pose P[MAXA];
    node_fk<MAXA, NEE, false>(sh, c, P);
    jax_res_jac<MAXA, NEE, false>(sh, P, tg, f, nullptr);
  }
  float e = 0.f;
#pragma unroll
  for (int i = 0; i < 6 * NEE; ++i) e += f[i] * f[i];
  return e;
}

struct JaxIkArgs {
  IkArgs ik;               // tables, targets, limits, seeds (P, S, A), outputs, max_iter, weights
  JlmOpts opts;
  int coarse;              // 1: hjcd_coarse_seed instead of jlm_seed
  const int* ee_of_seed;   // (S): the EE each seed's coarse search serves
  int k_max;
  float eps, nu;
};

// One lane per seed (p = blockIdx.y, s = blockIdx.x * 64 + lane).
template <int MAXA, int NEE>
__global__ __launch_bounds__(64) void jax_ik_kernel(JaxIkArgs ja) {
  extern __shared__ float4 smem4[];
  const IkArgs& a = ja.ik;
  const IkView<MAXA, NEE> v = ik_prologue<MAXA, NEE>(a, reinterpret_cast<char*>(smem4));
  const int p = blockIdx.y;
  const int s = blockIdx.x * kWave + threadIdx.x;
  if (s >= a.n_seeds) return;
  const long long gs = static_cast<long long>(p) * a.n_seeds + s;
  const IkShared<MAXA, NEE>& sh = *v.sh;
  if (sh.too_many_nodes) {  // reported as NaN errors, seeds unchanged
    for (int i = 0; i < a.n_act; ++i) a.out_cfgs[gs * a.n_act + i] = a.seeds[gs * a.n_act + i];
    a.out_err[gs] = NAN;
    return;
  }
  float tg[NEE][7];
  load_targets<MAXA, NEE>(a, p, tg);
  float q[MAXA];
  float err;
  if (ja.coarse) {
#pragma unroll
    for (int i = 0; i < MAXA; ++i) q[i] = (i < a.n_act) ? a.seeds[gs * a.n_act + i] : 0.f;
    const int ee = ja.ee_of_seed ? ja.ee_of_seed[s] : 0;
    err = hjcd_coarse_seed<MAXA, NEE>(sh, a, tg, ee, ja.k_max, ja.eps, ja.nu, q);
  } else {
    err = jlm_seed<MAXA, NEE>(sh, a, ja.opts, tg, a.seeds + gs * a.n_act, static_cast<unsigned long long>(gs), q);
  }
#pragma unroll
  for (int i = 0; i < MAXA; ++i)
    if (i < a.n_act) a.out_cfgs[gs * a.n_act + i] = q[i];
  a.out_err[gs] = err;
}

template <int MAXA, int NEE>
inline int launch_jax_ik(void* stream, const JaxIkArgs& ja) {
  using Sh = IkShared<MAXA, NEE>;
  const size_t lds = ik_lds_layout(sizeof(Sh), ja.ik.tables.n_joints, 0, 0).total;
  PK_REQUIRE(lds <= 150 * 1024, "IK tables need %zu bytes of LDS (limit 150 KiB)", lds);
  dim3 grid((ja.ik.n_seeds + kWave - 1) / kWave, ja.ik.n_problems);
  hipLaunchKernelGGL((jax_ik_kernel<MAXA, NEE>), grid, dim3(kWave), lds, static_cast<hipStream_t>(stream), ja);
  return check_launch("jax_ik_kernel");
}

#define PK_JAX_IK_DECLARE(m, e) int launch_jax_ik_##m##_##e(void* stream, const JaxIkArgs& a);
PK_IK_LAUNCHERS(PK_JAX_IK_DECLARE)

}  // namespace pk
