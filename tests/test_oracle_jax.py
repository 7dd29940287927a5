"""Pins of the FP64 oracle's JAX-convention IK (CPU).

* residual  -- log(T_ee^-1 T_tgt) against an independent SE(3) log: the
  rotation vector from scipy and rho = V(w)^-1 t with the SE(3) V matrix
  solved by numpy, on link transforms built from the URDF XML
  (``independent``).
* Jacobian  -- central finite differences of the residual (h = 1e-6),
  including a mimic/prismatic robot and two end effectors.
* LM loops  -- ``_ls_ik_single`` and ``_lm_refine_single`` (adaptive weights,
  joint-limit prior, stall kicks) and the HJCD coarse phase, restated here in
  numpy from the reference source on top of the pinned residual and Jacobian,
  agree with the C loops step for step.
* outcomes  -- the LM solves FK-generated (hence reachable) targets.
"""

import independent
import numpy as np
import pytest
from conftest import load_robot, random_cfgs
from scipy.spatial.transform import Rotation

from pyroffi_amd.optimization_engines._ls_ik import ee_tables

EE = {"panda": ("panda_hand",), "fetch": ("gripper_link",), "mimic_arm": ("finger_c",),
      "baxter": ("left_gripper_base", "right_gripper_base")}
ALPHAS = (1.0, 0.5, 0.25, 0.1, 0.025)


def _skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def _problem(oracle_mod, name, n, seed):
    robot, path = load_robot(name)
    t = oracle_mod.Tables.of(robot)
    links = [robot.links.names.index(x) for x in EE[name]]
    tj, _ = ee_tables(robot, links)
    lp = oracle_mod.fk_links(t, random_cfgs(robot, n, seed=seed))
    tgt = np.stack([lp[:, l] for l in links], axis=1)
    return robot, path, t, links, tj, tgt


@pytest.mark.parametrize("name", ("panda", "fetch", "mimic_arm", "baxter"))
def test_residual_matches_independent_log(oracle_mod, name):
    robot, path, t, links, tj, tgt = _problem(oracle_mod, name, 12, 3)
    q = random_cfgs(robot, 12, seed=4)
    f = oracle_mod.jres(t, q, tj, tgt)
    doc = independent.read(path)[0]
    for b in range(len(q)):
        mats = independent.link_matrices(path, q[b])
        for e, l in enumerate(links):
            Ta = mats[doc.index(robot.links.names[l])]
            Tt = independent.pose_to_matrix(tgt[b, e])
            Tr = np.linalg.inv(Ta) @ Tt
            w = Rotation.from_matrix(Tr[:3, :3]).as_rotvec()
            th = np.linalg.norm(w)
            W = _skew(w)
            V = np.eye(3) + (1 - np.cos(th)) / th ** 2 * W + (th - np.sin(th)) / th ** 3 * W @ W
            rho = np.linalg.solve(V, Tr[:3, 3])
            # the oracle takes the float32 tables the kernels take
            np.testing.assert_allclose(f[b, e], np.r_[rho, w], atol=5e-6)


@pytest.mark.parametrize("name", ("panda", "mimic_arm", "baxter"))
def test_jacobian_matches_finite_differences(oracle_mod, name):
    robot, path, t, links, tj, tgt = _problem(oracle_mod, name, 6, 5)
    q = random_cfgs(robot, 6, seed=6)
    _, J = oracle_mod.jres(t, q, tj, tgt, jac=True)
    h = 1e-6
    A = robot.joints.num_actuated_joints
    for a in range(A):
        e = np.zeros(A)
        e[a] = h
        fd = (oracle_mod.jres(t, q + e, tj, tgt) - oracle_mod.jres(t, q - e, tj, tgt)) / (2 * h)
        # The analytic columns assume unit quaternions; the float32 table
        # quaternions are unit to ~1e-7, which the finite differences see.
        np.testing.assert_allclose(J[:, :, a], fd.reshape(len(q), -1), rtol=0, atol=3e-6)
    assert np.abs(J).max() > 0.1


def _mix64(z):
    m = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def _kick(key, gid, it, j):
    h = _mix64(key ^ _mix64((gid << 20) ^ (it << 6) ^ j))
    u1 = ((h >> 40) + 1) / 16777216.0
    u2 = ((h >> 16) & 0xFFFFFF) / 16777216.0
    return np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)


def _lm_numpy(oracle_mod, t, tj, tgt, q0, lo, hi, fixed, max_iter, pw, ow, lam0, adaptive, prior, kick, key, gid):
    """_ls_ik_single / _lm_refine_single from the reference source, numpy."""
    n_ee = len(tj)
    A = len(q0)

    def res(q, jac=False):
        out = oracle_mod.jres(t, q[None], tj, tgt[None], jac=jac)
        return (out[0][0].reshape(-1), out[1][0]) if jac else out[0].reshape(-1)

    def weights(f):
        if not adaptive:
            return np.tile(np.r_[[pw] * 3, [ow] * 3], n_ee)
        ws = []
        for e in range(n_ee):
            fe = f[6 * e:6 * e + 6]
            s = np.clip((np.linalg.norm(fe[:3]) + 1e-8) / (np.linalg.norm(fe[3:]) + 1e-8), 0.05, 1.0)
            ws.append(np.r_[1, 1, 1, s, s, s])
        return np.concatenate(ws)

    mid, half = (lo + hi) / 2, (hi - lo) / 2 + 1e-8
    dpr = prior / half ** 2
    c = q0.copy()
    f = res(c)
    best, best_err = c.copy(), np.sum((weights(f) * f) ** 2)
    lam, stall = lam0, 0
    for it in range(max_iter):
        f, J = res(c, True)
        w = weights(f)
        fw, Jw = w * f, w[:, None] * J
        cur = fw @ fw
        col = np.linalg.norm(Jw, axis=0) + 1e-8
        Js = Jw / col
        M = Js.T @ Js + np.diag(lam + dpr / col ** 2)
        p = np.linalg.solve(M, -(Js.T @ fw + dpr * (c - mid) / col))
        d = np.where(fixed, 0.0, p / col)
        pe = max(np.linalg.norm(f[6 * e:6 * e + 3]) for e in range(n_ee))
        oe = max(np.linalg.norm(f[6 * e + 3:6 * e + 6]) for e in range(n_ee))
        R = 0.38 if (pe > 1e-2 or oe > 0.6) else 0.22 if (pe > 1e-3 or oe > 0.25) else \
            0.12 if (pe > 2e-4 or oe > 0.08) else 0.05
        dn = np.linalg.norm(d) + 1e-18
        if dn > R:
            d = d * R / dn
        errs = [np.sum((w * res(np.clip(c + a * d, lo, hi))) ** 2) for a in ALPHAS]
        k = int(np.argmin(errs))
        new_c = np.clip(c + ALPHAS[k] * d, lo, hi)
        improved = errs[k] < cur * (1 - 1e-4)
        lam = np.clip(lam * 0.5 if improved else lam * 3.0, 1e-10, 1e6)
        if errs[k] < best_err:
            best, best_err = new_c, errs[k]
        c = new_c if improved else c
        if kick > 0:
            stall = 0 if improved else stall + 1
            if stall >= 6:
                kv = np.array([0.0 if fixed[a] else _kick(key, gid, it, a) * kick for a in range(A)])
                c, lam, stall = np.clip(c + kv, lo, hi), lam0, 0
    fb = res(best)
    wsel = np.ones_like(fb) if adaptive else np.tile(np.r_[[pw] * 3, [ow] * 3], n_ee)
    return best, np.sum((wsel * fb) ** 2)


@pytest.mark.parametrize("name,hjcd", [("panda", False), ("panda", True), ("baxter", True), ("mimic_arm", False)])
def test_lm_loops_match_numpy_restatement(oracle_mod, name, hjcd):
    robot, path, t, links, tj, tgt = _problem(oracle_mod, name, 2, 7)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    A = robot.joints.num_actuated_joints
    seeds = np.random.default_rng(8).uniform(lo, hi, (2, 3, A))
    fixed = np.zeros(A, np.int32)
    fixed[1] = 1
    kw = dict(max_iter=25, pos_weight=50.0, ori_weight=10.0, lambda_init=5e-3, adaptive=hjcd,
              prior_weight=1e-4 if hjcd else 0.0, kick_scale=0.05 if hjcd else 0.0, kick_key=12345)
    out, err = oracle_mod.jlm(t, seeds, tj, tgt, lo, hi, fixed, **kw)
    for p in range(2):
        for s in range(3):
            q, e = _lm_numpy(oracle_mod, t, tj, tgt[p], seeds[p, s], lo, hi, fixed.astype(bool), 25, 50.0, 10.0, 5e-3,
                             hjcd, kw["prior_weight"], kw["kick_scale"], 12345, p * 3 + s)
            np.testing.assert_allclose(out[p, s], q, atol=1e-8)
            assert err[p, s] == pytest.approx(e, rel=1e-7, abs=1e-12)
            assert out[p, s, 1] == seeds[p, s, 1]  # fixed joint untouched


def test_kick_normal_distribution():
    z = np.array([_kick(99, g, it, j) for g in range(40) for it in range(10) for j in range(7)])
    assert abs(z.mean()) < 0.08 and abs(z.std() - 1.0) < 0.06


def _coarse_numpy(oracle_mod, t, tj, tgt, q0, ee, lo, hi, fixed, k_max, eps, nu):
    c, mom = q0.copy(), np.zeros_like(q0)
    for _ in range(k_max):
        f, J = oracle_mod.jres(t, c[None], tj[ee:ee + 1], tgt[None, ee:ee + 1], jac=True)
        f, J = f[0, 0], J[0]
        if np.linalg.norm(f[:3]) < eps and np.linalg.norm(f[3:]) < nu:
            continue
        s = np.clip((np.linalg.norm(f[:3]) + 1e-8) / (np.linalg.norm(f[3:]) + 1e-8), 0.05, 1.0)
        gate = float(np.linalg.norm(f[:3]) < 1e-3)
        w = np.r_[1, 1, 1, s * gate, s * gate, s * gate]
        Jw, fw = w[:, None] * J, w * f
        g = Jw.T @ fw
        nsq = np.sum(Jw ** 2, axis=0) + 1e-8
        score = np.where(fixed, -1.0, np.abs(g) / np.sqrt(nsq))
        j = int(np.argmax(score))
        mom[j] = 0.9 * mom[j] - 0.35 * g[j] / nsq[j]
        d = np.zeros_like(c)
        d[j] = mom[j]
        c = np.clip(c + np.where(fixed, 0.0, d), lo, hi)
    f = oracle_mod.jres(t, c[None], tj, tgt[None])
    return c, np.sum(f ** 2)


@pytest.mark.parametrize("name", ("panda", "baxter"))
def test_coarse_matches_numpy_restatement(oracle_mod, name):
    robot, path, t, links, tj, tgt = _problem(oracle_mod, name, 2, 9)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    A = robot.joints.num_actuated_joints
    S = 6
    seeds = np.random.default_rng(10).uniform(lo, hi, (2, S, A))
    ee_of_seed = np.array([(s * len(tj)) // S for s in range(S)], np.int32)  # equal split over EEs
    fixed = np.zeros(A, np.int32)
    out, err = oracle_mod.hjcd_coarse(t, seeds, tj, tgt, ee_of_seed, lo, hi, fixed, k_max=20)
    for p in range(2):
        for s in range(S):
            q, e = _coarse_numpy(oracle_mod, t, tj, tgt[p], seeds[p, s], ee_of_seed[s], lo, hi, fixed.astype(bool),
                                 20, 0.02, np.pi / 2)
            np.testing.assert_allclose(out[p, s], q, atol=1e-10)
            assert err[p, s] == pytest.approx(e, rel=1e-9)
    # coarse descent closes the position gap of most seeds (orientation is
    # gated out until the position is within 1 mm)
    own = np.tile(ee_of_seed, 2)  # each seed descends on its own EE

    def pos(x):
        f = oracle_mod.jres(t, x.reshape(-1, A), tj, np.repeat(tgt, S, 0))
        return np.linalg.norm(f[np.arange(len(own)), own, :3], axis=1)

    assert np.mean(pos(out) < pos(seeds)) > 0.7


def test_lm_solves_reachable_targets(oracle_mod):
    robot, path, t, links, tj, tgt = _problem(oracle_mod, "panda", 16, 11)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    seeds = np.random.default_rng(12).uniform(lo, hi, (16, 32, 7))
    out, err = oracle_mod.jlm(t, seeds, tj, tgt, lo, hi, max_iter=60)
    best = out[np.arange(16), err.argmin(1)]
    f = oracle_mod.jres(t, best, tj, tgt)
    assert np.all(out >= lo - 1e-12) and np.all(out <= hi + 1e-12)
    ok = (np.linalg.norm(f[:, 0, :3], axis=1) < 1e-3) & (np.linalg.norm(f[:, 0, 3:], axis=1) < 1e-2)
    assert ok.mean() >= 0.9, ok
