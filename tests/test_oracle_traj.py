"""Pins of the FP64 trajectory oracle (CPU): cost terms, gradient, two-loop
recursion and the direct L-BFGS loop.

* smoothness -- closed forms: the 5-point stencil returns the exact second
  derivative of polynomials up to degree 5 (quadratic: constant acc, no jerk;
  cubic: jerk 6 per step), so a wrong coefficient, index or scale fails.
* limits     -- a trajectory outside its limits by known amounts.
* collision  -- distances from ``independent`` (URDF XML + scipy cdist, no
  oracle code), through the reference's activation functions; both modes'
  piecewise pieces meet at d = 0 and d = margin.
* gradient   -- central finite differences of the oracle cost.
* two-loop   -- the dense BFGS inverse-Hessian recursion H+ = (I - rho s y^T)
  H (I - rho y s^T) + rho s s^T from H0 = gamma I.
* L-BFGS     -- on a collision-free quadratic (smoothness only, inside the
  limits) it converges to the minimum-energy path from numpy's linear solve;
  in a colliding and in an empty world its accepted steps, last iterate and
  tracked best equal a dense-algebra restatement of the loop step for step.
"""

import independent
import numpy as np
import pytest
from conftest import load_robot, random_cfgs
from scipy.spatial import distance

from pyroffi_amd.collision import RobotCollisionSpherized

W_ALL = (1.0, 1.0, 0.5, 0.1, 1.0, 1.0)


def _model(oracle_mod, name="panda", M=40, margin=0.05, mode=1, seed=0, pairs=True):
    robot, path = load_robot(name)
    coll = RobotCollisionSpherized.from_urdf(path)
    t = oracle_mod.Tables.of(robot)
    rng = np.random.default_rng(seed)
    world = np.concatenate([rng.uniform(-0.8, 0.8, (M, 3)), rng.uniform(0.02, 0.1, (M, 1))], 1)
    pi = np.asarray(coll.active_idx_i if pairs else [], np.int32)
    pj = np.asarray(coll.active_idx_j if pairs else [], np.int32)
    md = oracle_mod.TrajModel(t, oracle_mod.padded_spheres(coll), pi, pj, world, margin, mode)
    return robot, path, coll, t, md, world


def _line(robot, T, seed):
    q0, q1 = random_cfgs(robot, 2, seed=seed)
    s = np.linspace(0, 1, T)[:, None]
    return (1 - s) * q0 + s * q1


def test_smoothness_closed_forms(oracle_mod):
    robot, _, _, t, md, _ = _model(oracle_mod, M=0, pairs=False)
    md.world = np.array([[100.0, 100.0, 100.0, 0.01]])  # far away: no collision term
    A, T = 7, 12
    lo, hi = np.full(A, -1e3), np.full(A, 1e3)
    tt = np.arange(T, dtype=np.float64)[:, None]
    c = np.linspace(0.01, 0.07, A)[None]
    quad = c * tt ** 2
    cubic = c * tt ** 3
    comps, _ = oracle_mod.traj_cost(md, [quad], lo, hi, (1, 0, 1, 0, 0, 0), grad=False)
    assert comps[0, 0] == pytest.approx((T - 4) * np.sum((2 * c) ** 2), rel=1e-12)
    comps, _ = oracle_mod.traj_cost(md, [cubic], lo, hi, (1, 0, 0, 1, 0, 0), grad=False)
    assert comps[0, 0] == pytest.approx((T - 5) * np.sum((6 * c) ** 2), rel=1e-10)
    comps, _ = oracle_mod.traj_cost(md, [cubic], lo, hi, (1, 0, 1, 0, 0, 0), grad=False)
    acc = 6 * c * (tt[: T - 4] + 2)  # second derivative at the stencil centre
    assert comps[0, 0] == pytest.approx(np.sum(acc ** 2), rel=1e-10)
    comps, _ = oracle_mod.traj_cost(md, [quad], lo, hi, (1, 1, 0, 0, 0, 0), grad=False)
    vel = c * (2 * tt[: T - 1] + 1)
    assert comps[0, 0] == pytest.approx(np.sum(vel ** 2), rel=1e-12)


def test_limits_closed_form(oracle_mod):
    robot, _, _, t, md, _ = _model(oracle_mod, M=0, pairs=False)
    md.world = np.array([[100.0, 100.0, 100.0, 0.01]])
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    x = np.tile((lo + hi) / 2, (6, 1))
    x[1, 0] = hi[0] + 0.3
    x[4, 3] = lo[3] - 0.2
    comps, _ = oracle_mod.traj_cost(md, [x], lo, hi, W_ALL, grad=False)
    assert comps[0, 1] == pytest.approx(0.3 ** 2 + 0.2 ** 2, rel=1e-9)


@pytest.mark.parametrize("mode", [0, 1])
def test_collision_term_matches_independent(oracle_mod, mode):
    margin = 0.05
    robot, path, coll, t, md, world = _model(oracle_mod, M=60, margin=margin, mode=mode, seed=3)
    x = random_cfgs(robot, 6, seed=4)
    comps, _ = oracle_mod.traj_cost(md, [x], robot.joints.lower_limits, robot.joints.upper_limits, W_ALL,
                                    grad=False)

    def f(d):
        if mode == 0:
            return np.maximum(margin - d, 0.0) ** 2
        d2 = np.minimum(d, margin)
        v = np.where(d2 < 0, d2 - 0.5 * margin, -0.5 / (margin + 1e-6) * (d2 - margin) ** 2)
        return -np.minimum(v, 0.0)

    pi, pj = np.asarray(coll.active_idx_i), np.asarray(coll.active_idx_j)
    total = 0.0
    for q in x:
        links, ws = independent.world_spheres(path, q)
        for l in links:
            s = ws[l]
            if len(s):
                d = (distance.cdist(s[:, :3], world[:, :3]) - s[:, 3:4] - world[None, :, 3]).min(0)
                total += f(d).sum()
        for i, j in zip(pi, pj):
            a, b = ws[links[i]], ws[links[j]]
            n = min(len(a), len(b))  # same-slot pairs: slot s of link i with slot s of link j
            if n:
                d = np.linalg.norm(a[:n, :3] - b[:n, :3], axis=1) - a[:n, 3] - b[:n, 3]
                total += f(d.min())
    assert total > 0
    # float32 tables and the +1e-6 bias of the reference distance
    assert comps[0, 2] == pytest.approx(total, rel=1e-5)


def test_activation_pieces_meet():
    """colldist: value m/2 and slope -1 at d = 0; value 0 and slope 0 at d = m."""
    m = 0.05
    inner = lambda d: 0.5 / (m + 1e-6) * (d - m) ** 2  # noqa: E731
    outer = lambda d: 0.5 * m - d  # noqa: E731
    assert inner(0.0) == pytest.approx(outer(0.0), rel=1e-4)
    h = 1e-7
    assert (inner(h) - inner(0.0)) / h == pytest.approx(-1.0, rel=1e-4)
    assert inner(m) == 0.0


@pytest.mark.parametrize("mode", [0, 1])
def test_gradient_matches_finite_differences(oracle_mod, mode):
    robot, path, coll, t, md, world = _model(oracle_mod, M=50, margin=0.08, mode=mode, seed=7)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    x = _line(robot, 9, seed=8) + np.random.default_rng(9).normal(0, 0.05, (9, 7))
    x[2, 1] = hi[1] + 0.05  # a limit violation
    w = (1.3, 1.0, 0.5, 0.1, 2.0, 5.0)
    comps, g = oracle_mod.traj_cost(md, [x], lo, hi, w)
    assert comps[0, 2] > 0
    h = 1e-6
    fd = np.zeros_like(x)
    for i in range(x.shape[0]):
        for a in range(7):
            xp, xm = x.copy(), x.copy()
            xp[i, a] += h
            xm[i, a] -= h
            cp, _ = oracle_mod.traj_cost(md, [xp], lo, hi, w, grad=False)
            cm, _ = oracle_mod.traj_cost(md, [xm], lo, hi, w, grad=False)
            tot = lambda c: w[0] * c[0, 0] + w[4] * c[0, 1] + w[5] * c[0, 2]  # noqa: E731
            fd[i, a] = (tot(cp) - tot(cm)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-6)


def test_two_loop_matches_dense_bfgs(oracle_mod):
    rng = np.random.default_rng(11)
    n, m = 12, 6
    Q = rng.standard_normal((n, n))
    H_true = Q @ Q.T + n * np.eye(n)
    s_list = [rng.standard_normal(n) for _ in range(4)]
    y_list = [H_true @ s for s in s_list]  # curvature pairs of a quadratic: s.y > 0
    sb, yb, rho = np.zeros((m, n)), np.zeros((m, n)), np.zeros(m)
    newest = 0
    for k, (s, y) in enumerate(zip(s_list, y_list)):
        newest = (newest + 1) % m
        sb[newest], yb[newest], rho[newest] = s, y, 1 / (s @ y)
    g = rng.standard_normal(n)
    d = oracle_mod.two_loop(g, sb, yb, rho, m_used=4, newest=newest)
    gamma = s_list[-1] @ y_list[-1] / (y_list[-1] @ y_list[-1])
    H = gamma * np.eye(n)
    for s, y in zip(s_list, y_list):
        r = 1 / (s @ y)
        V = np.eye(n) - r * np.outer(y, s)
        H = V.T @ H @ V + r * np.outer(s, s)
    np.testing.assert_allclose(d, -H @ g, rtol=1e-10, atol=1e-12)


def test_lbfgs_converges_to_minimum_energy_path(oracle_mod):
    robot, _, _, t, md, _ = _model(oracle_mod, M=0, pairs=False)
    md.world = np.array([[100.0, 100.0, 100.0, 0.01]])
    lo, hi = np.full(7, -50.0), np.full(7, 50.0)
    T = 14
    rng = np.random.default_rng(5)
    x0 = _line(robot, T, seed=6) + np.concatenate([np.zeros((1, 7)), rng.normal(0, 0.3, (T - 2, 7)),
                                                   np.zeros((1, 7))])
    best, cost, _ = oracle_mod.traj_lbfgs(md, [x0], lo, hi, n_iters=200, w_collision=0.0, w_collision_max=0.0)
    # Minimum of the quadratic smoothness with both endpoints fixed, per joint.
    w = (1.0, 1.0, 0.5, 0.1, 0.0, 0.0)
    _, g0 = oracle_mod.traj_cost(md, [np.zeros((T, 7))], lo, hi, w)
    Hm = np.zeros((T * 7, T * 7))
    for k in range(T * 7):
        e = np.zeros(T * 7)
        e[k] = 1.0
        _, gk = oracle_mod.traj_cost(md, [e.reshape(T, 7)], lo, hi, w)
        Hm[:, k] = (gk - g0).ravel()  # the gradient of a quadratic is linear
    free = np.arange(7, (T - 1) * 7)
    fixed = np.r_[np.arange(7), np.arange((T - 1) * 7, T * 7)]
    xs = x0.ravel().copy()
    xs[free] = np.linalg.solve(Hm[np.ix_(free, free)], -Hm[np.ix_(free, fixed)] @ xs[fixed])
    np.testing.assert_allclose(best, xs.reshape(T, 7), atol=1e-5)
    comps, _ = oracle_mod.traj_cost(md, [best], lo, hi, w, grad=False)
    assert cost[0] == pytest.approx(comps[0, 0], rel=1e-9)


def _lbfgs_dense(oracle_mod, md, x0, lo, hi, n_iters, w_coll0, w_coll_max, scale=2.0, esc=20, m=6):
    """The direct L-BFGS loop written from the reference's description with dense
    algebra: the inverse Hessian is rebuilt each iteration from gamma I and the
    last m accepted curvature pairs (oldest first), and the cost and gradient
    come from the pinned ``traj_cost``.  Returns (last, best, best cost, picks)."""
    alphas = (1.0, 0.5, 0.25, 0.1, 0.025)
    T, A = x0.shape
    n = T * A
    mask = np.ones(n)
    mask[:A] = 0
    mask[n - A:] = 0

    def total(x, wc, grad):
        w = (1.0, 1.0, 0.5, 0.1, 1.0, wc)
        c, g = oracle_mod.traj_cost(md, [x.reshape(T, A)], lo, hi, w, grad=grad)
        v = c[0, 0] + c[0, 1] + wc * c[0, 2]
        return (v, g.ravel() * mask) if grad else v

    x = x0.ravel().copy()
    wc = w_coll0
    best_cost, _ = total(x, wc, True)
    best = x.copy()
    xp, gp = x.copy(), None
    pairs, picks, fallback = [], [], 0
    for k in range(n_iters):
        if k > 0 and k % esc == 0:
            wc = min(wc * scale, w_coll_max)
        cost, g = total(x, wc, True)
        if k > 0:
            s, y = x - xp, g - gp
            if s @ y > 1e-10 * (y @ y) + 1e-30:
                pairs = (pairs + [(s, y)])[-m:]
        if pairs:
            s_n, y_n = pairs[-1]
            H = (s_n @ y_n) / (y_n @ y_n + 1e-18) * np.eye(n)
            for s, y in pairs:
                r = 1.0 / (s @ y + 1e-30)
                V = np.eye(n) - r * np.outer(y, s)
                H = V.T @ H @ V + r * np.outer(s, s)
            d = -H @ g
        else:
            d = -g / (np.linalg.norm(g) + 1e-18)
        d *= mask
        trials = [total(x + a * d, wc, False) for a in alphas]
        ok = [i for i, v in enumerate(trials) if v < cost * (1 - 1e-4)]
        pick = ok[0] if ok else int(np.argmin(trials))
        fallback += not ok
        picks.append(pick)
        xp, gp = x.copy(), g
        x = x + alphas[pick] * d
        ev = total(x, w_coll_max, False)
        if ev < best_cost:
            best_cost, best = ev, x.copy()
    return x.reshape(T, A), best.reshape(T, A), best_cost, picks, fallback


@pytest.mark.parametrize("noise,M,margin,mode,w_coll", [(0.05, 40, 0.05, 0, 30.0), (0.05, 40, 0.1, 1, 30.0),
                                                        (0.2, 0, 0.02, 1, 1.0)])
def test_lbfgs_iterates_match_dense_restatement(oracle_mod, noise, M, margin, mode, w_coll):
    """Last iterate, accepted steps, best iterate and best cost of the oracle's
    L-BFGS equal a dense-algebra restatement step for step over 45 iterations
    (escalations at 20 and 40).  A colliding world (the best stays the start,
    as in the reference) and an empty one (the best moves with the iterate)."""
    robot, _, _, t, md, world = _model(oracle_mod, M=max(M, 1), margin=margin, mode=mode, seed=4)
    if M == 0:
        md.world = np.array([[100.0, 100.0, 100.0, 0.01]])
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    T = 10
    rng = np.random.default_rng(8)
    x0 = _line(robot, T, seed=9)
    x0[1:-1] += rng.normal(0, noise, (T - 2, 7))
    last, best, bc, picks, fallback = _lbfgs_dense(oracle_mod, md, x0, lo, hi, 45, w_coll, 100.0)
    obest, ocost, _, olast, trace = oracle_mod.traj_lbfgs(md, [x0], lo, hi, n_iters=45, w_collision=w_coll,
                                                          iterates=True)
    print("picks", np.bincount(picks, minlength=5), "fallback", fallback)
    assert [int(p) for p in trace[0, :, 1]] == picks
    # dense and two-loop algebra round differently: 1e-8 after 45 iterations
    np.testing.assert_allclose(olast, last, atol=1e-7, rtol=0)
    np.testing.assert_allclose(obest, best, atol=1e-7, rtol=0)
    assert ocost[0] == pytest.approx(bc, rel=1e-9)
    if M > 0:
        assert set(picks) == {0, 1, 2, 3, 4}  # every step size of the list is taken
    else:
        assert np.abs(best - x0).max() > 1e-2  # the tracked best moved
