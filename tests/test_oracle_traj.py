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
* CHOMP      -- the truncated CG solves the dense pentadiagonal precision
  operator (numpy) when run long enough and stops where its tolerance says;
  one CHOMP step (clip, endpoint rows, covariant or plain direction, six
  clamped trials, first-minimum pick, accept rule) equals a dense numpy step
  on the pinned cost and gradient; early stopping and monotone accepted costs.
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


# ----------------------------------------------------------------- CHOMP
def _chomp_matrix(n, reg):
    """Dense interior precision operator of _chomp_metric_apply_1d: the pentadiagonal (1, -4, 6, -4, 1) + reg I."""
    M = (6.0 + reg) * np.eye(n)
    for k, v in ((1, -4.0), (2, 1.0)):
        M += v * (np.eye(n, k=k) + np.eye(n, k=-k))
    return M


@pytest.mark.parametrize("n,reg", [(5, 1e-3), (30, 1e-3), (126, 0.5)])
def test_chomp_cg_solves_the_precision_operator(oracle_mod, n, reg):
    """Enough CG steps solve M x = b (scaled so the reference's 1e-9 guards are negligible)."""
    b = np.random.default_rng(n).normal(size=n) * 1e3
    x = oracle_mod.chomp_cg(b, reg, 4 * n + 50, 1e-15)
    ref = np.linalg.solve(_chomp_matrix(n, reg), b)
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-6 * np.abs(ref).max())


def test_chomp_cg_truncation(oracle_mod):
    """n steps of CG on an n x n system reach the solution, fewer steps do not; a
    loose tolerance stops early (the stopping test compares |r|^2 with (tol |b|)^2)."""
    n = 8
    b = np.random.default_rng(1).normal(size=n) * 1e3
    M = _chomp_matrix(n, 0.5)
    ref = np.linalg.solve(M, b)
    full = oracle_mod.chomp_cg(b, 0.5, n, 1e-15)
    np.testing.assert_allclose(full, ref, atol=1e-6 * np.abs(ref).max())
    part = oracle_mod.chomp_cg(b, 0.5, 2, 1e-15)
    assert np.abs(part - ref).max() > 1e-3 * np.abs(ref).max()
    loose = oracle_mod.chomp_cg(b, 0.5, 50, 0.5)
    assert np.linalg.norm(M @ loose - b) <= 0.5 * np.linalg.norm(b) + 1e-9
    assert np.abs(loose - ref).max() > 1e-6 * np.abs(ref).max()


def _chomp_one_step(oracle_mod, md, x0, lo, hi, covariant, step=0.05, clip=10.0, max_delta=0.05, w_coll=3.0):
    """One CHOMP iteration in dense numpy around the oracle's pinned cost and gradient."""
    w = (1.0, 0.0, 0.5, 0.1, 1.0, w_coll)
    comps, g = oracle_mod.traj_cost(md, [x0], lo, hi, w, grad=True)
    cost = comps[0] @ [1.0, 1.0, w_coll]
    gn = np.linalg.norm(g)
    if clip > 0 and gn > clip:
        g = g * clip / (gn + 1e-12)
    g[0] = g[-1] = 0.0
    if covariant:
        T = len(x0)
        d = np.zeros_like(g)
        d[1:-1] = -np.linalg.solve(_chomp_matrix(T - 2, 1e-3), g[1:-1])
    else:
        d = -g
    trials = []
    for a in (1.0, 0.5, 0.25, 0.1, 0.05, 0.0):
        xt = x0 + np.clip(step * a * d, -max_delta, max_delta)
        c, _ = oracle_mod.traj_cost(md, [xt], lo, hi, w, grad=False)
        trials.append((c[0] @ [1.0, 1.0, w_coll], xt))
    k = int(np.argmin([c for c, _ in trials]))
    improved = trials[k][0] < cost - 1e-5
    return (trials[k][1] if improved else x0), k, improved


@pytest.mark.parametrize("covariant", [False, True])
@pytest.mark.parametrize("noise", [0.0, 0.05])
def test_chomp_step_matches_dense_restatement(oracle_mod, covariant, noise):
    """The oracle's first iterate, picked trial and accept decision equal a dense
    numpy step built on the (separately pinned) cost and gradient, with the
    covariant direction from numpy's linear solve (CG run to convergence)."""
    robot, _, _, t, md, _ = _model(oracle_mod, M=60, margin=0.05, mode=0, seed=3)
    rng = np.random.default_rng(4)
    x0 = _line(robot, 40, seed=5) + rng.normal(0, noise, (40, 7))
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    want, k, imp = _chomp_one_step(oracle_mod, md, x0, lo, hi, covariant)
    got, _, _, tr = oracle_mod.traj_chomp(md, [x0], lo, hi, n_iters=1, use_covariant_update=covariant,
                                          n_cg_iters=400, cg_tol=1e-14, trace=True)
    assert tr[0, 0, 1] == k and tr[0, 0, 2] == imp
    moved = np.abs(want - x0).max()
    assert imp and moved > 1e-3
    # The reference's CG guards (+1e-9 in both divisions) stall its residual near
    # |r|^2 ~ 1e-9, so the covariant step matches the exact solve to ~1e-3 of the step.
    np.testing.assert_allclose(got, want, atol=(2e-3 * moved if covariant else 1e-9), rtol=0)


def test_chomp_early_stop_and_weight_schedule(oracle_mod):
    """A path that cannot improve (straight, in free space) is stopped after
    `patience` iterations; on a colliding path the accepted costs fall while the
    collision weight is held (scale 1)."""
    robot, _, _, t, md, _ = _model(oracle_mod, M=0, pairs=False)
    md.world = np.array([[100.0, 100.0, 100.0, 0.01]])
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    q0 = np.array([0.0, -0.785, 0.0, -2.356, 0.0, 1.571, 0.785])
    s = np.linspace(0, 1, 30)[:, None]
    line = (1 - s) * q0 + s * (q0 + 0.2)
    out, _, _, tr = oracle_mod.traj_chomp(md, [line], lo, hi, n_iters=10, early_stop_patience=3, trace=True)
    np.testing.assert_array_equal(tr[0, :3, 2], 0)
    np.testing.assert_array_equal(tr[0, 3:], -1)
    np.testing.assert_array_equal(out, line)
    robot, _, _, t, md, _ = _model(oracle_mod, M=80, margin=0.05, mode=0, seed=7)
    x0 = _line(robot, 32, seed=8)
    _, _, _, tr = oracle_mod.traj_chomp(md, [x0], lo, hi, n_iters=25, collision_penalty_scale=1.0, trace=True)
    c = tr[0, :, 0]
    run = c[tr[0, :, 2] >= 0]
    assert np.all(np.diff(run) <= 1e-12)


# ----------------------------------------------------------------- SCO
def _sco_model(oracle_mod, M=120, seed=11):
    robot, path, coll, t, md, world = _model(oracle_mod, M=M, margin=0.01, mode=1, seed=seed)
    gstart = np.array([0, M // 2, M], np.int32)
    return robot, coll, t, md, world, gstart


def test_sco_linearize_gradient_matches_finite_differences(oracle_mod):
    robot, _, _, md, _, gs = _sco_model(oracle_mod)
    q = random_cfgs(robot, 6, seed=12)
    d, J = oracle_mod.sco_linearize(md, q, gs, 0.05)
    assert d.shape == (6, 3) and J.shape == (6, 3, 7)
    h = 1e-6
    for a in range(7):
        qp, qm = q.copy(), q.copy()
        qp[:, a] += h
        qm[:, a] -= h
        fd = (oracle_mod.sco_linearize(md, qp, gs, 0.05)[0] - oracle_mod.sco_linearize(md, qm, gs, 0.05)[0]) / (2 * h)
        np.testing.assert_allclose(J[:, :, a], fd, atol=1e-6 * max(1.0, np.abs(J).max()))


def test_sco_smooth_min_bounds_and_limit(oracle_mod):
    """dmin - T log(n) <= smooth min <= dmin, and it tends to the group minimum of the
    JAX distances (same-slot self pairs; closest sphere per link and world sphere)."""
    robot, coll, t, md, world, gs = _sco_model(oracle_mod)
    q = random_cfgs(robot, 5, seed=13)
    lp = oracle_mod.fk_links(t, q)
    sp = oracle_mod.padded_spheres(coll)
    selfd = oracle_mod.self_dist(lp, sp, coll.active_idx_i, coll.active_idx_j, mode=1)
    full, _ = oracle_mod.world_dist(lp, sp, world, bias=True)
    mins = np.stack([selfd.min(1), full[:, :, :60].min((1, 2)), full[:, :, 60:].min((1, 2))], 1)
    counts = np.array([selfd.shape[1], full.shape[1] * 60, full.shape[1] * 60])
    for tau in (0.05, 0.01):
        d, _ = oracle_mod.sco_linearize(md, q, gs, tau)
        assert np.all(d <= mins + 1e-12) and np.all(d >= mins - tau * np.log(counts) - 1e-12)
    d, _ = oracle_mod.sco_linearize(md, q, gs, 1e-5)
    np.testing.assert_allclose(d, mins, atol=1e-5 * np.log(counts).max())


def test_sco_inner_without_contact_reaches_the_quadratic_minimum(oracle_mod):
    """With the linearised hinge inactive (a negative margin) the inner problem is
    smoothness + trust region, a quadratic with pinned endpoints: enough L-BFGS
    steps reach numpy's linear solve of its normal equations."""
    robot, _, _, md, _, gs = _sco_model(oracle_mod)
    T, A = 16, 7
    q0, q1 = random_cfgs(robot, 2, seed=14, margin=0.3)
    s = np.linspace(0, 1, T)[:, None]
    x0 = (1 - s) * q0 + s * q1 + np.random.default_rng(15).normal(0, 0.05, (T, A))
    x0[0], x0[-1] = q0, q1
    w_vel, w_acc, w_jerk, w_trust = 1.0, 0.5, 0.1, 0.5
    out, _, _ = oracle_mod.traj_sco(md, [x0], robot.joints.lower_limits, robot.joints.upper_limits, gs,
                                    n_outer_iters=1, n_inner_iters=300, collision_margin=-10.0, w_trust=w_trust,
                                    w_limits=0.0)
    # dense quadratic: smooth = x^T H x per joint, H from the stencils
    D1 = np.diff(np.eye(T), axis=0)
    D2 = np.zeros((T - 4, T))
    for i in range(T - 4):
        D2[i, i:i + 5] = np.array([-1, 16, -30, 16, -1]) / 12.0
    D3 = np.diff(D2, axis=0)
    H = w_vel * D1.T @ D1 + w_acc * D2.T @ D2 + w_jerk * D3.T @ D3 + w_trust * np.eye(T)
    want = x0.copy()
    inner = slice(1, T - 1)
    for a in range(A):
        rhs = w_trust * x0[:, a] - H[:, [0, T - 1]] @ x0[[0, T - 1], a]
        want[inner, a] = np.linalg.solve(H[inner, inner], rhs[inner])
    np.testing.assert_allclose(out, want, atol=1e-5)
