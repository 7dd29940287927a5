"""Pins of the FP64 oracle (CPU).

The reference ships no golden vectors or known-answer tests for this path
(its tests/ are benchmark scripts with runtime JAX-vs-CUDA comparisons), so
every oracle function is pinned here against something that shares no code
with it:

* FK        -- the Franka known answer at q = 0, and an independent FK built
               from the raw URDF XML with scipy rotations (4x4 matrices).
* distances -- scipy ``cdist`` on independently placed sphere centres.
* residual  -- scipy rotation vectors of R_ee R_tgt^T.
* Jacobian  -- central finite differences of the oracle FK (the geometric
               Jacobian is exact for position rows, first order for rotation).
* LM step   -- the damped normal-equation identity solved with numpy.linalg.
* LM solve  -- a 2-link planar arm with a closed-form IK, and success on
               FK-generated (hence reachable) targets.
"""

import independent
import numpy as np
import pytest
from conftest import GOLDEN, ROBOTS, load_robot, random_cfgs
from scipy.spatial.transform import Rotation

from pyroffi_amd.collision import RobotCollisionSpherized
from pyroffi_amd.optimization_engines._ls_ik import ee_tables


def test_fk_known_answer_panda(oracle_mod):
    robot, _ = load_robot("panda")
    t = oracle_mod.Tables.of(robot)
    lp = oracle_mod.fk_links(t, np.zeros(7))
    names = robot.links.names
    np.testing.assert_allclose(lp[names.index("panda_link8"), 4:], [0.088, 0.0, 0.926], atol=1e-6)
    np.testing.assert_allclose(lp[names.index("panda_hand"), 4:], [0.088, 0.0, 0.926], atol=1e-6)
    np.testing.assert_allclose(lp[names.index("panda_grasptarget"), 4:], [0.088, 0.0, 0.821], atol=1e-6)
    # flange points down: 180 degrees about x
    R = Rotation.from_quat(lp[names.index("panda_link8"), [1, 2, 3, 0]]).as_matrix()
    np.testing.assert_allclose(R, np.diag([1.0, -1.0, -1.0]), atol=1e-6)


@pytest.mark.parametrize("name", ROBOTS + ("mimic_arm",))
def test_fk_matches_independent(oracle_mod, name):
    robot, path = load_robot(name)
    t = oracle_mod.Tables.of(robot)
    q = random_cfgs(robot, 20, seed=1, margin=0.3)
    lp = oracle_mod.fk_links(t, q)
    # float32 parent quaternions are unit to ~1e-7
    np.testing.assert_allclose(np.linalg.norm(lp[..., :4], axis=-1), 1.0, atol=1e-6)
    for b in range(len(q)):
        ref = independent.link_matrices(path, q[b])
        got = independent.pose_to_matrix(lp[b])
        # the oracle takes float32 tables, as the kernels do
        np.testing.assert_allclose(got, ref, atol=2e-6)


def test_fk_joint_poses_and_link_gather(oracle_mod):
    robot, _ = load_robot("mimic_arm")
    t = oracle_mod.Tables.of(robot)
    q = random_cfgs(robot, 5, seed=2)
    jp = oracle_mod.fk_joints(t, q)
    lp = oracle_mod.fk_links(t, q)
    for l, j in enumerate(robot.links.parent_joint_indices):
        if j < 0:
            np.testing.assert_array_equal(lp[:, l], np.tile([1, 0, 0, 0, 0, 0, 0], (5, 1)))
        else:
            np.testing.assert_array_equal(lp[:, l], jp[:, j])


def _world(rng, m, lo=0.02, hi=0.15):
    return np.concatenate([rng.uniform(-1, 1, (m, 3)), rng.uniform(lo, hi, (m, 1))], axis=1)


@pytest.mark.parametrize("name", ("panda", "fetch", "mimic_arm"))
def test_world_distance_matches_cdist(oracle_mod, name):
    robot, path = load_robot(name)
    coll = RobotCollisionSpherized.from_urdf(path)
    t = oracle_mod.Tables.of(robot)
    rng = np.random.default_rng(5)
    q = random_cfgs(robot, 6, seed=3)
    world = _world(rng, 37)
    lp = oracle_mod.fk_links(t, q)
    out, mins = oracle_mod.world_dist(lp, oracle_mod.padded_spheres(coll), world)
    for b in range(len(q)):
        ref = independent.world_distance_matrix(path, q[b], world)
        np.testing.assert_allclose(out[b], ref, atol=2e-6)
        np.testing.assert_allclose(mins[b], ref.min(), atol=2e-6)


def test_world_distance_closed_form(oracle_mod):
    """Two spheres on the x axis: d = |dx| - r1 - r2, negative in penetration."""
    pose = np.array([[[1, 0, 0, 0, 0.0, 0, 0]]])
    spheres = np.array([[[0.5, 0, 0, 0.1], [0, 0, 0, -1e9]]])
    world = np.array([[2.0, 0, 0, 0.2], [0.55, 0, 0, 0.3]])
    out, mins = oracle_mod.world_dist(pose, spheres, world)
    np.testing.assert_allclose(out[0, 0], [1.5 - 0.3, 0.05 - 0.4])
    assert mins[0] == pytest.approx(-0.35)
    # JAX semantics: +1e-6 inside the norm on every axis, padding evaluated at r = -1e9
    out_j, _ = oracle_mod.world_dist(pose, spheres, world, bias=True)
    np.testing.assert_allclose(out_j[0, 0], [np.sqrt((1.5 + 1e-6) ** 2 + 2e-12) - 0.3,
                                             np.sqrt((0.05 + 1e-6) ** 2 + 2e-12) - 0.4], atol=1e-12)
    # a link without spheres reports the 1e9 sentinel (CUDA semantics)
    out_e, _ = oracle_mod.world_dist(pose, np.array([[[0, 0, 0, -1e9]]]), world)
    assert np.all(out_e == 1e9)


def test_world_distance_translation_invariant(oracle_mod):
    robot, path = load_robot("panda")
    coll = RobotCollisionSpherized.from_urdf(path)
    t = oracle_mod.Tables.of(robot)
    rng = np.random.default_rng(9)
    lp = oracle_mod.fk_links(t, random_cfgs(robot, 4, seed=4))
    world = _world(rng, 20)
    shift = np.array([0.3, -0.7, 1.1])
    lp2 = lp.copy()
    lp2[..., 4:] += shift
    w2 = world.copy()
    w2[:, :3] += shift
    a, _ = oracle_mod.world_dist(lp, oracle_mod.padded_spheres(coll), world)
    b, _ = oracle_mod.world_dist(lp2, oracle_mod.padded_spheres(coll), w2)
    np.testing.assert_allclose(a, b, atol=1e-12)


@pytest.mark.parametrize("name", ("panda", "ur5", "mimic_arm"))
def test_self_distance_matches_cdist(oracle_mod, name):
    robot, path = load_robot(name)
    coll = RobotCollisionSpherized.from_urdf(path)
    t = oracle_mod.Tables.of(robot)
    q = random_cfgs(robot, 5, seed=6)
    lp = oracle_mod.fk_links(t, q)
    out = oracle_mod.self_dist(lp, oracle_mod.padded_spheres(coll), coll.active_idx_i, coll.active_idx_j)
    for b in range(len(q)):
        ref = independent.self_distance(path, q[b], coll.active_idx_i, coll.active_idx_j)
        np.testing.assert_allclose(out[b], ref, atol=2e-6)


def test_self_distance_jax_diagonal_semantics(oracle_mod):
    """JAX path: sphere s of link i vs sphere s of link j only (+1e-6 bias)."""
    pose = np.tile(np.array([1, 0, 0, 0, 0.0, 0, 0]), (1, 2, 1))
    spheres = np.array([[[0, 0, 0, 0.1], [5, 0, 0, 0.1]], [[1, 0, 0, 0.1], [0.2, 0, 0, 0.1]]])
    full = oracle_mod.self_dist(pose, spheres, [1], [0], mode=0)
    diag = oracle_mod.self_dist(pose, spheres, [1], [0], mode=1)
    assert full[0, 0] == pytest.approx(0.0)  # sphere 0 of link 0 vs sphere 1 of link 1
    d0 = np.sqrt((1 - 1e-6) ** 2 + 2e-12) - 0.2
    d1 = np.sqrt((4.8 + 1e-6) ** 2 + 2e-12) - 0.2
    assert diag[0, 0] == pytest.approx(min(d0, d1), abs=1e-12)


def _targets(robot, oracle_mod, t, link, n, seed):
    q = random_cfgs(robot, n, seed=seed)
    return q, oracle_mod.fk_links(t, q)[:, link]


@pytest.mark.parametrize("name", ("panda", "mimic_arm"))
def test_residual_matches_rotvec(oracle_mod, name):
    robot, _ = load_robot(name)
    t = oracle_mod.Tables.of(robot)
    link = robot.links.num_links - 1 if name == "panda" else robot.links.names.index("finger_c")
    tj, masks = ee_tables(robot, (link,))
    _, tgt = _targets(robot, oracle_mod, t, link, 8, 11)
    q = random_cfgs(robot, 8, seed=12)
    for b in range(8):
        r, _ = oracle_mod.res_jac(t, q[b], tj, masks, tgt[b])
        cur = oracle_mod.fk_links(t, q[b])[link]
        np.testing.assert_allclose(r[:3], cur[4:] - tgt[b, 4:], atol=1e-12)
        Rc = Rotation.from_quat(cur[[1, 2, 3, 0]])
        Rt = Rotation.from_quat(tgt[b, [1, 2, 3, 0]])
        np.testing.assert_allclose(r[3:], (Rc * Rt.inv()).as_rotvec(), atol=1e-9)


@pytest.mark.parametrize("name", ("panda", "fetch", "mimic_arm"))
def test_jacobian_matches_finite_differences(oracle_mod, name):
    robot, _ = load_robot(name)
    t = oracle_mod.Tables.of(robot)
    link = robot.links.num_links - 1
    tj, masks = ee_tables(robot, (link,))
    q = random_cfgs(robot, 4, seed=13)
    h = 1e-6
    for b in range(4):
        tgt = oracle_mod.fk_links(t, q[b])[link]  # at the target: rotation residual is linear
        _, Jm = oracle_mod.res_jac(t, q[b], tj, masks, tgt)
        fd = np.zeros_like(Jm)
        for a in range(robot.joints.num_actuated_joints):
            dq = np.zeros_like(q[b])
            dq[a] = h
            rp, _ = oracle_mod.res_jac(t, q[b] + dq, tj, masks, tgt)
            rm, _ = oracle_mod.res_jac(t, q[b] - dq, tj, masks, tgt)
            fd[:, a] = (rp - rm) / (2 * h)
        # the reference uses the unit joint axis even for a non-unit URDF axis
        # (mimic_arm's j3 has |axis| = sqrt(2)), so its column is 1/sqrt(2) of d/dq
        if name == "mimic_arm":
            fd[:, 0] /= np.sqrt(2.0)
        np.testing.assert_allclose(Jm, fd, atol=2e-6)


def test_lm_direction_identity(oracle_mod):
    """delta solves (Js^T Js + lam I) p = -Js^T W r with Js = W J S^-1, delta = S^-1 p."""
    rng = np.random.default_rng(14)
    W = np.array([50, 50, 50, 10, 10, 10.0])
    for lam in (1e-2, 0.3, 5.0):
        Jm = rng.normal(size=(6, 7))
        r = rng.normal(size=6)
        fixed = np.array([0, 0, 1, 0, 0, 0, 0])
        delta, ok = oracle_mod.lm_direction(r, Jm, W, lam, fixed)
        assert ok
        Jw = W[:, None] * Jm
        s = np.linalg.norm(Jw, axis=0) + 1e-8
        Js = Jw / s
        A = Js.T @ Js + lam * np.eye(7)
        rhs = -Js.T @ (W * r)
        A[2, :] = 0
        A[:, 2] = 0
        A[2, 2] = 1
        rhs[2] = 0
        np.testing.assert_allclose(delta, np.linalg.solve(A, rhs) / s, rtol=1e-10, atol=1e-12)
        assert delta[2] == 0


def test_planar_2r_closed_form(oracle_mod, tmp_path):
    """2-link planar arm (l1 = 0.5, l2 = 0.3): LM reaches the analytic solutions."""
    urdf = tmp_path / "planar.urdf"
    urdf.write_text(
        """<robot name="p"><link name="b"/><link name="a"/><link name="c"/><link name="e"/>
        <joint name="q1" type="revolute"><parent link="b"/><child link="a"/><axis xyz="0 0 1"/>
          <limit lower="-3" upper="3" velocity="1"/></joint>
        <joint name="q2" type="revolute"><parent link="a"/><child link="c"/><origin xyz="0.5 0 0"/>
          <axis xyz="0 0 1"/><limit lower="-3" upper="3" velocity="1"/></joint>
        <joint name="tip" type="fixed"><parent link="c"/><child link="e"/><origin xyz="0.3 0 0"/></joint>
        </robot>"""
    )
    import pyroffi_amd as pk

    robot = pk.Robot.from_urdf(str(urdf))
    t = oracle_mod.Tables.of(robot)
    tj, masks = ee_tables(robot, (3,))
    rng = np.random.default_rng(15)
    for _ in range(5):
        x, y = rng.uniform(0.3, 0.7) * np.array([np.cos(a := rng.uniform(-3, 3)), np.sin(a)])
        c2 = (x * x + y * y - 0.25 - 0.09) / (2 * 0.5 * 0.3)
        q2 = np.arccos(c2)
        q1 = np.arctan2(y, x) - np.arctan2(0.3 * np.sin(q2), 0.5 + 0.3 * np.cos(q2))
        th = q1 + q2
        tgt = np.array([np.cos(th / 2), 0, 0, np.sin(th / 2), x, y, 0])
        seeds = rng.uniform(-3, 3, (1, 8, 2))
        cfgs, errs = oracle_mod.ls_ik(t, seeds, tj, masks, tgt[None, None], robot.joints.lower_limits,
                                      robot.joints.upper_limits, max_iter=60)
        best = cfgs[0, np.argmin(errs[0])]
        pos = oracle_mod.fk_links(t, best)[3, 4:6]
        np.testing.assert_allclose(pos, [x, y], atol=1e-6)
        # elbow-up or elbow-down closed-form solution
        sols = [np.array([q1, q2]), np.array([np.arctan2(y, x) + np.arctan2(0.3 * np.sin(q2), 0.5 + 0.3 * np.cos(q2)), -q2])]
        wrap = lambda v: (v + np.pi) % (2 * np.pi) - np.pi
        assert min(np.abs(wrap(best - s)).max() for s in sols) < 1e-4


def test_lm_success_on_reachable_targets(oracle_mod):
    """Panda, FK-generated targets: most problems reach 1 mm / 0.01 rad."""
    robot, _ = load_robot("panda")
    t = oracle_mod.Tables.of(robot)
    link = robot.links.names.index("panda_hand")
    tj, masks = ee_tables(robot, (link,))
    _, tgt = _targets(robot, oracle_mod, t, link, 6, 16)
    rng = np.random.default_rng(17)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    seeds = rng.uniform(lo, hi, (6, 16, 7))
    cfgs, errs = oracle_mod.ls_ik(t, seeds, tj, masks, tgt[:, None], lo, hi, max_iter=30, lambda_init=1e-2)
    best = cfgs[np.arange(6), np.argmin(errs, axis=1)]
    got = oracle_mod.fk_links(t, best)[:, link]
    pos = np.linalg.norm(got[:, 4:] - tgt[:, 4:], axis=1)
    rot = np.linalg.norm((Rotation.from_quat(tgt[:, [1, 2, 3, 0]]).inv() * Rotation.from_quat(got[:, [1, 2, 3, 0]])).as_rotvec(), axis=1)
    assert np.mean((pos < 1e-3) & (rot < 0.01)) >= 5 / 6
    assert np.all(best >= lo - 1e-12) and np.all(best <= hi + 1e-12)
    # the all-time best error never exceeds the seed's own error
    e0 = oracle_mod.ls_ik(t, seeds, tj, masks, tgt[:, None], lo, hi, max_iter=0)[1]
    assert np.all(errs <= e0 + 1e-12)


def test_pose_hinge_cost_matches_independent(oracle_mod):
    """Cost = |r|^2 (residual pinned above) + sum max(0, m - d)^2 with d from scipy cdist."""
    from scipy.spatial import distance

    robot, path = load_robot("panda")
    coll = RobotCollisionSpherized.from_urdf(path)
    t = oracle_mod.Tables.of(robot)
    link = robot.links.names.index("panda_hand")
    tj, masks = ee_tables(robot, (link,))
    rng = np.random.default_rng(21)
    world = np.concatenate([rng.uniform(-1, 1, (300, 3)), rng.uniform(0.02, 0.15, (300, 1))], 1)
    q = random_cfgs(robot, 5, seed=22)
    tgt = oracle_mod.fk_links(t, random_cfgs(robot, 5, seed=23))[:, link]
    margin = 0.05
    cost = oracle_mod.pose_hinge_cost(t, q, oracle_mod.padded_spheres(coll), world, margin, link, tgt,
                                      pos_weight=2.0, ori_weight=0.5)
    for b in range(5):
        r, _ = oracle_mod.res_jac(t, q[b], tj, masks, tgt[b])
        pose = 4.0 * np.sum(r[:3] ** 2) + 0.25 * np.sum(r[3:] ** 2)
        links, ws = independent.world_spheres(path, q[b])
        hinge = 0.0
        for l in links:
            s = ws[l]
            if len(s):
                d = distance.cdist(s[:, :3], world[:, :3]) - s[:, 3:4] - world[None, :, 3]
                hinge += np.sum(np.maximum(margin - d, 0.0) ** 2)
        # float32 tables (as the kernels take them) vs float64 URDF values
        assert cost[b] == pytest.approx(pose + hinge, rel=1e-6, abs=1e-9)


def test_pose_hinge_fd_gradient_sanity(oracle_mod):
    """The FD helper reproduces an exact gradient: the pose-only cost's position part
    2 J_pos^T r_pos (+ the rotation part) equals FD of |r|^2 near the target."""
    robot, _ = load_robot("panda")
    t = oracle_mod.Tables.of(robot)
    link = robot.links.names.index("panda_hand")
    tj, masks = ee_tables(robot, (link,))
    q = random_cfgs(robot, 3, seed=24)
    tgt = oracle_mod.fk_links(t, q + 0.01)[:, link]
    fd = oracle_mod.pose_hinge_fd_grad(t, q, np.zeros((13, 1, 4)) - 1e9, np.zeros((0, 4)), 0.05, link, tgt,
                                       ori_weight=0.0)
    for b in range(3):
        r, Jm = oracle_mod.res_jac(t, q[b], tj, masks, tgt[b])
        np.testing.assert_allclose(fd[b], 2 * Jm[:3].T @ r[:3], rtol=1e-6, atol=1e-9)


def test_obstacle_distances_independent(oracle_mod):
    """Sphere: scipy cdist.  Capsule: the minimum over 20001 points sampled
    along the segment.  Box: the distance to the clamped (closest) point
    outside, minus the distance to the nearest face inside.  Half-space: the
    plane offset.  Padding (r < 0) reads 1e9."""
    import worlds
    from scipy.spatial import distance

    rng = np.random.default_rng(12)
    ws, wc, wb, wh = (x.astype(np.float64) for x in worlds.random_world(rng, 7, 6, 9, 4))
    c = rng.uniform(-1.2, 1.2, (40, 3))
    r = rng.uniform(0.0, 0.2, 40)
    r[5] = -1e9
    got = oracle_mod.obstacle_dist(c, r, ws, wc, wb, wh)
    assert got.shape == (40, 26)
    np.testing.assert_array_equal(got[5], 1e9)
    keep = np.arange(40) != 5
    c, r, got = c[keep], r[keep], got[keep]
    np.testing.assert_allclose(got[:, :7], distance.cdist(c, ws[:, :3]) - r[:, None] - ws[None, :, 3], atol=1e-12)
    t = np.linspace(0, 1, 20001)
    for m, cap in enumerate(wc):
        pts = cap[None, :3] + t[:, None] * (cap[3:6] - cap[:3])[None]
        ref = distance.cdist(c, pts).min(1) - r - cap[6]
        np.testing.assert_allclose(got[:, 7 + m], ref, atol=1e-8)
    for m, box in enumerate(wb):
        R = np.stack([box[3:6], box[6:9], box[9:12]], 1)  # columns = axes
        h = box[12:15]
        loc = (c - box[:3]) @ R
        clamped = np.clip(loc, -h, h)
        outside = np.linalg.norm(loc - clamped, axis=1)
        inside = np.all(np.abs(loc) <= h, axis=1)
        face = np.min(h - np.abs(loc), axis=1)
        ref = np.where(inside, -face, outside) - r
        np.testing.assert_allclose(got[:, 13 + m], ref, atol=1e-12)
    for m, hs in enumerate(wh):
        ref = (c - hs[3:]) @ hs[:3] - r
        np.testing.assert_allclose(got[:, 22 + m], ref, atol=1e-12)
    # a centre inside a box: negative distance equal to minus the nearest face gap
    box = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0.2, 0.3]], np.float64)
    d = oracle_mod.obstacle_dist(np.array([[0.1, 0.05, -0.2]]), np.array([0.01]), boxes=box)
    assert d[0, 0] == pytest.approx(-0.1 - 0.01, abs=1e-12)


def test_capsule_distances_independent(oracle_mod):
    """Capsule robots: sphere and capsule obstacles against the minimum over
    points sampled along the segments; boxes against the clamped-point
    distance at the segment point nearest the box centre (found by sampling);
    half-spaces against the sampled minimum plane offset."""
    import worlds
    from scipy.spatial import distance

    rng = np.random.default_rng(14)
    ws, wc, wb, wh = (x.astype(np.float64) for x in worlds.random_world(rng, 5, 5, 6, 3))
    a = rng.uniform(-1, 1, (12, 3))
    caps = np.concatenate([a, a + rng.normal(0, 0.3, (12, 3)), rng.uniform(0.01, 0.1, (12, 1))], 1)
    caps[3, 3:6] = caps[3, :3]  # a degenerate (point) capsule
    got = oracle_mod.capsule_dist(caps, ws, wc, wb, wh)
    t = np.linspace(0, 1, 2001)
    for n, c in enumerate(caps):
        pts = c[None, :3] + t[:, None] * (c[3:6] - c[:3])[None]
        ref_s = distance.cdist(pts, ws[:, :3]).min(0) - c[6] - ws[:, 3]
        np.testing.assert_allclose(got[n, :5], ref_s, atol=1e-6)
        for m, o in enumerate(wc):
            q = o[None, :3] + t[:, None] * (o[3:6] - o[:3])[None]
            ref = distance.cdist(pts, q).min() - c[6] - o[6]
            np.testing.assert_allclose(got[n, 5 + m], ref, atol=1e-6)
        for m, box in enumerate(wb):
            R = np.stack([box[3:6], box[6:9], box[9:12]], 1)
            loc = (pts - box[:3]) @ R
            p = loc[np.argmin(np.linalg.norm(loc, axis=1))]
            h = box[12:15]
            inside = np.all(np.abs(p) <= h)
            ref = (-np.min(h - np.abs(p)) if inside else np.linalg.norm(p - np.clip(p, -h, h))) - c[6]
            np.testing.assert_allclose(got[n, 10 + m], ref, atol=2e-3)
        for m, hs in enumerate(wh):
            np.testing.assert_allclose(got[n, 16 + m], ((pts - hs[3:]) @ hs[:3]).min() - c[6], atol=1e-12)
