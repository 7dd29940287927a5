"""GPU parity of the cost residuals (``pyroffi_amd.costs``: ``pose_cost``,
``self_collision_cost``, ``world_collision_cost``, restating
``src/pyroffi/costs/_costs.py:11-29, 169-199``) against the FP64 oracle's
kinematics and distances composed with the host SE(3) log and the
``colldist_from_sdf`` transform (both pinned separately)."""

import numpy as np
import pytest
import torch
from conftest import load_robot, random_cfgs

from pyroffi_amd._se3 import se3_compose, se3_inverse, se3_log
from pyroffi_amd.collision import CUDARobotCollisionChecker, RobotCollisionSpherized, Sphere, colldist_from_sdf
from pyroffi_amd.costs import pose_cost, self_collision_cost, world_collision_cost

pytestmark = pytest.mark.gpu


def _pose_residual_fp64(oracle_mod, robot, q, targets, idx, pw, ow):
    lp = oracle_mod.fk_links(oracle_mod.Tables.of(robot), np.atleast_2d(q).astype(np.float64))
    if np.ndim(q) == 1:
        lp = lp[0]
    res = se3_log(se3_compose(se3_inverse(lp[..., idx, :]), targets))
    return np.concatenate([res[..., :3] * pw, res[..., 3:] * ow]).reshape(-1)


@pytest.mark.parametrize("links", ["panda_hand", ("panda_link4", "panda_hand")])
def test_pose_cost_matches_oracle(oracle_mod, links):
    robot, _ = load_robot("panda")
    names = (links,) if isinstance(links, str) else links
    idx = np.array([robot.links.names.index(n) for n in names], np.int32)
    idx = idx[0] if isinstance(links, str) else idx
    q = random_cfgs(robot, 6, seed=3).astype(np.float32)
    tq = random_cfgs(robot, 6, seed=4).astype(np.float32)
    targets = oracle_mod.fk_links(oracle_mod.Tables.of(robot), tq.astype(np.float64))[:, idx, :]
    qt = torch.as_tensor(q, device="cuda")
    got = pose_cost({"q": qt}, robot, "q", targets, idx, 50.0, 10.0).cpu().numpy()
    ref = _pose_residual_fp64(oracle_mod, robot, q, targets, idx, 50.0, 10.0)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, atol=1e-3, rtol=1e-4)  # position rows weighted by 50
    # the configuration given directly, and one configuration without a batch
    np.testing.assert_array_equal(pose_cost(qt, robot, None, targets, idx, 50.0, 10.0).cpu().numpy(), got)
    one = pose_cost(qt[0], robot, None, targets[0], idx, 1.0, 1.0).cpu().numpy()
    np.testing.assert_allclose(one, _pose_residual_fp64(oracle_mod, robot, q[0], targets[0], idx, 1.0, 1.0),
                               atol=2e-4, rtol=0)


def test_pose_cost_gradient(oracle_mod):
    """d/dq of |pose residual|^2 through the FK adjoint kernel vs FP64 central differences."""
    robot, _ = load_robot("panda")
    idx = np.int32(robot.links.names.index("panda_hand"))
    q = random_cfgs(robot, 4, seed=8)
    targets = oracle_mod.fk_links(oracle_mod.Tables.of(robot), random_cfgs(robot, 4, seed=9))[:, idx, :]
    qt = torch.as_tensor(q.astype(np.float32), device="cuda").requires_grad_(True)
    pose_cost(qt, robot, None, targets, idx, 1.0, 1.0).square().sum().backward()
    f = lambda x: np.sum(_pose_residual_fp64(oracle_mod, robot, x, targets, idx, 1.0, 1.0) ** 2)
    h = 1e-6
    fd = np.zeros_like(q)
    for b in range(q.shape[0]):
        for a in range(q.shape[1]):
            e = np.zeros_like(q)
            e[b, a] = h
            fd[b, a] = (f(q + e) - f(q - e)) / (2 * h)
    np.testing.assert_allclose(qt.grad.cpu().numpy(), fd, atol=2e-3 * np.abs(fd).max(), rtol=0)


def test_collision_costs_match_oracle(oracle_mod):
    robot, path = load_robot("panda")
    coll = RobotCollisionSpherized.from_urdf(path)
    chk = CUDARobotCollisionChecker(coll)
    q = random_cfgs(robot, 9, seed=21).astype(np.float32)
    rng = np.random.default_rng(22)
    world = np.concatenate([rng.uniform(-1, 1, (64, 3)), rng.uniform(0.02, 0.15, (64, 1))], 1).astype(np.float32)
    lp = oracle_mod.fk_links(oracle_mod.Tables.of(robot), q.astype(np.float64))
    sp = oracle_mod.padded_spheres(coll)
    full, _ = oracle_mod.world_dist(lp, sp, world.astype(np.float64))
    selfd = oracle_mod.self_dist(lp, sp, coll.active_idx_i, coll.active_idx_j)
    wg = Sphere.from_center_and_radius(world[:, :3], world[:, 3])
    qt = torch.as_tensor(q, device="cuda")
    got_w = world_collision_cost(qt, robot, chk, None, wg, 0.05, 3.0).cpu().numpy()
    ref_w = (np.asarray(colldist_from_sdf(full.astype(np.float32), 0.05)) * 3.0).reshape(-1)
    np.testing.assert_allclose(got_w, ref_w, atol=1e-5 * 3.0 * 2, rtol=1e-5)
    assert np.any(ref_w < 0), "the world touches the robot somewhere"
    got_s = self_collision_cost({"q": qt}, robot, chk, "q", 0.05, 2.0).cpu().numpy()
    ref_s = (np.asarray(colldist_from_sdf(selfd.astype(np.float32), 0.05)) * 2.0).reshape(-1)
    np.testing.assert_allclose(got_s, ref_s, atol=1e-5 * 2.0 * 2, rtol=1e-5)
