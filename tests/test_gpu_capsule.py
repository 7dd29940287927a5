"""Non-sphere geometry on the GPU against the FP64 oracle.

* capsule robots -- ``pk_capsules_place_f32`` (link-frame capsules moved by
  link poses), ``pk_collision_world_capsule_f32`` over the four obstacle kinds
  and ``pk_collision_self_capsule_f32`` over pair lists, each against
  ``oracle.capsule_dist`` on the same capsules;
* spherized robots in mixed worlds -- ``pk_sphere_world_links_f32`` and the
  checker (fine, minimum, coarse-first guard) against ``oracle.obstacle_dist``
  on the oracle's sphere centres;
* the LM penalty over capsules, boxes and half-spaces -- ``pk_ls_ik_f32`` at
  ``max_iter = 0`` (the error is the penalty plus the pose error) and after a
  short solve, against ``oracle.ls_ik`` with the same world.
"""

import numpy as np
import pytest
import torch
import worlds
from conftest import load_robot, random_cfgs

from pyroffi_amd.collision import (Box, Capsule, CUDARobotCollisionChecker, HalfSpace, RobotCollision,
                                   RobotCollisionSpherized, Sphere)
from pyroffi_amd.kernels import ops
from pyroffi_amd.optimization_engines._ls_ik import ee_tables, ls_ik_run, prepare_ls_collision_buffers

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _quat_rotate(q, v):
    w, u = q[..., :1], q[..., 1:4]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def _capsule_geom(wc):
    """``Capsule`` container for packed ``[a, b, r]`` rows."""
    from pyroffi_amd._se3 import SE3
    from pyroffi_amd.collision._geometry import _frame_with_z

    a, b = wc[:, :3].astype(np.float64), wc[:, 3:6].astype(np.float64)
    m = np.zeros((len(wc), 4, 4))
    m[:, :3, :3] = _frame_with_z(b - a)
    m[:, :3, 3] = 0.5 * (a + b)
    m[:, 3, 3] = 1.0
    return Capsule(SE3.from_matrix(m), wc[:, 6], np.linalg.norm(b - a, axis=1))


def _random_capsules(rng, B, N):
    a = rng.uniform(-0.8, 0.8, (B, N, 3))
    b = a + rng.normal(0, 0.15, (B, N, 3))
    b[:, ::5] = a[:, ::5]  # some degenerate (sphere-like) capsules
    r = rng.uniform(0.01, 0.1, (B, N, 1))
    return np.concatenate([a, b, r], -1).astype(np.float32)


@pytest.mark.parametrize("kinds", [(9, 0, 0, 0), (0, 7, 0, 0), (0, 0, 6, 0), (0, 0, 0, 3), (11, 5, 7, 2)])
@pytest.mark.parametrize("B,N", [(1, 1), (33, 13), (200, 7)])
def test_capsule_world_kernel(oracle_mod, kinds, B, N):
    rng = np.random.default_rng(B * 7 + N + sum(kinds))
    caps = _random_capsules(rng, B, N)
    ws, wc, wb, wh = worlds.random_world(rng, *kinds)
    soa = np.ascontiguousarray(np.transpose(caps, (2, 0, 1)))
    got = ops.collision_world_capsule(soa, ws, wc, wb, wh).cpu().numpy()
    ref = oracle_mod.capsule_dist(caps.reshape(-1, 7), ws, wc, wb, wh).reshape(B, N, -1)
    np.testing.assert_allclose(got, ref, atol=TOL, rtol=0)


@pytest.mark.parametrize("B,N,P", [(1, 2, 1), (45, 13, 40), (300, 9, 36)])
def test_capsule_self_kernel(oracle_mod, B, N, P):
    rng = np.random.default_rng(B + N + P)
    caps = _random_capsules(rng, B, N)
    ii, jj = np.tril_indices(N, k=-1)
    sel = rng.choice(len(ii), size=min(P, len(ii)), replace=False)
    pi, pj = ii[sel].astype(np.int32), jj[sel].astype(np.int32)
    got = ops.collision_self_capsule(caps, pi, pj).cpu().numpy()
    ref = np.stack([np.array([oracle_mod.capsule_dist(caps[b, i][None], capsules=caps[b, j][None])[0, 0]
                              for i, j in zip(pi, pj)]) for b in range(B)])
    np.testing.assert_allclose(got, ref, atol=TOL, rtol=0)


def _oracle_capsules(oracle_mod, robot, model, q):
    lp = oracle_mod.fk_links(oracle_mod.Tables.of(robot), q.astype(np.float64))  # (B, N, 7)
    loc = model.coll.as_array().astype(np.float64)  # (N, 7)
    a = _quat_rotate(lp[..., :4], loc[None, :, 0:3]) + lp[..., 4:]
    b = _quat_rotate(lp[..., :4], loc[None, :, 3:6]) + lp[..., 4:]
    return np.concatenate([a, b, np.broadcast_to(loc[None, :, 6:7], a.shape[:-1] + (1,))], -1)


@pytest.mark.parametrize("name", ["panda", "fetch", "baxter"])
def test_capsule_robot_model(oracle_mod, name):
    robot, path = load_robot(name)
    model = RobotCollision.from_urdf(path)
    q = random_cfgs(robot, 50, seed=3).astype(np.float32)
    want = _oracle_capsules(oracle_mod, robot, model, q)
    got = model.placed_capsules(robot, q).cpu().numpy()
    np.testing.assert_allclose(got, want, atol=TOL, rtol=0)
    soa, aos = ops.capsules_place(torch.as_tensor(
        oracle_mod.fk_links(oracle_mod.Tables.of(robot), q.astype(np.float64)).astype(np.float32), device="cuda"),
        model.coll.as_array())
    np.testing.assert_array_equal(np.transpose(soa.cpu().numpy(), (1, 2, 0)), aos.cpu().numpy())
    ws, wc, wb, wh = worlds.random_world(np.random.default_rng(4), 12, 5, 6, 2)
    world = [Sphere.from_center_and_radius(ws[:, :3], ws[:, 3]),
             _capsule_geom(wc), Box.from_center_and_half_lengths(wb[:, :3], wb[:, 12:15]),
             HalfSpace.from_point_and_normal(wh[:, 3:], wh[:, :3])]
    from pyroffi_amd.collision import world_arrays

    arrs = world_arrays(world)
    ref_w = oracle_mod.capsule_dist(want.reshape(-1, 7), arrs["spheres"], arrs["capsules"], arrs["boxes"],
                                    arrs["halfspaces"]).reshape(len(q), model.num_links, -1)
    got_w = model.compute_world_collision_distance(robot, q, world)
    np.testing.assert_allclose(got_w, ref_w, atol=TOL, rtol=0)
    ref_s = np.array([[oracle_mod.capsule_dist(want[b, i][None], capsules=want[b, j][None])[0, 0]
                       for i, j in zip(model.active_idx_i, model.active_idx_j)] for b in range(len(q))])
    got_s = model.compute_self_collision_distance(robot, q)
    np.testing.assert_allclose(got_s, ref_s, atol=TOL, rtol=0)
    chk = CUDARobotCollisionChecker(model)
    np.testing.assert_allclose(chk.compute_world_collision_distance(robot, q, world), ref_w, atol=TOL, rtol=0)
    np.testing.assert_allclose(chk.compute_min_world_distance(robot, q, world), ref_w.min(axis=(1, 2)),
                               atol=TOL, rtol=0)
    np.testing.assert_allclose(chk.compute_self_collision_distance(robot, q), ref_s, atol=TOL, rtol=0)


def _oracle_sphere_world(oracle_mod, robot, coll, q, arrs):
    lp = oracle_mod.fk_links(oracle_mod.Tables.of(robot), q.astype(np.float64))
    sp = oracle_mod.padded_spheres(coll)  # (L, S, 4)
    c = _quat_rotate(lp[:, :, None, :4], sp[None, :, :, :3]) + lp[:, :, None, 4:]  # (B, L, S, 3)
    r = np.broadcast_to(sp[None, :, :, 3], c.shape[:-1])
    d = oracle_mod.obstacle_dist(c.reshape(-1, 3), r.reshape(-1), arrs["spheres"], arrs["capsules"], arrs["boxes"],
                                 arrs["halfspaces"]).reshape(*c.shape[:-1], -1)
    d = np.where(r[..., None] < 0, np.inf, d).min(axis=2)
    return np.minimum(d, 1e9)


@pytest.mark.parametrize("kinds", [(0, 6, 0, 0), (0, 0, 5, 0), (0, 0, 0, 2), (20, 6, 8, 2)])
def test_sphere_robot_mixed_world(oracle_mod, kinds):
    robot, path = load_robot("panda")
    coll = RobotCollisionSpherized.from_urdf(path)
    q = random_cfgs(robot, 41, seed=sum(kinds)).astype(np.float32)
    ws, wc, wb, wh = worlds.random_world(np.random.default_rng(sum(kinds) + 1), *kinds)
    arrs = {"spheres": ws, "capsules": wc, "boxes": wb, "halfspaces": wh}
    ref = _oracle_sphere_world(oracle_mod, robot, coll, q, arrs)
    t = robot.device_tables()
    poses = ops.fk_links(torch.as_tensor(q, device="cuda"), *t.robot_buffers(), t.link_parent_joint)
    got, got_min = ops.sphere_world_links(poses, coll.device_tables(poses.device)["spheres_padded"], arrs)
    np.testing.assert_allclose(got.cpu().numpy(), ref, atol=TOL, rtol=0)
    np.testing.assert_allclose(got_min.cpu().numpy(), ref.min(axis=(1, 2)), atol=TOL, rtol=0)
    world = [Sphere.from_center_and_radius(ws[:, :3], ws[:, 3]), _capsule_geom(wc),
             Box.from_center_and_half_lengths(wb[:, :3], wb[:, 12:15]),
             HalfSpace.from_point_and_normal(wh[:, 3:], wh[:, :3])]
    world = [g for g in world if g.get_batch_axes()[0]]
    if not world:
        return
    from pyroffi_amd.collision import world_arrays

    a2 = world_arrays(world)
    ref2 = _oracle_sphere_world(oracle_mod, robot, coll, q, a2)
    chk = CUDARobotCollisionChecker(coll)
    np.testing.assert_allclose(chk.compute_world_collision_distance(robot, q, world), ref2, atol=TOL, rtol=0)
    np.testing.assert_allclose(coll.compute_world_collision_distance(robot, q, world, semantics="cuda"), ref2,
                               atol=TOL, rtol=0)


def test_coarse_first_guard_mixed_world(oracle_mod):
    """Coarse-first: all ones when the coarse model clears every obstacle, the fine
    distances otherwise (``_cuda_collision.py:544-593``)."""
    robot, path = load_robot("panda")
    fine = RobotCollisionSpherized.from_urdf(path)
    coarse = RobotCollision.from_urdf(path)
    chk = CUDARobotCollisionChecker(fine, coarse_inner=coarse)
    q = random_cfgs(robot, 16, seed=9).astype(np.float32)
    # a half-space is solid on the side opposite its normal
    far = [HalfSpace.from_point_and_normal(np.array([[0.0, 0.0, -3.0]]), np.array([[0.0, 0.0, 1.0]])),
           Box.from_center_and_half_lengths(np.array([[5.0, 5.0, 5.0]]), np.array([[0.1, 0.1, 0.1]]))]
    out = chk.compute_world_collision_distance(robot, q, far)
    np.testing.assert_array_equal(out, np.ones((16, fine.num_links, 2), np.float32))
    near = [HalfSpace.from_point_and_normal(np.array([[0.0, 0.0, 0.3]]), np.array([[0.0, 0.0, 1.0]])),
            Box.from_center_and_half_lengths(np.array([[5.0, 5.0, 5.0]]), np.array([[0.1, 0.1, 0.1]]))]
    from pyroffi_amd.collision import world_arrays

    ref = _oracle_sphere_world(oracle_mod, robot, fine, q, world_arrays(near))
    out = chk.compute_world_collision_distance(robot, q, near)
    np.testing.assert_allclose(out, ref, atol=TOL, rtol=0)


@pytest.mark.parametrize("kinds", [(0, 40, 0, 0), (0, 0, 40, 0), (0, 0, 0, 2), (10, 10, 10, 1)])
def test_ls_penalty_other_kinds(oracle_mod, kinds):
    robot, path = load_robot("panda")
    coll = RobotCollisionSpherized.from_urdf(path)
    rng = np.random.default_rng(sum(kinds) + 2)
    ws, wc, wb, wh = worlds.random_world(rng, *kinds, extent=0.6)
    arrs = {"spheres": ws, "capsules": wc, "boxes": wb, "halfspaces": wh}
    rs, rj, _, _ = prepare_ls_collision_buffers(robot, coll, np.zeros((1, 4), np.float32))
    ee = robot.links.names.index("panda_hand")
    tj, masks = ee_tables(robot, (ee,))
    t = oracle_mod.Tables.of(robot)
    P, S = 4, 32
    q = random_cfgs(robot, P * S, seed=5).astype(np.float32).reshape(P, S, -1)
    tgt = oracle_mod.fk_links(t, random_cfgs(robot, P, seed=6))[:, ee][:, None]
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    w, margin = 1e4, 0.05
    for iters in (0, 8):
        ref_q, ref_e = oracle_mod.ls_ik(t, q, tj, masks, tgt, lo, hi, max_iter=iters, lambda_init=1e-2,
                                        robot_spheres=rs, sphere_joint=rj, world=ws, world_capsules=wc,
                                        world_boxes=wb, world_halfspaces=wh, enable_collision=True,
                                        collision_weight=w, collision_margin=margin)
        got_q, got_e = ls_ik_run(robot, q, tgt.astype(np.float32), (ee,), max_iter=iters, lambda_init=1e-2,
                                 collision=(rs, rj, arrs, w, margin))
        got_q, got_e = got_q.cpu().numpy(), got_e.cpu().numpy()
        if iters == 0:
            assert np.count_nonzero(ref_e > 1e-2) > 0
            np.testing.assert_allclose(got_e, ref_e, rtol=2e-4, atol=1e-4)
        else:
            # float32 vs float64 LM: most seeds land on the same iterate
            close = np.max(np.abs(got_q - ref_q), axis=-1) < 1e-3
            assert close.mean() >= 0.8, close.mean()
