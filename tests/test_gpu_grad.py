"""GPU parity: pose + hinge collision cost and its hand-written adjoint.

* cost vs the FP64 oracle (relative 1e-5 per configuration);
* dcost/dq vs FP64 central finite differences of the oracle cost (h = 1e-4),
  norm-wise relative error <= 1e-4 per configuration;
* the torch.autograd.Function scales by the cotangent and matches.
"""

import numpy as np
import pytest
import torch
from conftest import load_robot, random_cfgs

from pyroffi_amd.collision import RobotCollisionSpherized
from pyroffi_amd.costs import pose_hinge_cost

pytestmark = pytest.mark.gpu

EE = {"panda": "panda_hand", "fetch": "gripper_link", "mimic_arm": "finger_c", "baxter": "left_gripper_base"}


def _setup(oracle_mod, name, B, M, seed):
    robot, path = load_robot(name)
    coll = RobotCollisionSpherized.from_urdf(path)
    t = oracle_mod.Tables.of(robot)
    link = robot.links.names.index(EE[name])
    rng = np.random.default_rng(seed)
    world = np.concatenate([rng.uniform(-1, 1, (M, 3)), rng.uniform(0.02, 0.15, (M, 1))], 1).astype(np.float32)
    q = random_cfgs(robot, B, seed=seed + 1).astype(np.float32)
    tgt = oracle_mod.fk_links(t, random_cfgs(robot, B, seed=seed + 2))[:, link]
    return robot, coll, t, link, world, q, tgt


@pytest.mark.parametrize("name,M", [("panda", 512), ("panda", 37), ("fetch", 256), ("mimic_arm", 100),
                                    ("baxter", 64)])
def test_cost_and_gradient_vs_fd(oracle_mod, name, M):
    robot, coll, t, link, world, q, tgt = _setup(oracle_mod, name, 48, M, seed=M)
    margin = 0.05
    qt = torch.as_tensor(q, device="cuda").requires_grad_(True)
    cost = pose_hinge_cost(robot, coll, qt, tgt.astype(np.float32), world, margin=margin, ee_link=link)
    (grad,) = torch.autograd.grad(cost.sum(), qt)
    sp = oracle_mod.padded_spheres(coll)
    ocost = oracle_mod.pose_hinge_cost(t, q.astype(np.float64), sp, world.astype(np.float64), margin, link, tgt)
    np.testing.assert_allclose(cost.detach().cpu().numpy(), ocost, rtol=1e-5, atol=1e-6)
    fd = oracle_mod.pose_hinge_fd_grad(t, q.astype(np.float64), sp, world.astype(np.float64), margin, link, tgt)
    g = grad.cpu().numpy()
    rel = np.linalg.norm(g - fd, axis=1) / np.maximum(np.linalg.norm(fd, axis=1), 1e-6)
    assert rel.max() <= 1e-4, (rel.max(), np.argmax(rel))


def test_reach_boundary_vs_fd(oracle_mod):
    """Obstacles 1-10 mm inside and outside the hinge reach r + R + margin of a
    robot sphere: the kernel's expanded reach test must hand every active pair
    to the exact hinge and skip only inactive ones."""
    robot, coll, t, link, _, q, tgt = _setup(oracle_mod, "panda", 16, 1, seed=77)
    margin = 0.05
    sp = oracle_mod.padded_spheres(coll)  # (N, S, 4)
    lp = oracle_mod.fk_links(t, q.astype(np.float64))
    li, si = np.nonzero(sp[:, :, 3] > 0)
    rng = np.random.default_rng(78)
    M = 256
    b, k = rng.integers(0, 16, M), rng.integers(0, len(li), M)
    qw, qv = lp[b, li[k], :1], lp[b, li[k], 1:4]
    v = sp[li[k], si[k], :3]
    tt = 2.0 * np.cross(qv, v)
    centre = v + qw * tt + np.cross(qv, tt) + lp[b, li[k], 4:]
    R = rng.uniform(0.02, 0.1, M)
    delta = np.geomspace(1e-3, 1e-2, M) * np.where(np.arange(M) % 2 == 0, 1.0, -1.0)
    u = rng.normal(size=(M, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    world = np.concatenate([centre + u * (sp[li[k], si[k], 3] + R + margin + delta)[:, None], R[:, None]], 1)
    world = world.astype(np.float32)
    qt = torch.as_tensor(q, device="cuda").requires_grad_(True)
    cost = pose_hinge_cost(robot, coll, qt, tgt.astype(np.float32), world, margin=margin, ee_link=-1)
    (grad,) = torch.autograd.grad(cost.sum(), qt)
    ocost = oracle_mod.pose_hinge_cost(t, q.astype(np.float64), sp, world.astype(np.float64), margin, -1, tgt)
    assert np.all(ocost > 0)
    np.testing.assert_allclose(cost.detach().cpu().numpy(), ocost, rtol=1e-5, atol=1e-9)
    fd = oracle_mod.pose_hinge_fd_grad(t, q.astype(np.float64), sp, world.astype(np.float64), margin, -1, tgt)
    g = grad.cpu().numpy()
    rel = np.linalg.norm(g - fd, axis=1) / np.maximum(np.linalg.norm(fd, axis=1), 1e-6)
    assert rel.max() <= 1e-4, (rel.max(), np.argmax(rel))


def test_cotangent_scaling_and_no_grad_path(oracle_mod):
    robot, coll, t, link, world, q, tgt = _setup(oracle_mod, "panda", 33, 512, seed=5)
    qt = torch.as_tensor(q, device="cuda").requires_grad_(True)
    cost = pose_hinge_cost(robot, coll, qt, tgt.astype(np.float32), world, ee_link=link)
    w = torch.linspace(-2, 3, 33, device="cuda")
    (g1,) = torch.autograd.grad((cost * w).sum(), qt)
    qt2 = torch.as_tensor(q, device="cuda").requires_grad_(True)
    (g0,) = torch.autograd.grad(pose_hinge_cost(robot, coll, qt2, tgt.astype(np.float32), world, ee_link=link).sum(), qt2)
    torch.testing.assert_close(g1, w[:, None] * g0)
    with torch.no_grad():
        c2 = pose_hinge_cost(robot, coll, torch.as_tensor(q, device="cuda"), tgt.astype(np.float32), world, ee_link=link)
    torch.testing.assert_close(c2, cost.detach())


def test_collision_only_gradient_is_zero_when_clear(oracle_mod):
    robot, coll, t, link, world, q, tgt = _setup(oracle_mod, "panda", 8, 16, seed=9)
    far = np.array([[30.0, 30.0, 30.0, 0.1]], np.float32)
    qt = torch.as_tensor(q, device="cuda").requires_grad_(True)
    cost = pose_hinge_cost(robot, coll, qt, tgt.astype(np.float32), far, ee_link=-1)
    (g,) = torch.autograd.grad(cost.sum(), qt)
    assert float(cost.abs().max()) == 0.0 and float(g.abs().max()) == 0.0


@pytest.mark.parametrize("name,B", [("panda", 200), ("mimic_arm", 65), ("fetch", 64), ("baxter", 33),
                                    ("ur5", 1)])
def test_fk_adjoint_vs_fd(oracle_mod, name, B):
    """The FK adjoint (J^T of the link poses on a random cotangent, through
    Robot.forward_kinematics' autograd) against FP64 central differences of the
    oracle FK (h = 1e-4): norm-wise relative error <= 1e-4 per configuration.
    The mimic arm covers mimic multipliers/offsets, a prismatic and a
    continuous joint; partial last wave at B = 65 and 33."""
    robot, _ = load_robot(name)
    t = oracle_mod.Tables.of(robot)
    q = random_cfgs(robot, B, seed=B + 7)
    rng = np.random.default_rng(B)
    L, A = robot.links.num_links, robot.joints.num_actuated_joints
    G = rng.standard_normal((B, L, 7))
    qt = torch.as_tensor(q.astype(np.float32), device="cuda").requires_grad_(True)
    poses = robot.forward_kinematics(qt)
    (poses * torch.as_tensor(G, dtype=torch.float32, device="cuda")).sum().backward()
    got = qt.grad.cpu().numpy().astype(np.float64)
    h = 1e-4
    fd = np.empty((B, A))
    q32 = q.astype(np.float32).astype(np.float64)  # differentiate at the point the GPU saw
    for a in range(A):
        e = np.zeros(A)
        e[a] = h
        fd[:, a] = np.sum(G * (oracle_mod.fk_links(t, q32 + e) - oracle_mod.fk_links(t, q32 - e)), axis=(1, 2)) / (2 * h)
    err = np.linalg.norm(got - fd, axis=1) / np.maximum(np.linalg.norm(fd, axis=1), 1e-6)
    print(f"{name}: max relative error {err.max():.2e}", flush=True)
    assert err.max() <= 1e-4, err.max()
    # the forward is the plain FK
    lp = oracle_mod.fk_links(t, q32)
    assert np.abs(poses.detach().cpu().numpy() - lp).max() < 1e-5


def test_fk_adjoint_root_link_and_zero_cotangent():
    robot, _ = load_robot("panda")
    q = torch.as_tensor(random_cfgs(robot, 70, seed=3).astype(np.float32), device="cuda").requires_grad_(True)
    poses = robot.forward_kinematics(q)
    g = torch.zeros_like(poses)
    g[:, 0] = 1.0  # the root link is the constant identity: no gradient
    (dq,) = torch.autograd.grad(poses, q, g)
    assert float(dq.abs().max()) == 0.0
