"""GPU parity of SCO trajectory optimisation (``pk_traj_sco_linearize_f32``,
``pk_traj_sco_inner_f32``) against the FP64 oracle (``oracle.sco_linearize``,
``oracle.traj_sco``: a restatement of ``_sco_optimization.py:224-503``).

* linearisation: per waypoint and group (self pairs, two world groups) the
  smooth minimum within 2e-5 m and its gradient within 1e-4 of the largest
  entry, on 1024 world spheres;
* one SCO round (linearise + 30 inner L-BFGS steps) lands on the oracle's path
  (bound below); ten rounds give final costs that agree as a batch;
* a captured graph replays to the eager result; ``sco_trajopt`` pins the
  endpoints and ranks by the oracle's nonlinear cost of its own paths.
"""

import dataclasses

import numpy as np
import pytest
import torch
from conftest import load_robot, random_cfgs

from pyroffi_amd.collision import RobotCollisionSpherized, Sphere
from pyroffi_amd.optimization_engines import ScoTrajOptConfig, sco_trajopt
from pyroffi_amd.optimization_engines._lbfgs_trajopt_optimization import flatten_trajs
from pyroffi_amd.optimization_engines._sco_optimization import BatchedSco

pytestmark = pytest.mark.gpu


def _setup(oracle_mod, seed, M=1024, split=600):
    robot, path = load_robot("panda")
    coll = RobotCollisionSpherized.from_urdf(path)
    rng = np.random.default_rng(seed)
    world = np.concatenate([rng.uniform(-1, 1, (M, 3)), rng.uniform(0.02, 0.1, (M, 1))], 1).astype(np.float32)
    geoms = (Sphere.from_center_and_radius(world[:split, :3], world[:split, 3]),
             Sphere.from_center_and_radius(world[split:, :3], world[split:, 3]))
    gstart = np.array([0, split, M], np.int32)
    t = oracle_mod.Tables.of(robot)
    om = oracle_mod.TrajModel(t, oracle_mod.padded_spheres(coll), coll.active_idx_i, coll.active_idx_j,
                              world.astype(np.float64), 0.01, 1)
    return robot, coll, geoms, gstart, om


def _trajs(robot, lens, seed, noise=0.02):
    rng = np.random.default_rng(seed)
    out = []
    for i, T in enumerate(lens):
        q0, q1 = random_cfgs(robot, 2, seed=seed * 100 + i)
        s = np.linspace(0, 1, T)[:, None]
        out.append(((1 - s) * q0 + s * q1 + rng.normal(0, noise, (T, len(q0)))).astype(np.float32))
    return out


def _oracle_kw(cfg):
    return {f.name: getattr(cfg, f.name) for f in dataclasses.fields(cfg)}


def test_sco_linearize_vs_oracle(oracle_mod):
    from pyroffi_amd.kernels._lib import check, lib
    from pyroffi_amd.kernels.ops import _ptr, _stream

    robot, coll, geoms, gstart, om = _setup(oracle_mod, seed=1)
    trajs = _trajs(robot, [17, 32, 64, 5], seed=2)
    opt = BatchedSco(robot, coll, geoms, trajs, ScoTrajOptConfig())
    W = opt.x0.shape[0]
    st = lib().pk_traj_sco_linearize_f32(_stream(), opt.model.struct_ref(), _ptr(opt.x0), _ptr(opt.start), opt.B, W,
                                         opt.max_len, 0, _ptr(opt.gstart), opt.G - 1, 0.05, _ptr(opt.d), _ptr(opt.J))
    check(st, "pk_traj_sco_linearize_f32")
    d, J = opt.d.cpu().numpy(), opt.J.cpu().numpy()
    rd, rJ = oracle_mod.sco_linearize(om, np.concatenate(trajs).astype(np.float64), gstart, 0.05)
    print("max |d - fp64|", np.abs(d - rd).max(), " max |J - fp64|", np.abs(J - rJ).max(), " max |J|",
          np.abs(rJ).max(), flush=True)
    np.testing.assert_allclose(d, rd, atol=2e-5, rtol=0)
    np.testing.assert_allclose(J, rJ, atol=1e-4 * np.abs(rJ).max(), rtol=0)


def test_sco_one_round_and_ten_rounds(oracle_mod):
    robot, coll, geoms, gstart, om = _setup(oracle_mod, seed=3)
    trajs = _trajs(robot, [32, 64, 128, 32, 64], seed=4)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    cfg1 = ScoTrajOptConfig(n_outer_iters=1)
    opt = BatchedSco(robot, coll, geoms, trajs, cfg1)
    x1, c1 = opt.solve()
    x1 = x1.cpu().numpy()
    s = opt.start.cpu().numpy()
    ro, rc, _ = oracle_mod.traj_sco(om, [t.astype(np.float64) for t in trajs], lo, hi, gstart, **_oracle_kw(cfg1))
    x0 = np.concatenate(trajs)
    for b in range(len(trajs)):
        sl = slice(s[b], s[b + 1])
        err = np.abs(x1[sl] - ro[sl]).max()
        moved = np.abs(ro[sl] - x0[sl]).max()
        print(f"one round traj {b}: max|gpu - fp64| {err:.2e}, moved {moved:.3f}", flush=True)
        assert err <= max(1e-4, 0.01 * moved), (b, err, moved)
    np.testing.assert_allclose(c1.cpu().numpy(), rc, rtol=1e-3)
    cfg = ScoTrajOptConfig()
    opt = BatchedSco(robot, coll, geoms, trajs, cfg)
    _, c = opt.solve()
    _, rc, _ = oracle_mod.traj_sco(om, [t.astype(np.float64) for t in trajs], lo, hi, gstart, **_oracle_kw(cfg))
    ratio = c.cpu().numpy() / rc
    print("ten rounds, cost ratio gpu / fp64:", np.round(ratio, 4), flush=True)
    assert 0.9 <= np.median(ratio) <= 1.1


def test_sco_graph_and_public_api(oracle_mod):
    robot, coll, geoms, gstart, om = _setup(oracle_mod, seed=5)
    trajs = _trajs(robot, [40, 40, 40], seed=6)
    cfg = ScoTrajOptConfig(n_outer_iters=3)
    opt = BatchedSco(robot, coll, geoms, trajs, cfg)
    xe, ce = opt.solve()
    xe, ce = xe.clone(), ce.clone()
    opt.capture()
    for _ in range(3):
        xg, cg = opt.solve()
        torch.cuda.synchronize()
        assert torch.equal(xg, xe) and torch.equal(cg, ce)
        opt.x.zero_()  # the replay recomputes everything from x0
    init = np.stack(trajs)
    q_s, q_g = init[0, 0] + 0.01, init[0, -1] - 0.01
    best, costs, final = sco_trajopt(init, q_s, q_g, robot, coll, geoms, cfg)
    assert best.shape == (40, 7) and costs.shape == (3,) and final.shape == (3, 40, 7)
    np.testing.assert_allclose(final[:, 0], np.broadcast_to(q_s, (3, 7)), atol=1e-6)
    np.testing.assert_allclose(final[:, -1], np.broadcast_to(q_g, (3, 7)), atol=1e-6)
    np.testing.assert_array_equal(best, final[int(np.argmin(costs))])
    c, _ = oracle_mod.traj_cost(om, [f.astype(np.float64) for f in final], robot.joints.lower_limits,
                                robot.joints.upper_limits, (1.0, 1.0, 0.5, 0.1, 1.0, cfg.w_collision_max), grad=False)
    np.testing.assert_allclose(costs, c @ [1.0, 1.0, cfg.w_collision_max], rtol=1e-4)
