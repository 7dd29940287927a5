"""The no-allocation entry point pk_traj_lbfgs_noalloc_f32 against the
workspace entry point pk_traj_lbfgs_f32, through the C ABI.

Both run the same kernel on the same inputs, so best paths and best costs must
be equal bit for bit: on two calls in a row, the second with a smaller input
(the static scratch keeps the first call's contents), and for a solve split
over two calls that resumes from the scratch.  Sizes beyond the scratch are
rejected on the host with a clear error.
"""

import ctypes

import numpy as np
import pytest
import torch
from conftest import load_robot, random_cfgs

from pyroffi_amd.collision import RobotCollisionSpherized
from pyroffi_amd.kernels._lib import lib
from pyroffi_amd.kernels.ops import _ptr, _stream
from pyroffi_amd.optimization_engines._lbfgs_trajopt_optimization import (
    BatchedLbfgsTrajOpt,
    LbfgsTrajOptConfig,
    TrajModel,
)

pytestmark = pytest.mark.gpu


def _model(seed):
    robot, path = load_robot("panda")
    coll = RobotCollisionSpherized.from_urdf(path)
    rng = np.random.default_rng(seed)
    world = np.concatenate([rng.uniform(-1, 1, (256, 3)), rng.uniform(0.02, 0.1, (256, 1))], 1).astype(np.float32)
    return robot, TrajModel(robot, coll, world, 0.01, 1)


def _trajs(robot, lens, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i, T in enumerate(lens):
        q0, q1 = random_cfgs(robot, 2, seed=seed * 100 + i)
        s = np.linspace(0, 1, T)[:, None]
        out.append(((1 - s) * q0 + s * q1 + rng.normal(0, 0.02, (T, 7))).astype(np.float32))
    return out


def _noalloc(opt, it0, it1, best, cost):
    st = lib().pk_traj_lbfgs_noalloc_f32(_stream(), opt.model.struct_ref(), _ptr(opt.x0), _ptr(opt.start), opt.B,
                                         opt.x0.shape[0], opt.max_len, 0, _ptr(None), 0, ctypes.byref(opt._cfg),
                                         it0, it1, _ptr(best), _ptr(cost))
    return st


def _check_call(model, trajs, cfg, split=None):
    ref = BatchedLbfgsTrajOpt(model, trajs, cfg, split_classes=False)
    rbest, rcost, _ = ref.solve()
    best = torch.full_like(ref.best, float("nan"))
    cost = torch.full_like(ref.best_cost, float("nan"))
    cuts = [0, cfg.n_iters] if split is None else [0, split, cfg.n_iters]
    for a, b in zip(cuts[:-1], cuts[1:]):
        st = _noalloc(ref, a, b, best, cost)
        assert st == 0, lib().pk_last_error()
    torch.cuda.synchronize()
    torch.testing.assert_close(best, rbest, rtol=0, atol=0)
    torch.testing.assert_close(cost, rcost, rtol=0, atol=0)
    assert bool(torch.isfinite(cost).all())


def test_noalloc_equals_workspace_entry_point_twice():
    robot, model = _model(1)
    cfg = LbfgsTrajOptConfig(n_iters=25)
    _check_call(model, _trajs(robot, [48, 64, 33, 64, 20, 57], seed=2), cfg)
    # second call, smaller input, same scratch
    _check_call(model, _trajs(robot, [16, 40], seed=3), cfg)
    # a solve split over two calls resumes from the scratch
    _check_call(model, _trajs(robot, [30, 64, 12], seed=4), cfg, split=11)


def test_noalloc_rejects_oversize_state():
    robot, model = _model(5)
    cfg = LbfgsTrajOptConfig(n_iters=5)
    opt = BatchedLbfgsTrajOpt(model, _trajs(robot, [16], seed=6), cfg, split_classes=False)
    best, cost = torch.empty_like(opt.best), torch.empty_like(opt.best_cost)
    # claim 50,000 waypoints of 7 joints: (3 + 2 * 6) * 350,000 floats > 4 Mi floats
    st = lib().pk_traj_lbfgs_noalloc_f32(_stream(), model.struct_ref(), _ptr(opt.x0), _ptr(opt.start), 1, 50000,
                                         16, 0, _ptr(None), 0, ctypes.byref(opt._cfg), 0, 5, _ptr(best), _ptr(cost))
    assert st == 1 and b"scratch" in lib().pk_last_error()
