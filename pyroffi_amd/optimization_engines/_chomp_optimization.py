"""CHOMP trajectory optimisation on the GPU (one launch for many trajectories).

Python entry points with the reference's names and signatures
(``src/pyroffi/optimization_engines/_chomp_optimization.py``):

* ``ChompTrajOptConfig`` -- ``:31-90`` (same fields and defaults)
* ``chomp_trajopt``      -- ``:351-416``: pins the endpoints, runs CHOMP on
  every trajectory of the batch (``_optimize_single_traj``, ``:232-330``), ranks
  them by the cost at the maximum collision weight (``_eval_cost``,
  ``:333-348``) and returns ``(best_traj, costs, final_trajs)``

and ``chomp_trajopt_uneven`` for trajectories of different lengths and
endpoints in one launch.  The whole solve -- cost and adjoint gradient, the
gradient clip, the covariant direction (truncated CG on the second-difference
precision operator per joint), the six clamped trial steps, early stopping and
collision-weight continuation -- runs in ``pk_traj_chomp_f32``, one workgroup
per trajectory; the collision term is the hinge^2 over the same JAX self and
world distances as the L-BFGS solver's cost (``TrajModel`` with mode 0).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from ._lbfgs_trajopt_optimization import TrajModel, flatten_trajs


@dataclass(frozen=True)
class ChompTrajOptConfig:
    """Hyper-parameters for CHOMP trajectory optimisation (``:31-90``)."""

    n_iters: int = 100
    step_size: float = 0.05
    w_smooth: float = 1.0
    w_vel: float = 1.0  # unused, as in the reference (CHOMP's smoothness has no velocity term)
    w_acc: float = 0.5
    w_jerk: float = 0.1
    w_collision: float = 3.0
    w_collision_max: float = 50.0
    collision_penalty_scale: float = 1.05
    collision_margin: float = 0.01
    w_limits: float = 1.0
    use_covariant_update: bool = True
    smoothness_reg: float = 1e-3
    n_cg_iters: int = 24
    cg_tol: float = 1e-4
    grad_clip_norm: float = 10.0
    max_delta_per_step: float = 0.05
    early_stop_patience: int = 15
    min_cost_improve: float = 1e-5


def _config_struct(cfg: ChompTrajOptConfig):
    from ..kernels._lib import ChompConfig

    c = ChompConfig()
    c.n_iters, c.use_covariant_update = int(cfg.n_iters), int(bool(cfg.use_covariant_update))
    c.n_cg_iters, c.early_stop_patience = int(cfg.n_cg_iters), int(cfg.early_stop_patience)
    for f in ("step_size", "w_smooth", "w_acc", "w_jerk", "w_collision", "w_collision_max",
              "collision_penalty_scale", "w_limits", "smoothness_reg", "cg_tol", "grad_clip_norm",
              "max_delta_per_step", "min_cost_improve"):
        setattr(c, f, float(getattr(cfg, f)))
    return c


def chomp_run(model: TrajModel, x: torch.Tensor, start: torch.Tensor, max_len: int, cfg: ChompTrajOptConfig,
              trace: bool = False, block_threads: int = 0):
    """One ``pk_traj_chomp_f32`` launch: ``(final (W, A), costs (B,), trace (B, n_iters, 3) or None)``."""
    from ..kernels._lib import check, lib
    from ..kernels.ops import _ptr, _stream

    if model.mode != 0:
        raise ValueError("CHOMP's collision cost is the hinge^2 (TrajModel mode 0)")
    B = start.shape[0] - 1
    out = torch.empty_like(x)
    cost = torch.empty((B,), dtype=torch.float32, device=x.device)
    tr = torch.empty((B, max(cfg.n_iters, 0), 3), dtype=torch.float32, device=x.device) if trace else None
    c = _config_struct(cfg)
    st = lib().pk_traj_chomp_f32(_stream(), model.struct_ref(), _ptr(x), _ptr(start), B, x.shape[0], int(max_len),
                                 int(block_threads), ctypes.byref(c), _ptr(out), _ptr(cost), _ptr(tr))
    check(st, "pk_traj_chomp_f32")
    return out, cost, tr


def chomp_trajopt_uneven(trajs, robot, robot_coll, world_geoms, opt_cfg: ChompTrajOptConfig = ChompTrajOptConfig()):
    """Trajectories of different lengths/endpoints (rows 0 and -1 are the pinned endpoints).

    Returns ``(list of optimised (T_b, A) tensors, ranking costs (B,))``.
    """
    model = TrajModel(robot, robot_coll, world_geoms, opt_cfg.collision_margin, mode=0)
    x, start, max_len = flatten_trajs(trajs, model.device)
    out, cost, _ = chomp_run(model, x, start, max_len, opt_cfg)
    s = start.cpu().tolist()
    return [out[s[b]:s[b + 1]] for b in range(len(s) - 1)], cost


def chomp_trajopt(init_trajs, start, goal, robot, robot_coll, world_geoms,
                  opt_cfg: ChompTrajOptConfig = ChompTrajOptConfig(), *, use_cuda: bool = True):
    """``(best_traj (T, A), costs (B,), final_trajs (B, T, A))`` like the reference (``:351-416``)."""
    x = torch.as_tensor(np.asarray(init_trajs) if not isinstance(init_trajs, torch.Tensor) else init_trajs,
                        dtype=torch.float32)
    dev = torch.device("cuda", torch.cuda.current_device())
    x = x.to(dev).clone()
    B, T, A = x.shape
    x[:, 0] = torch.as_tensor(np.asarray(start), dtype=torch.float32, device=dev)
    x[:, -1] = torch.as_tensor(np.asarray(goal), dtype=torch.float32, device=dev)
    model = TrajModel(robot, robot_coll, world_geoms, opt_cfg.collision_margin, mode=0)
    xf, st, max_len = flatten_trajs(x, dev)
    out, cost, _ = chomp_run(model, xf, st, max_len, opt_cfg)
    final = out.reshape(B, T, A)
    idx = int(torch.argmin(cost))
    res = (final[idx], cost, final)
    if isinstance(init_trajs, torch.Tensor):
        return res
    return tuple(o.cpu().numpy() for o in res)
