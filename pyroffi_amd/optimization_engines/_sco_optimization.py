"""Sequential convex optimisation (SCO) of trajectories on the GPU.

Python entry points with the reference's names and signatures
(``src/pyroffi/optimization_engines/_sco_optimization.py``):

* ``ScoTrajOptConfig`` (alias ``TrajOptConfig``) -- ``:49-110`` (same fields
  and defaults)
* ``sco_trajopt``     -- ``:450-548`` (``_sco_trajopt_jax`` and the entry point): pins the endpoints, runs
  ``n_outer_iters`` rounds of (linearise the collision groups at the current
  paths, solve the convex inner problem of every path with L-BFGS, raise the
  collision weight), ranks by the full nonlinear cost at the maximum weight
  (``_eval_cost``, ``:421-447``) and returns ``(best_traj, costs, final_trajs)``
* ``make_init_trajs`` -- ``:551-564`` (numpy's PCG64 in place of ``jax.random``)

A round is two launches: ``pk_traj_sco_linearize_f32`` (per waypoint and
group, the smooth minimum of the JAX distances and its gradient, FK and
sphere placement in the kernel) and ``pk_traj_sco_inner_f32`` (per
trajectory, the whole inner L-BFGS in LDS).  ``BatchedSco`` records every
launch of a solve into one HIP graph.  World groups: each geometry of the
``world_geoms`` tuple is one group, as in the reference.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from ._lbfgs_trajopt_optimization import TrajModel, flatten_trajs, traj_cost


@dataclass(frozen=True)
class ScoTrajOptConfig:
    """Hyper-parameters for the SCO trajectory optimiser (``:49-110``)."""

    n_outer_iters: int = 10
    n_inner_iters: int = 30
    m_lbfgs: int = 6
    w_smooth: float = 1.0
    w_vel: float = 1.0
    w_acc: float = 0.5
    w_jerk: float = 0.1
    w_collision: float = 1.0
    w_collision_max: float = 100.0
    penalty_scale: float = 3.0
    collision_margin: float = 0.01
    w_trust: float = 0.5
    w_limits: float = 1.0
    smooth_min_temperature: float = 0.05


TrajOptConfig = ScoTrajOptConfig


def _config_struct(cfg: ScoTrajOptConfig):
    from ..kernels._lib import ScoConfig

    c = ScoConfig()
    for f in ("n_outer_iters", "n_inner_iters", "m_lbfgs"):
        setattr(c, f, int(getattr(cfg, f)))
    for f in ("w_smooth", "w_vel", "w_acc", "w_jerk", "w_collision", "w_collision_max", "penalty_scale",
              "collision_margin", "w_trust", "w_limits", "smooth_min_temperature"):
        setattr(c, f, float(getattr(cfg, f)))
    return c


def world_groups(world_geoms) -> tuple[np.ndarray, np.ndarray]:
    """``world_geoms`` (a tuple of sphere geometries, one group each, or one geometry / ``(M, 4)``
    array) -> the concatenated ``(M, 4)`` spheres and the CSR group starts ``(n_groups + 1)``."""
    from ..collision._hip_collision import world_sphere_array

    items = list(world_geoms) if isinstance(world_geoms, (tuple, list)) else [world_geoms]
    parts = [world_sphere_array(g) for g in items]
    start = np.zeros(len(parts) + 1, np.int32)
    start[1:] = np.cumsum([len(p) for p in parts])
    ws = np.concatenate(parts, 0) if parts else np.zeros((0, 4), np.float32)
    return ws.astype(np.float32), start


class BatchedSco:
    """Many trajectories (any lengths) through SCO, all rounds on one stream, graph-capturable."""

    def __init__(self, robot, robot_coll, world_geoms, trajs, cfg: ScoTrajOptConfig = ScoTrajOptConfig()):
        ws, gstart = world_groups(world_geoms)
        if len(gstart) - 1 > 7:
            raise ValueError("at most 7 world geometry groups")
        # margin / mode serve the final ranking cost (colldist, _eval_cost)
        self.model = TrajModel(robot, robot_coll, ws, cfg.collision_margin, mode=1)
        dev = self.model.device
        self.cfg = cfg
        self._c = _config_struct(cfg)
        self.x0, self.start, self.max_len = flatten_trajs(trajs, dev)
        self.B = self.start.shape[0] - 1
        self.A = self.model.n_act
        self.G = len(gstart)
        self.gstart = torch.as_tensor(gstart, device=dev)
        W = self.x0.shape[0]
        self.x = torch.empty_like(self.x0)
        self.d = torch.empty((W, self.G), dtype=torch.float32, device=dev)
        self.J = torch.empty((W, self.G, self.A), dtype=torch.float32, device=dev)
        self.comps = torch.empty((self.B, 3), dtype=torch.float32, device=dev)
        self.lower = self.model.t["lower"]
        self.upper = self.model.t["upper"]
        w = cfg.w_collision
        self.w_sched = []
        for _ in range(cfg.n_outer_iters):
            self.w_sched.append(w)
            w = min(w * cfg.penalty_scale, cfg.w_collision_max)
        self.graph = None

    def _launches(self) -> None:
        from ..kernels._lib import check, lib
        from ..kernels.ops import _ptr, _stream

        st = _stream()
        W = self.x0.shape[0]
        self.x.copy_(self.x0)
        for w_coll in self.w_sched:
            check(lib().pk_traj_sco_linearize_f32(st, self.model.struct_ref(), _ptr(self.x), _ptr(self.start), self.B,
                                                  W, self.max_len, 0, _ptr(self.gstart), self.G - 1,
                                                  float(self.cfg.smooth_min_temperature), _ptr(self.d), _ptr(self.J)),
                  "pk_traj_sco_linearize_f32")
            check(lib().pk_traj_sco_inner_f32(st, _ptr(self.x), _ptr(self.start), self.B, W, self.max_len, self.A,
                                              self.G, _ptr(self.d), _ptr(self.J), _ptr(self.lower), _ptr(self.upper),
                                              ctypes.byref(self._c), float(w_coll), _ptr(self.x)),
                  "pk_traj_sco_inner_f32")
        c = self.cfg
        traj_cost(self.model, self.x, self.start, self.max_len,
                  (c.w_smooth, c.w_vel, c.w_acc, c.w_jerk, c.w_limits, c.w_collision_max), grad=False,
                  out_comps=self.comps)

    def capture(self) -> None:
        self._launches()  # warm-up outside the graph
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launches()
        self.graph = g

    def solve(self):
        """``(final (W, A), costs (B,))``; costs = the full cost at ``w_collision_max``."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launches()
        c = self.cfg
        cost = self.comps @ torch.tensor([c.w_smooth, c.w_limits, c.w_collision_max], device=self.comps.device)
        return self.x, cost


def sco_trajopt(init_trajs, start, goal, robot, robot_coll, world_geoms,
                opt_cfg: ScoTrajOptConfig = ScoTrajOptConfig(), *, use_cuda: bool = True):
    """``(best_traj (T, A), costs (B,), final_trajs (B, T, A))`` like the reference (``:505-548``)."""
    x = torch.as_tensor(np.asarray(init_trajs) if not isinstance(init_trajs, torch.Tensor) else init_trajs,
                        dtype=torch.float32)
    dev = torch.device("cuda", torch.cuda.current_device())
    x = x.to(dev).clone()
    B, T, A = x.shape
    x[:, 0] = torch.as_tensor(np.asarray(start), dtype=torch.float32, device=dev)
    x[:, -1] = torch.as_tensor(np.asarray(goal), dtype=torch.float32, device=dev)
    opt = BatchedSco(robot, robot_coll, world_geoms, x, opt_cfg)
    out, cost = opt.solve()
    final = out.reshape(B, T, A)
    idx = int(torch.argmin(cost))
    res = (final[idx], cost, final)
    if isinstance(init_trajs, torch.Tensor):
        return res
    return tuple(o.cpu().numpy() for o in res)


def make_init_trajs(start, goal, n_batch: int, n_timesteps: int, key, noise_scale: float = 0.05) -> np.ndarray:
    """Linear interpolations start -> goal plus N(0, noise_scale^2) on every waypoint (``:551-564``);
    ``key`` seeds numpy's PCG64 (``jax.random`` streams cannot be reproduced without JAX)."""
    rng = np.random.default_rng(int(np.asarray(key).ravel()[-1]) if np.ndim(key) else int(key))
    s = np.linspace(0.0, 1.0, n_timesteps)[:, None]
    base = (1 - s) * np.asarray(start, np.float64) + s * np.asarray(goal, np.float64)
    noise = rng.normal(0.0, noise_scale, (n_batch, n_timesteps, base.shape[1]))
    return (base[None] + noise).astype(np.float32)
