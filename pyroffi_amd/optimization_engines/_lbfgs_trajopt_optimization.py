"""Direct L-BFGS trajectory optimisation on the GPU (one launch for many trajectories).

Python entry points with the reference's names and signatures
(``src/pyroffi/optimization_engines/_lbfgs_trajopt_optimization.py``):

* ``LbfgsTrajOptConfig``  -- ``:36-70`` (same fields and defaults)
* ``lbfgs_trajopt``       -- ``:276-310``: pins the endpoints, optimises every
  trajectory of the batch, ranks them by the full cost at the maximum
  collision weight (``_eval_cost_lbfgs``, ``:213-231``) and returns
  ``(best_traj, costs, final_trajs)``

and, beyond the reference, ``lbfgs_trajopt_uneven`` / ``BatchedLbfgsTrajOpt``
for trajectories of different lengths and endpoints in one launch, with the
iterations captured in a graph.

The cost (smoothness + joint limits + collision over the JAX self and world
distances) and its gradient come from the HIP kernels (``pk_traj_cost_grad_f32``
and ``pk_traj_lbfgs_f32``); the host work here is the model compilation: the
fixed joints are folded into their moving ancestors ("nodes"), the sphere
table is re-expressed in node frames, and the same-slot self pairs, per-link
bounding spheres and world grids are built once.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from .._se3 import quat_mul, quat_rotate

_IDENTITY = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0])


@dataclass(frozen=True)
class LbfgsTrajOptConfig:
    """Hyper-parameters for the direct L-BFGS trajectory optimizer (``:36-70``)."""

    n_iters: int = 100
    m_lbfgs: int = 6
    w_smooth: float = 1.0
    w_vel: float = 1.0
    w_acc: float = 0.5
    w_jerk: float = 0.1
    w_collision: float = 1.0
    w_collision_max: float = 100.0
    penalty_scale: float = 2.0
    escalation_interval: int = 20
    collision_margin: float = 0.01
    w_limits: float = 1.0


def _compose(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    q = quat_mul(a[:4], b[:4])
    t = quat_rotate(a[:4], b[4:]) + a[4:]
    return np.concatenate([q, t])


def fold_robot(robot) -> dict:
    """Fold fixed joints into their nearest moving ancestor (FP64, host).

    Returns ``nodes (N, 20)`` float32 (layout of ``pk_traj_model.nodes``) and,
    per joint, the node it hangs from and its constant offset from that node.
    A joint is fixed when it has no actuated source and no mimic offset (the
    reference's FK rule, ``_fk_cuda_helpers.cuh:171-207``).
    """
    j = robot.joints
    J = int(j.num_joints)
    tw = np.asarray(j.twists, np.float64)
    ptf = np.asarray(j.parent_transforms, np.float64)
    par = np.asarray(j.parent_indices)
    act = np.asarray(j.actuated_indices)
    mul = np.asarray(j.mimic_multiplier, np.float64)
    off = np.asarray(j.mimic_offset, np.float64)
    mim = np.asarray(j.mimic_act_indices)
    order = np.asarray(j._topo_sort_inv)
    anchor = np.full(J, -1, np.int64)
    rel = np.tile(_IDENTITY, (J, 1))
    nodes = []
    for jj in order:
        jj = int(jj)
        src = int(mim[jj]) if int(mim[jj]) != -1 else int(act[jj])
        p = int(par[jj])
        base, brel = (-1, _IDENTITY) if p < 0 else (int(anchor[p]), rel[p])
        local = _compose(brel, ptf[jj])
        if src < 0 and off[jj] == 0.0:
            anchor[jj] = base
            rel[jj] = local
            continue
        lin, ang = tw[jj, :3], tw[jj, 3:]
        a2, l2 = float(ang @ ang), float(lin @ lin)
        kind = 2 if a2 == 0.0 else (1 if l2 == 0.0 else 3)
        row = np.zeros(20)
        row[:7], row[7:10], row[10:13] = local, lin, ang
        row[13], row[14], row[15], row[16], row[17] = mul[jj], off[jj], src, base, kind
        anchor[jj] = len(nodes)
        rel[jj] = _IDENTITY
        nodes.append(row)
    return {"nodes": np.asarray(nodes, np.float32).reshape(-1, 20), "anchor": anchor, "rel": rel}


def build_traj_tables(robot, robot_coll) -> dict:
    """Host tables of the trajectory model (numpy, float32/int32)."""
    fold = fold_robot(robot)
    anchor, rel = fold["anchor"], fold["rel"]
    lpj = np.asarray(robot.links.parent_joint_indices)
    L = int(robot.links.num_links)
    centers = np.asarray(robot_coll.coll.pose.translation(), np.float64).reshape(L, -1, 3)
    radii = np.asarray(robot_coll.coll.radius, np.float64).reshape(L, -1)
    node_of = np.array([-1 if int(lpj[l]) < 0 else int(anchor[int(lpj[l])]) for l in range(L)])
    # Links grouped by node (stable), so the kernel rebuilds a node's rotation
    # once per waypoint; the tables below are in this order.
    order = np.argsort(node_of, kind="stable")
    new_index = np.empty(L, np.int64)
    new_index[order] = np.arange(L)
    link_node = node_of[order].astype(np.int32)
    sph, start, bound = [], [0], np.zeros((L, 4), np.float32)
    slot_index = {}
    for nl, l in enumerate(order):
        l = int(l)
        pj = int(lpj[l])
        r0 = _IDENTITY if pj < 0 else rel[pj]
        rows = []
        for s in range(radii.shape[1]):
            r = radii[l, s]
            if not r > 0.0:  # padding (r = -1e9) and degenerate spheres
                continue
            c = quat_rotate(r0[:4], centers[l, s]) + r0[4:]
            slot_index[(l, s)] = len(sph) + len(rows)
            rows.append([*c, r])
        rows = np.asarray(rows, np.float64).reshape(-1, 4)
        if len(rows):
            lo, hi = rows[:, :3].min(0), rows[:, :3].max(0)
            cb = (lo + hi) / 2
            rb = np.max(np.linalg.norm(rows[:, :3] - cb, axis=1) + rows[:, 3])
            bound[nl] = [*cb, rb * (1 + 1e-6) + 1e-6]
        sph.extend(rows.tolist())
        start.append(len(sph))
    pi = np.asarray(robot_coll.active_idx_i)
    pj = np.asarray(robot_coll.active_idx_j)
    pair_links, pair_start, items = [], [0], []
    for i, j in zip(pi, pj):
        s = 0
        while (int(i), s) in slot_index and (int(j), s) in slot_index:
            items.append([slot_index[(int(i), s)], slot_index[(int(j), s)]])
            s += 1
        pair_links.append([int(new_index[int(i)]), int(new_index[int(j)])])
        pair_start.append(len(items))
    return {
        "nodes": fold["nodes"],
        "link_order": order.astype(np.int32),  # table link k is robot link link_order[k]
        "n_act": int(robot.joints.num_actuated_joints),
        "link_node": link_node,
        "link_sph_start": np.asarray(start, np.int32),
        "spheres": np.asarray(sph, np.float32).reshape(-1, 4),
        "link_bound": bound,
        "pair_links": np.asarray(pair_links, np.int32).reshape(-1, 2),
        "pair_start": np.asarray(pair_start, np.int32),
        "pair_items": np.asarray(items, np.int32).reshape(-1, 2),
        "lower": np.asarray(robot.joints.lower_limits, np.float32),
        "upper": np.asarray(robot.joints.upper_limits, np.float32),
    }


class TrajModel:
    """Device-resident trajectory model (robot + spheres + world + grids)."""

    def __init__(self, robot, robot_coll, world, margin: float, mode: int = 1, use_grid: bool = True,
                 device=None):
        from ..collision._hip_collision import world_sphere_array
        from ..kernels.ops import world_grid

        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        h = build_traj_tables(robot, robot_coll)
        self.n_act = h["n_act"]
        self.margin = float(margin)
        self.mode = int(mode)
        self.t = {k: torch.as_tensor(v, device=dev) for k, v in h.items() if isinstance(v, np.ndarray)}
        ws = world_sphere_array(world)
        self.t["world"] = torch.as_tensor(np.ascontiguousarray(ws, np.float32).reshape(-1, 4), device=dev)
        self.grids = []
        r = h["spheres"][:, 3]
        if use_grid and len(r) and len(ws):
            r_small, r_big = float(np.quantile(r, 0.9)), float(np.max(r))
            self.grids.append(world_grid(self.t["world"], r_small + self.margin))
            if r_big > r_small:
                self.grids.append(world_grid(self.t["world"], r_big + self.margin, max_cells=24))
        self.device = dev
        self._struct = self._make_struct()

    def _make_struct(self):
        from ..kernels._lib import TrajModel as CModel
        from ..kernels.ops import _grid_array

        t = self.t
        self._grid_arr, n_grids = _grid_array(self.grids or None)
        m = CModel()
        m.nodes = t["nodes"].data_ptr()
        m.n_nodes = t["nodes"].shape[0]
        m.n_act = self.n_act
        m.link_node = t["link_node"].data_ptr()
        m.link_sph_start = t["link_sph_start"].data_ptr()
        m.spheres = t["spheres"].data_ptr()
        m.link_bound = t["link_bound"].data_ptr()
        m.n_links = t["link_node"].shape[0]
        m.n_spheres = t["spheres"].shape[0]
        m.pair_links = t["pair_links"].data_ptr()
        m.pair_start = t["pair_start"].data_ptr()
        m.pair_items = t["pair_items"].data_ptr()
        m.n_pairs = t["pair_links"].shape[0]
        m.n_items = t["pair_items"].shape[0]
        m.world = t["world"].data_ptr()
        m.n_world = t["world"].shape[0]
        m.grids = ctypes.cast(self._grid_arr, ctypes.c_void_p).value if self._grid_arr is not None else None
        m.n_grids = n_grids
        m.lower = t["lower"].data_ptr()
        m.upper = t["upper"].data_ptr()
        m.margin = self.margin
        m.mode = self.mode
        return m

    def struct_ref(self):
        return ctypes.byref(self._struct)


def flatten_trajs(trajs, device) -> tuple[torch.Tensor, torch.Tensor, int]:
    """List of ``(T_b, A)`` or a ``(B, T, A)`` batch -> flat ``(W, A)`` float32, int64 starts, max T."""
    if isinstance(trajs, torch.Tensor) and trajs.dim() == 3:
        B, T, A = trajs.shape
        x = trajs.reshape(B * T, A).to(device, torch.float32).contiguous()
        start = torch.arange(0, (B + 1) * T, T, dtype=torch.int64, device=device)
        return x, start, int(T)
    if isinstance(trajs, np.ndarray) and trajs.ndim == 3:
        trajs = list(trajs)
    lens = [len(t) for t in trajs]
    x = torch.cat([torch.as_tensor(np.asarray(t) if not isinstance(t, torch.Tensor) else t, dtype=torch.float32)
                   .to(device) for t in trajs], 0).contiguous()
    start = torch.zeros(len(lens) + 1, dtype=torch.int64)
    start[1:] = torch.cumsum(torch.as_tensor(lens, dtype=torch.int64), 0)
    return x, start.to(device), int(max(lens))


def traj_cost(model: TrajModel, x: torch.Tensor, start: torch.Tensor, max_len: int, weights, grad: bool = True,
              out_comps=None, out_grad=None, block_threads: int = 0, subset: torch.Tensor | None = None):
    """Components ``(B, 3)`` = (smooth, limits, coll) and the gradient ``(W, A)`` of the weighted total.

    ``subset`` (int32 device tensor) restricts the launch to those trajectories;
    ``block_threads`` (128/256/512, 0 = from ``max_len``) sets the workgroup size.
    """
    from ..kernels._lib import check, lib
    from ..kernels.ops import _ptr, _stream

    B = start.shape[0] - 1
    W = x.shape[0]
    w = np.ascontiguousarray(weights, np.float32)
    if out_comps is None:
        out_comps = torch.empty((B, 3), dtype=torch.float32, device=x.device)
    if grad and out_grad is None:
        out_grad = torch.empty_like(x)
    st = lib().pk_traj_cost_grad_f32(_stream(), model.struct_ref(), _ptr(x), _ptr(start), B, W, int(max_len),
                                     int(block_threads), _ptr(subset), 0 if subset is None else subset.shape[0],
                                     w.ctypes.data_as(ctypes.c_void_p), _ptr(out_comps),
                                     _ptr(out_grad if grad else None))
    check(st, "pk_traj_cost_grad_f32")
    return out_comps, (out_grad if grad else None)


def _config_struct(cfg: LbfgsTrajOptConfig):
    from ..kernels._lib import LbfgsConfig

    c = LbfgsConfig()
    c.n_iters, c.m_hist, c.escalation_interval = cfg.n_iters, cfg.m_lbfgs, cfg.escalation_interval
    c.w_smooth, c.w_vel, c.w_acc, c.w_jerk, c.w_limits = cfg.w_smooth, cfg.w_vel, cfg.w_acc, cfg.w_jerk, cfg.w_limits
    c.w_collision, c.w_collision_max, c.penalty_scale = cfg.w_collision, cfg.w_collision_max, cfg.penalty_scale
    return c


# Workgroup size per trajectory length: each 8-lane group gets one to four
# waypoints, so long and short trajectories need about the same number of
# rounds per iteration.
# (longest length, workgroup threads): 8 lanes per waypoint, so trajectories
# up to 32 and 64 waypoints evaluate in one round; longer ones take several.
_LENGTH_CLASSES = ((32, 256), (64, 512), (1 << 30, 512))


class BatchedLbfgsTrajOpt:
    """Many trajectories (any lengths) through the direct L-BFGS in chunked launches.

    Trajectories are split into length classes (up to 32, 64, longer), each
    launched with its own workgroup size on a stream of its own, so the classes
    run concurrently.  The longest class sets the solve time and is launched
    first, so its workgroups are placed before the shorter classes fill the
    remaining CUs.  ``chunk`` iterations run per launch (0: all in one);
    ``capture()`` records every launch of a solve into one graph and
    ``solve()`` replays it.
    """

    def __init__(self, model: TrajModel, trajs, cfg: LbfgsTrajOptConfig = LbfgsTrajOptConfig(), chunk: int = 0,
                 split_classes: bool = True):
        from ..kernels._lib import lib

        self.model, self.cfg = model, cfg
        self.x0, self.start, self.max_len = flatten_trajs(trajs, model.device)
        self.B = self.start.shape[0] - 1
        W, A = self.x0.shape
        n_state = lib().pk_traj_lbfgs_state_floats(W, A, self.B, cfg.m_lbfgs)
        self.state = torch.empty((n_state,), dtype=torch.float32, device=model.device)
        self.best = torch.empty_like(self.x0)
        self.best_cost = torch.empty((self.B,), dtype=torch.float32, device=model.device)
        self.rank_comps = torch.empty((self.B, 3), dtype=torch.float32, device=model.device)
        self.chunk = int(chunk)
        self._cfg = _config_struct(cfg)
        self.graph = None
        lens = np.diff(self.start.cpu().numpy())
        self.classes = []  # (subset tensor or None, n, max_len, threads)
        if split_classes:
            lo = 0
            for hi, threads in _LENGTH_CLASSES:
                idx = np.nonzero((lens > lo) & (lens <= hi))[0].astype(np.int32)
                if len(idx):
                    self.classes.append((torch.as_tensor(idx, device=model.device), len(idx),
                                         int(lens[idx].max()), threads))
                lo = hi
            self.classes.reverse()  # longest first
        else:
            self.classes.append((None, self.B, self.max_len, 0))
        self._streams = None

    def _launch_iters(self, subset, count, max_len, threads, b0, b1):
        from ..kernels._lib import check, lib
        from ..kernels.ops import _ptr, _stream

        st = lib().pk_traj_lbfgs_f32(_stream(), self.model.struct_ref(), _ptr(self.x0), _ptr(self.start), self.B,
                                     self.x0.shape[0], max_len, threads, _ptr(subset), 0 if subset is None else count,
                                     ctypes.byref(self._cfg), b0, b1, _ptr(self.state), _ptr(self.best),
                                     _ptr(self.best_cost))
        check(st, "pk_traj_lbfgs_f32")

    def _per_class(self, fn):
        if len(self.classes) == 1:
            fn(*self.classes[0])
            return
        main = torch.cuda.current_stream()
        if self._streams is None:
            self._streams = [torch.cuda.Stream() for _ in self.classes]
        for c, s in zip(self.classes, self._streams):
            s.wait_stream(main)
            with torch.cuda.stream(s):
                fn(*c)
        for s in self._streams:
            main.wait_stream(s)

    def _launches(self):
        n = self.cfg.n_iters
        step = n if self.chunk <= 0 else self.chunk
        cfg = self.cfg
        rank_w = (cfg.w_smooth, cfg.w_vel, cfg.w_acc, cfg.w_jerk, cfg.w_limits, cfg.w_collision_max)

        def run_class(subset, count, max_len, threads):
            for b0 in range(0, max(n, 1), max(step, 1)):
                self._launch_iters(subset, count, max_len, threads, b0, min(b0 + step, n))
            traj_cost(self.model, self.best, self.start, max_len, rank_w, grad=False, out_comps=self.rank_comps,
                      block_threads=threads, subset=subset)

        self._per_class(run_class)

    def run_iterations(self, it0: int, it1: int) -> None:
        """Run iterations ``[it0, it1)`` of the solve (``it0 == 0`` starts from
        the initial paths, later ranges resume from the state buffer), without the
        final ranking; ``final_iterate()`` then holds iterate ``it1``."""
        if not 0 <= it0 <= it1 <= self.cfg.n_iters:
            raise ValueError(f"iteration range [{it0}, {it1}) outside [0, {self.cfg.n_iters}]")
        self._per_class(lambda *c: self._launch_iters(*c, it0, it1))

    def final_iterate(self) -> torch.Tensor:
        """The last L-BFGS iterate ``(W, A)`` (a view of the state buffer)."""
        W, A = self.x0.shape
        return self.state[: W * A].view(W, A)

    def capture(self) -> None:
        self._launches()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launches()
        self.graph = g

    def solve(self):
        """``(best (W, A), tracked best cost (B,), ranking cost (B,))`` as device tensors."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launches()
        cfg = self.cfg
        rank = (cfg.w_smooth * self.rank_comps[:, 0] + cfg.w_limits * self.rank_comps[:, 1]
                + cfg.w_collision_max * self.rank_comps[:, 2])
        return self.best, self.best_cost, rank


def lbfgs_trajopt_uneven(trajs, robot, robot_coll, world_geoms, opt_cfg: LbfgsTrajOptConfig = LbfgsTrajOptConfig(),
                         mode: int = 1):
    """Trajectories of different lengths/endpoints (rows 0 and -1 are the pinned endpoints).

    Returns ``(list of optimised (T_b, A) tensors, ranking costs (B,))``.
    """
    model = TrajModel(robot, robot_coll, world_geoms, opt_cfg.collision_margin, mode)
    opt = BatchedLbfgsTrajOpt(model, trajs, opt_cfg)
    best, _, rank = opt.solve()
    s = opt.start.cpu().tolist()
    return [best[s[b]:s[b + 1]] for b in range(opt.B)], rank


def lbfgs_trajopt(init_trajs, start, goal, robot, robot_coll, world_geoms,
                  opt_cfg: LbfgsTrajOptConfig = LbfgsTrajOptConfig(), *, use_cuda: bool = True):
    """``(best_traj (T, A), costs (B,), final_trajs (B, T, A))`` like the reference (``:276-310``)."""
    x = torch.as_tensor(np.asarray(init_trajs) if not isinstance(init_trajs, torch.Tensor) else init_trajs,
                        dtype=torch.float32)
    dev = torch.device("cuda", torch.cuda.current_device())
    x = x.to(dev).clone()
    B, T, A = x.shape
    x[:, 0] = torch.as_tensor(np.asarray(start), dtype=torch.float32, device=dev)
    x[:, -1] = torch.as_tensor(np.asarray(goal), dtype=torch.float32, device=dev)
    model = TrajModel(robot, robot_coll, world_geoms, opt_cfg.collision_margin, mode=1)
    opt = BatchedLbfgsTrajOpt(model, x, opt_cfg)
    best, _, rank = opt.solve()
    final = best.reshape(B, T, A)
    idx = int(torch.argmin(rank))
    out = (final[idx], rank, final)
    if isinstance(init_trajs, torch.Tensor):
        return out
    return tuple(o.cpu().numpy() for o in out)
