"""Uneven multi-robot LM-IK in persistent launches, captured in one graph.

The reference solves a batch of IK problems with ``ls_ik_solve_cuda_batch``
(``src/pyroffi/optimization_engines/_ls_ik.py:1002-1172``): one robot, one
end effector, the same seed count for every problem, one kernel launch per
call.  Problems that differ in robot, DOF, end-effector count or seed count
need one call each.  ``UnevenLsIkSolver`` takes any mix:

* problems are grouped by robot; each group's seeds are one flat list with
  CSR offsets per problem (``pk_ls_ik_batched_f32``), so seed counts can differ
  per problem and tiles of 64 seeds may straddle problems;
* each group runs as a persistent launch (a fixed number of workgroups
  walking the tiles), its iterations optionally split into chunks that resume
  from a per-seed state buffer;
* the collision penalty is evaluated over an exact world grid
  (``pk_world_grid_f32``) instead of every world sphere;
* all chunk launches and the per-problem winner selections of all groups are
  captured once into a graph (``torch.cuda.CUDAGraph`` drives HIP stream
  capture) and replayed; each group is a branch of its own, so the robots'
  launches run concurrently;
* an autotuner times candidate (workgroups, chunk length) pairs per group on
  the actual batch and caches the winners by batch signature.
"""

from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from ..kernels.ops import ls_ik_batched, ls_ik_hip, ls_ik_select, ls_ik_select_segments, world_grid
from ._ls_ik import ee_tables, prepare_ls_collision_buffers

_CACHE_PATH = os.path.join(os.path.expanduser("~"), ".cache", "pyroffi_amd", "autotune_ls_ik.json")

# Persistent-workgroup counts tried by the autotuner (0 = one workgroup per tile).
GRID_CANDIDATES = (0, 1024, 2048, 4096)
CHUNK_CANDIDATES = (0, 10, 25)  # 0 = all iterations in one launch


@dataclass
class RobotGroup:
    """All problems of one robot: flat seeds (S_total, A) and CSR offsets."""

    robot: object
    target_link_indices: tuple
    targets: np.ndarray  # (P, n_ee, 7)
    seeds: np.ndarray  # (S_total, A) float32
    seed_start: np.ndarray  # (P + 1,) int32
    collision: tuple | None = None  # (robot spheres (K,4), joint idx (K,), world (M,4))
    device: dict = field(default_factory=dict)
    tune: tuple = (0, 0)
    tune_ms: float = 0.0  # autotuned time of the group alone (launch order)


class UnevenLsIkSolver:
    """Multi-seed LM over problems of several robots with per-problem seed counts."""

    def __init__(self, groups: list[RobotGroup], *, max_iter: int = 50, pos_weight: float = 50.0,
                 ori_weight: float = 10.0, lambda_init: float = 5e-3, eps_pos: float = 1e-8,
                 eps_ori: float = 1e-8, collision_weight: float = 1e4, collision_margin: float = 0.02,
                 use_grid: bool = True):
        self.groups = groups
        self.kw = dict(max_iter=max_iter, pos_weight=pos_weight, ori_weight=ori_weight,
                       lambda_init=lambda_init, eps_pos=eps_pos, eps_ori=eps_ori,
                       collision_weight=collision_weight, collision_margin=collision_margin)
        self.use_grid = use_grid
        self.graph = None
        self._streams = None
        for g in groups:
            self._upload(g)

    # ------------------------------------------------------------------ setup
    def _upload(self, g: RobotGroup) -> None:
        dev = torch.device("cuda", torch.cuda.current_device())
        t = g.robot.device_tables(dev)
        tj, masks = ee_tables(g.robot, g.target_link_indices)
        d = {
            "tables": t.robot_buffers(),
            "lower": t.lower,
            "upper": t.upper,
            "seeds": torch.as_tensor(np.ascontiguousarray(g.seeds, np.float32), device=dev),
            "seed_start": torch.as_tensor(np.asarray(g.seed_start, np.int32), device=dev),
            "tj": torch.as_tensor(tj, device=dev),
            "masks": torch.as_tensor(masks, device=dev),
            "targets": torch.as_tensor(np.ascontiguousarray(g.targets, np.float32), device=dev),
            "fixed": torch.zeros(g.robot.joints.num_actuated_joints, dtype=torch.int32, device=dev),
        }
        S, A = g.seeds.shape
        P = len(g.seed_start) - 1
        d["out_cfgs"] = torch.empty((S, A), dtype=torch.float32, device=dev)
        d["out_err"] = torch.empty((S,), dtype=torch.float32, device=dev)
        d["state"] = torch.empty((S, 2 * A + 3), dtype=torch.float32, device=dev)
        d["best"] = torch.empty((P, A), dtype=torch.float32, device=dev)
        d["best_idx"] = torch.empty((P,), dtype=torch.int32, device=dev)
        if g.collision is not None:
            rs, rj, ws = g.collision
            d["rs"] = torch.as_tensor(rs, device=dev)
            d["rj"] = torch.as_tensor(rj, device=dev)
            d["ws"] = torch.as_tensor(ws, device=dev)
            if self.use_grid and len(rs) and len(ws):
                # A fine grid reaching the 90th-percentile robot sphere radius and,
                # when a few robot spheres are larger (a torso, a base plate), a
                # coarse one reaching the largest.  The kernel picks per sphere;
                # every choice gives the same penalty.  64 / 48 cells along the
                # longest side: on 256 world spheres in a 2 m cube they test
                # 15-20% fewer candidates than 48 / 24, and finer grids (2 MiB of
                # cells per robot at 64) lose more to the L2 than they save when
                # four robots run together.
                margin = self.kw["collision_margin"]
                r_small, r_big = float(np.quantile(rs[:, 3], 0.9)), float(np.max(rs[:, 3]))
                d["grid"] = [world_grid(d["ws"], r_small + margin, max_cells=64)]
                if r_big > r_small:
                    d["grid"].append(world_grid(d["ws"], r_big + margin, max_cells=48))
        else:
            d["rs"] = torch.zeros((0, 4), device=dev)
            d["rj"] = torch.zeros((0,), dtype=torch.int32, device=dev)
            d["ws"] = torch.zeros((0, 4), device=dev)
        g.device = d

    # ---------------------------------------------------------------- running
    def _run_group(self, g: RobotGroup, grid_waves: int, chunk: int, work=None) -> None:
        d = g.device
        kw = self.kw
        max_iter = kw["max_iter"]
        bounds = [0, max_iter] if chunk <= 0 or chunk >= max_iter else list(range(0, max_iter, chunk)) + [max_iter]
        for b0, b1 in zip(bounds[:-1], bounds[1:]):
            ls_ik_batched(
                d["seeds"], d["seed_start"], d["tables"], d["tj"], d["masks"], d["targets"], d["rs"], d["rj"],
                d["ws"], d["lower"], d["upper"], d["fixed"], max_iter=max_iter, iter_begin=b0, iter_end=b1,
                state=d["state"] if len(bounds) > 2 else None, grid=d.get("grid"),
                pos_weight=kw["pos_weight"], ori_weight=kw["ori_weight"], lambda_init=kw["lambda_init"],
                eps_pos=kw["eps_pos"], eps_ori=kw["eps_ori"], enable_collision=g.collision is not None,
                collision_weight=kw["collision_weight"], collision_margin=kw["collision_margin"],
                grid_waves=grid_waves, out_cfgs=d["out_cfgs"], out_err=d["out_err"], work=work,
            )
        ls_ik_select_segments(d["out_cfgs"], d["out_err"], d["seed_start"], out_best=d["best"],
                              out_idx=d["best_idx"])

    def run_eager(self) -> None:
        """Every group on a stream of its own (forked from and joined to the current one).

        Groups are launched longest first (by their autotuned time): the longest
        sets the batch time, and its workgroups take the CUs before the others
        fill the remaining ones.
        """
        if len(self.groups) == 1:
            self._run_group(self.groups[0], *self.groups[0].tune)
            return
        main = torch.cuda.current_stream()
        if self._streams is None:
            self._streams = [torch.cuda.Stream() for _ in self.groups]
        order = sorted(self.groups, key=lambda g: -g.tune_ms)
        for g, s in zip(order, self._streams):
            s.wait_stream(main)
            with torch.cuda.stream(s):
                self._run_group(g, *g.tune)
        for s in self._streams:
            main.wait_stream(s)

    def _signature(self, g: RobotGroup) -> str:
        n = g.robot.joints.num_actuated_joints
        return json.dumps([n, g.robot.joints.num_joints, len(g.target_link_indices), int(g.seeds.shape[0]),
                           int(len(g.seed_start) - 1), g.collision is not None, self.kw["max_iter"],
                           self.use_grid])

    def autotune(self, reps: int = 2, cache: bool = True) -> dict:
        """Time every (workgroups, chunk) candidate per group; keep the fastest."""
        saved = {}
        if cache and os.path.exists(_CACHE_PATH):
            try:
                with open(_CACHE_PATH) as f:
                    saved = json.load(f)
            except (OSError, ValueError):
                saved = {}
        report = {}
        for g in self.groups:
            key = self._signature(g)
            if key in saved:
                g.tune = tuple(saved[key][:2])
                g.tune_ms = float(saved[key][2]) if len(saved[key]) > 2 else 0.0
                report[key] = {"choice": g.tune, "cached": True}
                continue
            timings = {}
            for gw in GRID_CANDIDATES:
                for ch in CHUNK_CANDIDATES:
                    self._run_group(g, gw, ch)  # warm-up
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    for _ in range(reps):
                        self._run_group(g, gw, ch)
                    torch.cuda.synchronize()
                    timings[(gw, ch)] = (time.perf_counter() - t0) / reps
            g.tune = min(timings, key=timings.get)
            g.tune_ms = timings[g.tune] * 1e3
            saved[key] = [*g.tune, g.tune_ms]
            report[key] = {"choice": g.tune, "ms": {f"{k[0]}/{k[1]}": v * 1e3 for k, v in timings.items()}}
        if len(self.groups) > 1:
            report["joint"] = self._tune_lead(saved, reps)
        if cache:
            try:
                os.makedirs(os.path.dirname(_CACHE_PATH), exist_ok=True)
                with open(_CACHE_PATH, "w") as f:
                    json.dump(saved, f)
            except OSError:
                pass
        self.graph = None
        return report

    def _tune_lead(self, saved: dict, reps: int) -> dict:
        """Workgroups of the longest group, timed with every group running.

        Launched first, the longest group's workgroups take every CU they can
        hold, and the others wait for its first round.  Fewer of its
        workgroups, each looping over two or three tiles, finish it in the
        same number of rounds while leaving CUs to the others (on four robots
        with 4096 targets and 355 k seeds, one box: 81.9 -> 75.4 ms).  The
        choice is the fastest whole run.
        """
        lead = max(self.groups, key=lambda g: g.tune_ms)
        key = "lead:" + json.dumps([self._signature(g) for g in self.groups])
        if key in saved:
            lead.tune = tuple(saved[key])
            return {"choice": lead.tune, "cached": True}
        n_tiles = -(-int(lead.seeds.shape[0]) // 64)
        cands = [lead.tune] + [(-(-n_tiles // k), lead.tune[1]) for k in (2, 3) if n_tiles >= 2 * k]
        timings = {}
        for c in dict.fromkeys(cands):
            lead.tune = c
            self.run_eager()  # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                self.run_eager()
            torch.cuda.synchronize()
            timings[c] = (time.perf_counter() - t0) / reps
        lead.tune = min(timings, key=timings.get)
        saved[key] = list(lead.tune)
        return {"choice": lead.tune, "ms": {f"{k[0]}/{k[1]}": v * 1e3 for k, v in timings.items()}}

    def capture(self, keep_faster: bool = False, reps: int = 3) -> str:
        """Capture every group's chunk launches and selections into one graph.

        With ``keep_faster`` the replay is timed against the same launches made
        directly and the graph is dropped when it is slower: the direct
        launches start the groups longest-first, and a graph's replay need
        not launch its independent branches in that order.  Returns the mode
        ``solve`` will use, ``"graph"`` or ``"eager"``.
        """
        self.run_eager()  # warm-up outside the capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run_eager()
        self.graph = g
        if keep_faster and len(self.groups) > 1:
            def timed(fn):
                fn()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn()
                torch.cuda.synchronize()
                return time.perf_counter() - t0
            if timed(self.run_eager) < timed(g.replay):
                self.graph = None
        return "graph" if self.graph is not None else "eager"

    def count_work(self) -> list[int]:
        """Per group, the sphere-pair tests its collision penalty executes in one
        solve (one eager run with the kernel's work counter)."""
        out = []
        for g in self.groups:
            w = torch.zeros(1, dtype=torch.int64, device=g.device["seeds"].device)
            self._run_group(g, *g.tune, work=w)
            out.append(int(w.item()))
        return out

    def solve(self) -> list[tuple[torch.Tensor, torch.Tensor]]:
        """Run (replay the graph if captured); per group ``(winners (P, A), winner index (P,))``."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self.run_eager()
        return [(g.device["best"], g.device["best_idx"]) for g in self.groups]

    def solve_per_call(self) -> list[torch.Tensor]:
        """The plain path: one single-problem LM launch plus one selection per problem.

        Each problem's seeds go through ``pk_ls_ik_f32`` against every world
        sphere (no grid), as ``ls_ik_solve_cuda`` does per call; the inputs are
        already on the device, so this is the launch path without host work.
        Returns per group the winners ``(P, A)``.
        """
        kw = self.kw
        outs = []
        for g in self.groups:
            d = g.device
            A = g.seeds.shape[1]
            starts = [int(v) for v in g.seed_start]
            no = torch.zeros((0, 1), device=d["seeds"].device)
            best = torch.empty((len(starts) - 1, A), dtype=torch.float32, device=d["seeds"].device)
            for p in range(len(starts) - 1):
                cfgs, errs = ls_ik_hip(
                    d["seeds"][starts[p]:starts[p + 1]].unsqueeze(0), *d["tables"], d["tj"], d["masks"],
                    d["targets"][p:p + 1], d["rs"], d["rj"], d["ws"], no, no, no, d["lower"], d["upper"],
                    d["fixed"], max_iter=kw["max_iter"], pos_weight=kw["pos_weight"],
                    ori_weight=kw["ori_weight"], lambda_init=kw["lambda_init"], eps_pos=kw["eps_pos"],
                    eps_ori=kw["eps_ori"], enable_collision=g.collision is not None,
                    collision_weight=kw["collision_weight"], collision_margin=kw["collision_margin"])
                b, _ = ls_ik_select(cfgs, errs, d["seeds"][starts[p]:starts[p] + 1])
                best[p:p + 1].copy_(b)
            outs.append(best)
        return outs


    def solve_batch_padded(self, n_problems: int | None = None) -> list[torch.Tensor]:
        """The reference's batch call shape: per robot, one ``pk_ls_ik_f32`` launch
        over its first ``n_problems`` problems (all by default) with every problem's
        seeds padded to the largest count, brute-force collision penalty, then one
        selection -- ``ls_ik_solve_cuda_batch`` (``_ls_ik.py:1002-1172``) takes a
        single seed count for all problems.  Padding rows repeat the problem's
        first seed, which ties with it and never wins (first index on ties), so
        the winners are those of the problem's own seeds.  Returns per group the
        winners ``(n, A)``.
        """
        kw = self.kw
        outs = []
        for g in self.groups:
            d = g.device
            A = g.seeds.shape[1]
            starts = np.asarray(g.seed_start, np.int64)
            n = len(starts) - 1 if n_problems is None else min(int(n_problems), len(starts) - 1)
            counts = np.diff(starts[:n + 1])
            smax = int(counts.max())
            rows = np.concatenate([np.r_[np.arange(starts[p], starts[p + 1]), np.full(smax - counts[p], starts[p])]
                                   for p in range(n)])
            idx = torch.as_tensor(rows, device=d["seeds"].device)
            seeds = d["seeds"][idx].reshape(n, smax, A)
            no = torch.zeros((0, 1), device=d["seeds"].device)
            cfgs, errs = ls_ik_hip(
                seeds, *d["tables"], d["tj"], d["masks"], d["targets"][:n], d["rs"], d["rj"], d["ws"], no, no, no,
                d["lower"], d["upper"], d["fixed"], max_iter=kw["max_iter"], pos_weight=kw["pos_weight"],
                ori_weight=kw["ori_weight"], lambda_init=kw["lambda_init"], eps_pos=kw["eps_pos"],
                eps_ori=kw["eps_ori"], enable_collision=g.collision is not None,
                collision_weight=kw["collision_weight"], collision_margin=kw["collision_margin"])
            best, _ = ls_ik_select(cfgs, errs, seeds[:, 0])
            outs.append(best)
        return outs


def log_uniform_seed_counts(rng: np.random.Generator, n: int, lo: int = 16, hi: int = 256) -> np.ndarray:
    """Seeds per problem: floor(exp(U(ln lo, ln hi)))."""
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(np.int64)


def make_group(robot, target_link_indices, targets: np.ndarray, seed_counts: np.ndarray,
               rng: np.random.Generator, previous: np.ndarray | None = None, collision_checker=None,
               world=None) -> RobotGroup:
    """One robot's uneven problems.

    Seeds are uniform in the joint limits, or, given ``previous`` (P, A), half
    N(previous, 0.05^2) clipped to the limits and half uniform (the reference's
    warm-start rule, ``_ls_ik.py:452-489``).
    """
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    A = robot.joints.num_actuated_joints
    P = len(seed_counts)
    rows = []
    for p, n in enumerate(seed_counts):
        if previous is None:
            rows.append(rng.uniform(lo, hi, (int(n), A)))
            continue
        n_warm = max(1, int(n) // 2)
        warm = np.clip(previous[p] + rng.standard_normal((n_warm, A)) * 0.05, lo, hi)
        rand = rng.uniform(lo, hi, (int(n) - n_warm, A))
        rows.append(np.concatenate([warm, rand], 0))
    seeds = np.concatenate(rows, 0).astype(np.float32)
    start = np.zeros(P + 1, np.int32)
    start[1:] = np.cumsum(seed_counts)
    coll = None
    if collision_checker is not None and world is not None:
        rs, rj, ws, ok = prepare_ls_collision_buffers(robot, collision_checker, world)
        if ok and isinstance(ws, dict):
            raise NotImplementedError("the uneven batch solver's penalty grid covers sphere worlds; use "
                                      "ls_ik_solve_cuda for capsule, box or half-space obstacles")
        if ok:
            coll = (rs, rj, ws)
    return RobotGroup(robot=robot, target_link_indices=tuple(target_link_indices),
                      targets=np.asarray(targets, np.float64).reshape(P, len(target_link_indices), 7),
                      seeds=seeds, seed_start=start, collision=coll)
