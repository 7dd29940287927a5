"""Multi-seed Levenberg-Marquardt ("LS") IK on the GPU.

Python entry points with the reference's signatures:

* ``ls_ik_solve_cuda``       -- ``src/pyroffi/optimization_engines/_ls_ik.py:667-855``
* ``ls_ik_solve_cuda_batch`` -- ``src/pyroffi/optimization_engines/_ls_ik.py:1002-1172``

Host work here is the reference's host work: seed generation (half warm
starts N(previous, 0.05^2) clipped to the limits, half uniform in the limits,
fixed joints pinned to ``previous``; ``:452-489``, ``:919-938``), ancestor
masks of the target joints (``:775-792``) and the joint-frame robot sphere
table for the in-kernel collision penalty (``:51-186``).  The LM iterations
and the per-problem winner selection run in the HIP kernels.

Seeds come from numpy's PCG64 instead of ``jax.random`` (threefry); results
agree with the reference statistically (success rates), not bit for bit.
Python constraint callables (evaluated by JAX in the reference) are not
supported.
"""

from __future__ import annotations

import hashlib

from typing import Any

import numpy as np
import torch

from .._se3 import SE3


def rng_from_key(rng_key) -> np.random.Generator:
    """Accept a numpy Generator, an int seed, or any integer array (a "key")."""
    if isinstance(rng_key, np.random.Generator):
        return rng_key
    if rng_key is None:
        return np.random.default_rng(0)
    vals = np.asarray(rng_key).astype(np.int64).ravel()
    return np.random.default_rng([int(v) & 0xFFFFFFFF for v in vals] if vals.size else 0)


def make_seeds(rng: np.random.Generator, previous: np.ndarray, lower: np.ndarray, upper: np.ndarray,
               num_seeds: int, fixed_mask: np.ndarray | None = None, multi_ee: bool = False) -> np.ndarray:
    """Seeds ``(P, num_seeds, A)`` float32 for ``previous`` of shape ``(P, A)``."""
    previous = np.atleast_2d(np.asarray(previous, dtype=np.float64))
    P, A = previous.shape
    n_warm = max(1, num_seeds // 2)
    n_rand = num_seeds - n_warm
    if multi_ee:
        n_tight = max(1, n_warm // 2)
        sig = np.concatenate([np.full(n_tight, 0.05), np.full(n_warm - n_tight, 0.3)])[None, :, None]
    else:
        sig = 0.05
    warm = np.clip(previous[:, None, :] + rng.standard_normal((P, n_warm, A)) * sig, lower, upper)
    rand = rng.uniform(lower, upper, size=(P, n_rand, A))
    seeds = np.concatenate([warm, rand], axis=1)
    if fixed_mask is not None:
        fm = np.asarray(fixed_mask, dtype=bool)
        seeds = np.where(fm[None, None, :], previous[:, None, :], seeds)
    return seeds.astype(np.float32)


def ee_tables(robot, target_link_indices) -> tuple[np.ndarray, np.ndarray]:
    """Target joint per EE and the ancestor mask of each (``(n_ee,)``, ``(n_ee, J)``)."""
    plink = np.asarray(robot.links.parent_joint_indices)
    pj = np.asarray(robot.joints.parent_indices)
    n_j = robot.joints.num_joints
    tj = np.zeros(len(target_link_indices), dtype=np.int32)
    masks = np.zeros((len(target_link_indices), n_j), dtype=np.int32)
    for e, li in enumerate(target_link_indices):
        j = int(plink[li])
        tj[e] = j
        while j >= 0:
            masks[e, j] = 1
            j = int(pj[j])
    return tj, masks


def prepare_ls_collision_buffers(robot, collision_checker: Any | None, world_geom: Any | None):
    """Robot spheres in joint frames + the world (``_ls_ik.py:77-186``).

    Base-link spheres (no parent joint) and padding are dropped, exactly as in
    the reference.  Returns ``(spheres (K, 4), joint_idx (K,), world, enabled)``
    where ``world`` is the ``(M, 4)`` sphere array for a sphere-only world and
    otherwise the packed arrays by kind (``spheres``, ``capsules``, ``boxes``,
    ``halfspaces``; ``_cuda_collision.py:98-168``).
    """
    from ..collision._hip_collision import world_arrays

    empty = (np.zeros((0, 4), np.float32), np.zeros((0,), np.int32), np.zeros((0, 4), np.float32), False)
    if collision_checker is None or world_geom is None:
        return empty
    inner = getattr(collision_checker, "_inner", collision_checker)
    coll = getattr(inner, "coll", None)
    if coll is None:
        return empty
    centers = np.asarray(coll.pose.wxyz_xyz)[..., 4:7].astype(np.float32)
    radii = np.asarray(coll.radius, dtype=np.float32)
    parent_joint = np.asarray(robot.links.parent_joint_indices, dtype=np.int32)
    rows, jidx = [], []
    for li in range(min(centers.shape[0], parent_joint.shape[0])):
        j = int(parent_joint[li])
        if j < 0:
            continue
        for si in range(centers.shape[1]):
            r = float(radii[li, si])
            if r <= 0.0 or r < -1e8:
                continue
            rows.append([*centers[li, si], r])
            jidx.append(j)
    if not rows:
        return empty
    arrs = world_arrays(world_geom)
    mixed = any(v.shape[0] for k, v in arrs.items() if k != "spheres")
    world = arrs if mixed else arrs["spheres"]
    return np.asarray(rows, np.float32), np.asarray(jidx, np.int32), world, True


def _target_array(target_poses) -> np.ndarray:
    if isinstance(target_poses, SE3):
        return target_poses.wxyz_xyz
    if isinstance(target_poses, (tuple, list)):
        return np.stack([_target_array(t) for t in target_poses], axis=-2)
    return np.asarray(target_poses, dtype=np.float64)


def _like(ref, out: torch.Tensor):
    if isinstance(ref, torch.Tensor):
        return out if ref.is_cuda else out.cpu()
    return out.cpu().numpy()


def ls_ik_run(robot, seeds, target_T, target_link_indices, *, max_iter=60, pos_weight=50.0,
              ori_weight=10.0, lambda_init=5e-3, eps_pos=1e-8, eps_ori=1e-8, fixed_joint_mask=None,
              collision=None):
    """Run the LM kernel on explicit seeds ``(P, S, A)`` and targets ``(P, n_ee, 7)``.

    ``collision`` is ``(spheres, joint_idx, world, weight, margin)`` or None; ``world``
    is an ``(M, 4)`` sphere array or packed arrays by kind.
    Returns device tensors ``(cfgs (P, S, A), errs (P, S))``.
    """
    from ..kernels.ops import ls_ik_hip

    t = robot.device_tables()
    dev = t.twists.device
    n_act = robot.joints.num_actuated_joints
    links = tuple(int(i) for i in target_link_indices)
    fm = np.zeros(n_act, np.int32) if fixed_joint_mask is None else np.asarray(fixed_joint_mask).astype(np.int32)
    # Device copies of the small per-call tables, built once per content.
    tj_d, masks_d, fm_d = t.derived(
        ("ls_ik", links, fm.tobytes()),
        lambda: (*(torch.as_tensor(x, device=dev) for x in ee_tables(robot, links)), torch.as_tensor(fm, device=dev)))
    if collision is not None:
        rs, rj, ws, w, m = collision
        enable = True
    else:
        rs, rj, ws, w, m = np.zeros((0, 4), np.float32), np.zeros(0, np.int32), np.zeros((0, 4), np.float32), 0.0, 0.02
        enable = False
    rs = np.ascontiguousarray(rs, np.float32)
    rj = np.ascontiguousarray(rj, np.int32)
    kinds = (("spheres", 4), ("capsules", 7), ("boxes", 15), ("halfspaces", 6))
    wd = ws if isinstance(ws, dict) else {"spheres": ws}
    warr = [np.ascontiguousarray(wd.get(k, np.zeros((0, n))), np.float32).reshape(-1, n) for k, n in kinds]
    ckey = ("ls_ik_coll", hashlib.sha1(b"|".join([rs.tobytes(), rj.tobytes()] + [a.tobytes() for a in warr])).digest(),
            rs.shape, rj.shape, *(a.shape for a in warr))
    rs_d, rj_d, ws_d, wc_d, wb_d, wh_d = t.derived(
        ckey, lambda: (torch.as_tensor(rs, device=dev), torch.as_tensor(rj, device=dev),
                       *(torch.as_tensor(a, device=dev) for a in warr)))
    seeds_t = seeds if isinstance(seeds, torch.Tensor) else torch.as_tensor(np.asarray(seeds, np.float32))
    seeds_t = seeds_t.to(dev, torch.float32).contiguous()
    tT = target_T if isinstance(target_T, torch.Tensor) else torch.as_tensor(np.asarray(target_T, np.float32))
    return ls_ik_hip(
        seeds_t, *t.robot_buffers(), tj_d, masks_d,
        tT.to(dev, torch.float32), rs_d, rj_d, ws_d, wc_d, wb_d, wh_d, t.lower, t.upper, fm_d,
        max_iter=max_iter, pos_weight=pos_weight, ori_weight=ori_weight, lambda_init=lambda_init,
        eps_pos=eps_pos, eps_ori=eps_ori, enable_collision=enable, collision_weight=w, collision_margin=m,
    )


def _unsupported_constraints(constraints) -> None:
    if constraints:
        raise NotImplementedError(
            "Python constraint callables are not supported; use collision_free=True for the "
            "in-kernel collision penalty"
        )


def ls_ik_solve(robot, target_link_indices, target_poses, rng_key, previous_cfg, num_seeds: int = 32,
                max_iter: int = 60, pos_weight: float = 50.0, ori_weight: float = 10.0, lambda_init: float = 5e-3,
                continuity_weight: float = 0.0, fixed_joint_mask=None, constraint_fns=(), constraint_args=(),
                constraint_weights=None):
    """The reference's pure-JAX multi-seed LM (``_ls_ik.py:361-518``); returns ``(n_act,)``.

    Residual ``log(T_ee^-1 T_tgt)`` (jaxlie ``SE3.log``) with its exact
    Jacobian, weights W = (pw, pw, pw, ow, ow, ow), every seed refined by
    ``_ls_ik_single`` (``:192-350``) in ``pk_ik_jax_lm_f32``; the winner has the
    lowest sum over EEs of |W f|^2 plus ``continuity_weight * |q - prev|^2``.
    Seeds: half near ``previous_cfg`` (sigma 0.05; tight/loose halves for
    several EEs), half uniform, from numpy's PCG64.
    """
    _unsupported_constraints(constraint_fns)
    if isinstance(target_link_indices, int):
        target_link_indices = (target_link_indices,)
    links = tuple(int(i) for i in target_link_indices)
    tT = _target_array(target_poses).reshape(1, len(links), 7)
    prev = previous_cfg.detach().cpu().numpy() if isinstance(previous_cfg, torch.Tensor) else np.asarray(previous_cfg)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    seeds = make_seeds(rng_from_key(rng_key), prev[None], lo, hi, num_seeds, fixed_joint_mask,
                       multi_ee=len(links) > 1)
    n_act = robot.joints.num_actuated_joints
    fm = np.zeros(n_act, np.int32) if fixed_joint_mask is None else np.asarray(fixed_joint_mask).astype(np.int32)
    from ..kernels.ops import jax_ik_lm, ls_ik_select

    t = robot.device_tables()
    tj, masks = ee_tables(robot, links)
    cfgs, errs = jax_ik_lm(seeds, t.robot_buffers(), tj, masks, tT.astype(np.float32), t.lower, t.upper, fm,
                           max_iter=max_iter, pos_weight=pos_weight, ori_weight=ori_weight, lambda_init=lambda_init)
    best, _ = ls_ik_select(cfgs, errs, torch.as_tensor(prev[None], dtype=torch.float32, device=cfgs.device),
                           continuity_weight)
    return _like(previous_cfg, best[0])


def ls_ik_solve_cuda(robot, target_link_indices, target_poses, rng_key, previous_cfg, num_seeds: int = 32,
                     max_iter: int = 60, pos_weight: float = 50.0, ori_weight: float = 10.0,
                     lambda_init: float = 5e-3, eps_pos: float = 1e-8, eps_ori: float = 1e-8,
                     continuity_weight: float = 0.0, fixed_joint_mask=None, constraints=None,
                     constraint_args=None, constraint_weights=None, collision_constraint_indices=None,
                     collision_free: bool = False, collision_checker=None, collision_world=None,
                     collision_weight: float = 1e4, collision_margin: float = 0.02,
                     constraint_refine_iters: int = 12):
    """Solve one IK problem (one or more EEs) with ``num_seeds`` LM seeds; returns ``(n_act,)``."""
    _unsupported_constraints(constraints)
    if isinstance(target_link_indices, int):
        target_link_indices = (target_link_indices,)
    tT = _target_array(target_poses).reshape(1, len(target_link_indices), 7)
    prev = previous_cfg.detach().cpu().numpy() if isinstance(previous_cfg, torch.Tensor) else np.asarray(previous_cfg)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    seeds = make_seeds(rng_from_key(rng_key), prev[None], lo, hi, num_seeds, fixed_joint_mask,
                       multi_ee=False)
    coll = None
    if collision_free:
        rs, rj, ws, ok = prepare_ls_collision_buffers(robot, collision_checker, collision_world)
        if ok:
            coll = (rs, rj, ws, collision_weight, collision_margin)
    cfgs, errs = ls_ik_run(robot, seeds, tT, target_link_indices, max_iter=max_iter, pos_weight=pos_weight,
                           ori_weight=ori_weight, lambda_init=lambda_init, eps_pos=eps_pos, eps_ori=eps_ori,
                           fixed_joint_mask=fixed_joint_mask, collision=coll)
    from ..kernels.ops import ls_ik_select

    best, _ = ls_ik_select(cfgs, errs, torch.as_tensor(prev[None], dtype=torch.float32, device=cfgs.device),
                           continuity_weight)
    return _like(previous_cfg, best[0])


def ls_ik_solve_cuda_batch(robot, target_link_indices, target_poses, rng_key, previous_cfgs,
                           num_seeds: int = 32, max_iter: int = 60, pos_weight: float = 50.0,
                           ori_weight: float = 10.0, lambda_init: float = 5e-3, eps_pos: float = 1e-8,
                           eps_ori: float = 1e-8, continuity_weight: float = 0.0, fixed_joint_mask=None,
                           constraints=None, constraint_args=None, constraint_weights=None,
                           collision_constraint_indices=None, collision_free: bool = False,
                           collision_checker=None, collision_world=None, collision_weight: float = 1e4,
                           collision_margin: float = 0.02, constraint_refine_iters: int = 12):
    """Solve ``P`` single-EE problems in one launch; returns ``(P, n_act)``."""
    _unsupported_constraints(constraints)
    if isinstance(target_link_indices, int):
        target_link_indices = (target_link_indices,)
    target_link_indices = tuple(target_link_indices)[:1]  # the reference batch path is single-EE
    tT = _target_array(target_poses).reshape(-1, 1, 7)
    prev = (previous_cfgs.detach().cpu().numpy() if isinstance(previous_cfgs, torch.Tensor)
            else np.asarray(previous_cfgs, dtype=np.float64))
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    seeds = make_seeds(rng_from_key(rng_key), prev, lo, hi, num_seeds, fixed_joint_mask)
    coll = None
    if collision_free:
        rs, rj, ws, ok = prepare_ls_collision_buffers(robot, collision_checker, collision_world)
        if ok:
            coll = (rs, rj, ws, collision_weight, collision_margin)
    cfgs, errs = ls_ik_run(robot, seeds, tT, target_link_indices, max_iter=max_iter, pos_weight=pos_weight,
                           ori_weight=ori_weight, lambda_init=lambda_init, eps_pos=eps_pos, eps_ori=eps_ori,
                           fixed_joint_mask=fixed_joint_mask, collision=coll)
    from ..kernels.ops import ls_ik_select

    best, _ = ls_ik_select(cfgs, errs, torch.as_tensor(prev, dtype=torch.float32, device=cfgs.device),
                           continuity_weight)
    return _like(previous_cfgs, best)
