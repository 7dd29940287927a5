"""HJCD-IK (Hamiltonian Jacobian coordinate descent + LM refinement) on the GPU.

``hjcd_solve`` has the signature and algorithm of the reference's pure-JAX
solver (``src/pyroffi/optimization_engines/_hjcd_ik.py:395-634``), which
``Robot.inverse_kinematics`` calls (``src/pyroffi/_robot.py:111-178``):

1. seeds: the first ``top_k`` near ``previous_cfg`` (sigma 0.05; for several
   end effectors a tight sigma-0.05 and a loose sigma-0.3 half), the rest
   uniform in the limits, fixed joints pinned to ``previous_cfg``;
2. coarse phase (``pk_hjcd_coarse_f32``): ``coarse_max_iter`` greedy
   Hamiltonian coordinate-descent steps per seed on one end effector (seeds
   split equally over the end effectors), scored by the summed squared
   residual of all end effectors;
3. refinement seeds (``pk_hjcd_refine_seeds_f32``): the ``top_k`` best coarse
   results (stable order) tiled ``repeats`` times, every copy after the
   first ``top_k`` perturbed by N(0, 1) * 0.15 * (upper - lower);
4. LM refinement (``pk_ik_jax_lm_f32`` with per-EE adaptive weights, the soft
   joint-limit prior and stall kicks) on all end effectors;
5. winner (``pk_ls_ik_select_f32``): lowest sum of squared residuals plus
   ``continuity_weight * |q - previous_cfg|^2``.

The residual is the JAX one, ``log(T_ee^-1 T_tgt)`` (jaxlie ``SE3.log``), with
its exact Jacobian.  Random numbers: the seeds and perturbations come from
numpy's PCG64 on the host (``jax.random`` threefry streams cannot be
reproduced without JAX), the stall kicks from a counter-based generator keyed
by the same stream; results agree with the reference in distribution, not bit
for bit.  Python constraint callables (evaluated by JAX in the reference) are
not supported.
"""

from __future__ import annotations

import numpy as np
import torch

from ._ls_ik import _like, _target_array, _unsupported_constraints, ee_tables, rng_from_key


def _get_refine_schedule(num_seeds: int) -> tuple[int, int]:
    """``(top_k, repeats)`` for a coarse batch of ``num_seeds`` (``_hjcd_ik.py:368-385``)."""
    if num_seeds <= 8:
        return num_seeds, 24
    if num_seeds <= 16:
        return max(1, num_seeds // 2), 16
    if num_seeds <= 128:
        return max(1, int(num_seeds * 0.2)), 5
    if num_seeds <= 1024:
        return max(1, int(num_seeds * 0.02)), 5
    if num_seeds <= 2048:
        return max(1, int(num_seeds * 0.01)), 5
    return max(1, int(num_seeds * 0.01)), 2


def hjcd_seeds(rng: np.random.Generator, previous: np.ndarray, lower, upper, num_seeds: int, n_ee: int,
               fixed_mask=None) -> np.ndarray:
    """Coarse-phase seeds ``(P, num_seeds, A)`` for ``previous`` ``(P, A)`` (``_hjcd_ik.py:516-553``)."""
    previous = np.atleast_2d(np.asarray(previous, np.float64))
    P, A = previous.shape
    top_k, _ = _get_refine_schedule(num_seeds)
    n_warm = min(top_k, num_seeds)
    if n_ee > 1:
        n_tight = max(1, n_warm // 2)
        sig = np.concatenate([np.full(n_tight, 0.05), np.full(n_warm - n_tight, 0.3)])[None, :, None]
    else:
        sig = 0.05
    warm = np.clip(previous[:, None, :] + rng.standard_normal((P, n_warm, A)) * sig, lower, upper)
    rand = rng.uniform(lower, upper, size=(P, num_seeds - n_warm, A))
    seeds = np.concatenate([warm, rand], axis=1)
    if fixed_mask is not None:
        seeds = np.where(np.asarray(fixed_mask, bool)[None, None, :], previous[:, None, :], seeds)
    return seeds.astype(np.float32)


def hjcd_solve_batch(robot, target_link_indices, target_poses, rng_key, previous_cfgs, num_seeds: int = 32,
                     coarse_max_iter: int = 20, lm_max_iter: int = 40, epsilon: float = 0.02,
                     nu: float = float(np.pi / 2), eps_pos: float = 1e-8, eps_ori: float = 1e-8,
                     lambda_init: float = 5e-3, continuity_weight: float = 0.0, limit_prior_weight: float = 1e-4,
                     kick_scale: float = 0.05, fixed_joint_mask=None, constraint_fns=(), constraint_args=(),
                     constraint_weights=None, return_details: bool = False):
    """``P`` independent HJCD-IK problems in one set of launches; returns ``(P, n_act)``.

    ``target_poses`` is an SE3 / wxyz_xyz array of shape ``(P, n_ee, 7)`` (or
    ``(P, 7)`` for one end effector); ``previous_cfgs`` is ``(P, n_act)``.
    ``eps_pos`` / ``eps_ori`` are accepted for signature parity: the reference
    runs every LM iteration regardless.
    """
    del eps_pos, eps_ori
    _unsupported_constraints(constraint_fns)
    if isinstance(target_link_indices, int):
        target_link_indices = (target_link_indices,)
    links = tuple(int(i) for i in target_link_indices)
    n_ee = len(links)
    tT = _target_array(target_poses).reshape(-1, n_ee, 7)
    P = tT.shape[0]
    A = robot.joints.num_actuated_joints
    prev = (previous_cfgs.detach().cpu().numpy() if isinstance(previous_cfgs, torch.Tensor)
            else np.asarray(previous_cfgs, np.float64)).reshape(P, A)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    fm = np.zeros(A, np.int32) if fixed_joint_mask is None else np.asarray(fixed_joint_mask).astype(np.int32)
    rng = rng_from_key(rng_key)
    top_k, repeats = _get_refine_schedule(num_seeds)
    seeds = hjcd_seeds(rng, prev, lo, hi, num_seeds, n_ee, fm.astype(bool))
    noise = rng.standard_normal((P, top_k * repeats, A)).astype(np.float32)
    kick_key = int(rng.integers(0, 2 ** 63))
    ee_of_seed = np.array([(s * n_ee) // num_seeds for s in range(num_seeds)], np.int32)

    from ..kernels.ops import hjcd_coarse, hjcd_refine_seeds, jax_ik_lm, ls_ik_select

    t = robot.device_tables()
    dev = t.twists.device
    tj, masks = ee_tables(robot, links)
    tabs = t.robot_buffers()
    tT_d = torch.as_tensor(tT.astype(np.float32), device=dev)
    coarse, cerr = hjcd_coarse(seeds, tabs, tj, masks, tT_d, ee_of_seed, t.lower, t.upper, fm, k_max=coarse_max_iter,
                               epsilon=epsilon, nu=nu)
    refine = hjcd_refine_seeds(coarse, cerr, top_k, repeats, noise, t.lower, t.upper, fm)
    cfgs, errs = jax_ik_lm(refine, tabs, tj, masks, tT_d, t.lower, t.upper, fm, max_iter=lm_max_iter,
                           lambda_init=lambda_init, adaptive=True, prior_weight=limit_prior_weight,
                           kick_scale=kick_scale, kick_key=kick_key)
    best, idx = ls_ik_select(cfgs, errs, torch.as_tensor(prev.astype(np.float32), device=dev), continuity_weight)
    if return_details:
        return best, dict(seeds=seeds, coarse=coarse, coarse_err=cerr, refine=refine, cfgs=cfgs, errs=errs,
                          noise=noise, kick_key=kick_key, ee_of_seed=ee_of_seed, top_k=top_k, repeats=repeats)
    return _like(previous_cfgs, best)


def hjcd_solve(robot, target_link_indices, target_poses, rng_key, previous_cfg, num_seeds: int = 32,
               coarse_max_iter: int = 20, lm_max_iter: int = 40, epsilon: float = 0.02,
               nu: float = float(np.pi / 2), eps_pos: float = 1e-8, eps_ori: float = 1e-8,
               lambda_init: float = 5e-3, continuity_weight: float = 0.0, limit_prior_weight: float = 1e-4,
               kick_scale: float = 0.05, fixed_joint_mask=None, constraint_fns=(), constraint_args=(),
               constraint_weights=None):
    """Solve one IK problem (one or more end effectors) with HJCD-IK; returns ``(n_act,)``."""
    if isinstance(target_link_indices, int):
        target_link_indices = (target_link_indices,)
    n_ee = len(target_link_indices)
    tT = _target_array(target_poses).reshape(1, n_ee, 7)
    prev = previous_cfg.detach().cpu().numpy() if isinstance(previous_cfg, torch.Tensor) else np.asarray(previous_cfg)
    best = hjcd_solve_batch(robot, target_link_indices, tT, rng_key, prev[None], num_seeds=num_seeds,
                            coarse_max_iter=coarse_max_iter, lm_max_iter=lm_max_iter, epsilon=epsilon, nu=nu,
                            lambda_init=lambda_init, continuity_weight=continuity_weight,
                            limit_prior_weight=limit_prior_weight, kick_scale=kick_scale,
                            fixed_joint_mask=fixed_joint_mask, constraint_fns=constraint_fns,
                            constraint_args=constraint_args, constraint_weights=constraint_weights)
    return _like(previous_cfg, torch.as_tensor(np.asarray(best))[0] if not isinstance(best, torch.Tensor) else best[0])
