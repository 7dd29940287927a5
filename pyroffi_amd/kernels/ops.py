"""Tensor-level wrappers of the C ABI, mirroring the reference's FFI wrappers.

Each function takes torch tensors that live on the GPU (any float/int dtype is
cast on the device, like the reference's ``astype`` calls), allocates its
outputs with torch, and enqueues one library call on torch's current HIP
stream.  Names and argument order follow the reference:

* ``fk_cuda``                      -- ``src/pyroffi/cuda_kernels/_fk_cuda.py:52-126``
* ``collision_world_sphere_reduced`` -- ``src/pyroffi/cuda_kernels/_collision_cuda_ffi.py:109-151``
* ``collision_self_sphere``        -- ``src/pyroffi/cuda_kernels/_collision_cuda_ffi.py:191-216``
* ``ls_ik_cuda``                   -- ``src/pyroffi/cuda_kernels/_ls_ik_cuda.py:76-183``

``fk_links`` and ``sphere_sweep`` are the fused entry points the
higher-level API uses.
"""

from __future__ import annotations

import ctypes
import math

import torch

from ._lib import MAX_GRIDS, PkInvalidArgument, WorldGrid, check, lib


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(x, dtype: torch.dtype, device=None) -> torch.Tensor:
    """Contiguous device tensor of `dtype` (host inputs are uploaded)."""
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x)
    if not x.is_cuda:
        x = x.to(device if device is not None else "cuda")
    if x.dtype != dtype:
        x = x.to(dtype)
    return x.contiguous()


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _tables(twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx, topo_inv, device):
    f, i = torch.float32, torch.int32
    return (
        _dev(twists, f, device),
        _dev(parent_tf, f, device),
        _dev(parent_idx, i, device),
        _dev(act_idx, i, device),
        _dev(mimic_mul, f, device),
        _dev(mimic_off, f, device),
        _dev(mimic_act_idx, i, device),
        _dev(topo_inv, i, device),
    )


def fk_hip(cfg, twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx,
           topo_inv, fk_level_starts, fk_level_joints) -> torch.Tensor:
    """World pose of every joint, ``(*batch, n_joints, 7)`` wxyz_xyz."""
    cfg = _dev(cfg, torch.float32)
    dev = cfg.device
    batch_axes = tuple(cfg.shape[:-1])
    n_act = cfg.shape[-1]
    batch = math.prod(batch_axes)
    tb = _tables(twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx, topo_inv, dev)
    n_joints = tb[0].shape[0]
    starts = _dev(fk_level_starts, torch.int32, dev)
    joints = _dev(fk_level_joints, torch.int32, dev)
    out = torch.empty((batch, n_joints, 7), dtype=torch.float32, device=dev)
    st = lib().pk_fk_f32(
        _stream(), _ptr(cfg), batch, n_act, *[_ptr(t) for t in tb], _ptr(starts), _ptr(joints),
        n_joints, max(0, starts.shape[0] - 1), _ptr(out),
    )
    check(st, "pk_fk_f32")
    return out.reshape(*batch_axes, n_joints, 7)


fk_cuda = fk_hip


def fk_links(cfg, twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx,
             topo_inv, link_parent_joint) -> torch.Tensor:
    """World pose of every link, ``(*batch, n_links, 7)`` (FK fused with the link gather)."""
    cfg = _dev(cfg, torch.float32)
    dev = cfg.device
    batch_axes = tuple(cfg.shape[:-1])
    n_act = cfg.shape[-1]
    batch = math.prod(batch_axes)
    tb = _tables(twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx, topo_inv, dev)
    lpj = _dev(link_parent_joint, torch.int32, dev)
    n_links = lpj.shape[0]
    out = torch.empty((batch, n_links, 7), dtype=torch.float32, device=dev)
    st = lib().pk_fk_links_f32(
        _stream(), _ptr(cfg), batch, n_act, *[_ptr(t) for t in tb], tb[0].shape[0], _ptr(lpj),
        n_links, _ptr(out),
    )
    check(st, "pk_fk_links_f32")
    return out.reshape(*batch_axes, n_links, 7)


def fk_vjp(cfg, cot_links, twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx,
           topo_inv, link_parent_joint) -> torch.Tensor:
    """``(*batch, n_act)``: the FK adjoint, J^T of the link poses applied to ``cot_links``."""
    cfg = _dev(cfg, torch.float32)
    dev = cfg.device
    batch_axes = tuple(cfg.shape[:-1])
    n_act = cfg.shape[-1]
    batch = math.prod(batch_axes)
    tb = _tables(twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx, topo_inv, dev)
    lpj = _dev(link_parent_joint, torch.int32, dev)
    n_links = lpj.shape[0]
    cot = _dev(cot_links, torch.float32, dev)
    if tuple(cot.shape) != batch_axes + (n_links, 7):
        raise ValueError(f"cotangent shape {tuple(cot.shape)} != {batch_axes + (n_links, 7)}")
    out = torch.empty((batch, n_act), dtype=torch.float32, device=dev)
    st = lib().pk_fk_vjp_f32(
        _stream(), _ptr(cfg), batch, n_act, *[_ptr(t) for t in tb], tb[0].shape[0], _ptr(lpj),
        n_links, _ptr(cot), _ptr(out),
    )
    check(st, "pk_fk_vjp_f32")
    return out.reshape(*batch_axes, n_act)


def collision_world_sphere_reduced(sphere_centers, sphere_radii, world_spheres, world_capsules,
                                   world_boxes, world_halfspaces, n: int) -> torch.Tensor:
    """``[B, N, M]`` per-link world distances from SoA centres ``[3, B, K]``."""
    sc = _dev(sphere_centers, torch.float32)
    dev = sc.device
    sr = _dev(sphere_radii, torch.float32, dev)
    ws = _dev(world_spheres, torch.float32, dev).reshape(-1, 4)
    wc = _dev(world_capsules, torch.float32, dev).reshape(-1, 7)
    wb = _dev(world_boxes, torch.float32, dev).reshape(-1, 15)
    wh = _dev(world_halfspaces, torch.float32, dev).reshape(-1, 6)
    _, B, K = sc.shape
    M = ws.shape[0] + wc.shape[0] + wb.shape[0] + wh.shape[0]
    out = torch.empty((B, n, M), dtype=torch.float32, device=dev)
    st = lib().pk_collision_world_sphere_reduced_f32(
        _stream(), _ptr(sc), _ptr(sr), B, K, _ptr(ws), ws.shape[0], _ptr(wc), wc.shape[0],
        _ptr(wb), wb.shape[0], _ptr(wh), wh.shape[0], n, _ptr(out),
    )
    check(st, "pk_collision_world_sphere_reduced_f32")
    return out


def collision_world_sphere(sphere_centers, sphere_radii, world_spheres, world_capsules, world_boxes,
                           world_halfspaces) -> torch.Tensor:
    """``[B, K, M]`` per-sphere world distances from SoA centres ``[3, B, K]``
    (``M`` = spheres, capsules, boxes, half-spaces in that order)."""
    sc = _dev(sphere_centers, torch.float32)
    dev = sc.device
    sr = _dev(sphere_radii, torch.float32, dev)
    ws = _dev(world_spheres, torch.float32, dev).reshape(-1, 4)
    wc = _dev(world_capsules, torch.float32, dev).reshape(-1, 7)
    wb = _dev(world_boxes, torch.float32, dev).reshape(-1, 15)
    wh = _dev(world_halfspaces, torch.float32, dev).reshape(-1, 6)
    _, B, K = sc.shape
    M = ws.shape[0] + wc.shape[0] + wb.shape[0] + wh.shape[0]
    out = torch.empty((B, K, M), dtype=torch.float32, device=dev)
    st = lib().pk_collision_world_sphere_f32(
        _stream(), _ptr(sc), _ptr(sr), B, K, _ptr(ws), ws.shape[0], _ptr(wc), wc.shape[0],
        _ptr(wb), wb.shape[0], _ptr(wh), wh.shape[0], _ptr(out),
    )
    check(st, "pk_collision_world_sphere_f32")
    return out


def row_min(dist) -> torch.Tensor:
    """``(B,)`` minimum over everything after the first axis of ``dist`` (1e9 when empty)."""
    d = _dev(dist, torch.float32)
    B = d.shape[0]
    row = d[0].numel() if B else 0
    out = torch.empty((B,), dtype=torch.float32, device=d.device)
    check(lib().pk_row_min_f32(_stream(), _ptr(d), B, row, _ptr(out)), "pk_row_min_f32")
    return out


def capsules_place(link_poses, capsules_local, *, soa: bool = True, aos: bool = True):
    """Capsule robot at link poses ``(B, N, 7)``: ``(caps [7, B, N] | None, caps [B, N, 7] | None)``
    from the link-frame ``(N, 7)`` ``[a, b, r]`` table."""
    lp = _dev(link_poses, torch.float32)
    dev = lp.device
    B, N, _ = lp.shape
    cl = _dev(capsules_local, torch.float32, dev).reshape(N, 7)
    o_soa = torch.empty((7, B, N), dtype=torch.float32, device=dev) if soa else None
    o_aos = torch.empty((B, N, 7), dtype=torch.float32, device=dev) if aos else None
    st = lib().pk_capsules_place_f32(_stream(), _ptr(lp), B, N, _ptr(cl), _ptr(o_soa), _ptr(o_aos))
    check(st, "pk_capsules_place_f32")
    return o_soa, o_aos


def collision_world_capsule(caps, world_spheres, world_capsules, world_boxes, world_halfspaces) -> torch.Tensor:
    """``[B, N, M]`` capsule-robot world distances from component-major capsules ``[7, B, N]``."""
    c = _dev(caps, torch.float32)
    dev = c.device
    ws = _dev(world_spheres, torch.float32, dev).reshape(-1, 4)
    wc = _dev(world_capsules, torch.float32, dev).reshape(-1, 7)
    wb = _dev(world_boxes, torch.float32, dev).reshape(-1, 15)
    wh = _dev(world_halfspaces, torch.float32, dev).reshape(-1, 6)
    _, B, N = c.shape
    M = ws.shape[0] + wc.shape[0] + wb.shape[0] + wh.shape[0]
    out = torch.empty((B, N, M), dtype=torch.float32, device=dev)
    st = lib().pk_collision_world_capsule_f32(_stream(), _ptr(c), B, N, _ptr(ws), ws.shape[0], _ptr(wc), wc.shape[0],
                                              _ptr(wb), wb.shape[0], _ptr(wh), wh.shape[0], _ptr(out))
    check(st, "pk_collision_world_capsule_f32")
    return out


def collision_self_capsule(caps, pair_i, pair_j) -> torch.Tensor:
    """``[B, P]`` capsule-robot self distances from AoS capsules ``[B, N, 7]``."""
    c = _dev(caps, torch.float32)
    dev = c.device
    pi, pj = _dev(pair_i, torch.int32, dev), _dev(pair_j, torch.int32, dev)
    B, N, _ = c.shape
    out = torch.empty((B, pi.shape[0]), dtype=torch.float32, device=dev)
    st = lib().pk_collision_self_capsule_f32(_stream(), _ptr(c), B, N, _ptr(pi), _ptr(pj), pi.shape[0], _ptr(out))
    check(st, "pk_collision_self_capsule_f32")
    return out


def sphere_world_links(link_poses, spheres_padded, world: dict, *, per_link: bool = True):
    """World distances of a spherized robot over the four obstacle kinds with the CUDA
    kernels' semantics: ``(out (B, L, M), out_min (B,))`` from link poses ``(B, L, 7)``,
    the padded ``(L, S, 4)`` link-frame sphere table and ``world`` = packed arrays by kind
    (``spheres (Ms,4)``, ``capsules (Mc,7)``, ``boxes (Mb,15)``, ``halfspaces (Mh,6)``)."""
    lp = _dev(link_poses, torch.float32)
    dev = lp.device
    B, L, _ = lp.shape
    sp = _dev(spheres_padded, torch.float32, dev).reshape(L, -1, 4)
    w = {k: _dev(world[k], torch.float32, dev).reshape(-1, n)
         for k, n in (("spheres", 4), ("capsules", 7), ("boxes", 15), ("halfspaces", 6))}
    M = sum(v.shape[0] for v in w.values())
    out = torch.empty((B, L, M), dtype=torch.float32, device=dev)
    out_min = torch.empty((B,), dtype=torch.float32, device=dev)
    st = lib().pk_sphere_world_links_f32(
        _stream(), _ptr(lp), B, L, _ptr(sp), sp.shape[1],
        _ptr(w["spheres"]), w["spheres"].shape[0], _ptr(w["capsules"]), w["capsules"].shape[0],
        _ptr(w["boxes"]), w["boxes"].shape[0], _ptr(w["halfspaces"]), w["halfspaces"].shape[0],
        _ptr(out), _ptr(out_min),
    )
    check(st, "pk_sphere_world_links_f32")
    return out, out_min


def sphere_distances_jax(link_poses, spheres_padded, world_spheres, pair_i, pair_j, *, world: bool = True,
                         self_pairs: bool = True):
    """The JAX path's conventions: ``(out_world (B, L, M) | None, out_self (B, P) | None)``
    from link poses ``(B, L, 7)`` and the padded ``(L, S, 4)`` link-frame sphere table."""
    lp = _dev(link_poses, torch.float32)
    dev = lp.device
    B, L, _ = lp.shape
    sp = _dev(spheres_padded, torch.float32, dev).reshape(L, -1, 4)
    ws = _dev(world_spheres, torch.float32, dev).reshape(-1, 4)
    pi = _dev(pair_i, torch.int32, dev)
    pj = _dev(pair_j, torch.int32, dev)
    ow = torch.empty((B, L, ws.shape[0]), dtype=torch.float32, device=dev) if world else None
    osf = torch.empty((B, pi.shape[0]), dtype=torch.float32, device=dev) if self_pairs else None
    st = lib().pk_sphere_distances_jax_f32(
        _stream(), _ptr(lp), B, L, _ptr(sp), sp.shape[1], _ptr(ws), ws.shape[0], _ptr(pi), _ptr(pj),
        pi.shape[0], _ptr(ow), _ptr(osf),
    )
    check(st, "pk_sphere_distances_jax_f32")
    return ow, osf


def collision_self_sphere(centers, radii, pair_i, pair_j) -> torch.Tensor:
    """``[B, P]`` link-pair distances from AoS centres ``[B, S, N, 3]``."""
    c = _dev(centers, torch.float32)
    dev = c.device
    r = _dev(radii, torch.float32, dev)
    pi = _dev(pair_i, torch.int32, dev)
    pj = _dev(pair_j, torch.int32, dev)
    B, S, N, _ = c.shape
    P = pi.shape[0]
    out = torch.empty((B, P), dtype=torch.float32, device=dev)
    st = lib().pk_collision_self_sphere_f32(
        _stream(), _ptr(c), _ptr(r), B, S, N, _ptr(pi), _ptr(pj), P, _ptr(out)
    )
    check(st, "pk_collision_self_sphere_f32")
    return out


def sphere_sweep(link_poses, spheres_local, link_sphere_start, world_spheres, self_entries,
                 n_pairs: int, *, world: bool = False, world_min: bool = False,
                 self_pairs: bool = False):
    """Fused distances from link poses ``(B, L, 7)``.

    Returns ``(out_world (B, L, M) | None, out_world_min (B,) | None,
    out_self (B, P) | None)``.
    """
    lp = _dev(link_poses, torch.float32)
    dev = lp.device
    B, L, _ = lp.shape
    sl = _dev(spheres_local, torch.float32, dev).reshape(-1, 4)
    ls = _dev(link_sphere_start, torch.int32, dev)
    ws = _dev(world_spheres, torch.float32, dev).reshape(-1, 4)
    ents = _dev(self_entries, torch.int32, dev)
    M = ws.shape[0]
    ow = torch.empty((B, L, M), dtype=torch.float32, device=dev) if world else None
    om = torch.empty((B,), dtype=torch.float32, device=dev) if world_min else None
    osf = torch.empty((B, n_pairs), dtype=torch.float32, device=dev) if self_pairs else None
    st = lib().pk_sphere_sweep_f32(
        _stream(), _ptr(lp), B, L, _ptr(sl), _ptr(ls), sl.shape[0], _ptr(ws), M, _ptr(ents),
        ents.shape[0], n_pairs, _ptr(ow), _ptr(om), _ptr(osf),
    )
    check(st, "pk_sphere_sweep_f32")
    return ow, om, osf


def ls_ik_hip(seeds, twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx,
              topo_inv, target_jnts, ancestor_masks, target_T, robot_spheres_local,
              robot_sphere_joint_idx, world_spheres, world_capsules, world_boxes,
              world_halfspaces, lower, upper, fixed_mask, *, max_iter: int, pos_weight: float,
              ori_weight: float, lambda_init: float, eps_pos: float, eps_ori: float,
              enable_collision: bool = False, collision_weight: float = 0.0,
              collision_margin: float = 0.02):
    """Multi-seed LM; returns ``(cfgs (P, S, A), errs (P, S))``."""
    sd = _dev(seeds, torch.float32)
    dev = sd.device
    P, S, A = sd.shape
    tb = _tables(twists, parent_tf, parent_idx, act_idx, mimic_mul, mimic_off, mimic_act_idx, topo_inv, dev)
    tj = _dev(target_jnts, torch.int32, dev).reshape(-1)
    n_ee = tj.shape[0]
    am = _dev(ancestor_masks, torch.int32, dev).reshape(n_ee, -1)
    tT = _dev(target_T, torch.float32, dev).reshape(P, n_ee, 7)
    rs = _dev(robot_spheres_local, torch.float32, dev).reshape(-1, 4)
    rj = _dev(robot_sphere_joint_idx, torch.int32, dev).reshape(-1)
    ws = _dev(world_spheres, torch.float32, dev).reshape(-1, 4)
    wc = _dev(world_capsules, torch.float32, dev).reshape(-1, 7)
    wb = _dev(world_boxes, torch.float32, dev).reshape(-1, 15)
    wh = _dev(world_halfspaces, torch.float32, dev).reshape(-1, 6)
    lo = _dev(lower, torch.float32, dev)
    hi = _dev(upper, torch.float32, dev)
    fm = _dev(fixed_mask, torch.int32, dev)
    out = torch.empty((P, S, A), dtype=torch.float32, device=dev)
    err = torch.empty((P, S), dtype=torch.float32, device=dev)
    st = lib().pk_ls_ik_f32(
        _stream(), _ptr(sd), P, S, A, *[_ptr(t) for t in tb], tb[0].shape[0],
        _ptr(tj), _ptr(am), n_ee, _ptr(tT), _ptr(rs), _ptr(rj), rs.shape[0],
        _ptr(ws), ws.shape[0], _ptr(wc), wc.shape[0], _ptr(wb), wb.shape[0], _ptr(wh), wh.shape[0],
        _ptr(lo), _ptr(hi), _ptr(fm), int(max_iter), float(pos_weight), float(ori_weight),
        float(lambda_init), float(eps_pos), float(eps_ori), int(bool(enable_collision)),
        float(collision_weight), float(collision_margin), _ptr(out), _ptr(err),
    )
    check(st, "pk_ls_ik_f32")
    return out, err


ls_ik_cuda = ls_ik_hip


def jax_ik_lm(seeds, tables, target_jnts, ancestor_masks, target_T, lower, upper, fixed_mask, *, max_iter: int,
              pos_weight: float = 50.0, ori_weight: float = 10.0, lambda_init: float = 5e-3, adaptive: bool = False,
              prior_weight: float = 0.0, kick_scale: float = 0.0, kick_key: int = 0):
    """JAX-convention multi-seed LM (``pk_ik_jax_lm_f32``): ``(best cfgs (P, S, A), winner errors (P, S))``.

    ``tables`` is the robot's 8 FFI buffers (``_DeviceTables.robot_buffers()``).
    """
    sd = _dev(seeds, torch.float32)
    dev = sd.device
    P, S, A = sd.shape
    tb = _tables(*tables, dev)
    tj = _dev(target_jnts, torch.int32, dev).reshape(-1)
    n_ee = tj.shape[0]
    am = _dev(ancestor_masks, torch.int32, dev).reshape(n_ee, -1)
    tT = _dev(target_T, torch.float32, dev).reshape(P, n_ee, 7)
    lo, hi = _dev(lower, torch.float32, dev), _dev(upper, torch.float32, dev)
    fm = _dev(fixed_mask, torch.int32, dev)
    out = torch.empty((P, S, A), dtype=torch.float32, device=dev)
    err = torch.empty((P, S), dtype=torch.float32, device=dev)
    st = lib().pk_ik_jax_lm_f32(
        _stream(), _ptr(sd), P, S, A, *[_ptr(t) for t in tb], tb[0].shape[0], _ptr(tj), _ptr(am), n_ee, _ptr(tT),
        _ptr(lo), _ptr(hi), _ptr(fm), int(max_iter), float(pos_weight), float(ori_weight), float(lambda_init),
        int(bool(adaptive)), float(prior_weight), float(kick_scale), ctypes.c_uint64(int(kick_key) & (2 ** 64 - 1)),
        _ptr(out), _ptr(err),
    )
    check(st, "pk_ik_jax_lm_f32")
    return out, err


def hjcd_coarse(seeds, tables, target_jnts, ancestor_masks, target_T, ee_of_seed, lower, upper, fixed_mask, *,
                k_max: int = 20, epsilon: float = 0.02, nu: float = math.pi / 2):
    """HJCD coarse coordinate descent (``pk_hjcd_coarse_f32``): ``(cfgs (P, S, A), sum |f|^2 (P, S))``."""
    sd = _dev(seeds, torch.float32)
    dev = sd.device
    P, S, A = sd.shape
    tb = _tables(*tables, dev)
    tj = _dev(target_jnts, torch.int32, dev).reshape(-1)
    n_ee = tj.shape[0]
    am = _dev(ancestor_masks, torch.int32, dev).reshape(n_ee, -1)
    tT = _dev(target_T, torch.float32, dev).reshape(P, n_ee, 7)
    es = _dev(ee_of_seed, torch.int32, dev).reshape(S)
    lo, hi = _dev(lower, torch.float32, dev), _dev(upper, torch.float32, dev)
    fm = _dev(fixed_mask, torch.int32, dev)
    out = torch.empty((P, S, A), dtype=torch.float32, device=dev)
    err = torch.empty((P, S), dtype=torch.float32, device=dev)
    st = lib().pk_hjcd_coarse_f32(
        _stream(), _ptr(sd), P, S, A, *[_ptr(t) for t in tb], tb[0].shape[0], _ptr(tj), _ptr(am), n_ee, _ptr(tT),
        _ptr(es), _ptr(lo), _ptr(hi), _ptr(fm), int(k_max), float(epsilon), float(nu), _ptr(out), _ptr(err),
    )
    check(st, "pk_hjcd_coarse_f32")
    return out, err


def hjcd_refine_seeds(coarse_cfgs, coarse_err, top_k: int, repeats: int, noise, lower, upper, fixed_mask):
    """Top-k coarse results tiled ``repeats`` times with perturbations: ``(P, top_k * repeats, A)``."""
    c = _dev(coarse_cfgs, torch.float32)
    dev = c.device
    P, S, A = c.shape
    e = _dev(coarse_err, torch.float32, dev)
    R = top_k * repeats
    nz = _dev(noise, torch.float32, dev).reshape(P, R, A)
    lo, hi = _dev(lower, torch.float32, dev), _dev(upper, torch.float32, dev)
    fm = _dev(fixed_mask, torch.int32, dev)
    out = torch.empty((P, R, A), dtype=torch.float32, device=dev)
    st = lib().pk_hjcd_refine_seeds_f32(_stream(), _ptr(c), _ptr(e), P, S, A, int(top_k), int(repeats), _ptr(nz),
                                        _ptr(lo), _ptr(hi), _ptr(fm), _ptr(out))
    check(st, "pk_hjcd_refine_seeds_f32")
    return out


def ls_ik_select(cfgs, errs, previous, continuity_weight: float = 0.0):
    """Per-problem winner ``(best (P, A), index (P,))``."""
    c = _dev(cfgs, torch.float32)
    dev = c.device
    e = _dev(errs, torch.float32, dev)
    P, S, A = c.shape
    prev = _dev(previous, torch.float32, dev).reshape(P, A)
    best = torch.empty((P, A), dtype=torch.float32, device=dev)
    idx = torch.empty((P,), dtype=torch.int32, device=dev)
    st = lib().pk_ls_ik_select_f32(
        _stream(), _ptr(c), _ptr(e), _ptr(prev), P, S, A, float(continuity_weight), _ptr(best), _ptr(idx)
    )
    check(st, "pk_ls_ik_select_f32")
    return best, idx


def pose_hinge_cost(link_poses, spheres_local, link_sphere_start, world_spheres, margin: float,
                    ee_link: int, targets, twists, parent_idx, act_idx, mimic_act_idx, mimic_mul,
                    topo_inv, link_parent_joint, n_act: int, *, pos_weight: float = 1.0,
                    ori_weight: float = 1.0, want_grad: bool = True):
    """Cost ``(B,)`` and, when ``want_grad``, ``dcost/dq (B, n_act)`` from link poses."""
    lp = _dev(link_poses, torch.float32)
    dev = lp.device
    B, L, _ = lp.shape
    sl = _dev(spheres_local, torch.float32, dev).reshape(-1, 4)
    ls = _dev(link_sphere_start, torch.int32, dev)
    ws = _dev(world_spheres, torch.float32, dev).reshape(-1, 4)
    tg = _dev(targets, torch.float32, dev).reshape(-1, 7) if ee_link >= 0 else None
    tw = _dev(twists, torch.float32, dev)
    tabs = [_dev(x, torch.int32, dev) for x in (parent_idx, act_idx, mimic_act_idx)]
    mm = _dev(mimic_mul, torch.float32, dev)
    ti = _dev(topo_inv, torch.int32, dev)
    lpj = _dev(link_parent_joint, torch.int32, dev)
    cost = torch.empty((B,), dtype=torch.float32, device=dev)
    grad = torch.empty((B, n_act), dtype=torch.float32, device=dev) if want_grad else None
    st = lib().pk_pose_hinge_cost_f32(
        _stream(), _ptr(lp), B, L, _ptr(sl), _ptr(ls), sl.shape[0], _ptr(ws), ws.shape[0], float(margin),
        int(ee_link), _ptr(tg), float(pos_weight), float(ori_weight), _ptr(tw), _ptr(tabs[0]), _ptr(tabs[1]),
        _ptr(tabs[2]), _ptr(mm), _ptr(ti), _ptr(lpj), tw.shape[0], int(n_act), _ptr(cost), _ptr(grad),
    )
    check(st, "pk_pose_hinge_cost_f32")
    return cost, grad


def world_grid(world_spheres, pad: float, cell: float | None = None, max_cells: int = 48):
    """Build the exact-culling CSR grid of a sphere world (``pk_world_grid_count/fill_f32``).

    ``pad`` = the robot sphere radius the grid serves + cost margin (kernels
    test larger robot spheres against every world sphere).  The grid box is
    the world's bounding box grown by ``pad``; ``cell`` defaults to the box's
    longest side / ``max_cells``.  Returns a dict with ``cells (n_cells, 2)``
    (start, count), ``items`` (world sphere indices) and the grid geometry.
    """
    import numpy as np

    ws = _dev(world_spheres, torch.float32).reshape(-1, 4)
    dev = ws.device
    w = ws.detach().cpu().numpy().astype(np.float64)
    if w.shape[0] == 0:
        lo, hi = np.zeros(3), np.ones(3)
    else:
        lo = (w[:, :3] - w[:, 3:4]).min(0) - pad
        hi = (w[:, :3] + w[:, 3:4]).max(0) + pad
    if cell is None:
        cell = float(max(hi - lo)) / max_cells
    dims = np.maximum(1, np.ceil((hi - lo) / cell).astype(np.int64))
    n_cells = int(np.prod(dims))
    geo = (float(pad), float(lo[0]), float(lo[1]), float(lo[2]), float(cell), int(dims[0]), int(dims[1]),
           int(dims[2]))
    cells = torch.zeros((n_cells, 2), dtype=torch.int32, device=dev)
    check(lib().pk_world_grid_count_f32(_stream(), _ptr(ws), ws.shape[0], *geo, _ptr(cells)),
          "pk_world_grid_count_f32")
    counts = cells[:, 1].to(torch.int64)
    starts = torch.cumsum(counts, 0) - counts
    cells[:, 0] = starts.to(torch.int32)
    n_items = int(counts.sum().item())
    items = torch.empty((max(n_items, 1),), dtype=torch.int32, device=dev)
    check(lib().pk_world_grid_fill_f32(_stream(), _ptr(ws), ws.shape[0], *geo, _ptr(cells), _ptr(items), n_items),
          "pk_world_grid_fill_f32")
    return {"cells": cells, "items": items, "pad": float(pad), "lo": tuple(float(v) for v in lo),
            "cell": float(cell), "dims": tuple(int(v) for v in dims)}


def _grid_array(grid):
    """``pk_world_grid[]`` from None, one ``world_grid`` dict or a list of them (increasing pad)."""
    levels = [] if grid is None else ([grid] if isinstance(grid, dict) else list(grid))
    if len(levels) > MAX_GRIDS:
        raise PkInvalidArgument(f"at most {MAX_GRIDS} grids, got {len(levels)}")
    arr = (WorldGrid * max(1, len(levels)))()
    for i, g in enumerate(levels):
        arr[i].cells = g["cells"].data_ptr()
        arr[i].items = g["items"].data_ptr()
        arr[i].lo[:] = [float(v) for v in g["lo"]]
        arr[i].cell = float(g["cell"])
        arr[i].dims[:] = [int(v) for v in g["dims"]]
        arr[i].pad = float(g["pad"])
    return (arr if levels else None), len(levels)


def ls_ik_batched(seeds, problem_seed_start, tables, target_jnts, ancestor_masks, target_T,
                  robot_spheres_local, robot_sphere_joint_idx, world_spheres, lower, upper, fixed_mask, *,
                  max_iter: int, iter_begin: int = 0, iter_end: int | None = None, state=None,
                  grid: dict | None = None, pos_weight: float = 50.0, ori_weight: float = 10.0,
                  lambda_init: float = 5e-3, eps_pos: float = 1e-8, eps_ori: float = 1e-8,
                  enable_collision: bool = False, collision_weight: float = 1e4,
                  collision_margin: float = 0.02, grid_waves: int = 0, out_cfgs=None, out_err=None, work=None):
    """Uneven batched LM over flat seeds ``(S_total, A)``; ``tables`` = the 8 robot buffers.

    Writes into (or allocates) ``out_cfgs (S_total, A)`` and ``out_err (S_total,)``.
    ``work`` (a one-element int64 device tensor, or None) accumulates the
    number of sphere-pair tests the collision penalty executed.
    """
    sd = _dev(seeds, torch.float32)
    dev = sd.device
    S_total, A = sd.shape
    ss = _dev(problem_seed_start, torch.int32, dev)
    P = ss.shape[0] - 1
    tb = [_dev(x, torch.float32 if x.dtype.is_floating_point else torch.int32, dev) for x in tables]
    tj = _dev(target_jnts, torch.int32, dev).reshape(-1)
    n_ee = tj.shape[0]
    am = _dev(ancestor_masks, torch.int32, dev).reshape(n_ee, -1)
    tT = _dev(target_T, torch.float32, dev).reshape(P, n_ee, 7)
    rs = _dev(robot_spheres_local, torch.float32, dev).reshape(-1, 4)
    rj = _dev(robot_sphere_joint_idx, torch.int32, dev).reshape(-1)
    ws = _dev(world_spheres, torch.float32, dev).reshape(-1, 4)
    lo = _dev(lower, torch.float32, dev)
    hi = _dev(upper, torch.float32, dev)
    fm = _dev(fixed_mask, torch.int32, dev)
    if out_cfgs is None:
        out_cfgs = torch.empty((S_total, A), dtype=torch.float32, device=dev)
    if out_err is None:
        out_err = torch.empty((S_total,), dtype=torch.float32, device=dev)
    iter_end = max_iter if iter_end is None else iter_end
    grids = _grid_array(grid)
    st = lib().pk_ls_ik_batched_f32(
        _stream(), _ptr(sd), _ptr(ss), P, S_total, A, *[_ptr(t) for t in tb], tb[0].shape[0],
        _ptr(tj), _ptr(am), n_ee, _ptr(tT), _ptr(rs), _ptr(rj), rs.shape[0], _ptr(ws), ws.shape[0],
        grids[0], grids[1], _ptr(lo), _ptr(hi), _ptr(fm), int(max_iter),
        int(iter_begin), int(iter_end), _ptr(state), float(pos_weight), float(ori_weight), float(lambda_init),
        float(eps_pos), float(eps_ori), int(bool(enable_collision)), float(collision_weight),
        float(collision_margin), int(grid_waves), _ptr(out_cfgs), _ptr(out_err), _ptr(work),
    )
    check(st, "pk_ls_ik_batched_f32")
    return out_cfgs, out_err


def ls_ik_select_segments(cfgs, errs, problem_seed_start, previous=None, continuity_weight: float = 0.0,
                          out_best=None, out_idx=None):
    """Per-problem winner over CSR seed ranges: ``(best (P, A), index within the problem (P,))``."""
    c = _dev(cfgs, torch.float32)
    dev = c.device
    e = _dev(errs, torch.float32, dev)
    ss = _dev(problem_seed_start, torch.int32, dev)
    P, A = ss.shape[0] - 1, c.shape[1]
    prev = None if previous is None else _dev(previous, torch.float32, dev).reshape(P, A)
    if out_best is None:
        out_best = torch.empty((P, A), dtype=torch.float32, device=dev)
    if out_idx is None:
        out_idx = torch.empty((P,), dtype=torch.int32, device=dev)
    st = lib().pk_ls_ik_select_segments_f32(
        _stream(), _ptr(c), _ptr(e), _ptr(prev), _ptr(ss), P, A, float(continuity_weight), _ptr(out_best),
        _ptr(out_idx),
    )
    check(st, "pk_ls_ik_select_segments_f32")
    return out_best, out_idx
