"""FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

A double-precision restatement of the reference algorithms (``oracle.c``,
each function citing the reference file:line it follows), used as the checker
for the HIP kernels.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the
``pyroffi_amd`` package never does.

Each function is pinned by ``tests/test_oracle.py`` against checks that share
no code with it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
LIB_F32 = os.path.join(HERE, "_build", "liboracle_f32.so")

_R = np.float64  # the arithmetic type of the loaded build (see single_precision)
_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i = ctypes.POINTER(ctypes.c_int32)
_I = ctypes.c_int
_L = ctypes.c_int64
_D = ctypes.c_double


def _empty(shape):
    return np.empty(shape, _R)


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        h = ctypes.CDLL(LIB)
        h.or_fk.argtypes = [_d, _L, _I, _f, _f, _i, _i, _f, _f, _i, _i, _I, _d, _I]
        h.or_link_poses.argtypes = [_d, _L, _I, _i, _I, _d]
        h.or_world_dist.argtypes = [_d, _L, _I, _d, _I, _d, _I, _I, _d, _d, _I]
        h.or_self_dist.argtypes = [_d, _L, _I, _d, _I, _i, _i, _I, _I, _d, _I]
        h.or_ls_ik.argtypes = (
            [_d, _I, _I, _I, _f, _f, _i, _i, _f, _f, _i, _i, _I, _i, _i, _I, _d, _f, _i, _I, _f, _I,
             _d, _d, _i, _I, _D, _D, _D, _D, _D, _I, _D, _D, _d, _d, _I, _f, _f, _f, _i]
        )
        h.or_res_jac.argtypes = [_d, _I, _f, _f, _i, _i, _f, _f, _i, _i, _I, _i, _i, _I, _d, _d, _d]
        h.or_lm_direction.argtypes = [_d, _d, _I, _I, _d, _D, _i, _d]
        h.or_lm_direction.restype = ctypes.c_int
        h.or_pose_hinge_cost.argtypes = [_d, _L, _I, _f, _f, _i, _i, _f, _f, _i, _i, _I, _i, _I, _d, _I, _d, _I,
                                         _D, _I, _d, _D, _D, _d, _I]
        h.or_pose_hinge_cost.restype = None
        traj_model = [_f, _f, _i, _i, _f, _f, _i, _i, _I, _I, _i, _I, _d, _I, _i, _i, _I, _d, _I, _D, _I]
        h.or_traj_cost.argtypes = traj_model + [_d, ctypes.POINTER(ctypes.c_int64), _L, _d, _d, _d, _d, _d, _I]
        h.or_traj_lbfgs.argtypes = traj_model + [_d, ctypes.POINTER(ctypes.c_int64), _L, _d, _d, _i, _d, _d, _d,
                                                 _d, _d, _I]
        h.or_two_loop.argtypes = [_d, _d, _d, _d, _I, _I, _I, _I, _d]
        h.or_traj_chomp.argtypes = traj_model + [_d, ctypes.POINTER(ctypes.c_int64), _L, _d, _d, _i, _d, _d, _d,
                                                 _d, _I]
        h.or_traj_chomp.restype = None
        h.or_chomp_cg.argtypes = [_d, _I, _D, _I, _D, _d]
        h.or_sco_linearize.argtypes = traj_model + [_d, _L, _i, _I, _D, _d, _d, _I]
        h.or_sco_linearize.restype = None
        h.or_traj_sco.argtypes = traj_model + [_d, ctypes.POINTER(ctypes.c_int64), _L, _d, _d, _i, _I, _i, _d, _d,
                                               _d, _I]
        h.or_traj_sco.restype = None
        h.or_chomp_cg.restype = None
        h.or_obstacle_dist.argtypes = [_d, _d, _L, _d, _I, _d, _I, _d, _I, _d, _I, _d, _I]
        h.or_obstacle_dist.restype = None
        h.or_capsule_dist.argtypes = [_d, _L, _d, _I, _d, _I, _d, _I, _d, _I, _d, _I]
        h.or_capsule_dist.restype = None
        robot8 = [_f, _f, _i, _i, _f, _f, _i, _i, _I]
        h.or_jres.argtypes = [_d, _L, _I] + robot8 + [_i, _I, _d, _d, _d]
        h.or_jres.restype = None
        h.or_jlm.argtypes = [_d, _L, _I, _I] + robot8 + [_i, _I, _d, _d, _d, _i, _i, _d, ctypes.c_uint64, _d, _d,
                                                         _I]
        h.or_jlm.restype = None
        h.or_hjcd_coarse.argtypes = [_d, _L, _I, _I] + robot8 + [_i, _I, _d, _i, _d, _d, _i, _I, _D, _D, _d, _d, _I]
        h.or_hjcd_coarse.restype = None
        for fn in (h.or_fk, h.or_link_poses, h.or_world_dist, h.or_self_dist, h.or_ls_ik, h.or_res_jac,
                   h.or_traj_cost, h.or_traj_lbfgs, h.or_two_loop):
            fn.restype = None
        _lib = h
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a, ct):
    return a.ctypes.data_as(ct)


class Tables:
    """float32/int32 robot tables exactly as the kernels receive them."""

    def __init__(self, joints, links):
        self.twists = _c(joints.twists, np.float32)
        self.parent_tf = _c(joints.parent_transforms, np.float32)
        self.parent_idx = _c(joints.parent_indices, np.int32)
        self.act_idx = _c(joints.actuated_indices, np.int32)
        self.mimic_mul = _c(joints.mimic_multiplier, np.float32)
        self.mimic_off = _c(joints.mimic_offset, np.float32)
        self.mimic_act_idx = _c(joints.mimic_act_indices, np.int32)
        self.topo_inv = _c(joints._topo_sort_inv, np.int32)
        self.link_parent_joint = _c(links.parent_joint_indices, np.int32)
        self.J = int(joints.num_joints)
        self.A = int(joints.num_actuated_joints)
        self.L = int(links.num_links)

    @staticmethod
    def of(robot) -> "Tables":
        return Tables(robot.joints, robot.links)

    def ptrs(self):
        return (
            _p(self.twists, _f), _p(self.parent_tf, _f), _p(self.parent_idx, _i), _p(self.act_idx, _i),
            _p(self.mimic_mul, _f), _p(self.mimic_off, _f), _p(self.mimic_act_idx, _i), _p(self.topo_inv, _i),
        )


def fk_joints(tables: Tables, cfg, threads: int = 0) -> np.ndarray:
    """World pose of every joint, FP64, ``(*batch, J, 7)``."""
    cfg = _c(cfg, _R)
    batch = cfg.shape[:-1]
    flat = cfg.reshape(-1, tables.A)
    out = _empty((flat.shape[0], tables.J, 7))
    t = tables.ptrs()
    lib().or_fk(_p(flat, _d), flat.shape[0], tables.A, t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7],
                tables.J, _p(out, _d), threads)
    return out.reshape(*batch, tables.J, 7)


def fk_links(tables: Tables, cfg, threads: int = 0) -> np.ndarray:
    """World pose of every link, FP64, ``(*batch, L, 7)``."""
    jp = fk_joints(tables, cfg, threads)
    batch = jp.shape[:-2]
    flat = _c(jp.reshape(-1, tables.J, 7), _R)
    out = _empty((flat.shape[0], tables.L, 7))
    lib().or_link_poses(_p(flat, _d), flat.shape[0], tables.J, _p(tables.link_parent_joint, _i), tables.L,
                        _p(out, _d))
    return out.reshape(*batch, tables.L, 7)


def padded_spheres(coll) -> np.ndarray:
    """``(L, S, 4)`` link-frame sphere table with the model's -1e9 padding."""
    c = np.asarray(coll.coll.pose.translation(), dtype=_R)
    r = np.asarray(coll.coll.radius, dtype=_R)
    return _c(np.concatenate([c, r[..., None]], axis=-1), _R)


def world_dist(link_poses, spheres, world, bias: bool = False, want_full: bool = True, threads: int = 0):
    """``(out (B, L, M) | None, per-config min (B,))`` from FP64 link poses."""
    lp = _c(link_poses, _R)
    B, L, _ = lp.shape
    sp = _c(spheres, _R)
    S = sp.shape[1]
    w = _c(world, _R).reshape(-1, 4)
    out = _empty((B, L, w.shape[0])) if want_full else None
    mins = _empty(B)
    lib().or_world_dist(_p(lp, _d), B, L, _p(sp, _d), S, _p(w, _d), w.shape[0], int(bias),
                        None if out is None else _p(out, _d), _p(mins, _d), threads)
    return out, mins


def self_dist(link_poses, spheres, pair_i, pair_j, mode: int = 0, threads: int = 0) -> np.ndarray:
    """``(B, P)``; mode 0 = CUDA semantics (all sphere pairs), 1 = JAX diagonal pairs."""
    lp = _c(link_poses, _R)
    B, L, _ = lp.shape
    sp = _c(spheres, _R)
    pi = _c(pair_i, np.int32)
    pj = _c(pair_j, np.int32)
    out = _empty((B, pi.shape[0]))
    lib().or_self_dist(_p(lp, _d), B, L, _p(sp, _d), sp.shape[1], _p(pi, _i), _p(pj, _i), pi.shape[0], mode,
                       _p(out, _d), threads)
    return out


def ls_ik(tables: Tables, seeds, target_jnts, ancestor_masks, target_T, lower, upper, fixed_mask=None, *,
          max_iter=60, pos_weight=50.0, ori_weight=10.0, lambda_init=5e-3, eps_pos=1e-8, eps_ori=1e-8,
          robot_spheres=None, sphere_joint=None, world=None, enable_collision=False, collision_weight=1e4,
          collision_margin=0.02, threads: int = 0, world_capsules=None, world_boxes=None, world_halfspaces=None):
    """FP64 restatement of the LM kernel; returns ``(cfgs (P, S, A), errs (P, S))``."""
    sd = _c(seeds, _R)
    P, S, A = sd.shape
    tj = _c(target_jnts, np.int32).reshape(-1)
    n_ee = tj.shape[0]
    am = _c(ancestor_masks, np.int32).reshape(n_ee, tables.J)
    tT = _c(target_T, _R).reshape(P, n_ee, 7)
    rs = _c(np.zeros((0, 4)) if robot_spheres is None else robot_spheres, np.float32).reshape(-1, 4)
    rj = _c(np.zeros(0) if sphere_joint is None else sphere_joint, np.int32).reshape(-1)
    ws = _c(np.zeros((0, 4)) if world is None else world, np.float32).reshape(-1, 4)
    lo = _c(lower, _R)
    hi = _c(upper, _R)
    fm = _c(np.zeros(A) if fixed_mask is None else fixed_mask, np.int32)
    out = _empty((P, S, A))
    err = _empty((P, S))
    t = tables.ptrs()
    others = [_c(np.zeros((0, w)) if x is None else x, np.float32).reshape(-1, w)
              for x, w in ((world_capsules, 7), (world_boxes, 15), (world_halfspaces, 6))]
    counts = _c([o.shape[0] for o in others], np.int32)
    oth = [_p(o, _f) for o in others]
    lib().or_ls_ik(_p(sd, _d), P, S, A, *t, tables.J, _p(tj, _i), _p(am, _i), n_ee, _p(tT, _d), _p(rs, _f),
                   _p(rj, _i), rs.shape[0], _p(ws, _f), ws.shape[0], _p(lo, _d), _p(hi, _d), _p(fm, _i),
                   int(max_iter), pos_weight, ori_weight, lambda_init, eps_pos, eps_ori, int(bool(enable_collision)),
                   collision_weight, collision_margin, _p(out, _d), _p(err, _d), threads, *oth, _p(counts, _i))
    return out, err


def res_jac(tables: Tables, q, target_jnts, ancestor_masks, target_T):
    """Residual ``(6 n_ee,)`` and geometric Jacobian ``(6 n_ee, A)`` at one configuration."""
    qq = _c(q, _R)
    tj = _c(target_jnts, np.int32).reshape(-1)
    n_ee = tj.shape[0]
    am = _c(ancestor_masks, np.int32).reshape(n_ee, tables.J)
    tT = _c(target_T, _R).reshape(n_ee, 7)
    r = _empty(6 * n_ee)
    jm = _empty((6 * n_ee, tables.A))
    lib().or_res_jac(_p(qq, _d), tables.A, *tables.ptrs(), tables.J, _p(tj, _i), _p(am, _i), n_ee, _p(tT, _d),
                     _p(r, _d), _p(jm, _d))
    return r, jm


def lm_direction(r, jac, weights6, lam, fixed_mask=None):
    """One LM direction (before the trust region) and the Cholesky success flag."""
    rr = _c(r, _R)
    jm = _c(jac, _R)
    R, A = jm.shape
    w = _c(weights6, _R)
    fm = _c(np.zeros(A) if fixed_mask is None else fixed_mask, np.int32)
    out = _empty(A)
    ok = lib().or_lm_direction(_p(rr, _d), _p(jm, _d), R, A, _p(w, _d), float(lam), _p(fm, _i), _p(out, _d))
    return out, bool(ok)


def pose_hinge_cost(tables: Tables, cfg, spheres, world, margin, ee_link, targets, pos_weight=1.0, ori_weight=1.0,
                    threads: int = 0) -> np.ndarray:
    """FP64 pose + hinge cost ``(B,)``; ``spheres`` is the padded ``(L, S, 4)`` table."""
    q = _c(cfg, _R).reshape(-1, tables.A)
    sp = _c(spheres, _R)
    w = _c(world, _R).reshape(-1, 4)
    tg = _c(np.zeros((q.shape[0], 7)) if targets is None else targets, _R).reshape(-1, 7)
    out = _empty(q.shape[0])
    lib().or_pose_hinge_cost(_p(q, _d), q.shape[0], tables.A, *tables.ptrs(), tables.J,
                             _p(tables.link_parent_joint, _i), tables.L, _p(sp, _d), sp.shape[1], _p(w, _d),
                             w.shape[0], float(margin), int(ee_link), _p(tg, _d), float(pos_weight),
                             float(ori_weight), _p(out, _d), threads)
    return out


def pose_hinge_fd_grad(tables: Tables, cfg, spheres, world, margin, ee_link, targets, h=1e-4, **kw) -> np.ndarray:
    """FP64 central finite-difference gradient ``(B, A)`` of ``pose_hinge_cost``."""
    q = _c(cfg, _R).reshape(-1, tables.A)
    B, A = q.shape
    pert = np.repeat(q[:, None, None, :], 2, axis=1).repeat(A, axis=2)  # (B, 2, A, A)
    eye = np.eye(A) * h
    pert[:, 0] += eye
    pert[:, 1] -= eye
    tg = np.repeat(np.asarray(targets, _R).reshape(B, 1, 7), 2 * A, axis=1).reshape(-1, 7)
    c = pose_hinge_cost(tables, pert.reshape(-1, A), spheres, world, margin, ee_link, tg, **kw).reshape(B, 2, A)
    return (c[:, 0] - c[:, 1]) / (2 * h)


class TrajModel:
    """Robot + collision model of the trajectory cost (padded ``(L, S, 4)`` spheres, active pairs, world)."""

    def __init__(self, tables: Tables, spheres, pair_i, pair_j, world, margin: float, mode: int):
        self.t = tables
        self.sph = _c(spheres, _R)
        self.pi = _c(pair_i, np.int32)
        self.pj = _c(pair_j, np.int32)
        self.world = _c(world, _R).reshape(-1, 4)
        self.margin = float(margin)
        self.mode = int(mode)

    def args(self):
        t = self.t
        return (*t.ptrs(), t.J, t.A, _p(t.link_parent_joint, _i), t.L, _p(self.sph, _d), self.sph.shape[1],
                _p(self.pi, _i), _p(self.pj, _i), self.pi.shape[0], _p(self.world, _d), self.world.shape[0],
                self.margin, self.mode)


def _traj_flat(trajs):
    """List of (T_b, A) arrays or a (B, T, A) array -> flat (W, A) + int64 CSR starts (B + 1)."""
    if isinstance(trajs, np.ndarray) and trajs.ndim == 3:
        trajs = list(trajs)
    x = _c(np.concatenate([np.asarray(t, _R) for t in trajs], 0), _R)
    start = np.zeros(len(trajs) + 1, np.int64)
    start[1:] = np.cumsum([len(t) for t in trajs])
    return x, start


def traj_cost(model: TrajModel, trajs, lower, upper, weights, grad: bool = True, threads: int = 0):
    """``(comps (B, 3) = smooth, limits, coll; grad (W, A) of the weighted total or None)``.

    ``weights`` = (w_smooth, w_vel, w_acc, w_jerk, w_limits, w_coll).
    """
    x, start = _traj_flat(trajs)
    B = len(start) - 1
    comps = _empty((B, 3))
    g = np.empty_like(x) if grad else None
    lo, hi, w = _c(lower, _R), _c(upper, _R), _c(weights, _R)
    lib().or_traj_cost(*model.args(), _p(x, _d), start.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), B,
                       _p(lo, _d), _p(hi, _d), _p(w, _d), _p(comps, _d), None if g is None else _p(g, _d), threads)
    return comps, g


def traj_lbfgs(model: TrajModel, trajs, lower, upper, *, n_iters=100, m_lbfgs=6, escalation_interval=20,
               w_smooth=1.0, w_vel=1.0, w_acc=0.5, w_jerk=0.1, w_limits=1.0, w_collision=1.0,
               w_collision_max=100.0, penalty_scale=2.0, threads: int = 0, iterates: bool = False):
    """FP64 direct L-BFGS per trajectory: ``(best (W, A) flat, best tracked cost (B,), starts (B + 1))``.

    With ``iterates=True`` two more results follow: the last iterate ``(W, A)``
    and the per-iteration trace ``(B, n_iters, 3)`` = (cost of the current
    iterate at the current collision weight, accepted step index, accepted
    point's cost at the maximum collision weight).
    """
    x, start = _traj_flat(trajs)
    B = len(start) - 1
    best = np.empty_like(x)
    cost = _empty(B)
    last = np.empty_like(x) if iterates else None
    trace = _empty((B, max(n_iters, 0), 3)) if iterates else None
    cfg = _c([n_iters, m_lbfgs, escalation_interval], np.int32)
    fcfg = _c([w_smooth, w_vel, w_acc, w_jerk, w_limits, w_collision, w_collision_max, penalty_scale], _R)
    lo, hi = _c(lower, _R), _c(upper, _R)
    lib().or_traj_lbfgs(*model.args(), _p(x, _d), start.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), B,
                        _p(lo, _d), _p(hi, _d), _p(cfg, _i), _p(fcfg, _d), _p(best, _d), _p(cost, _d),
                        None if last is None else _p(last, _d), None if trace is None else _p(trace, _d), threads)
    if iterates:
        return best, cost, start, last, trace
    return best, cost, start


def traj_chomp(model: TrajModel, trajs, lower, upper, *, n_iters=100, step_size=0.05, w_smooth=1.0, w_acc=0.5,
               w_jerk=0.1, w_collision=3.0, w_collision_max=50.0, collision_penalty_scale=1.05, w_limits=1.0,
               use_covariant_update=True, smoothness_reg=1e-3, n_cg_iters=24, cg_tol=1e-4, grad_clip_norm=10.0,
               max_delta_per_step=0.05, early_stop_patience=15, min_cost_improve=1e-5, threads: int = 0,
               trace: bool = False):
    """FP64 CHOMP per trajectory: ``(final (W, A) flat, cost at w_collision_max (B,), starts (B + 1))``,
    plus the per-iteration trace ``(B, n_iters, 3)`` = (current cost, picked trial, accepted) with ``trace``."""
    x, start = _traj_flat(trajs)
    B = len(start) - 1
    out = np.empty_like(x)
    cost = _empty(B)
    tr = _empty((B, max(n_iters, 0), 3)) if trace else None
    icfg = _c([n_iters, int(bool(use_covariant_update)), n_cg_iters, early_stop_patience], np.int32)
    fcfg = _c([step_size, w_smooth, w_acc, w_jerk, w_collision, w_collision_max, collision_penalty_scale, w_limits,
               smoothness_reg, cg_tol, grad_clip_norm, max_delta_per_step, min_cost_improve], _R)
    lo, hi = _c(lower, _R), _c(upper, _R)
    lib().or_traj_chomp(*model.args(), _p(x, _d), start.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), B,
                        _p(lo, _d), _p(hi, _d), _p(icfg, _i), _p(fcfg, _d), _p(out, _d), _p(cost, _d),
                        None if tr is None else _p(tr, _d), threads)
    if trace:
        return out, cost, start, tr
    return out, cost, start


def sco_linearize(model: TrajModel, q, world_group_start, temperature: float = 0.05, threads: int = 0):
    """SCO's per-waypoint group smooth minima: ``(d (W, G), J (W, G, A))`` for configurations ``q (W, A)``;
    G = 1 (self pairs) + one per world group (CSR ``world_group_start`` over the model's world spheres)."""
    qq = _c(q, _R).reshape(-1, model.t.A)
    gs = _c(world_group_start, np.int32)
    G = gs.shape[0]  # 1 + n_world_groups = len(CSR)
    W = qq.shape[0]
    d = _empty((W, G))
    J = _empty((W, G, model.t.A))
    lib().or_sco_linearize(*model.args(), _p(qq, _d), W, _p(gs, _i), G - 1, float(temperature),
                           _p(d, _d), _p(J, _d), threads)
    return d, J


def traj_sco(model: TrajModel, trajs, lower, upper, world_group_start, *, n_outer_iters=10, n_inner_iters=30,
             m_lbfgs=6, w_smooth=1.0, w_vel=1.0, w_acc=0.5, w_jerk=0.1, w_collision=1.0, w_collision_max=100.0,
             penalty_scale=3.0, collision_margin=0.01, w_trust=0.5, w_limits=1.0, smooth_min_temperature=0.05,
             threads: int = 0):
    """FP64 SCO per trajectory: ``(final (W, A) flat, final cost at w_collision_max (B,), starts (B + 1))``.
    The model's margin/mode serve the final cost (mode 1, colldist)."""
    x, start = _traj_flat(trajs)
    B = len(start) - 1
    out = np.empty_like(x)
    cost = _empty(B)
    gs = _c(world_group_start, np.int32)
    icfg = _c([n_outer_iters, n_inner_iters, m_lbfgs], np.int32)
    fcfg = _c([w_smooth, w_vel, w_acc, w_jerk, w_limits, w_collision, w_collision_max, penalty_scale,
               collision_margin, w_trust, smooth_min_temperature], _R)
    lo, hi = _c(lower, _R), _c(upper, _R)
    lib().or_traj_sco(*model.args(), _p(x, _d), start.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), B,
                      _p(lo, _d), _p(hi, _d), _p(gs, _i), gs.shape[0] - 1, _p(icfg, _i), _p(fcfg, _d), _p(out, _d),
                      _p(cost, _d), threads)
    return out, cost, start


def chomp_cg(b, reg: float = 1e-3, n_iters: int = 24, tol: float = 1e-4):
    """Truncated CG of CHOMP's covariant step: ``x`` with ``M x ~= b`` for one joint's interior timeline."""
    bb = _c(b, _R).reshape(-1)
    out = _empty(bb.shape[0])
    lib().or_chomp_cg(_p(bb, _d), bb.shape[0], float(reg), int(n_iters), float(tol), _p(out, _d))
    return out


def two_loop(g, s_buf, y_buf, rho, m_used: int, newest: int):
    """L-BFGS direction -H g from a masked ring buffer (``(m, n)`` s and y)."""
    gg = _c(g, _R)
    sb, yb, r = _c(s_buf, _R), _c(y_buf, _R), _c(rho, _R)
    m, n = sb.shape
    out = _empty(n)
    lib().or_two_loop(_p(gg, _d), _p(sb, _d), _p(yb, _d), _p(r, _d), int(m_used), int(newest), m, n, _p(out, _d))
    return out


def obstacle_dist(centres, radii, spheres=None, capsules=None, boxes=None, halfspaces=None, threads: int = 0):
    """``(N, M)`` FP64 signed distances of spheres ``(N, 3)`` / ``(N,)`` to the packed world
    (spheres ``(Ms, 4)``, capsules ``(Mc, 7)``, boxes ``(Mb, 15)``, half-spaces ``(Mh, 6)``)."""
    c = _c(centres, _R).reshape(-1, 3)
    r = _c(radii, _R).reshape(-1)
    parts = [_c(np.zeros((0, w)) if a is None else a, _R).reshape(-1, w)
             for a, w in ((spheres, 4), (capsules, 7), (boxes, 15), (halfspaces, 6))]
    M = sum(p.shape[0] for p in parts)
    out = _empty((c.shape[0], M))
    args = []
    for p in parts:
        args += [_p(p, _d), p.shape[0]]
    lib().or_obstacle_dist(_p(c, _d), _p(r, _d), c.shape[0], *args, _p(out, _d), threads)
    return out


def jres(tables: Tables, cfg, target_jnts, target_T, jac: bool = False):
    """JAX residual ``log(T_ee^-1 T_tgt)`` per configuration: ``f (B, n_ee, 6)`` and,
    with ``jac``, the exact Jacobian ``(B, 6 n_ee, A)``.  ``target_T`` is ``(B, n_ee, 7)``."""
    q = _c(cfg, _R).reshape(-1, tables.A)
    B = q.shape[0]
    tj = _c(target_jnts, np.int32).reshape(-1)
    n_ee = tj.shape[0]
    tT = _c(np.broadcast_to(np.asarray(target_T, _R).reshape(-1, n_ee, 7), (B, n_ee, 7)), _R)
    f = _empty((B, n_ee, 6))
    J = _empty((B, 6 * n_ee, tables.A)) if jac else None
    t = tables.ptrs()
    lib().or_jres(_p(q, _d), B, tables.A, *t, tables.J, _p(tj, _i), n_ee, _p(tT, _d), _p(f, _d),
                  None if J is None else _p(J, _d))
    return (f, J) if jac else f


def jlm(tables: Tables, seeds, target_jnts, target_T, lower, upper, fixed_mask=None, *, max_iter=60,
        pos_weight=50.0, ori_weight=10.0, lambda_init=5e-3, adaptive=False, prior_weight=0.0, kick_scale=0.0,
        kick_key=0, threads: int = 0):
    """FP64 JAX-convention LM per seed: ``(cfgs (P, S, A), winner errors (P, S))``."""
    sd = _c(seeds, _R)
    P, S, A = sd.shape
    tj = _c(target_jnts, np.int32).reshape(-1)
    n_ee = tj.shape[0]
    tT = _c(target_T, _R).reshape(P, n_ee, 7)
    fm = _c(np.zeros(A) if fixed_mask is None else fixed_mask, np.int32)
    iopt = _c([max_iter, int(bool(adaptive))], np.int32)
    dopt = _c([pos_weight, ori_weight, lambda_init, prior_weight, kick_scale], _R)
    out = _empty((P, S, A))
    err = _empty((P, S))
    lo, hi = _c(lower, _R), _c(upper, _R)
    lib().or_jlm(_p(sd, _d), P, S, A, *tables.ptrs(), tables.J, _p(tj, _i), n_ee, _p(tT, _d), _p(lo, _d),
                 _p(hi, _d), _p(fm, _i), _p(iopt, _i), _p(dopt, _d), ctypes.c_uint64(int(kick_key)), _p(out, _d),
                 _p(err, _d), threads)
    return out, err


def hjcd_coarse(tables: Tables, seeds, target_jnts, target_T, ee_of_seed, lower, upper, fixed_mask=None, *,
                k_max=20, epsilon=0.02, nu=np.pi / 2, threads: int = 0):
    """FP64 HJCD coarse phase per seed: ``(cfgs (P, S, A), sum over EEs of |f|^2 (P, S))``."""
    sd = _c(seeds, _R)
    P, S, A = sd.shape
    tj = _c(target_jnts, np.int32).reshape(-1)
    n_ee = tj.shape[0]
    tT = _c(target_T, _R).reshape(P, n_ee, 7)
    es = _c(ee_of_seed, np.int32).reshape(S)
    fm = _c(np.zeros(A) if fixed_mask is None else fixed_mask, np.int32)
    out = _empty((P, S, A))
    err = _empty((P, S))
    lo, hi = _c(lower, _R), _c(upper, _R)
    lib().or_hjcd_coarse(_p(sd, _d), P, S, A, *tables.ptrs(), tables.J, _p(tj, _i), n_ee, _p(tT, _d), _p(es, _i),
                         _p(lo, _d), _p(hi, _d), _p(fm, _i), int(k_max), float(epsilon), float(nu), _p(out, _d),
                         _p(err, _d), threads)
    return out, err


def capsule_dist(caps, spheres=None, capsules=None, boxes=None, halfspaces=None, threads: int = 0):
    """``(N, M)`` signed distances of capsules ``(N, 7)`` ``[a, b, r]`` to the packed world."""
    c = _c(caps, _R).reshape(-1, 7)
    parts = [_c(np.zeros((0, w)) if a is None else a, _R).reshape(-1, w)
             for a, w in ((spheres, 4), (capsules, 7), (boxes, 15), (halfspaces, 6))]
    out = _empty((c.shape[0], sum(p.shape[0] for p in parts)))
    args = []
    for p in parts:
        args += [_p(p, _d), p.shape[0]]
    lib().or_capsule_dist(_p(c, _d), c.shape[0], *args, _p(out, _d), threads)
    return out


_F32 = None


def single_precision():
    """This module's functions bound to ``_build/liboracle_f32.so``: the same
    restatement compiled with every ``double`` replaced by ``float`` (and the
    type-generic math functions), for the CPU baseline's single-precision leg.
    Arrays in and out are float32."""
    global _F32
    if _F32 is None:
        import importlib.util

        spec = importlib.util.spec_from_file_location("oracle_f32", __file__)
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m._R = np.float32
        m._d = ctypes.POINTER(ctypes.c_float)
        m._D = ctypes.c_float
        m.LIB = m.LIB_F32
        m._lib = None
        m._F32 = m
        _F32 = m
    return _F32
