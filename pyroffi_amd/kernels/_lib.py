"""ctypes binding of ``libpyroffi_amd.so`` (declared in ``include/pyroffi_amd.h``).

This is the ROCm counterpart of the reference's loader
(``src/pyroffi/cuda_kernels/_fk_cuda.py:26-49``): the reference opens its
``.so`` with ``ctypes.CDLL`` and hands the handler symbol to XLA; here the
same library is called directly with device pointers taken from torch tensors
and the current HIP stream.  A missing library raises ``RuntimeError`` -- there
is no fallback implementation.
"""

from __future__ import annotations

import ctypes
import functools
import os

_LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libpyroffi_amd.so")

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float



class WorldGrid(ctypes.Structure):
    """``pk_world_grid``: one world grid (device arrays + geometry), passed by host pointer."""

    _fields_ = [
        ("cells", ctypes.c_void_p),
        ("items", ctypes.c_void_p),
        ("lo", ctypes.c_float * 3),
        ("cell", ctypes.c_float),
        ("dims", ctypes.c_int32 * 3),
        ("pad", ctypes.c_float),
    ]


MAX_GRIDS = 2


class TrajModel(ctypes.Structure):
    """``pk_traj_model``: device pointers + sizes of a trajectory model."""

    _fields_ = [
        ("nodes", ctypes.c_void_p),
        ("n_nodes", ctypes.c_int32),
        ("n_act", ctypes.c_int32),
        ("link_node", ctypes.c_void_p),
        ("link_sph_start", ctypes.c_void_p),
        ("spheres", ctypes.c_void_p),
        ("link_bound", ctypes.c_void_p),
        ("n_links", ctypes.c_int32),
        ("n_spheres", ctypes.c_int32),
        ("pair_links", ctypes.c_void_p),
        ("pair_start", ctypes.c_void_p),
        ("pair_items", ctypes.c_void_p),
        ("n_pairs", ctypes.c_int32),
        ("n_items", ctypes.c_int32),
        ("world", ctypes.c_void_p),
        ("n_world", ctypes.c_int32),
        ("grids", ctypes.c_void_p),
        ("n_grids", ctypes.c_int32),
        ("lower", ctypes.c_void_p),
        ("upper", ctypes.c_void_p),
        ("margin", ctypes.c_float),
        ("mode", ctypes.c_int32),
    ]


class ChompConfig(ctypes.Structure):
    """``pk_chomp_config``."""

    _fields_ = [(n, ctypes.c_int32) for n in ("n_iters", "use_covariant_update", "n_cg_iters",
                                               "early_stop_patience")] + \
               [(n, ctypes.c_float) for n in ("step_size", "w_smooth", "w_acc", "w_jerk", "w_collision",
                                               "w_collision_max", "collision_penalty_scale", "w_limits",
                                               "smoothness_reg", "cg_tol", "grad_clip_norm", "max_delta_per_step",
                                               "min_cost_improve")]


class ScoConfig(ctypes.Structure):
    """``pk_sco_config``."""

    _fields_ = [(n, ctypes.c_int32) for n in ("n_outer_iters", "n_inner_iters", "m_lbfgs")] + \
               [(n, ctypes.c_float) for n in ("w_smooth", "w_vel", "w_acc", "w_jerk", "w_collision",
                                               "w_collision_max", "penalty_scale", "collision_margin", "w_trust",
                                               "w_limits", "smooth_min_temperature")]


class LbfgsConfig(ctypes.Structure):
    """``pk_lbfgs_config``."""

    _fields_ = [
        ("n_iters", ctypes.c_int32),
        ("m_hist", ctypes.c_int32),
        ("escalation_interval", ctypes.c_int32),
        ("w_smooth", ctypes.c_float),
        ("w_vel", ctypes.c_float),
        ("w_acc", ctypes.c_float),
        ("w_jerk", ctypes.c_float),
        ("w_limits", ctypes.c_float),
        ("w_collision", ctypes.c_float),
        ("w_collision_max", ctypes.c_float),
        ("penalty_scale", ctypes.c_float),
    ]

_SIGNATURES = {
    "pk_last_error": (ctypes.c_char_p, []),
    "pk_abi_version": (_i32, []),
    "pk_fk_f32": (_i32, [_p, _p, _i64, _i32] + [_p] * 10 + [_i32, _i32, _p]),
    "pk_fk_links_f32": (_i32, [_p, _p, _i64, _i32] + [_p] * 8 + [_i32, _p, _i32, _p]),
    "pk_fk_vjp_f32": (_i32, [_p, _p, _i64, _i32] + [_p] * 8 + [_i32, _p, _i32, _p, _p]),
    "pk_collision_world_sphere_reduced_f32": (
        _i32,
        [_p, _p, _p, _i64, _i32, _p, _i32, _p, _i32, _p, _i32, _p, _i32, _i32, _p],
    ),
    "pk_collision_world_sphere_f32": (
        _i32,
        [_p, _p, _p, _i64, _i32, _p, _i32, _p, _i32, _p, _i32, _p, _i32, _p],
    ),
    "pk_collision_self_sphere_f32": (_i32, [_p, _p, _p, _i64, _i32, _i32, _p, _p, _i32, _p]),
    "pk_sphere_distances_jax_f32": (_i32, [_p, _p, _i64, _i32, _p, _i32, _p, _i32, _p, _p, _i32, _p, _p]),
    "pk_collision_world_capsule_f32": (_i32, [_p, _p, _i64, _i32, _p, _i32, _p, _i32, _p, _i32, _p, _i32, _p]),
    "pk_collision_self_capsule_f32": (_i32, [_p, _p, _i64, _i32, _p, _p, _i32, _p]),
    "pk_capsules_place_f32": (_i32, [_p, _p, _i64, _i32, _p, _p, _p]),
    "pk_row_min_f32": (_i32, [_p, _p, _i64, _i64, _p]),
    "pk_sphere_world_links_f32": (
        _i32,
        [_p, _p, _i64, _i32, _p, _i32, _p, _i32, _p, _i32, _p, _i32, _p, _i32, _p, _p],
    ),
    "pk_sphere_sweep_f32": (
        _i32,
        [_p, _p, _i64, _i32, _p, _p, _i32, _p, _i32, _p, _i32, _i32, _p, _p, _p],
    ),
    "pk_ls_ik_f32": (
        _i32,
        [_p, _p, _i32, _i32, _i32]  # stream, seeds, P, S, A
        + [_p] * 8 + [_i32]  # robot tables, n_joints
        + [_p, _p, _i32, _p]  # target_jnts, ancestor_masks, n_ee, target_T
        + [_p, _p, _i32]  # robot spheres
        + [_p, _i32, _p, _i32, _p, _i32, _p, _i32]  # world primitives
        + [_p, _p, _p]  # lower, upper, fixed_mask
        + [_i32, _f32, _f32, _f32, _f32, _f32, _i32, _f32, _f32]  # attributes
        + [_p, _p],  # outputs
    ),
    "pk_ls_ik_select_f32": (_i32, [_p, _p, _p, _p, _i32, _i32, _i32, _f32, _p, _p]),
    "pk_ik_jax_lm_f32": (
        _i32,
        [_p, _p, _i32, _i32, _i32]  # stream, seeds, P, S, A
        + [_p] * 8 + [_i32]  # robot tables, n_joints
        + [_p, _p, _i32, _p]  # target_jnts, ancestor_masks, n_ee, target_T
        + [_p, _p, _p]  # lower, upper, fixed_mask
        + [_i32, _f32, _f32, _f32, _i32, _f32, _f32, ctypes.c_uint64]  # max_iter .. kick_key
        + [_p, _p],
    ),
    "pk_hjcd_coarse_f32": (
        _i32,
        [_p, _p, _i32, _i32, _i32] + [_p] * 8 + [_i32] + [_p, _p, _i32, _p, _p] + [_p, _p, _p]
        + [_i32, _f32, _f32, _p, _p],
    ),
    "pk_hjcd_refine_seeds_f32": (_i32, [_p, _p, _p, _i32, _i32, _i32, _i32, _i32, _p, _p, _p, _p, _p]),
    "pk_world_grid_count_f32": (_i32, [_p, _p, _i32, _f32, _f32, _f32, _f32, _f32, _i32, _i32, _i32, _p]),
    "pk_world_grid_fill_f32": (_i32, [_p, _p, _i32, _f32, _f32, _f32, _f32, _f32, _i32, _i32, _i32, _p, _p, _i64]),
    "pk_ls_ik_batched_f32": (
        _i32,
        [_p, _p, _p, _i32, _i64, _i32]  # stream, seeds, seed_start, P, total, A
        + [_p] * 8 + [_i32]  # robot tables, n_joints
        + [_p, _p, _i32, _p]  # targets
        + [_p, _p, _i32]  # robot spheres
        + [_p, _i32]  # world
        + [_p, _i32]  # grids (pk_world_grid*), n_grids
        + [_p, _p, _p]  # limits
        + [_i32, _i32, _i32, _p]  # max_iter, iter range, state
        + [_f32, _f32, _f32, _f32, _f32, _i32, _f32, _f32, _i32]  # attrs, grid_waves
        + [_p, _p, _p],  # outputs, work counter
    ),
    "pk_traj_cost_grad_f32": (_i32, [_p, _p, _p, _p, _i32, _i64, _i32, _i32, _p, _i32, _p, _p, _p]),
    "pk_traj_lbfgs_state_floats": (_i64, [_i64, _i32, _i32, _i32]),
    "pk_traj_lbfgs_f32": (_i32, [_p, _p, _p, _p, _i32, _i64, _i32, _i32, _p, _i32, _p, _i32, _i32, _p, _p, _p]),
    "pk_traj_sco_linearize_f32": (_i32, [_p, _p, _p, _p, _i32, _i64, _i32, _i32, _p, _i32, _f32, _p, _p]),
    "pk_traj_sco_inner_f32": (_i32, [_p, _p, _p, _i32, _i64, _i32, _i32, _i32, _p, _p, _p, _p, _p, _f32, _p]),
    "pk_traj_chomp_f32": (_i32, [_p, _p, _p, _p, _i32, _i64, _i32, _i32, _p, _p, _p, _p]),
    "pk_traj_lbfgs_noalloc_f32": (_i32, [_p, _p, _p, _p, _i32, _i64, _i32, _i32, _p, _i32, _p, _i32, _i32, _p, _p]),
    "pk_ls_ik_select_segments_f32": (_i32, [_p, _p, _p, _p, _p, _i32, _i32, _f32, _p, _p]),
    "pk_pose_hinge_cost_f32": (
        _i32,
        [_p, _p, _i64, _i32, _p, _p, _i32, _p, _i32, _f32, _i32, _p, _f32, _f32]
        + [_p] * 7 + [_i32, _i32, _p, _p],
    ),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)


def library_path() -> str:
    return _LIB_PATH


@functools.lru_cache(maxsize=1)
def lib() -> ctypes.CDLL:
    """Load the library once; raise ``RuntimeError`` when it is not built."""
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(
            f"HIP library not found at {_LIB_PATH}.\n"
            "Build it first:  python -m pyroffi_amd._build"
        )
    handle = ctypes.CDLL(_LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    return handle


class PkError(RuntimeError):
    """A C-ABI call failed with a HIP error (the reference's kInternal)."""


class PkInvalidArgument(PkError, ValueError):
    """A C-ABI call rejected its arguments (the reference's kInvalidArgument)."""


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().pk_last_error().decode(errors="replace")
        if status == 1:
            raise PkInvalidArgument(f"{what}: invalid argument: {msg}")
        raise PkError(f"{what}: HIP error: {msg}")
