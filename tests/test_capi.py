"""C ABI of libpyroffi_amd.so, checked without a GPU.

The library must load and export every function ``include/pyroffi_amd.h``
declares; argument validation happens on the host before any HIP call, so
invalid arguments and empty batches can be exercised here.
"""

import ctypes
import os
import re

import pytest
from conftest import ROOT

from pyroffi_amd.kernels import _lib

HEADER = os.path.join(ROOT, "include", "pyroffi_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(pk_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 9
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED_SYMBOLS)
    assert lib.pk_abi_version() == 1


def _null():
    return ctypes.c_void_p(0)


def test_invalid_arguments_are_rejected_on_host():
    lib = _lib.lib()
    n = _null()
    # n_joints above the limit (the reference FK handler's kInvalidArgument)
    st = lib.pk_fk_f32(n, n, 4, 7, *([n] * 10), 65, 1, n)
    assert st == 1 and b"n_joints" in lib.pk_last_error()
    # n_act above the limit
    st = lib.pk_ls_ik_f32(n, n, 1, 1, 17, *([n] * 8), 12, n, n, 1, n, n, n, 0, n, 0, n, 0, n, 0, n, 0,
                          n, n, n, 10, 50.0, 10.0, 1e-2, 1e-8, 1e-8, 0, 0.0, 0.02, n, n)
    assert st == 1 and b"n_act" in lib.pk_last_error()
    # n_ee above the limit
    st = lib.pk_ls_ik_f32(n, n, 1, 1, 7, *([n] * 8), 12, n, n, 5, n, n, n, 0, n, 0, n, 0, n, 0, n, 0,
                          n, n, n, 10, 50.0, 10.0, 1e-2, 1e-8, 1e-8, 0, 0.0, 0.02, n, n)
    assert st == 1 and b"n_ee" in lib.pk_last_error()
    # obstacle counts with no array behind them
    one = ctypes.c_void_p(8)
    st = lib.pk_collision_world_sphere_reduced_f32(n, one, one, 2, 26, n, 0, n, 3, n, 0, n, 0, 13, one)
    assert st == 1 and b"null world" in lib.pk_last_error()
    st = lib.pk_collision_world_sphere_f32(n, one, one, 2, 26, one, 4, n, 0, n, 2, n, 0, one)
    assert st == 1 and b"null world" in lib.pk_last_error()
    # FK adjoint: n_act above the limit
    st = lib.pk_fk_vjp_f32(n, n, 4, 17, *([n] * 8), 12, n, 13, n, n)
    assert st == 1 and b"n_act" in lib.pk_last_error()
    # K not a multiple of the link count
    st = lib.pk_collision_world_sphere_reduced_f32(n, n, n, 2, 25, n, 4, n, 0, n, 0, n, 0, 13, n)
    assert st == 1
    # self-entry encoding limits
    st = lib.pk_sphere_sweep_f32(n, n, 2, 13, n, n, 2000, n, 0, n, 0, 0, n, n, n)
    assert st == 1
    st = lib.pk_sphere_sweep_f32(n, n, 2, 13, n, n, 59, n, 0, n, 1, 2048, n, n, n)
    assert st == 1 and b"2047" in lib.pk_last_error()
    # self entries with no pair slot to publish to
    st = lib.pk_sphere_sweep_f32(n, one, 2, 13, one, one, 59, n, 0, one, 5, 0, n, n, one)
    assert st == 1 and b"n_pairs=0" in lib.pk_last_error()
    # null output pointer with work to do
    st = lib.pk_fk_links_f32(n, n, 4, 7, *([n] * 8), 12, n, 13, n)
    assert st == 1


def test_empty_batches_are_noops():
    lib = _lib.lib()
    n = _null()
    assert lib.pk_fk_f32(n, n, 0, 7, *([n] * 10), 12, 10, n) == 0
    assert lib.pk_fk_links_f32(n, n, 0, 7, *([n] * 8), 12, n, 13, n) == 0
    assert lib.pk_collision_world_sphere_reduced_f32(n, n, n, 0, 26, n, 4, n, 0, n, 0, n, 0, 13, n) == 0
    assert lib.pk_collision_self_sphere_f32(n, n, n, 0, 2, 13, n, n, 5, n) == 0
    assert lib.pk_sphere_sweep_f32(n, n, 0, 13, n, n, 59, n, 256, n, 0, 0, n, n, n) == 0
    assert lib.pk_ls_ik_select_f32(n, n, n, n, 0, 4, 7, 0.0, n, n) == 0
    assert lib.pk_collision_world_sphere_f32(n, n, n, 0, 26, n, 4, n, 0, n, 0, n, 0, n) == 0
    assert lib.pk_fk_vjp_f32(n, n, 0, 7, *([n] * 8), 12, n, 13, n, n) == 0
    assert lib.pk_sphere_distances_jax_f32(n, n, 0, 13, n, 18, n, 4, n, n, 5, n, n) == 0
    assert lib.pk_last_error() is not None


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB_PATH", str(tmp_path / "missing.so"))
    _lib.lib.cache_clear()
    try:
        with pytest.raises(RuntimeError, match="not found"):
            _lib.lib()
    finally:
        monkeypatch.undo()
        _lib.lib.cache_clear()
