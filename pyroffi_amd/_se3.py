"""Host-side SE(3) helpers in ``wxyz_xyz`` storage (numpy, float64).

The reference stores every rigid transform as jaxlie's ``wxyz_xyz`` 7-vector
(pinned ``jaxlie==1.5.0``, ``pyproject.toml:27``): a unit Hamilton quaternion
followed by a translation.  The kinematics path uses four jaxlie operations on
the host side: ``SE3.from_matrix`` (URDF joint origins,
``src/pyroffi/_robot_urdf_parser.py:500``), ``SE3.exp`` / composition (the JAX FK,
``src/pyroffi/_robot.py:212-217``) and ``inverse``/``log`` (the IK residual,
``src/pyroffi/optimization_engines/_ik_primitives.py:102-103``).

This module restates those operations from their published definitions:

* rotation matrix -> quaternion: Mike Day, "Converting a Rotation Matrix to a
  Quaternion" (Insomniac Games, 2015) -- four candidate formulas, one picked from
  the signs of the diagonal exactly as jaxlie does, so the quaternion *sign*
  matches the reference's tables.
* exp/log on SO(3) and SE(3): closed forms (Rodrigues; left Jacobian ``V``) with
  small-angle Taylor branches below ``theta^2 < 1e-10`` (float64 epsilon used by
  jaxlie).

The ``SE3`` class mirrors the subset of ``jaxlie.SE3`` that the pyroffi API
passes around (targets for IK), so user code written against the reference keeps
working.  Device kernels never call into this module.
"""

from __future__ import annotations

import numpy as np

_EPS64 = 1e-10


def quat_mul(q1: np.ndarray, q2: np.ndarray) -> np.ndarray:
    """Hamilton product of wxyz quaternions (broadcasting over leading axes)."""
    w1, x1, y1, z1 = np.moveaxis(np.asarray(q1, dtype=np.float64), -1, 0)
    w2, x2, y2, z2 = np.moveaxis(np.asarray(q2, dtype=np.float64), -1, 0)
    return np.stack(
        [
            w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
            w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
            w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
            w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
        ],
        axis=-1,
    )


def quat_conj(q: np.ndarray) -> np.ndarray:
    q = np.asarray(q, dtype=np.float64)
    return q * np.array([1.0, -1.0, -1.0, -1.0])


def quat_to_matrix(q: np.ndarray) -> np.ndarray:
    q = np.asarray(q, dtype=np.float64)
    w, x, y, z = np.moveaxis(q, -1, 0)
    r = np.empty(q.shape[:-1] + (3, 3))
    r[..., 0, 0] = 1 - 2 * (y * y + z * z)
    r[..., 0, 1] = 2 * (x * y - w * z)
    r[..., 0, 2] = 2 * (x * z + w * y)
    r[..., 1, 0] = 2 * (x * y + w * z)
    r[..., 1, 1] = 1 - 2 * (x * x + z * z)
    r[..., 1, 2] = 2 * (y * z - w * x)
    r[..., 2, 0] = 2 * (x * z - w * y)
    r[..., 2, 1] = 2 * (y * z + w * x)
    r[..., 2, 2] = 1 - 2 * (x * x + y * y)
    return r


def matrix_to_quat(m: np.ndarray) -> np.ndarray:
    """Rotation matrix -> wxyz quaternion (Day 2015, jaxlie's case selection)."""
    m = np.asarray(m, dtype=np.float64)
    m00, m01, m02 = m[..., 0, 0], m[..., 0, 1], m[..., 0, 2]
    m10, m11, m12 = m[..., 1, 0], m[..., 1, 1], m[..., 1, 2]
    m20, m21, m22 = m[..., 2, 0], m[..., 2, 1], m[..., 2, 2]
    # Candidate where x dominates.
    t0 = 1 + m00 - m11 - m22
    q0 = np.stack([m21 - m12, t0, m10 + m01, m02 + m20], axis=-1)
    # y dominates.
    t1 = 1 - m00 + m11 - m22
    q1 = np.stack([m02 - m20, m10 + m01, t1, m21 + m12], axis=-1)
    # z dominates.
    t2 = 1 - m00 - m11 + m22
    q2 = np.stack([m10 - m01, m02 + m20, m21 + m12, t2], axis=-1)
    # w dominates.
    t3 = 1 + m00 + m11 + m22
    q3 = np.stack([t3, m21 - m12, m02 - m20, m10 - m01], axis=-1)
    c0 = m22 < 0
    c1 = m00 > m11
    c2 = m00 < -m11
    t = np.where(c0, np.where(c1, t0, t1), np.where(c2, t2, t3))
    q = np.where(
        c0[..., None],
        np.where(c1[..., None], q0, q1),
        np.where(c2[..., None], q2, q3),
    )
    return q * (0.5 / np.sqrt(t))[..., None]


def quat_rotate(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    return np.einsum("...ij,...j->...i", quat_to_matrix(q), np.asarray(v, dtype=np.float64))


def so3_exp(omega: np.ndarray) -> np.ndarray:
    """Axis-angle vector -> wxyz quaternion."""
    omega = np.asarray(omega, dtype=np.float64)
    th2 = np.sum(omega * omega, axis=-1)
    small = th2 < _EPS64
    th = np.sqrt(np.where(small, 1.0, th2))
    real = np.where(small, 1.0 - th2 / 8.0 + th2 * th2 / 384.0, np.cos(0.5 * th))
    imag = np.where(small, 0.5 - th2 / 48.0 + th2 * th2 / 3840.0, np.sin(0.5 * th) / th)
    return np.concatenate([real[..., None], imag[..., None] * omega], axis=-1)


def so3_log(q: np.ndarray) -> np.ndarray:
    """wxyz quaternion -> axis-angle vector (shortest rotation, |theta| <= pi)."""
    q = np.asarray(q, dtype=np.float64)
    w = q[..., 0]
    v = q[..., 1:]
    nv2 = np.sum(v * v, axis=-1)
    small = nv2 < _EPS64
    nv = np.sqrt(np.where(small, 1.0, nv2))
    # atan of (|v| / w) folded into (-pi/2, pi/2] keeps the shorter rotation.
    safe_w = np.where(np.abs(w) < 1e-300, 1e-300, w)
    ang = np.where(w < 0, np.arctan2(-nv, -w), np.arctan2(nv, w))
    scale = np.where(small, 2.0 / safe_w - 2.0 / 3.0 * nv2 / safe_w**3, 2.0 * ang / nv)
    return scale[..., None] * v


def _skew(w: np.ndarray) -> np.ndarray:
    z = np.zeros(w.shape[:-1])
    return np.stack(
        [
            np.stack([z, -w[..., 2], w[..., 1]], axis=-1),
            np.stack([w[..., 2], z, -w[..., 0]], axis=-1),
            np.stack([-w[..., 1], w[..., 0], z], axis=-1),
        ],
        axis=-2,
    )


def se3_exp(tangent: np.ndarray) -> np.ndarray:
    """[v; omega] twist -> wxyz_xyz transform (jaxlie tangent convention)."""
    tangent = np.asarray(tangent, dtype=np.float64)
    v = tangent[..., :3]
    w = tangent[..., 3:]
    q = so3_exp(w)
    th2 = np.sum(w * w, axis=-1)
    small = th2 < _EPS64
    th2s = np.where(small, 1.0, th2)
    th = np.sqrt(th2s)
    k = _skew(w)
    eye = np.broadcast_to(np.eye(3), k.shape)
    vmat = np.where(
        small[..., None, None],
        quat_to_matrix(q),
        eye
        + ((1.0 - np.cos(th)) / th2s)[..., None, None] * k
        + ((th - np.sin(th)) / (th2s * th))[..., None, None] * (k @ k),
    )
    t = np.einsum("...ij,...j->...i", vmat, v)
    return np.concatenate([q, t], axis=-1)


def se3_compose(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    q = quat_mul(a[..., :4], b[..., :4])
    t = quat_rotate(a[..., :4], b[..., 4:]) + a[..., 4:]
    return np.concatenate([q, t], axis=-1)


def se3_inverse(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    qi = quat_conj(a[..., :4])
    return np.concatenate([qi, -quat_rotate(qi, a[..., 4:])], axis=-1)


def se3_from_matrix(m: np.ndarray) -> np.ndarray:
    m = np.asarray(m, dtype=np.float64)
    return np.concatenate([matrix_to_quat(m[..., :3, :3]), m[..., :3, 3]], axis=-1)


def se3_as_matrix(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    out = np.zeros(a.shape[:-1] + (4, 4))
    out[..., :3, :3] = quat_to_matrix(a[..., :4])
    out[..., :3, 3] = a[..., 4:]
    out[..., 3, 3] = 1.0
    return out


def se3_log(a: np.ndarray) -> np.ndarray:
    """wxyz_xyz -> [v; omega] twist (inverse of ``se3_exp``)."""
    a = np.asarray(a, dtype=np.float64)
    w = so3_log(a[..., :4])
    th2 = np.sum(w * w, axis=-1)
    small = th2 < _EPS64
    th2s = np.where(small, 1.0, th2)
    th = np.sqrt(th2s)
    k = _skew(w)
    eye = np.broadcast_to(np.eye(3), k.shape)
    half = th / 2.0
    coef = (1.0 - th * np.cos(half) / (2.0 * np.sin(np.where(small, 1.0, half)))) / th2s
    vinv = np.where(
        small[..., None, None],
        eye - 0.5 * k + (k @ k) / 12.0,
        eye - 0.5 * k + coef[..., None, None] * (k @ k),
    )
    v = np.einsum("...ij,...j->...i", vinv, a[..., 4:])
    return np.concatenate([v, w], axis=-1)


class SO3:
    """Minimal stand-in for ``jaxlie.SO3`` (wxyz storage)."""

    def __init__(self, wxyz):
        self.wxyz = np.asarray(wxyz, dtype=np.float64)

    def inverse(self) -> "SO3":
        return SO3(quat_conj(self.wxyz))

    def __matmul__(self, other: "SO3") -> "SO3":
        return SO3(quat_mul(self.wxyz, other.wxyz))

    def log(self) -> np.ndarray:
        return so3_log(self.wxyz)

    def as_matrix(self) -> np.ndarray:
        return quat_to_matrix(self.wxyz)

    @staticmethod
    def exp(omega) -> "SO3":
        return SO3(so3_exp(omega))

    @staticmethod
    def from_matrix(m) -> "SO3":
        return SO3(matrix_to_quat(m))


class SE3:
    """Minimal stand-in for ``jaxlie.SE3`` holding a ``wxyz_xyz`` array.

    Supports leading batch axes like jaxlie.  Used for IK targets and the
    collision world; kernels receive the raw ``wxyz_xyz`` float32 buffer.
    """

    def __init__(self, wxyz_xyz):
        self.wxyz_xyz = np.asarray(wxyz_xyz, dtype=np.float64)
        assert self.wxyz_xyz.shape[-1] == 7

    def __repr__(self) -> str:
        return f"SE3(wxyz_xyz={self.wxyz_xyz!r})"

    def __matmul__(self, other: "SE3") -> "SE3":
        return SE3(se3_compose(self.wxyz_xyz, other.wxyz_xyz))

    def inverse(self) -> "SE3":
        return SE3(se3_inverse(self.wxyz_xyz))

    def log(self) -> np.ndarray:
        return se3_log(self.wxyz_xyz)

    def translation(self) -> np.ndarray:
        return self.wxyz_xyz[..., 4:]

    def rotation(self) -> SO3:
        return SO3(self.wxyz_xyz[..., :4])

    def as_matrix(self) -> np.ndarray:
        return se3_as_matrix(self.wxyz_xyz)

    def get_batch_axes(self) -> tuple[int, ...]:
        return self.wxyz_xyz.shape[:-1]

    @staticmethod
    def identity(batch_axes: tuple[int, ...] = ()) -> "SE3":
        out = np.zeros(batch_axes + (7,))
        out[..., 0] = 1.0
        return SE3(out)

    @staticmethod
    def exp(tangent) -> "SE3":
        return SE3(se3_exp(tangent))

    @staticmethod
    def from_matrix(m) -> "SE3":
        return SE3(se3_from_matrix(m))

    @staticmethod
    def from_translation(t) -> "SE3":
        t = np.asarray(t, dtype=np.float64)
        out = np.zeros(t.shape[:-1] + (7,))
        out[..., 0] = 1.0
        out[..., 4:] = t
        return SE3(out)

    @staticmethod
    def from_rotation_and_translation(rotation: SO3, translation) -> "SE3":
        q = np.asarray(rotation.wxyz, dtype=np.float64)
        t = np.asarray(translation, dtype=np.float64)
        batch = np.broadcast_shapes(q.shape[:-1], t.shape[:-1])
        return SE3(
            np.concatenate(
                [np.broadcast_to(q, batch + (4,)), np.broadcast_to(t, batch + (3,))], axis=-1
            )
        )
