"""Geometry containers of ``src/pyroffi/collision/_geometry.py``: spheres, capsules,
boxes and half-spaces.

Each holds a batch of SE(3) poses (like the reference's ``CollGeom.pose``) and
its size parameters; ``as_array()`` gives the packed float32 rows the kernels
read (the reference's world arrays, ``src/pyroffi/collision/_cuda_collision.py:98-168``):
sphere ``[x, y, z, r]``, capsule ``[a, b, r]``, box ``[centre, 3 axes, half
lengths]``, half-space ``[normal, point]``.
"""

from __future__ import annotations

import numpy as np

from .._se3 import SE3


class Sphere:
    def __init__(self, pose: SE3, radius):
        self.pose = pose
        self.radius = np.asarray(radius, dtype=np.float64)
        assert self.pose.get_batch_axes() == self.radius.shape

    @staticmethod
    def from_center_and_radius(center, radius) -> "Sphere":
        center = np.asarray(center, dtype=np.float64)
        radius = np.asarray(radius, dtype=np.float64)
        batch = np.broadcast_shapes(center.shape[:-1], radius.shape)
        center = np.broadcast_to(center, batch + (3,))
        return Sphere(SE3.from_translation(center), np.broadcast_to(radius, batch))

    def get_batch_axes(self) -> tuple[int, ...]:
        return self.radius.shape

    def as_array(self) -> np.ndarray:
        """``(*batch, 4)`` float32 ``[x, y, z, r]``."""
        c = self.pose.translation()
        return np.concatenate([c, self.radius[..., None]], axis=-1).astype(np.float32)


def _frame_with_z(normal: np.ndarray) -> np.ndarray:
    """Rotation matrices ``(*batch, 3, 3)`` whose third column is the unit ``normal``."""
    z = normal / np.linalg.norm(normal, axis=-1, keepdims=True)
    helper = np.where(np.abs(z[..., :1]) < 0.9, np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]))
    x = np.cross(helper, z)
    x /= np.linalg.norm(x, axis=-1, keepdims=True)
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=-1)


class Capsule:
    """Capsules: a pose (the axis is its local z) with radius and full height of the
    cylindrical part (the reference's ``Capsule``, ``src/pyroffi/collision/_geometry.py:382-432``)."""

    def __init__(self, pose: SE3, radius, height):
        self.pose = pose
        self.radius = np.asarray(radius, dtype=np.float64)
        self.height = np.asarray(height, dtype=np.float64)

    @staticmethod
    def from_radius_height(radius, height, position=None, wxyz=None) -> "Capsule":
        position = np.zeros(3) if position is None else np.asarray(position, np.float64)
        wxyz = np.array([1.0, 0.0, 0.0, 0.0]) if wxyz is None else np.asarray(wxyz, np.float64)
        radius, height = np.asarray(radius, np.float64), np.asarray(height, np.float64)
        batch = np.broadcast_shapes(position.shape[:-1], wxyz.shape[:-1], radius.shape, height.shape)
        pose = SE3(np.concatenate([np.broadcast_to(wxyz, batch + (4,)), np.broadcast_to(position, batch + (3,))], -1))
        return Capsule(pose, np.broadcast_to(radius, batch), np.broadcast_to(height, batch))

    @staticmethod
    def from_spheres(centers, radii) -> "Capsule":
        """A capsule bounding a set of spheres: the axis is the principal direction of
        the centres, the height spans the spheres' extent along it and the radius
        their largest reach off it -- a bounding cylinder turned into a capsule of the
        same radius and height, as ``Capsule.from_trimesh`` does with trimesh's
        minimum cylinder (``src/pyroffi/collision/_geometry.py:433-450``).  No spheres
        give the reference's empty capsule (identity pose, zero size)."""
        c = np.asarray(centers, np.float64).reshape(-1, 3)
        r = np.asarray(radii, np.float64).reshape(-1)
        if c.shape[0] == 0:
            return Capsule.from_radius_height(0.0, 0.0)
        mean = c.mean(axis=0)
        axis = np.array([0.0, 0.0, 1.0])
        if c.shape[0] > 1:
            w, v = np.linalg.eigh((c - mean).T @ (c - mean))
            if w[-1] > 1e-12:
                axis = v[:, -1] * (1.0 if v[2, -1] >= 0 else -1.0)
        t = (c - mean) @ axis
        lo, hi = float(np.min(t - r)), float(np.max(t + r))
        perp = np.linalg.norm((c - mean) - t[:, None] * axis, axis=1)
        radius = float(np.max(perp + r))
        m = np.eye(4)
        m[:3, :3] = _frame_with_z(axis)
        m[:3, 3] = mean + axis * (0.5 * (lo + hi))
        return Capsule(SE3.from_matrix(m), radius, hi - lo)

    @property
    def axis(self) -> np.ndarray:
        return self.pose.rotation().as_matrix()[..., :, 2]

    def get_batch_axes(self) -> tuple[int, ...]:
        return self.radius.shape

    def as_array(self) -> np.ndarray:
        """``(*batch, 7)`` float32 ``[a, b, r]``: the segment ends centre -/+ axis * height / 2."""
        c = self.pose.translation()
        half = self.axis * (0.5 * self.height)[..., None]
        return np.concatenate([c - half, c + half, self.radius[..., None]], axis=-1).astype(np.float32)


class Box:
    """Boxes: a pose and half lengths along its local axes (the reference's ``Box``)."""

    def __init__(self, pose: SE3, half_lengths):
        self.pose = pose
        self.half_lengths = np.asarray(half_lengths, dtype=np.float64)

    @staticmethod
    def from_center_and_half_lengths(center, half_lengths, wxyz=None) -> "Box":
        center, hl = np.asarray(center, np.float64), np.asarray(half_lengths, np.float64)
        wxyz = np.array([1.0, 0.0, 0.0, 0.0]) if wxyz is None else np.asarray(wxyz, np.float64)
        batch = np.broadcast_shapes(center.shape[:-1], hl.shape[:-1], wxyz.shape[:-1])
        pose = SE3(np.concatenate([np.broadcast_to(wxyz, batch + (4,)), np.broadcast_to(center, batch + (3,))], -1))
        return Box(pose, np.broadcast_to(hl, batch + (3,)))

    @staticmethod
    def from_center_and_dimensions(center, length, width, height, wxyz=None) -> "Box":
        hl = 0.5 * np.stack(np.broadcast_arrays(np.asarray(length, np.float64), np.asarray(width, np.float64),
                                                np.asarray(height, np.float64)), axis=-1)
        return Box.from_center_and_half_lengths(center, hl, wxyz)

    def get_batch_axes(self) -> tuple[int, ...]:
        return self.half_lengths.shape[:-1]

    def as_array(self) -> np.ndarray:
        """``(*batch, 15)`` float32 ``[centre, axis 1, axis 2, axis 3, half lengths]``."""
        R = self.pose.rotation().as_matrix()
        return np.concatenate([self.pose.translation(), R[..., :, 0], R[..., :, 1], R[..., :, 2],
                               self.half_lengths], axis=-1).astype(np.float32)


class HalfSpace:
    """Half-spaces: a point on the boundary plane and the outward normal (the local z)."""

    def __init__(self, pose: SE3):
        self.pose = pose

    @staticmethod
    def from_point_and_normal(point, normal) -> "HalfSpace":
        point, normal = np.asarray(point, np.float64), np.asarray(normal, np.float64)
        batch = np.broadcast_shapes(point.shape[:-1], normal.shape[:-1])
        R = _frame_with_z(np.broadcast_to(normal, batch + (3,)))
        m = np.zeros(batch + (4, 4))
        m[..., :3, :3] = R
        m[..., :3, 3] = np.broadcast_to(point, batch + (3,))
        m[..., 3, 3] = 1.0
        return HalfSpace(SE3.from_matrix(m))

    @property
    def normal(self) -> np.ndarray:
        return self.pose.rotation().as_matrix()[..., :, 2]

    def get_batch_axes(self) -> tuple[int, ...]:
        return self.pose.get_batch_axes()

    def as_array(self) -> np.ndarray:
        """``(*batch, 6)`` float32 ``[unit normal, point]``."""
        return np.concatenate([self.normal, self.pose.translation()], axis=-1).astype(np.float32)
