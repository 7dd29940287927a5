"""Sphere geometry container (the sphere case of ``src/pyroffi/collision/_geometry.py``).

A ``Sphere`` holds a batch of centres (as SE(3) poses, like the reference's
``CollGeom.pose``) and radii.  Robot sphere tables and sphere worlds both use
it; the kernels read the float32 ``[x, y, z, r]`` rows of ``as_array()``.
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
