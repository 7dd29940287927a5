"""Collision-cost transforms on signed distances."""

from __future__ import annotations

import numpy as np
import torch


def colldist_from_sdf(_dist, activation_dist):
    """Smooth collision cost from a signed distance (``src/pyroffi/collision/_collision.py:136-167``).

    d' = min(d, a);  d' - a/2 if d' < 0, else -0.5/(a + 1e-6) (d' - a)^2;  then min(., 0).
    Returns values <= 0: 0 at and beyond the activation distance, increasingly
    negative with penetration.  Torch tensors are transformed on their own
    device; other array-likes are uploaded, transformed on the GPU and returned
    as numpy.
    """
    if _dist is None:
        _dist = activation_dist
    if not isinstance(_dist, torch.Tensor):
        return colldist_from_sdf(torch.as_tensor(np.asarray(_dist), device="cuda"), activation_dist).cpu().numpy()
    d = torch.clamp(_dist, max=activation_dist)
    v = torch.where(d < 0, d - 0.5 * activation_dist,
                    -0.5 / (activation_dist + 1e-6) * (d - activation_dist) ** 2)
    return torch.clamp(v, max=0.0)
