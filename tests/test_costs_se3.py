"""The cost module's device SE(3) log (torch) against the FP64 host
restatement ``pyroffi_amd._se3.se3_log`` (itself pinned by the exp/log round
trip in test_model_tables.py), on the CPU in float64: generic poses, rotations
below the small-angle threshold, rotations near pi from both hemispheres, and
the exp -> log round trip."""

import numpy as np
import torch

from pyroffi_amd._se3 import se3_exp, se3_log
from pyroffi_amd.costs._costs import _inv_compose, se3_log_t


def _random_poses(rng, n, max_angle=np.pi):
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    ang = rng.uniform(-max_angle, max_angle, n)
    q = np.concatenate([np.cos(ang / 2)[:, None], np.sin(ang / 2)[:, None] * axis], 1)
    q *= rng.choice([-1, 1], (n, 1))  # both hemispheres
    return np.concatenate([q, rng.uniform(-1, 1, (n, 3))], 1)


def test_se3_log_matches_host_restatement():
    rng = np.random.default_rng(0)
    poses = np.concatenate([_random_poses(rng, 200), _random_poses(rng, 50, 1e-9), _random_poses(rng, 50, 1e-4)])
    near_pi = _random_poses(rng, 50, np.pi)
    near_pi[:, 0] = rng.uniform(-1e-4, 1e-4, 50)  # |w| tiny: angle close to pi
    near_pi[:, 1:4] /= np.linalg.norm(near_pi[:, 1:4], axis=1, keepdims=True) / np.sqrt(1 - near_pi[:, :1] ** 2)
    poses = np.concatenate([poses, near_pi])
    got = se3_log_t(torch.as_tensor(poses)).numpy()
    np.testing.assert_allclose(got, se3_log(poses), atol=1e-9, rtol=0)


def test_se3_log_inverts_exp_and_relative_pose():
    rng = np.random.default_rng(1)
    w = rng.normal(size=(300, 3))
    w *= rng.uniform(0, 3.0, (300, 1)) / np.linalg.norm(w, axis=1, keepdims=True)  # |omega| < pi: no wrap
    tangent = np.concatenate([rng.uniform(-1, 1, (300, 3)), w], 1)
    T = se3_exp(tangent)
    np.testing.assert_allclose(se3_log_t(torch.as_tensor(T)).numpy(), tangent, atol=1e-9)
    # a^-1 o a is the identity; a^-1 o (a o exp(x)) = exp(x)
    a = torch.as_tensor(_random_poses(rng, 300))
    np.testing.assert_allclose(se3_log_t(_inv_compose(a, a)).numpy(), 0.0, atol=1e-12)
    from pyroffi_amd._se3 import se3_compose
    b = torch.as_tensor(se3_compose(a.numpy(), T))
    np.testing.assert_allclose(se3_log_t(_inv_compose(a, b)).numpy(), tangent, atol=1e-9)
