"""Random packed worlds of the reference's four obstacle kinds (tests only).

Rows as in the reference's world arrays: sphere [c, R], capsule [a, b, R],
box [c, axis1, axis2, axis3, half lengths] with the axes the columns of a
random rotation, half-space [unit normal, point].
"""

import numpy as np
from scipy.spatial.transform import Rotation


def random_world(rng, ms=0, mc=0, mb=0, mh=0, extent=1.0):
    ws = np.concatenate([rng.uniform(-extent, extent, (ms, 3)), rng.uniform(0.02, 0.15, (ms, 1))], 1)
    a = rng.uniform(-extent, extent, (mc, 3))
    b = a + rng.normal(0, 0.2, (mc, 3))
    wc = np.concatenate([a, b, rng.uniform(0.02, 0.1, (mc, 1))], 1)
    R = Rotation.random(mb, random_state=rng).as_matrix().reshape(mb, 3, 3)
    wb = np.concatenate([rng.uniform(-extent, extent, (mb, 3)), R[:, :, 0], R[:, :, 1], R[:, :, 2],
                         rng.uniform(0.02, 0.3, (mb, 3))], 1)
    n = rng.normal(size=(mh, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    wh = np.concatenate([n, rng.uniform(-extent, extent, (mh, 3))], 1)
    f = lambda x, w: np.ascontiguousarray(x, np.float32).reshape(-1, w)  # noqa: E731
    return f(ws, 4), f(wc, 7), f(wb, 15), f(wh, 6)
