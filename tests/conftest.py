import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (ROOT, os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
ROBOTS = ("panda", "ur5", "fetch", "baxter")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


def load_robot(name):
    import pyroffi_amd as pk

    path = os.path.join(GOLDEN, "mimic_arm.urdf") if name == "mimic_arm" else pk.bundled_urdf_path(name)
    return pk.Robot.from_urdf(path), path


def random_cfgs(robot, n, seed=0, margin=0.0):
    rng = np.random.default_rng(seed)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    return rng.uniform(lo - margin, hi + margin, size=(n, robot.joints.num_actuated_joints))
