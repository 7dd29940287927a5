"""The on-disk model format (CPU): a saved robot and collision model load back
with identical tables, the same FK (FP64 oracle on the loaded tables) and the
same collision tables; corrupt or foreign files are rejected before any kernel."""

import json

import numpy as np
import pytest
from conftest import ROBOTS, load_robot, random_cfgs

from pyroffi_amd._model_file import load_model, save_model
from pyroffi_amd.collision import RobotCollision, RobotCollisionSpherized

JOINT_FIELDS = ("twists", "parent_transforms", "parent_indices", "actuated_indices", "lower_limits", "upper_limits",
                "velocity_limits", "mimic_multiplier", "mimic_offset", "mimic_act_indices", "_topo_sort_inv",
                "fk_level_starts", "fk_level_joints")


@pytest.mark.parametrize("name", ROBOTS + ("mimic_arm",))
def test_round_trip_tables_and_fk(oracle_mod, tmp_path, name):
    robot, path = load_robot(name)
    coll = RobotCollisionSpherized.from_urdf(path, semantics="cuda")
    f = tmp_path / f"{name}.npz"
    save_model(f, robot, coll)
    r2, c2 = load_model(f)
    for k in JOINT_FIELDS:
        np.testing.assert_array_equal(getattr(r2.joints, k), getattr(robot.joints, k))
    assert r2.joints.names == robot.joints.names and r2.links.names == robot.links.names
    assert r2.joints.num_fk_levels == robot.joints.num_fk_levels
    np.testing.assert_array_equal(r2.links.parent_joint_indices, robot.links.parent_joint_indices)
    q = random_cfgs(robot, 20, seed=1)
    np.testing.assert_array_equal(oracle_mod.fk_links(oracle_mod.Tables.of(r2), q),
                                  oracle_mod.fk_links(oracle_mod.Tables.of(robot), q))
    assert isinstance(c2, RobotCollisionSpherized) and c2.semantics == "cuda"
    for k in ("spheres_local", "link_sphere_start", "self_entries", "active_idx_i", "active_idx_j"):
        np.testing.assert_array_equal(getattr(c2, k), getattr(coll, k))
    np.testing.assert_array_equal(c2.coll.as_array(), coll.coll.as_array())


def test_round_trip_capsules_and_no_collision(tmp_path):
    robot, path = load_robot("panda")
    cap = RobotCollision.from_urdf(path)
    save_model(tmp_path / "c.npz", robot, cap)
    _, c2 = load_model(tmp_path / "c.npz")
    assert isinstance(c2, RobotCollision)
    np.testing.assert_array_equal(c2.coll.as_array(), cap.coll.as_array())
    np.testing.assert_array_equal(c2.active_idx_i, cap.active_idx_i)
    save_model(tmp_path / "r.npz", robot)
    _, none = load_model(tmp_path / "r.npz")
    assert none is None


def test_rejects_foreign_and_corrupt_files(tmp_path):
    robot, _ = load_robot("panda")
    np.savez(tmp_path / "x.npz", a=np.zeros(3))
    with pytest.raises(ValueError, match="no header"):
        load_model(tmp_path / "x.npz")
    save_model(tmp_path / "m.npz", robot)
    with np.load(tmp_path / "m.npz", allow_pickle=False) as z:
        a = {k: z[k] for k in z.files}
    hdr = json.loads(bytes(a["header"]).decode())
    hdr["version"] = 99
    np.savez(tmp_path / "v.npz", **{**a, "header": np.frombuffer(json.dumps(hdr).encode(), np.uint8)})
    with pytest.raises(ValueError, match="version"):
        load_model(tmp_path / "v.npz")
    bad = dict(a)
    bad["joints/parent_indices"] = a["joints/parent_indices"].copy()
    bad["joints/parent_indices"][3] = 99
    np.savez(tmp_path / "b.npz", **bad)
    with pytest.raises(ValueError, match="parent_indices"):
        load_model(tmp_path / "b.npz")
    bad = dict(a)
    bad["joints/_topo_sort_inv"] = np.zeros_like(a["joints/_topo_sort_inv"])
    np.savez(tmp_path / "t.npz", **bad)
    with pytest.raises(ValueError, match="permutation"):
        load_model(tmp_path / "t.npz")
