"""Host side of the trajectory model (CPU): the folded node tree reproduces the
FP64 oracle's link poses, the node-frame sphere table reproduces the oracle's
world sphere centres, and the same-slot pair lists and bounding spheres are
what the JAX self-distance semantics and the culling test need.
"""

import numpy as np
import pytest
from conftest import load_robot, random_cfgs

from pyroffi_amd._se3 import quat_rotate, se3_compose, se3_exp
from pyroffi_amd.collision import RobotCollisionSpherized
from pyroffi_amd.optimization_engines._lbfgs_trajopt_optimization import build_traj_tables, fold_robot


def _node_poses(nodes, q):
    P = []
    for r in nodes.astype(np.float64):
        src, par = int(r[15]), int(r[16])
        qj = (q[src] if src >= 0 else 0.0) * r[13] + r[14]
        local = se3_compose(r[:7], se3_exp(np.concatenate([r[7:10], r[10:13]]) * qj))
        P.append(local if par < 0 else se3_compose(P[par], local))
    return P


@pytest.mark.parametrize("name", ["panda", "fetch", "mimic_arm", "baxter"])
def test_folded_tree_reproduces_link_poses(oracle_mod, name):
    robot, path = load_robot(name)
    fold = fold_robot(robot)
    t = oracle_mod.Tables.of(robot)
    lpj = np.asarray(robot.links.parent_joint_indices)
    assert fold["nodes"].shape[0] <= 16
    for q in random_cfgs(robot, 4, seed=3):
        P = _node_poses(fold["nodes"], q)
        ref = oracle_mod.fk_links(t, q[None])[0]
        for l in range(robot.links.num_links):
            j = lpj[l]
            if j < 0:
                continue
            n = fold["anchor"][j]
            got = fold["rel"][j] if n < 0 else se3_compose(P[n], fold["rel"][j])
            if got[0] * ref[l, 0] < 0:
                got = np.concatenate([-got[:4], got[4:]])
            np.testing.assert_allclose(got, ref[l], atol=2e-6)


@pytest.mark.parametrize("name", ["panda", "fetch"])
def test_node_frame_spheres_and_pairs(oracle_mod, name):
    robot, path = load_robot(name)
    coll = RobotCollisionSpherized.from_urdf(path)
    h = build_traj_tables(robot, coll)
    t = oracle_mod.Tables.of(robot)
    sp = oracle_mod.padded_spheres(coll)
    q = random_cfgs(robot, 1, seed=5)[0]
    P = _node_poses(h["nodes"], q)
    lp = oracle_mod.fk_links(t, q[None])[0]
    start = h["link_sph_start"]
    order = h["link_order"]
    assert np.all(np.diff(h["link_node"]) >= 0)  # grouped by node
    for l in range(robot.links.num_links):
        n = h["link_node"][l]
        rl = order[l]
        real = sp[rl][sp[rl, :, 3] > 0]
        rows = h["spheres"][start[l]:start[l + 1]].astype(np.float64)
        assert len(rows) == len(real)
        for s in range(len(rows)):
            c = rows[s, :3] if n < 0 else quat_rotate(P[n][:4], rows[s, :3]) + P[n][4:]
            ref = quat_rotate(lp[rl, :4], real[s, :3]) + lp[rl, 4:]
            np.testing.assert_allclose(c, ref, atol=2e-6)
            assert rows[s, 3] == pytest.approx(real[s, 3])
            # the bounding sphere covers the sphere
            b = h["link_bound"][l].astype(np.float64)
            assert np.linalg.norm(rows[s, :3] - b[:3]) + rows[s, 3] <= b[3] + 1e-6
    counts = np.diff(start)
    for p, (i, j) in enumerate(h["pair_links"]):
        k0, k1 = h["pair_start"][p], h["pair_start"][p + 1]
        assert k1 - k0 == min(counts[i], counts[j])
        for s, (a, b) in enumerate(h["pair_items"][k0:k1]):
            assert a == start[i] + s and b == start[j] + s
    np.testing.assert_array_equal(order[h["pair_links"][:, 0]], coll.active_idx_i)
    np.testing.assert_array_equal(order[h["pair_links"][:, 1]], coll.active_idx_j)
