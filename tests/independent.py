"""Independent restatements used to pin the oracle (tests only).

These read the URDF XML directly and use scipy's rotation and distance
routines, so they share no code with the oracle or the package: a wrong sign,
index or composition order in either one shows up as a mismatch.
"""

from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np
from scipy.spatial import distance
from scipy.spatial.transform import Rotation


def _vec(text, default):
    return np.array([float(v) for v in text.split()]) if text else np.array(default, dtype=float)


def _origin(elem):
    T = np.eye(4)
    if elem is not None:
        T[:3, :3] = Rotation.from_euler("xyz", _vec(elem.get("rpy"), (0, 0, 0))).as_matrix()
        T[:3, 3] = _vec(elem.get("xyz"), (0, 0, 0))
    return T


def read(path):
    root = ET.parse(path).getroot()
    links = [l.get("name") for l in root.findall("link")]
    joints = []
    for j in root.findall("joint"):
        m = j.find("mimic")
        ax = j.find("axis")
        joints.append(
            dict(
                name=j.get("name"),
                type=j.get("type"),
                parent=j.find("parent").get("link"),
                child=j.find("child").get("link"),
                origin=_origin(j.find("origin")),
                axis=_vec(None if ax is None else ax.get("xyz"), (1, 0, 0)),
                mimic=None
                if m is None
                else (m.get("joint"), float(m.get("multiplier", 1.0)), float(m.get("offset", 0.0))),
            )
        )
    spheres = {}
    for l in root.findall("link"):
        rows = []
        for c in l.findall("collision"):
            s = c.find("geometry/sphere")
            if s is not None:
                rows.append([*_origin(c.find("origin"))[:3, 3], float(s.get("radius"))])
        spheres[l.get("name")] = np.array(rows).reshape(-1, 4)
    return links, joints, spheres


def link_matrices(path, q):
    """World 4x4 transform of every link (document order) for actuated values q."""
    links, joints, _ = read(path)
    actuated = [j["name"] for j in joints if j["type"] != "fixed" and j["mimic"] is None]
    by_child = {j["child"]: j for j in joints}
    value = {}
    for j in joints:
        if j["name"] in actuated:
            value[j["name"]] = q[actuated.index(j["name"])]
    for j in joints:
        if j["mimic"] is not None:
            src, mul, off = j["mimic"]
            value[j["name"]] = value.get(src, 0.0) * mul + off
    cache = {}

    def world(link):
        if link in cache:
            return cache[link]
        j = by_child.get(link)
        if j is None:
            T = np.eye(4)
        else:
            v = value.get(j["name"], 0.0)
            M = np.eye(4)
            if j["type"] in ("revolute", "continuous"):
                M[:3, :3] = Rotation.from_rotvec(j["axis"] * v).as_matrix()
            elif j["type"] == "prismatic":
                M[:3, 3] = j["axis"] * v
            T = world(j["parent"]) @ j["origin"] @ M
        cache[link] = T
        return T

    return np.stack([world(l) for l in links])


def pose_to_matrix(p):
    p = np.asarray(p, dtype=float)
    T = np.zeros(p.shape[:-1] + (4, 4))
    T[..., :3, :3] = Rotation.from_quat(p[..., [1, 2, 3, 0]].reshape(-1, 4)).as_matrix().reshape(p.shape[:-1] + (3, 3))
    T[..., :3, 3] = p[..., 4:]
    T[..., 3, 3] = 1
    return T


def world_spheres(path, q):
    """World-frame sphere centres and radii per link (dict link -> (k, 4))."""
    links, _, spheres = read(path)
    T = link_matrices(path, q)
    out = {}
    for i, l in enumerate(links):
        s = spheres[l]
        c = (T[i, :3, :3] @ s[:, :3].T).T + T[i, :3, 3] if len(s) else np.zeros((0, 3))
        out[l] = np.concatenate([c, s[:, 3:]], axis=1) if len(s) else np.zeros((0, 4))
    return links, out


def world_distance_matrix(path, q, world):
    """(L, M) per link: min over its spheres of |c - o| - r - R (1e9 without spheres)."""
    links, ws = world_spheres(path, q)
    out = np.full((len(links), len(world)), 1e9)
    for i, l in enumerate(links):
        s = ws[l]
        if len(s):
            d = distance.cdist(s[:, :3], world[:, :3]) - s[:, 3:4] - world[None, :, 3]
            out[i] = d.min(axis=0)
    return out


def self_distance(path, q, pair_i, pair_j):
    links, ws = world_spheres(path, q)
    out = np.full(len(pair_i), 1e9)
    for p, (i, j) in enumerate(zip(pair_i, pair_j)):
        a, b = ws[links[i]], ws[links[j]]
        if len(a) and len(b):
            out[p] = (distance.cdist(a[:, :3], b[:, :3]) - a[:, 3:4] - b[None, :, 3]).min()
    return out
