"""Dependency-free URDF reader.

The reference loads robots through ``yourdfpy.URDF`` (pinned ``yourdfpy==0.0.58``,
``pyproject.toml:31``) and reads only a handful of its fields on the kinematics
path: ``joint_map`` and ``link_map`` in document order, ``actuated_joints``,
``base_link``, each joint's ``type/parent/child/origin/axis/limit/mimic`` and each
link's sphere ``collision`` elements (``src/pyroffi/_robot_urdf_parser.py:163-500``,
``src/pyroffi/collision/_robot_collision.py:551-580``).  This module parses the
same fields straight from the XML with ``xml.etree`` so that model ingestion needs
nothing outside the standard library and numpy.

Conventions followed (yourdfpy's published behaviour):

* ``origin`` = 4x4 homogeneous transform ``T = [Rz(yaw) Ry(pitch) Rx(roll) | xyz]``
  (URDF fixed-axis roll/pitch/yaw); a missing ``<origin>`` is ``None``.
* ``axis`` defaults to ``(1, 0, 0)`` and is not normalised.
* ``actuated_joints`` = joints whose type is not ``fixed`` and that carry no
  ``<mimic>``, in document order.
* ``base_link`` = the first link (document order) that is no joint's child.
"""

from __future__ import annotations

import dataclasses
import os
import xml.etree.ElementTree as ET

import numpy as np


@dataclasses.dataclass
class Limit:
    lower: float | None
    upper: float | None
    velocity: float | None
    effort: float | None


@dataclasses.dataclass
class Mimic:
    joint: str
    multiplier: float | None
    offset: float | None


@dataclasses.dataclass
class Joint:
    name: str
    type: str
    parent: str
    child: str
    origin: np.ndarray | None
    axis: np.ndarray
    limit: Limit | None
    mimic: Mimic | None


@dataclasses.dataclass
class SphereCollision:
    """One ``<collision><geometry><sphere>`` entry of a link."""

    radius: float
    origin: np.ndarray  # 4x4, identity when the element has no <origin>


@dataclasses.dataclass
class Link:
    name: str
    spheres: list[SphereCollision]
    n_non_sphere_collisions: int


def _floats(text: str | None, n: int, default: tuple[float, ...]) -> np.ndarray:
    if text is None:
        return np.array(default, dtype=np.float64)
    vals = [float(t) for t in text.split()]
    if len(vals) != n:
        raise ValueError(f"expected {n} numbers, got {text!r}")
    return np.array(vals, dtype=np.float64)


def _opt_float(text: str | None) -> float | None:
    return None if text is None else float(text)


def rpy_xyz_to_matrix(rpy: np.ndarray, xyz: np.ndarray) -> np.ndarray:
    """4x4 transform for a URDF ``<origin rpy xyz>``: R = Rz(y) @ Ry(p) @ Rx(r)."""
    r, p, y = (float(v) for v in rpy)
    cr, sr = np.cos(r), np.sin(r)
    cp, sp = np.cos(p), np.sin(p)
    cy, sy = np.cos(y), np.sin(y)
    rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    out = np.eye(4)
    out[:3, :3] = rz @ ry @ rx
    out[:3, 3] = xyz
    return out


def _parse_origin(elem: ET.Element | None) -> np.ndarray | None:
    if elem is None:
        return None
    xyz = _floats(elem.get("xyz"), 3, (0.0, 0.0, 0.0))
    rpy = _floats(elem.get("rpy"), 3, (0.0, 0.0, 0.0))
    return rpy_xyz_to_matrix(rpy, xyz)


class URDF:
    """Parsed URDF exposing the yourdfpy fields the kinematics path reads."""

    def __init__(self, joints: list[Joint], links: list[Link], name: str, path: str | None):
        self.name = name
        self.path = path
        self.joint_map: dict[str, Joint] = {j.name: j for j in joints}
        self.link_map: dict[str, Link] = {l.name: l for l in links}
        self.actuated_joints: list[Joint] = [
            j for j in joints if j.type != "fixed" and j.mimic is None
        ]
        children = {j.child for j in joints}
        roots = [l.name for l in links if l.name not in children]
        if not roots:
            raise ValueError("URDF has no root link")
        self.base_link: str = roots[0]

    @staticmethod
    def load(path: str | os.PathLike) -> "URDF":
        path = os.fspath(path)
        with open(path, "rb") as f:
            return URDF.from_xml_string(f.read(), path=path)

    @staticmethod
    def from_xml_string(xml: str | bytes, path: str | None = None) -> "URDF":
        root = ET.fromstring(xml)
        if root.tag != "robot":
            raise ValueError(f"not a URDF <robot> document (root tag {root.tag!r})")
        links: list[Link] = []
        for le in root.findall("link"):
            spheres: list[SphereCollision] = []
            n_other = 0
            for ce in le.findall("collision"):
                ge = ce.find("geometry")
                se = None if ge is None else ge.find("sphere")
                if se is None:
                    n_other += 1
                    continue
                origin = _parse_origin(ce.find("origin"))
                spheres.append(
                    SphereCollision(
                        radius=float(se.get("radius")),
                        origin=np.eye(4) if origin is None else origin,
                    )
                )
            links.append(Link(name=le.get("name"), spheres=spheres, n_non_sphere_collisions=n_other))
        joints: list[Joint] = []
        for je in root.findall("joint"):
            le = je.find("limit")
            me = je.find("mimic")
            ae = je.find("axis")
            joints.append(
                Joint(
                    name=je.get("name"),
                    type=je.get("type"),
                    parent=je.find("parent").get("link"),
                    child=je.find("child").get("link"),
                    origin=_parse_origin(je.find("origin")),
                    axis=_floats(None if ae is None else ae.get("xyz"), 3, (1.0, 0.0, 0.0)),
                    limit=None
                    if le is None
                    else Limit(
                        lower=_opt_float(le.get("lower")),
                        upper=_opt_float(le.get("upper")),
                        velocity=_opt_float(le.get("velocity")),
                        effort=_opt_float(le.get("effort")),
                    ),
                    mimic=None
                    if me is None
                    else Mimic(
                        joint=me.get("joint"),
                        multiplier=_opt_float(me.get("multiplier")),
                        offset=_opt_float(me.get("offset")),
                    ),
                )
            )
        return URDF(joints, links, name=root.get("name", ""), path=path)


_RES_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "resources")

_BUNDLED = {
    "panda": ("panda/panda_spherized.urdf", "panda/panda.srdf"),
    "ur5": ("ur5/ur5_spherized.urdf", "ur5/ur5.srdf"),
    "fetch": ("fetch/fetch_spherized.urdf", "fetch/fetch.srdf"),
    "baxter": ("baxter/baxter_spherized.urdf", "baxter/baxter.srdf"),
}


def bundled_robot_names() -> tuple[str, ...]:
    return tuple(_BUNDLED)


def bundled_urdf_path(name: str) -> str:
    return os.path.join(_RES_DIR, _BUNDLED[name][0])


def bundled_srdf_path(name: str) -> str | None:
    rel = _BUNDLED[name][1]
    return None if rel is None else os.path.join(_RES_DIR, rel)


def load_bundled_urdf(name: str) -> URDF:
    """Load one of the spherized robot models shipped with the package."""
    return URDF.load(bundled_urdf_path(name))
