"""SRDF reader: the ``<disable_collisions>`` pairs that shrink the self-collision
pair list (``src/pyroffi/_robot_srdf_parser.py:4-19``)."""

from __future__ import annotations

from xml.etree import ElementTree as ET


def read_disabled_collisions_from_srdf(srdf_path) -> list[dict[str, str]]:
    root = ET.parse(srdf_path).getroot()
    return [
        {"link1": e.get("link1"), "link2": e.get("link2"), "reason": e.get("reason", "Unknown")}
        for e in root.findall("disable_collisions")
    ]
