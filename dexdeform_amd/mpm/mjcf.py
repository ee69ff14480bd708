"""Reading the hand's MuJoCo model (MJCF) into the tables of the hand kinematics.

Produces what mpm/mujoco_parser.py:prepare (lines 378-550) produces from the same files,
in the same structure, so a hand configuration built here drives ``HandSimulator``
exactly as the reference's does:

* ``root_info``: 4x4 frames of the world placement and of the named root bodies;
* ``site_computations``: for every fingertip site, the chain from the hand's root body to
  the site as a list of ("joint_idx", joint) and ("body_idx", k) steps, the k-th fixed
  frame of that site's chain being ``mats[site][k]`` (fixed bodies between two joints
  folded into one frame);
* ``joint_info``: (pos, axis) of every hinge joint, in the given joint order;
* ``geometries``: the visual (group 1) and collision (group 4) geoms, each with the name of
  the joint (or body, if the body has no joint) it hangs from and its 4x4 frame in that body.

Format handling (MJCF, https://mujoco.readthedocs.io/en/stable/XMLreference.html):
``<include file>`` elements are replaced by the included file's top-level children; a
``class`` attribute fills the element's missing attributes from that ``<default class>``'s
entry for the element's tag. Orientations: ``euler`` (radians, MuJoCo's default sequence:
intrinsic x, y, z, i.e. R = Rx Ry Rz), ``axisangle`` (axis, angle) and ``quat`` (w, x, y, z,
normalised). Positions of bodies, geoms and sites and geom sizes are multiplied by
``scale``; joint positions are not. Two conventions of the reference's reader are kept
because hand poses depend on them: a geom without a ``pos`` attribute gets the identity
frame (its orientation attributes are then not read), and ``childclass`` defaults are not
applied (the reference defines them but never applies them; for the Shadow hand they only
carry contact parameters).
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as et
from typing import Dict, List, Optional, Tuple

import numpy as np

from .robots import FINGERTIP_SITE_NAMES, JOINTS

ROOT_BODY = "robot0:hand mount"
ROOT_POS = np.array([1.0, 1.25, 0.15])
ROOT_EULER = np.array([np.pi / 2, 0.0, np.pi])
_ROOT_PARENT = "NONE"


# ---------------------------------------------------------------- orientations
def euler_xyz_to_matrix(e) -> np.ndarray:
    """MuJoCo's euler convention: rotations about the body's x, then new y, then new z
    axes (R = Rx(e0) Ry(e1) Rz(e2))."""
    a, b, c = (float(v) for v in e)
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return rx @ ry @ rz


def quat_to_matrix(q) -> np.ndarray:
    """Rotation of a (w, x, y, z) quaternion of any non-zero norm (identity for ~zero)."""
    q = np.asarray(q, np.float64)
    n = float(q @ q)
    if n <= np.finfo(np.float64).eps:
        return np.eye(3)
    w, x, y, z = q
    s = 2.0 / n
    return np.array([
        [1 - s * (y * y + z * z), s * (x * y - w * z), s * (x * z + w * y)],
        [s * (x * y + w * z), 1 - s * (x * x + z * z), s * (y * z - w * x)],
        [s * (x * z - w * y), s * (y * z + w * x), 1 - s * (x * x + y * y)],
    ])


def axis_angle_quat(axis, angle) -> np.ndarray:
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    q = np.concatenate([[np.cos(angle / 2)], np.sin(angle / 2) * axis])
    return q / np.linalg.norm(q)


def homogeneous(pos, rot) -> np.ndarray:
    m = np.eye(4)
    m[:3, :3] = rot
    m[:3, 3] = pos
    return m


def _floats(s: Optional[str]) -> np.ndarray:
    return np.array([float(v) for v in s.split()], np.float64)


def element_frame(x: et.Element, scale: float = 1.0) -> np.ndarray:
    """4x4 frame of a body / geom / site relative to its parent body."""
    pos = _floats(x.attrib["pos"]) * scale
    if "euler" in x.attrib:
        rot = euler_xyz_to_matrix(_floats(x.attrib["euler"]))
    elif "axisangle" in x.attrib:
        aa = _floats(x.attrib["axisangle"])
        rot = quat_to_matrix(axis_angle_quat(aa[:3], aa[3]))
    elif "quat" in x.attrib:
        rot = quat_to_matrix(_floats(x.attrib["quat"]))
    else:
        rot = np.eye(3)
    return homogeneous(pos, rot)


# ---------------------------------------------------------------- document
def load(path: str) -> et.Element:
    """Root element of an MJCF file with its includes expanded (paths relative to the
    including file) and class defaults applied."""
    root = _load_expanded(path)
    _apply_classes(root)
    return root


def _load_expanded(path: str) -> et.Element:
    with open(path) as f:
        root = et.parse(f).getroot()
    base = os.path.dirname(os.path.abspath(path))
    parents = [p for p in root.iter() if any(k.tag == "include" for k in p)]
    for parent in parents:
        kids = list(parent)
        new = []
        for k in kids:
            if k.tag == "include":
                new.extend(list(_load_expanded(os.path.join(base, k.attrib["file"]))))
            else:
                new.append(k)
        for k in kids:
            parent.remove(k)
        parent.extend(new)
    return root


def _apply_classes(root: et.Element):
    defaults = root.find("default")
    classes: Dict[str, Dict[str, dict]] = {}
    if defaults is not None:
        for d in defaults.findall("default"):
            classes[d.attrib["class"]] = {e.tag: dict(e.attrib) for e in d}
    wb = root.find("worldbody")
    if wb is None:
        return
    for x in wb.iter():
        cls = x.attrib.get("class")
        if cls is None or cls not in classes:
            continue
        for k, v in classes[cls].get(x.tag, {}).items():
            x.attrib.setdefault(k, v)


# ---------------------------------------------------------------- hand tables
class MujocoGeom:
    """One geom: the joint (or body) it hangs from, its frame in that body, whether it is
    a mesh, and its data ({"type", "size"} or the mesh file)."""

    def __init__(self, parent_name: str, matrix: np.ndarray, is_mesh: bool, data: dict):
        self.parent_name = parent_name
        self.matrix = matrix
        self.is_mesh = is_mesh
        self.data = data


def prepare(xml_path: str, scale: float = 1.0, root_body: str = ROOT_BODY,
            target_sites: List[str] = FINGERTIP_SITE_NAMES, joint_names: List[str] = JOINTS):
    """(root_info, site_computations, joint_info, geometries) of a hand model; see the
    module docstring."""
    doc = load(xml_path)
    site_idx = {n: i for i, n in enumerate(target_sites)}
    joint_idx = {n: i for i, n in enumerate(joint_names)}
    site_info: List[Optional[Tuple[np.ndarray, str]]] = [None] * len(target_sites)
    joint_info: List[Optional[Tuple[np.ndarray, np.ndarray]]] = [None] * len(joint_names)
    body_info: Dict[str, Tuple[np.ndarray, str]] = {}
    body_joint: Dict[str, str] = {}
    vis, col = [], []

    def visit(body: et.Element, parent: str):
        name = body.attrib.get("name", "noname_body_%d" % len(body_info))
        body_info[name] = (element_frame(body, scale), parent)
        jname = None
        j = body.find("joint")
        if j is not None:
            jname = j.attrib.get("name", "")
            if jname in joint_idx:
                if j.attrib.get("type", "hinge") != "hinge":
                    raise ValueError(f"joint {jname}: only hinge joints are supported")
                if name in body_joint:
                    raise ValueError(f"body {name} has more than one driven joint")
                joint_info[joint_idx[jname]] = (_floats(j.attrib["pos"]), _floats(j.attrib["axis"]))
                body_joint[name] = jname
        for g in body.findall("geom"):
            frame = element_frame(g, scale) if g.attrib.get("pos", "") else np.eye(4)
            mesh = g.attrib.get("mesh", "")
            if mesh:
                data = {"mesh": mesh, "scale": g.attrib.get("scale", "")}
            else:
                data = {"type": g.attrib.get("type", ""), "size": [float(v) * scale for v in g.attrib["size"].split()]}
            geom = MujocoGeom(jname if jname is not None else name, frame, bool(mesh), data)
            group = g.attrib.get("group", "")
            if group == "1":
                vis.append(geom)
            elif group == "4":
                col.append(geom)
        for s in body.findall("site"):
            if s.attrib.get("name", "") in site_idx:
                site_info[site_idx[s.attrib["name"]]] = (element_frame(s, scale), name)
        for b in body.findall("body"):
            visit(b, name)

    wb = doc.find("worldbody")
    roots = [b for b in wb.findall("body") if b.attrib.get("name", "") == root_body]
    if not roots:
        raise ValueError(f"no body named {root_body!r} at the top of the worldbody")
    visit(roots[0], _ROOT_PARENT)
    missing = [n for n, s in zip(target_sites, site_info) if s is None] + \
        [n for n, j in zip(joint_names, joint_info) if j is None]
    if missing:
        raise ValueError(f"model lacks {missing}")

    # per site: walk from the site up to the root body; a body with a joint closes a fixed
    # frame (everything between it and the previous joint) and adds the joint
    inds: List[List[Tuple[str, int]]] = []
    mats: List[List[np.ndarray]] = []
    for frame, body in site_info:
        steps, fixed = [], []
        m = frame
        while body != _ROOT_PARENT:
            pm, parent = body_info[body]
            jn = body_joint.get(body, "")
            if jn:
                steps.append(("body_idx", len(fixed)))
                fixed.append(m)
                steps.append(("joint_idx", joint_idx[jn]))
                m = pm
            else:
                m = pm @ m
            body = parent
        inds.append(list(reversed(steps)))
        mats.append(fixed)

    def body_frame(name):
        for b in wb.iter("body"):
            if b.attrib.get("name", "") == name:
                return element_frame(b, scale)
        raise ValueError(f"no body {name!r}")

    root_info = {"root_matrix": homogeneous(ROOT_POS, euler_xyz_to_matrix(ROOT_EULER))}
    for name in (root_body, "robot0:forearm", "robot0:wrist"):
        root_info[name] = body_frame(name)
    return root_info, {"mats": mats, "inds": inds}, joint_info, {"vis": vis, "col": col}
