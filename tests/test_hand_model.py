"""The hand model on the CPU: the MJCF reader and the hand kinematics against an
independent composition of the model file's transforms, the joint-velocity control
against its closed form, and the Shadow-hand tables against the model file.

The independent composition walks the XML with ElementTree and scipy rotations: a body's
frame is its parent's frame times Trans(pos) Rot(orientation) (MuJoCo's intrinsic x-y-z
euler, (w, x, y, z) quaternions, axis-angle), then times Trans(joint pos) Rot(axis, q) if it
has a driven joint; a collision geom's frame is its body's frame times the geom's frame,
capsules turned by pi/2 about x (the kernel's capsule axis is y). Where the Shadow-hand files
are absent (they belong to the reference checkout), the tests that read them are skipped;
the stored configuration (tests/golden/shadow_rh_scale1p5.npz) is then checked alone.
"""
import os
import xml.etree.ElementTree as et

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from dexdeform_amd.mpm import hand as HM
from dexdeform_amd.mpm import mjcf, robots

ASSETS = "/root/reference/mpm/assets/robots"
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shadow_rh_scale1p5.npz")
have_assets = os.path.exists(os.path.join(ASSETS, "shadow", "right_hand", "shadow_hand.xml"))
needs_assets = pytest.mark.skipif(not have_assets, reason="Shadow-hand MJCF files not present")


def _expanded_root(path):
    root = et.parse(path).getroot()
    base = os.path.dirname(path)
    for parent in [p for p in root.iter() if any(k.tag == "include" for k in p)]:
        kids = list(parent)
        for k in kids:
            parent.remove(k)
        for k in kids:
            if k.tag == "include":
                parent.extend(list(_expanded_root(os.path.join(base, k.attrib["file"]))))
            else:
                parent.append(k)
    return root


def _frame(x, scale):
    m = np.eye(4)
    m[:3, 3] = np.array([float(v) for v in x.attrib["pos"].split()]) * scale
    if "euler" in x.attrib:
        m[:3, :3] = Rotation.from_euler("XYZ", [float(v) for v in x.attrib["euler"].split()]).as_matrix()
    elif "quat" in x.attrib:
        w, a, b, c = (float(v) for v in x.attrib["quat"].split())
        m[:3, :3] = Rotation.from_quat([a, b, c, w]).as_matrix()
    elif "axisangle" in x.attrib:
        v = np.array([float(t) for t in x.attrib["axisangle"].split()])
        m[:3, :3] = Rotation.from_rotvec(v[:3] / np.linalg.norm(v[:3]) * v[3]).as_matrix()
    return m


def _independent_geoms(path, scale, q, wrist_frame):
    """World frames of the collision geoms (group 4, not on the forearm) with the wrist
    body at `wrist_frame` and joint angles q (JOINTS order)."""
    root = _expanded_root(path)
    classes = {d.attrib["class"]: {e.tag: e.attrib for e in d} for d in root.find("default").findall("default")}
    jidx = {n: i for i, n in enumerate(robots.JOINTS)}
    out = []

    def walk(body, frame):
        j = body.find("joint")
        if j is not None and j.attrib.get("name") in jidx:
            t = np.eye(4)
            t[:3, 3] = [float(v) for v in j.attrib["pos"].split()]
            axis = np.array([float(v) for v in j.attrib["axis"].split()])
            t[:3, :3] = Rotation.from_rotvec(axis * q[jidx[j.attrib["name"]]]).as_matrix()
            frame = frame @ t
        for g in body.findall("geom"):
            attrib = dict(classes.get(g.attrib.get("class"), {}).get("geom", {}))
            attrib.update(g.attrib)
            if attrib.get("group") != "4" or attrib.get("type") not in ("capsule", "box"):
                continue
            gf = _frame(g, scale)
            if attrib["type"] == "capsule":
                gf[:3, :3] = gf[:3, :3] @ Rotation.from_euler("x", np.pi / 2).as_matrix()
            out.append(frame @ gf)
        for b in body.findall("body"):
            walk(b, frame @ _frame(b, scale))

    wrist = next(b for b in root.iter("body") if b.attrib.get("name") == "robot0:wrist")
    walk(wrist, wrist_frame)
    return np.stack(out)


@needs_assets
@pytest.mark.parametrize("zero", [True, False])
def test_kinematics_match_independent_composition_of_the_model_file(zero):
    scale = 1.5
    prims, cfg = HM.hand_config("rh", ASSETS, scale=scale)
    kin = HM.HandKinematics(cfg, device="cpu", dtype=torch.float64)
    rng = np.random.default_rng(0)
    q = np.zeros(24) if zero else rng.uniform(-0.5, 0.5, 24)
    wrist = HM.root_matrix([0.5, 0.2, 0.3], [0.3, -0.2, np.pi])
    pos, rot = kin.forward_kinematics(torch.tensor(wrist)[None, None], torch.tensor(q)[None, None])
    ref = _independent_geoms(os.path.join(ASSETS, "shadow", "right_hand", "shadow_hand.xml"), scale, q, wrist)
    assert ref.shape[0] == 19 == pos.shape[1]
    np.testing.assert_allclose(pos[0].numpy(), ref[:, :3, 3], atol=1e-12)
    got_r = Rotation.from_quat(rot[0].numpy()[:, [1, 2, 3, 0]]).as_matrix()
    np.testing.assert_allclose(got_r, ref[:, :3, :3], atol=1e-12)
    # primitive sizes: capsules (radius, half-length), boxes (half extents), times the scale
    shapes = [p["shape"] for p in prims]
    assert shapes.count("Capsule") == 16 and shapes.count("Box") == 3
    assert np.allclose(prims[0]["size"], [0.015 * scale, 0.01 * scale])


@needs_assets
def test_joint_limits_lie_in_the_model_files_joint_ranges():
    root = _expanded_root(os.path.join(ASSETS, "shadow", "right_hand", "shadow_hand.xml"))
    ranges = {j.attrib["name"]: [float(v) for v in j.attrib["range"].split()] for j in root.iter("joint")
              if "range" in j.attrib}
    for name in robots.JOINTS:
        lo, hi = robots.JOINT_LIMITS[name]
        assert abs(lo - ranges[name][0]) < 2.5e-3 and abs(hi - ranges[name][1]) < 2.5e-3, name


@needs_assets
def test_stored_configuration_matches_the_model_file():
    p1, c1 = HM.hand_config("rh", ASSETS, scale=1.5, hand_friction=0.9)
    p2, c2 = HM.load_hand_config(GOLDEN)
    assert p1 == p2
    k1, k2 = HM.HandKinematics(c1, device="cpu"), HM.HandKinematics(c2, device="cpu")
    q = torch.rand(4, 1, 24)
    b = torch.tensor(c1["root_frame"], dtype=torch.float32)[None].expand(4, -1, -1, -1)
    for a, c in zip(k1.forward_kinematics(b, q), k2.forward_kinematics(b, q)):
        assert torch.equal(a, c)


@needs_assets
def test_two_hand_configuration_stacks_both_hands():
    prims, cfg = HM.hand_config("dual", ASSETS, scale=1.0)
    assert cfg["n_hands"] == 2 and len(prims) == 38
    kin = HM.HandKinematics(cfg, device="cpu", dtype=torch.float64)
    base = torch.tensor(np.stack([HM.root_matrix([0.3, 0.2, 0.5], [0, 0, 0]), HM.root_matrix([0.7, 0.2, 0.5],
                                                                                                [0, 0, 0])]))[None]
    pos, _ = kin.forward_kinematics(base, torch.zeros(1, 2, 24, dtype=torch.float64))
    assert pos.shape == (1, 38, 3)
    assert pos[0, :19, 0].mean() < 0.45 < pos[0, 19:, 0].mean()


def test_stored_configuration_loads_without_the_model_files():
    prims, cfg = HM.load_hand_config(GOLDEN)
    n, kw = HM.hand_bodies(prims)
    assert n == 19 and kw["types"].count(1) == 16 and kw["types"].count(0) == 3
    assert all(len(a) == 4 for a in kw["args"]) and set(kw["softness"]) == {666.0}
    kin = HM.HandKinematics(cfg, device="cpu")
    pos, rot = kin.forward_kinematics(torch.tensor(cfg["root_frame"], dtype=torch.float32)[None],
                                      torch.zeros(1, 1, 24))
    assert pos.shape == (1, 19, 3) and torch.allclose(rot.norm(dim=-1), torch.ones(1, 19))


def test_joint_velocity_control_closed_form():
    """Joints move by clamp(a, -1, 1)[actuator] * scale * (i + 1) at substep i (coupled
    joints share their actuator's action), clamped to the joint limits; the wrist ramps its
    6-DOF action linearly over the step, reaching the full action at the last substep."""
    _, cfg = HM.load_hand_config(GOLDEN)
    kin = HM.HandKinematics(cfg, device="cpu", dtype=torch.float64)
    S = 40
    scale = torch.tensor([0.33 * 0.002] * 20 + [0.01] * 3 + [0.015] * 3, dtype=torch.float64)
    q0 = torch.tensor([robots.DEFAULT_INITIAL_QPOS[j] for j in robots.JOINTS], dtype=torch.float64)[None]
    rng = np.random.default_rng(1)
    a = torch.tensor(rng.uniform(-2, 2, size=(1, 26)))
    a[0, 3] = 50.0  # FFJ2 runs into its upper limit
    base0 = torch.tensor(HM.root_matrix([0.5, 0.2, 0.3], [0, 0, np.pi]))[None]
    base, q = kin.joint_velocity(base0, q0, a, scale, S)
    amap = robots.get_actuator_mapping()
    for i in (0, 17, S - 1):
        for j, name in enumerate(robots.JOINTS):
            lo, hi = robots.JOINT_LIMITS[name]
            want = q0[0, j] + float(np.clip(a[0, amap[j]], -1, 1)) * 0.33 * 0.002 * (i + 1)
            assert abs(float(q[i, 0, j]) - float(np.clip(want, lo, hi))) < 1e-12
    assert float(q[0, 0, 3]) == pytest.approx(min(q0[0, 3] + 0.33 * 0.002, robots.JOINT_LIMITS["robot0:FFJ2"][1]))
    # coupled joints (FFJ1, FFJ0) move together
    assert torch.allclose(q[:, 0, 4] - q0[0, 4], q[:, 0, 5] - q0[0, 5])
    # wrist: translation a * scale * (i + 1) / S, full rotation a * scale at the last substep
    for i in (0, S - 1):
        np.testing.assert_allclose(base[i, 0, :3, 3].numpy(), (base0[0, :3, 3] + a[0, 20:23] * scale[20:23] *
                                                                (i + 1) / S).numpy(), atol=1e-12)
    # the rotation increment is applied in the world frame: R_S = dR R_0
    dr = Rotation.from_matrix(base[S - 1, 0, :3, :3].numpy()) * Rotation.from_matrix(base0[0, :3, :3].numpy()).inv()
    np.testing.assert_allclose(dr.as_rotvec(), (a[0, 23:] * scale[23:]).numpy(), atol=1e-9)
    # 20 actions only: the wrist stays
    base2, _ = kin.joint_velocity(base0, q0, a[:, :20], scale, S)
    assert torch.equal(base2[S - 1], base0)
