"""Drop-in ``HandSimulator`` for DexDeform's mpm/hand.py: the Shadow hand's joint-velocity
control, its batched forward kinematics in torch on the device, and the hand state
(wrist frame and joint angles per environment step) carried beside the particle state.

* ``rigid_body_motion_hand`` -- wrist ramp (mpm/hand.py:20-65), from ``kinematics``;
* ``HandKinematics``         -- HandSimulator.hand_forward_kinematics / JointVel_Fk
                                (mpm/hand.py:347-428) on the tables of a hand configuration;
* ``HandSimulator``          -- mpm/hand.py:68-344, 383-434 on the fused HIP substep;
* ``hand_config`` / ``hand_bodies`` / ``root_matrix`` -- what HandEnv.parse_sim_cfg,
  CudaEnv.parse_tools and HandEnv.get_root_matrix build (mpm/hand.py:578-650,
  mpm/cuda_env.py:51-102), without the yacs configuration objects.

Everything runs in torch on the simulator's device (FP32, as the reference); the quaternion
conversions are those of ``kinematics`` (the reference takes them from pytorch3d).
"""
from __future__ import annotations

import math
import os
from typing import List, Optional, Sequence

import numpy as np
import torch

from .kinematics import axis_angle_to_matrix, matrix_to_quaternion, rigid_body_motion_hand  # noqa: F401
from .mjcf import homogeneous, prepare
from .robots import ACTUATOR_JOINT_MAPPING, DEFAULT_INITIAL_QPOS, JOINT_LIMITS, JOINTS, get_actuator_mapping
from .simulator import MPMSimulator, _cfg_namespace

GEOMS_PER_HAND = 19
LH_PRIM_RANGE = list(range(0, 19))
RH_PRIM_RANGE = list(range(19, 38))
SINGLE_PRIM_RANGE = list(range(0, 19))
DUAL_PRIM_RANGE = list(range(0, 38))


def _capsule_frame_fix() -> np.ndarray:
    """pi/2 about x: the kernel's capsule axis is y, MuJoCo's is z (mpm/hand.py:160-162)."""
    c, s = math.cos(math.pi / 2), math.sin(math.pi / 2)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


class HandKinematics:
    """Forward kinematics of ``n_hands`` hands from a hand configuration (the dictionary
    ``hand_config`` returns, or the reference's HandEnv builds).

    Joint j's frame is its chain predecessor's frame times the fixed frames between them
    times T_j = [axis_angle(axis_j * q_j) | pos_j]; each collision geom's frame is its
    joint's frame times the geom's frame (capsules turned by pi/2 about x). The reference
    walks ``site_computations`` site by site; the chains share their prefixes (the wrist
    and palm joints open every finger's chain), so here the joints are grouped by depth in
    the tree the chains describe and each depth is one batched product, with the same
    association per joint ((parent frame @ fixed frames) @ T_j, left to right)."""

    def __init__(self, hand_cfg: dict, device="cuda:0", dtype=torch.float32):
        self.device = torch.device(device)
        self.dtype = dtype
        f = lambda a: torch.as_tensor(np.asarray(a), dtype=dtype, device=self.device)  # noqa: E731
        self.n_hands = int(hand_cfg["n_hands"])
        self.n_joints = len(JOINTS)
        sc = hand_cfg["site_computations"]
        # chains: per site, a list of ("mat", [n_hands, 4, 4]) / ("joint", j) steps from the
        # wrist frame; the fixed frame after a site's last joint (the site itself) moves nothing
        self.chains = []
        for site, inds in enumerate(sc["inds"]):
            steps = []
            for kind, k in inds:
                if kind == "joint_idx":
                    steps.append(("joint", int(k)))
                else:
                    steps.append(("mat", f(sc["mats"][site][k]).reshape(self.n_hands, 4, 4)))
            while steps and steps[-1][0] == "mat":
                steps.pop()
            self.chains.append(steps)
        # the tree: parent joint (-1: the wrist frame) and fixed frames before each joint,
        # identical in every chain that passes through it; joints grouped by depth
        parent, pre = [None] * self.n_joints, [None] * self.n_joints
        for steps in self.chains:
            p, mats = -1, []
            for kind, v in steps:
                if kind == "mat":
                    mats.append(v)
                    continue
                if parent[v] is None:
                    parent[v], pre[v] = p, mats
                elif parent[v] != p or len(pre[v]) != len(mats) or any(
                        not torch.equal(a, b) for a, b in zip(pre[v], mats)):
                    raise ValueError(f"joint {v} is reached along different paths")
                p, mats = v, []
        missing = [j for j in range(self.n_joints) if parent[j] is None]
        if missing:
            raise ValueError(f"joints {missing} are in no chain")
        depth = [0] * self.n_joints
        for j in range(self.n_joints):  # chains list ancestors before descendants
            d, k = 0, j
            while parent[k] >= 0:
                d, k = d + 1, parent[k]
            depth[j] = d
        if any(len(m) > 1 for m in pre):
            raise ValueError("more than one fixed frame between two joints is not supported")
        eye = torch.eye(4, dtype=dtype, device=self.device).expand(self.n_hands, 4, 4)
        # per depth: (joints, slots of their parents in the frame stack, fixed frames
        # [n_hands, n, 4, 4]); the stack holds the wrist frame at slot 0, then each depth's
        # joints in order
        self.levels = []
        slot = {-1: 0}
        for d in range(max(depth) + 1):
            js = [j for j in range(self.n_joints) if depth[j] == d]
            par = [slot[parent[j]] for j in js]
            M = torch.stack([pre[j][0] if pre[j] else eye for j in js], 1)
            self.levels.append((torch.as_tensor(js, dtype=torch.long, device=self.device),
                                torch.as_tensor(par, dtype=torch.long, device=self.device), M))
            first = len(slot)
            for i, j in enumerate(js):
                slot[j] = first + i
        self.frame_slot = torch.as_tensor([slot[j] for j in range(self.n_joints)], dtype=torch.long, device=self.device)
        ji = hand_cfg["joint_info"]
        self.joint_pos = f(np.stack([np.asarray(p) for p, _ in ji])).reshape(self.n_joints, self.n_hands, 3)
        self.joint_pos = self.joint_pos.transpose(0, 1).contiguous()
        self.joint_axis = f(np.stack([np.asarray(a) for _, a in ji])).reshape(self.n_joints, self.n_hands, 3)
        self.joint_axis = self.joint_axis.transpose(0, 1).contiguous()
        mats, index = [], []
        for g in hand_cfg["geometries"]["col"]:
            if "forearm" in g.parent_name:
                continue  # forearm geoms do not move with the hand (mpm/hand.py:150-154)
            if g.is_mesh:
                raise ValueError(f"collision geom of {g.parent_name} is a mesh; only capsules and boxes are simulated")
            index.append(JOINTS.index(g.parent_name))
            m = np.array(g.matrix, dtype=np.float64)
            if g.data["type"] == "capsule":
                m[..., :3, :3] = m[..., :3, :3] @ _capsule_frame_fix()
            mats.append(m)
        per_hand = len(mats) // self.n_hands
        self.geom_mats = f(np.stack(mats)).reshape(self.n_hands, per_hand, 4, 4)
        self.geom_index = torch.as_tensor(index[:per_hand], dtype=torch.long, device=self.device)
        self.n_geoms = per_hand * self.n_hands
        self.n_actuators = len(ACTUATOR_JOINT_MAPPING)
        self.q_lower = f([JOINT_LIMITS[j][0] for j in JOINTS])[None, None, :]
        self.q_upper = f([JOINT_LIMITS[j][1] for j in JOINTS])[None, None, :]
        self.action_map = torch.as_tensor(get_actuator_mapping(), dtype=torch.long, device=self.device)

    def joint_velocity(self, curr_base: torch.Tensor, curr_q: torch.Tensor, actions: torch.Tensor,
                       action_scale: torch.Tensor, substeps: int):
        """Wrist frames [S, n_hands, 4, 4] and joint angles [S, n_hands, 24] of the substeps
        of one environment step (mpm/hand.py:397-418): with 6 wrist actions after the 20
        actuator actions, the wrist follows rigid_body_motion_hand (ramp (i + 1) / S); joint j
        moves by clamp(a, -1, 1)[actuator(j)] * scale * (i + 1), not divided by S, and is
        clamped to its limits."""
        if actions.shape[1] == self.n_actuators + 6:
            next_base = rigid_body_motion_hand(curr_base, actions[:, -6:] * action_scale[None, -6:], substeps)
        else:
            next_base = curr_base[None, ...].expand(substeps, -1, -1, -1)
        a = (actions[..., :self.n_actuators].clamp(-1.0, 1.0) * action_scale[None, :self.n_actuators])[:, self.action_map]
        ramp = torch.arange(substeps, device=actions.device)[:, None, None] + 1
        next_q = (curr_q[None, :] + a[None, :] * ramp).clamp(self.q_lower.to(a.dtype), self.q_upper.to(a.dtype))
        return next_base, next_q

    def forward_kinematics(self, base_pose: torch.Tensor, q: torch.Tensor):
        """base_pose [S, n_hands, 4, 4] wrist frames, q [S, n_hands, 24] joint angles ->
        (pos [S, n_hands * geoms, 3], rot [S, n_hands * geoms, 4]) of the collision geoms
        (mpm/hand.py:347-381)."""
        S = base_pose.shape[0]
        T = torch.zeros(S, self.n_hands, self.n_joints, 4, 4, dtype=q.dtype, device=q.device)
        T[..., :3, :3] = axis_angle_to_matrix(self.joint_axis.to(q.dtype)[None] * q[..., None])
        T[..., :3, 3] = self.joint_pos.to(q.dtype)
        T[..., 3, 3] = 1.0
        # frame stack: slot 0 the wrist, then the joints depth by depth
        F = base_pose[..., None, :, :]
        for js, par, M in self.levels:
            f = (F.index_select(-3, par) @ M.to(q.dtype)) @ T.index_select(-3, js)
            F = torch.cat([F, f], -3)
        J = F.index_select(-3, self.frame_slot)
        G = J[..., self.geom_index, :, :] @ self.geom_mats.to(q.dtype)
        G = G.reshape(S, -1, 4, 4)
        return G[..., :3, 3], matrix_to_quaternion(G[..., :3, :3])


def _to_tensor(a, device, dtype=torch.float32):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype)
    return torch.tensor(np.asarray(a), device=device, dtype=dtype)


class HandSimulator(MPMSimulator):
    """The hand-driven simulator (mpm/hand.py:68-434): substep dt = 5e-5 / quality,
    dx = 1 / (64 quality), 2 ms of simulated time per environment step; joint-velocity
    control (``ctrl_type`` "vel"); the wrist frame ``base_pose[f]`` and joint angles
    ``joint_rot[f]`` of environment-step boundaries f are kept per state index."""

    def __init__(self, n_bodies, hand_cfg, cfg=None, quality=None, action_scale=None, device="cuda:0", mode=None,
                 scale=None, ctrl_type=None, hand_friction=None, fixed_base=False, **kwargs):
        c = _cfg_namespace(cfg, dict(quality=quality, ctrl_type=ctrl_type, fixed_base=fixed_base))
        quality = float(c.quality if c.quality is not None else 1)
        fixed_base = bool(getattr(c, "fixed_base", fixed_base))
        dt = 0.5e-4 / quality
        dx = 1.0 / (64 * quality)
        substeps = int(np.ceil(2e-3 / dt))
        # the step sizes follow from quality whatever cfg says (explicit arguments win over
        # the configuration, as in the reference's configurable classes)
        base = {k: v for k, v in vars(c).items() if k not in ("dt", "dx", "substeps")}
        super().__init__(n_bodies, cfg=base, dt=dt, dx=dx, substeps=substeps, device=device, **kwargs)
        self.ctrl_type = getattr(c, "ctrl_type", None) or "vel"
        if self.ctrl_type != "vel":
            raise ValueError(f"ctrl type {self.ctrl_type} is not supported!")
        self.compute_forward_kinematics = self.JointVel_Fk
        self.n_hands = int(hand_cfg["n_hands"])
        self.n_joints_per_hand = len(JOINTS)
        self.n_actuators_per_hand = len(ACTUATOR_JOINT_MAPPING)
        self.kin = HandKinematics(hand_cfg, device=self.device)
        if self.kin.n_geoms != n_bodies:
            raise ValueError(f"hand configuration has {self.kin.n_geoms} collision geoms, simulator {n_bodies} bodies")
        self.base_pose: List[Optional[torch.Tensor]] = [None] * self.max_steps
        self.joint_rot: List[Optional[torch.Tensor]] = [None] * self.max_steps
        self.q_lower, self.q_upper, self.action_map = self.kin.q_lower, self.kin.q_upper, self.kin.action_map
        if action_scale is None:
            action_scale = [0.33 * 0.002] * self.n_actuators_per_hand
            action_scale = action_scale + ([0.0] * 6 if fixed_base else [0.01] * 3 + [0.015] * 3)
        self.torch_action_scale = self.togpu(np.array(action_scale))
        self.base_pose[0] = self.togpu(hand_cfg["root_frame"]).reshape(self.n_hands, 4, 4)
        self.joint_rot[0] = self.togpu([DEFAULT_INITIAL_QPOS[j] for j in JOINTS])[None, :].expand(self.n_hands, -1)

    def togpu(self, x, dtype=torch.float32):
        return torch.tensor(np.array(x), dtype=dtype, device=self.device)

    # ------------------------------------------------------------------ kinematics
    def hand_forward_kinematics(self, base_pose, q):
        """Batched forward kinematics: base_pose [S, n_hands, 4, 4], q [S, n_hands, 24] ->
        geom positions [S, n_bodies, 3] and rotations [S, n_bodies, 4] (mpm/hand.py:347-381)."""
        return self.kin.forward_kinematics(base_pose, q)

    def mat2pos_rot(self, mat):
        return mat[..., :3, 3], matrix_to_quaternion(mat[..., :3, :3])

    def JointVel_Fk(self, f, actions, pos_rot=None):
        """Per-substep geom poses of one environment step from joint-velocity actions
        [n_hands, 20] (+ 6 wrist actions, ramped over the step): joint j moves by
        clamp(a, -1, 1)[actuator(j)] * scale * (i + 1) at substep i -- not divided by the
        substep count -- clamped to the joint limits (mpm/hand.py:383-428). Returns
        (pos [S, n_bodies, 3], rot [S, n_bodies, 4], (wrist frame, joint angles) at the end
        of the step), which are also stored at index f + substeps."""
        device = self.device
        if pos_rot is None:
            curr_base, curr_q = self.base_pose[f], self.joint_rot[f]
        else:
            curr_base, curr_q = pos_rot
        actions = _to_tensor(actions, device) if not isinstance(actions, torch.Tensor) else actions.to(device)
        if actions.shape[0] != self.n_hands:
            raise ValueError(f"actions for {actions.shape[0]} hands, simulator has {self.n_hands}")
        next_base, next_q = self.kin.joint_velocity(curr_base, curr_q, actions, self.torch_action_scale,
                                                    self.substeps)
        geom_pos, geom_rot = self.hand_forward_kinematics(next_base, next_q)
        nb, nq = next_base[-1], next_q[-1]
        if f + self.substeps < self.max_steps:
            self.base_pose[f + self.substeps] = nb
            self.joint_rot[f + self.substeps] = nq
        return geom_pos, geom_rot, (nb, nq)

    # ------------------------------------------------------------------ state
    def download_pos_rot(self, cur, device):
        """(wrist frame, joint angles) at index 0; forgets the later ones (mpm/hand.py:179-184)."""
        assert cur == 0
        for i in range(1, self.max_steps):
            self.base_pose[i] = self.joint_rot[i] = None
        return self.base_pose[cur].to(device), self.joint_rot[cur].to(device)

    def get_state(self, index):
        """(x, v, F, C, pose per body ..., wrist frames [n_hands, 4, 4], joint angles
        [n_hands, 24]) at an environment-step boundary (mpm/hand.py:186-191)."""
        assert index % self.substeps == 0
        return super().get_state(index) + (self.base_pose[index].detach().cpu().numpy(),
                                           self.joint_rot[index].detach().cpu().numpy())

    def set_state(self, index, state):
        """Sets particles, wrist frames and joint angles; the geom poses follow from the
        kinematics (mpm/hand.py:193-211)."""
        assert index % self.substeps == 0
        self.base_pose[index] = torch.tensor(np.asarray(state[-2]), device=self.device, dtype=torch.float32)
        self.joint_rot[index] = torch.tensor(np.asarray(state[-1]), device=self.device, dtype=torch.float32)
        pos, rot = self.hand_forward_kinematics(self.base_pose[index][None], self.joint_rot[index][None])
        pos = pos[0].detach().cpu().numpy()
        rot = rot[0].detach().cpu().numpy()
        particles = [np.float32(a) for a in state[:4]]
        super().set_state(index, particles + [np.r_[p, r] for p, r in zip(pos, rot)])

    def get_state_render_only(self, f=0, device="numpy"):
        return self.get_x(f), self.base_pose[f].cpu().numpy(), self.joint_rot[f].cpu().numpy()

    def set_state_render_only(self, p, base_pose, joint_rot, f=0):
        base_pose = torch.tensor(np.asarray(base_pose), device=self.device, dtype=torch.float32)
        joint_rot = torch.tensor(np.asarray(joint_rot), device=self.device, dtype=torch.float32)
        pos, rot = self.hand_forward_kinematics(base_pose[None], joint_rot[None])
        self.states[f].x.upload(self._to_internal(np.float32(p)))
        self.states[f].body_pos.upload(pos[0])
        self.states[f].body_rot.upload(rot[0])

    def step(self, action, q_state=None):
        super().step(action, q_state)
        self.base_pose[0] = self.base_pose[self.substeps]
        self.joint_rot[0] = self.joint_rot[self.substeps]


# ---------------------------------------------------------------- configuration helpers
def asset_path(asset_dir: str, mode: str) -> List[str]:
    names = {"lh": ["left_hand"], "rh": ["right_hand"], "dual": ["left_hand", "right_hand"],
             "lh+rh": ["left_hand", "right_hand"]}
    if mode not in names:
        raise ValueError(f"incorrect hand mode: {mode}")
    return [os.path.join(asset_dir, "shadow", n, "shadow_hand.xml") for n in names[mode]]


def hand_config(mode: str, asset_dir: str, scale: float = 1.0, hand_friction: float = 0.9):
    """(primitives, hand_cfg) for hand mode "lh", "rh" or "dual" from the Shadow-hand MJCF
    files under ``asset_dir`` (``<asset_dir>/shadow/{left,right}_hand/shadow_hand.xml``),
    as HandEnv.parse_sim_cfg builds them (mpm/hand.py:578-650): one primitive per
    capsule/box collision geom (friction ``hand_friction``, round 0), per-hand tables with
    a leading hand axis, the arm frame root * hand mount and the wrist frame arm * wrist."""
    parts = [prepare(p, scale=scale) for p in asset_path(asset_dir, mode)]
    n_hands = len(parts)
    root_info = {k: np.stack([p[0][k] for p in parts]) for k in parts[0][0]}
    sc0 = parts[0][1]
    site_computations = {"inds": sc0["inds"],
                         "mats": [[np.stack([p[1]["mats"][i][j] for p in parts]) for j in range(len(sc0["mats"][i]))]
                                  for i in range(len(sc0["mats"]))]}
    joint_info = [tuple(np.stack([p[2][j][k] for p in parts]) for k in range(2)) for j in range(len(JOINTS))]
    geometries = {t: sum((p[3][t] for p in parts), []) for t in ("vis", "col")}
    primitives = []
    for g in geometries["col"]:
        if "type" not in g.data:
            continue
        shape = {"capsule": "Capsule", "box": "Box"}.get(g.data["type"])
        if shape is None:
            continue
        primitives.append({"shape": shape, "size": list(g.data["size"]), "round": 0.0, "friction": hand_friction,
                           "action": {"scale": ()}, "init_pos": (0.0, 0.0, 0.0), "init_rot": (1.0, 0.0, 0.0, 0.0)})
    arm_frame = root_info["root_matrix"] @ root_info["robot0:hand mount"]
    root_frame = arm_frame @ root_info["robot0:wrist"]
    hand_cfg = {"n_hands": n_hands, "arm_frame": arm_frame, "root_frame": root_frame,
                "site_computations": site_computations, "joint_info": joint_info, "geometries": geometries}
    return primitives, hand_cfg


def hand_bodies(primitives: Sequence[dict]):
    """(n_bodies, init_bodies keyword arguments) of a primitive list (mpm/cuda_env.py:51-102):
    box type 0 with args (half extents, 0), capsule type 1 with (radius, half-length, 0, 0),
    softness 666."""
    kw = {k: [] for k in ("types", "softness", "mu", "round", "args", "action_scales", "pos", "rot")}
    for p in primitives:
        if p["shape"] == "Box":
            kw["types"].append(0)
            kw["args"].append([*p["size"], 0])
        elif p["shape"] == "Capsule":
            kw["types"].append(1)
            kw["args"].append([*p["size"], 0, 0])
        else:
            raise ValueError(f"unknown primitive shape {p['shape']}")
        kw["action_scales"].append(p.get("action", {}).get("scale", ()))
        kw["softness"].append(666.0)
        kw["mu"].append(p["friction"])
        kw["round"].append(p["round"])
        kw["pos"].append(p.get("init_pos", (0.0, 0.0, 0.0)))
        kw["rot"].append(p.get("init_rot", (1.0, 0.0, 0.0, 0.0)))
    return len(primitives), kw


def root_matrix(pos, rot) -> np.ndarray:
    """Wrist placement from a position and static-axis x-y-z euler angles (R = Rz Ry Rx),
    as HandEnv.get_root_matrix (mpm/hand.py:652-654) builds it with transforms3d."""
    a, b, c = (float(v) for v in rot)
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return homogeneous(np.asarray(pos, np.float64), rz @ ry @ rx)


# ---------------------------------------------------------------- configuration files
def save_hand_config(path: str, primitives: Sequence[dict], hand_cfg: dict):
    """Writes a hand configuration (the output of ``hand_config``) as plain arrays (.npz,
    no pickled objects), so that a simulator can be built where the MJCF files are absent."""
    sc = hand_cfg["site_computations"]
    kinds, ks, site_of, mats = [], [], [], []
    for site, inds in enumerate(sc["inds"]):
        for kind, k in inds:
            kinds.append(1 if kind == "joint_idx" else 0)
            ks.append(int(k))
            site_of.append(site)
    n_mats = [len(m) for m in sc["mats"]]
    for m in sc["mats"]:
        mats.extend(np.asarray(x, np.float64) for x in m)
    col = hand_cfg["geometries"]["col"]
    sizes = np.zeros((len(col), 3))
    for i, g in enumerate(col):
        if not g.is_mesh:
            sizes[i, :len(g.data["size"])] = g.data["size"]
    np.savez(path, n_hands=int(hand_cfg["n_hands"]), arm_frame=np.asarray(hand_cfg["arm_frame"], np.float64),
             root_frame=np.asarray(hand_cfg["root_frame"], np.float64), step_kind=np.int32(kinds),
             step_k=np.int32(ks), step_site=np.int32(site_of), site_n_mats=np.int32(n_mats), site_mats=np.stack(mats),
             joint_pos=np.stack([np.asarray(p, np.float64) for p, _ in hand_cfg["joint_info"]]),
             joint_axis=np.stack([np.asarray(a, np.float64) for _, a in hand_cfg["joint_info"]]),
             geom_parent=np.array([g.parent_name for g in col]), geom_matrix=np.stack([g.matrix for g in col]),
             geom_is_mesh=np.array([g.is_mesh for g in col]),
             geom_type=np.array([g.data.get("type", "") for g in col]), geom_size=sizes,
             geom_size_len=np.int32([0 if g.is_mesh else len(g.data["size"]) for g in col]),
             prim_friction=np.float64([p["friction"] for p in primitives]))


def load_hand_config(path: str):
    """(primitives, hand_cfg) written by ``save_hand_config``."""
    from .mjcf import MujocoGeom
    z = np.load(path, allow_pickle=False)
    n_sites = len(z["site_n_mats"])
    inds = [[] for _ in range(n_sites)]
    for kind, k, site in zip(z["step_kind"], z["step_k"], z["step_site"]):
        inds[int(site)].append(("joint_idx" if kind else "body_idx", int(k)))
    mats, off = [], 0
    for n in z["site_n_mats"]:
        mats.append([z["site_mats"][off + j] for j in range(int(n))])
        off += int(n)
    col = []
    for name, m, mesh, t, size, sl in zip(z["geom_parent"], z["geom_matrix"], z["geom_is_mesh"], z["geom_type"],
                                          z["geom_size"], z["geom_size_len"]):
        data = {"mesh": ""} if mesh else {"type": str(t), "size": [float(v) for v in size[:int(sl)]]}
        col.append(MujocoGeom(str(name), np.array(m), bool(mesh), data))
    hand_cfg = {"n_hands": int(z["n_hands"]), "arm_frame": z["arm_frame"], "root_frame": z["root_frame"],
                "site_computations": {"inds": inds, "mats": mats},
                "joint_info": [(p, a) for p, a in zip(z["joint_pos"], z["joint_axis"])],
                "geometries": {"vis": [], "col": col}}
    primitives, fi = [], 0
    for g in col:
        if g.is_mesh or g.data["type"] not in ("capsule", "box"):
            continue
        primitives.append({"shape": "Capsule" if g.data["type"] == "capsule" else "Box", "size": g.data["size"],
                           "round": 0.0, "friction": float(z["prim_friction"][fi]), "action": {"scale": ()},
                           "init_pos": (0.0, 0.0, 0.0), "init_rot": (1.0, 0.0, 0.0, 0.0)})
        fi += 1
    return primitives, hand_cfg
