"""Action -> per-substep primitive poses, in torch (differentiable, any device).

Restates the reference's host-side kinematics without pytorch3d:
* ``rigid_body_motion``        mpm/simulator.py:19-43
* ``primitive_poses``          MPMSimulator.compute_forward_kinematics, mpm/simulator.py:597-624
* ``rigid_body_motion_hand``   mpm/hand.py:20-65
* ``quaternion_to_matrix`` / ``matrix_to_quaternion`` / ``axis_angle_to_matrix``: the
  conversions mpm/hand.py takes from pytorch3d 0.7.2 (not installed here), written from
  their mathematical definitions; quaternions are (w, x, y, z).
"""
from __future__ import annotations

import copy

import numpy as np
import torch


def axis_angle_quat(rot: torch.Tensor) -> torch.Tensor:
    """Unit quaternion of an axis-angle vector with the reference's guards
    (w = sqrt(|r|^2 + 1e-16), r / clamp(w, 1e-7, 1e9); mpm/simulator.py:31-32)."""
    w = torch.sqrt((rot * rot).sum(dim=-1, keepdim=True) + 1e-16)
    return torch.cat((torch.cos(w / 2), (rot / torch.clamp(w, 1e-7, 1e9)) * torch.sin(w / 2)), -1)


def quat_mul(q: torch.Tensor, r: torch.Tensor) -> torch.Tensor:
    """The reference's quaternion composition from outer-product terms
    (mpm/simulator.py:35-41): the Hamilton product r * q, i.e. the increment r applied
    in the world frame after the orientation q."""
    t = q[..., :, None] * r[..., None, :]
    w = t[..., 0, 0] - t[..., 1, 1] - t[..., 2, 2] - t[..., 3, 3]
    x = t[..., 0, 1] + t[..., 1, 0] - t[..., 2, 3] + t[..., 3, 2]
    y = t[..., 0, 2] + t[..., 1, 3] + t[..., 2, 0] - t[..., 3, 1]
    z = t[..., 0, 3] - t[..., 1, 2] + t[..., 2, 1] + t[..., 3, 0]
    return torch.stack([w, x, y, z], -1)


def rigid_body_motion(states, actions: torch.Tensor):
    """states = (pos [B,3], quat [B,4]); actions [T,B,6] -> (pos [T,B,3], rot [T,B,4])."""
    T, B = actions.shape[:2]
    pos, q = states
    pos = pos[None, :].expand(T, -1, -1).reshape(-1, 3)
    q = q[None, :].expand(T, -1, -1).reshape(-1, 4)
    actions = actions.reshape(-1, 6)
    dq = axis_angle_quat(actions[:, 3:])
    next_pos = pos + actions[:, :3]
    out = quat_mul(q, dq)
    next_rot = out / torch.linalg.norm(out, dim=-1, keepdim=True)
    return next_pos.reshape(T, B, 3), next_rot.reshape(T, B, 4)


def primitive_poses(pos: torch.Tensor, rot: torch.Tensor, action: torch.Tensor, scales: torch.Tensor,
                    substeps: int):
    """Per-substep poses for primitive actions: a = clamp(a, -1, 1) * scale, substep i
    uses a * (i + 1) / substeps (mpm/simulator.py:616-623)."""
    device = pos.device
    action = action.reshape(-1, 6).clamp(-1.0, 1.0) * scales
    ramp = (torch.arange(substeps, device=device)[:, None, None] + 1) / substeps
    return rigid_body_motion((pos, rot), action[None, :].expand(substeps, -1, -1) * ramp)


# ---------------------------------------------------------------- rotation conversions
def quaternion_to_matrix(q: torch.Tensor) -> torch.Tensor:
    """Rotation matrix of a (w, x, y, z) quaternion (normalised by 2/|q|^2)."""
    r, i, j, k = torch.unbind(q, -1)
    two_s = 2.0 / (q * q).sum(-1)
    o = torch.stack((
        1 - two_s * (j * j + k * k), two_s * (i * j - k * r), two_s * (i * k + j * r),
        two_s * (i * j + k * r), 1 - two_s * (i * i + k * k), two_s * (j * k - i * r),
        two_s * (i * k - j * r), two_s * (j * k + i * r), 1 - two_s * (i * i + j * j),
    ), -1)
    return o.reshape(q.shape[:-1] + (3, 3))


def _sqrt_positive_part(x: torch.Tensor) -> torch.Tensor:
    """sqrt(max(x, 0)) with a zero gradient where x <= 0 (elementwise selects, no
    boolean indexing: that would synchronise the host with the device)."""
    pos = x > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, x, torch.ones_like(x))), torch.zeros_like(x))


def matrix_to_quaternion(m: torch.Tensor) -> torch.Tensor:
    """(w, x, y, z) quaternion of a rotation matrix.

    Uses the four candidate formulas (one per dominant diagonal term) and picks the
    best-conditioned one, so the result is stable for every rotation. The sign of the
    result is not normalised; q and -q rotate identically in the kernels."""
    batch = m.shape[:-2]
    m00, m01, m02, m10, m11, m12, m20, m21, m22 = torch.unbind(m.reshape(batch + (9,)), -1)
    q_abs = _sqrt_positive_part(torch.stack([
        1.0 + m00 + m11 + m22, 1.0 + m00 - m11 - m22, 1.0 - m00 + m11 - m22, 1.0 - m00 - m11 + m22], -1))
    cands = torch.stack([
        torch.stack([q_abs[..., 0] ** 2, m21 - m12, m02 - m20, m10 - m01], -1),
        torch.stack([m21 - m12, q_abs[..., 1] ** 2, m10 + m01, m02 + m20], -1),
        torch.stack([m02 - m20, m10 + m01, q_abs[..., 2] ** 2, m12 + m21], -1),
        torch.stack([m10 - m01, m20 + m02, m21 + m12, q_abs[..., 3] ** 2], -1),
    ], -2)
    flr = torch.tensor(0.1, dtype=q_abs.dtype, device=q_abs.device)
    cands = cands / (2.0 * torch.max(q_abs[..., None], flr))
    idx = q_abs.argmax(-1)
    return torch.gather(cands, -2, idx[..., None, None].expand(batch + (1, 4))).squeeze(-2)


def axis_angle_to_matrix(aa: torch.Tensor) -> torch.Tensor:
    """Rotation matrix of an axis-angle vector (via the half-angle quaternion)."""
    ang = torch.norm(aa, p=2, dim=-1, keepdim=True)
    half = ang * 0.5
    eps = 1e-6
    small = ang.abs() < eps
    sinc = torch.where(small, 0.5 - (ang * ang) / 48, torch.sin(half) / torch.where(small, torch.ones_like(ang), ang))
    q = torch.cat([torch.cos(half), aa * sinc], -1)
    return quaternion_to_matrix(q)


def rigid_body_motion_hand(state: torch.Tensor, actions: torch.Tensor, T: int) -> torch.Tensor:
    """Wrist pose ramp (mpm/hand.py:20-65): state [n_hands,4,4], actions [n_hands,6]
    -> [T, n_hands, 4, 4]; substep i applies actions * (i + 1) / T."""
    state = state[None].expand(T, -1, -1, -1).clone()
    actions = actions[None].expand(T, -1, -1).clone() * (torch.arange(T, device=actions.device)[:, None, None] + 1) / T
    state[..., :3, 3] += actions[..., :3]
    q = matrix_to_quaternion(state[..., :3, :3])
    dq = axis_angle_quat(actions[..., 3:])
    out = quat_mul(q, dq)
    out = out / torch.linalg.norm(out, dim=-1, keepdim=True)
    state[..., :3, :3] = quaternion_to_matrix(out)
    return state


# ---------------------------------------------------------------- articulated capsule hands
class HingeTree:
    """Kinematic tree of hinge joints carrying one collision primitive per link.

    Joint-velocity control as in HandSimulator.JointVel_Fk (mpm/hand.py:383-428):
    q_i = clamp(q0 + (clamp(a, -1, 1) * scale)[action_map] * (i + 1), lower, upper) for
    substep i (no division by the substep count, mpm/hand.py:414-415), then batched
    forward kinematics T_link = T_parent * Trans(offset) * Rot(axis, q) and
    geom = T_link * geom_offset (mpm/hand.py:347-381). Link order must be topological
    (parent index < link index, -1 = the base frame).
    """

    def __init__(self, parents, offsets, axes, geom_offsets, lower, upper, base, scale, action_map=None,
                 device="cpu", dtype=torch.float32):
        f = lambda a: torch.as_tensor(a, dtype=dtype, device=device)
        self.parents = list(int(p) for p in parents)
        self.n = len(self.parents)
        assert all(p < i for i, p in enumerate(self.parents))
        self.offsets = f(offsets).reshape(self.n, 3)
        self.axes = f(axes).reshape(self.n, 3)
        self.axes = self.axes / torch.linalg.norm(self.axes, dim=-1, keepdim=True)
        self.geom = f(geom_offsets).reshape(self.n, 4, 4)
        self.lower = f(lower).reshape(self.n)
        self.upper = f(upper).reshape(self.n)
        self.base = f(base).reshape(4, 4)
        self.scale = f(scale).reshape(-1)
        self.action_map = (torch.as_tensor(action_map, dtype=torch.long, device=device) if action_map is not None
                           else torch.arange(self.n, device=device))

    def fk(self, q: torch.Tensor):
        """q [..., n] joint angles -> (pos [..., n, 3], quat [..., n, 4]) of the primitives."""
        R = axis_angle_to_matrix(self.axes * q[..., None])
        batch = q.shape[:-1]
        T = torch.zeros(batch + (self.n, 4, 4), dtype=q.dtype, device=q.device)
        T[..., :3, :3] = R
        T[..., :3, 3] = self.offsets
        T[..., 3, 3] = 1.0
        frames = []
        for i, p in enumerate(self.parents):
            parent = self.base.expand(batch + (4, 4)) if p < 0 else frames[p]
            frames.append(parent @ T[..., i, :, :])
        G = torch.stack(frames, -3) @ self.geom
        return G[..., :3, 3], matrix_to_quaternion(G[..., :3, :3])

    def joint_targets(self, q0: torch.Tensor, action: torch.Tensor, substeps: int):
        a = (action.reshape(-1).clamp(-1.0, 1.0) * self.scale)[self.action_map]
        ramp = torch.arange(substeps, device=action.device, dtype=action.dtype)[:, None] + 1
        return torch.minimum(torch.maximum(q0[None] + a[None] * ramp, self.lower), self.upper)

    def poses(self, q0: torch.Tensor, action: torch.Tensor, substeps: int):
        """Per-substep primitive poses for one joint-velocity action; returns
        (pos [S, n, 3], rot [S, n, 4], q at the last substep)."""
        q = self.joint_targets(q0, action, substeps)
        pos, rot = self.fk(q)
        return pos, rot, q[-1]

    def trajectory(self, q0: torch.Tensor, actions: torch.Tensor, substeps: int):
        """Poses of every substep for a sequence of joint-velocity actions [E, n_act],
        each env step starting from the joints the previous one ended at:
        (pos [E*S+1, n, 3], rot [E*S+1, n, 4]); row 0 is the pose at q0."""
        p0, r0 = self.fk(q0)
        P, R = [p0[None]], [r0[None]]
        q = q0
        for a in actions:
            p, r, q = self.poses(q, a, substeps)
            P.append(p)
            R.append(r)
        return torch.cat(P), torch.cat(R)

    def to(self, dtype=None, device=None) -> "HingeTree":
        """Copy of the tree with its tensors converted (e.g. float64 for finite differences)."""
        out = copy.copy(self)
        for k in ("offsets", "axes", "geom", "lower", "upper", "base", "scale"):
            setattr(out, k, getattr(self, k).to(dtype=dtype, device=device))
        out.action_map = self.action_map.to(device=device)
        return out


def synthetic_hand(seed: int = 4, palm=(0.5, 0.34, 0.5), device="cpu"):
    """A hand-like hinge tree of 22 capsule links: 2 palm links hanging from a fixed base
    and 5 fingers of 4 links each, pointing down (-y). Capsule radius U[0.008, 0.015],
    half-length U[0.01, 0.025] (segment length = 2 * half-length); each capsule starts at
    its joint and the next link's joint sits at the end of its segment.
    Returns (tree, capsule args [22, 4], initial joint angles [22])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 22
    r = rng.uniform(0.008, 0.015, n)
    hl = rng.uniform(0.01, 0.025, n)
    parents = [-1, -1]
    offsets = [[0.0, 0.0, -0.02], [0.0, 0.0, 0.02]]
    axes = [[1.0, 0.0, 0.0], [1.0, 0.0, 0.0]]
    for fi in range(5):
        prev = fi % 2
        for li in range(4):
            parents.append(prev)
            offsets.append([(fi - 2) * 0.025, -0.01, 0.0] if li == 0 else [0.0, -2 * hl[prev], 0.0])
            axes.append([0.0, 0.0, 1.0] if li == 0 else [1.0, 0.0, 0.0])
            prev = len(parents) - 1
    geoms = np.tile(np.eye(4), (n, 1, 1))
    geoms[:, 1, 3] = -hl
    base = np.eye(4)
    base[:3, 3] = palm
    tree = HingeTree(parents, np.asarray(offsets), np.asarray(axes), geoms, np.full(n, -0.6), np.full(n, 0.6), base,
                     np.full(n, 0.33 * 0.002), device=device)
    args = np.zeros((n, 4), np.float32)
    args[:, 0] = r
    args[:, 1] = 2 * hl
    q0 = torch.as_tensor(rng.uniform(-0.2, 0.2, size=n), dtype=torch.float32, device=device)
    return tree, args, q0
