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

import torch


def axis_angle_quat(rot: torch.Tensor) -> torch.Tensor:
    """Unit quaternion of an axis-angle vector with the reference's guards
    (w = sqrt(|r|^2 + 1e-16), r / clamp(w, 1e-7, 1e9); mpm/simulator.py:31-32)."""
    w = torch.sqrt((rot * rot).sum(dim=-1, keepdim=True) + 1e-16)
    return torch.cat((torch.cos(w / 2), (rot / torch.clamp(w, 1e-7, 1e9)) * torch.sin(w / 2)), -1)


def quat_mul(q: torch.Tensor, r: torch.Tensor) -> torch.Tensor:
    """Hamilton product q * r, written as the reference's outer-product terms
    (mpm/simulator.py:35-41)."""
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
    ret = torch.zeros_like(x)
    pos = x > 0
    ret[pos] = torch.sqrt(x[pos])
    return ret


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
