"""Shared test helpers: build the oracle's view of a workload scene, compare arrays."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def oracle_scene(sc, dtype=np.float32):
    s = sc.spec
    return O.Scene(grid_dim=tuple(int(g) for g in s.grid_dim), dx=float(np.float32(s.dx)),
                   dt=float(np.float32(s.dt)), gravity=tuple(np.float32(s.gravity)), mass=np.asarray(s.mass),
                   vol=np.asarray(s.vol), mly=np.asarray(s.mu_lam_yield), tfsr=np.asarray(s.body_prim),
                   args=np.asarray(s.body_args), ground_friction=float(s.ground_friction),
                   ground_height=float(s.ground_height))


def initial_state(sc, rng=None, v_scale=0.0, f_scale=0.0, c_scale=0.0):
    n = sc.n
    rng = rng or np.random.default_rng(0)
    v = (rng.normal(size=(n, 3)) * v_scale).astype(np.float32)
    F = (np.eye(3)[None] + rng.normal(size=(n, 3, 3)) * f_scale).astype(np.float32)
    C = (rng.normal(size=(n, 3, 3)) * c_scale).astype(np.float32)
    return O.State(sc.x0.copy(), v, F, C, sc.body_pos.copy(), sc.body_rot.copy())


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    if den == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / den)


def poses_from_actions(sc, actions, substeps, scales):
    """Per-substep body poses [T+1, nb, 3|4] from primitive actions (torch CPU, fp32)."""
    import torch
    from dexdeform_amd.mpm.kinematics import primitive_poses
    pos = torch.tensor(sc.body_pos)
    rot = torch.tensor(sc.body_rot)
    P, R = [pos.numpy().copy()], [rot.numpy().copy()]
    sc_t = torch.tensor(np.asarray(scales, np.float32))
    for a in actions:
        p, r = primitive_poses(pos, rot, torch.tensor(a), sc_t, substeps)
        for t in range(substeps):
            P.append(p[t].numpy().copy())
            R.append(r[t].numpy().copy())
        pos, rot = p[-1], r[-1]
    return np.stack(P).astype(np.float32), np.stack(R).astype(np.float32)
