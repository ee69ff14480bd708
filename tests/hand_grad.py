"""Action gradients of the hand-contact rollout (22-link capsule hand pressing into an
ellipsoid; joint-velocity actions; L2 loss of the final positions to a seeded target):

* gpu_action_grad: HIP rollout with recompute checkpointing, body-pose gradients chained
  through the torch hinge-tree kinematics on the device;
* oracle_action_grad: the CPU oracle's adjoint (FP32 or FP64) through the same kinematics;
* fd_action_grad: FP64 central finite differences of the oracle's FP64 forward pass
  (exact-Jacobi SVD, so the loss is smooth in the actions).

Shared by tests/test_gpu_hand.py (reduced size) and tests/check_hand_gradients.py
(the full 50k-particle, 200-substep configuration).
"""
from __future__ import annotations

import numpy as np
import torch

from oracle import oracle as O
from tests.helpers import initial_state, oracle_scene
from dexdeform_amd.mpm import workloads as W

SUBSTEPS = 10  # substeps per env step


def setup(n_particles, n_env_steps):
    sc, tree, q0 = W.hand_scene(n_particles)
    sc.spec.adjoint_mode = 1  # exact adjoint: finite differences are the reference
    acts = W.hand_actions(n_env_steps)
    target = W.plain_target(sc.x0)
    return sc, tree, q0, acts, target


def gpu_action_grad(sc, tree, q0, acts, target, device, checkpoint_every=10):
    from dexdeform_amd.mpm.engine import Engine
    from dexdeform_amd.mpm.rollout import Rollout
    T = len(acts) * SUBSTEPS
    eng = Engine([sc.spec], device=device)
    ro = Rollout(eng, T, checkpoint_every=checkpoint_every)
    ro.set_initial(torch.tensor(sc.x0, device=device))
    a = torch.tensor(acts, device=device, requires_grad=True)
    P, R = tree.to(device=device).trajectory(q0.to(device), a, SUBSTEPS)
    ro.set_poses(P.detach(), R.detach())
    xT = ro.forward().x
    d = xT - torch.tensor(target, device=device)
    loss = (d * d).sum(-1).mean()
    ro.backward(x_grad=(2.0 / sc.n) * d)
    ((P * ro.pos_grad).sum() + (R * ro.rot_grad).sum()).backward()
    torch.cuda.synchronize()
    return a.grad.double().cpu().numpy(), float(loss.item())


def oracle_action_grad(sc, tree, q0, acts, target, dtype):
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    a = torch.tensor(acts, dtype=tdt, requires_grad=True)
    P, R = tree.to(dtype=tdt).trajectory(q0.to(tdt), a, SUBSTEPS)
    osim = O.OracleSim(oracle_scene(sc), dtype, adjoint="exact")
    states = osim.rollout(initial_state(sc).astype(dtype), P.detach().numpy(), R.detach().numpy())
    gx = (2.0 / sc.n) * (states[-1].x.astype(np.float64) - target)
    nb = sc.nb
    g_last = O.Grads(gx.astype(dtype), np.zeros((sc.n, 3), dtype), np.zeros((sc.n, 3, 3), dtype),
                     np.zeros((sc.n, 3, 3), dtype), np.zeros((nb, 3), dtype), np.zeros((nb, 4), dtype))
    _, pg, rg = osim.rollout_grad(states, g_last)
    ((P * torch.tensor(pg, dtype=tdt)).sum() + (R * torch.tensor(rg, dtype=tdt)).sum()).backward()
    return a.grad.double().numpy()


def _fd_loss(sc, tree64, q0, a64, target):
    P, R = tree64.trajectory(q0.double(), a64, SUBSTEPS)
    osim = O.OracleSim(oracle_scene(sc), np.float64, adjoint="exact")
    st = initial_state(sc).astype(np.float64)
    for t in range(1, len(P)):
        st = osim.substep(st, P[t].numpy(), R[t].numpy())
    d = st.x - target.astype(np.float64)
    return float((d * d).sum(-1).mean())


def fd_action_grad(sc, tree, q0, acts, target, coords, eps=1e-4):
    tree64 = tree.to(dtype=torch.float64)
    base = torch.tensor(acts, dtype=torch.float64)
    out = []
    with O.exact_svd():
        for (e, k) in coords:
            ap, am = base.clone(), base.clone()
            ap[e, k] += eps
            am[e, k] -= eps
            out.append((_fd_loss(sc, tree64, q0, ap, target) - _fd_loss(sc, tree64, q0, am, target)) / (2 * eps))
    return np.array(out)


def random_coords(n_env_steps, n_joints, k, seed=5):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(n_env_steps)), int(rng.integers(n_joints))) for _ in range(k)]
