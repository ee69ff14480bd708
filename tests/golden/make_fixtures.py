"""Writes the FP64-oracle fixtures that the full-size GPU gradient tests compare against.
Runs on the CPU with the oracle built (minutes of OpenMP work):

    python tests/golden/make_fixtures.py [hand] [large]

hand_c2_fp64.npz -- the hand-contact configuration at full size: 50,000 particles in the
  seeded ellipsoid, the 22-link capsule hand, 20 env steps x 10 substeps, exact adjoint mode,
  L2 loss of the final positions to the seeded target. Holds the FP64 oracle's action
  gradient [20, 22] (every state kept, no checkpoints: the adjoint is the same), 16 seeded
  action coordinates, and FP64 central finite differences of the oracle's forward pass (exact
  Jacobi SVD, so the loss is smooth in the actions) at those coordinates for eps 1e-4 and
  2.5e-5.
large_c5_30k_fp64.npz -- the large-scene configuration (128^3 grid, particles in the union
  of 3 seeded boxes, the 22-link hand, 50 env steps x 10 substeps) at 30,000 particles: the
  FP64 oracle's action gradient [50, 22] and the loss.

The fixtures are inputs/outputs of this build's own oracle (a restatement of the reference,
whose CUDA code cannot run here); the GPU tests regenerate the same scenes from the seeds.
"""
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from tests import hand_grad as H  # noqa: E402
from tests.helpers import initial_state, oracle_scene  # noqa: E402
from dexdeform_amd.mpm import workloads as W  # noqa: E402

SUB = 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fp64_adjoint(sc, tree, q0, acts, target):
    """FP64 oracle (OpenMP) forward storing every state, then its adjoint, chained through
    the hinge-tree kinematics in FP64. Returns (loss, action gradient)."""
    A = torch.tensor(acts, dtype=torch.float64, requires_grad=True)
    P, R = tree.to(dtype=torch.float64).trajectory(q0.double(), A, SUB)
    osim = O.OracleSim(oracle_scene(sc), np.float64, adjoint="exact", omp=True)
    states = osim.rollout(initial_state(sc).astype(np.float64), P.detach().numpy(), R.detach().numpy())
    d = states[-1].x - target.astype(np.float64)
    loss = float((d * d).sum(-1).mean())
    z = lambda *s: np.zeros(s, np.float64)  # noqa: E731
    gl = O.Grads((2.0 / sc.n) * d, z(sc.n, 3), z(sc.n, 3, 3), z(sc.n, 3, 3), z(sc.nb, 3), z(sc.nb, 4))
    _, pg, rg = osim.rollout_grad(states, gl)
    ((P * torch.tensor(pg)).sum() + (R * torch.tensor(rg)).sum()).backward()
    return loss, A.grad.numpy().copy()


def fp64_loss(sc, tree64, q0, a64, target):
    P, R = tree64.trajectory(q0.double(), a64, SUB)
    osim = O.OracleSim(oracle_scene(sc), np.float64, adjoint="exact", omp=True)
    st = initial_state(sc).astype(np.float64)
    for t in range(1, len(P)):
        st = osim.substep(st, P[t].numpy(), R[t].numpy())
    d = st.x - target.astype(np.float64)
    return float((d * d).sum(-1).mean())


def make_hand():
    E = 20
    sc, tree, q0, acts, target = H.setup(50000, E)
    t0 = time.time()
    loss, g = fp64_adjoint(sc, tree, q0, acts, target)
    log(f"hand: FP64 adjoint done in {time.time() - t0:.0f} s, loss {loss:.6e}")
    coords = np.array(H.random_coords(E, 22, 16), np.int64)
    tree64 = tree.to(dtype=torch.float64)
    base = torch.tensor(acts, dtype=torch.float64)
    fd = {}
    with O.exact_svd():
        for eps in (1e-4, 2.5e-5):
            vals = []
            for i, (e, k) in enumerate(coords):
                ap, am = base.clone(), base.clone()
                ap[e, k] += eps
                am[e, k] -= eps
                vals.append((fp64_loss(sc, tree64, q0, ap, target) - fp64_loss(sc, tree64, q0, am, target)) / (2 * eps))
                log(f"hand: FD eps {eps:g} {i + 1}/16 ({time.time() - t0:.0f} s)")
            fd[eps] = np.array(vals)
    pick = g[coords[:, 0], coords[:, 1]]
    for eps, v in fd.items():
        log(f"hand: FP64 adjoint vs FD(eps {eps:g}): {np.linalg.norm(pick - v) / np.linalg.norm(v):.3e}")
    np.savez(os.path.join(HERE, "hand_c2_fp64.npz"), n_particles=50000, env_steps=E, substeps_per_env_step=SUB,
             loss=loss, grad=g, coords=coords, fd_eps_1e4=fd[1e-4], fd_eps_2p5e5=fd[2.5e-5])


def make_large():
    E, N = 50, 30000
    sc, tree, q0 = W.large_scene(N)
    sc.spec.adjoint_mode = 1
    acts = W.hand_actions(E)
    target = W.plain_target(sc.x0)
    t0 = time.time()
    loss, g = fp64_adjoint(sc, tree, q0, acts, target)
    log(f"large: FP64 adjoint done in {time.time() - t0:.0f} s, loss {loss:.6e}")
    np.savez(os.path.join(HERE, "large_c5_30k_fp64.npz"), n_particles=N, env_steps=E, substeps_per_env_step=SUB,
             loss=loss, grad=g)


if __name__ == "__main__":
    which = sys.argv[1:] or ["hand", "large"]
    if "large" in which:
        make_large()
    if "hand" in which:
        make_hand()
