"""Full-size hand-contact gradient check (not collected by pytest; needs a GPU and a few
minutes of CPU time for the FP64 finite differences):

    python tests/check_hand_gradients.py [--particles 50000] [--env-steps 20] [--coords 16]

50,000 particles in a seeded ellipsoid, the 22-link capsule hand, 20 env steps x 10
substeps = 200 substeps with recompute checkpoints every 10, exact adjoint. Prints the
relative error of the GPU action gradients against FP64 central finite differences
(eps = 1e-4) on random action coordinates, next to the oracle's FP32 and FP64 adjoints,
as one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from tests import hand_grad as H  # noqa: E402
from tests.helpers import rel_err  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--particles", type=int, default=50000)
    ap.add_argument("--env-steps", type=int, default=20)
    ap.add_argument("--coords", type=int, default=16)
    ap.add_argument("--checkpoint", type=int, default=10)
    ap.add_argument("--no-oracle-adjoint", action="store_true")
    a = ap.parse_args()
    sc, tree, q0, acts, target = H.setup(a.particles, a.env_steps)
    coords = H.random_coords(a.env_steps, 22, a.coords)
    pick = lambda g: np.array([g[e, k] for e, k in coords])
    t0 = time.time()
    g_gpu, loss = H.gpu_action_grad(sc, tree, q0, acts, target, torch.device("cuda:0"), a.checkpoint)
    print(f"gpu rollout + adjoint done in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    out = {"particles": a.particles, "substeps": a.env_steps * H.SUBSTEPS, "checkpoint_every": a.checkpoint,
           "coords": coords, "loss": loss}
    if not a.no_oracle_adjoint:
        for name, dt in (("oracle_fp32", np.float32), ("oracle_fp64", np.float64)):
            t0 = time.time()
            out[name] = pick(H.oracle_action_grad(sc, tree, q0, acts, target, dt))
            print(f"{name} adjoint done in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    t0 = time.time()
    fd = []
    for i, c in enumerate(coords):  # one coordinate at a time, with a progress line each
        fd.append(H.fd_action_grad(sc, tree, q0, acts, target, [c])[0])
        print(f"finite difference {i + 1}/{len(coords)} {c}: {fd[-1]:+.6e} ({time.time() - t0:.0f} s)",
              file=sys.stderr, flush=True)
    fd = np.array(fd)
    res = {"gpu_vs_fd": rel_err(pick(g_gpu), fd)}
    for name in ("oracle_fp32", "oracle_fp64"):
        if name in out:
            res[name + "_vs_fd"] = rel_err(out.pop(name), fd)
    out.update(res)
    out["gpu"] = pick(g_gpu).tolist()
    out["fd"] = fd.tolist()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
