"""Hand-contact differentiation on the GPU (reduced size of the 50k-particle, 200-substep
configuration: 5,000 particles, 4 env steps x 10 substeps, checkpoints every 10).

Gradients of the loss w.r.t. joint-velocity actions, from the HIP adjoint chained through
the torch hand kinematics, against FP64 central finite differences (eps = 1e-4) of the
oracle on 8 random action coordinates. The FP32 adjoint of the reference algorithm is
itself only as close to the finite differences as its rounding allows (the SVD adjoint
divides by squared singular-value gaps), so the bound is max(1e-3, 2 x the distance of
the oracle's FP32 adjoint from the finite differences); the oracle's FP64 adjoint must
match the finite differences within 1e-3 (about 1.5e-4 at this size: contact switches
and the friction branch make the loss only piecewise smooth at eps = 1e-4), which pins
the adjoint formulas.
"""
import numpy as np
import pytest
import torch

from tests import hand_grad as H
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def test_hand_action_gradients_vs_fp64_finite_differences():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    E = 4
    sc, tree, q0, acts, target = H.setup(5000, E)
    coords = H.random_coords(E, 22, 8)
    g_gpu, loss = H.gpu_action_grad(sc, tree, q0, acts, target, torch.device("cuda:0"))
    g32 = H.oracle_action_grad(sc, tree, q0, acts, target, np.float32)
    g64 = H.oracle_action_grad(sc, tree, q0, acts, target, np.float64)
    fd = H.fd_action_grad(sc, tree, q0, acts, target, coords)
    pick = lambda g: np.array([g[e, k] for e, k in coords])
    e_gpu, e32, e64 = rel_err(pick(g_gpu), fd), rel_err(pick(g32), fd), rel_err(pick(g64), fd)
    e_all = rel_err(g_gpu, g64)
    noise = rel_err(g32, g64)
    print(f"loss {loss:.6e}; vs FD: gpu {e_gpu:.2e}, oracle fp32 {e32:.2e}, oracle fp64 {e64:.2e}; "
          f"all actions gpu vs oracle fp64 {e_all:.2e} (oracle fp32 vs fp64 {noise:.2e})")
    assert e64 <= 1e-3
    assert e_gpu <= max(1e-3, 2 * e32)
    assert e_all <= max(1e-3, 2 * noise)
