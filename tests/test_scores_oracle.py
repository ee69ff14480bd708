"""Pins of the shape-score oracle (oracle/sinkhorn.py) on the CPU, by what the mathematics
fixes rather than by the oracle's own formulas:

* the debiased Sinkhorn divergence of a cloud with itself is 0;
* for p = 1 and a small blur it is the Wasserstein-1 distance: |t| for a translated cloud,
  and, with the iterations run to convergence, the optimal assignment cost of two small
  clouds of equal size (scipy's linear_sum_assignment solves that problem exactly);
* its gradient equals automatic differentiation of the last extrapolation step (torch,
  FP64), and, converged, twice the central finite differences of S (both extrapolated
  potentials depend on x);
* the Chamfer distance and its gradient against a k-d tree's nearest neighbours.
"""
import numpy as np
import pytest
import torch
from scipy.optimize import linear_sum_assignment
from scipy.spatial import cKDTree

from oracle import sinkhorn as SK


def _clouds(n=60, m=60, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, (n, 3)), rng.uniform(0, 1, (m, 3))


def test_divergence_of_a_cloud_with_itself_is_zero():
    x, _ = _clouds()
    for p in (1, 2):
        assert abs(SK.sinkhorn(x, x.copy(), p=p, blur=1e-4 if p == 1 else 0.05)[0]) < 1e-12


def test_translated_cloud_gives_the_translation_length():
    x, _ = _clouds()
    t = np.array([0.05, -0.02, 0.03])
    S = SK.sinkhorn(x, x + t, p=1, blur=1e-4)[0]
    assert abs(S - np.linalg.norm(t)) <= 1e-9 * np.linalg.norm(t)


def test_converged_divergence_equals_the_exact_assignment_cost():
    x, y = _clouds(seed=1)
    C = SK.cost(x, y, 1)
    r, c = linear_sum_assignment(C)
    w1 = C[r, c].mean()
    S = SK.sinkhorn(x, y, p=1, blur=1e-3, extra_iters=5000)[0]
    assert abs(S - w1) <= 1e-3 * w1, (S, w1)


def _autograd_last_step(x, y, p, blur):
    """S and dS/dx by torch autograd of the last extrapolation step, its input potentials
    (from the oracle's loop) held fixed."""
    n, m = len(x), len(y)
    a_log, b_log = np.full(n, -np.log(n)), np.full(m, -np.log(m))
    Cs = {k: SK.cost(u, v, p) for k, (u, v) in dict(xy=(x, y), yx=(y, x), xx=(x, x), yy=(y, y)).items()}
    eps_list = SK.epsilon_schedule(p, SK.diameter(x, y), blur, 0.5)
    eps = eps_list[0]
    f_aa, g_bb = SK.softmin(eps, Cs["xx"], a_log), SK.softmin(eps, Cs["yy"], b_log)
    g_ab, f_ba = SK.softmin(eps, Cs["yx"], a_log), SK.softmin(eps, Cs["xy"], b_log)
    for eps in eps_list:
        ft = SK.softmin(eps, Cs["xy"], b_log + g_ab / eps), SK.softmin(eps, Cs["yx"], a_log + f_ba / eps)
        dt = SK.softmin(eps, Cs["xx"], a_log + f_aa / eps), SK.softmin(eps, Cs["yy"], b_log + g_bb / eps)
        f_ba, g_ab = 0.5 * (f_ba + ft[0]), 0.5 * (g_ab + ft[1])
        f_aa, g_bb = 0.5 * (f_aa + dt[0]), 0.5 * (g_bb + dt[1])
    xt, yt = torch.tensor(x, requires_grad=True), torch.tensor(y)

    def C(u, v):
        d2 = ((u[:, None, :] - v[None, :, :]) ** 2).sum(-1)
        if p == 2:
            return 0.5 * d2
        # |x - y| with a zero derivative at coincident points (the diagonal of C(x, x))
        return torch.where(d2 > 0, torch.sqrt(torch.where(d2 > 0, d2, torch.ones_like(d2))), torch.zeros_like(d2))

    def sm(Cm, h):
        return -eps * torch.logsumexp(torch.tensor(h)[None, :] - Cm / eps, 1)

    S = ((sm(C(xt, yt), b_log + g_ab / eps) - sm(C(xt, xt), a_log + f_aa / eps)).mean() +
         (sm(C(yt, xt), a_log + f_ba / eps) - sm(C(yt, yt), b_log + g_bb / eps)).mean())
    S.backward()
    return float(S), xt.grad.numpy()


@pytest.mark.parametrize("p,blur", [(1, 1e-4), (1, 0.05), (2, 0.05)])
def test_gradient_is_autograd_of_the_last_step(p, blur):
    x, y = _clouds(40, 50, seed=2)
    S, g, _ = SK.sinkhorn(x, y, p=p, blur=blur)
    S2, g2 = _autograd_last_step(x, y, p, blur)
    assert abs(S - S2) <= 1e-12 * abs(S2)
    assert np.linalg.norm(g - g2) <= 1e-10 * np.linalg.norm(g2)


@pytest.mark.parametrize("p", [1, 2])
def test_converged_gradient_is_twice_the_finite_differences(p):
    x, y = _clouds(30, 40, seed=3)
    blur, extra, h = 0.05, 400, 1e-6
    d = SK.diameter(x, y)
    g = SK.sinkhorn(x, y, p=p, blur=blur, extra_iters=extra)[1]
    fd = np.zeros((4, 3))
    for i in range(4):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd[i, k] = (SK.sinkhorn(xp, y, p=p, blur=blur, diam=d, extra_iters=extra)[0] -
                        SK.sinkhorn(xm, y, p=p, blur=blur, diam=d, extra_iters=extra)[0]) / (2 * h)
    assert np.linalg.norm(g[:4] - 2 * fd) <= 0.05 * np.linalg.norm(2 * fd)


def test_chamfer_against_kdtree_nearest_neighbours():
    x, y = _clouds(200, 300, seed=4)
    val, g = SK.chamfer(x, y)
    dx, ix = cKDTree(y).query(x)
    dy, iy = cKDTree(x).query(y)
    assert abs(val - ((dx ** 2).mean() + (dy ** 2).mean())) <= 1e-12
    g2 = 2.0 / len(x) * (x - y[ix])
    for j, i in enumerate(iy):
        g2[i] += 2.0 / len(y) * (x[i] - y[j])
    np.testing.assert_allclose(g, g2, atol=1e-14)


def test_device_score_schedule_and_argument_checks():
    """The device implementation's epsilon schedule is the oracle's; CPU tensors and
    unsupported options are rejected (there is no CPU implementation)."""
    from dexdeform_amd.mpm import losses as L
    for p, d, blur in ((1, 1.7, 1e-4), (2, 3.2, 0.05), (1, 1e-4, 1e-4)):
        assert L.epsilon_schedule(p, d, blur, 0.5) == SK.epsilon_schedule(p, d, blur, 0.5)
    x = torch.zeros(10, 3)
    with pytest.raises(ValueError):
        L.SamplesLoss("sinkhorn", p=1, blur=1e-4)(x, x)
    with pytest.raises(ValueError):
        L.chamfer_distance(x, x)
    with pytest.raises(ValueError):
        L.SamplesLoss("energy")
    with pytest.raises(ValueError):
        L.SamplesLoss("sinkhorn", p=3)
