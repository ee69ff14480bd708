"""CPU oracle (test infrastructure) for the shape scores: the debiased Sinkhorn divergence
and the Chamfer distance, dense FP64 numpy on small point clouds.

DexDeform scores shapes with the geomloss package (tutorials/4_computing_score.ipynb cell 1:
``SamplesLoss(loss="sinkhorn", p=1, blur=0.0001)``; cell 7 scores the simulator's particles
against a target cloud). geomloss is a third-party dependency installed unpinned
(``pip install geomloss``, README.md:14-19) and is not present here; this restates its
published single-scale algorithm (J. Feydy et al., AISTATS 2019; J. Feydy's PhD thesis,
Alg. 3.1-3.2, as geomloss 0.2's ``sinkhorn_divergence`` implements it):

* cost C(x, y) = |x - y| for p = 1, |x - y|^2 / 2 for p = 2; uniform weights a = 1/n, b = 1/m;
* epsilon schedule: diameter^p, exp(e) for e in arange(p log diameter, p log blur,
  p log scaling), blur^p; diameter = norm of the joint bounding box's extent;
* softmin(eps, C, h)_i = -eps log sum_j exp(h_j - C_ij / eps);
* initialisation at the first epsilon without potentials, then per epsilon one symmetric
  update of (f_ba, g_ab) and of the debiasing (f_aa, g_bb), each the average of the old
  potential and its Sinkhorn update; a last extrapolation step with inputs held fixed;
* S = <a, f_ba - f_aa> + <b, g_ab - g_bb>, and its gradient w.r.t. x through the last step
  with the step's input potentials held fixed (what automatic differentiation of geomloss's
  code gives). Both extrapolated potentials f_ba and g_ab depend on x through the cost, so
  at convergence this gradient is twice the derivative of S (and the debiasing term's
  likewise).

No golden vector of the reference covers these (the tutorial's numbers come from random
inputs). The oracle is pinned in tests/test_scores_oracle.py by what the mathematics fixes:
S(x, x) = 0, the Wasserstein-1 distance of a translated cloud (|t|) and, with the iterations
converged, of small clouds solved exactly as an assignment problem; the gradient against
automatic differentiation of the last step and, converged, against central finite
differences of S (the factor 2 above).
"""
from __future__ import annotations

import numpy as np


def cost(x: np.ndarray, y: np.ndarray, p: int) -> np.ndarray:
    d2 = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    return np.sqrt(d2) if p == 1 else 0.5 * d2


def softmin(eps: float, C: np.ndarray, h: np.ndarray) -> np.ndarray:
    t = h[None, :] - C / eps
    mx = t.max(1, keepdims=True)
    return -eps * (mx[:, 0] + np.log(np.exp(t - mx).sum(1)))


def epsilon_schedule(p: int, diameter: float, blur: float, scaling: float):
    return ([diameter ** p] + [float(np.exp(e)) for e in np.arange(p * np.log(diameter), p * np.log(blur),
                                                                   p * np.log(scaling))] + [blur ** p])


def diameter(x: np.ndarray, y: np.ndarray) -> float:
    lo = np.minimum(x.min(0), y.min(0))
    hi = np.maximum(x.max(0), y.max(0))
    return float(np.linalg.norm(hi - lo))


def sinkhorn(x, y, p=1, blur=1e-4, scaling=0.5, diam=None, extra_iters=0):
    """Returns (S, gradient w.r.t. x, eps) for uniform point clouds x [n, 3], y [m, 3].
    extra_iters: further updates at the final epsilon before the last step (not part of
    geomloss's schedule; used to check the converged limit)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n, m = len(x), len(y)
    a_log = np.full(n, -np.log(n))
    b_log = np.full(m, -np.log(m))
    Cxy, Cyx, Cxx, Cyy = cost(x, y, p), cost(y, x, p), cost(x, x, p), cost(y, y, p)
    d = max(diameter(x, y) if diam is None else diam, blur)
    eps_list = epsilon_schedule(p, d, blur, scaling) + [blur ** p] * int(extra_iters)
    eps = eps_list[0]
    f_aa, g_bb = softmin(eps, Cxx, a_log), softmin(eps, Cyy, b_log)
    g_ab, f_ba = softmin(eps, Cyx, a_log), softmin(eps, Cxy, b_log)
    for eps in eps_list:
        ft_ba = softmin(eps, Cxy, b_log + g_ab / eps)
        gt_ab = softmin(eps, Cyx, a_log + f_ba / eps)
        ft_aa = softmin(eps, Cxx, a_log + f_aa / eps)
        gt_bb = softmin(eps, Cyy, b_log + g_bb / eps)
        f_ba, g_ab = 0.5 * (f_ba + ft_ba), 0.5 * (g_ab + gt_ab)
        f_aa, g_bb = 0.5 * (f_aa + ft_aa), 0.5 * (g_bb + gt_bb)
    f0, g0, faa0 = f_ba, g_ab, f_aa
    F = softmin(eps, Cxy, b_log + g_ab / eps)
    G = softmin(eps, Cyx, a_log + f_ba / eps)
    Faa = softmin(eps, Cxx, a_log + f_aa / eps)
    Gbb = softmin(eps, Cyy, b_log + g_bb / eps)
    S = float((F - Faa).mean() + (G - Gbb).mean())

    def dcost(u, v):  # dC(u_i, v_j)/du_i, [len(u), len(v), 3]
        diff = u[:, None, :] - v[None, :, :]
        if p == 2:
            return diff
        r = np.sqrt((diff ** 2).sum(-1, keepdims=True))
        return np.where(r > 0, diff / np.where(r > 0, r, 1.0), 0.0)

    def plan(alpha, C, beta):
        return np.exp(alpha[:, None] + beta[None, :] - C / eps)

    la, lb = a_log[0], b_log[0]
    Dxy, Dxx = dcost(x, y), dcost(x, x)
    t1 = (plan(F / eps + la, Cxy, g0 / eps + lb)[..., None] * Dxy).sum(1)
    t2 = (plan(f0 / eps + la, Cxy, G / eps + lb)[..., None] * Dxy).sum(1)
    t3 = (plan(Faa / eps + la, Cxx, faa0 / eps + la)[..., None] * Dxx).sum(1)
    t4 = (plan(faa0 / eps + la, Cxx, Faa / eps + la)[..., None] * Dxx).sum(1)
    return S, t1 + t2 - t3 - t4, eps


def chamfer(x, y):
    """(mean_i min_j |x_i - y_j|^2 + mean_j min_i |x_i - y_j|^2, its gradient w.r.t. x)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    d2 = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    ix, iy = d2.argmin(1), d2.argmin(0)
    val = float(d2.min(1).mean() + d2.min(0).mean())
    g = 2.0 / len(x) * (x - y[ix])
    np.add.at(g, iy, 2.0 / len(y) * (x[iy] - y))
    return val, g
