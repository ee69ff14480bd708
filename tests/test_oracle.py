"""Pins for the CPU oracle (oracle/mpm_oracle.cpp), independent of the HIP code.

The reference ships no tests, golden vectors or fixtures for the simulator, so each
restated function is pinned by what the mathematics fixes:
* svd3 (svd.h)             : numpy.linalg.svd singular values, reconstruction, orthogonality, ordering
* von_mises (cu:42-67)      : identity below yield; projected deviatoric log-strain norm = yield/(2 mu)
* quaternion ops (quat.h)   : scipy Rotation; inverse transform round trip
* SDFs (shape.h)            : closed-form distances; gradient = finite-difference gradient of the SDF
* p2g / g2p (cu:312-394,1059-1109): mass and momentum conservation; exact APIC transfer of an affine field
* particle2mass (cu:239-310): total mass; its adjoint vs finite differences
* every adjoint (cu:110-186,396-627,779-1057,1527-1614, shape.h:104-151, quat.h:25-80):
                              FP64 central finite differences of the FP64 forward (exact mode)
* reference adjoint mode     : equal to exact mode where the reference's dropped terms vanish
"""
import numpy as np
import pytest

from oracle import oracle as O

RNG = np.random.default_rng


# ------------------------------------------------------------------ svd3
def test_svd3_matches_numpy_and_reconstructs():
    rng = RNG(0)
    A = rng.normal(size=(400, 3, 3))
    A[:100] = np.eye(3) + 0.05 * rng.normal(size=(100, 3, 3))  # near-identity deformation gradients
    U, s, V = O.svd3(A)
    rec = U @ (s[:, :, None] * np.swapaxes(V, 1, 2))
    assert np.abs(rec - A).max() < 2e-5 * np.abs(A).max()
    eye = np.eye(3)
    assert np.abs(U @ np.swapaxes(U, 1, 2) - eye).max() < 1e-12
    # V comes from an unnormalised Jacobi quaternion: the float-rounded cos/sin(pi/8)
    # constants (svd.h:18-19) make it drift from unit norm by ~1e-7 (svd.h:81-109)
    assert np.abs(V @ np.swapaxes(V, 1, 2) - eye).max() < 2e-6
    assert np.all(s[:, 0] >= s[:, 1]) and np.all(s[:, 1] >= np.abs(s[:, 2]))
    sv = np.linalg.svd(A, compute_uv=False)
    assert np.abs(np.sort(np.abs(s), 1) - np.sort(sv, 1)).max() < 2e-5 * sv.max()
    # sign convention: det(U) = det(V) = +1 (rotations), negative determinant shows in sigma_3
    assert np.allclose(np.linalg.det(U), 1) and np.allclose(np.linalg.det(V), 1, atol=4e-6)
    assert np.all(np.sign(s[:, 2]) == np.sign(np.linalg.det(A)))


def test_svd3_fp32_demotion():
    A = RNG(1).normal(size=(50, 3, 3)).astype(np.float32)
    U32, s32, V32 = O.svd3(A, dtype=np.float32)
    U64, s64, V64 = O.svd3(A.astype(np.float64))
    assert np.abs(s32 - s64).max() < 1e-6 * np.abs(s64).max()
    assert np.abs(U32 - U64).max() < 1e-6


# ------------------------------------------------------------------ von Mises
def test_von_mises_return_map():
    rng = RNG(2)
    n = 200
    mu = 5e3 / 2.4
    scale = np.geomspace(1e-3, 0.3, n)[:, None, None]  # elastic and plastic particles
    F = np.eye(3) + scale * rng.normal(size=(n, 3, 3))
    U, s, V = O.svd3(F)
    mly = np.tile([mu, 1000.0, 50.0], (n, 1))
    Fn, J = O.von_mises(F, U, s, V, mly)
    sc = np.maximum(s, 0.05)
    eps = np.log(sc)
    eh = eps - eps.mean(1, keepdims=True)
    ehn = np.sqrt((eh * eh).sum(1) + 1e-8)
    plastic = ehn - 50.0 / (2 * mu) > 0
    assert plastic.any() and (~plastic).any()
    assert np.allclose(Fn[~plastic], F[~plastic])
    assert np.allclose(J[~plastic], np.prod(s[~plastic], 1))
    # plastic: new singular values e^(eps_new) with deviatoric norm yield/(2 mu)
    sn = np.linalg.svd(Fn[plastic], compute_uv=False)
    le = np.log(sn)
    dn = np.linalg.norm(le - le.mean(1, keepdims=True), axis=1)
    expect = (ehn - (ehn - 50.0 / (2 * mu))) * np.linalg.norm(eh, axis=1) / ehn
    # (U, V from svd3 are orthogonal to ~1e-6, which bounds the agreement)
    assert np.allclose(dn, expect[plastic], rtol=1e-4)
    assert np.allclose(dn, 50.0 / (2 * mu), rtol=1e-4)
    assert np.allclose(J[plastic], np.prod(sn, 1), rtol=1e-5)


# ------------------------------------------------------------------ quaternions
def test_quaternion_rotation_matches_scipy():
    from scipy.spatial.transform import Rotation
    rng = RNG(3)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        p = rng.normal(size=3)
        pts = rng.normal(size=(10, 3))
        got = O.xform(p, q, pts)
        R = Rotation.from_quat([q[1], q[2], q[3], q[0]])
        assert np.allclose(got, R.apply(pts) + p, atol=1e-12)
        back = O.xform(p, q, got, inverse=True)
        assert np.allclose(back, pts, atol=1e-12)


# ------------------------------------------------------------------ SDFs
def test_box_and_capsule_sdf_closed_forms():
    rng = RNG(4)
    pts = rng.uniform(-0.3, 0.3, size=(500, 3))
    half = np.array([0.1, 0.05, 0.08])
    sdf, grad = O.shape_sdf([0, 0.5, 666, 0.0], [*half, 0], pts)
    q = np.abs(pts) - half
    expect = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(1), 0)
    assert np.allclose(sdf, expect, atol=1e-12)
    r, h = 0.03, 0.1
    sdf, grad = O.shape_sdf([1, 0.9, 666, 0.01], [r, h, 0, 0], pts)
    yc = np.clip(pts[:, 1], -h / 2, h / 2)
    seg = np.stack([np.zeros(len(pts)), yc, np.zeros(len(pts))], 1)
    expect = np.linalg.norm(pts - seg, axis=1) - r - 0.01
    assert np.allclose(sdf, expect, atol=1e-12)


@pytest.mark.parametrize("prim,args", [([0, 0.5, 666, 0.0], [0.1, 0.05, 0.08, 0]),
                                       ([1, 0.9, 666, 0.0], [0.03, 0.1, 0, 0])])
def test_sdf_gradient_is_derivative(prim, args):
    rng = RNG(5)
    pts = rng.uniform(-0.25, 0.25, size=(300, 3))
    sdf, grad = O.shape_sdf(prim, args, pts)
    e = 1e-7
    fd = np.zeros_like(pts)
    for c in range(3):
        d = np.zeros(3)
        d[c] = e
        fd[:, c] = (O.shape_sdf(prim, args, pts + d)[0] - O.shape_sdf(prim, args, pts - d)[0]) / (2 * e)
    # the box gradient inside is the normal of the nearest face: that is the derivative too
    assert np.allclose(grad, fd, atol=1e-5)


# ------------------------------------------------------------------ transfers
def _scene(n, grid=16, nb=0, yield_stress=50.0, E=5e3, gf=0.0, dt=2e-3, tfsr=None, args=None, vol=None):
    dx = 1.0 / grid
    v = (dx / 2) ** 2 if vol is None else vol
    m, vol_, mly = O.material(n, E=E, yield_stress=yield_stress, vol=v, mass=(dx / 2) ** 2)
    if tfsr is None:
        tfsr = np.zeros((nb, 4))
        args = np.zeros((nb, 4))
    return O.Scene((grid, grid, grid), dx, dt, (0, -294.0, 0), m, vol_, mly, np.asarray(tfsr, float),
                   np.asarray(args, float), ground_friction=gf)


def test_p2g_conserves_mass_and_momentum():
    rng = RNG(6)
    n = 300
    sc = _scene(n)
    sim = O.OracleSim(sc, np.float64)
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    v = rng.normal(size=(n, 3))
    F = np.eye(3) + 0.1 * rng.normal(size=(n, 3, 3))
    C = rng.normal(size=(n, 3, 3))
    tmp = sim._temps(n)
    st = O.State(x, v, F, C, np.zeros((0, 3)), np.zeros((0, 4)))
    sim._fwd_to_grid(st, st.body_pos, st.body_rot, tmp)
    assert np.isclose(tmp["grid_m"].sum(), sim.mass.sum(), rtol=1e-13)
    assert np.allclose(tmp["grid_mv"].sum(0), (sim.mass[:, None] * v).sum(0), rtol=1e-10, atol=1e-15)


def test_apic_transfers_affine_field_exactly():
    """P2G without stress followed by G2P (no grid update) reproduces v = b + A x and C = A."""
    rng = RNG(7)
    n = 4000
    sc = _scene(n, vol=0.0)  # zero volume -> zero stress contribution
    sim = O.OracleSim(sc, np.float64)
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    v = b + x @ A.T
    C = np.tile(A, (n, 1, 1))
    st = O.State(x, v, np.tile(np.eye(3), (n, 1, 1)), C, np.zeros((0, 3)), np.zeros((0, 4)))
    tmp = sim._temps(n)
    sim._c("compute_svd", n, st.F, st.C, sim._F(0.0), tmp["Ftmp"], tmp["U"], tmp["V"], tmp["sig"])
    Fn = np.zeros_like(st.F)
    sim._c("p2g", n, x, v, sim.mass, sim.vol, tmp["Ftmp"], tmp["U"], tmp["sig"], tmp["V"], C, sim.mly, sim.grid,
           sim._F(sc.dx), sim._F(1 / sc.dx), sim._F(0.0), Fn, tmp["grid_mv"], tmp["grid_m"])
    act = tmp["grid_m"] > 1e-12
    gv = np.zeros((sim.G, 3))
    gv[act] = tmp["grid_mv"][act] / tmp["grid_m"][act, None]
    idx = np.arange(sim.G)
    node = np.stack([idx // 256, (idx // 16) % 16, idx % 16], 1) * sc.dx
    assert np.allclose(gv[act], b + node[act] @ A.T, atol=1e-11)
    vn, Cn, xn = np.zeros_like(v), np.zeros_like(C), np.zeros_like(x)
    sim._c("g2p", n, x, gv, sim.grid, sim._F(sc.dx), sim._F(1 / sc.dx), sim._F(0.0), sim._F(3.0), vn, Cn, xn)
    assert np.allclose(vn, v, atol=1e-11)
    assert np.allclose(Cn, C, atol=1e-9)
    assert np.allclose(xn, x, atol=1e-15)


def test_particle2mass_total_and_adjoint():
    rng = RNG(8)
    n = 200
    sc = _scene(n)
    sim = O.OracleSim(sc, np.float64)
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    gm = sim.grid_mass(x)
    assert np.isclose(gm.sum(), sim.mass.sum(), rtol=1e-13)
    W = rng.normal(size=gm.shape)
    g = np.zeros((n, 3))
    sim.grid_mass_grad(x, W, g)
    e = 1e-7
    for _ in range(10):
        p, c = rng.integers(n), rng.integers(3)
        xp, xm = x.copy(), x.copy()
        xp[p, c] += e
        xm[p, c] -= e
        fd = ((sim.grid_mass(xp) - sim.grid_mass(xm)) * W).sum() / (2 * e)
        assert np.isclose(g[p, c], fd, rtol=1e-6, atol=1e-9)


# ------------------------------------------------------------------ adjoints vs FP64 finite differences
@pytest.fixture
def exact_svd():
    """Finite differences differentiate the exact SVD; the reference's approximate one
    jumps by ~1e-6 where its Givens branch switches (svd.h:120-123)."""
    with O.exact_svd():
        yield


def _fd_setup(seed, yield_stress, gf, with_bodies=True, dt=2e-3, c_scale=2.0):
    rng = RNG(seed)
    n = 64
    tfsr = [[1, 0.9, 666, 0], [0, 0.5, 666, 0]] if with_bodies else None
    args = [[0.05, 0.1, 0, 0], [0.05, 0.03, 0.04, 0]] if with_bodies else None
    sc = _scene(n, nb=2 if with_bodies else 0, yield_stress=yield_stress, gf=gf, tfsr=tfsr, args=args, dt=dt)
    lo = [0.35, 0.10 if gf else 0.22, 0.4]
    x = rng.uniform(lo, [0.65, 0.4, 0.6], size=(n, 3))
    F0 = np.eye(3) + 0.05 * rng.normal(size=(n, 3, 3))
    C0 = c_scale * rng.normal(size=(n, 3, 3))
    v0 = 0.3 * rng.normal(size=(n, 3))
    if gf:
        v0[:, 1] -= 1.5
    if with_bodies:
        bp = np.array([[0.503, 0.431, 0.497], [0.421, 0.302, 0.493]])
        q = np.array([0.9, 0.1, 0.2, 0.1])
        br = np.array([[1.0, 0, 0, 0], q / np.linalg.norm(q)])
        nbp = bp + [[0.001, -0.002, 0.0005], [0.0, 0.001, 0]]
    else:
        bp, br, nbp = np.zeros((0, 3)), np.zeros((0, 4)), np.zeros((0, 3))
    st = O.State(x, v0, F0, C0, bp, br)
    return rng, sc, st, nbp, br.copy()


@pytest.mark.parametrize("yield_stress,gf", [(50.0, 0.0), (1e9, 0.0), (50.0, 0.4)])
def test_substep_adjoint_matches_central_differences(yield_stress, gf, exact_svd):
    rng, sc, st, npos, nrot = _fd_setup(11, yield_stress, gf)
    sim = O.OracleSim(sc, np.float64, adjoint="exact")
    n, nb = len(st.x), 2
    W = [rng.normal(size=a.shape) for a in (st.x, st.v, st.F, st.C)]

    def loss(s, p=npos, r=nrot):
        s1 = sim.substep(s, p, r)
        return sum((w * a).sum() for w, a in zip(W, (s1.x, s1.v, s1.F, s1.C)))

    s1 = sim.substep(st, npos, nrot)
    gn = O.Grads(*(w.copy() for w in W), np.zeros((nb, 3)), np.zeros((nb, 4)))
    gc = O.Grads.zeros(n, nb, np.float64)
    sim.substep_grad(st, s1, gn, gc)
    e = 1e-6
    checks = []
    for name in ["x", "v", "F", "C"]:
        arr = getattr(st, name)
        for _ in range(6):
            idx = tuple(int(rng.integers(0, s)) for s in arr.shape)
            sp, sm = st.copy(), st.copy()
            getattr(sp, name)[idx] += e
            getattr(sm, name)[idx] -= e
            checks.append(((loss(sp) - loss(sm)) / (2 * e), getattr(gc, name)[idx], name))
    for name in ["body_pos", "body_rot"]:
        for b in range(nb):
            for c in range(3 if name == "body_pos" else 4):
                sp, sm = st.copy(), st.copy()
                getattr(sp, name)[b, c] += e
                getattr(sm, name)[b, c] -= e
                checks.append(((loss(sp) - loss(sm)) / (2 * e), getattr(gc, name)[b, c], name))
                p, r = npos.copy(), nrot.copy()
                tgt = p if name == "body_pos" else r
                tgt[b, c] += e
                lp = loss(st, p, r)
                tgt[b, c] -= 2 * e
                lm = loss(st, p, r)
                checks.append(((lp - lm) / (2 * e), getattr(gn, name)[b, c], "next_" + name))
    for fd, ad, name in checks:
        assert abs(fd - ad) <= 2e-4 * max(1.0, abs(fd), abs(ad)), (name, fd, ad)


def test_multistep_adjoint_matches_central_differences(exact_svd):
    # a stable multi-substep setting (dt below the elastic CFL limit of the 16^3 grid)
    rng, sc, st, _, _ = _fd_setup(12, 50.0, 0.0, dt=1e-4, c_scale=0.5)
    sim = O.OracleSim(sc, np.float64, adjoint="exact")
    T = 6
    P = np.stack([st.body_pos + t * np.array([[0.0002, -0.0004, 0.0], [0.0, 0.0002, 0.0002]]) for t in range(T + 1)])
    R = np.stack([st.body_rot] * (T + 1))
    wx = rng.normal(size=st.x.shape)

    def loss(s0, P_=P):
        return (sim.rollout(s0, P_, R)[-1].x * wx).sum()

    states = sim.rollout(st, P, R)
    n = len(st.x)
    g_last = O.Grads(wx.copy(), np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3)), np.zeros((2, 3)),
                     np.zeros((2, 4)))
    g0, pg, _ = sim.rollout_grad(states, g_last)
    e = 1e-6
    for _ in range(6):
        p, c = int(rng.integers(n)), int(rng.integers(3))
        sp, sm = st.copy(), st.copy()
        sp.x[p, c] += e
        sm.x[p, c] -= e
        fd = (loss(sp) - loss(sm)) / (2 * e)
        assert abs(fd - g0.x[p, c]) <= 2e-4 * max(1, abs(fd)), (fd, g0.x[p, c])
    for t in (1, 3, 6):
        for b, c in ((0, 1), (1, 0)):
            Pp, Pm = P.copy(), P.copy()
            Pp[t, b, c] += e
            Pm[t, b, c] -= e
            fd = (loss(st, Pp) - loss(st, Pm)) / (2 * e)
            assert abs(fd - pg[t, b, c]) <= 2e-4 * max(1, abs(fd)), (t, b, c, fd, pg[t, b, c])


def test_compute_dist_adjoint():
    rng, sc, st, _, _ = _fd_setup(13, 50.0, 0.0)
    sim = O.OracleSim(sc, np.float64)
    W = rng.normal(size=(len(st.x), 2))
    g = O.Grads.zeros(len(st.x), 2, np.float64)
    sim.compute_dist_grad(st, W, g)
    e = 1e-7
    for name, shape in (("x", st.x.shape), ("body_pos", (2, 3)), ("body_rot", (2, 4))):
        for _ in range(5):
            idx = tuple(int(rng.integers(0, s)) for s in shape)
            sp, sm = st.copy(), st.copy()
            getattr(sp, name)[idx] += e
            getattr(sm, name)[idx] -= e
            fd = ((sim.compute_dist(sp) - sim.compute_dist(sm)) * W).sum() / (2 * e)
            assert abs(fd - getattr(g, name)[idx]) <= 1e-5 * max(1, abs(fd)), (name, fd)


def test_reference_adjoint_equals_exact_without_capsule_contact():
    """The reference's two adjoint departures (capsule normal term, sticky ground) only
    matter for capsule contact and ground_friction >= 99: box-only scenes agree bitwise."""
    rng = RNG(14)
    n = 64
    sc = _scene(n, nb=1, tfsr=[[0, 0.5, 666, 0]], args=[[0.05, 0.03, 0.04, 0]], gf=0.3)
    x = rng.uniform([0.35, 0.1, 0.4], [0.65, 0.4, 0.6], size=(n, 3))
    st = O.State(x, rng.normal(size=(n, 3)) - [0, 1, 0], np.tile(np.eye(3), (n, 1, 1)), rng.normal(size=(n, 3, 3)),
                 np.array([[0.45, 0.3, 0.5]]), np.array([[0.96, 0.1, 0.2, 0.17]]) / 1.0)
    st.body_rot /= np.linalg.norm(st.body_rot)
    nb_pos = st.body_pos + 0.001
    outs = []
    for mode in ("reference", "exact"):
        sim = O.OracleSim(sc, np.float64, adjoint=mode)
        s1 = sim.substep(st, nb_pos, st.body_rot)
        gn = O.Grads(np.ones((n, 3)), np.ones((n, 3)), np.ones((n, 3, 3)), np.ones((n, 3, 3)), np.zeros((1, 3)),
                     np.zeros((1, 4)))
        gc = O.Grads.zeros(n, 1, np.float64)
        sim.substep_grad(st, s1, gn, gc)
        outs.append((gc, gn))
    for a, b in zip(outs[0][0].__dict__.values(), outs[1][0].__dict__.values()):
        assert np.array_equal(a, b)
    # with a capsule in contact, the reference mode drops the normal-direction term
    sc2 = _scene(n, nb=1, tfsr=[[1, 0.9, 666, 0]], args=[[0.05, 0.1, 0, 0]])
    st2 = O.State(x, np.zeros((n, 3)), np.tile(np.eye(3), (n, 1, 1)), np.zeros((n, 3, 3)),
                  np.array([[0.503, 0.3, 0.497]]), np.array([[1.0, 0, 0, 0]]))
    res = []
    for mode in ("reference", "exact"):
        sim = O.OracleSim(sc2, np.float64, adjoint=mode)
        s1 = sim.substep(st2, st2.body_pos + 0.001, st2.body_rot)
        gn = O.Grads(rng.normal(size=(n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3)),
                     np.zeros((1, 3)), np.zeros((1, 4)))
        gc = O.Grads.zeros(n, 1, np.float64)
        sim.substep_grad(st2, s1, gn, gc)
        res.append(gc.body_pos.copy())
    assert not np.allclose(res[0], res[1])


def test_fp32_oracle_within_rounding_of_fp64():
    rng, sc, st, npos, nrot = _fd_setup(15, 50.0, 0.0)
    s64 = O.OracleSim(sc, np.float64).substep(st, npos, nrot)
    s32 = O.OracleSim(sc, np.float32).substep(st, npos, nrot)
    for a, b in zip((s32.x, s32.v, s32.F, s32.C), (s64.x, s64.v, s64.F, s64.C)):
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)


# ------------------------------------------------------------------ grid update forward (cu:647-777)
# Closed-form cases: every expected value below comes from the geometry the test builds
# (node position, body pose, body-frame point, face normal), not from the oracle's SDFs.
GD = (16, 16, 16)
GDX = 1.0 / 16
GDT = 1e-3
GRAV = np.array([0.3, -2.0, 0.1])


def _gidx(i, j, k):
    return (i * GD[1] + j) * GD[2] + k


def _rotq(axis, angle):
    """(w, x, y, z) quaternion and matrix of a rotation, from scipy."""
    from scipy.spatial.transform import Rotation
    r = Rotation.from_rotvec(np.asarray(axis, float) / np.linalg.norm(axis) * angle)
    x, y, z, w = r.as_quat()
    return np.array([w, x, y, z]), r.as_matrix()


def _one_node(node, v_in, bodies, gf=0.0, m=2.0):
    """Grid with one active node whose velocity after gravity is v_in. bodies: list of
    (tfsr, args, pos, quat, next_pos, next_quat)."""
    G = int(np.prod(GD))
    gm = np.zeros(G)
    gmv = np.zeros((G, 3))
    t = _gidx(*node)
    gm[t] = m
    gmv[t] = (np.asarray(v_in) - GDT * GRAV) * m
    cols = list(zip(*bodies)) if bodies else [[]] * 6
    out, tape = O.grid_op(gm, gmv, GD, cols[2], cols[3], cols[4], cols[5], cols[0], cols[1], GDX, GDT, GRAV,
                          ground_friction=gf, out_fill=7.0)
    return out[t], tape[t], out, t


def _placed(node, R, q, x_b, next_R=None, next_q=None, shift=(0, 0, 0)):
    """Body pose putting grid node `node` at body-frame point x_b; optionally a next pose."""
    xw = np.asarray(node, float) * GDX
    p = xw - R @ x_b
    if next_R is None:
        return p, q, p, q
    return p, q, p + np.asarray(shift), next_q


def test_grid_op_free_node_and_activity_threshold():
    G = int(np.prod(GD))
    gm = np.zeros(G)
    gmv = np.zeros((G, 3))
    a, b, c = _gidx(8, 8, 8), _gidx(8, 9, 8), _gidx(8, 10, 8)
    gm[a], gmv[a] = 2.0, [0.4, -0.2, 0.6]
    gm[b], gmv[b] = 1e-13, [1.0, 1.0, 1.0]      # below the 1e-12 activity threshold: skipped
    gm[c], gmv[c] = 4e-12, [4e-12, 0.0, 0.0]    # just above it: updated
    far_box = ([0, 0.0, 666, 0], [0.02, 0.02, 0.02, 0], [0.1, 0.1, 0.1], [1, 0, 0, 0])
    out, tape = O.grid_op(gm, gmv, GD, [far_box[2]], [far_box[3]], [far_box[2]], [far_box[3]], [far_box[0]],
                          [far_box[1]], GDX, GDT, GRAV, out_fill=7.0)
    expect = np.array([0.2, -0.1, 0.3]) + GDT * GRAV
    assert np.allclose(out[a], expect, rtol=0, atol=1e-15)
    assert np.all(out[b] == 7.0)
    assert np.allclose(out[c], np.array([1.0, 0, 0]) + GDT * GRAV, atol=1e-12)
    # tape: velocity entering body 0, then after the last body (no contact here: equal)
    assert np.allclose(tape[a, 0], expect, atol=1e-15) and np.allclose(tape[a, 1], expect, atol=1e-15)
    untouched = np.ones(G, bool)
    untouched[[a, c]] = False
    assert np.all(out[untouched] == 7.0)


@pytest.mark.parametrize("approach", [True, False])
def test_grid_op_static_rotated_box_frictionless(approach):
    """Node inside a static rotated box (influence 1), friction 0: the inward normal
    component is removed, a separating velocity passes through unchanged."""
    half = np.array([0.05, 0.04, 0.06])
    q, R = _rotq([1, 2, 3], 0.7)
    x_b = np.array([0.3, 0.8, -0.2]) * half  # nearest face: +y of the body
    n = R @ np.array([0.0, 1.0, 0.0])
    t = np.cross(n, [1.0, 0, 0])
    t *= 0.3 / np.linalg.norm(t)
    v = (-0.7 if approach else 0.5) * n + t
    node = (8, 8, 8)
    p, q0, pn, qn = _placed(node, R, q, x_b)
    far = ([1, 0.0, 666, 0], [0.01, 0.02, 0, 0], [0.9, 0.9, 0.9], [1, 0, 0, 0], [0.9, 0.9, 0.9], [1, 0, 0, 0])
    box = ([0, 0.0, 666, 0], [*half, 0], p, q0, pn, qn)
    v_out, tape, _, _ = _one_node(node, v, [box, far])
    expect = t if approach else v
    assert np.allclose(v_out, expect, atol=1e-10)
    assert np.allclose(tape[0], v, atol=1e-12)      # entering body 0
    assert np.allclose(tape[1], expect, atol=1e-10)  # entering body 1 = after body 0
    assert np.allclose(tape[2], expect, atol=1e-10)  # after the last body


def _capsule_case(mu, v_rel_n, v_rel_t, moving=True, round_=0.0):
    r, h = 0.03, 0.1
    q, R = _rotq([0.2, -1, 0.4], 0.9)
    x_b = np.array([0.01, 0.02, 0.005])        # inside: segment point (0, 0.02, 0), radial (0.01, 0, 0.005)
    nb_ = np.array([0.01, 0.0, 0.005]) / np.hypot(0.01, 0.005)
    node = (7, 9, 8)
    xw = np.asarray(node, float) * GDX
    if moving:
        q1, R1 = _rotq([1, 0.5, -0.3], 0.02)
        q1 = np.array([q1[0] * q[0] - q1[1:] @ q[1:], *(q1[0] * q[1:] + q[0] * q1[1:] + np.cross(q1[1:], q[1:]))])
        R1 = R1 @ R
        shift = np.array([1e-4, -2e-4, 5e-5])
    else:
        q1, R1, shift = q, R, np.zeros(3)
    p = xw - R @ x_b
    pn = p + shift
    bv = (pn + R1 @ x_b - xw) / GDT             # body velocity at the node
    n = R @ nb_
    t = np.cross(n, [0, 0, 1.0])
    t /= np.linalg.norm(t)
    v = bv + v_rel_n * n + v_rel_t * t
    cap = ([1, mu, 666, round_], [r, h, 0, 0], p, q, pn, q1)
    v_out, _, _, _ = _one_node(node, v, [cap])
    return v_out, bv, n, t


def test_grid_op_moving_capsule_sticks_at_high_friction():
    """Approaching a moving capsule with friction 10: the node takes the body's velocity
    (p' + R' x_b - x) / dt at that point."""
    v_out, bv, n, t = _capsule_case(10.0, -1.0, 0.2)
    assert np.linalg.norm(bv) > 0.05
    assert np.allclose(v_out, bv, atol=1e-9)


def test_grid_op_coulomb_friction_on_moving_capsule():
    """Friction 0.5, normal speed -1, tangential 0.8: the tangential relative speed
    shrinks to 0.8 - 0.5 = 0.3 and the normal component is removed."""
    v_out, bv, n, t = _capsule_case(0.5, -1.0, 0.8)
    assert np.allclose(v_out, bv + 0.3 * t, atol=1e-9)
    # separating from the moving capsule: relative velocity kept
    v_out, bv, n, t = _capsule_case(0.5, 0.4, 0.8)
    assert np.allclose(v_out, bv + 0.4 * n + 0.8 * t, atol=1e-9)


@pytest.mark.parametrize("d,round_", [(0.002, 0.0), (0.003, 0.001), (0.004, 0.0)])
def test_grid_op_influence_shell_beyond_capsule_cap(d, round_):
    """Node on the capsule axis beyond the segment end y = h/2, at distance d from the
    surface (the segment spans y in [-h/2, h/2]; h is its full length). Inside the shell
    (exp(-(d - round) * 666) > 0.1) the contact is blended with the influence; outside it
    the node is untouched."""
    r, h, mu = 0.03, 0.1, 0.25
    q, R = _rotq([0.5, 0.1, 1.0], 1.3)
    x_b = np.array([0.0, h / 2 + r + d, 0.0])
    node = (8, 7, 9)
    p, q0, pn, qn = _placed(node, R, q, x_b)
    n = R @ np.array([0.0, 1.0, 0.0])
    t = np.cross(n, [1.0, 0, 0])
    t /= np.linalg.norm(t)
    v = -0.6 * n + 0.4 * t
    cap = ([1, mu, 666, round_], [r, h, 0, 0], p, q0, pn, qn)
    v_out, _, _, _ = _one_node(node, v, [cap])
    infl = np.exp(-(d - round_) * 666)
    if infl > 0.1:
        vt = (0.4 - 0.6 * mu) * t
        expect = v * (1 - infl) + vt * infl
        assert 0.1 < infl < 1
    else:
        expect = v
    assert np.allclose(v_out, expect, atol=1e-10)


def test_grid_op_domain_walls_are_asymmetric():
    """Low side: index < 3 stops outward motion; high side: index > dim - 3, i.e. only the
    last two planes (integrator.cu:735-774)."""
    d = GD[0]
    cases = []
    for i, vx, stop in [(1, -0.5, True), (2, -0.5, True), (3, -0.5, False), (1, 0.5, False),
                        (d - 3, 0.5, False), (d - 2, 0.5, True), (d - 1, 0.5, True), (d - 2, -0.5, False)]:
        cases.append(((i, 8, 8), np.array([vx, 0.25, -0.125]), (0, stop)))
        cases.append(((8, 8, i), np.array([0.125, 0.25, vx]), (2, stop)))
        if i > 3:  # the low y side is the ground (tested separately)
            cases.append(((8, i, 8), np.array([0.125, vx, 0.25]), (1, stop)))
    for node, v, (ax, stop) in cases:
        v_out, _, _, _ = _one_node(node, v, [])
        expect = v.copy()
        if stop:
            expect[ax] = 0.0
        assert np.allclose(v_out, expect, atol=1e-12), (node, v, v_out)


@pytest.mark.parametrize("gf", [0.0, 0.4, 100.0])
def test_grid_op_ground_friction(gf):
    """Ground (y index < 3, moving down): friction 0 only zeroes v_y; 0 < friction < 99
    scales the tangential velocity by max(1 + gf v_y / sqrt(|v_t|^2 + 1e-8), 0); >= 99
    sticks (v = 0). Moving up: untouched."""
    for j in (0, 1, 2):
        for vy, vx in ((-1.5, 1.0), (-3.0, 1.0), (0.5, 1.0)):
            v = np.array([vx, vy, 0.0])
            v_out, _, _, _ = _one_node((8, j, 8), v, [], gf=gf)
            if vy > 0:
                expect = v
            elif gf == 0.0:
                expect = np.array([vx, 0.0, 0.0])
            elif gf >= 99:
                expect = np.zeros(3)
            else:
                expect = np.array([vx * max(1 + gf * vy / np.sqrt(vx * vx + 1e-8), 0.0), 0.0, 0.0])
            assert np.allclose(v_out, expect, atol=1e-12), (gf, j, vy, v_out)
    v = np.array([1.0, -1.5, 0.0])
    v_out, _, _, _ = _one_node((8, 3, 8), v, [], gf=gf)
    assert np.allclose(v_out, v, atol=1e-12)
    if 0 < gf < 99:
        v_out, _, _, _ = _one_node((8, 1, 8), np.array([0.6, -1.5, 0.8]), [], gf=gf)
        assert np.allclose(v_out, np.array([0.6, 0, 0.8]) * 0.4, atol=1e-8)


def test_full_size_hand_fixture_adjoint_matches_its_finite_differences():
    """The FP64 oracle adjoint stored for the full-size hand-contact configuration (50,000
    particles, 200 substeps; tests/golden/make_fixtures.py) against FP64 central differences
    of the oracle's forward pass stored beside it: within 1e-3 at eps 2.5e-5. At eps 1e-4
    one coordinate's difference crosses the contact-influence threshold (influence > 0.1,
    a jump of the reference's contact model) and the relative error of the 16 is 0.33."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "hand_c2_fp64.npz"), allow_pickle=False)
    g = z["grad"][z["coords"][:, 0], z["coords"][:, 1]]
    fd = z["fd_eps_2p5e5"]
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) <= 1e-3
