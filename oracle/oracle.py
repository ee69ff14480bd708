"""ORACLE — test infrastructure only.

Python driver for the C++ CPU restatement in ``mpm_oracle.cpp``. It runs the
reference's per-substep kernel sequences with numpy buffers:

* forward ``substep``      = ``MPMSimulator.substep``      (mpm/simulator.py:561-571):
  temp.clear -> compute_svd -> p2g -> grid_op_v2 -> g2p
* backward ``substep_grad`` = ``MPMSimulator.substep_grad`` (mpm/simulator.py:573-585):
  temp.clear + clear_grad -> recompute compute_svd, p2g, grid_op_v2 ->
  g2p_grad -> grid_op_v2_grad -> p2g_grad -> compute_svd_grad

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import this module; the product (``dexdeform_amd``) never does.

Each restated function is pinned by ``tests/test_oracle.py``: closed forms and
invariants (SVD reconstruction, von Mises projection, SDF distances, mass and momentum
conservation, exact APIC transfer), numpy and scipy routines (SVD, rotations) and FP64
central finite differences of every adjoint. Parity unpinned: the exact FP32 bits of the
reference's NVIDIA build (FMA contraction, atomic order, expf/logf ulps); the
restatement matches it up to rounding only.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

ADJ_REFERENCE = 0
ADJ_EXACT = 1


def build(force: bool = False) -> None:
    """Compile the oracle libraries with g++ (``oracle/Makefile``)."""
    args = ["make", "-s", "-C", _HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)


def _lib(omp: bool = False):
    key = "omp" if omp else "serial"
    if key not in _LIBS:
        name = "libmpm_oracle_omp.so" if omp else "libmpm_oracle.so"
        path = os.path.join(_HERE, "lib", name)
        if not os.path.exists(path):
            build()
        _LIBS[key] = ctypes.CDLL(path)
    return _LIBS[key]


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle buffers must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _call(name, dtype, *args, omp=False):
    sfx = "f32" if dtype == np.float32 else "f64"
    fn = getattr(_lib(omp), f"orc_{name}_{sfx}")
    scalar = ctypes.c_float if dtype == np.float32 else ctypes.c_double
    cargs = []
    for a in args:
        if isinstance(a, np.ndarray):
            cargs.append(_p(a))
        elif a is None:
            cargs.append(None)
        elif isinstance(a, bool):
            cargs.append(ctypes.c_int(int(a)))
        elif isinstance(a, (int, np.integer)):
            cargs.append(ctypes.c_int(int(a)))
        elif isinstance(a, _F):
            cargs.append(scalar(a.v))
        else:
            raise TypeError(f"unsupported oracle argument {type(a)}")
    fn.restype = None
    fn(*cargs)


class _F:
    """Marks a scalar passed as the oracle's value type."""

    __slots__ = ("v",)

    def __init__(self, v):
        self.v = float(v)


class exact_svd:
    """Context manager: use the exact (smooth) Jacobi SVD in the oracle instead of the
    reference's approximate one. For finite-difference checks only."""

    def __enter__(self):
        for lib in (_lib(False), _lib(True)):
            lib.orc_set_svd_mode(1)
        return self

    def __exit__(self, *exc):
        for lib in (_lib(False), _lib(True)):
            lib.orc_set_svd_mode(0)
        return False


def num_threads() -> int:
    fn = _lib(True).orc_num_threads
    fn.restype = ctypes.c_int
    return int(fn())


# ------------------------------------------------------------------------ single functions
def svd3(A, dtype=np.float64):
    A = np.ascontiguousarray(A, dtype=dtype).reshape(-1, 3, 3)
    n = len(A)
    U = np.zeros_like(A)
    V = np.zeros_like(A)
    s = np.zeros((n, 3), dtype)
    _call("svd3", dtype, n, A, U, s, V)
    return U, s, V


def von_mises(F, U, sig, V, mly, dtype=np.float64):
    F, U, V = (np.ascontiguousarray(a, dtype=dtype).reshape(-1, 3, 3) for a in (F, U, V))
    sig = np.ascontiguousarray(sig, dtype=dtype).reshape(-1, 3)
    mly = np.ascontiguousarray(mly, dtype=dtype).reshape(-1, 3)
    n = len(F)
    Fo = np.zeros_like(F)
    J = np.zeros(n, dtype)
    _call("von_mises", dtype, n, F, U, sig, V, mly, Fo, J)
    return Fo, J


def shape_sdf(tfsr, args, pts, dtype=np.float64):
    pts = np.ascontiguousarray(pts, dtype=dtype).reshape(-1, 3)
    tfsr = np.ascontiguousarray(tfsr, dtype=dtype).reshape(4)
    args = np.ascontiguousarray(args, dtype=dtype).reshape(4)
    sdf = np.zeros(len(pts), dtype)
    grad = np.zeros_like(pts)
    _call("shape_sdf", dtype, len(pts), tfsr, args, pts, sdf, grad)
    return sdf, grad


def xform(pos, rot, pts, inverse=False, dtype=np.float64):
    pts = np.ascontiguousarray(pts, dtype=dtype).reshape(-1, 3)
    pos = np.ascontiguousarray(pos, dtype=dtype).reshape(3)
    rot = np.ascontiguousarray(rot, dtype=dtype).reshape(4)
    out = np.zeros_like(pts)
    _call("xform", dtype, len(pts), pos, rot, pts, out, int(inverse))
    return out


def grid_op(grid_m, grid_mv, grid_dim, body_pos, body_rot, next_pos, next_rot, tfsr, args, dx, dt, gravity,
            ground_friction=0.0, out_fill=0.0, dtype=np.float64):
    """One call of the grid update (integrator.cu:647-777) on a dense x-major grid.
    Returns (grid_v_out [G,3], tape [G, nb+1, 3]); nodes the update skips (m <= 1e-12)
    keep ``out_fill``. The tape holds the node velocity entering each body's contact,
    then the velocity after the last body (the reference's grid_body_v_in layout)."""
    G = int(np.prod(grid_dim))
    nb = len(tfsr)
    f = lambda a, w: np.ascontiguousarray(a, dtype=dtype).reshape(-1, w) if nb else np.zeros((1, w), dtype)
    m = np.ascontiguousarray(grid_m, dtype=dtype).reshape(G)
    mv = np.ascontiguousarray(grid_mv, dtype=dtype).reshape(G, 3)
    out = np.full((G, 3), out_fill, dtype)
    tape = np.zeros((G, nb + 1, 3), dtype)
    grid = np.ascontiguousarray(grid_dim, dtype=np.int32)
    _call("grid_op", dtype, m, mv, tape, f(body_pos, 3), f(body_rot, 4), f(next_pos, 3), f(next_rot, 4),
          f(tfsr, 4), f(args, 4), out, grid, nb, _F(dx), _F(dt), _F(ground_friction),
          np.ascontiguousarray(gravity, dtype=dtype))
    return out, tape


# ------------------------------------------------------------------------ scene + simulator
@dataclass
class Scene:
    """Constants of one scene, in the units the kernels see.

    ``gravity`` is the value the kernel adds (the reference uploads the configured
    gravity multiplied by 30, mpm/simulator.py:385). ``tfsr`` rows are
    (type, friction, softness, round) and ``args`` rows the shape arguments
    (mpm/cuda_env.py:80-101, mpm/simulator.py:396-398).
    """

    grid_dim: tuple
    dx: float
    dt: float
    gravity: tuple
    mass: np.ndarray
    vol: np.ndarray
    mly: np.ndarray  # [N,3] (mu, lambda, yield)
    tfsr: np.ndarray  # [nb,4]
    args: np.ndarray  # [nb,4]
    ground_friction: float = 0.0
    ground_height: float = 3.0

    @property
    def n_bodies(self):
        return int(len(self.tfsr))

    @property
    def n_particles(self):
        return int(len(self.mass))


@dataclass
class State:
    x: np.ndarray
    v: np.ndarray
    F: np.ndarray
    C: np.ndarray
    body_pos: np.ndarray
    body_rot: np.ndarray

    def astype(self, dtype):
        return State(*(np.ascontiguousarray(a, dtype=dtype) for a in
                       (self.x, self.v, self.F, self.C, self.body_pos, self.body_rot)))

    def copy(self):
        return State(*(a.copy() for a in (self.x, self.v, self.F, self.C, self.body_pos, self.body_rot)))


@dataclass
class Grads:
    x: np.ndarray
    v: np.ndarray
    F: np.ndarray
    C: np.ndarray
    body_pos: np.ndarray
    body_rot: np.ndarray

    @staticmethod
    def zeros(n, nb, dtype):
        return Grads(np.zeros((n, 3), dtype), np.zeros((n, 3), dtype), np.zeros((n, 3, 3), dtype),
                     np.zeros((n, 3, 3), dtype), np.zeros((nb, 3), dtype), np.zeros((nb, 4), dtype))


class OracleSim:
    """Reference kernel sequences on numpy buffers (one scene).

    dtype float32 follows the reference arithmetic; float64 is the FP64 restatement.
    ``omp=True`` uses the OpenMP build (atomic scatter, nondeterministic order).
    """

    def __init__(self, scene: Scene, dtype=np.float32, adjoint: str = "reference", omp: bool = False):
        self.s = scene
        self.dtype = np.dtype(dtype).type
        self.adjoint = ADJ_EXACT if adjoint == "exact" else ADJ_REFERENCE
        self.omp = omp
        dt_ = self.dtype
        self.grid = np.ascontiguousarray(scene.grid_dim, dtype=np.int32)
        G = int(np.prod(scene.grid_dim))
        self.G = G
        self.mass = np.ascontiguousarray(scene.mass, dt_)
        self.vol = np.ascontiguousarray(scene.vol, dt_)
        self.mly = np.ascontiguousarray(scene.mly, dt_).reshape(-1, 3)
        nb = max(scene.n_bodies, 1)
        self.tfsr = np.ascontiguousarray(scene.tfsr, dt_).reshape(-1, 4) if scene.n_bodies else np.zeros((nb, 4), dt_)
        self.args = np.ascontiguousarray(scene.args, dt_).reshape(-1, 4) if scene.n_bodies else np.zeros((nb, 4), dt_)
        self.gravity = np.ascontiguousarray(scene.gravity, dt_)

    # scalar helpers
    def _F(self, v):
        return _F(v)

    @property
    def inv_dx(self):
        return 1.0 / self.s.dx

    def _c(self, name, *args):
        _call(name, self.dtype, *args, omp=self.omp)

    def _temps(self, n):
        dt_ = self.dtype
        return dict(Ftmp=np.zeros((n, 3, 3), dt_), U=np.zeros((n, 3, 3), dt_), V=np.zeros((n, 3, 3), dt_),
                    sig=np.zeros((n, 3), dt_), grid_m=np.zeros(self.G, dt_), grid_mv=np.zeros((self.G, 3), dt_),
                    grid_v=np.zeros((self.G, 3), dt_))

    def _fwd_to_grid(self, st: State, nxt_pos, nxt_rot, tmp):
        s = self.s
        n = len(st.x)
        self._c("compute_svd", n, st.F, st.C, self._F(s.dt), tmp["Ftmp"], tmp["U"], tmp["V"], tmp["sig"])
        Fn = np.zeros_like(st.F)
        self._c("p2g", n, st.x, st.v, self.mass, self.vol, tmp["Ftmp"], tmp["U"], tmp["sig"], tmp["V"], st.C,
                self.mly, self.grid, self._F(s.dx), self._F(self.inv_dx), self._F(s.dt), Fn, tmp["grid_mv"],
                tmp["grid_m"])
        self._c("grid_op", tmp["grid_m"], tmp["grid_mv"], None, st.body_pos, st.body_rot, nxt_pos, nxt_rot,
                self.tfsr, self.args, tmp["grid_v"], self.grid, s.n_bodies, self._F(s.dx), self._F(s.dt),
                self._F(s.ground_friction), self.gravity)
        return Fn

    def substep(self, st: State, next_pos, next_rot) -> State:
        """One forward substep from ``st`` (poses of substep f) to f+1 (poses next_*)."""
        st = st.astype(self.dtype)
        next_pos = np.ascontiguousarray(next_pos, self.dtype).reshape(-1, 3) if self.s.n_bodies else st.body_pos
        next_rot = np.ascontiguousarray(next_rot, self.dtype).reshape(-1, 4) if self.s.n_bodies else st.body_rot
        s = self.s
        n = len(st.x)
        tmp = self._temps(n)
        Fn = self._fwd_to_grid(st, next_pos, next_rot, tmp)
        xn = np.zeros_like(st.x)
        vn = np.zeros_like(st.v)
        Cn = np.zeros_like(st.C)
        self._c("g2p", n, st.x, tmp["grid_v"], self.grid, self._F(s.dx), self._F(self.inv_dx), self._F(s.dt),
                self._F(s.ground_height), vn, Cn, xn)
        return State(xn, vn, Fn, Cn, next_pos.copy(), next_rot.copy())

    def substep_grad(self, st: State, nxt: State, g_next: Grads, g_cur: Grads) -> None:
        """Adjoint of one substep. Adds into ``g_cur`` (state f grads and body grads of
        f) and into ``g_next.body_pos/body_rot`` (body grads of f+1), as the reference
        accumulates into states[f] and states[f+1] (mpm/simulator.py:573-585)."""
        dt_ = self.dtype
        st = st.astype(dt_)
        nxt = nxt.astype(dt_)
        s = self.s
        n = len(st.x)
        tmp = self._temps(n)
        self._fwd_to_grid(st, nxt.body_pos, nxt.body_rot, tmp)
        G = self.G
        grid_v_grad = np.zeros((G, 3), dt_)
        grid_mv_grad = np.zeros((G, 3), dt_)
        grid_m_grad = np.zeros(G, dt_)
        self._c("g2p_grad", n, st.x, tmp["grid_v"], self.grid, self._F(s.dx), self._F(self.inv_dx), self._F(s.dt),
                self._F(s.ground_height), nxt.v, g_cur.x, grid_v_grad, g_next.v, g_next.C, g_next.x)
        self._c("grid_op_grad", tmp["grid_m"], tmp["grid_mv"], st.body_pos, st.body_rot, nxt.body_pos,
                nxt.body_rot, self.tfsr, self.args, grid_m_grad, grid_mv_grad, g_cur.body_pos, g_cur.body_rot,
                g_next.body_pos, g_next.body_rot, grid_v_grad, self.grid, s.n_bodies, self._F(s.dx),
                self._F(s.dt), self._F(s.ground_friction), self.gravity, self.adjoint)
        Ft_grad = np.zeros((n, 3, 3), dt_)
        U_grad = np.zeros((n, 3, 3), dt_)
        V_grad = np.zeros((n, 3, 3), dt_)
        sig_grad = np.zeros((n, 3), dt_)
        self._c("p2g_grad", n, st.x, st.v, self.mass, self.vol, tmp["Ftmp"], tmp["U"], tmp["sig"], tmp["V"], st.C,
                self.mly, self.grid, self._F(s.dx), self._F(self.inv_dx), self._F(s.dt), g_cur.x, g_cur.v, Ft_grad,
                g_cur.C, U_grad, sig_grad, V_grad, g_next.F, grid_mv_grad, grid_m_grad)
        self._c("compute_svd_grad", n, st.F, st.C, tmp["U"], tmp["V"], tmp["sig"], Ft_grad, U_grad, V_grad,
                sig_grad, g_cur.F, g_cur.C, self._F(s.dt))

    # ---------------------------------------------------------------- observations
    def compute_dist(self, st: State):
        n, nb = len(st.x), self.s.n_bodies
        st = st.astype(self.dtype)
        d = np.zeros((n, nb), self.dtype)
        self._c("compute_dist", n, st.x, st.body_pos, st.body_rot, self.tfsr, self.args, d, nb, None, None, None,
                None, 0)
        return d

    def compute_dist_grad(self, st: State, dist_grad, g: Grads):
        n, nb = len(st.x), self.s.n_bodies
        st = st.astype(self.dtype)
        dg = np.ascontiguousarray(dist_grad, self.dtype).reshape(n, nb)
        self._c("compute_dist", n, st.x, st.body_pos, st.body_rot, self.tfsr, self.args, None, nb, g.x,
                g.body_pos, g.body_rot, dg, 1)

    def grid_mass(self, x, ids=None, id=-1):
        x = np.ascontiguousarray(x, self.dtype).reshape(-1, 3)
        gm = np.zeros(self.G, self.dtype)
        ids_ = np.ascontiguousarray(ids if ids is not None else np.zeros(len(x)), np.int32)
        self._c("particle2mass", len(x), x, self.mass, self.grid, self._F(self.s.dx), self._F(self.inv_dx), gm,
                None, None, ids_, id, 0)
        return gm.reshape(self.s.grid_dim)

    def grid_mass_grad(self, x, grid_m_grad, x_grad, ids=None, id=-1):
        x = np.ascontiguousarray(x, self.dtype).reshape(-1, 3)
        ids_ = np.ascontiguousarray(ids if ids is not None else np.zeros(len(x)), np.int32)
        gmg = np.ascontiguousarray(grid_m_grad, self.dtype).reshape(-1)
        self._c("particle2mass", len(x), x, self.mass, self.grid, self._F(self.s.dx), self._F(self.inv_dx), None,
                gmg, x_grad, ids_, id, 1)

    # ---------------------------------------------------------------- rollouts
    def rollout(self, s0: State, poses_pos, poses_rot):
        """Run len(poses)-1 substeps; poses[t] are the body poses of substep t
        (poses[0] must equal s0's). Returns the list of T+1 states."""
        states = [s0.astype(self.dtype)]
        for t in range(1, len(poses_pos)):
            states.append(self.substep(states[-1], poses_pos[t], poses_rot[t]))
        return states

    def rollout_grad(self, states, g_last: Grads):
        """Backward through a stored rollout. Returns (grads of state 0,
        pose grads [T+1, nb, 3], [T+1, nb, 4])."""
        T = len(states) - 1
        n, nb = len(states[0].x), max(self.s.n_bodies, 1)
        gs = [Grads.zeros(n, nb, self.dtype) for _ in range(T + 1)]
        gs[T] = Grads(*(np.ascontiguousarray(a, self.dtype).copy() for a in
                        (g_last.x, g_last.v, g_last.F, g_last.C, g_last.body_pos, g_last.body_rot)))
        for t in range(T - 1, -1, -1):
            self.substep_grad(states[t], states[t + 1], gs[t + 1], gs[t])
        pos_g = np.stack([g.body_pos for g in gs])
        rot_g = np.stack([g.body_rot for g in gs])
        return gs[0], pos_g, rot_g


# ------------------------------------------------------------------------ scene helpers
def material(n, E=5e3, nu=0.2, yield_stress=50.0, vol=(1.0 / 64 / 2) ** 2, mass=(1.0 / 64 / 2) ** 2):
    """Per-particle arrays as MPMSimulator.__init__ builds them (mpm/simulator.py:211-217)."""
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return (np.zeros(n) + mass, np.zeros(n) + vol, np.zeros((n, 3)) + np.array([mu, lam, yield_stress]))
