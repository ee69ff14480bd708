"""Drop-in ``MPMSimulator`` for DexDeform's mpm/simulator.py on MI355X.

Same classes, constructor keywords, attributes and methods as the reference
(mpm/simulator.py:46-634). Differences in mechanism, not in results:

* buffers are torch tensors on the GPU (``DeviceArray`` keeps the reference's
  ``upload/download/zero`` surface, mpm/types.py:294-400);
* ``substep``/``substep_grad`` run the fused HIP entry points
  ``ddm_substep_forward``/``ddm_substep_backward`` (one launch per stage, block-sparse
  grid, no dense memsets); the per-stage methods (``compute_svd``, ``p2g``,
  ``grid_op``, ``g2p`` and their ``*_grad``) call the reference-shaped launchers of the
  same library on dense ``TempState`` buffers;
* poses stay on the device (``set_pose`` is a device copy, no host round trip);
* everything is enqueued on torch's current HIP stream; ``stream0``/``stream1`` are
  kept as attributes for compatibility and refer to that stream.

There is no CPU fallback: constructing a simulator without the HIP library raises.
"""
from __future__ import annotations

import ctypes
from types import SimpleNamespace
from typing import Optional

import numpy as np
import torch

from .. import _lib
from .engine import Engine, SceneSpec, StateBuffers, lame
from .kinematics import primitive_poses, rigid_body_motion  # noqa: F401  (re-exported as in the reference)

grid_lower_zero = np.array([[0, 0, 0]], dtype=np.int32)


class DeviceArray:
    """Thin stand-in for the reference's device ``array`` (mpm/types.py:294-400)."""

    def __init__(self, tensor: torch.Tensor):
        self.tensor = tensor

    @property
    def data_ptr(self) -> int:
        return int(self.tensor.data_ptr())

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    @property
    def length(self) -> int:
        return int(self.tensor.shape[0])

    def _rows(self, a):
        if isinstance(a, torch.Tensor):
            t = a.detach().to(device=self.tensor.device, dtype=self.tensor.dtype)
        else:
            t = torch.as_tensor(np.asarray(a), dtype=self.tensor.dtype, device=self.tensor.device)
        row = self.tensor.shape[1:]
        return t.reshape((-1,) + tuple(row))

    def upload(self, a, strict: bool = False):
        t = self._rows(a)
        if strict and t.shape[0] != self.tensor.shape[0]:
            raise ValueError(f"strict upload of {t.shape[0]} rows into {self.tensor.shape[0]}")
        if t.shape[0] > self.tensor.shape[0]:
            raise ValueError(f"upload of {t.shape[0]} rows exceeds capacity {self.tensor.shape[0]}")
        self.tensor[: t.shape[0]].copy_(t)

    def upload_async(self, a, stream=None):
        self.upload(a)

    def download(self, n: Optional[int] = None, device="numpy", stream=None):
        t = self.tensor if n is None else self.tensor[:n]
        if device == "numpy":
            return t.detach().cpu().numpy().copy()
        return t.detach().to(device).clone()

    def zero(self, stream=None):
        self.tensor.zero_()

    def zero_async(self, stream=None):
        self.tensor.zero_()

    def cuda_add(self, a):
        t = self._rows(a)
        self.tensor[: t.shape[0]] += t


class TempState:
    """Dense per-stage scratch of the reference (mpm/simulator.py:46-85). Only the
    per-stage methods use it; the fused substep keeps its own block-sparse grid."""

    def __init__(self, n_particles, grid_dim, n_bodies, device="cuda:0"):
        G = int(grid_dim[0] * grid_dim[1] * grid_dim[2])
        z = lambda *s: DeviceArray(torch.zeros(*s, dtype=torch.float32, device=device))
        self.grid_m, self.grid_v_in, self.grid_v_out = z(G), z(G, 3), z(G, 3)
        self.grid_m_grad, self.grid_v_in_grad, self.grid_v_out_grad = z(G), z(G, 3), z(G, 3)
        self.grid_body_v_in = z(G * (n_bodies + 1), 3)
        self.F, self.U, self.V, self.sig = z(n_particles, 9), z(n_particles, 9), z(n_particles, 9), z(n_particles, 3)
        self.F_grad, self.U_grad = z(n_particles, 9), z(n_particles, 9)
        self.V_grad, self.sig_grad = z(n_particles, 9), z(n_particles, 3)
        self.grid_lower = DeviceArray(torch.zeros(1, 3, dtype=torch.int32, device=device))

    def clear(self, stream=None):
        for a in (self.grid_m, self.grid_v_in, self.grid_v_out):
            a.zero()

    def clear_grad(self, stream=None):
        for a in (self.grid_v_in_grad, self.grid_v_out_grad, self.grid_m_grad, self.sig_grad, self.F_grad,
                  self.U_grad, self.V_grad):
            a.zero()


class State:
    """One substep's particle state, body poses and their adjoints (mpm/simulator.py:88-145)."""

    def __init__(self, n_particles, n_bodies, device="cuda:0"):
        self.n_particles = int(n_particles)
        self.n_bodies = int(n_bodies)
        self.buf = StateBuffers(self.n_particles, self.n_bodies, device, identity_F=False)
        self.grad = StateBuffers(self.n_particles, self.n_bodies, device, identity_F=False)
        b, g = self.buf, self.grad
        self.x, self.v = DeviceArray(b.x), DeviceArray(b.v)
        self.F, self.C = DeviceArray(b.F.view(-1, 9)), DeviceArray(b.C.view(-1, 9))
        self.x_grad, self.v_grad = DeviceArray(g.x), DeviceArray(g.v)
        self.F_grad, self.C_grad = DeviceArray(g.F.view(-1, 9)), DeviceArray(g.C.view(-1, 9))
        self.body_pos, self.body_rot = DeviceArray(b.body_pos), DeviceArray(b.body_rot)
        self.body_pos_grad, self.body_rot_grad = DeviceArray(g.body_pos), DeviceArray(g.body_rot)

    def reset(self):
        self.x.zero()
        self.v.zero()
        self.C.zero()
        self.buf.F[:] = torch.eye(3, device=self.buf.F.device)

    def get_state(self, n=None):
        x, v = self.x.download(n), self.v.download(n)
        F, C = self.F.download(n), self.C.download(n)
        return x, v, F.reshape(-1, 3, 3), C.reshape(-1, 3, 3), self.body_pos.download(), self.body_rot.download()

    def set_state(self, state):
        self.x.upload(state[0])
        self.v.upload(state[1])
        self.F.upload(np.asarray(state[2]).reshape(-1, 9))
        self.C.upload(np.asarray(state[3]).reshape(-1, 9))
        pos, rot = state[4:]
        self.body_pos.upload(pos)
        self.body_rot.upload(rot)

    def clear_grad(self, stream=None):
        self.grad.zero_()


def _cfg_namespace(cfg, defaults: dict):
    ns = dict(defaults)
    if cfg is not None:
        items = cfg.items() if hasattr(cfg, "items") else vars(cfg).items()
        for k, v in items:
            k2 = k.lower() if k.lower() in ns else k
            ns[k2] = v
    return SimpleNamespace(**ns)


class MPMSimulator:
    """Single-scene differentiable MLS-MPM simulator (mpm/simulator.py:151-634)."""

    def __init__(self, n_bodies, cfg=None, ground_friction=0., gravity=(0., -1., 0.), n_particles=20000,
                 dx=1. / 64, dt=0.0001, grid_size=(1., 1., 1.), max_steps=30, substeps=20, yield_stress=30.,
                 vol=(1. / 64 / 2) ** 2, mass=(1. / 64 / 2) ** 2, E=5000., nu=0.2, device="cuda:0",
                 adjoint="reference", sort_particles=True):
        """sort_particles: particles given in random order (as the reference's samplers
        produce them) are kept in spatial order inside the simulator (the particle
        kernels are 2-4x faster on it); every per-particle input and output of the
        public methods stays in the caller's order. ``states[i]`` buffers then hold the
        internal order."""
        kw = dict(ground_friction=ground_friction, gravity=gravity, n_particles=n_particles, dx=dx, dt=dt,
                  grid_size=grid_size, max_steps=max_steps, substeps=substeps, yield_stress=yield_stress, vol=vol,
                  mass=mass, E=E, nu=nu)
        self._cfg = c = _cfg_namespace(cfg, kw)
        self.device = torch.device(device)
        self.dx = float(c.dx)
        self.inv_dx = 1.0 / self.dx
        self.dt = float(c.dt)
        self.max_particles = self.n_particles = int(c.n_particles)
        self.n_bodies = int(n_bodies)
        self.substeps = int(c.substeps)
        grid_dim = (np.ceil(np.array(c.grid_size, dtype=np.float64) / self.dx / 4).astype(int) * 4)
        self.grid_dim_np = grid_dim.astype(np.int32)
        self.grid_dim = (ctypes.c_int * 3)(*[int(g) for g in grid_dim])
        self.grid_length = int(np.prod(grid_dim))
        self.n_grid = int(grid_dim[0])
        self.adjoint = adjoint
        self.max_steps = int(c.max_steps)
        mu, lam = lame(float(c.E), float(c.nu))
        N = self.n_particles
        self._spec = SceneSpec(
            n_particles=N, grid_dim=tuple(int(g) for g in grid_dim), dx=self.dx, dt=self.dt,
            gravity=tuple(np.float32(np.array(c.gravity, dtype=np.float64) * 30)),
            mass=np.zeros(N) + c.mass, vol=np.zeros(N) + c.vol,
            mu_lam_yield=np.zeros((N, 3)) + np.array([mu, lam, float(c.yield_stress)]),
            body_prim=np.zeros((self.n_bodies, 4), np.float32), body_args=np.zeros((self.n_bodies, 4), np.float32),
            ground_friction=float(c.ground_friction), ground_height=self.get_ground_height(),
            adjoint_mode=1 if adjoint == "exact" else 0)
        self.engine = Engine([self._spec], device=self.device, with_grad=True)
        self.states = [State(N, self.n_bodies, self.device) for _ in range(self.max_steps)]
        self._temp = None
        self.stream0 = self.stream1 = None  # all work runs on torch's current stream
        self.particle_color = DeviceArray(torch.zeros(N, dtype=torch.int32, device=self.device))
        self.particle_ids = DeviceArray(torch.zeros(N, dtype=torch.int32, device=self.device))
        self.dists = DeviceArray(torch.zeros(N * max(self.n_bodies, 1), dtype=torch.float32, device=self.device))
        self.dists_grad = DeviceArray(torch.zeros_like(self.dists.tensor))
        self.gravity = DeviceArray(torch.tensor(np.float32([np.array(c.gravity) * 30]), device=self.device))
        self.particle_vol = DeviceArray(self.engine.vol)
        self.particle_mass = DeviceArray(self.engine.mass)
        self.particle_mu_lam_yield = DeviceArray(self.engine.mu_lam_yield)
        self.body_type_mu_softness_round = DeviceArray(self.engine.body_prim)
        self.body_args = DeviceArray(self.engine.body_args)
        self._grid_lower = torch.zeros(1, 3, dtype=torch.int32, device=self.device)
        self.object_id = None
        # particle order: internal row i holds the caller's particle _perm[i] (None: same order)
        self.sort_particles = bool(sort_particles)
        self._perm = self._inv = self._perm_np = self._inv_np = None
        self._caller_particle_arrays = {}  # caller-order copies re-uploaded when the order changes
        self.torch_scale = None
        self.action_scales = None
        self.cur = 0
        for s in self.states:
            s.reset()

    # ------------------------------------------------------------------ properties
    @property
    def temp(self) -> TempState:
        if self._temp is None:
            self._temp = TempState(self.max_particles, self.grid_dim_np, self.n_bodies, self.device)
        return self._temp

    @property
    def lib(self):
        return self.engine.lib

    def _stream(self):
        return _lib.stream_ptr()

    # ------------------------------------------------------------------ particle order
    def _to_internal(self, a):
        """Rows in the caller's particle order -> the internal order."""
        if self._perm is None:
            return a
        if isinstance(a, torch.Tensor):
            return a[self._perm.to(a.device)]
        return np.asarray(a)[self._perm_np]

    def _to_caller(self, a):
        """Rows in the internal particle order -> the caller's order."""
        if self._perm is None:
            return a
        if isinstance(a, torch.Tensor):
            return a[self._inv.to(a.device)]
        return np.asarray(a)[self._inv_np]

    def _choose_order(self, x):
        """Keep the caller's order unless it is far from spatial (more than 10 % of
        neighbouring particles out of block/cell order, e.g. randomly sampled)."""
        from .workloads import needs_sort, spatial_order
        x = np.asarray(x, np.float32)
        if not self.sort_particles or not needs_sort(x, self.dx):
            self._perm = self._inv = self._perm_np = self._inv_np = None
        else:
            self._perm_np = spatial_order(x, self.dx)
            self._inv_np = np.argsort(self._perm_np)
            self._perm = torch.as_tensor(self._perm_np, device=self.device)
            self._inv = torch.as_tensor(self._inv_np, device=self.device)
        for name, a in self._caller_particle_arrays.items():  # per-particle inputs in the new order
            getattr(self, name).upload(self._to_internal(a))

    def _set_n(self, n):
        n = int(n)
        if n > self.max_particles:
            raise ValueError(f"{n} particles exceed the simulator capacity {self.max_particles}")
        self.n_particles = n
        self.engine.set_counts([n])

    # ------------------------------------------------------------------ state API
    def clear_grad(self, max_steps):
        for i in range(max_steps):
            self.states[i].clear_grad()

    def get_state(self, index):
        s = self.states[index].get_state(self.n_particles)
        return tuple(self._to_caller(a) for a in s[:4]) + tuple(np.r_[a, b] for a, b in zip(s[4], s[5]))

    def set_state(self, index, state):
        """(x, v, F, C, pose per body). At index 0, or for a new particle count, the
        particle order is chosen from x (see sort_particles); otherwise the current order
        applies (e.g. to a state returned by get_state)."""
        state = [np.float32(i) for i in state]
        if index == 0 or len(state[0]) != self.n_particles or (self._perm is None and self.sort_particles):
            self._choose_order(state[0])
        self._set_n(len(state[0]))
        pos = np.float32([i[:3] for i in state[4:]]).reshape(-1, 3)
        rot = np.float32([i[3:] for i in state[4:]]).reshape(-1, 4)
        self.states[index].set_state([self._to_internal(a) for a in state[:4]] + [pos, rot])
        self.cur = index

    def set_object_id(self, object_id):
        self.object_id = object_id
        assert len(object_id) == self.n_particles, f"{len(object_id)} {self.n_particles}"
        self._caller_particle_arrays["particle_ids"] = np.int32(object_id)
        self.particle_ids.upload(self._to_internal(np.int32(object_id)))

    def get_object_id(self, device):
        assert device == "numpy"
        return self.object_id

    def set_color(self, inp):
        color = self._to_caller(self.particle_color.download(self.n_particles))
        color[:] = inp
        self._caller_particle_arrays["particle_color"] = color
        self.particle_color.upload(self._to_internal(color))

    def set_softness(self, softness):
        pp = self.body_type_mu_softness_round.download()
        pp[:, 2] = softness
        self.body_type_mu_softness_round.upload(pp)

    def get_softness(self):
        return self.body_type_mu_softness_round.download()[:, 2]

    def get_x(self, index, device="numpy"):
        return self._to_caller(self.states[index].x.download(self.n_particles, device=device))

    def get_v(self, index, device="numpy"):
        return self._to_caller(self.states[index].v.download(self.n_particles, device=device))

    def get_tool_state(self, index=0, device=None):
        pos = self.states[index].body_pos.download()
        rot = self.states[index].body_rot.download()
        state = [i for i in np.concatenate((pos, rot), axis=1)]
        if device == "numpy":
            return state
        return [torch.tensor(i).to(device) for i in state]

    def set_tool_state(self, index, pose):
        pose = np.array(pose)
        assert pose.shape == (self.n_bodies, 7), f"{pose.shape}, {self.n_bodies}"
        self.states[index].body_pos.upload(pose[:, :3])
        self.states[index].body_rot.upload(pose[:, 3:7])

    def get_action_scales(self):
        return np.array(self.action_scales)

    def get_ground_friction(self):
        return float(self._cfg.ground_friction)

    def get_ground_height(self):
        return float(3.)

    def init_particles(self, vol, mass, mu_lam_yield):
        assert vol.shape == (self.n_particles,)
        assert mass.shape == (self.n_particles,)
        assert mu_lam_yield.shape == (self.n_particles, 3)
        for name, a in (("particle_vol", vol), ("particle_mass", mass), ("particle_mu_lam_yield", mu_lam_yield)):
            self._caller_particle_arrays[name] = np.float32(a)
            getattr(self, name).upload(self._to_internal(np.float32(a)))

    def init_bodies(self, types, softness, mu, round, args, action_scales, pos=None, rot=None):
        assert len(mu) == len(args) == self.n_bodies
        prim = np.stack((np.float32(types), np.float32(mu), np.float32(softness), np.float32(round)), 1)
        self.body_type_mu_softness_round.upload(prim.reshape(-1, 4))
        self.body_args.upload(np.float32(args).reshape(-1, 4))
        self.action_scales = action_scales
        self.torch_scale = None
        if pos is not None:
            self.states[0].body_pos.upload(np.array(pos).reshape(self.n_bodies, 3))
        if rot is not None:
            self.states[0].body_rot.upload(np.array(rot).reshape(self.n_bodies, 4))

    def sync(self):
        torch.cuda.current_stream().synchronize()

    # ------------------------------------------------------------------ observations
    def get_dists(self, f, grad=None, device="cuda:0"):
        state = self.states[f]
        n, nb = self.n_particles, self.n_bodies
        need_grad = grad is not None
        if need_grad:
            assert tuple(grad.shape) == (n, nb)
            g = self._to_internal(torch.as_tensor(grad, dtype=torch.float32, device=self.device))
            self.dists_grad.tensor[: n * nb].copy_(g.reshape(-1))
        _lib.check(self.lib.compute_dist(
            state.x.data_ptr, state.body_pos.data_ptr, state.body_rot.data_ptr,
            self.body_type_mu_softness_round.data_ptr, self.body_args.data_ptr, self.dists.data_ptr, nb,
            state.x_grad.data_ptr, state.body_pos_grad.data_ptr, state.body_rot_grad.data_ptr,
            self.dists_grad.data_ptr, int(need_grad), n, self._stream()), "compute_dist")
        if not need_grad:
            return self._to_caller(self.dists.download(n * nb, device=device).reshape(-1, nb))

    def compute_grid_mass(self, f, id=-1, device="numpy", backward_grad=None):
        need_grad = backward_grad is not None
        temp = self.temp
        if isinstance(f, int):
            n = self.n_particles
            state = self.states[f]
        else:
            x = torch.as_tensor(f, dtype=torch.float32) if not isinstance(f, torch.Tensor) else f.detach().float()
            assert x.shape[-1] == 3 and id == -1
            if len(x) == self.n_particles:  # positions of the scene's particles, in the caller's order
                x = self._to_internal(x.to(self.device))
            state = self.states[-1]
            state.x.upload(x)
            n = len(x)
            if need_grad:
                state.clear_grad()
        if need_grad:
            temp.grid_m_grad.tensor.copy_(torch.as_tensor(backward_grad, dtype=torch.float32).reshape(-1))
        temp.grid_m.zero()
        _lib.check(self.lib.particle2mass(
            state.x.data_ptr, self.particle_mass.data_ptr, int(self._grid_lower.data_ptr()), self.grid_dim, self.dx,
            self.inv_dx, temp.grid_m.data_ptr, temp.grid_m_grad.data_ptr, state.x_grad.data_ptr,
            self.particle_ids.data_ptr, int(id), int(need_grad), n, self._stream()), "particle2mass")
        if not need_grad:
            return temp.grid_m.download(device=device).reshape(*[int(g) for g in self.grid_dim_np])

    # ------------------------------------------------------------------ per-stage (reference launchers)
    def compute_svd(self, state: State, temp: TempState):
        _lib.check(self.lib.compute_svd(state.F.data_ptr, state.C.data_ptr, temp.F.data_ptr, temp.U.data_ptr,
                                        temp.V.data_ptr, temp.sig.data_ptr, self.dt, self.n_particles,
                                        self._stream()), "compute_svd")

    def compute_svd_grad(self, state: State, temp: TempState):
        _lib.check(self.lib.compute_svd_grad(
            state.F.data_ptr, state.C.data_ptr, temp.U.data_ptr, temp.V.data_ptr, temp.sig.data_ptr,
            temp.F_grad.data_ptr, temp.U_grad.data_ptr, temp.V_grad.data_ptr, temp.sig_grad.data_ptr,
            state.F_grad.data_ptr, state.C_grad.data_ptr, self.dt, self.n_particles, self._stream()),
            "compute_svd_grad")

    def p2g(self, state1: State, temp: TempState, state2: State):
        _lib.check(self.lib.p2g(
            state1.x.data_ptr, state1.v.data_ptr, self.particle_mass.data_ptr, self.particle_vol.data_ptr,
            temp.F.data_ptr, temp.U.data_ptr, temp.sig.data_ptr, temp.V.data_ptr, state1.C.data_ptr,
            self.particle_mu_lam_yield.data_ptr, temp.grid_lower.data_ptr, self.grid_dim, self.dx, self.inv_dx,
            self.dt, state2.F.data_ptr, temp.grid_v_in.data_ptr, temp.grid_m.data_ptr, self.n_particles,
            self._stream()), "p2g")

    def p2g_grad(self, state1: State, temp: TempState, state2: State):
        _lib.check(self.lib.p2g_grad(
            state1.x.data_ptr, state1.v.data_ptr, self.particle_mass.data_ptr, self.particle_vol.data_ptr,
            temp.F.data_ptr, temp.U.data_ptr, temp.sig.data_ptr, temp.V.data_ptr, state1.C.data_ptr,
            self.particle_mu_lam_yield.data_ptr, temp.grid_lower.data_ptr, self.grid_dim, self.dx, self.inv_dx,
            self.dt, state2.F.data_ptr, temp.grid_v_in.data_ptr, temp.grid_m.data_ptr,
            state1.x_grad.data_ptr, state1.v_grad.data_ptr, temp.F_grad.data_ptr, state1.C_grad.data_ptr,
            temp.U_grad.data_ptr, temp.sig_grad.data_ptr, temp.V_grad.data_ptr, state2.F_grad.data_ptr,
            temp.grid_v_in_grad.data_ptr, temp.grid_m_grad.data_ptr, self.n_particles, self._stream()), "p2g_grad")

    def grid_op(self, state1: State, temp: TempState, state2: State):
        _lib.check(self.lib.grid_op_v2(
            temp.grid_m.data_ptr, temp.grid_v_in.data_ptr, temp.grid_body_v_in.data_ptr, temp.grid_lower.data_ptr,
            self.gravity.data_ptr, state1.body_pos.data_ptr, state1.body_rot.data_ptr, state2.body_pos.data_ptr,
            state2.body_rot.data_ptr, self.body_type_mu_softness_round.data_ptr, self.body_args.data_ptr, self.dx,
            self.inv_dx, self.dt, self.get_ground_friction(), temp.grid_v_out.data_ptr, self.grid_dim,
            self.n_bodies, self._stream()), "grid_op_v2")

    def grid_op_grad(self, state1: State, temp: TempState, state2: State):
        _lib.check(self.lib.grid_op_v2_grad(
            temp.grid_m.data_ptr, temp.grid_v_in.data_ptr, temp.grid_body_v_in.data_ptr, temp.grid_lower.data_ptr,
            self.gravity.data_ptr, state1.body_pos.data_ptr, state1.body_rot.data_ptr, state2.body_pos.data_ptr,
            state2.body_rot.data_ptr, self.body_type_mu_softness_round.data_ptr, self.body_args.data_ptr,
            temp.grid_m_grad.data_ptr, temp.grid_v_in_grad.data_ptr, state1.body_pos_grad.data_ptr,
            state1.body_rot_grad.data_ptr, state2.body_pos_grad.data_ptr, state2.body_rot_grad.data_ptr, self.dx,
            self.inv_dx, self.dt, self.get_ground_friction(), temp.grid_v_out.data_ptr,
            temp.grid_v_out_grad.data_ptr, self.grid_dim, self.n_bodies, self._stream()), "grid_op_v2_grad")

    def g2p(self, state1: State, temp: TempState, state2: State):
        _lib.check(self.lib.g2p(
            state1.x.data_ptr, temp.grid_v_out.data_ptr, temp.grid_lower.data_ptr, self.dx, self.inv_dx, self.dt,
            self.grid_dim, state2.v.data_ptr, self.get_ground_height(), state2.C.data_ptr, state2.x.data_ptr,
            self.n_particles, self._stream()), "g2p")

    def g2p_grad(self, state1: State, temp: TempState, state2: State):
        _lib.check(self.lib.g2p_grad(
            state1.x.data_ptr, temp.grid_v_out.data_ptr, temp.grid_lower.data_ptr, self.dx, self.inv_dx, self.dt,
            self.grid_dim, state2.v.data_ptr, self.get_ground_height(), state2.C.data_ptr, state2.x.data_ptr,
            self.n_particles, state1.x_grad.data_ptr, temp.grid_v_out_grad.data_ptr, state2.v_grad.data_ptr,
            state2.C_grad.data_ptr, state2.x_grad.data_ptr, self._stream()), "g2p_grad")

    def substep_stages(self, f):
        """Forward substep through the per-stage launchers (the reference's exact
        sequence, mpm/simulator.py:561-571)."""
        temp, cur, nxt = self.temp, self.states[f], self.states[f + 1]
        temp.clear()
        self.compute_svd(cur, temp)
        self.p2g(cur, temp, nxt)
        self.grid_op(cur, temp, nxt)
        self.g2p(cur, temp, nxt)

    def substep_grad_stages(self, f):
        """Adjoint substep through the per-stage launchers (mpm/simulator.py:573-585)."""
        temp, cur, nxt = self.temp, self.states[f], self.states[f + 1]
        temp.clear()
        temp.clear_grad()
        self.compute_svd(cur, temp)
        self.p2g(cur, temp, nxt)
        self.grid_op(cur, temp, nxt)
        self.g2p_grad(cur, temp, nxt)
        self.grid_op_grad(cur, temp, nxt)
        self.p2g_grad(cur, temp, nxt)
        self.compute_svd_grad(cur, temp)

    # ------------------------------------------------------------------ fused substeps
    def set_pose(self, state: State, pos, rot, stream=None):
        state.body_pos.upload(pos)
        state.body_rot.upload(rot)

    def substep(self, f, clear_grad=False):
        cur, nxt = self.states[f], self.states[f + 1]
        self.engine.forward(cur.buf, nxt.buf)
        if clear_grad:
            nxt.clear_grad()

    def substep_grad(self, f):
        cur, nxt = self.states[f], self.states[f + 1]
        self.engine.backward(cur.buf, nxt.buf, cur.grad, nxt.grad)

    def download_pos_rot(self, cur, device):
        s = self.states[cur]
        return s.body_pos.tensor.detach().to(device).clone(), s.body_rot.tensor.detach().to(device).clone()

    def compute_forward_kinematics(self, f, action, pos_rot=None):
        device = action.device if isinstance(action, torch.Tensor) else self.device
        if pos_rot is None:
            pos, rot = self.download_pos_rot(f, device)
        else:
            pos, rot = pos_rot
        if not isinstance(action, torch.Tensor):
            action = torch.tensor(np.array(action, np.float32), device=device)
        else:
            action = action.to(device)
        if self.torch_scale is None or self.torch_scale.device != torch.device(device):
            self.torch_scale = torch.tensor(np.array(self.action_scales), device=device, dtype=torch.float32)
        pos, rot = primitive_poses(pos, rot, action, self.torch_scale, self.substeps)
        # the carried kinematic state is the pose after the last substep (the reference
        # returns the whole trajectory here, mpm/simulator.py:624, which GradModel.forward
        # cannot feed back in for the next env step)
        return pos, rot, (pos[-1], rot[-1])

    def step(self, action, pos_rot=None):
        pos, rot, q_state = self.compute_forward_kinematics(self.cur, action, pos_rot=pos_rot)
        for f in range(self.substeps):
            self.set_pose(self.states[f + 1], pos[f], rot[f])
            self.substep(f)
        self.sync()
        self.states[0], self.states[self.substeps] = self.states[self.substeps], self.states[0]
