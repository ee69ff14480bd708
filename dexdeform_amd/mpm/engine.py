"""Device-resident engine for one or many MLS-MPM scenes.

An :class:`Engine` owns everything the fused kernels need for a fixed set of scenes:
per-particle material arrays, primitive parameters, the block-sparse grid work arrays
(4x4x4-blocked, float4 per node) and the scene/chunk tables. States are
:class:`StateBuffers` (particle x, v, F, C and body poses) allocated by the caller, so
a single-scene simulator, a batched rollout and the rollout server all drive the same
two entry points, ``ddm_substep_forward`` and ``ddm_substep_backward``.

The forward substep replaces the reference's ``compute_svd -> p2g -> grid_op_v2 -> g2p``
sequence (mpm/simulator.py:561-571); the backward replaces
``substep_grad`` (mpm/simulator.py:573-585) with the same recompute-then-adjoint order.
"""
from __future__ import annotations

import contextlib
import ctypes
import gc
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from .. import _lib

CHUNK = 256  # particles per workgroup in the particle kernels
# floats per particle of one substep's SVD temporaries (U and V as quaternions: 4 each,
# sigma: 3); the adjoint recomputes F_tmp = (I + dt C) F from the state
TEMP_FLOATS = 11


@dataclass
class SceneSpec:
    """Constants of one scene, in kernel units.

    gravity is the acceleration the grid update adds (the reference uploads the configured
    gravity multiplied by 30, mpm/simulator.py:385). body_prim rows are
    (type, friction, softness, round) and body_args rows the shape arguments
    (mpm/simulator.py:396-398, mpm/cuda_env.py:80-101).
    """

    n_particles: int
    grid_dim: Sequence[int]
    dx: float
    dt: float
    gravity: Sequence[float]
    mass: np.ndarray
    vol: np.ndarray
    mu_lam_yield: np.ndarray
    body_prim: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    body_args: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    ground_friction: float = 0.0
    ground_height: float = 3.0
    adjoint_mode: int = 0  # 0 reference adjoint, 1 exact

    @property
    def n_bodies(self) -> int:
        return int(len(self.body_prim))

    def validate(self):
        gd = tuple(int(g) for g in self.grid_dim)
        if len(gd) != 3 or any(g < 8 or g > 1024 or g % 4 for g in gd):
            raise ValueError(f"grid dims must be multiples of 4 in [8, 1024], got {gd}")
        if self.n_bodies > _lib.DDM_MAX_BODIES:
            raise ValueError(f"at most {_lib.DDM_MAX_BODIES} bodies per scene, got {self.n_bodies}")
        for name in ("mass", "vol"):
            if np.asarray(getattr(self, name)).shape != (self.n_particles,):
                raise ValueError(f"{name} must have shape ({self.n_particles},)")
        if np.asarray(self.mu_lam_yield).shape != (self.n_particles, 3):
            raise ValueError("mu_lam_yield must have shape (n_particles, 3)")


def lame(E: float, nu: float):
    """(mu, lambda) as MPMSimulator.__init__ computes them (mpm/simulator.py:211-212)."""
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


class StateBuffers:
    """Particle state (x, v, F, C) of every scene of an engine plus body poses, stored
    concatenated (scene s owns rows p_begin[s] : p_begin[s] + capacity[s])."""

    def __init__(self, n_particles: int, n_bodies: int, device, identity_F: bool = True):
        z = lambda *s: torch.zeros(*s, dtype=torch.float32, device=device)
        self.x = z(n_particles, 3)
        self.v = z(n_particles, 3)
        self.F = z(n_particles, 3, 3)
        if identity_F and n_particles:
            self.F[:] = torch.eye(3, device=device)
        self.C = z(n_particles, 3, 3)
        self.body_pos = z(max(n_bodies, 0), 3)
        self.body_rot = z(max(n_bodies, 0), 4)
        if n_bodies:
            self.body_rot[:, 0] = 1.0

    def tensors(self):
        return (self.x, self.v, self.F, self.C, self.body_pos, self.body_rot)

    def zero_(self):
        for t in self.tensors():
            t.zero_()
        return self

    def copy_(self, other: "StateBuffers"):
        for a, b in zip(self.tensors(), other.tensors()):
            a.copy_(b)
        return self


@contextlib.contextmanager
def gc_paused():
    """Python's automatic cyclic garbage collection paused (for a HIP graph capture): a
    collection during a capture could free a dead graph held in a reference cycle, and a
    graph destroyed while a stream is capturing fails (hipErrorStreamCaptureUnsupported,
    which the graph's destructor turns into std::terminate). Dead cycles are collected
    after the capture, as usual."""
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def chunk_size_for(n_particles: int) -> int:
    """Particles per chunk for a launch of n_particles: 256 (one per thread) when that
    still gives >= 160 chunks, else 128 or 64, so that small launches spread over more
    CUs (a 10k-particle rollout is 28 % faster with 64-particle chunks; from ~40k
    particles 256 is fastest)."""
    for c in (CHUNK, CHUNK // 2):
        if n_particles >= 160 * c:
            return c
    return CHUNK // 4


def chunk_table(counts: Sequence[int], scene_ids: Optional[Sequence[int]] = None, chunk: int = CHUNK) -> np.ndarray:
    """(scene, first, end, 0) rows covering each scene's particles (local indices) in
    <= chunk pieces (chunk <= CHUNK: one workgroup of CHUNK threads per row)."""
    if not 0 < chunk <= CHUNK:
        raise ValueError(f"chunk must be in (0, {CHUNK}]")
    ids = range(len(counts)) if scene_ids is None else scene_ids
    parts = []
    for s, n in zip(ids, counts):
        first = np.arange(0, int(n), chunk, dtype=np.int32)
        t = np.zeros((len(first), 4), np.int32)
        t[:, 0], t[:, 1], t[:, 2] = s, first, np.minimum(first + chunk, int(n))
        parts.append(t)
    return np.concatenate(parts) if parts else np.zeros((0, 4), np.int32)


def state_table(records, device) -> torch.Tensor:
    """Device array of ddm_state records from (x, v, F, C, body_pos, body_rot) address tuples."""
    tab = np.zeros(len(records), dtype=_lib.STATE_DTYPE)
    for i, r in enumerate(records):
        tab[i] = tuple(int(a) for a in r)
    return torch.as_tensor(tab.view(np.uint8).copy(), device=device)


class Engine:
    def __init__(self, scenes: List[SceneSpec], device="cuda:0", with_grad: bool = True,
                 deterministic: bool = False):
        """deterministic=True: every substep (forward and adjoint) uses sorted reductions
        instead of atomics, so results are bitwise reproducible and identical for a scene
        alone and in any batch (slower; needs a workspace of about 150 B per particle
        and 512 B per node block plus 64 B per block and body)."""
        if not scenes:
            raise ValueError("an engine needs at least one scene")
        for s in scenes:
            s.validate()
        self.lib = _lib.load()
        self.device = torch.device(device)
        self.scenes = scenes
        self.with_grad = with_grad
        caps = [int(s.n_particles) for s in scenes]
        self.p_begin = np.concatenate([[0], np.cumsum(caps)[:-1]]).astype(np.int64)
        self.capacity = caps
        self.n_particles_total = int(sum(caps))
        nbs = [s.n_bodies for s in scenes]
        self.b_begin = np.concatenate([[0], np.cumsum(nbs)[:-1]]).astype(np.int64)
        self.n_bodies_total = int(sum(nbs))
        nblk = [(int(s.grid_dim[0]) // 4) * (int(s.grid_dim[1]) // 4) * (int(s.grid_dim[2]) // 4) for s in scenes]
        self.blk_begin = np.concatenate([[0], np.cumsum(nblk)[:-1]]).astype(np.int64)
        self.n_blocks = int(sum(nblk))
        if self.n_blocks * 64 >= 2 ** 31 or self.n_particles_total * 9 >= 2 ** 31:
            raise ValueError("batch too large for 32-bit node/particle indices")

        dev = self.device
        f32 = lambda a, shape: torch.as_tensor(np.asarray(a, np.float32).reshape(shape), device=dev)
        P, NB = self.n_particles_total, self.n_bodies_total
        self.mass = f32(np.concatenate([np.asarray(s.mass, np.float32) for s in scenes]), (P,))
        self.vol = f32(np.concatenate([np.asarray(s.vol, np.float32) for s in scenes]), (P,))
        self.mu_lam_yield = f32(np.concatenate([np.asarray(s.mu_lam_yield, np.float32).reshape(-1, 3)
                                                for s in scenes]), (P, 3))
        prim = [np.asarray(s.body_prim, np.float32).reshape(-1, 4) for s in scenes]
        args = [np.asarray(s.body_args, np.float32).reshape(-1, 4) for s in scenes]
        self.body_prim = f32(np.concatenate(prim) if NB else np.zeros((0, 4)), (NB, 4))
        self.body_args = f32(np.concatenate(args) if NB else np.zeros((0, 4)), (NB, 4))

        G = self.n_blocks * 64
        z = lambda *s: torch.zeros(*s, dtype=torch.float32, device=dev)
        self.grid_mmv = z(G, 4)
        self.grid_v = z(G, 4)
        self.block_flags = torch.zeros(self.n_blocks, dtype=torch.int32, device=dev)
        self.block_list = torch.zeros(self.n_blocks, 2, dtype=torch.int32, device=dev)
        self.block_count = torch.zeros(1, dtype=torch.int32, device=dev)
        if with_grad:
            self.grid_dv = z(G, 4)
            self.grid_dmmv = z(G, 4)
            self.tmp_qu = z(P, 4)
            self.tmp_qv = z(P, 4)
            self.tmp_sig = z(P, 3)
        else:
            self.grid_dv = self.grid_dmmv = self.tmp_qu = self.tmp_qv = self.tmp_sig = None

        self.deterministic = bool(deterministic)
        self._det = None
        if self.deterministic:
            slots = max(sum((c + CHUNK - 1) // CHUNK for c in caps), 1) * CHUNK
            i32 = lambda n: torch.zeros(int(n), dtype=torch.int32, device=dev)
            d = dict(cell_count=i32(G), cell_start=i32(G), order=i32(slots), order_p=i32(slots),
                     key_slot=i32(2 * slots), records=z(slots, 16), sorted=z(slots, 16),
                     body_partial=z(self.n_blocks * max(max(nbs), 1), 16),
                     range_count=i32((self.n_blocks + 1023) // 1024), counter=i32(1))
            self.det_buffers = d
            self._det = _lib.DdmDet(*[_lib.ptr(d[k]) for k in ("cell_count", "cell_start", "order", "order_p",
                                                               "key_slot", "records", "sorted", "body_partial",
                                                               "range_count", "counter")])

        tab = np.zeros(len(scenes), dtype=_lib.SCENE_DTYPE)
        f4 = 4
        for i, s in enumerate(scenes):
            pb, bb = int(self.p_begin[i]), int(self.b_begin[i])
            tab[i]["p_count"] = s.n_particles
            tab[i]["blk_begin"] = self.blk_begin[i]
            tab[i]["n_blocks"] = nblk[i]
            tab[i]["dim"] = [int(g) for g in s.grid_dim]
            tab[i]["n_bodies"] = s.n_bodies
            tab[i]["adjoint_mode"] = int(s.adjoint_mode)
            tab[i]["dx"] = s.dx
            tab[i]["inv_dx"] = 1.0 / float(s.dx)
            tab[i]["dt"] = s.dt
            tab[i]["ground_friction"] = s.ground_friction
            tab[i]["ground_height"] = s.ground_height
            tab[i]["gravity"] = np.asarray(s.gravity, np.float32)
            tab[i]["mass"] = self.mass.data_ptr() + pb * f4
            tab[i]["vol"] = self.vol.data_ptr() + pb * f4
            tab[i]["mu_lam_yield"] = self.mu_lam_yield.data_ptr() + 3 * pb * f4
            if NB:
                tab[i]["body_prim"] = self.body_prim.data_ptr() + 4 * bb * f4
                tab[i]["body_args"] = self.body_args.data_ptr() + 4 * bb * f4
            if with_grad:
                tab[i]["tmp_qu"] = self.tmp_qu.data_ptr() + 4 * pb * f4
                tab[i]["tmp_qv"] = self.tmp_qv.data_ptr() + 4 * pb * f4
                tab[i]["tmp_sig"] = self.tmp_sig.data_ptr() + 3 * pb * f4
        self._scene_np = tab
        self.scene_table = torch.as_tensor(tab.view(np.uint8).copy(), device=dev)
        self.counts = list(caps)
        self.chunks = None
        self._tables = {}
        self._set_chunks(self.counts)
        self._batch = None
        self._rebuild_batch()

    # ------------------------------------------------------------------ tables
    def _set_chunks(self, counts):
        # the deterministic workspace is sized for full chunks
        chunk = CHUNK if self.deterministic else chunk_size_for(sum(counts))
        tab = chunk_table(counts, chunk=chunk)
        self.n_chunks = int(len(tab))
        self.chunks = torch.as_tensor(tab if len(tab) else np.zeros((1, 4), np.int32), device=self.device)

    def set_counts(self, counts: Sequence[int]):
        """Number of live particles per scene (<= capacity); the rest are ignored."""
        counts = [int(c) for c in counts]
        for c, cap in zip(counts, self.capacity):
            if c < 0 or c > cap:
                raise ValueError(f"particle count {c} outside [0, {cap}]")
        if counts != self.counts:
            self.counts = counts
            self._set_chunks(counts)
            self._rebuild_batch()

    def _rebuild_batch(self):
        """New launch record. Bumps ``generation``: launch tables and chunk tensors of
        earlier generations may be freed, so graphs captured before must not be replayed."""
        p = _lib.ptr
        self.generation = getattr(self, "generation", 0) + 1
        self._batch = _lib.DdmBatch(
            len(self.scenes), self.n_chunks, self.n_blocks, max(s.n_bodies for s in self.scenes),
            p(self.scene_table), p(self.chunks), p(self.grid_mmv), p(self.grid_v), p(self.grid_dv), p(self.grid_dmmv),
            p(self.block_flags), p(self.block_list), p(self.block_count), None,
            ctypes.pointer(self._det) if self._det is not None else None, self._options())

    def _options(self, extra: int = 0) -> int:
        return (_lib.DDM_DETERMINISTIC if self.deterministic else 0) | extra

    def table(self, x, v, F, C, body_pos, body_rot) -> torch.Tensor:
        """Per-scene ddm_state records for concatenated tensors (cached by address)."""
        key = tuple(int(t.data_ptr()) for t in (x, v, F, C, body_pos, body_rot))
        tab = self._tables.get(key)
        if tab is None:
            recs = []
            for i in range(len(self.scenes)):
                pb, bb = int(self.p_begin[i]), int(self.b_begin[i])
                recs.append((key[0] + 12 * pb, key[1] + 12 * pb, key[2] + 36 * pb, key[3] + 36 * pb,
                             key[4] + 12 * bb if key[4] else 0, key[5] + 16 * bb if key[5] else 0))
            tab = state_table(recs, self.device)
            self._tables[key] = tab
        return tab

    def table_of(self, buf) -> torch.Tensor:
        return self.table(*buf.tensors())

    def update_body_params(self, body_prim: torch.Tensor = None, body_args: torch.Tensor = None):
        if body_prim is not None:
            self.body_prim.copy_(torch.as_tensor(body_prim, dtype=torch.float32).reshape(self.body_prim.shape))
        if body_args is not None:
            self.body_args.copy_(torch.as_tensor(body_args, dtype=torch.float32).reshape(self.body_args.shape))

    # ------------------------------------------------------------------ states
    def new_state(self) -> StateBuffers:
        return StateBuffers(self.n_particles_total, self.n_bodies_total, self.device)

    def new_grad(self) -> StateBuffers:
        return StateBuffers(self.n_particles_total, self.n_bodies_total, self.device, identity_F=False)

    def scene_slice(self, i: int, n: Optional[int] = None) -> slice:
        b = int(self.p_begin[i])
        return slice(b, b + (self.counts[i] if n is None else n))

    def body_slice(self, i: int) -> slice:
        b = int(self.b_begin[i])
        return slice(b, b + self.scenes[i].n_bodies)

    # ------------------------------------------------------------------ substeps
    def forward(self, cur, nxt, stream=None):
        """One forward substep of every scene; cur/nxt are StateBuffers (or anything
        with x, v, F, C, body_pos, body_rot tensors laid out like them)."""
        self.forward_tables(self.table_of(cur), self.table_of(nxt), stream)

    def backward(self, cur, nxt, gcur, gnxt, stream=None, overwrite=False):
        self.backward_tables(self.table_of(cur), self.table_of(nxt), self.table_of(gcur), self.table_of(gnxt), stream,
                             overwrite)

    def forward_tables(self, cur_tab, nxt_tab, stream=None, prev_tab=None, finish=True):
        """One forward substep. In a sequence of substeps: finish=False leaves this
        substep's G2P to the next call, which passes prev_tab (this call's cur_tab) and
        runs it first, inside its own P2G launch for small launches."""
        if prev_tab is not None or not finish:
            _lib.check(self.lib.ddm_substep_forward_ahead(
                ctypes.byref(self._batch), cur_tab.data_ptr(), nxt_tab.data_ptr(),
                prev_tab.data_ptr() if prev_tab is not None else None, int(bool(finish)), 0,
                _lib.stream_ptr(stream)), "ddm_substep_forward_ahead")
            return
        _lib.check(self.lib.ddm_substep_forward(ctypes.byref(self._batch), cur_tab.data_ptr(), nxt_tab.data_ptr(),
                                                _lib.stream_ptr(stream)), "ddm_substep_forward")

    def backward_tables(self, cur_tab, nxt_tab, gcur_tab, gnxt_tab, stream=None, overwrite=False):
        """Adjoint substep; adds into the current state's gradients, or writes them
        (overwrite=True: no clearing needed beforehand)."""
        if not self.with_grad:
            raise RuntimeError("engine was created with with_grad=False")
        b = self._batch_with(options=_lib.DDM_GRAD_OVERWRITE if overwrite else 0)
        _lib.check(self.lib.ddm_substep_backward(ctypes.byref(b), cur_tab.data_ptr(), nxt_tab.data_ptr(),
                                                 gcur_tab.data_ptr(), gnxt_tab.data_ptr(), _lib.stream_ptr(stream)),
                   "ddm_substep_backward")

    # ------------------------------------------------------------------ taped adjoint
    # a snapshot block holds (m, mv) and v of 64 nodes: 2 KiB
    TAPE_BLOCK_BYTES = 2048
    # whole-grid snapshots for every scene when they take at most this much per substep
    TAPE_COMPLETE_BYTES = 64 << 20

    def grid_blocks(self) -> List[int]:
        return [(int(s.grid_dim[0]) // 4) * (int(s.grid_dim[1]) // 4) * (int(s.grid_dim[2]) // 4) for s in self.scenes]

    def tape_capacities(self, whole_grid: Optional[bool] = None) -> List[int]:
        """Snapshot size per scene (blocks). whole_grid=True: every block of each grid (the
        adjoint then never recomputes, see tape_complete); False: the whole grid of a small
        one, else particles / 32 + 64 blocks (a substep that lists more is recomputed by its
        adjoint); None (default): whole grids when all scenes' whole grids take at most
        TAPE_COMPLETE_BYTES per substep."""
        blocks = self.grid_blocks()
        if whole_grid is None:
            whole_grid = sum(blocks) * self.TAPE_BLOCK_BYTES <= self.TAPE_COMPLETE_BYTES
        if whole_grid:
            return blocks
        return [min(b, int(c) // 32 + 64) for b, c in zip(blocks, self.capacity)]

    def tape_complete(self, caps) -> bool:
        """Whether snapshots of these capacities can hold every block each scene's grid has
        (no scene's substep can overflow: the taped adjoint skips its recompute kernels)."""
        return len(caps) == len(self.scenes) and all(int(c) >= b for c, b in zip(caps, self.grid_blocks()))

    def taped_table(self, temps, tape_mmv, tape_v, tape_list, tape_status, caps) -> torch.Tensor:
        """Scene table of one taped substep: each scene's SVD temporaries in ``temps``
        (TEMP_FLOATS * n_particles_total floats: U and V as quaternions | sigma) and its grid
        snapshot in the next caps[i] blocks of tape_mmv / tape_v / tape_list (scenes in
        order), with two status ints per scene in ``tape_status``."""
        P = self.n_particles_total
        if temps.numel() < TEMP_FLOATS * P or temps.dtype != torch.float32 or not temps.is_contiguous():
            raise ValueError(f"temps must be a contiguous float32 tensor of {TEMP_FLOATS} * n_particles_total elements")
        need = int(sum(caps))
        if tape_mmv.numel() < need * 256 or tape_v.numel() < need * 256 or tape_list.numel() < need:
            raise ValueError("tape buffers smaller than the scenes' capacities")
        base, f4 = int(temps.data_ptr()), 4
        tab = self._scene_np.copy()
        off = 0
        for i in range(len(tab)):
            pb = int(self.p_begin[i])
            tab[i]["tmp_qu"] = base + 4 * pb * f4
            tab[i]["tmp_qv"] = base + (4 * P + 4 * pb) * f4
            tab[i]["tmp_sig"] = base + (8 * P + 3 * pb) * f4
            tab[i]["tape_mmv"] = int(tape_mmv.data_ptr()) + off * 256 * f4
            tab[i]["tape_v"] = int(tape_v.data_ptr()) + off * 256 * f4
            tab[i]["tape_list"] = int(tape_list.data_ptr()) + off * 4
            tab[i]["tape_status"] = int(tape_status.data_ptr()) + 2 * i * 4
            tab[i]["tape_capacity"] = int(caps[i])
            off += int(caps[i])
        return torch.as_tensor(tab.view(np.uint8).copy(), device=self.device)

    def _batch_with(self, scene_tab=None, options=0):
        b = _lib.DdmBatch.from_buffer_copy(self._batch)
        if scene_tab is not None:
            b.scenes = scene_tab.data_ptr()
        b.options = self._options(options)
        return b

    def _batch_with_scenes(self, scene_tab):
        return self._batch_with(scene_tab)

    def forward_taped(self, cur_tab, nxt_tab, scene_tab, stream=None, prev_tab=None, finish=True):
        """Forward substep recording, per scene, the snapshot and temporaries that the
        scene table from taped_table points to. prev_tab / finish as in forward_tables."""
        b = self._batch_with_scenes(scene_tab)
        if prev_tab is not None or not finish:
            _lib.check(self.lib.ddm_substep_forward_ahead(
                ctypes.byref(b), cur_tab.data_ptr(), nxt_tab.data_ptr(),
                prev_tab.data_ptr() if prev_tab is not None else None, int(bool(finish)), 1,
                _lib.stream_ptr(stream)), "ddm_substep_forward_ahead")
            return
        _lib.check(self.lib.ddm_substep_forward_taped(ctypes.byref(b), cur_tab.data_ptr(), nxt_tab.data_ptr(),
                                                      _lib.stream_ptr(stream)), "ddm_substep_forward_taped")

    def backward_taped(self, cur_tab, nxt_tab, gcur_tab, gnxt_tab, scene_tab, stream=None, overwrite=False,
                       complete=False, restored=False, prev_scene_tab=None):
        """Adjoint substep from what forward_taped recorded with the same scene table.
        complete=True: every scene's snapshot holds its whole grid (tape_complete), so the
        recompute kernels for overflowing scenes are not launched. prev_scene_tab (the
        previous substep's scene table): this substep's P2G adjoint launch also puts that
        substep's snapshots back, and its own call then passes restored=True (not in the
        deterministic mode)."""
        if not self.with_grad:
            raise RuntimeError("engine was created with with_grad=False")
        opts = (_lib.DDM_GRAD_OVERWRITE if overwrite else 0) | (_lib.DDM_TAPE_COMPLETE if complete else 0)
        b = self._batch_with(scene_tab, opts)
        if restored or prev_scene_tab is not None:
            prev = prev_scene_tab.data_ptr() if prev_scene_tab is not None else None
            _lib.check(self.lib.ddm_substep_backward_taped_ahead(
                ctypes.byref(b), cur_tab.data_ptr(), nxt_tab.data_ptr(), gcur_tab.data_ptr(), gnxt_tab.data_ptr(),
                int(bool(restored)), prev, _lib.stream_ptr(stream)), "ddm_substep_backward_taped_ahead")
            return
        _lib.check(self.lib.ddm_substep_backward_taped(ctypes.byref(b), cur_tab.data_ptr(), nxt_tab.data_ptr(),
                                                       gcur_tab.data_ptr(), gnxt_tab.data_ptr(),
                                                       _lib.stream_ptr(stream)), "ddm_substep_backward_taped")

    def reset_grid(self, stream=None):
        _lib.check(self.lib.ddm_batch_reset(ctypes.byref(self._batch), _lib.stream_ptr(stream)), "ddm_batch_reset")

    def nbytes(self) -> int:
        ts = [self.mass, self.vol, self.mu_lam_yield, self.body_prim, self.body_args, self.grid_mmv, self.grid_v,
              self.block_flags, self.block_list, self.block_count, self.grid_dv, self.grid_dmmv, self.tmp_qu, self.tmp_qv,
              self.tmp_sig]
        if self._det is not None:
            ts += list(self.det_buffers.values())
        return int(sum(t.numel() * t.element_size() for t in ts if t is not None))
