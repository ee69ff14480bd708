"""Continuous-batching server for differentiable MPM rollouts.

Each request is a trajectory-optimisation rollout of one scene: ``T`` forward substeps
with given per-substep primitive poses, an L2 loss of the final particle positions to a
target, and the adjoint pass producing d loss / d poses and d loss / d initial state.
Requests of different particle count, grid size and horizon share the GPU:

* admission: a request is admitted when its grid blocks fit in the grid arena and its
  device memory (states, checkpoints, temporaries) fits the memory budget; until then
  it waits in FIFO order. Slots and memory are returned when a request finishes, and
  waiting requests are admitted at the next tick (continuous batching, not a queue in
  front of a one-at-a-time executor).
* execution: every tick advances every admitted request by one substep operation.
  All forward-type operations (forward substeps and recomputation of checkpointed
  segments) run in one fused launch sequence over all those scenes, and all adjoint
  substeps run in one fused adjoint launch sequence (``ddm_substep_forward`` /
  ``ddm_substep_backward`` with per-scene state tables).
* state stays on the device for the whole request; only the small scene/chunk/state
  tables are uploaded per tick, from a pinned host ring with one asynchronous copy into
  tables at fixed device addresses, so the host prepares the next tick while the GPU
  runs the current one (a synchronous pageable upload would wait for all queued kernels
  first).
* graphs: a tick's two launch sequences are captured once per launch size (chunk counts
  rounded up to powers of two) into a HIP graph and replayed on later ticks of that
  size; the kernels read the live chunk and scene counts from the uploaded tables and
  the chunks beyond them exit at once. ``use_graphs=False`` launches every tick eagerly,
  sized exactly.
* gradients: the adjoint launches write (not add to) the state-t gradients
  (DDM_GRAD_OVERWRITE), so no clearing kernels run between ticks.

* persistent loop: ``start()`` runs the tick loop in a background thread; requests
  submitted with ``submit_async`` from any thread while it runs join at the next tick and
  each returns a future resolved with its ``Result`` (``poll()`` runs one iteration for a
  caller that drives the loop itself; ``run`` replays a fixed arrival trace).

Latency of a request = completion (an event recorded after its last operation) minus
arrival; throughput = particle-substeps (forward + adjoint) per second of wall time.
"""
from __future__ import annotations

import ctypes
import heapq
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from .. import _lib
from .engine import CHUNK, TEMP_FLOATS, SceneSpec, chunk_size_for, chunk_table, gc_paused
from .workloads import needs_sort, spatial_order


@dataclass
class Request:
    rid: int
    spec: SceneSpec
    x0: np.ndarray               # [n, 3]
    pose_pos: np.ndarray         # [T+1, nb, 3]
    pose_rot: np.ndarray         # [T+1, nb, 4]
    target: np.ndarray           # [n, 3]
    arrival: float = 0.0         # seconds after the start of the stream
    checkpoint_every: Optional[int] = None

    @property
    def T(self) -> int:
        return int(len(self.pose_pos) - 1)

    @property
    def n(self) -> int:
        return int(len(self.x0))


@dataclass
class Result:
    rid: int
    loss: float
    x_final: torch.Tensor
    pos_grad: torch.Tensor
    rot_grad: torch.Tensor
    x0_grad: torch.Tensor
    arrival: float
    admitted: float
    finished: float
    finite: bool = True  # loss and gradients finite; a failing request does not affect others

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class BlockArena:
    """First-fit allocator of contiguous node-block ranges in the grid arena."""

    def __init__(self, n_blocks: int):
        self.n = n_blocks
        self.free = [(0, n_blocks)]  # sorted (start, length)

    def alloc(self, k: int) -> Optional[int]:
        for i, (s, l) in enumerate(self.free):
            if l >= k:
                if l == k:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + k, l - k)
                return s
        return None

    def release(self, s: int, k: int):
        self.free.append((s, k))
        self.free.sort()
        merged = []
        for a, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((a, l))
        self.free = merged

    @property
    def free_blocks(self) -> int:
        return sum(l for _, l in self.free)


class _View:
    """x, v, F, C views of one particle-state set inside a request's state block."""

    def __init__(self, flat, n):
        self.flat = flat  # contiguous x | v | F | C
        self.x = flat[0:3 * n].view(n, 3)
        self.v = flat[3 * n:6 * n].view(n, 3)
        self.F = flat[6 * n:15 * n].view(n, 3, 3)
        self.C = flat[15 * n:24 * n].view(n, 3, 3)

    def tensors(self):
        return (self.x, self.v, self.F, self.C)

    def copy_(self, o):
        self.flat.copy_(o.flat)  # one copy for the four arrays

    def zero_(self):
        self.flat.zero_()  # one fill kernel for the four arrays


class _StateSets:
    """count particle-state sets in one uninitialised allocation (every set is written
    before it is read: by a substep, a checkpoint copy or an explicit initialisation).
    Views are made on first access; addresses() gives the x, v, F, C addresses of every
    set without making them."""

    def __init__(self, n, count, device, start=0, block=None):
        self.n, self.start, self.count = n, start, count
        self.block = torch.empty(count, 24 * n, dtype=torch.float32, device=device) if block is None else block
        self._views = {}

    def __len__(self):
        return self.count

    def __getitem__(self, i):
        if isinstance(i, slice):
            a, b, step = i.indices(self.count)
            assert step == 1
            return _StateSets(self.n, b - a, None, self.start + a, self.block)
        if i < 0:
            i += self.count
        if not 0 <= i < self.count:
            raise IndexError(i)
        v = self._views.get(i)
        if v is None:
            v = self._views[i] = _View(self.block[self.start + i], self.n)
        return v

    def addresses(self) -> np.ndarray:
        """[count, 4] int64: x, v, F, C addresses of each set."""
        base = int(self.block.data_ptr()) + self.start * 24 * self.n * 4
        rows = base + np.arange(self.count, dtype=np.int64)[:, None] * (24 * self.n * 4)
        return rows + np.array([0, 3, 6, 15], dtype=np.int64)[None] * (self.n * 4)


def _state_sets(n, count, device):
    return _StateSets(n, count, device)


def _to_device(a, device, dtype=np.float32) -> torch.Tensor:
    """Asynchronous upload through freshly allocated pinned memory (the caching host
    allocator keeps it until the copy has run)."""
    h = torch.from_numpy(np.ascontiguousarray(a, dtype)).pin_memory()
    return h.to(device, non_blocking=True)


def _upload_packed(arrays, device) -> List[torch.Tensor]:
    """Upload float32 arrays through one pinned host buffer and one asynchronous copy;
    returns device views of the arrays' shapes, each starting on a 256-byte boundary."""
    arrays = [np.asarray(a, np.float32) for a in arrays]
    offs = np.cumsum([0] + [(a.size + 63) // 64 * 64 for a in arrays])
    host = torch.empty(int(offs[-1]), dtype=torch.float32, pin_memory=True)
    hn = host.numpy()
    for a, o in zip(arrays, offs):
        hn[o:o + a.size] = a.reshape(-1)
    d = host.to(device, non_blocking=True)
    return [d[o:o + a.size].view(a.shape) for a, o in zip(arrays, offs)]


class _TickTables:
    """Per-tick launch tables at fixed device addresses, so that a tick's launches can be
    captured once in a HIP graph and replayed with each tick's contents.

    A tick's key -- (forward scenes, forward chunks, adjoint scenes, adjoint chunks), the
    launch sizes, which the live counts may undercut -- fixes a compact layout: for the
    forward launch, then the adjoint launch, the scene records, the state records (current,
    next; for the adjoint also their gradients), the live counts (chunks, scenes) and the
    chunk table, each 64-byte aligned and sized for the key. A tick's contents are packed
    into one slot of a pinned host ring at its key's offsets and copied to the device in
    one asynchronous transfer of the layout's span (about 1.4 KB for one request of up to
    64 chunks) on the launch stream; a host slot is reused only after the copy that read it
    completed. Launches on the same stream consume the device tables in order, so one
    device region suffices."""

    ALIGN = 64
    PARTS = {"F": ("recs", "cur", "nxt", "live", "chunks"),
             "B": ("recs", "cur", "nxt", "gcur", "gnxt", "live", "chunks")}

    def __init__(self, device, max_scenes, max_chunks, slots=8):
        self.device = device
        self.max_scenes = int(max_scenes)
        self.max_chunks = int(max_chunks)
        self.nbytes = self._offsets((self.max_scenes, self.max_chunks, self.max_scenes, self.max_chunks))[1]
        self.dev = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        self.base = self.dev.data_ptr()
        self.host = [torch.empty(self.nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(slots)]
        self.events = [None] * slots
        self.i = 0
        self._layouts = {}

    @staticmethod
    def _part_bytes(part, scenes, chunks):
        if part == "recs":
            return _lib.SCENE_DTYPE.itemsize * scenes
        if part == "live":
            return 8
        if part == "chunks":
            return 16 * chunks
        return 48 * scenes  # a ddm_state record per scene

    @classmethod
    def _offsets(cls, key):
        fs, fc, bs, bc = key
        off, cap, o = {}, {}, 0
        for kind, sc, ch in (("F", fs, fc), ("B", bs, bc)):
            if not sc:
                continue
            for part in cls.PARTS[kind]:
                off[kind, part] = o
                cap[kind, part] = cls._part_bytes(part, sc, ch)
                o += (cap[kind, part] + cls.ALIGN - 1) // cls.ALIGN * cls.ALIGN
        return off, o, cap

    def fits(self, key) -> bool:
        return self._offsets(key)[1] <= self.nbytes

    def layout(self, key):
        """(device addresses, offsets, span, part capacities) of a key's tables."""
        lay = self._layouts.get(key)
        if lay is None:
            off, span, cap = self._offsets(key)
            lay = ({k: self.base + o for k, o in off.items()}, off, span, cap)
            self._layouts[key] = lay
        return lay

    def stage(self, key, contents):
        """contents: {(kind, part): numpy array}; packed at the key's offsets and copied."""
        _, off, span, cap = self.layout(key)
        slot = self.i
        self.i = (self.i + 1) % len(self.host)
        if self.events[slot] is not None:
            self.events[slot].synchronize()
        h = self.host[slot].numpy()
        for k, arr in contents.items():
            b = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
            if b.size > cap[k]:
                raise ValueError(f"{k} table exceeds its capacity")
            h[off[k]:off[k] + b.size] = b
        self.dev[:span].copy_(self.host[slot][:span], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self.events[slot] = ev


def _bucket(n: int, least: int) -> int:
    """Launch capacity of a captured tick: the power of two >= n (at least `least`)."""
    b = least
    while b < n:
        b *= 2
    return b


def _state_records(records) -> np.ndarray:
    """ddm_state records (six device addresses each) from lists of ints."""
    return np.array(records, dtype=np.int64).reshape(-1, 6)


def tape_capacity(req: Request) -> int:
    """Grid blocks one substep's snapshot of the request holds (a substep that lists more
    is recomputed by its adjoint)."""
    gd = req.spec.grid_dim
    return min((gd[0] // 4) * (gd[1] // 4) * (gd[2] // 4), req.n // 32 + 64)


def request_bytes(req: Request, K: Optional[int] = None) -> int:
    """Device bytes a request holds while active, with checkpoints every K substeps
    (default: the request's interval; K = T keeps every state and tapes every substep)."""
    T = req.T
    K = K or req.checkpoint_every or T
    n_states = (T // K + 1 if K != T else 0) + (K + 1) + 2
    per_state = req.n * 24 * 4
    temps = req.n * 5 * 4 + K * (req.n * TEMP_FLOATS * 4 + tape_capacity(req) * (2 * 1024 + 4) + 8)
    nb = max(req.spec.n_bodies, 1)
    poses = (T + 1) * nb * 7 * 4 * 2
    gd = req.spec.grid_dim
    blocks = (gd[0] // 4) * (gd[1] // 4) * (gd[2] // 4)
    return n_states * per_state + temps + poses + blocks * 64 * 16 * 4


class _Active:
    """Device-side state and operation schedule of an admitted request."""

    def __init__(self, server: "RolloutServer", req: Request, blk_begin: int, now: float, K: int):
        dev = server.device
        self.req = req
        self.admitted = now
        sp = req.spec
        self.blk_begin = blk_begin
        gd = [int(g) for g in sp.grid_dim]
        self.n_blocks = (gd[0] // 4) * (gd[1] // 4) * (gd[2] // 4)
        n, nb, T = req.n, sp.n_bodies, req.T
        self.K = int(K)
        if T % self.K:
            raise ValueError("checkpoint_every must divide the horizon")
        self.bytes = request_bytes(req, self.K)
        # particles far from spatial order (4^3 block, then cell) are sorted for the kernels;
        # results are returned in the request's order
        self.order = spatial_order(req.x0, sp.dx) if needs_sort(req.x0, sp.dx) else None
        srt = (lambda a: np.asarray(a)[self.order]) if self.order is not None else (lambda a: a)
        self.inv = _to_device(np.argsort(self.order), dev, np.int64) if self.order is not None else None
        up = [srt(sp.mass), srt(sp.vol), srt(np.asarray(sp.mu_lam_yield).reshape(-1, 3)),
              srt(np.asarray(req.x0).reshape(-1, 3)), srt(np.asarray(req.target).reshape(-1, 3))]
        if nb:
            up += [np.asarray(sp.body_prim).reshape(-1, 4), np.asarray(sp.body_args).reshape(-1, 4),
                   np.asarray(req.pose_pos).reshape(T + 1, nb, 3), np.asarray(req.pose_rot).reshape(T + 1, nb, 4)]
        dv = _upload_packed(up, dev)
        self.mass, self.vol, self.mly, x0, self.target = dv[:5]
        self.prim, self.args = (dv[5], dv[6]) if nb else (None, None)
        # taped adjoint: per substep of a segment, the SVD temporaries and a grid snapshot
        # (written by the segment's taped forward pass, read by its adjoint substeps)
        self.tape_cap = tape_capacity(req)
        self.tape_temps = torch.empty(self.K, TEMP_FLOATS * n, dtype=torch.float32, device=dev)
        self.tape_mmv = torch.empty(self.K, self.tape_cap * 256, dtype=torch.float32, device=dev)
        self.tape_v = torch.empty_like(self.tape_mmv)
        self.tape_list = torch.empty(self.K, self.tape_cap, dtype=torch.int32, device=dev)
        self.tape_status = torch.zeros(self.K, 2, dtype=torch.int32, device=dev)
        if nb:
            self.pose_pos, self.pose_rot = dv[7], dv[8]
        else:
            self.pose_pos = torch.zeros(T + 1, 1, 3, device=dev)
            self.pose_rot = torch.zeros(T + 1, 1, 4, device=dev)
            self.pose_rot[..., 0] = 1
        self.pos_grad = torch.zeros_like(self.pose_pos)
        self.rot_grad = torch.zeros_like(self.pose_rot)
        self.full = self.K == T
        n_ckpt = 0 if self.full else T // self.K + 1
        sets = _state_sets(n, n_ckpt + self.K + 1 + 2, dev)
        self.ckpt = sets[:n_ckpt]
        self.seg = sets[n_ckpt:n_ckpt + self.K + 1]
        self.g = sets[n_ckpt + self.K + 1:]
        s0 = self.seg[0]
        s0.x.copy_(x0)
        s0.v.zero_()
        s0.C.zero_()
        s0.F.zero_()
        s0.F.diagonal(dim1=1, dim2=2).fill_(1.0)
        if not self.full:
            self.ckpt[0].copy_(self.seg[0])
        self.record = np.zeros(1, dtype=_lib.SCENE_DTYPE)[0]
        r = self.record
        r["p_count"] = n
        r["blk_begin"] = blk_begin
        r["n_blocks"] = self.n_blocks
        r["dim"] = gd
        r["n_bodies"] = nb
        r["adjoint_mode"] = int(sp.adjoint_mode)
        r["dx"], r["inv_dx"], r["dt"] = sp.dx, 1.0 / sp.dx, sp.dt
        r["ground_friction"], r["ground_height"] = sp.ground_friction, sp.ground_height
        r["gravity"] = np.asarray(sp.gravity, np.float32)
        r["mass"], r["vol"], r["mu_lam_yield"] = self.mass.data_ptr(), self.vol.data_ptr(), self.mly.data_ptr()
        r["body_prim"] = self.prim.data_ptr() if nb else 0
        r["body_args"] = self.args.data_ptr() if nb else 0
        self._temps_of(r, 0)  # untaped forward substeps write no temporaries; any slot will do
        # (0, first, end, 0) rows; column 0 = scene slot. Smaller chunks for small requests,
        # which often run alone at low load
        self.chunks = chunk_table([n], chunk=chunk_size_for(n)).astype(np.int32)
        self.ops = self._schedule()
        self._addresses()
        self.pc = 0
        self.done_event = None
        self.loss = None

    def _temps_of(self, r, j):
        base, n, f4 = int(self.tape_temps[j].data_ptr()), self.req.n, 4
        r["tmp_qu"], r["tmp_qv"], r["tmp_sig"] = base, base + 4 * n * f4, base + 8 * n * f4

    # ---- operation schedule: ("pre", callable) run before the tick's launches,
    #      ("F", j, t, taped) forward substep seg[j] -> seg[j+1] at substep t (taped: it
    #      records slot j's tape for the adjoint), ("B", j, t, gcur, gnxt) adjoint substep
    #      from slot j's tape.
    def _schedule(self):
        T, K = self.req.T, self.K
        ops: List[tuple] = []
        nseg = T // K
        for c in range(nseg):
            if not self.full and c > 0:
                ops.append(("pre", lambda c=c: self.seg[0].copy_(self.ckpt[c]), c))
            for j in range(K):
                ops.append(("F", j, c * K + j, self.full))
            if not self.full:
                ops.append(("post", lambda c=c: self.ckpt[c + 1].copy_(self.seg[K]), c))
        ops.append(("loss", None, None))
        cur = 0
        for c in range(nseg - 1, -1, -1):
            if not self.full:
                ops.append(("pre", lambda c=c: self.seg[0].copy_(self.ckpt[c]), c))
                for j in range(K):
                    ops.append(("F", j, c * K + j, True))
            for j in range(K - 1, -1, -1):
                nxt, cur = cur, 1 - cur
                ops.append(("B", j, c * K + j, cur, nxt))
        self.final_grad_slot = cur
        return ops

    def _final(self):
        return self.seg[self.K] if self.full else self.ckpt[-1]

    def loss_and_seed(self):
        x = self._final().x
        d = x - self.target
        self.loss = (d * d).sum(-1).mean()
        # read back through pinned memory once the request's completion event has fired;
        # .item() here would synchronize the whole stream
        self.loss_host = torch.empty((), dtype=torch.float32, pin_memory=True)
        self.loss_host.copy_(self.loss, non_blocking=True)
        g = self.g[0]
        g.zero_()
        g.x.copy_(2.0 / self.req.n * d)
        self.pos_grad.zero_()
        self.rot_grad.zero_()
        self.x_final = x.clone()

    def check_finite(self):
        """Enqueue the check that the loss and the gradients are finite; read back through
        pinned memory once the completion event has fired (no synchronisation here)."""
        g0 = self.g[self.final_grad_slot]
        ok = (torch.isfinite(self.loss) & torch.isfinite(self.pos_grad).all() & torch.isfinite(self.rot_grad).all()
              & torch.isfinite(g0.x).all())
        self.finite_host = torch.empty((), dtype=torch.bool, pin_memory=True)
        self.finite_host.copy_(ok, non_blocking=True)

    def _addresses(self):
        """Device addresses behind the per-launch state records and the per-slot taped scene
        records (fixed for the request), computed from the allocations' base addresses."""
        self.seg_addr = self.seg.addresses().tolist()
        self.g_addr = self.g.addresses().tolist()
        rows = self.pose_pos.shape[1]
        self.pose_addr = (int(self.pose_pos.data_ptr()), int(self.pose_rot.data_ptr()), rows * 12, rows * 16)
        self.pgrad_addr = (int(self.pos_grad.data_ptr()), int(self.rot_grad.data_ptr()))
        self.rec_bytes = np.array([self.record], dtype=_lib.SCENE_DTYPE).view(np.uint8)
        K, n, f4 = self.K, self.req.n, 4
        j = np.arange(K, dtype=np.uint64)
        recs = np.repeat(np.array([self.record], dtype=_lib.SCENE_DTYPE), K)
        slot = lambda t: np.uint64(t.data_ptr()) + j * np.uint64(t.stride(0) * t.element_size())
        temps = slot(self.tape_temps)
        recs["tmp_qu"], recs["tmp_qv"] = temps, temps + np.uint64(4 * n * f4)
        recs["tmp_sig"] = temps + np.uint64(8 * n * f4)
        recs["tape_mmv"], recs["tape_v"] = slot(self.tape_mmv), slot(self.tape_v)
        recs["tape_list"], recs["tape_status"] = slot(self.tape_list), slot(self.tape_status)
        recs["tape_capacity"] = self.tape_cap
        self.rec_taped = recs.view(np.uint8).reshape(K, -1)

    def record_for(self, op):
        """Scene record of a kernel operation: slot j's taped record, or the plain one."""
        taped = op[0] == "B" or op[3]
        return self.rec_taped[op[1]] if taped else self.rec_bytes

    def state_record(self, j, t):
        pp, pr, sp, sr = self.pose_addr
        return self.seg_addr[j] + [pp + t * sp, pr + t * sr]

    def grad_record(self, slot, t):
        _, _, sp, sr = self.pose_addr
        gp, gr = self.pgrad_addr
        return self.g_addr[slot] + [gp + t * sp, gr + t * sr]

    def next_kernel_op(self):
        """Run host-side pre/post operations up to the next kernel operation and return
        it (or None when the schedule is finished)."""
        while self.pc < len(self.ops):
            op = self.ops[self.pc]
            if op[0] in ("pre", "post"):
                op[1]()
                self.pc += 1
            elif op[0] == "loss":
                self.loss_and_seed()
                self.pc += 1
            else:
                return op
        return None


class RolloutServer:
    def __init__(self, device="cuda:0", arena_blocks: int = 64 * 13824, memory_budget: Optional[int] = None,
                 max_active: int = 256, use_graphs: bool = True, full_tape_bytes: Optional[int] = None):
        """memory_budget: device bytes the admitted requests may hold (default: 60 % of the
        device's memory). full_tape_bytes: a request whose whole trajectory (every state,
        every substep taped) takes at most this many bytes, admitted while at most half
        the budget is in use, runs without checkpoints -- 2T substep operations instead of
        3T, same results up to FP32 rounding; otherwise it uses its checkpoint interval
        (more requests fit). Default: an eighth of the budget; 0 always honours the
        request's interval."""
        self.lib = _lib.load()
        self.device = torch.device(device)
        self.arena = BlockArena(arena_blocks)
        G = arena_blocks * 64
        z = lambda *s: torch.zeros(*s, dtype=torch.float32, device=self.device)
        self.grid_mmv, self.grid_v, self.grid_dv, self.grid_dmmv = z(G, 4), z(G, 4), z(G, 4), z(G, 4)
        self.block_flags = torch.zeros(arena_blocks, dtype=torch.int32, device=self.device)
        self.block_list = torch.zeros(arena_blocks, 2, dtype=torch.int32, device=self.device)
        self.block_count = torch.zeros(1, dtype=torch.int32, device=self.device)
        if memory_budget is None:
            memory_budget = int(0.6 * torch.cuda.get_device_properties(self.device).total_memory)
        self.memory_budget = int(memory_budget)
        self.full_tape_bytes = self.memory_budget // 8 if full_tape_bytes is None else int(full_tape_bytes)
        self.max_active = max_active
        self.waiting: List[Request] = []
        self.active: List[_Active] = []
        self.results: Dict[int, Result] = {}
        self.used_bytes = 0
        # per-tick tables; chunk capacity grows (and captured graphs are dropped) on demand
        self.tables = _TickTables(self.device, max_active, 4096)
        self.use_graphs = use_graphs
        self.capture_stream = torch.cuda.Stream(self.device)
        self.graphs: Dict[tuple, torch.cuda.CUDAGraph] = {}
        self.host_s = {"admit": 0.0, "launch": 0.0, "tick": 0.0}  # host seconds spent per phase
        self.t0 = None
        self.ticks = 0
        self.particle_substeps = 0
        self.replays = 0  # ticks that replayed a captured graph
        # persistent loop: submissions from other threads, futures of unfinished requests
        self._inbox: "queue.SimpleQueue" = queue.SimpleQueue()
        self._futures: Dict[int, Future] = {}
        self._thread: Optional[threading.Thread] = None
        self._stop = threading.Event()
        self._outstanding = set()  # futures not yet resolved
        self._outstanding_lock = threading.Lock()

    # ------------------------------------------------------------------ admission
    def submit(self, req: Request):
        """Queues a request of a trace (arrival relative to the serving loop's start);
        call from the thread that runs the loop (see submit_async for other threads)."""
        self._validate(req)
        self.waiting.append(req)

    def submit_async(self, req: Request) -> Future:
        """Queues a request from any thread while the loop runs (``start``/``poll``); it
        arrives now and joins at the next tick. Returns a future resolved with its Result
        (or with the exception that stopped the loop). Invalid requests raise here."""
        self._validate(req)
        fut: Future = Future()
        with self._outstanding_lock:
            self._outstanding.add(fut)
        fut.add_done_callback(self._forget)
        self._inbox.put((req, fut))
        return fut

    def _forget(self, fut):
        with self._outstanding_lock:
            self._outstanding.discard(fut)

    def _validate(self, req: Request):
        if req.T < 1:
            raise ValueError(f"request {req.rid}: the horizon must be at least one substep")
        if req.checkpoint_every is not None and (req.checkpoint_every < 1 or req.T % req.checkpoint_every):
            raise ValueError(f"request {req.rid}: checkpoint_every={req.checkpoint_every} must divide "
                             f"the horizon {req.T}")
        if len(req.pose_rot) != len(req.pose_pos) or len(req.target) != req.n:
            raise ValueError(f"request {req.rid}: pose or target arrays do not match the horizon / particles")
        if request_bytes(req) > self.memory_budget:
            raise ValueError(f"request {req.rid} needs more memory than the whole budget")
        gd = req.spec.grid_dim
        if (gd[0] // 4) * (gd[1] // 4) * (gd[2] // 4) > self.arena.n:
            raise ValueError(f"request {req.rid} grid exceeds the arena")
        req.spec.validate()

    def interval(self, req: Request) -> int:
        """Checkpoint interval the request would run with if admitted now (see
        full_tape_bytes)."""
        full = request_bytes(req, req.T)
        if full <= self.full_tape_bytes and self.used_bytes + full <= self.memory_budget // 2:
            return req.T
        return req.checkpoint_every or req.T

    def _now(self):
        return time.perf_counter() - self.t0

    def _admit(self, now):
        # admission is host work (allocations, uploads, tables); spend at most ~2 ms of it
        # per tick while requests are running, so the device keeps receiving ticks
        t_start = time.perf_counter()
        while self.waiting and len(self.active) < self.max_active:
            if self.active and time.perf_counter() - t_start > 2e-3:
                break
            req = self.waiting[0]
            if req.arrival > now:
                break
            K = self.interval(req)
            need = request_bytes(req, K)
            if self.used_bytes + need > self.memory_budget:
                break
            gd = req.spec.grid_dim
            nblk = (gd[0] // 4) * (gd[1] // 4) * (gd[2] // 4)
            b0 = self.arena.alloc(nblk)
            if b0 is None:
                break
            t0 = time.perf_counter()
            try:
                act = _Active(self, req, b0, now, K)
            except Exception:
                # nothing was committed yet: give the blocks back and leave the request
                # queued for the caller to inspect (the running requests are unaffected)
                self.arena.release(b0, nblk)
                raise
            self.waiting.pop(0)
            self.used_bytes += need
            self.active.append(act)
            self.host_s["admit"] += time.perf_counter() - t0

    # ------------------------------------------------------------------ launches
    def _contents(self, entries, kind):
        """Tables of one fused launch over entries [(active, op)]: scene records, chunk
        table (column 0 = position of the scene in the launch), state records."""
        recs = np.concatenate([a.record_for(op) for a, op in entries])
        chunks = np.concatenate([a.chunks for a, _ in entries])
        k = 0
        for si, (a, _) in enumerate(entries):
            chunks[k:k + len(a.chunks), 0] = si
            k += len(a.chunks)
        c = {(kind, "recs"): recs, (kind, "chunks"): chunks,
             (kind, "cur"): _state_records([a.state_record(op[1], op[2]) for a, op in entries]),
             (kind, "nxt"): _state_records([a.state_record(op[1] + 1, op[2] + 1) for a, op in entries]),
             (kind, "live"): np.array([len(chunks), len(entries)], dtype=np.int32)}
        if kind == "B":
            c[kind, "gcur"] = _state_records([a.grad_record(op[3], op[2]) for a, op in entries])
            c[kind, "gnxt"] = _state_records([a.grad_record(op[4], op[2] + 1) for a, op in entries])
        return c

    def _launch(self, kind, n_scenes, n_chunks, live, A):
        """One fused launch sequence over the staged tables of `kind` ("F": forward-type
        substeps, "B": adjoint substeps) at addresses A. With `live`, the launch is sized
        for n_scenes / n_chunks and its kernels read the counts actually used from the
        staged tables (what a captured tick replays); without, it is sized exactly."""
        batch = _lib.DdmBatch(n_scenes, n_chunks, self.arena.n, _lib.DDM_MAX_BODIES, A[kind, "recs"],
                              A[kind, "chunks"], self.grid_mmv.data_ptr(), self.grid_v.data_ptr(),
                              self.grid_dv.data_ptr(), self.grid_dmmv.data_ptr(), self.block_flags.data_ptr(),
                              self.block_list.data_ptr(), self.block_count.data_ptr(),
                              live=A[kind, "live"] if live else None, options=_lib.DDM_GRAD_OVERWRITE)
        st = _lib.stream_ptr()
        if kind == "F":  # scenes without a tape in their record run a plain forward substep
            _lib.check(self.lib.ddm_substep_forward_taped(ctypes.byref(batch), A["F", "cur"], A["F", "nxt"], st),
                       "ddm_substep_forward_taped")
        else:  # the adjoint substep writes the state-t gradients: nothing to clear first
            _lib.check(self.lib.ddm_substep_backward_taped(ctypes.byref(batch), A["B", "cur"], A["B", "nxt"],
                                                           A["B", "gcur"], A["B", "gnxt"], st),
                       "ddm_substep_backward_taped")

    def _launch_tick(self, key, live):
        """key: (forward scenes, forward chunks, adjoint scenes, adjoint chunks); 0 = no launch."""
        fs, fc, bs, bc = key
        A = self.tables.layout(key)[0]
        if fs:
            self._launch("F", fs, fc, live, A)
        if bs:
            self._launch("B", bs, bc, live, A)

    def _run_tick(self, fwd, bwd):
        contents = {}
        if fwd:
            contents.update(self._contents(fwd, "F"))
        if bwd:
            contents.update(self._contents(bwd, "B"))
        nf, nb = len(fwd), len(bwd)
        cf = len(contents["F", "chunks"]) if fwd else 0
        cb = len(contents["B", "chunks"]) if bwd else 0
        if self.use_graphs:
            # launch sizes of the captured tick: scene and chunk counts rounded up to powers
            # of two (one graph per key); the live counts carry the exact ones
            key = (_bucket(nf, 1) if fwd else 0, _bucket(cf, 64) if fwd else 0,
                   _bucket(nb, 1) if bwd else 0, _bucket(cb, 64) if bwd else 0)
        else:
            key = (nf, cf, nb, cb)
        if not self.tables.fits(key):
            torch.cuda.synchronize()  # the old tables may still be read by queued launches
            self.tables = _TickTables(self.device, _bucket(max(self.tables.max_scenes, nf, nb), 1),
                                      _bucket(max(self.tables.max_chunks, cf, cb), 64))
            self.graphs.clear()
        self.tables.stage(key, contents)
        if not self.use_graphs:
            self._launch_tick(key, live=False)
            return
        g = self.graphs.get(key)
        if g is None:
            # first tick of this shape: capture its launches on a side stream (capture runs
            # nothing, and the launches allocate nothing, so queued work need not drain)
            g = torch.cuda.CUDAGraph()
            with gc_paused(), torch.cuda.stream(self.capture_stream):
                g.capture_begin(capture_error_mode="thread_local")
                try:
                    self._launch_tick(key, live=True)
                finally:  # a failed launch must not leave the stream capturing
                    g.capture_end()
            self.graphs[key] = g
        g.replay()
        self.replays += 1

    def tick(self):
        t_tick = time.perf_counter()
        now = self._now()
        self._admit(now)
        fwd, bwd = [], []
        finished = []
        for a in self.active:
            if a.done_event is not None:
                continue
            op = a.next_kernel_op()
            if op is None:
                a.check_finite()
                a.done_event = torch.cuda.Event()
                a.done_event.record()
                finished.append(a)
                continue
            (fwd if op[0] == "F" else bwd).append((a, op))
        t0 = time.perf_counter()
        if fwd or bwd:
            self._run_tick(fwd, bwd)
        self.host_s["launch"] += time.perf_counter() - t0
        for a, _ in bwd:
            self.particle_substeps += a.req.n
        for a, _ in fwd + bwd:
            a.pc += 1
        if fwd or bwd:
            self.ticks += 1  # ticks that launched work
        self._collect()
        self.host_s["tick"] += time.perf_counter() - t_tick
        return bool(fwd or bwd)

    def _collect(self, block=False):
        still = []
        for a in self.active:
            if a.done_event is not None and (block or a.done_event.query()):
                a.done_event.synchronize()
                fin = self._now()
                g0 = a.g[a.final_grad_slot]
                x_final, x0_grad = a.x_final, g0.x.clone()
                if a.inv is not None:  # back to the request's particle order
                    x_final, x0_grad = x_final[a.inv], x0_grad[a.inv]
                res = Result(a.req.rid, float(a.loss_host.item()), x_final, a.pos_grad, a.rot_grad, x0_grad,
                             a.req.arrival, a.admitted, fin, bool(a.finite_host.item()))
                self.results[a.req.rid] = res
                fut = self._futures.pop(a.req.rid, None)
                if fut is not None:
                    fut.set_result(res)
                self.used_bytes -= a.bytes
                self.arena.release(a.blk_begin, a.n_blocks)
            else:
                still.append(a)
        self.active = still

    def run(self, requests: List[Request], timeout: float = 600.0) -> Dict[int, Result]:
        """Serve a stream of requests (arrival times relative to the call) until all
        have finished."""
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        self.t0 = time.perf_counter()
        while self.waiting or self.active:
            if not self.active and self.waiting and self.waiting[0].arrival > self._now():
                time.sleep(min(1e-3, self.waiting[0].arrival - self._now()))
            if not self.tick() and self.active and all(a.done_event is not None for a in self.active):
                self._collect(block=True)  # nothing left to launch: wait for the stream
            if self._now() > timeout:
                raise TimeoutError("rollout server did not drain in time")
        torch.cuda.synchronize()
        return self.results


    # ------------------------------------------------------------------ persistent loop
    def _drain_inbox(self, first=None):
        now = self._now()
        while True:
            if first is not None:
                (req, fut), first = first, None
            else:
                try:
                    req, fut = self._inbox.get_nowait()
                except queue.Empty:
                    return
            if req.rid in self._futures or any(w.rid == req.rid for w in self.waiting) or \
                    any(a.req.rid == req.rid for a in self.active):
                fut.set_exception(ValueError(f"request id {req.rid} is already being served"))
                continue
            req.arrival = now
            self.waiting.append(req)
            self._futures[req.rid] = fut

    def poll(self) -> bool:
        """One iteration of the serving loop: takes submitted requests, admits, launches one
        tick and collects finished requests (resolving their futures). Returns whether any
        request is still waiting or running."""
        if self.t0 is None:
            self.t0 = time.perf_counter()
        self._drain_inbox()
        if not self.tick() and self.active and all(a.done_event is not None for a in self.active):
            self._collect(block=True)
        return bool(self.waiting or self.active)

    def serve_forever(self, idle_wait: float = 1e-3):
        """Runs ``poll`` until ``stop`` is called; sleeps on the inbox while idle. An
        exception fails every unresolved future and ends the loop."""
        try:
            while not self._stop.is_set():
                if not self.poll():
                    try:
                        item = self._inbox.get(timeout=idle_wait)
                    except queue.Empty:
                        continue
                    self._drain_inbox(first=item)
        except BaseException as e:  # noqa: BLE001 (handed to the waiting callers)
            self._fail_all(e)
            raise

    def _fail_all(self, e):
        for fut in list(self._futures.values()):
            if not fut.done():
                fut.set_exception(e)
        self._futures.clear()
        while True:
            try:
                _, fut = self._inbox.get_nowait()
            except queue.Empty:
                break
            fut.set_exception(e)

    def start(self) -> "RolloutServer":
        """Starts the serving loop in a background thread (see submit_async)."""
        if self._thread is not None and self._thread.is_alive():
            raise RuntimeError("the serving loop is already running")
        self._stop.clear()
        self._thread = threading.Thread(target=self.serve_forever, name="rollout-server", daemon=True)
        self._thread.start()
        return self

    def stop(self, drain: bool = True, timeout: float = 600.0):
        """Stops the background loop, after the submitted requests have finished if drain."""
        if self._thread is None:
            return
        if drain:
            t_end = time.perf_counter() + timeout
            while self._thread.is_alive() and self._outstanding:
                if time.perf_counter() > t_end:
                    raise TimeoutError("rollout server did not drain in time")
                time.sleep(1e-3)
        self._stop.set()
        self._thread.join()
        self._thread = None
        torch.cuda.synchronize(self.device)


def latency_stats(results: Dict[int, Result]):
    lat = np.array([r.latency for r in results.values()]) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "n": int(len(lat))}
