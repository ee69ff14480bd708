"""Device-resident differentiable rollouts with substep checkpointing.

A :class:`Rollout` runs ``T`` forward substeps of every scene of an :class:`Engine`
and then the adjoint substeps in reverse, producing the gradients of a loss on the
final state with respect to every substep's body poses (and the initial state).

Memory: with ``checkpoint_every = K`` only every K-th state is kept during the
forward pass (the reference keeps all of them, mpm/simulator.py:205-206); the backward
pass recomputes one K-substep segment at a time from its checkpoint into a K+1 state
buffer and runs the adjoint through it. States then cost (T/K + K + 1) instead of T+1
copies, at the price of one extra forward substep per substep.

A whole forward + backward pass can be captured once into a HIP graph
(:meth:`Rollout.capture`) and replayed (:meth:`Rollout.replay`): every buffer, table and
launch of a rollout is fixed once the engine and horizon are, so a replay re-runs the
same kernels on whatever the initial state, poses and loss seed hold at that time.

Body poses of all substeps live in ``pose_pos [T+1, NB, 3]`` / ``pose_rot [T+1, NB, 4]``;
the states' pose pointers alias those rows, so setting a trajectory's poses is one
device copy and nothing crosses the host per substep. Pose gradients accumulate into
``pos_grad`` / ``rot_grad`` with the same shapes.
"""
from __future__ import annotations

from typing import List, Optional

import torch

from .. import _lib
from .engine import TEMP_FLOATS, Engine, gc_paused


class _Particles:
    """x, v, F, C of one particle state in one contiguous allocation, so that a clear or
    a copy of the whole state is a single kernel."""

    def __init__(self, n, device, identity=True):
        self.flat = torch.zeros(24 * n, dtype=torch.float32, device=device)
        self.x = self.flat[0:3 * n].view(n, 3)
        self.v = self.flat[3 * n:6 * n].view(n, 3)
        self.F = self.flat[6 * n:15 * n].view(n, 3, 3)
        self.C = self.flat[15 * n:24 * n].view(n, 3, 3)
        if identity and n:
            self.F.diagonal(dim1=1, dim2=2).fill_(1.0)

    def tensors(self):
        return (self.x, self.v, self.F, self.C)

    def copy_(self, o):
        self.flat.copy_(o.flat)

    def zero_(self):
        self.flat.zero_()


class _Tapes:
    """Per-substep tapes of one rollout segment: the SVD temporaries (30 floats per
    particle) and, per scene, a snapshot of its listed grid blocks (up to its capacity;
    beyond it the adjoint recomputes that scene's substep), reached through a scene table
    per slot."""

    def __init__(self, engine: Engine, slots: int, caps):
        dev = engine.device
        P = engine.n_particles_total
        self.caps = [int(c) for c in caps]
        self.complete = engine.tape_complete(self.caps)  # no substep can overflow its snapshot
        nblk = sum(self.caps)
        self.temps = torch.empty(slots, TEMP_FLOATS * P, dtype=torch.float32, device=dev)
        self.grid_mmv = torch.empty(slots, nblk * 256, dtype=torch.float32, device=dev)
        self.grid_v = torch.empty_like(self.grid_mmv)
        self.block_list = torch.empty(slots, nblk, dtype=torch.int32, device=dev)
        self.status = torch.zeros(slots, len(self.caps), 2, dtype=torch.int32, device=dev)
        self.scene_tabs = [engine.taped_table(self.temps[i], self.grid_mmv[i], self.grid_v[i], self.block_list[i],
                                              self.status[i], self.caps) for i in range(slots)]

    def nbytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in (self.temps, self.grid_mmv, self.grid_v, self.block_list,
                                                          self.status))


class Rollout:
    def __init__(self, engine: Engine, T: int, checkpoint_every: Optional[int] = None, tape: bool = False,
                 tape_blocks: Optional[int] = None):
        """tape=True keeps, for every substep of a segment, the SVD temporaries and the
        grid of the forward pass, so the adjoint substeps skip recomputing P2G and the
        grid update (44 B per particle plus 2 KiB per active grid block per substep).
        tape_blocks bounds the blocks one scene's snapshot holds per substep (an int for
        every scene or a list per scene; default: every scene's whole grid when the tape's
        snapshots then take at most a quarter of the device's memory, else
        Engine.tape_capacities); a scene whose substep lists more is recomputed by its
        adjoint instead (decided on the device)."""
        self.engine = engine
        self.T = int(T)
        dev = engine.device
        P, NB = engine.n_particles_total, engine.n_bodies_total
        self.K = int(checkpoint_every) if checkpoint_every else self.T
        if self.K <= 0 or self.T % self.K:
            raise ValueError("checkpoint_every must divide the horizon")
        self.full = self.K == self.T
        self._last_segment_taped = False  # forward() ran the last segment taped
        nb_rows = max(NB, 1)
        self.pose_pos = torch.zeros(self.T + 1, nb_rows, 3, device=dev)
        self.pose_rot = torch.zeros(self.T + 1, nb_rows, 4, device=dev)
        self.pose_rot[..., 0] = 1.0
        self.pos_grad = torch.zeros_like(self.pose_pos)
        self.rot_grad = torch.zeros_like(self.pose_rot)
        # checkpoints at t = 0, K, 2K, ... (all states when K == T)
        n_ckpt = self.T // self.K + 1
        self.ckpt = [_Particles(P, dev) for _ in range(n_ckpt)] if not self.full else []
        # segment buffer: K+1 states (holds the whole trajectory when not checkpointing)
        self.seg = [_Particles(P, dev) for _ in range(self.K + 1)]
        self.g = [_Particles(P, dev, identity=False) for _ in range(2)]
        pp, pr, gp, gr = self.pose_pos, self.pose_rot, self.pos_grad, self.rot_grad

        # per-scene state tables: segment slot j with the poses of absolute substep t
        def sview(j, t):
            s = self.seg[j]
            return engine.table(s.x, s.v, s.F, s.C, pp[t], pr[t])

        def gview(j, t):
            s = self.g[j]
            return engine.table(s.x, s.v, s.F, s.C, gp[t], gr[t])

        self._sview, self._gview = sview, gview
        self.tape = None
        if tape:
            if tape_blocks is None:
                # whole grids (no substep can overflow, so the adjoint launches no recompute
                # kernels) when the K slots of them take at most a quarter of the device
                whole = engine.tape_capacities(whole_grid=True)
                budget = torch.cuda.get_device_properties(engine.device).total_memory // 4
                fits = self.K * sum(whole) * engine.TAPE_BLOCK_BYTES <= budget
                caps = whole if fits else engine.tape_capacities()
            elif isinstance(tape_blocks, (list, tuple)):
                caps = [int(c) for c in tape_blocks]
            else:
                caps = [int(tape_blocks)] * len(engine.scenes)
            self.tape = _Tapes(engine, self.K, caps)
        self.tape_complete = self.tape is not None and self.tape.complete

    # ------------------------------------------------------------------ inputs / outputs
    def set_initial(self, x, v=None, F=None, C=None):
        s0 = self.seg[0]
        s0.x.copy_(x)
        s0.v.copy_(v) if v is not None else s0.v.zero_()
        if F is not None:
            s0.F.copy_(F)
        else:
            s0.F[:] = torch.eye(3, device=s0.F.device)
        s0.C.copy_(C) if C is not None else s0.C.zero_()
        if not self.full:
            self.ckpt[0].copy_(s0)

    def set_poses(self, pos, rot):
        """pos [T+1, NB, 3], rot [T+1, NB, 4] (row t = poses of substep t)."""
        if self.engine.n_bodies_total:
            self.pose_pos.copy_(pos)
            self.pose_rot.copy_(rot)

    def final_state(self) -> _Particles:
        return self.seg[self.K] if self.full else self._final

    # ------------------------------------------------------------------ passes
    def _fwd(self, j, t, taped=False, ahead=False, finish=True):
        """Forward substep t from segment state j. ahead: the G2P of substep t-1 (state j-1
        -> j) was left to this call; finish=False leaves this substep's G2P to the next."""
        prev = self._sview(j - 1, t - 1) if ahead else None
        if taped:
            self.engine.forward_taped(self._sview(j, t), self._sview(j + 1, t + 1), self.tape.scene_tabs[j],
                                      prev_tab=prev, finish=finish)
        else:
            self.engine.forward_tables(self._sview(j, t), self._sview(j + 1, t + 1), prev_tab=prev, finish=finish)

    def forward(self):
        """T forward substeps from the initial state."""
        # within a run of substeps each G2P runs inside the next substep's P2G launch
        # (ahead=True there, finish=False here); the run's last substep finishes its own
        if self.full:
            for t in range(self.T):
                self._fwd(t, t, taped=self.tape is not None, ahead=t > 0, finish=t == self.T - 1)
            return self.seg[self.T]
        # checkpointed: run each segment in the segment buffer and keep its end state;
        # segment 0 restarts from checkpoint 0 too, since a previous pass left a later
        # segment's start in the buffer. The last segment runs taped when there is a tape:
        # the backward pass starts with it and need not run it again.
        n_seg = self.T // self.K
        for c in range(n_seg):
            self.seg[0].copy_(self.ckpt[c])
            last = c == n_seg - 1 and self.tape is not None
            for j in range(self.K):
                self._fwd(j, c * self.K + j, taped=last, ahead=j > 0, finish=j == self.K - 1)
            self.ckpt[c + 1].copy_(self.seg[self.K])
        self._last_segment_taped = self.tape is not None
        self._final = self.ckpt[-1]
        return self._final

    def backward(self, x_grad=None, v_grad=None, F_grad=None, C_grad=None):
        """Adjoint of the whole rollout for the given gradient of the final state.
        Afterwards ``pos_grad``/``rot_grad`` hold d loss / d poses of every substep and
        ``initial_grad()`` the gradient of the initial particle state."""
        eng = self.engine
        self.pos_grad.zero_()
        self.rot_grad.zero_()
        gn = self.g[0]
        for a, b in zip(gn.tensors(), (x_grad, v_grad, F_grad, C_grad)):
            a.copy_(b) if b is not None else a.zero_()
        cur_slot = 0
        n_seg = self.T // self.K
        for c in range(n_seg - 1, -1, -1):
            base = c * self.K
            # the segment buffer and the tape still hold the last segment from forward()
            # until an earlier segment is run again into them
            if not self.full and not (c == n_seg - 1 and self._last_segment_taped):
                self.seg[0].copy_(self.ckpt[c])
                for j in range(self.K):
                    self._fwd(j, base + j, taped=self.tape is not None, ahead=j > 0, finish=j == self.K - 1)
                self._last_segment_taped = False
            restored = False
            for j in range(self.K - 1, -1, -1):
                t = base + j
                nxt_slot = cur_slot
                cur_slot = 1 - cur_slot
                tabs = (self._sview(j, t), self._sview(j + 1, t + 1), self._gview(cur_slot, t),
                        self._gview(nxt_slot, t + 1))
                # the adjoint substep writes (not adds to) the state-t gradient: no clearing
                if self.tape is not None:
                    # within a segment, each substep's P2G adjoint launch also restores the
                    # snapshot of the substep before (whose call then skips its restore)
                    prev = self.tape.scene_tabs[j - 1] if j > 0 and not eng.deterministic else None
                    eng.backward_taped(*tabs, self.tape.scene_tabs[j], overwrite=True, complete=self.tape_complete,
                                       restored=restored, prev_scene_tab=prev)
                    restored = prev is not None
                else:
                    eng.backward_tables(*tabs, overwrite=True)
        self._g0 = self.g[cur_slot]
        return self._g0

    def initial_grad(self) -> _Particles:
        return self._g0

    # ------------------------------------------------------------------ HIP graph
    def capture(self, x_grad: torch.Tensor):
        """Capture forward() + backward(x_grad=x_grad) into a HIP graph. x_grad is kept
        and read at every replay, so the loss seed may change between replays; so may
        the initial state and the poses. The graph holds only kernel launches and
        device copies (grid and gradient buffers are cleared by kernels)."""
        self._seed = x_grad
        side = torch.cuda.Stream()  # default priority
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):  # populates the table caches: no uploads during capture
            self.forward()
            self.backward(x_grad=x_grad)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with gc_paused(), torch.cuda.graph(g):
            self.forward()
            self.backward(x_grad=x_grad)
        self._graph = g
        self._graph_generation = self.engine.generation
        return g

    def replay(self):
        """One captured forward + backward pass (see capture)."""
        if getattr(self, "_graph", None) is None:
            raise RuntimeError("no captured graph: call capture() first")
        if self._graph_generation != self.engine.generation:
            raise RuntimeError("the engine's launch tables changed since capture (set_counts or a new batch): "
                               "capture again")
        self._graph.replay()

    def state_bytes(self) -> int:
        per = self.engine.n_particles_total * (3 + 3 + 9 + 9) * 4
        return per * (len(self.ckpt) + len(self.seg) + len(self.g)) + (self.tape.nbytes() if self.tape else 0)

    def tape_recomputed(self) -> int:
        """(Scene, substep) pairs of the last backward segment that had to recompute
        because the scene's snapshot was full."""
        return int(self.tape.status[:, :, 1].sum().item()) if self.tape is not None else 0
