"""Autograd bridge over the simulator (mirrors mpm/torch_wrapper.py:7-184).

``GradModel.forward(s, action, *past_obs)`` runs one environment step (``substeps``
substeps) as one autograd node; ``loss.backward()`` runs the hand-written adjoint
substeps in reverse and returns pose gradients that flow into the torch kinematics
(``compute_forward_kinematics``) and from there to the actions.

As in the reference, the particle-state adjoint is carried between environment steps
inside the simulator's per-state gradient buffers, and the observation gradients of
past observations are returned as zeros. Unlike the reference, observations and their
gradients never leave the device (no numpy round trips).
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from .engine import TEMP_FLOATS, gc_paused


class GradModel:
    def __init__(self, env, return_grid=(-1,), return_svd=False, graphs=True, tape=None):
        """graphs: replay each environment step's substeps (forward, and the adjoint
        substeps of backward) from a HIP graph captured on first use, instead of one
        Python call per substep; the results are the same. Graphs are keyed by the step's
        state buffers and the simulator's launch tables, and recaptured when those change.

        tape (with graphs): the forward substeps also keep the SVD temporaries and the
        grid snapshot of every substep, and the adjoint substeps restore them instead of
        recomputing P2G and the grid update (44 B per particle plus ~2 KiB per active node
        block, per substep of the simulator's max_steps). None: on when that memory is at
        most 16 GiB and a quarter of the device. The adjoint of an environment step reads
        what its own forward wrote, as it reads the states that forward wrote."""
        self.graphs = bool(graphs)
        self.tape = tape
        self._tapes = None
        self._graphs = {}
        self._graph_generation = None
        self._pose_in = None
        self._pose_grad_out = None
        self.env = env
        self.sim = env.simulator if hasattr(env, "simulator") else env
        self.dim = 3
        assert len(return_grid) == 0 or return_grid[0] == -1
        self.primitives = list(range(self.sim.n_bodies))
        self._forward_func = None
        self._set_pose_func = None
        self.return_grid = return_grid
        self.return_svd = return_svd
        self.device = self.sim.device
        self.pos_rot = None

    @property
    def substeps(self):
        return self.sim.substeps

    def zero_grad(self, return_grid=None, return_svd=None, **kwargs):
        self.sim.states[0].clear_grad()
        if return_grid is not None:
            self.return_grid = return_grid
        if return_svd is not None:
            self.return_svd = return_svd or self.return_svd

    def wrap_obs(self, obs):
        output = {"pos": obs[0][:, :3], "vel": obs[0][:, 3:6], "tool": obs[1], "dist": obs[0][:, 6:]}
        obs = obs[2:]
        if len(self.return_grid) > 0:
            ng = len(self.return_grid)
            grid, obs = obs[:ng], obs[ng:]
            output["grid"] = {k: v for k, v in zip(self.return_grid, grid)}
        assert len(obs) == 0
        return output

    def get_obs(self, s, device):
        """(particles [N, 6+nb] = x|v|dists, tools [nb, 7], grid mass) at env step s
        (mpm/torch_wrapper.py:54-77)."""
        sim = self.sim
        f = s * sim.substeps
        st = sim.states[f]
        n = sim.n_particles
        x = sim._to_caller(st.x.tensor[:n].detach().clone())
        v = sim._to_caller(st.v.tensor[:n].detach().clone())
        c = torch.cat((st.body_pos.tensor, st.body_rot.tensor), 1).detach().clone()
        outs = []
        if sim.n_bodies:
            dists = sim.get_dists(f, device=device)
            outs.append(torch.cat((x, v, dists.to(x.device)), 1))
        else:
            outs.append(torch.cat((x, v), 1))
        outs.append(c)
        for i in self.return_grid:
            outs.append(sim.compute_grid_mass(f, i, device=device))
        return tuple(o.to(device) for o in outs)

    def set_obs_grad(self, s, particle_grad, tool_grad, *args):
        """Adds observation gradients into the state-s adjoint buffers
        (mpm/torch_wrapper.py:79-105)."""
        sim = self.sim
        f = s * sim.substeps
        for idx, i in enumerate(self.return_grid):
            if args[idx] is not None:
                sim.compute_grid_mass(f, i, backward_grad=args[idx])
        nb = sim.n_bodies
        st = sim.states[f]
        n = sim.n_particles
        if particle_grad is not None:
            particle_grad = particle_grad.to(st.x.tensor.device, torch.float32)
            if nb:
                sim.get_dists(f, particle_grad[:, -nb:].contiguous())
                particle_grad = particle_grad[:, :-nb]
            pg = sim._to_internal(particle_grad.reshape(-1, self.dim * 2))
            st.x_grad.tensor[:n] += pg[:, :3]
            st.v_grad.tensor[:n] += pg[:, 3:]
        if tool_grad is not None and nb:
            tg = tool_grad.to(st.x.tensor.device, torch.float32).reshape(nb, 7)
            st.body_pos_grad.tensor += tg[:, :3]
            st.body_rot_grad.tensor += tg[:, 3:]

    # ------------------------------------------------------------------ HIP graphs
    def _graph_run(self, kind, f, fn):
        """Run fn() -- launches and device copies on fixed buffers for the environment
        step starting at substep f -- eagerly the first time (which also fills the launch
        table caches), then from a captured graph."""
        sim = self.sim
        if not self.graphs or not sim.device.type == "cuda":
            fn()
            return
        gen = sim.engine.generation
        if gen != self._graph_generation:
            # the engine rebuilt its launch record (e.g. a new particle count): graphs
            # captured before point at freed chunk tables and hold stale sizes
            self._graphs.clear()
            self._graph_generation = gen
        key = (kind, f) + tuple(id(s) for s in sim.states[f:f + self.substeps + 1])
        g = self._graphs.get(key)
        if g is not None:
            g.replay()
            return
        fn()
        g = torch.cuda.CUDAGraph()
        with gc_paused(), torch.cuda.graph(g):
            fn()
        self._graphs[key] = g

    def _tape_slots(self):
        """Per-substep tapes of the simulator's states (allocated on first use), or None."""
        if not self.graphs or self.tape is False:
            return None
        if self._tapes is None:
            from .rollout import _Tapes
            eng, sim = self.sim.engine, self.sim
            P = eng.n_particles_total
            slots = sim.max_steps - 1
            need = lambda caps: slots * (TEMP_FLOATS * P * 4 + sum(caps) * (2 * 256 * 4 + 4) + 8)  # noqa: E731
            budget = min(16 << 30, torch.cuda.get_device_properties(sim.device).total_memory // 4)
            caps = eng.tape_capacities()
            if need(caps) > budget:  # whole-grid snapshots for every substep do not fit: bounded ones
                caps = eng.tape_capacities(whole_grid=False)
            if self.tape is None:
                self.tape = need(caps) <= budget
                if not self.tape:
                    return None
            self._tapes = _Tapes(eng, slots, caps)
        return self._tapes

    def _pose_buffers(self, pos):
        S, nb = self.substeps, self.sim.n_bodies
        if self._pose_in is None or self._pose_in[0].device != pos.device:
            z = lambda *sh: torch.zeros(*sh, dtype=torch.float32, device=self.sim.device)
            self._pose_in = (z(S, nb, 3), z(S, nb, 4))
            self._pose_grad_out = (z(S, nb, 3), z(S, nb, 4))
        return self._pose_in

    def _substeps_forward(self, f, pos, rot):
        """Substeps f .. f+S-1 with the given poses. Gradient buffers: the adjoint substeps
        of backward write (DDM_GRAD_OVERWRITE) the particle gradients of states f .. f+S-1,
        so forward clears only their pose gradients (accumulated by the contact adjoint)
        and all gradients of the environment step's last state, which may be the rollout's
        end and then receives only observation gradients (the reference clears every
        gradient of every substep, mpm/simulator.py:570-571: same values)."""
        sim, S = self.sim, self.substeps
        states = sim.states[f + 1:f + S + 1]
        if not self.graphs:
            for i in range(S):
                sim.set_pose(states[i], pos[i], rot[i])
                sim.substep(f + i, clear_grad=True)
            return
        pin, rin = self._pose_buffers(pos)
        pin.copy_(pos.reshape(pin.shape))
        rin.copy_(rot.reshape(rin.shape))
        dst = [s.body_pos.tensor for s in states] + [s.body_rot.tensor for s in states]
        src = list(pin) + list(rin)
        clear = ([s.grad.body_pos for s in states[:-1]] + [s.grad.body_rot for s in states[:-1]] +
                 list(states[-1].grad.tensors()))

        tapes = self._tape_slots()
        eng = sim.engine

        def run():
            torch._foreach_copy_(dst, src)  # one launch for the environment step's poses
            torch._foreach_zero_(clear)
            for i in range(S):
                # each G2P runs inside the next substep's P2G launch; the step's last
                # substep finishes its own
                cur, nxt = eng.table_of(sim.states[f + i].buf), eng.table_of(states[i].buf)
                prev = eng.table_of(sim.states[f + i - 1].buf) if i > 0 else None
                if tapes is None:
                    eng.forward_tables(cur, nxt, prev_tab=prev, finish=i == S - 1)
                else:
                    eng.forward_taped(cur, nxt, tapes.scene_tabs[f + i], prev_tab=prev, finish=i == S - 1)
        self._graph_run("F", f, run)

    def _substeps_backward(self, f):
        """Adjoint substeps f+S-1 .. f; returns the pose gradients of states f+1 .. f+S."""
        sim, S = self.sim, self.substeps
        if not self.graphs:
            pos, rot = [], []
            for i in range(f + S - 1, f - 1, -1):
                sim.substep_grad(i)
                st = sim.states[i + 1]
                pos.append(st.body_pos_grad.tensor.clone())
                rot.append(st.body_rot_grad.tensor.clone())
            return torch.stack(pos[::-1]), torch.stack(rot[::-1])
        self._pose_buffers(sim.states[f].body_pos.tensor)
        pg, rg = self._pose_grad_out
        states = sim.states[f + 1:f + S + 1]
        dst = list(pg) + list(rg)
        src = [s.body_pos_grad.tensor for s in states] + [s.body_rot_grad.tensor for s in states]

        tapes = self._tape_slots()
        eng = sim.engine

        def run():
            restored = False
            for i in range(f + S - 1, f - 1, -1):
                cur, nxt = sim.states[i], sim.states[i + 1]
                if tapes is None:
                    eng.backward(cur.buf, nxt.buf, cur.grad, nxt.grad, overwrite=True)
                else:
                    # each substep's P2G adjoint launch also restores the step's substep before
                    prev = tapes.scene_tabs[i - 1] if i > f and not eng.deterministic else None
                    eng.backward_taped(eng.table_of(cur.buf), eng.table_of(nxt.buf), eng.table_of(cur.grad),
                                       eng.table_of(nxt.grad), tapes.scene_tabs[i], overwrite=True,
                                       complete=tapes.complete, restored=restored, prev_scene_tab=prev)
                    restored = prev is not None
            # the pose gradients of states f+1 .. f+S are complete once the step's adjoint
            # substeps have run (later adjoint substeps touch state f and earlier only)
            torch._foreach_copy_(dst, src)
        self._graph_run("B", f, run)
        return pg.clone(), rg.clone()

    @property
    def diff_forward(self):
        if self._forward_func is None:
            model = self

            class Forward(Function):
                @staticmethod
                def forward(ctx, s, pos, rot, *past_obs):
                    ctx.s = s
                    ctx.save_for_backward(*[torch.zeros_like(i) for i in past_obs])
                    f = s * model.substeps
                    model._substeps_forward(f, pos, rot)
                    return model.get_obs(s + 1, pos.device)

                @staticmethod
                def backward(ctx, *obs_grad):
                    zero_grads = ctx.saved_tensors
                    s = ctx.s
                    model.set_obs_grad(s + 1, *obs_grad)
                    f = s * model.substeps
                    pos_grad, rot_grad = model._substeps_backward(f)
                    pos_grad, rot_grad = pos_grad.to(model.device), rot_grad.to(model.device)
                    return (None, pos_grad, rot_grad) + tuple(zero_grads)

            self._forward_func = Forward
        return self._forward_func.apply

    @property
    def diff_set_pose(self):
        if self._set_pose_func is None:
            model = self

            class SetPose(Function):
                @staticmethod
                def forward(ctx, s, pos, rot):
                    ctx.s = s
                    model.sim.set_pose(model.sim.states[s * model.substeps], pos, rot)
                    return model.get_obs(s, pos.device)

                @staticmethod
                def backward(ctx, *obs_grad):
                    s = ctx.s
                    model.set_obs_grad(s, *obs_grad)
                    st = model.sim.states[s * model.substeps]
                    pos_grad = st.body_pos_grad.tensor.clone()
                    rot_grad = st.body_rot_grad.tensor.clone()
                    st.body_pos_grad.zero()
                    st.body_rot_grad.zero()
                    return (None, pos_grad, rot_grad)

            self._set_pose_func = SetPose.apply
        return self._set_pose_func

    def forward(self, s, action, *past_obs, pos_rot=None):
        if s == 0 and pos_rot is None:
            self.pos_rot = self.sim.download_pos_rot(0, action.device)
        if pos_rot is not None:
            assert isinstance(pos_rot[0], torch.Tensor)
            past_obs = self.diff_set_pose(s, *pos_rot)
        pos, rot, q_state = self.sim.compute_forward_kinematics(s * self.substeps, action, pos_rot=self.pos_rot)
        self.pos_rot = q_state
        return self.diff_forward(s, pos, rot, *past_obs)
