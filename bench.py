"""Benchmark: differentiable MLS-MPM substep (forward + adjoint) on one MI355X.

Metric: particle-substeps/s, forward+backward (one particle advanced one substep and
its adjoint substep), batched uneven scenes vs the same scenes run one by one through
the single-scene path.

Workload (default, "batched"): 64 scenes, particle counts U[5k, 60k], 64^3 grid each,
yield stress U[20, 100], 2-22 capsules per scene with constant random velocities,
100 substeps forward then 100 adjoint substeps, all scenes in one launch per stage,
the whole rollout captured once into a HIP graph and replayed. The adjoint reads the
forward pass's tape (SVD temporaries and grid snapshots) instead of recomputing P2G and
the grid update; the recomputing adjoint (the reference's scheme) is measured too. A step = one such
rollout (forward + backward) of all 64 scenes. Also reported: the same rollout with
eager launches, and the scenes run one at a time (plain; eager and graph-replayed).

Prints one JSON line (rank 0). --gpus N > 1 is refused: the benchmark runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle per kernel launch (state, material and temporaries the
# kernel must read or write; grid traffic excluded), itemised per entry.
KERNEL_BYTES = {
    "p2g_fwd": 152,      # read x,v,F,C + mass,vol,mu/lam/yield (116) ; write F' (36)
    "grid_fwd": 0,
    "g2p_fwd": 72,       # read x (12) ; write x',v',C' (60)
    "p2g_recompute": 160,  # read x,v,F,C + material (116) ; write U,V (quaternions), sigma (44)
    "grid_recompute": 0,
    "g2p_adj": 108,      # read x,v',dx',dv',dC' (84) ; read+write dx (24)
    "grid_adj": 0,
    "p2g_adj": 388,      # read x,v,F,C,material,U,V,sigma,dF' (196) ; read+write dx,dv,dF,dC (192)
}
# HBM bytes per particle per launch measured on the default workload (64 scenes, seed 7)
# with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads
# for a 256 MiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0). Re-measured whenever a
# kernel changes; reported as roofline.traffic for the default workload only.
TRAFFIC_BYTES = {
    "p2g_fwd": 220, "grid_fwd": 10, "g2p_fwd": 75, "p2g_recompute": 229, "grid_recompute": 7, "g2p_adj": 148,
    "grid_adj": 18, "p2g_adj": 332,
    # taped rollout
    "taped_p2g_fwd": 266, "taped_grid_fwd": 17, "taped_g2p_fwd": 75, "taped_restore": 13, "taped_g2p_adj": 148,
    "taped_grid_adj": 18, "taped_p2g_adj": 332,
}
FWD_NAMES = ["p2g_fwd", "grid_fwd", "g2p_fwd"]
BWD_NAMES = ["p2g_recompute", "grid_recompute", "g2p_adj", "grid_adj", "p2g_adj"]
# the taped rollout: its forward P2G also writes the temporaries (44 B per particle) and
# its adjoint restores the grid snapshot (no recompute unless the snapshot overflowed).
# The tape is an implementation choice, not part of the path's algorithmic traffic: like
# checkpoint recompute it is overhead, so the taped P2G is credited the same 152 B.
KERNEL_BYTES_TAPED = dict(KERNEL_BYTES, restore=0)
# VALU issue: each of a CU's 4 SIMDs issues one wave64 VALU instruction per 4 cycles, at
# the 2.4 GHz peak engine clock on 256 CUs (vector FP64 issues at the same rate: 78.6
# TFLOP/s is 16 FMA lanes per SIMD per cycle)
VALU_PEAK_WAVE_INSTR_S = 256 * 4 * 2.4e9 / 4
# VALU wave-instructions per wave (64 particles) and the VALUBusy fraction per kernel,
# from rocprofv3 SQ_INSTS_VALU / SQ_WAVES and SQ_ACTIVE_INST_VALU on the default workload
VALU_INSTR_PER_WAVE = {"p2g_fwd": 4047, "g2p_fwd": 1259, "g2p_adj": 2959, "p2g_adj": 2147}
VALU_BUSY = {"p2g_fwd": 0.60, "g2p_fwd": 0.92, "g2p_adj": 0.52, "p2g_adj": 0.58}
FWD_NAMES_TAPED = ["p2g_fwd", "grid_fwd", "g2p_fwd"]
BWD_NAMES_TAPED = ["restore", "g2p_adj", "grid_adj", "p2g_adj"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_batched(n_scenes, seed, T):
    from dexdeform_amd.mpm import workloads as W
    scenes = W.batched_scenes(n_scenes=n_scenes, seed=seed)
    return scenes, W.constant_velocity_poses(scenes, T, seed + 1)


def build_rollout(scenes, poses, T, device, checkpoint_every=None, tape=False, deterministic=False, gx=None):
    import torch
    from dexdeform_amd.mpm.engine import Engine
    from dexdeform_amd.mpm.rollout import Rollout
    eng = Engine([s.spec for s in scenes], device=device, with_grad=True, deterministic=deterministic)
    ro = Rollout(eng, T, checkpoint_every=checkpoint_every, tape=tape)
    x0 = torch.tensor(np.concatenate([s.x0 for s in scenes]), device=device)
    ro.set_initial(x0)
    if eng.n_bodies_total:
        ro.set_poses(torch.tensor(np.concatenate([p[0] for p in poses], 1), device=device),
                     torch.tensor(np.concatenate([p[1] for p in poses], 1), device=device))
    if gx is None:  # loss seed dL/dx_T
        rng = np.random.Generator(np.random.PCG64(11))
        gx = torch.tensor(rng.normal(size=(eng.n_particles_total, 3)).astype(np.float32) / eng.n_particles_total,
                          device=device)
    return eng, ro, gx


def serve_benchmark(dev, rates, seconds, seed, graphs=True):
    """Served rollouts: for each Poisson arrival rate, a stream of uneven requests
    (particles U[5k,100k], grid 48^3/64^3/96^3, horizon U[20,200] with checkpoints every
    10 substeps, forward + backward each) through the continuous-batching server; then
    200 such requests arriving at once ("burst": throughput with the server saturated).
    Latency = completion - arrival; throughput = particle-substeps (forward + adjoint
    counted once) per second of wall time from the first arrival to the last completion."""
    import torch
    from dexdeform_amd.mpm import workloads as W
    from dexdeform_amd.mpm.serve import Request, RolloutServer, latency_stats
    out = []
    # the rates, then a burst of 200 requests arriving at once (the server saturated)
    for k, rate in enumerate(list(rates) + ["burst"]):
        burst = rate == "burst"
        n_requests = 200 if burst else max(24, int(round(rate * seconds)))
        stream = W.served_stream(200.0 if burst else rate, n_requests, seed=seed + 1000 * k)
        reqs = [Request(i, d["scene"].spec, d["scene"].x0, d["pose_pos"], d["pose_rot"], d["target"],
                        arrival=0.0 if burst else d["arrival"], checkpoint_every=d["checkpoint_every"])
                for i, d in enumerate(stream)]
        srv = RolloutServer(device=dev, use_graphs=graphs)
        t0 = time.perf_counter()
        res = srv.run(reqs)
        wall = time.perf_counter() - t0
        st = latency_stats(res)
        ps = sum(r.n * r.T for r in reqs)
        out.append({"rate_req_s": rate, "requests": n_requests, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                    "throughput": ps / wall, "wall_s": wall, "ticks": srv.ticks, "graph_replays": srv.replays,
                    "graphs_captured": len(srv.graphs),
                    "host_s": {k: round(v, 4) for k, v in srv.host_s.items()}})
        log(f"serve rate {rate}{'' if burst else '/s'}: p50 {st['p50_ms']:.1f} ms p99 {st['p99_ms']:.1f} ms "
            f"throughput {ps / wall:.3e} particle-substeps/s")
        del srv, res
        torch.cuda.empty_cache()
    return out


def named_rollouts(dev, steps, large=True, only=None):
    """The single-scene workloads, each one forward + backward rollout through the
    single-scene path (taped adjoint): the plain 10k-particle scene (50 substeps, two
    capsules with constant velocities), the 50k-particle hand-contact scene (22 links,
    200 substeps, checkpoints every 10, seeded joint-velocity actions) and the
    1M-particle 128^3 scene (500 substeps, checkpoints every 25). The first two are
    replayed from HIP graphs. Reports particle-substeps/s (forward + adjoint counted
    once) and ms per rollout."""
    import torch
    from dexdeform_amd.mpm import workloads as W
    from dexdeform_amd.mpm.engine import Engine
    from dexdeform_amd.mpm.rollout import Rollout
    out = {}

    def run(name, sc, T, K, poses, graph):
        eng = Engine([sc.spec], device=dev)
        ro = Rollout(eng, T, checkpoint_every=K, tape=True)
        ro.set_initial(torch.tensor(sc.x0, device=dev))
        ro.set_poses(*poses)
        target = torch.tensor(W.plain_target(sc.x0), device=dev)
        if graph:
            d = ro.forward().x - target
            seed = (2.0 / sc.n) * d
            ro.capture(seed)
            step = ro.replay
        else:
            def step():
                d = ro.forward().x - target
                ro.backward(x_grad=(2.0 / sc.n) * d)
        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
        out[name] = {"particles": sc.n, "substeps": T, "checkpoint_every": K, "graph": graph,
                     "ms_per_rollout": sec * 1e3, "particle_substeps_per_s": sc.n * T / sec}
        log(f"{name}: {sc.n * T / sec:.3e} particle-substeps/s, {sec * 1e3:.2f} ms per forward + backward rollout")
        del ro, eng
        torch.cuda.empty_cache()

    want = lambda name: only is None or name in only  # noqa: E731
    if want("plain_10k"):
        sc = W.plain_scene(n_particles=10000)
        T = 50
        vel = np.random.Generator(np.random.PCG64(1)).uniform(-0.5, 0.5, size=(sc.nb, 3)).astype(np.float32)
        t = np.arange(T + 1, dtype=np.float32)[:, None, None] * sc.spec.dt
        run("plain_10k", sc, T, None, (torch.tensor(sc.body_pos[None] + t * vel[None], device=dev),
                                        torch.tensor(np.repeat(sc.body_rot[None], T + 1, 0), device=dev)), True)
    if want("hand_50k"):
        sc, tree, q0 = W.hand_scene(50000)
        P, R = tree.to(device=dev).trajectory(q0.to(dev), torch.tensor(W.hand_actions(20), device=dev), 10)
        run("hand_50k", sc, 200, 10, (P.detach(), R.detach()), True)
    if large and want("large_1m"):
        sc, tree, q0 = W.large_scene(1_000_000)
        P, R = tree.to(device=dev).trajectory(q0.to(dev), torch.tensor(W.hand_actions(50), device=dev), 10)
        run("large_1m", sc, 500, 25, (P.detach(), R.detach()), False)
    return out


def trajectory_optimisation(dev, iters=2):
    """One trajectory-optimisation iteration through the drop-in MPMSimulator + GradModel
    (eager, one Python call per substep, as a reference user runs it): the lift-box
    configuration -- 10,000 particles sampled in random order in a 0.09 box, dx 1/64,
    dt 5e-5, 40 substeps per env step, 50 env steps, gravity (0,-2,0), ground friction
    0.3 -- with 19 capsule primitives driven by per-step velocity actions; L2 loss of the
    final positions to a target, backward, one Adam step. Seconds per iteration."""
    import torch
    from dexdeform_amd.mpm import workloads as W
    from dexdeform_amd.mpm.simulator import MPMSimulator
    from dexdeform_amd.mpm.torch_wrapper import GradModel
    n, nb, S, E = 10000, 19, 40, 50
    rng = np.random.Generator(np.random.PCG64(21))
    x0 = rng.uniform([0.445, 0.175, 0.405], [0.535, 0.265, 0.495], size=(n, 3)).astype(np.float32)
    pos, rot, args = W.random_capsules(rng, nb, [0.49, 0.31, 0.45], spread=0.05, r_range=(0.008, 0.015),
                                       h_range=(0.02, 0.05))
    sim = MPMSimulator(n_bodies=nb, n_particles=n, dx=1 / 64, dt=5e-5, substeps=S, max_steps=E * S + 1,
                       gravity=(0., -2., 0.), ground_friction=0.3, yield_stress=50., E=5e3, nu=0.2,
                       vol=(1 / 128) ** 2, mass=(1 / 128) ** 2, device=dev)
    sim.init_bodies(types=[W.CAPSULE] * nb, softness=[666.] * nb, mu=[0.9] * nb, round=[0.] * nb, args=args,
                    action_scales=np.tile(np.float32([S * 5e-5] * 3 + [0.0] * 3), (nb, 1)))
    state = (x0, np.zeros_like(x0), np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)), np.zeros((n, 3, 3), np.float32))
    state += tuple(np.r_[p, r] for p, r in zip(pos, rot))
    target = torch.tensor(x0 + np.float32([0, 0.05, 0]), device=dev)
    acts = torch.zeros(E, nb * 6, device=dev, requires_grad=True)
    with torch.no_grad():
        acts[:, 1::6] = -0.3  # press down
    opt = torch.optim.Adam([acts], lr=0.01)
    model = GradModel(sim, return_grid=())

    def iteration():
        sim.set_state(0, state)
        model.zero_grad()
        opt.zero_grad()
        obs = model.get_obs(0, dev)
        for e in range(E):
            obs = model.forward(e, acts[e], *obs)
        loss = ((obs[0][:, :3] - target) ** 2).sum(-1).mean()
        loss.backward()
        opt.step()
        return float(loss.item())

    iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        loss = iteration()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / iters
    log(f"trajectory optimisation (drop-in simulator, lift-box shape): {sec:.3f} s per iteration, loss {loss:.5f}")
    return {"seconds_per_iteration": sec, "particle_substeps_per_s": n * E * S / sec, "substeps": E * S,
            "particles": n, "bodies": nb}


def score_benchmark(dev, n=100000, m=200000, reps=3):
    """The shape score of tutorials/4_computing_score.ipynb cell 3 on the device: debiased
    Sinkhorn divergence (p = 1, blur 1e-4) between 100,000 and 200,000 Gaussian points, value
    and gradient w.r.t. the first cloud; and the Chamfer distance of the same clouds. Reports
    seconds per call and cost-pair evaluations per second (each softmin over an n x m pair
    set evaluates n m pairs; the Sinkhorn loop runs 4 softmins per epsilon value plus the
    initial and final steps, the gradient 4 plan reductions)."""
    import torch
    from dexdeform_amd.mpm import losses as L
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.randn(n, 3, generator=g).to(dev).requires_grad_(True)
    y = torch.randn(m, 3, generator=g).to(dev)
    loss = L.SamplesLoss("sinkhorn", p=1, blur=1e-4)
    n_eps = len(L.epsilon_schedule(1, L._diameter(x.detach(), y), 1e-4, 0.5))
    sweep = 2 * n * m + n * n + m * m  # pairs of one epsilon's four softmins
    pairs_fwd = sweep * (n_eps + 2)
    pairs_bwd = 2 * n * m + 2 * n * n
    s = loss(x, y)
    s.backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        x.grad = None
        s = loss(x, y)
        torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        x.grad = None
        s = loss(x, y)
        s.backward()
        torch.cuda.synchronize()
    t_all = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        c = L.chamfer_distance(x, y)
        torch.cuda.synchronize()
    t_ch = (time.perf_counter() - t0) / reps
    out = {"points": [n, m], "p": 1, "blur": 1e-4, "epsilon_steps": n_eps, "sinkhorn": float(s.item()),
           "seconds_value": t_fwd, "seconds_value_and_grad": t_all,
           "pairs_per_s_value": pairs_fwd / t_fwd, "pairs_per_s_value_and_grad": (pairs_fwd + pairs_bwd) / t_all,
           "chamfer": float(c.item()), "chamfer_seconds": t_ch, "chamfer_pairs_per_s": 2 * n * m / t_ch}
    log(f"score: sinkhorn {n}x{m} {t_fwd:.3f} s value, {t_all:.3f} s with gradient "
        f"({out['pairs_per_s_value']:.3e} pairs/s); chamfer {t_ch * 1e3:.1f} ms")
    return out


def render_benchmark(dev, steps=20):
    """Random-action stepping with a render per step, the loop of
    tutorials/2_environment_loading.ipynb cell 5 (reference: 11.52 it/s on its GPU, 100 steps
    of the folding scene: 10,000 particles, 40 substeps per step, a 128 x 128 image with
    spp 10 per step through Viewer.render): the drop-in HandSimulator with the stored
    Shadow hand (19 geoms, scale 1.5), a flat cylinder of 10,000 particles (r 0.16, h 0.01,
    the folding shape), E 4000, yield stress 250, ground friction 500, gravity (0, -10, 0);
    per step one random action, 40 forward substeps, bake + render (voxel box 168^3, two
    bounces) and the uint8 image on the host. Also the render alone at 128^2 spp 10 and at
    the Renderer's defaults (512^2, spp 3), and the bake alone."""
    import os
    import torch
    from dexdeform_amd.mpm import hand as HM
    from dexdeform_amd.mpm.renderer import Renderer
    cfg_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "shadow_rh_scale1p5.npz")
    prims, hand_cfg = HM.load_hand_config(cfg_path)
    nb, kw = HM.hand_bodies(prims)
    n = 10000
    sim = HM.HandSimulator(nb, hand_cfg, cfg=dict(n_particles=n, E=4e3, nu=0.2, yield_stress=250.0,
                                                  ground_friction=500.0, quality=1, max_steps=steps * 40 + 81,
                                                  gravity=(0.0, -10.0, 0.0), fixed_base=True), device=dev)
    sim.init_bodies(**kw)
    rng = np.random.Generator(np.random.PCG64(5))
    r = 0.16 * np.sqrt(rng.uniform(0, 1, n))
    a = rng.uniform(0, 2 * np.pi, n)
    x0 = np.stack([0.5 + r * np.cos(a), 0.056875 + rng.uniform(0, 0.01, n), 0.6 + r * np.sin(a)], 1)
    base = HM.root_matrix([0.5, 0.2, 0.3], [0.0, 0.0, 0.0])
    sim.set_state(0, (x0.astype(np.float32), np.zeros((n, 3), np.float32),
                      np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)), np.zeros((n, 3, 3), np.float32), base[None],
                      np.zeros((1, 24))))
    sim.set_color(np.zeros(n, np.int64) + 0xFFFFFF)
    side = dict(center=(1.2, 0.35, 0.6), theta=np.pi / 8, phi=-np.pi / 2, radius=1.5)  # Viewer's side view
    small = Renderer(image_res=(128, 128), device=dev)
    small.lookat(**side)
    acts = torch.tensor(rng.uniform(0, 1, (steps + 2, 1, 26)), dtype=torch.float32, device=dev)

    def step(i):
        sim.step(acts[i])
        return np.uint8(small.render(sim, sim.states[0], spp=10)[..., :3].clip(0, 1) * 255)

    step(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        img = step(i + 1)
    torch.cuda.synchronize()
    it_s = steps / (time.perf_counter() - t0)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    big = Renderer(device=dev)
    big.lookat(**side)
    out = {"scene": "stored Shadow hand (19 geoms) + 10k-particle cylinder, 40 substeps per step",
           "step_and_render_it_per_s": it_s, "reference_it_per_s": 11.52,
           "render_128_spp10_ms": 1e3 * timed(lambda: small.render(sim, sim.states[0], to_numpy=False, spp=10), 10),
           "render_512_spp3_ms": 1e3 * timed(lambda: big.render(sim, sim.states[0], to_numpy=False), 5),
           "bake_168_ms": 1e3 * timed(lambda: small.bake_particles(sim, sim.states[0]), 10),
           "image_mean": float(img.mean())}
    log(f"render: step + 128^2 spp-10 render {it_s:.1f} it/s (reference 11.52); render alone "
        f"{out['render_128_spp10_ms']:.2f} ms (128^2 spp 10), {out['render_512_spp3_ms']:.2f} ms (512^2 spp 3), "
        f"bake {out['bake_168_ms']:.2f} ms")
    return out


def run_step(ro, gx):
    ro.forward()
    ro.backward(x_grad=gx)


class KernelTimer:
    """Per-kernel durations from hipEvents recorded between the stage kernels of one
    rollout, on the stream the kernels run on."""

    def __init__(self, ro):
        import torch
        self.ro = ro
        self.torch = torch

    def timed_step(self, gx):
        import ctypes
        from dexdeform_amd import _lib
        torch = self.torch
        ro = self.ro
        eng = ro.engine
        lib = eng.lib
        st = _lib.stream_ptr()
        fwd_ev, bwd_ev = [], []

        def events(n):
            evs = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
            for e in evs:
                e.record()  # creates the underlying hipEvent
            arr = (ctypes.c_void_p * n)(*[e.cuda_event for e in evs])
            return evs, arr

        assert ro.full, "kernel timing uses the non-checkpointed rollout"
        taped = ro.tape is not None
        fwd_names, bwd_names = (FWD_NAMES_TAPED, BWD_NAMES_TAPED) if taped else (FWD_NAMES, BWD_NAMES)
        batches = [eng._batch_with_scenes(ro.tape.scene_tabs[t]) for t in range(ro.T)] if taped else None
        # the adjoint substeps write the state-t gradients, and with a complete tape skip
        # the recompute kernels (as Rollout.backward does)
        ow = _lib.DDM_GRAD_OVERWRITE | (_lib.DDM_TAPE_COMPLETE if taped and ro.tape_complete else 0)
        bwd_batches = ([eng._batch_with(ro.tape.scene_tabs[t], ow) for t in range(ro.T)] if taped else
                       [eng._batch_with(None, ow)] * ro.T)
        for t in range(ro.T):
            evs, arr = events(4)
            if taped:
                _lib.check(lib.ddm_substep_forward_taped_timed(
                    ctypes.byref(batches[t]), ro._sview(t, t).data_ptr(), ro._sview(t + 1, t + 1).data_ptr(), st, arr))
            else:
                _lib.check(lib.ddm_substep_forward_timed(ctypes.byref(eng._batch), ro._sview(t, t).data_ptr(),
                                                         ro._sview(t + 1, t + 1).data_ptr(), st, arr))
            fwd_ev.append((evs, arr))
        gn = ro.g[0]
        gn.x.copy_(gx)
        gn.v.zero_()
        gn.F.zero_()
        gn.C.zero_()
        ro.pos_grad.zero_()
        ro.rot_grad.zero_()
        cur = 0
        for t in range(ro.T - 1, -1, -1):
            nxt, cur = cur, 1 - cur
            tabs = (ro._sview(t, t).data_ptr(), ro._sview(t + 1, t + 1).data_ptr(), ro._gview(cur, t).data_ptr(),
                    ro._gview(nxt, t + 1).data_ptr())
            if taped:
                evs, arr = events(5)
                _lib.check(lib.ddm_substep_backward_taped_timed(ctypes.byref(bwd_batches[t]), *tabs, st, arr))
            else:
                evs, arr = events(6)
                _lib.check(lib.ddm_substep_backward_timed(ctypes.byref(bwd_batches[t]), *tabs, st, arr))
            bwd_ev.append((evs, arr))
        torch.cuda.synchronize()
        tot = {k: 0.0 for k in fwd_names + bwd_names}
        for evs, _ in fwd_ev:
            for i, k in enumerate(fwd_names):
                tot[k] += evs[i].elapsed_time(evs[i + 1])
        for evs, _ in bwd_ev:
            for i, k in enumerate(bwd_names):
                tot[k] += evs[i].elapsed_time(evs[i + 1])
        return {k: v / ro.T for k, v in tot.items()}  # ms per launch


def cpu_baseline(budget_s=20.0):
    """The oracle (OpenMP build, FP32) timed on the host cores: repeated forward +
    adjoint passes of the plain 10k-particle scene (50 substeps each, as the plain
    workload) for about two thirds of the budget, then forward substeps of the
    50k-particle hand-contact scene for the rest."""
    from oracle import oracle as O
    from dexdeform_amd.mpm import workloads as W
    sc = W.plain_scene(n_particles=10000)
    s = sc.spec
    osc = O.Scene(grid_dim=tuple(s.grid_dim), dx=s.dx, dt=s.dt, gravity=tuple(s.gravity), mass=s.mass, vol=s.vol,
                  mly=s.mu_lam_yield, tfsr=s.body_prim, args=s.body_args)
    osim = O.OracleSim(osc, np.float32, omp=True)
    n = sc.n
    st0 = O.State(sc.x0, np.zeros((n, 3), np.float32), np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)),
                  np.zeros((n, 3, 3), np.float32), sc.body_pos, sc.body_rot)
    T = 50
    g = O.Grads(np.ones((n, 3), np.float32) / n, np.zeros((n, 3), np.float32), np.zeros((n, 3, 3), np.float32),
                np.zeros((n, 3, 3), np.float32), np.zeros((2, 3), np.float32), np.zeros((2, 4), np.float32))
    passes = 0
    t0 = time.perf_counter()
    while passes == 0 or time.perf_counter() - t0 < budget_s * 2 / 3:
        states = [st0]
        for _ in range(T):
            states.append(osim.substep(states[-1], sc.body_pos, sc.body_rot))
        gs = [O.Grads.zeros(n, 2, np.float32) for _ in range(T + 1)]
        gs[-1] = g
        for t in range(T - 1, -1, -1):
            osim.substep_grad(states[t], states[t + 1], gs[t + 1], gs[t])
        passes += 1
    dt = time.perf_counter() - t0
    out = {"value": n * T * passes / dt, "unit": "particle-substeps/s", "cores": O.num_threads(), "kind": "port",
           "sample": f"oracle OpenMP fp32 on {O.num_threads()} host threads: plain scene (10k particles, 64^3, "
                     f"2 capsules), {passes} passes of {T} substeps forward + {T} adjoint substeps ({dt:.1f} s)"}
    # forward only, hand-contact scene (50k particles, 22 capsule links)
    hs, tree, q0 = W.hand_scene(50000)
    hsp = hs.spec
    hosc = O.Scene(grid_dim=tuple(hsp.grid_dim), dx=hsp.dx, dt=hsp.dt, gravity=tuple(hsp.gravity), mass=hsp.mass,
                   vol=hsp.vol, mly=hsp.mu_lam_yield, tfsr=hsp.body_prim, args=hsp.body_args,
                   ground_friction=float(hsp.ground_friction))
    hsim = O.OracleSim(hosc, np.float32, omp=True)
    nh = hs.n
    st = O.State(hs.x0, np.zeros((nh, 3), np.float32), np.tile(np.eye(3, dtype=np.float32), (nh, 1, 1)),
                 np.zeros((nh, 3, 3), np.float32), hs.body_pos, hs.body_rot)
    k = 0
    t1 = time.perf_counter()
    while k == 0 or (time.perf_counter() - t1 < budget_s / 3 and k < 400):
        st = hsim.substep(st, hs.body_pos, hs.body_rot)
        k += 1
    dh = time.perf_counter() - t1
    out["forward_only_hand_50k"] = {"value": nh * k / dh, "unit": "particle-substeps/s (forward only)",
                                    "sample": f"hand-contact scene, 50k particles, 22 links, {k} forward substeps "
                                              f"({dh:.1f} s)"}
    return out


def _oracle_scene(O, spec):
    return O.Scene(grid_dim=tuple(int(g) for g in spec.grid_dim), dx=float(np.float32(spec.dx)),
                   dt=float(np.float32(spec.dt)), gravity=tuple(np.float32(spec.gravity)), mass=np.asarray(spec.mass),
                   vol=np.asarray(spec.vol), mly=np.asarray(spec.mu_lam_yield), tfsr=np.asarray(spec.body_prim),
                   args=np.asarray(spec.body_args), ground_friction=float(spec.ground_friction),
                   ground_height=float(spec.ground_height))


def oracle_check(scene, pose, gx_slice, x_final, g_x0):
    """Part of the CPU-baseline leg: one scene of the benchmarked batch run by the oracle
    (OpenMP build) for the same substeps, FP32 and FP64, with the adjoint from the same
    loss seed dL/dx_T. Relative L2 errors of the GPU's final positions and initial-position
    gradient (taken from the batched rollout) against the FP32 oracle, with the FP32
    oracle's own distance from the FP64 oracle beside them (the FP32 rounding noise)."""
    from oracle import oracle as O
    T = len(pose[0]) - 1
    n, nb = scene.n, scene.nb
    rel = lambda a, b: float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))  # noqa: E731
    res = {}
    t0 = time.perf_counter()
    for dtype in (np.float32, np.float64):
        osim = O.OracleSim(_oracle_scene(O, scene.spec), dtype, omp=True)
        st = O.State(scene.x0, np.zeros((n, 3)), np.tile(np.eye(3), (n, 1, 1)), np.zeros((n, 3, 3)),
                     pose[0][0], pose[1][0]).astype(dtype)
        states = osim.rollout(st, pose[0], pose[1])
        g_last = O.Grads(np.ascontiguousarray(gx_slice, dtype), np.zeros((n, 3), dtype), np.zeros((n, 3, 3), dtype),
                         np.zeros((n, 3, 3), dtype), np.zeros((nb, 3), dtype), np.zeros((nb, 4), dtype))
        g0, _, _ = osim.rollout_grad(states, g_last)
        res[dtype] = (states[-1].x.astype(np.float64), g0.x.astype(np.float64))
    (x32, g32), (x64, g64) = res[np.float32], res[np.float64]
    return {"scene_particles": n, "bodies": nb, "substeps": T,
            "traj_rel_err": rel(x_final, x32), "grad_rel_err": rel(g_x0, g32),
            "oracle_fp32_vs_fp64_traj": rel(x32, x64), "oracle_fp32_vs_fp64_grad": rel(g32, g64),
            "gpu_vs_fp64_traj": rel(x_final, x64), "gpu_vs_fp64_grad": rel(g_x0, g64),
            "seconds": time.perf_counter() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--substeps", type=int, default=100)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-score", action="store_true", help="skip the shape-score (Sinkhorn, Chamfer) leg")
    ap.add_argument("--no-render", action="store_true", help="skip the step + render leg")
    ap.add_argument("--only", default=None, metavar="LEGS",
                    help="diagnostic: run only these legs (comma list of score, render, trajopt) and print "
                         "one JSON line of them (for profiling a leg)")
    ap.add_argument("--no-tape", action="store_true",
                    help="adjoint recomputes P2G and the grid update (as the reference does) instead of "
                         "reading the forward pass's tape")
    ap.add_argument("--no-recompute", action="store_true",
                    help="skip the extra measurement of the recompute adjoint")
    ap.add_argument("--no-det", action="store_true",
                    help="skip the extra measurement of the deterministic sorted-reduction mode")
    ap.add_argument("--serve-rates", default="10,50,100,200", help="request arrival rates (req/s) to sweep")
    ap.add_argument("--serve-eager", action="store_true", help="serve without capturing ticks in HIP graphs")
    ap.add_argument("--no-named", action="store_true",
                    help="skip the single-scene plain 10k / hand 50k / large 1M rollouts")
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="arrival window per rate: rate x seconds requests (at least 24)")
    ap.add_argument("--named-only", default=None, metavar="NAMES",
                    help="diagnostic: run only these single-scene rollouts (comma list of plain_10k, "
                         "hand_50k, large_1m) and print their figures")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 256 MiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the benchmark runs on one GPU; multi-GPU runs are not supported (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    if args.only:  # diagnostic: only the listed legs, one JSON line of them
        legs = {"score": score_benchmark, "render": render_benchmark, "trajopt": trajectory_optimisation}
        print(json.dumps({name: legs[name](dev) for name in args.only.split(",")}), flush=True)
        return
    if args.named_only:  # diagnostic: only the named single-scene rollouts, one JSON line of them
        print(json.dumps({"single_scene_rollouts": named_rollouts(dev, max(1, args.steps),
                                                                  only=args.named_only.split(","))}), flush=True)
        return
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 268435456 bytes read, 268435456 bytes written")
        del src, dst
    T = args.substeps
    scenes, poses = make_batched(args.scenes, args.seed, T)
    n_total = sum(s.n for s in scenes)
    log(f"batched: {len(scenes)} scenes, {n_total} particles, {sum(s.nb for s in scenes)} bodies, T={T}")

    # ---------------- batched: eager launches, then the same rollout replayed from a HIP graph
    tape = not args.no_tape
    eng, ro, gx = build_rollout(scenes, poses, T, dev, tape=tape)
    for _ in range(args.warmup):
        run_step(ro, gx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        run_step(ro, gx)
        torch.cuda.synchronize()
        log(f"batched eager step {i + 1}/{args.steps} done")
    t1 = time.perf_counter()
    eager_value = n_total * T / ((t1 - t0) / args.steps)
    ro.capture(gx)
    for _ in range(args.warmup):
        ro.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ro.replay()
        torch.cuda.synchronize()
        log(f"batched graph step {i + 1}/{args.steps} done")
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = n_total * T / ((t1 - t0) / args.steps)
    # results of the timed rollout, compared below with each scene run alone and with the oracle
    x_final_b = ro.final_state().x.detach().clone()
    g_x0_b = ro.initial_grad().x.detach().clone()
    if tape:
        log(f"taped adjoint: {ro.tape_recomputed()} substeps of the last segment recomputed; "
            f"tape {ro.tape.nbytes() / 2**30:.1f} GiB")
    # the same replayed rollout with its inputs handed over in pageable host memory (as a
    # caller holding numpy arrays would) and its results read back: initial positions and
    # poses uploaded, final positions and the initial-position gradient downloaded
    x0_host = np.concatenate([s.x0 for s in scenes])
    pos_host = np.ascontiguousarray(np.concatenate([p[0] for p in poses], 1))
    rot_host = np.ascontiguousarray(np.concatenate([p[1] for p in poses], 1))
    pcie_bytes = x0_host.nbytes + pos_host.nbytes + rot_host.nbytes + 2 * x0_host.nbytes
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ro.set_initial(torch.from_numpy(x0_host))
        if eng.n_bodies_total:
            ro.set_poses(torch.from_numpy(pos_host), torch.from_numpy(rot_host))
        ro.replay()
        x_back, g_back = ro.final_state().x.cpu(), ro.initial_grad().x.cpu()
    torch.cuda.synchronize()
    pcie_sec = (time.perf_counter() - t0) / args.steps
    pcie_value = n_total * T / pcie_sec
    log(f"batched graph with host inputs and outputs: {pcie_value:.3e} particle-substeps/s "
        f"({pcie_bytes / 1e6:.0f} MB over PCIe per rollout, {pcie_sec * 1e3 - ms_step:.1f} ms more)")
    del x_back, g_back
    recompute_value = None
    if tape and not args.no_recompute:
        # the same rollout with the reference's recompute-in-backward adjoint
        del ro._graph
        _, rr, gr_ = build_rollout(scenes, poses, T, dev)
        rr.capture(gr_)
        rr.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            rr.replay()
            torch.cuda.synchronize()
        recompute_value = n_total * T / ((time.perf_counter() - t0) / args.steps)
        log(f"batched graph, recompute adjoint: {recompute_value:.3e} particle-substeps/s")
        del rr, gr_
        torch.cuda.empty_cache()
    det_value = None
    if not args.no_det:
        # the same rollout in the deterministic sorted-reduction mode (bitwise reproducible)
        _, rd, gd_ = build_rollout(scenes, poses, T, dev, tape=tape, deterministic=True)
        rd.capture(gd_)
        rd.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            rd.replay()
            torch.cuda.synchronize()
        det_value = n_total * T / ((time.perf_counter() - t0) / args.steps)
        log(f"batched graph, deterministic mode: {det_value:.3e} particle-substeps/s")
        del rd, gd_
        torch.cuda.empty_cache()
    kt = KernelTimer(ro).timed_step(gx)
    log("per-kernel ms/launch: " + ", ".join(f"{k}={v:.4f}" for k, v in kt.items()))
    kb = KERNEL_BYTES_TAPED if tape else KERNEL_BYTES
    bytes_per = {k: kb[k] * n_total for k in kt}
    dom = max(kt, key=kt.get)
    achieved = bytes_per[dom] / (kt[dom] * 1e-3) / 1e9 if kt[dom] > 0 else 0.0
    sum_ms = sum(kt.values())
    alg = 480.0 * n_total / (sum_ms * 1e-3) / 1e9
    waves = eng.n_chunks * 4  # 256-thread workgroups, one chunk each
    valu = None
    if dom in VALU_INSTR_PER_WAVE and kt[dom] > 0:
        v_ach = VALU_INSTR_PER_WAVE[dom] * waves / (kt[dom] * 1e-3)
        valu = {"bound": "valu", "kernel": dom, "achieved": v_ach, "peak": VALU_PEAK_WAVE_INSTR_S,
                "unit": "wave64 VALU instr/s", "frac": v_ach / VALU_PEAK_WAVE_INSTR_S,
                "instr_per_wave": VALU_INSTR_PER_WAVE[dom], "waves_per_launch": waves,
                "valu_busy_counter": VALU_BUSY.get(dom)}

    # ---------------- plain: the same scenes one at a time through the single-scene path,
    # eager and replayed from a per-scene HIP graph
    plain_value = plain_eager_value = batched_vs_alone = None
    if not args.no_plain:
        tot_e, tot_g, tot_ps = 0.0, 0.0, 0
        traj_errs, grad_errs = [], []
        for i, (sc, po) in enumerate(zip(scenes, poses)):
            sl = eng.scene_slice(i)
            e1, r1, g1 = build_rollout([sc], [po], T, dev, tape=tape, gx=gx[sl].clone())
            run_step(r1, g1)
            torch.cuda.synchronize()
            a = time.perf_counter()
            run_step(r1, g1)
            torch.cuda.synchronize()
            tot_e += time.perf_counter() - a
            r1.capture(g1)
            r1.replay()
            torch.cuda.synchronize()
            a = time.perf_counter()
            r1.replay()
            torch.cuda.synchronize()
            tot_g += time.perf_counter() - a
            tot_ps += sc.n * T
            traj_errs.append(float(torch.linalg.norm(r1.final_state().x - x_final_b[sl]) /
                                   torch.linalg.norm(x_final_b[sl])))
            grad_errs.append(float(torch.linalg.norm(r1.initial_grad().x - g_x0_b[sl]) /
                                   torch.linalg.norm(g_x0_b[sl])))
            del e1, r1, g1
            if (i + 1) % 16 == 0:
                log(f"plain {i + 1}/{len(scenes)} scenes done")
        plain_value = tot_ps / tot_g
        plain_eager_value = tot_ps / tot_e
        batched_vs_alone = {"scenes": len(scenes), "traj_rel_err_max": max(traj_errs),
                            "traj_rel_err_median": float(np.median(traj_errs)), "grad_rel_err_max": max(grad_errs),
                            "grad_rel_err_median": float(np.median(grad_errs))}
        log(f"batched vs alone: trajectory rel err max {max(traj_errs):.2e}, gradient rel err max "
            f"{max(grad_errs):.2e}")

    named = None if args.no_named else named_rollouts(dev, max(1, min(args.steps, 3)))
    tutorial = None if args.no_named else trajectory_optimisation(dev)

    serve = None
    if not args.no_serve:
        serve = serve_benchmark(dev, [float(r) for r in args.serve_rates.split(",")], args.serve_seconds, args.seed,
                                not args.serve_eager)

    score = None if args.no_score else score_benchmark(dev)
    render = None if args.no_render else render_benchmark(dev)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(args.cpu_budget)
        k = int(np.argmin([sc.n for sc in scenes]))
        sl = eng.scene_slice(k)
        cpu["oracle_check"] = oracle_check(scenes[k], poses[k], gx[sl].cpu().numpy(), x_final_b[sl].cpu().numpy(),
                                           g_x0_b[sl].cpu().numpy())
        log("oracle check (smallest scene): " + json.dumps(cpu["oracle_check"]))
    chk = cpu["oracle_check"] if cpu else None

    out = {
        "metric": "particle-substeps/s fwd+bwd, batched vs plain at <=1e-4 err; p50/p99 rollout ms",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched uneven scenes: {len(scenes)} scenes, particles U[5k,60k] "
                               f"(total {n_total}), 64^3 grid each, 2-22 capsules, {T} substeps fwd+bwd",
                   "global_batch": len(scenes), "substeps": T, "particles": n_total, "parallelism": "none"},
        "adjoint": "taped" if tape else "recompute",
        "recompute_value": recompute_value,
        "deterministic_value": det_value,
        "eager_value": eager_value,
        "pcie_inclusive_value": pcie_value,
        "pcie_bytes_per_rollout": pcie_bytes,
        "plain_value": plain_value,
        "plain_eager_value": plain_eager_value,
        "batched_over_plain": (value / plain_value) if plain_value else None,
        # accuracy of the benchmarked rollout: one scene against the oracle (FP32; the FP32
        # oracle's distance from its FP64 run is the rounding-noise reference), and every
        # scene batched against the same scene run alone
        "traj_rel_err": chk["traj_rel_err"] if chk else None,
        "grad_rel_err": chk["grad_rel_err"] if chk else None,
        "batched_vs_alone": batched_vs_alone,
        "kernel_ms_per_launch": kt,
        "kernel_time_source": "hipEvent pairs recorded around each stage kernel on its launch stream, averaged "
                              f"over the {T} launches of one timed rollout",
        "algorithmic_fwd_bwd_GBps": alg,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_particle": kb[dom],
                     "particles_per_launch": n_total, "kernel_ms": kt[dom],
                     "traffic": (TRAFFIC_BYTES[("taped_" if tape else "") + dom] * n_total
                                 if (args.scenes, args.seed) == (64, 7) and
                                 ("taped_" if tape else "") + dom in TRAFFIC_BYTES else None)},
        "roofline_valu": valu,
        "single_scene_rollouts": named,
        "trajectory_optimisation": tutorial,
        "serve": serve,
        "score": score,
        "render": render,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
