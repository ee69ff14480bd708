"""Batched rollouts, checkpointing and the continuous-batching server on the GPU.

* a checkpointed rollout (recompute every K substeps) gives the same gradients as the
  rollout that stores every state;
* a multi-scene batched rollout gives each scene the result it gets alone;
* the server gives each request the same loss and gradients alone and in a mix of
  concurrently running requests of different sizes, grids and horizons, and the same as
  a direct rollout of that request.
Tolerance: 1e-4 relative for states and losses, 1e-3 for gradients (FP32 atomic order
differs between runs).
"""
import copy

import numpy as np
import pytest
import torch

from tests.helpers import rel_err
from dexdeform_amd.mpm import workloads as W
from dexdeform_amd.mpm.engine import Engine
from dexdeform_amd.mpm.rollout import Rollout
from dexdeform_amd.mpm.serve import Request, RolloutServer, latency_stats

pytestmark = pytest.mark.gpu


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def _scene(n, grid, seed, nb=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 1.0 / grid
    x0 = W.box_particles(n, [0.42, 0.1, 0.42], [0.58, 0.22, 0.58], seed, dx)
    pos, rot, args = W.random_capsules(rng, nb, [0.5, 0.27, 0.5], spread=0.05, r_range=(0.02, 0.04),
                                       h_range=(0.04, 0.08))
    return W.make_scene(x0, grid, dx, 5e-5, W.capsule_prims(nb), args, pos, rot)


def _poses(sc, T, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    vel = rng.uniform(-0.5, 0.5, size=(sc.nb, 3)).astype(np.float32)
    t = np.arange(T + 1, dtype=np.float32)[:, None, None] * sc.spec.dt
    return (sc.body_pos[None] + t * vel[None]).astype(np.float32), np.repeat(sc.body_rot[None], T + 1, 0)


def _rollout(scenes, T, K=None, pose_seeds=None, grad_scale=None, tape=False, tape_blocks=None):
    dev = _dev()
    eng = Engine([s.spec for s in scenes], device=dev)
    ro = Rollout(eng, T, checkpoint_every=K, tape=tape, tape_blocks=tape_blocks)
    ro.set_initial(torch.tensor(np.concatenate([s.x0 for s in scenes]), device=dev))
    seeds = pose_seeds or [100 + i for i in range(len(scenes))]
    P = [_poses(s, T, sd) for s, sd in zip(scenes, seeds)]
    ro.set_poses(torch.tensor(np.concatenate([p[0] for p in P], 1), device=dev),
                 torch.tensor(np.concatenate([p[1] for p in P], 1), device=dev))
    xT = ro.forward().x.clone()
    gx = torch.ones_like(xT) / (grad_scale or len(xT))
    ro.backward(x_grad=gx)
    torch.cuda.synchronize()
    return eng, ro, xT.cpu().numpy(), ro.pos_grad.cpu().numpy(), ro.initial_grad().x.cpu().numpy()


def test_checkpointed_rollout_matches_full_storage():
    sc = _scene(4000, 48, seed=1)
    _, _, x_full, pg_full, g0_full = _rollout([sc], 40)
    _, ro, x_ck, pg_ck, g0_ck = _rollout([sc], 40, K=10)
    e = [rel_err(x_ck, x_full), rel_err(pg_ck, pg_full), rel_err(g0_ck, g0_full)]
    print(f"checkpointed vs full: x {e[0]:.2e} pose grad {e[1]:.2e} x0 grad {e[2]:.2e}")
    assert e[0] <= 1e-4 and e[1] <= 1e-3 and e[2] <= 1e-3


def test_batched_scenes_match_single_scene_runs():
    scenes = [_scene(3000, 48, seed=2), _scene(6000, 64, seed=3, nb=5), _scene(1500, 32, seed=4, nb=2)]
    eng, ro, xT, pg, g0 = _rollout(scenes, 20)
    for i, sc in enumerate(scenes):
        _, _, x1, pg1, g01 = _rollout([sc], 20, pose_seeds=[100 + i], grad_scale=len(xT))
        sl = eng.scene_slice(i)
        bs = eng.body_slice(i)
        ex = rel_err(x1, xT[sl])
        eg = rel_err(pg1[:, :sc.nb], pg[:, bs])
        e0 = rel_err(g01, g0[sl])
        print(f"scene {i}: batched vs alone x {ex:.2e} pose grad {eg:.2e} x0 grad {e0:.2e}")
        assert ex <= 1e-4 and eg <= 1e-3 and e0 <= 1e-3


def _requests():
    specs = [(3000, 48, 20, None), (8000, 64, 40, 10), (2000, 32, 30, 10), (5000, 48, 20, None),
             (4000, 64, 40, 20)]
    reqs = []
    for i, (n, g, T, K) in enumerate(specs):
        sc = _scene(n, g, seed=10 + i, nb=2 + i)
        pos, rot = _poses(sc, T, 200 + i)
        target = W.plain_target(sc.x0, seed=300 + i)
        reqs.append(Request(i, sc.spec, sc.x0, pos, rot, target, arrival=0.0, checkpoint_every=K))
    return reqs


@pytest.mark.parametrize("full_tape_bytes", [0, None])
def test_server_results_alone_equal_mixed(full_tape_bytes):
    """0: every request runs with its checkpoint interval (recomputed segments); None:
    these small requests keep their whole trajectory (no recompute)."""
    reqs = _requests()
    alone = {}
    for r in reqs:
        srv = RolloutServer(device=_dev(), arena_blocks=8192, memory_budget=8 << 30, full_tape_bytes=full_tape_bytes)
        alone[r.rid] = srv.run([r])[r.rid]
    # mixed: staggered arrivals so requests join while others are mid-rollout
    mixed_reqs = _requests()
    for k, r in enumerate(mixed_reqs):
        r.arrival = 0.002 * k
    # the arena holds every request's grid at once (1728 + 4096 + 512 + 1728 + 4096 blocks)
    srv = RolloutServer(device=_dev(), arena_blocks=16384, memory_budget=8 << 30, full_tape_bytes=full_tape_bytes)
    assert all((srv.interval(r) == r.T) == (full_tape_bytes is None) for r in reqs if r.checkpoint_every)
    mixed = srv.run(mixed_reqs)
    stats = latency_stats(mixed)
    print("mixed latency", stats, "ticks", srv.ticks)
    # requests overlap on the device: several are admitted and unfinished at once, and the
    # server needs fewer ticks than one-at-a-time execution (the sum of their operations)
    ops = [2 * r.T + (r.T if srv.interval(r) < r.T else 0) for r in reqs]
    events = sorted([(r.admitted, 1) for r in mixed.values()] + [(r.finished, -1) for r in mixed.values()])
    live, most = 0, 0
    for _, d in events:
        live += d
        most = max(most, live)
    print(f"at most {most} requests in flight; {srv.ticks} ticks vs {sum(ops)} serial operations")
    assert most >= 3 and srv.ticks < sum(ops), "requests did not overlap"
    for rid, a in alone.items():
        b = mixed[rid]
        el = abs(a.loss - b.loss) / abs(a.loss)
        ex = rel_err(b.x_final.cpu().numpy(), a.x_final.cpu().numpy())
        eg = rel_err(b.pos_grad.cpu().numpy(), a.pos_grad.cpu().numpy())
        e0 = rel_err(b.x0_grad.cpu().numpy(), a.x0_grad.cpu().numpy())
        print(f"request {rid}: loss {a.loss:.6f} | mixed vs alone: loss {el:.1e} x {ex:.1e} pose grad {eg:.1e} "
              f"x0 grad {e0:.1e}")
        assert el <= 1e-4 and ex <= 1e-4 and eg <= 1e-3 and e0 <= 1e-3


def test_server_graph_ticks_match_eager_ticks():
    """Ticks replayed from captured HIP graphs (launches sized to power-of-two buckets,
    live counts read on the device) give the results of exactly sized eager launches."""
    results, servers = {}, {}
    for graphs in (True, False):
        reqs = _requests()
        for k, r in enumerate(reqs):
            r.arrival = 0.002 * k
        srv = RolloutServer(device=_dev(), arena_blocks=16384, memory_budget=8 << 30, use_graphs=graphs)
        results[graphs] = srv.run(reqs)
        servers[graphs] = srv
    g = servers[True]
    print(f"graph server: {g.ticks} ticks, {g.replays} replayed, {len(g.graphs)} graphs captured")
    assert g.replays == g.ticks, "every tick replays a captured graph"
    assert len(g.graphs) < g.ticks // 4, "graphs are reused across ticks"
    assert servers[False].replays == 0
    for rid, a in results[False].items():
        b = results[True][rid]
        el = abs(a.loss - b.loss) / abs(a.loss)
        eg = rel_err(b.pos_grad.cpu().numpy(), a.pos_grad.cpu().numpy())
        e0 = rel_err(b.x0_grad.cpu().numpy(), a.x0_grad.cpu().numpy())
        print(f"request {rid}: graph vs eager: loss {el:.1e} pose grad {eg:.1e} x0 grad {e0:.1e}")
        assert el <= 1e-5 and eg <= 1e-3 and e0 <= 1e-3


def test_server_sorts_particles_given_in_random_order():
    """A request whose particles come in random order (as the reference's samplers make
    them) is sorted at admission and answered in its own order: the same loss, final
    positions and gradients as the spatially ordered request."""
    r = _requests()[1]
    perm = np.random.default_rng(8).permutation(r.n)
    sp = copy.deepcopy(r.spec)
    sp.mass, sp.vol, sp.mu_lam_yield = sp.mass[perm], sp.vol[perm], np.asarray(sp.mu_lam_yield)[perm]
    rs = Request(r.rid, sp, r.x0[perm], r.pose_pos, r.pose_rot, r.target[perm], checkpoint_every=r.checkpoint_every)
    a = RolloutServer(device=_dev(), arena_blocks=8192, memory_budget=8 << 30).run([r])[r.rid]
    b = RolloutServer(device=_dev(), arena_blocks=8192, memory_budget=8 << 30).run([rs])[r.rid]
    el = abs(a.loss - b.loss) / abs(a.loss)
    ex = rel_err(b.x_final.cpu().numpy(), a.x_final.cpu().numpy()[perm])
    e0 = rel_err(b.x0_grad.cpu().numpy(), a.x0_grad.cpu().numpy()[perm])
    eg = rel_err(b.pos_grad.cpu().numpy(), a.pos_grad.cpu().numpy())
    print(f"random-order request vs ordered: loss {el:.1e} x {ex:.1e} x0 grad {e0:.1e} pose grad {eg:.1e}")
    assert el <= 1e-5 and ex <= 1e-5 and e0 <= 1e-3 and eg <= 1e-3


def test_failing_request_is_flagged_and_isolated():
    """A request whose loss is not finite (a NaN in its target, injected from the host:
    positions stay finite, only its loss and gradients are NaN) is reported with
    finite=False; the requests served with it get their stand-alone results."""
    reqs = _requests()[:3]
    bad = reqs[1]
    bad.target = bad.target.copy()
    bad.target[7, 1] = np.nan
    srv = RolloutServer(device=_dev(), arena_blocks=16384, memory_budget=8 << 30)
    res = srv.run(reqs)
    assert not res[bad.rid].finite and not np.isfinite(res[bad.rid].loss)
    for r in (reqs[0], reqs[2]):
        alone = RolloutServer(device=_dev(), arena_blocks=8192, memory_budget=8 << 30).run([_requests()[r.rid]])[r.rid]
        b = res[r.rid]
        assert b.finite and alone.finite
        el = abs(alone.loss - b.loss) / abs(alone.loss)
        eg = rel_err(b.pos_grad.cpu().numpy(), alone.pos_grad.cpu().numpy())
        print(f"request {r.rid} next to a failing request: loss {el:.1e} pose grad {eg:.1e}")
        assert el <= 1e-4 and eg <= 1e-3


def test_server_matches_direct_rollout():
    r = _requests()[1]
    srv = RolloutServer(device=_dev(), arena_blocks=8192, memory_budget=8 << 30)
    res = srv.run([r])[r.rid]
    dev = _dev()
    eng = Engine([r.spec], device=dev)
    ro = Rollout(eng, r.T, checkpoint_every=r.checkpoint_every)
    ro.set_initial(torch.tensor(r.x0, device=dev))
    ro.set_poses(torch.tensor(r.pose_pos, device=dev), torch.tensor(r.pose_rot, device=dev))
    xT = ro.forward().x.clone()
    d = xT - torch.tensor(r.target, device=dev)
    loss = (d * d).sum(-1).mean()
    ro.backward(x_grad=2.0 / r.n * d)
    torch.cuda.synchronize()
    assert abs(loss.item() - res.loss) <= 1e-4 * abs(loss.item())
    eg = rel_err(res.pos_grad.cpu().numpy()[:, :r.spec.n_bodies], ro.pos_grad.cpu().numpy()[:, :r.spec.n_bodies])
    print(f"server vs direct rollout: pose grad {eg:.2e}")
    assert eg <= 1e-3


def test_server_admission_waits_for_grid_blocks():
    """An arena smaller than the sum of the requests' grids: the last request is admitted
    only after earlier ones release their blocks, and still gets its stand-alone result."""
    reqs = _requests()
    srv = RolloutServer(device=_dev(), arena_blocks=8192, memory_budget=8 << 30)
    res = srv.run(reqs)
    first_finish = min(r.finished for r in res.values())
    assert res[4].admitted >= first_finish, "request 4 cannot fit before another request finished"
    ref = RolloutServer(device=_dev(), arena_blocks=8192, memory_budget=8 << 30).run([_requests()[4]])[4]
    assert abs(res[4].loss - ref.loss) <= 1e-4 * abs(ref.loss)
    eg = rel_err(res[4].pos_grad.cpu().numpy(), ref.pos_grad.cpu().numpy())
    print(f"queued request vs alone: pose grad {eg:.2e}")
    assert eg <= 1e-3


def test_captured_rollout_graph_replays_match_eager():
    """A batched, checkpointed rollout captured into one HIP graph: three replays, with
    eager rollouts and synchronizations in between, each match the eager pass; the
    second and third replays run with a doubled loss seed, which the graph must read."""
    dev = _dev()
    scenes = [_scene(3000, 48, seed=2), _scene(6000, 64, seed=3, nb=5), _scene(1500, 32, seed=4, nb=2)]
    T, K = 12, 4
    eng = Engine([s.spec for s in scenes], device=dev)

    def make():
        ro = Rollout(eng, T, checkpoint_every=K)
        ro.set_initial(torch.tensor(np.concatenate([s.x0 for s in scenes]), device=dev))
        P = [_poses(s, T, 100 + i) for i, s in enumerate(scenes)]
        ro.set_poses(torch.tensor(np.concatenate([p[0] for p in P], 1), device=dev),
                     torch.tensor(np.concatenate([p[1] for p in P], 1), device=dev))
        return ro

    def outputs(ro):
        torch.cuda.synchronize()
        return (ro.final_state().x.cpu().numpy(), ro.pos_grad.cpu().numpy(), ro.initial_grad().x.cpu().numpy())

    n = eng.n_particles_total
    seed = torch.full((n, 3), 1.0 / n, device=dev)
    eager = make()
    eager.forward()
    eager.backward(x_grad=seed)
    ref = outputs(eager)

    ro = make()
    ro.capture(seed)
    other = make()
    for rep in range(3):
        if rep == 1:
            seed.mul_(2.0)
        scale = 1.0 if rep == 0 else 2.0
        ro.replay()
        x, pg, g0 = outputs(ro)
        ex, eg, e0 = rel_err(x, ref[0]), rel_err(pg, scale * ref[1]), rel_err(g0, scale * ref[2])
        print(f"replay {rep}: x {ex:.2e} pose grad {eg:.2e} x0 grad {e0:.2e}")
        assert ex <= 1e-5 and eg <= 1e-4 and e0 <= 1e-4
        other.forward()  # eager work between replays
        other.backward(x_grad=seed)
        _, opg, _ = outputs(other)
        assert rel_err(opg, scale * ref[1]) <= 1e-4


def test_repeated_checkpointed_forward_passes_restart_from_the_initial_state():
    """Forward-only passes of a checkpointed rollout (as finite differences make them)
    each start from the initial state: the second pass reproduces the first."""
    dev = _dev()
    sc = _scene(3000, 48, seed=5)
    eng = Engine([sc.spec], device=dev)
    ro = Rollout(eng, 20, checkpoint_every=5)
    ro.set_initial(torch.tensor(sc.x0, device=dev))
    pos, rot = _poses(sc, 20, 7)
    ro.set_poses(torch.tensor(pos, device=dev), torch.tensor(rot, device=dev))
    a = ro.forward().x.clone()
    b = ro.forward().x.clone()
    torch.cuda.synchronize()
    e = rel_err(b.cpu().numpy(), a.cpu().numpy())
    print(f"second forward pass vs first: x {e:.2e}")
    assert e <= 1e-6


@pytest.mark.parametrize("K", [None, 10])
def test_taped_adjoint_matches_recompute(K):
    """The taped adjoint (grid snapshot + temporaries of the forward pass) gives the
    recomputing adjoint's gradients; with a snapshot too small for the substeps' blocks,
    every adjoint substep falls back to recomputing, with the same result."""
    scenes = [_scene(3000, 48, seed=2), _scene(6000, 64, seed=3, nb=5)]
    _, _, x_ref, pg_ref, g0_ref = _rollout(scenes, 20, K=K)
    _, ro, x_tp, pg_tp, g0_tp = _rollout(scenes, 20, K=K, tape=True)
    assert ro.tape_recomputed() == 0
    e = [rel_err(x_tp, x_ref), rel_err(pg_tp, pg_ref), rel_err(g0_tp, g0_ref)]
    print(f"taped vs recompute (K={K}): x {e[0]:.2e} pose grad {e[1]:.2e} x0 grad {e[2]:.2e}")
    assert e[0] <= 1e-6 and e[1] <= 1e-4 and e[2] <= 1e-4
    _, ro, x_of, pg_of, g0_of = _rollout(scenes, 20, K=K, tape=True, tape_blocks=4)
    assert ro.tape_recomputed() == 2 * ro.K, "4-block snapshots cannot hold these scenes' substeps"
    e = [rel_err(pg_of, pg_ref), rel_err(g0_of, g0_ref)]
    print(f"overflowing tapes (K={K}): pose grad {e[0]:.2e} x0 grad {e[1]:.2e}")
    assert e[0] <= 1e-4 and e[1] <= 1e-4
    # one scene restored, the other recomputed, in the same launches
    _, ro, x_mx, pg_mx, g0_mx = _rollout(scenes, 20, K=K, tape=True, tape_blocks=[1728, 4])
    assert ro.tape_recomputed() == ro.K, "only the second scene's 4-block snapshots overflow"
    e = [rel_err(pg_mx, pg_ref), rel_err(g0_mx, g0_ref)]
    print(f"mixed restore/recompute (K={K}): pose grad {e[0]:.2e} x0 grad {e[1]:.2e}")
    assert e[0] <= 1e-4 and e[1] <= 1e-4


def test_second_backward_after_one_checkpointed_taped_forward():
    """forward() runs the last segment taped and backward() starts from that tape; a
    second backward() without a forward pass runs that segment again (the first one
    overwrote the tape with earlier segments) and gives the same gradients."""
    dev = _dev()
    scenes = [_scene(3000, 48, seed=2), _scene(6000, 64, seed=3, nb=5)]
    eng = Engine([s.spec for s in scenes], device=dev)
    ro = Rollout(eng, 20, checkpoint_every=5, tape=True)
    ro.set_initial(torch.tensor(np.concatenate([s.x0 for s in scenes]), device=dev))
    P = [_poses(s, 20, 100 + i) for i, s in enumerate(scenes)]
    ro.set_poses(torch.tensor(np.concatenate([p[0] for p in P], 1), device=dev),
                 torch.tensor(np.concatenate([p[1] for p in P], 1), device=dev))
    xT = ro.forward().x.clone()
    gx = torch.ones_like(xT) / len(xT)
    out = []
    for _ in range(2):
        ro.backward(x_grad=gx)
        torch.cuda.synchronize()
        out.append((ro.pos_grad.cpu().numpy(), ro.initial_grad().x.cpu().numpy()))
    e = [rel_err(out[1][0], out[0][0]), rel_err(out[1][1], out[0][1])]
    print(f"second backward vs first: pose grad {e[0]:.2e} x0 grad {e[1]:.2e}")
    assert e[0] <= 1e-4 and e[1] <= 1e-4


def test_forward_ahead_matches_separate_substeps():
    """A rollout runs each G2P inside the next substep's P2G launch (small launch: the fused
    kernel); stepping the engine one plain substep at a time gives the same states."""
    from dexdeform_amd.mpm.engine import StateBuffers
    dev = _dev()
    sc = _scene(3000, 48, seed=4)
    T = 12
    pos, rot = _poses(sc, T, 9)
    eng = Engine([sc.spec], device=dev)
    ro = Rollout(eng, T)
    ro.set_initial(torch.tensor(sc.x0, device=dev))
    ro.set_poses(torch.tensor(pos, device=dev), torch.tensor(rot, device=dev))
    xa = ro.forward()
    xa = [t.clone() for t in (xa.x, xa.v, xa.F, xa.C)]
    bufs = [StateBuffers(sc.n, sc.nb, dev) for _ in range(2)]
    bufs[0].x.copy_(torch.tensor(sc.x0, device=dev))
    for t in range(T):
        cur, nxt = bufs[t % 2], bufs[(t + 1) % 2]
        cur.body_pos.copy_(torch.tensor(pos[t], device=dev))
        cur.body_rot.copy_(torch.tensor(rot[t], device=dev))
        nxt.body_pos.copy_(torch.tensor(pos[t + 1], device=dev))
        nxt.body_rot.copy_(torch.tensor(rot[t + 1], device=dev))
        eng.forward(cur, nxt)
    torch.cuda.synchronize()
    fin = bufs[T % 2]
    e = [rel_err(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(xa, (fin.x, fin.v, fin.F, fin.C))]
    print(f"forward ahead vs separate substeps: x {e[0]:.1e} v {e[1]:.1e} F {e[2]:.1e} C {e[3]:.1e}")
    # the same operations; the grid's float atomics add in another order in each run, which
    # twelve substeps amplify most in C (measured x 2e-9, v 1.5e-6, F 1.6e-8, C 2.8e-5)
    assert e[0] <= 1e-6 and max(e[1:]) <= 1e-4, e


def test_persistent_loop_serves_requests_submitted_while_it_runs():
    """The serving loop runs in a background thread; requests are submitted from this
    thread one by one while earlier ones are on the GPU, each future resolves with the
    result the request gets alone, and the requests overlap on the device."""
    import time as _time
    reqs = _requests()
    srv = RolloutServer(device=_dev(), arena_blocks=16384, memory_budget=8 << 30).start()
    try:
        futs = []
        for r in reqs:  # 0.5 ms apart: each request takes a few ms on the device
            futs.append(srv.submit_async(r))
            _time.sleep(0.0005)
        dup = srv.submit_async(copy.copy(reqs[-1]))  # same id while the first is being served
        with pytest.raises(ValueError):
            srv.submit_async(Request(99, reqs[0].spec, reqs[0].x0, reqs[0].pose_pos[:1], reqs[0].pose_rot[:1],
                                     reqs[0].target))  # horizon 0: rejected in the caller's thread
        results = {r.rid: f.result(timeout=300) for r, f in zip(reqs, futs)}
        with pytest.raises(ValueError):
            dup.result(timeout=60)
    finally:
        srv.stop()
    events = sorted([(r.admitted, 1) for r in results.values()] + [(r.finished, -1) for r in results.values()])
    live, most = 0, 0
    for _, d in events:
        live += d
        most = max(most, live)
    print(f"persistent loop: at most {most} requests in flight, {srv.ticks} ticks, "
          f"latencies {[round(r.latency * 1e3, 1) for r in results.values()]} ms")
    assert most >= 2, "requests submitted while others ran did not overlap"
    for r in reqs:
        fresh = Request(r.rid, r.spec, r.x0, r.pose_pos, r.pose_rot, r.target, arrival=0.0,
                        checkpoint_every=r.checkpoint_every)
        alone = RolloutServer(device=_dev(), arena_blocks=8192, memory_budget=8 << 30).run([fresh])[r.rid]
        b = results[r.rid]
        el = abs(alone.loss - b.loss) / abs(alone.loss)
        ex = rel_err(b.x_final.cpu().numpy(), alone.x_final.cpu().numpy())
        eg = rel_err(b.pos_grad.cpu().numpy(), alone.pos_grad.cpu().numpy())
        e0 = rel_err(b.x0_grad.cpu().numpy(), alone.x0_grad.cpu().numpy())
        print(f"request {r.rid}: loop vs alone loss {el:.2e} x {ex:.2e} pose grad {eg:.2e} x0 grad {e0:.2e}")
        assert el <= 1e-4 and ex <= 1e-4 and eg <= 1e-3 and e0 <= 1e-3
