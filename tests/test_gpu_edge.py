"""Edge cases of the fused multi-scene path on the GPU.

* a batch holding a scene with no live particles and a scene with one particle: the
  empty scene launches nothing and disturbs nothing, and every scene gets its
  stand-alone result (1e-4 states, 1e-3 gradients);
* the far blocks of a large grid: a scene near the high corner of a 256^3 grid (block
  indices above 200000) matches the same scene translated by 64 cells in x and z into a
  160^3 grid with the same cell size (translation invariance away from the walls; the
  positions lie in [0.5, 1), where the translation by 0.25 is exact in FP32, so only
  the atomic order differs: 1e-4 / 1e-3);
* the DDM_EINVAL path: a batch asking for the deterministic mode without its workspace
  is rejected before any launch;
* a HIP graph capture while a dead graph sits in a reference cycle: the capture pauses
  automatic garbage collection, and its replay matches the eager pass.
"""
import ctypes

import numpy as np
import pytest
import torch

from tests.helpers import rel_err
from dexdeform_amd import _lib
from dexdeform_amd.mpm import workloads as W
from dexdeform_amd.mpm.engine import Engine
from dexdeform_amd.mpm.rollout import Rollout

pytestmark = pytest.mark.gpu


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def _scene(n, grid, dx, lo, hi, body_center, seed, nb=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    x0 = W.box_particles(n, lo, hi, seed, dx)
    pos, rot, args = W.random_capsules(rng, nb, body_center, spread=0.02, r_range=(0.006, 0.01),
                                       h_range=(0.01, 0.03))
    return W.make_scene(x0, grid, dx, 2e-5, W.capsule_prims(nb), args, pos, rot)


def _run(scenes, T, counts=None, K=None, tape=False):
    dev = _dev()
    eng = Engine([s.spec for s in scenes], device=dev)
    if counts is not None:
        eng.set_counts(counts)
    ro = Rollout(eng, T, checkpoint_every=K, tape=tape)
    ro.set_initial(torch.tensor(np.concatenate([s.x0 for s in scenes]), device=dev))
    vel = [np.float32([0.3, -0.2, 0.1]) for _ in scenes]
    t = np.arange(T + 1, dtype=np.float32)[:, None, None]
    pos = np.concatenate([(s.body_pos[None] + t * s.spec.dt * v).astype(np.float32) for s, v in zip(scenes, vel)], 1)
    rot = np.concatenate([np.repeat(s.body_rot[None], T + 1, 0) for s in scenes], 1)
    ro.set_poses(torch.tensor(pos, device=dev), torch.tensor(rot, device=dev))
    xT = ro.forward().x.clone()
    ro.backward(x_grad=torch.full_like(xT, 1e-3))
    torch.cuda.synchronize()
    return eng, xT.cpu().numpy(), ro.pos_grad.cpu().numpy(), ro.initial_grad().x.cpu().numpy()


def test_empty_and_single_particle_scenes_in_a_batch():
    dx = 1 / 64
    a = _scene(3000, 64, dx, [0.42, 0.1, 0.42], [0.58, 0.2, 0.58], [0.5, 0.24, 0.5], seed=1)
    b = _scene(2000, 48, dx, [0.3, 0.1, 0.3], [0.4, 0.2, 0.4], [0.35, 0.24, 0.35], seed=2)
    c = _scene(1, 32, dx, [0.25, 0.15, 0.25], [0.2501, 0.1501, 0.2501], [0.25, 0.2, 0.25], seed=3, nb=1)
    T = 8
    eng, x, pg, g0 = _run([a, b, c], T, counts=[3000, 0, 1], K=4, tape=True)
    assert not np.any(pg[:, eng.body_slice(1)]), "a scene without particles has no contacts"
    for i, sc in ((0, a), (2, c)):
        e1, xa, pga, g0a = _run([sc], T, K=4, tape=True)
        ps, bs = eng.scene_slice(i), eng.body_slice(i)
        ex, eg, e0 = rel_err(x[ps], xa), rel_err(pg[:, bs], pga), rel_err(g0[ps], g0a)
        print(f"scene {i}: batched with an empty scene vs alone: x {ex:.1e} pose grad {eg:.1e} x0 grad {e0:.1e}")
        assert ex <= 1e-4 and eg <= 1e-3 and e0 <= 1e-3
    # every scene empty: the launches have no chunks and return at once
    _, _, pg, _ = _run([a, b], 2, counts=[0, 0])
    assert not np.any(pg)


def test_far_blocks_of_a_256_grid_match_a_translated_smaller_grid():
    dx = 1 / 256
    lo, hi, body = [0.55, 0.02, 0.55], [0.59, 0.05, 0.59], [0.57, 0.06, 0.57]
    small = _scene(20000, 160, dx, lo, hi, body, seed=4)
    shift = np.float32([0.25, 0.0, 0.25])  # 64 cells in x and z, exact for x, z in [0.5, 0.75)
    assert np.array_equal((small.x0 + shift) - shift, small.x0)
    big = W.make_scene(small.x0 + shift, 256, dx, small.spec.dt, small.spec.body_prim, small.spec.body_args,
                       small.body_pos + shift, small.body_rot)
    T = 10
    e_big, xb, pgb, g0b = _run([big], T)
    _, xs, pgs, g0s = _run([small], T)
    ex = rel_err(xb - shift, xs)
    eg, e0 = rel_err(pgb, pgs), rel_err(g0b, g0s)
    print(f"256^3 grid ({e_big.n_blocks} blocks), scene near x=z=0.82 vs 160^3 at x=z=0.57: x {ex:.1e} "
          f"pose grad {eg:.1e} x0 grad {e0:.1e}")
    assert ex <= 1e-4 and eg <= 1e-3 and e0 <= 1e-3


def test_deterministic_option_without_workspace_is_rejected():
    sc = _scene(500, 32, 1 / 32, [0.4, 0.1, 0.4], [0.6, 0.2, 0.6], [0.5, 0.25, 0.5], seed=5)
    eng = Engine([sc.spec], device=_dev())
    cur, nxt = eng.new_state(), eng.new_state()
    b = _lib.DdmBatch.from_buffer_copy(eng._batch)
    b.options = _lib.DDM_DETERMINISTIC
    st = eng.lib.ddm_substep_forward(ctypes.byref(b), eng.table_of(cur).data_ptr(), eng.table_of(nxt).data_ptr(),
                                     _lib.stream_ptr())
    assert st == _lib.DDM_EINVAL


def test_capture_with_a_dead_graph_in_a_reference_cycle():
    """A dead reference cycle that holds a captured HIP graph is not collected during a
    later capture (the graph's destructor fails while a stream is capturing and
    terminates the process): captures pause Python's automatic garbage collection. The
    cycle is dropped inside the capture with the collection threshold at 1, so an
    unpaused collector would run there."""
    import gc
    dev = _dev()
    sc = _scene(2000, 32, 1 / 32, [0.4, 0.1, 0.4], [0.6, 0.2, 0.6], [0.5, 0.25, 0.5], seed=6)
    T = 4

    def rollout():
        eng = Engine([sc.spec], device=dev)
        ro = Rollout(eng, T, tape=True)
        ro.set_initial(torch.tensor(sc.x0, device=dev))
        t = np.arange(T + 1, dtype=np.float32)[:, None, None]
        pos = (sc.body_pos[None] + t * sc.spec.dt * np.float32([0.3, -0.2, 0.1])).astype(np.float32)
        rot = np.repeat(sc.body_rot[None], T + 1, 0)
        ro.set_poses(torch.tensor(pos, device=dev), torch.tensor(rot, device=dev))
        return eng, ro

    seed = torch.full((sc.n, 3), 1e-3, device=dev)
    e1, r1 = rollout()
    r1.capture(seed)
    cycle = [r1._graph]
    cycle.append(cycle)
    r1._graph = None
    del e1, r1
    e2, r2 = rollout()
    fwd = r2.forward
    dropped = []

    def forward_dropping_the_cycle():
        nonlocal cycle
        if cycle is not None and torch.cuda.is_current_stream_capturing():
            cycle = None  # the graph is now garbage in a cycle
            dropped.append([[] for _ in range(100)])  # allocations that trigger collections
        return fwd()

    r2.forward = forward_dropping_the_cycle
    thresholds = gc.get_threshold()
    gc.set_threshold(1, 1, 1)
    try:
        r2.capture(seed)
    finally:
        gc.set_threshold(*thresholds)
    assert dropped, "the cycle was not dropped inside the capture"
    gc.collect()  # outside the capture: the dead graph goes now
    r2.replay()
    x_graph = r2.final_state().x.clone()
    g_graph = r2.initial_grad().x.clone()
    x_eager = fwd().x.clone()
    r2.backward(x_grad=seed)
    torch.cuda.synchronize()
    ex, eg = rel_err(x_graph.cpu().numpy(), x_eager.cpu().numpy()), rel_err(g_graph.cpu().numpy(),
                                                                              r2.initial_grad().x.cpu().numpy())
    print(f"replay after a capture with a dead graph cycle vs eager: x {ex:.1e} x0 grad {eg:.1e}")
    assert ex <= 1e-5 and eg <= 1e-4
