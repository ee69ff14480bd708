"""Deterministic sorted-reduction mode (Engine(deterministic=True), options bit
DDM_DETERMINISTIC): particles binned by stencil-base node each substep, grid nodes summed
in a fixed order instead of through atomics.

* parity with the CPU oracle at the default mode's tolerances (one forward substep
  <= 1e-5 per array; one adjoint substep <= max(1e-3, 2 x the oracle's FP32-vs-FP64
  discrepancy) per array);
* bitwise reproducibility: the same substeps run twice give identical bits, forward
  and adjoint (the default mode's atomics make the last bits differ between runs);
* bitwise identity of a scene alone and in a batch (sums run in scene-local order);
* bitwise identity of the taped and recomputing adjoints (the restored grid is the
  recomputed one, bit for bit), also when snapshots overflow.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import initial_state, oracle_scene, poses_from_actions, rel_err
from tests.test_gpu_parity import _fp32_noise_bound, _np, _random_grads, _to_buf
from tests.test_gpu_serve import _poses, _scene
from dexdeform_amd.mpm import workloads as W
from dexdeform_amd.mpm.engine import Engine
from dexdeform_amd.mpm.rollout import Rollout

pytestmark = pytest.mark.gpu

SUB = 10


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def plain():
    sc = W.plain_scene(n_particles=10000)
    acts = W.plain_actions(n_env_steps=5)
    scales = np.tile(np.float32([SUB * sc.spec.dt] * 3 + [0.0] * 3), (2, 1))
    P, R = poses_from_actions(sc, acts, SUB, scales)
    osim = O.OracleSim(oracle_scene(sc), np.float32)
    states = [initial_state(sc).astype(np.float32)]
    for t in range(1, 21):
        states.append(osim.substep(states[-1], P[t], R[t]))
    return dict(sc=sc, P=P, R=R, osim=osim, states=states)


def _bits_equal(a, b):
    return all(np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32)) for x, y in zip(a, b))


def test_det_forward_matches_oracle_and_repeats_bitwise(plain):
    sc = plain["sc"]
    sc.spec.adjoint_mode = 0
    eng = Engine([sc.spec], device=_dev(), deterministic=True)
    st = plain["states"][20]
    ref = plain["osim"].substep(st, plain["P"][21], plain["R"][21])
    cur, nxt = eng.new_state(), eng.new_state()
    _to_buf(st, cur)
    nxt.body_pos.copy_(torch.from_numpy(np.ascontiguousarray(plain["P"][21], np.float32)))
    nxt.body_rot.copy_(torch.from_numpy(np.ascontiguousarray(plain["R"][21], np.float32)))
    eng.forward(cur, nxt)
    torch.cuda.synchronize()
    for name, g, r in zip("xvFC", _np(nxt), (ref.x, ref.v, ref.F, ref.C)):
        e = rel_err(g, r)
        print(f"deterministic forward {name}: rel err vs oracle {e:.2e}")
        assert e <= 1e-5, (name, e)
    # ten substeps, twice, from the same state: identical bits
    runs = []
    for _ in range(2):
        a, b = eng.new_state(), eng.new_state()
        _to_buf(plain["states"][10], a)
        for t in range(10, 20):
            b.body_pos.copy_(torch.from_numpy(np.ascontiguousarray(plain["P"][t + 1], np.float32)))
            b.body_rot.copy_(torch.from_numpy(np.ascontiguousarray(plain["R"][t + 1], np.float32)))
            eng.forward(a, b)
            a, b = b, a
        torch.cuda.synchronize()
        runs.append(_np(a))
    assert _bits_equal(runs[0], runs[1]), "deterministic forward substeps differ between runs"


def test_det_backward_matches_oracle_and_repeats_bitwise(plain):
    sc = plain["sc"]
    sc.spec.adjoint_mode = 0
    st = plain["states"][20]
    nx = plain["osim"].substep(st, plain["P"][21], plain["R"][21])
    g_next = _random_grads(sc.n, sc.nb)
    ref = _fp32_noise_bound(sc, st, nx, g_next, 0)
    g32, gn32 = ref[np.float32]
    g64, gn64 = ref[np.float64]
    eng = Engine([sc.spec], device=_dev(), deterministic=True)
    cur, nxt, gnxt = eng.new_state(), eng.new_state(), eng.new_grad()
    _to_buf(st, cur)
    _to_buf(nx, nxt)
    outs = []
    for _ in range(2):
        _to_buf(O.State(g_next.x, g_next.v, g_next.F, g_next.C, g_next.body_pos, g_next.body_rot), gnxt)
        gcur = eng.new_grad()
        eng.backward(cur, nxt, gcur, gnxt)
        torch.cuda.synchronize()
        outs.append(_np(gcur) + [gcur.body_pos.cpu().numpy(), gcur.body_rot.cpu().numpy(),
                                 gnxt.body_pos.cpu().numpy(), gnxt.body_rot.cpu().numpy()])
    got = outs[0]
    r32 = [g32.x, g32.v, g32.F, g32.C, g32.body_pos, g32.body_rot, gn32.body_pos, gn32.body_rot]
    r64 = [g64.x, g64.v, g64.F, g64.C, g64.body_pos, g64.body_rot, gn64.body_pos, gn64.body_rot]
    names = ["x", "v", "F", "C", "body_pos", "body_rot", "next_body_pos", "next_body_rot"]
    for name, g, a, b in zip(names, got, r32, r64):
        e = rel_err(g, a)
        tol = max(1e-3, 2 * rel_err(a, b))
        print(f"deterministic adjoint {name}: rel err vs fp32 oracle {e:.3e} (tol {tol:.1e})")
        assert e <= tol, (name, e)
    assert _bits_equal(outs[0], outs[1]), "deterministic adjoint substeps differ between runs"


def _det_rollout(scenes, T, K=None, tape=False, tape_blocks=None, seeds=None):
    dev = _dev()
    eng = Engine([s.spec for s in scenes], device=dev, deterministic=True)
    ro = Rollout(eng, T, checkpoint_every=K, tape=tape, tape_blocks=tape_blocks)
    ro.set_initial(torch.tensor(np.concatenate([s.x0 for s in scenes]), device=dev))
    P = [_poses(s, T, sd) for s, sd in zip(scenes, seeds)]
    ro.set_poses(torch.tensor(np.concatenate([p[0] for p in P], 1), device=dev),
                 torch.tensor(np.concatenate([p[1] for p in P], 1), device=dev))
    xT = ro.forward().x.clone()
    ro.backward(x_grad=torch.full_like(xT, 1e-4))
    torch.cuda.synchronize()
    return ro, xT.cpu().numpy(), ro.pos_grad.cpu().numpy(), ro.initial_grad().x.cpu().numpy()


def test_det_scene_alone_equals_batched_bitwise():
    scenes = [_scene(3000, 48, seed=2), _scene(6000, 64, seed=3, nb=5), _scene(1500, 32, seed=4, nb=2)]
    seeds = [100, 101, 102]
    ro, x, pg, g0 = _det_rollout(scenes, 12, K=4, seeds=seeds)
    eng = ro.engine
    for i, sc in enumerate(scenes):
        _, xa, pga, g0a = _det_rollout([sc], 12, K=4, seeds=[seeds[i]])
        ps, bs = eng.scene_slice(i), eng.body_slice(i)
        same = _bits_equal([x[ps], pg[:, bs], g0[ps]], [xa, pga, g0a])
        print(f"scene {i}: alone vs batched bitwise identical: {same}; x0 grad rel diff "
              f"{rel_err(g0[ps], g0a):.1e}")
        assert same, f"scene {i} differs alone and batched"


@pytest.mark.parametrize("tape_blocks", [None, 4, "mixed"])
def test_det_taped_equals_recompute_bitwise(tape_blocks):
    scenes = [_scene(3000, 48, seed=2), _scene(6000, 64, seed=3, nb=5)]
    seeds = [100, 101]
    _, x_r, pg_r, g0_r = _det_rollout(scenes, 20, K=10, seeds=seeds)
    tb = [1728, 4] if tape_blocks == "mixed" else tape_blocks
    ro, x_t, pg_t, g0_t = _det_rollout(scenes, 20, K=10, tape=True, tape_blocks=tb, seeds=seeds)
    expect = {None: 0, 4: 2 * ro.K, "mixed": ro.K}[tape_blocks]
    assert ro.tape_recomputed() == expect
    same = _bits_equal([x_t, pg_t, g0_t], [x_r, pg_r, g0_r])
    print(f"tape {tape_blocks}: taped vs recompute bitwise identical: {same}")
    assert same
