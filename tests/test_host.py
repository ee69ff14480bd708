"""Host-side logic that runs without a GPU: kinematics (torch, CPU), chunk and state
tables, the rollout server's block arena and memory accounting, workload generators."""
import numpy as np
import pytest
import torch

from dexdeform_amd.mpm import kinematics as K
from dexdeform_amd.mpm import workloads as W
from dexdeform_amd.mpm.engine import CHUNK, chunk_table, lame


def _rand_quat(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def test_quat_mul_matches_scipy_composition():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(0)
    q, r = _rand_quat(rng, 20), _rand_quat(rng, 20)
    got = K.quat_mul(torch.tensor(q), torch.tensor(r)).numpy()
    Rq = Rotation.from_quat(q[:, [1, 2, 3, 0]])
    Rr = Rotation.from_quat(r[:, [1, 2, 3, 0]])
    exp = (Rr * Rq).as_quat()[:, [3, 0, 1, 2]]  # world-frame increment r after q
    sign = np.sign((got * exp).sum(1, keepdims=True))
    assert np.allclose(got, exp * sign, atol=1e-12)


def test_quaternion_matrix_round_trip_and_scipy():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(1)
    q = _rand_quat(rng, 50)
    M = K.quaternion_to_matrix(torch.tensor(q)).numpy()
    assert np.allclose(M, Rotation.from_quat(q[:, [1, 2, 3, 0]]).as_matrix(), atol=1e-12)
    q2 = K.matrix_to_quaternion(torch.tensor(M)).numpy()
    assert np.allclose(np.abs((q2 * q).sum(1)), 1.0, atol=1e-12)  # same rotation up to sign


def test_axis_angle_matches_scipy_rotvec():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(2)
    aa = rng.normal(size=(40, 3))
    aa[:5] *= 1e-8  # small-angle branch
    M = K.axis_angle_to_matrix(torch.tensor(aa)).numpy()
    assert np.allclose(M, Rotation.from_rotvec(aa).as_matrix(), atol=1e-12)


def test_rigid_body_motion_composes_rotation():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(3)
    B, T = 3, 4
    pos = torch.tensor(rng.normal(size=(B, 3)))
    q = torch.tensor(_rand_quat(rng, B))
    act = torch.tensor(rng.normal(size=(T, B, 6)) * 0.3)
    p2, r2 = K.rigid_body_motion((pos, q), act)
    assert np.allclose(p2.numpy(), (pos[None] + act[..., :3]).numpy())
    exp = (Rotation.from_rotvec(act[..., 3:].reshape(-1, 3).numpy()) *
           Rotation.from_quat(q.numpy()[None].repeat(T, 0).reshape(-1, 4)[:, [1, 2, 3, 0]])).as_quat()[:, [3, 0, 1, 2]]
    got = r2.reshape(-1, 4).numpy()
    assert np.allclose(np.abs((got * exp).sum(1)), 1.0, atol=1e-10)


def test_primitive_poses_ramp():
    pos = torch.zeros(2, 3)
    rot = torch.tensor([[1.0, 0, 0, 0]] * 2)
    a = torch.tensor([[2.0, -0.5, 0.25, 0, 0, 0], [0.1, 0.0, 0.0, 0, 0, 0]])
    scales = torch.tensor([[1e-3] * 6] * 2)
    p, r = K.primitive_poses(pos, rot, a, scales, 10)
    # clamp to [-1, 1], then substep i moves a * (i + 1) / substeps
    assert torch.allclose(p[-1, 0], torch.tensor([1e-3, -0.5e-3, 0.25e-3]))
    assert torch.allclose(p[0, 1], torch.tensor([0.1e-4, 0, 0]))
    assert torch.allclose(r, rot.expand_as(r))


def test_hand_rigid_body_motion_translation_and_unit_rotation():
    rng = np.random.default_rng(4)
    base = torch.eye(4)[None].repeat(2, 1, 1)
    base[:, :3, :3] = K.quaternion_to_matrix(torch.tensor(_rand_quat(rng, 2), dtype=torch.float32))
    acts = torch.tensor(rng.normal(size=(2, 6)) * 0.01, dtype=torch.float32)
    out = K.rigid_body_motion_hand(base, acts, 5)
    assert out.shape == (5, 2, 4, 4)
    assert torch.allclose(out[-1, :, :3, 3], base[:, :3, 3] + acts[:, :3], atol=1e-6)
    R = out[..., :3, :3]
    assert torch.allclose(R @ R.transpose(-1, -2), torch.eye(3).expand_as(R), atol=1e-5)


def test_hinge_tree_kinematics_and_gradients():
    tree, args, q0 = K.synthetic_hand()
    assert args.shape == (22, 4) and np.all((args[:, 0] >= 0.008) & (args[:, 0] <= 0.015))
    a = torch.zeros(22, requires_grad=True)
    pos, rot, qe = tree.poses(q0, a, 8)
    assert pos.shape == (8, 22, 3) and rot.shape == (8, 22, 4)
    assert torch.allclose(torch.linalg.norm(rot, dim=-1), torch.ones(8, 22), atol=1e-5)
    # joint velocity ramp without /substeps (mpm/hand.py:414-415), clamped to limits
    a2 = torch.ones(22)
    q = tree.joint_targets(q0, a2, 8)
    assert torch.allclose(q[3], torch.clamp(q0 + 0.33 * 0.002 * 4, -0.6, 0.6))
    (pos.sum() + rot.sum()).backward()
    assert a.grad is not None and torch.isfinite(a.grad).all()


def test_hinge_tree_fk_two_link_arm_closed_form():
    """Planar two-link arm (hinges about z, link lengths L1, L2) under a rotated, shifted
    base: the tip of link 2 sits at p + Rb (L1 (cos a, sin a, 0) + L2 (cos(a+b), sin(a+b), 0))
    and its orientation is Rb Rz(a + b) (FK order T_parent @ Trans(offset) @ Rot(axis, q)
    @ geom_offset, mpm/hand.py:347-381)."""
    from scipy.spatial.transform import Rotation
    L1, L2 = 0.07, 0.05
    base = np.eye(4)
    Rb = Rotation.from_rotvec([0.4, -0.3, 0.9]).as_matrix()
    base[:3, :3] = Rb
    base[:3, 3] = [0.5, 0.3, 0.45]
    geoms = np.tile(np.eye(4), (2, 1, 1))
    geoms[0, 0, 3] = L1 / 2   # link-1 geom at its midpoint
    geoms[1, 0, 3] = L2       # link-2 geom at its tip
    tree = K.HingeTree([-1, 0], [[0, 0, 0], [L1, 0, 0]], [[0, 0, 1], [0, 0, 1]], geoms, [-3, -3], [3, 3],
                       base, [1.0, 1.0], dtype=torch.float64)
    for a, b in ((0.3, -0.8), (1.2, 0.5), (-2.0, 2.5)):
        pos, quat = tree.fk(torch.tensor([a, b], dtype=torch.float64))
        p0 = base[:3, 3] + Rb @ (L1 / 2 * np.array([np.cos(a), np.sin(a), 0]))
        p1 = base[:3, 3] + Rb @ (L1 * np.array([np.cos(a), np.sin(a), 0]) +
                                 L2 * np.array([np.cos(a + b), np.sin(a + b), 0]))
        assert np.allclose(pos[0].numpy(), p0, rtol=0, atol=1e-12)
        assert np.allclose(pos[1].numpy(), p1, rtol=0, atol=1e-12)
        for link, ang in ((0, a), (1, a + b)):
            R = Rb @ Rotation.from_rotvec([0, 0, ang]).as_matrix()
            x, y, z, w = Rotation.from_matrix(R).as_quat()
            qe = np.array([w, x, y, z])
            qg = quat[link].numpy()
            assert min(np.abs(qg - qe).max(), np.abs(qg + qe).max()) < 1e-12
    # the joint ramp feeds the same FK: substep i of action u sees q0 + u * scale * (i + 1)
    q0 = torch.tensor([0.1, 0.2], dtype=torch.float64)
    pos, rot, qe = tree.poses(q0, torch.tensor([0.5, -0.25], dtype=torch.float64), 3)
    assert torch.allclose(qe, q0 + torch.tensor([0.5, -0.25], dtype=torch.float64) * 3)
    assert torch.allclose(pos[1], tree.fk(q0 + torch.tensor([0.5, -0.25], dtype=torch.float64) * 2)[0])


def test_chunk_table_covers_every_particle_once():
    counts = [1, 255, 256, 257, 1000, 0]
    tab = chunk_table(counts)
    assert np.all(tab[:, 2] - tab[:, 1] <= CHUNK) and np.all(tab[:, 2] > tab[:, 1])
    for s, n in enumerate(counts):
        rows = tab[tab[:, 0] == s]
        covered = np.concatenate([np.arange(b, e) for _, b, e, _ in rows]) if len(rows) else np.zeros(0)
        assert np.array_equal(np.sort(covered), np.arange(n))


def test_lame_parameters():
    mu, lam = lame(5e3, 0.2)
    assert np.isclose(mu, 5e3 / 2.4) and np.isclose(lam, 5e3 * 0.2 / (1.2 * 0.6))


def test_block_arena_alloc_release_merge():
    from dexdeform_amd.mpm.serve import BlockArena
    ar = BlockArena(100)
    a = ar.alloc(30)
    b = ar.alloc(50)
    assert (a, b) == (0, 30) and ar.alloc(30) is None
    ar.release(a, 30)
    c = ar.alloc(10)
    assert c == 0
    ar.release(b, 50)
    ar.release(c, 10)
    assert ar.free == [(0, 100)]


def test_request_bytes_scales_with_checkpointing():
    from dexdeform_amd.mpm.serve import Request, request_bytes
    sc = W.plain_scene(n_particles=2000, grid=32)
    T = 40
    pos = np.repeat(sc.body_pos[None], T + 1, 0)
    rot = np.repeat(sc.body_rot[None], T + 1, 0)
    full = Request(0, sc.spec, sc.x0, pos, rot, sc.x0)
    ck = Request(1, sc.spec, sc.x0, pos, rot, sc.x0, checkpoint_every=10)
    assert request_bytes(ck) < request_bytes(full)


def test_workloads_shapes_and_order():
    sc = W.plain_scene(n_particles=3000)
    assert sc.x0.shape == (3000, 3) and sc.nb == 2
    lo, hi = np.array([0.4, 0.05, 0.4]), np.array([0.6, 0.25, 0.6])
    assert np.all(sc.x0 >= lo) and np.all(sc.x0 <= hi)
    assert np.isclose(sc.spec.gravity[1], -9.8 * 30)
    scenes = W.batched_scenes(n_scenes=4, seed=3, n_range=(100, 200))
    assert len(scenes) == 4 and all(100 <= s.n <= 200 for s in scenes)
    assert all(2 <= s.nb <= 22 for s in scenes)
    x = W.ellipsoid_particles(500, [0.5, 0.2, 0.5], [0.12, 0.08, 0.1], seed=3)
    r = ((x - [0.5, 0.2, 0.5]) / [0.12, 0.08, 0.1]) ** 2
    assert np.all(r.sum(1) <= 1 + 1e-5)


def test_scene_validation_rejects_bad_grids_and_shapes():
    """Grid dims must be multiples of 4 in [8, 1024] (stencil bases are packed in 10 bits
    per axis), at most DDM_MAX_BODIES bodies, per-particle arrays of the right shape."""
    import pytest
    from dexdeform_amd.mpm import workloads as W
    sc = W.plain_scene(n_particles=100)
    sc.spec.validate()
    for gd in ((64, 64, 62), (4, 64, 64), (1028, 64, 64)):
        sc.spec.grid_dim = gd
        with pytest.raises(ValueError):
            sc.spec.validate()
    sc.spec.grid_dim = (64, 64, 64)
    sc.spec.mass = sc.spec.mass[:-1]
    with pytest.raises(ValueError):
        sc.spec.validate()


def test_chunk_size_for_small_and_large_launches():
    from dexdeform_amd.mpm.engine import chunk_size_for, chunk_table
    assert chunk_size_for(10_000) == 64
    assert chunk_size_for(30_000) == 128
    assert chunk_size_for(50_000) == 256
    assert chunk_size_for(2_000_000) == 256
    for c in (64, 128, 256):
        tab = chunk_table([1000, 0, 300], chunk=c)
        assert (tab[:, 2] - tab[:, 1]).max() <= c
        for s, n in ((0, 1000), (2, 300)):
            rows = tab[tab[:, 0] == s]
            assert rows[0, 1] == 0 and rows[-1, 2] == n and np.array_equal(rows[1:, 1], rows[:-1, 2])


def test_spatial_order_groups_blocks_then_cells():
    """spatial_order sorts by 4^3 cell block, then by cell within the block; a random
    permutation of ordered particles is detected as unordered and sorted back into
    block/cell order (ties keep their input order)."""
    from dexdeform_amd.mpm import workloads as W
    dx = 1 / 64
    x = W.box_particles(5000, [0.3, 0.1, 0.3], [0.5, 0.3, 0.5], 3, dx)
    assert W.spatially_ordered(x, dx)
    xs = x[np.random.default_rng(0).permutation(len(x))]
    assert not W.spatially_ordered(xs, dx)
    k = W.spatial_key(xs[W.spatial_order(xs, dx)], dx)
    assert np.all(k[1:] >= k[:-1])
    c = np.floor(xs / dx).astype(np.int64)
    blk = (c >> 2)[W.spatial_order(xs, dx)]
    # consecutive particles change block only when the (x, y, z) block index increases
    bk = (blk[:, 0] * 1024 + blk[:, 1]) * 1024 + blk[:, 2]
    assert np.all(bk[1:] >= bk[:-1])
    # needs_sort: random order is flagged; sorted input, and sorted input with a few
    # neighbours swapped (under 10 % of adjacent pairs out of order), is not
    assert W.needs_sort(xs, dx)
    assert not W.needs_sort(x, dx)
    xn = x.copy()
    i = np.arange(0, len(x) - 1, 40)
    xn[i], xn[i + 1] = x[i + 1], x[i]
    assert not W.needs_sort(xn, dx)
    assert not W.needs_sort(x[:1], dx)


def test_tick_graph_buckets():
    """Captured server ticks are sized to powers of two (at least the minimum)."""
    from dexdeform_amd.mpm.serve import _bucket
    assert _bucket(1, 64) == 64 and _bucket(64, 64) == 64 and _bucket(65, 64) == 128
    assert _bucket(1000, 64) == 1024 and _bucket(3, 4) == 4 and _bucket(5, 4) == 8


def test_tick_table_layout_is_compact_aligned_and_disjoint():
    """A tick key's tables: every part 64-byte aligned, parts disjoint and in order,
    capacities as the launch sizes need (records 168 B and state records 48 B per scene,
    16 B per chunk), and the span a few KB for one small request."""
    from dexdeform_amd import _lib
    from dexdeform_amd.mpm.serve import _TickTables
    for key in [(1, 64, 0, 0), (0, 0, 1, 64), (1, 64, 1, 64), (4, 256, 2, 128), (256, 4096, 256, 4096)]:
        off, span, cap = _TickTables._offsets(key)
        parts = sorted(off.items(), key=lambda kv: kv[1])
        for (k, o), nxt in zip(parts, parts[1:] + [(None, span)]):
            assert o % 64 == 0
            assert o + cap[k] <= nxt[1]
        fs, fc, bs, bc = key
        for kind, sc, ch in (("F", fs, fc), ("B", bs, bc)):
            if not sc:
                assert not any(k[0] == kind for k in off)
                continue
            assert cap[kind, "recs"] == _lib.SCENE_DTYPE.itemsize * sc
            assert cap[kind, "cur"] == cap[kind, "nxt"] == 48 * sc
            assert cap[kind, "chunks"] == 16 * ch and cap[kind, "live"] == 8
            if kind == "B":
                assert cap["B", "gcur"] == cap["B", "gnxt"] == 48 * sc
    assert _TickTables._offsets((1, 64, 1, 64))[1] < 4096


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus N > 1 (or a multi-rank launch) exits with status 2
    and says so, before touching a device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for extra_env, args in (({}, ["--gpus", "2"]), ({"WORLD_SIZE": "2"}, [])):
        env = dict(os.environ, **extra_env)
        r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), *args], capture_output=True, text=True,
                           env=env, timeout=120)
        assert r.returncode == 2
        assert "one GPU" in r.stderr


def test_server_submit_validates_and_admission_rolls_back(monkeypatch):
    """submit() rejects a checkpoint interval that does not divide the horizon (and empty
    horizons) before queueing; if building a request's device state fails at admission,
    its arena blocks and memory are given back and the request stays queued."""
    from dexdeform_amd.mpm import serve as S
    sc = W.plain_scene(n_particles=500, grid=32)
    T = 12
    pos = np.repeat(sc.body_pos[None], T + 1, 0)
    rot = np.repeat(sc.body_rot[None], T + 1, 0)
    srv = S.RolloutServer.__new__(S.RolloutServer)  # host bookkeeping only, no device
    srv.memory_budget, srv.full_tape_bytes = 1 << 30, 0
    srv.arena, srv.waiting, srv.active, srv.max_active = S.BlockArena(4096), [], [], 8
    srv.used_bytes, srv.host_s = 0, {"admit": 0.0}
    for bad in (5, 0, -2):
        with pytest.raises(ValueError):
            srv.submit(S.Request(0, sc.spec, sc.x0, pos, rot, sc.x0, checkpoint_every=bad))
    with pytest.raises(ValueError):
        srv.submit(S.Request(1, sc.spec, sc.x0, pos[:1], rot[:1], sc.x0))
    assert srv.waiting == []
    srv.submit(S.Request(2, sc.spec, sc.x0, pos, rot, sc.x0, checkpoint_every=4))

    def boom(*a, **k):
        raise RuntimeError("device allocation failed")
    monkeypatch.setattr(S, "_Active", boom)
    with pytest.raises(RuntimeError):
        srv._admit(now=1.0)
    assert srv.used_bytes == 0 and srv.arena.free == [(0, 4096)] and len(srv.waiting) == 1 and srv.active == []
