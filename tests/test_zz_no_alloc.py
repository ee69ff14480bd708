"""The allocation-free forward substep, ddm_substep_forward_scratch, against the allocating
entry point ddm_substep_forward, both called through the C ABI on the same inputs.

Two calls in a row share the library's one grid workspace on the GPU: a 10k-particle scene
on a 64^3 grid (4096 node blocks), then a smaller one (2k particles, 32^3 grid, 512 blocks),
so the second call runs on whatever the first left in the workspace. Each output (x, v, F,
C of the next state) must match ddm_substep_forward's to 1e-5 relative L2, the bar of a
forward substep (FP32 atomic accumulation order is the only difference between the two
launches). Sizes beyond the workspace and the deterministic mode are rejected on the host.
"""
import ctypes

import numpy as np
import pytest
import torch

from dexdeform_amd import _lib
from dexdeform_amd.mpm import workloads as W
from dexdeform_amd.mpm.engine import Engine
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def _scratch_batch(eng):
    """The engine's launch record with its grid workspace pointers removed."""
    b = _lib.DdmBatch.from_buffer_copy(eng._batch)
    b.grid_mmv = b.grid_v = b.grid_dv = b.grid_dmmv = None
    b.block_flags = b.block_list = b.block_count = None
    return b


def _one_call(n_particles, grid, seed):
    dev = _dev()
    sc = W.plain_scene(n_particles=n_particles, grid=grid, seed=seed)
    print(f"scratch forward: {n_particles} particles, {grid}^3 grid", flush=True)
    eng = Engine([sc.spec], device=dev, with_grad=False)
    rng = np.random.default_rng(seed)
    cur = eng.new_state()
    cur.x.copy_(torch.from_numpy(sc.x0))
    cur.v.copy_(torch.from_numpy((rng.normal(size=(sc.n, 3)) * 0.2).astype(np.float32)))
    cur.F.copy_(torch.from_numpy((np.eye(3)[None] + rng.normal(size=(sc.n, 3, 3)) * 0.02).astype(np.float32)))
    cur.body_pos.copy_(torch.from_numpy(sc.body_pos))
    cur.body_rot.copy_(torch.from_numpy(sc.body_rot))
    ref, got = eng.new_state(), eng.new_state()
    npos = torch.from_numpy(sc.body_pos + np.float32([0.0, -2e-4, 0.0]))
    for nxt in (ref, got):
        nxt.body_pos.copy_(npos)
        nxt.body_rot.copy_(torch.from_numpy(sc.body_rot))
    lib = _lib.load()
    cur_tab = eng.table_of(cur)
    s = _lib.stream_ptr()
    _lib.check(lib.ddm_substep_forward(ctypes.byref(eng._batch), cur_tab.data_ptr(), eng.table_of(ref).data_ptr(), s),
               "ddm_substep_forward")
    b = _scratch_batch(eng)
    assert b.n_blocks <= _lib.DDM_SCRATCH_BLOCKS
    _lib.check(lib.ddm_substep_forward_scratch(ctypes.byref(b), cur_tab.data_ptr(), eng.table_of(got).data_ptr(), s),
               "ddm_substep_forward_scratch")
    torch.cuda.synchronize()
    moved = float((ref.x - cur.x).abs().max())
    assert moved > 0.0, "the substep moved no particle"
    for name in "xvFC":
        r = getattr(ref, name).cpu().numpy()
        g = getattr(got, name).cpu().numpy()
        assert np.isfinite(g).all(), name
        e = rel_err(g, r)
        print(f"  {name}: rel err vs ddm_substep_forward {e:.2e}", flush=True)
        assert e <= 1e-5, (name, e)


def test_scratch_forward_matches_allocating_forward_twice_in_a_row():
    _one_call(10000, 64, seed=0)
    _one_call(2000, 32, seed=1)


def test_scratch_forward_rejects_oversized_and_deterministic_launches():
    dev = _dev()
    sc = W.plain_scene(n_particles=1000, grid=32)
    eng = Engine([sc.spec], device=dev, with_grad=False)
    cur, nxt = eng.new_state(), eng.new_state()
    lib = _lib.load()
    b = _scratch_batch(eng)
    b.n_blocks = _lib.DDM_SCRATCH_BLOCKS + 1
    st = lib.ddm_substep_forward_scratch(ctypes.byref(b), eng.table_of(cur).data_ptr(), eng.table_of(nxt).data_ptr(),
                                         _lib.stream_ptr())
    assert st == _lib.DDM_ESCRATCH
    b = _scratch_batch(eng)
    b.options = _lib.DDM_DETERMINISTIC
    st = lib.ddm_substep_forward_scratch(ctypes.byref(b), eng.table_of(cur).data_ptr(), eng.table_of(nxt).data_ptr(),
                                         _lib.stream_ptr())
    assert st == _lib.DDM_EINVAL
    torch.cuda.synchronize()
