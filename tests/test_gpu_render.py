"""The renderer on the GPU (libdexdeform_mpm's particle_sdf and render, called through the
C ABI and through the drop-in Renderer) against the CPU oracle (oracle/render.py).

* particle_sdf on 3,000 particles in a 64^3 voxel box (bake size 7, inv_dx 150): the packed
  int64 volume, the unpacked sdf and colour and the smoothed sdf equal the oracle's bit
  for bit (integer min; the same FP32 operations in the same order);
* render of that volume with the ground and three bodies (box, capsule, sphere), two
  bounces: image sizes that are not multiples of the 8 x 8 tile, spp 5 (one pass of
  sample waves) and spp 20 (two passes); colour and depth equal the oracle's on at least
  99.5% of the pixels to 1e-5 and the mean difference is below 1e-5 (the operations are the
  same; a pixel can differ only where a rounding difference flips a hit test);
* the drop-in Renderer on a simulator state: bake box, [H, W, 4] output, seed advance,
  and its image against the oracle's on the same inputs;
* argument errors are rejected.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import render as R
from dexdeform_amd import _lib
from dexdeform_amd.mpm.renderer import EMPTY_VOXEL, Renderer, intrinsic, lookat

pytestmark = pytest.mark.gpu


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def _t(a, dtype=torch.float32):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=_dev())


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _particles(n=3000, seed=0):
    rng = np.random.default_rng(seed)
    # a flat cylinder on the ground (the folding scene's shape) plus a small cube above it
    k = n * 2 // 3
    r = 0.08 * np.sqrt(rng.uniform(0, 1, k))
    a = rng.uniform(0, 2 * np.pi, k)
    cyl = np.stack([0.5 + r * np.cos(a), 0.05 + rng.uniform(0, 0.02, k), 0.5 + r * np.sin(a)], 1)
    cube = rng.uniform([0.46, 0.09, 0.46], [0.54, 0.15, 0.54], size=(n - k, 3))
    x = np.concatenate([cyl, cube]).astype(np.float32)
    color = np.where(np.arange(n) < k, (200 << 16) + (120 << 8) + 60, (60 << 16) + (160 << 8) + 220).astype(np.int32)
    return x, color


def _bake_gpu(x, color, lo, bake, grid, inv_dx):
    lib = _lib.load()
    V = int(np.prod(grid))
    vol = torch.full((V,), EMPTY_VOXEL, dtype=torch.int64, device=_dev())
    sdf = torch.zeros(V, device=_dev())
    tmp = torch.zeros(V, device=_dev())
    rgb = torch.zeros(V, 3, device=_dev())
    xd, cd, lod = _t(x), _t(color, torch.int32), _t(lo)
    hid = _t(lo + np.float32(grid) / np.float32(inv_dx))
    _lib.check(lib.particle_sdf(_lib.ptr(vol), _lib.ptr(xd), _lib.ptr(cd), _lib.ptr(lod), _lib.ptr(hid), bake,
                                _i3(grid), inv_dx, _lib.ptr(sdf), _lib.ptr(rgb), _lib.ptr(tmp), len(x),
                                _lib.stream_ptr()), "particle_sdf")
    torch.cuda.synchronize()
    g = tuple(grid)
    return vol.cpu().numpy().reshape(g), sdf.cpu().numpy().reshape(g), rgb.cpu().numpy().reshape(g + (3,))


GRID = (64, 64, 64)
INV_DX = 150.0


def _scene():
    x, color = _particles()
    lo = (np.floor(x * np.float32(INV_DX) - 6) / np.float32(INV_DX)).min(0).astype(np.float32)
    return x, color, lo


def test_particle_sdf_bit_exact():
    x, color, lo = _scene()
    print(f"particle_sdf: {len(x)} particles, grid {GRID}, bake 7", flush=True)
    vol, sdf, rgb = _bake_gpu(x, color, lo, 7, GRID, INV_DX)
    rvol, rsdf, rrgb = R.particle_sdf(x, color, lo, 7, GRID, INV_DX)
    assert (rvol != EMPTY_VOXEL).sum() > 20000
    assert np.array_equal(vol, rvol), int((vol != rvol).sum())
    assert np.array_equal(rgb, rrgb)
    assert np.array_equal(sdf, rsdf), float(np.abs(sdf - rsdf).max())


def _bodies():
    pos = np.float32([[0.42, 0.12, 0.5], [0.58, 0.16, 0.47], [0.5, 0.25, 0.56]])
    rot = np.float32([[0.92388, 0.0, 0.38268, 0.0], [0.96593, 0.25882, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0]])
    tfsr = np.float32([[0, 0.9, 0, 0.002], [1, 0.9, 0, 0], [1, 0.9, 0, 0]])  # box, capsule, sphere
    args = np.float32([[0.02, 0.03, 0.025, 0], [0.015, 0.06, 0, 0], [0.03, 0.0, 0, 0]])
    return pos, rot, tfsr, args


def _render_both(H, W, spp, seed, depth=2, shape=True, n_bodies=3):
    x, color, lo = _scene()
    vol, sdf, rgb = R.particle_sdf(x, color, lo, 7, GRID, INV_DX)
    hi = (lo.astype(np.float64) + np.array(GRID) / INV_DX).astype(np.float32)
    pos, rot, tfsr, args = _bodies()
    Rm, T = lookat((0.5, 0.12, 0.5), np.pi / 8, -np.pi / 2, 0.45)
    K = intrinsic((H, W))
    light = np.float32([1.0, 1.0, 0.0])
    print(f"render: {H}x{W}, spp {spp}, depth {depth}, bodies {n_bodies}, shape {shape}", flush=True)
    lib = _lib.load()
    col = torch.zeros(H * W, 3, device=_dev())
    dep = torch.zeros(H * W, device=_dev())
    d_sdf, d_rgb, d_lo, d_hi = _t(sdf), _t(rgb), _t(lo), _t(hi)
    d_pos, d_rot, d_tfsr, d_args = _t(pos), _t(rot), _t(tfsr), _t(args)
    d_R, d_T, d_K, d_L = _t(np.float32(Rm)), _t(np.float32(T)), _t(np.float32(K)), _t(light)
    _lib.check(lib.render(_lib.ptr(d_sdf), _lib.ptr(d_lo), _lib.ptr(d_hi), _lib.ptr(d_rgb), _lib.ptr(d_pos),
                          _lib.ptr(d_rot), _lib.ptr(d_tfsr), _lib.ptr(d_args), _lib.ptr(d_R), _lib.ptr(d_T),
                          _lib.ptr(d_K), _lib.ptr(col), _lib.ptr(dep), 0.414, _i3(GRID), n_bodies, shape,
                          _i3((H, W, 3)), depth, spp, 0.0, seed, _lib.ptr(d_L), _lib.stream_ptr()), "render")
    torch.cuda.synchronize()
    rc, rd = R.render(sdf, lo, hi, rgb, pos, rot, tfsr, args, np.float32(Rm), np.float32(T), np.float32(K), (H, W),
                      sdf_threshold=0.414, n_bodies=n_bodies, visualize_shape=shape, max_ray_depth=depth, spp=spp,
                      ground_height=0.0, seed=seed, light_direction=light)
    return col.cpu().numpy().reshape(H, W, 3), dep.cpu().numpy().reshape(H, W), rc, rd


def _compare(name, got, ref):
    diff = np.abs(got - ref)
    per_pixel = diff.reshape(diff.shape[0], diff.shape[1], -1).max(-1)
    exact = float((per_pixel == 0).mean())
    close = float((per_pixel <= 1e-5).mean())
    print(f"  {name}: bit-equal pixels {exact:.4f}, within 1e-5 {close:.4f}, mean |diff| {diff.mean():.2e}, "
          f"max {diff.max():.2e}", flush=True)
    assert close >= 0.995 and diff.mean() <= 1e-5, (name, close, diff.mean())


@pytest.mark.parametrize("H,W,spp", [(30, 45, 5), (16, 24, 20)])
def test_render_matches_oracle(H, W, spp):
    col, dep, rc, rd = _render_both(H, W, spp, seed=7)
    assert np.isfinite(col).all() and np.isfinite(dep).all()
    near = dep < 50
    assert near.mean() > 0.5, "most pixels should see the scene, not the sky"
    _compare("colour", col, rc)
    _compare("depth", np.where(near, dep, 0), np.where(rd < 50, rd, 0))


def test_render_without_bodies_or_volume():
    col, dep, rc, rd = _render_both(20, 20, 3, seed=1, depth=1, shape=False, n_bodies=0)
    _compare("ground only colour", col, rc)
    _compare("ground only depth", dep, rd)


def test_render_rejects_bad_arguments():
    lib = _lib.load()
    buf = torch.zeros(4096, device=_dev())
    p = _lib.ptr(buf)
    s = _lib.stream_ptr()
    args = lambda nb, spp, img: (p, p, p, p, p, p, p, p, p, p, p, p, p, 0.414, _i3((4, 4, 4)), nb, True,  # noqa: E731
                                 _i3(img), 2, spp, 0.0, 0, p, s)
    assert lib.render(*args(65, 1, (4, 4, 3))) == _lib.DDM_ETOOMANYBODIES
    assert lib.render(*args(1, 0, (4, 4, 3))) == _lib.DDM_EINVAL
    assert lib.render(*args(1, 1, (0, 4, 3))) == _lib.DDM_EINVAL
    assert lib.particle_sdf(p, p, p, p, p, 0, _i3((4, 4, 4)), 150.0, p, p, p, 1, s) == _lib.DDM_EINVAL
    torch.cuda.synchronize()


def test_dropin_renderer_on_a_simulator_state():
    from dexdeform_amd.mpm.simulator import MPMSimulator
    x, color, _ = _scene()
    n = len(x)
    sim = MPMSimulator(3, cfg=dict(n_particles=n, max_steps=2), device=_dev())
    pos, rot, tfsr, args = _bodies()
    sim.init_bodies(types=tfsr[:, 0], softness=tfsr[:, 2], mu=tfsr[:, 1], round=tfsr[:, 3], args=args,
                    action_scales=np.zeros((3, 6)), pos=pos, rot=rot)
    sim.states[0].x.upload(x)
    sim.set_color(color)
    H, W = 24, 32
    ren = Renderer(image_res=(H, W), voxel_res=GRID, spp=4, device=_dev())
    ren.lookat((0.5, 0.12, 0.5), np.pi / 8, -np.pi / 2, 0.45)
    img = ren.render(sim, sim.states[0])
    assert img.shape == (H, W, 4) and ren.render_seed == 1
    lo = ren.bbox_min.cpu().numpy().reshape(3)
    np.testing.assert_array_equal(lo, (np.floor(x * np.float32(INV_DX) - 6) / np.float32(INV_DX)).min(0))
    vol, sdf, rgb = R.particle_sdf(x, color, lo, 7, GRID, INV_DX)
    hi = ren.bbox_max.cpu().numpy().reshape(3)
    rc, rd = R.render(sdf, lo, hi, rgb, pos, rot, tfsr, args, ren.camera_rot.cpu().numpy(),
                      ren.camera_pos.cpu().numpy(), ren.camera_intrinsic.cpu().numpy(), (H, W), n_bodies=3, spp=4,
                      max_ray_depth=2, ground_height=sim.get_ground_height() * sim.dx, seed=0)
    _compare("drop-in colour", img[..., :3], rc)
    img2 = ren.render(sim, sim.states[0], primitive=0, shape=0, spp=2)
    assert ren.render_seed == 2 and not np.array_equal(img2, img)


def test_texture_volume_helpers():
    """create_volume / destroy_volume (the reference's texture helpers): the float4 host
    volume is copied to the device (gfx950 has no texture objects: the handle is 0) and
    released."""
    lib = _lib.load()
    data = np.arange(5 * 4 * 3 * 4, dtype=np.float32)
    vol = lib.create_volume(data.ctypes.data, 5, 4, 3)
    assert vol.volume and vol.texture == 0
    back = np.zeros_like(data)
    _lib.check(lib.cuda_download(back.ctypes.data, vol.volume, data.nbytes), "cuda_download")
    np.testing.assert_array_equal(back, data)
    assert lib.destroy_volume(vol) == 0
    bad = lib.create_volume(data.ctypes.data, 0, 4, 3)
    assert not bad.volume
    torch.cuda.synchronize()
