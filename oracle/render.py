"""ORACLE — test infrastructure only.

Python driver for the renderer restatement in ``mpm_oracle.cpp`` (FP32): the particle
volume bake (``particle_sdf``, mpm/csrc/integrator.cu:1438-1525,1908-1933) and the path
tracer (``render``, integrator.cu:1111-1436,1863-1906; volume queries array.h:6-75), with
the random numbers of the HIP renderer (Philox4x32-10, one stream per pixel and sample)
in place of the reference's curand XORWOW stream.

Only ``tests/`` may import this module. Pins (tests/test_render_oracle.py): the Philox
block function against the published known-answer vectors and rocRAND's implementation;
the uniform mapping's range; sin/cos against numpy; the bake against a brute-force
minimum over particles and hand-computed voxels; the smoothing against closed forms;
first-hit depths of the ground plane, a sphere and a box against ray-intersection
formulas; the ground's checkerboard colour; the volume march against the analytic
distance to a baked ball.
"""
from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from oracle import oracle as _O

EMPTY_VOXEL = (1 << 33) - 1


def _lib():
    return _O._lib(False)


def _p(a):
    return _O._p(a)


def philox4x32_10(counter, key):
    """Philox4x32-10 block function: 4 counter words, 2 key words -> 4 output words."""
    c = np.ascontiguousarray(counter, np.uint32).reshape(4)
    k = np.ascontiguousarray(key, np.uint32).reshape(2)
    out = np.zeros(4, np.uint32)
    fn = _lib().orc_philox4x32_10
    fn.restype = None
    fn(_p(c), _p(k), _p(out))
    return out


def uniforms(seed, pixel, sample, n):
    """The first n uniforms in (0, 1] of the (pixel, sample) stream."""
    out = np.zeros(n, np.float32)
    fn = _lib().orc_uniforms
    fn.restype = None
    fn(ctypes.c_uint32(seed), ctypes.c_uint32(pixel), ctypes.c_uint32(sample), ctypes.c_int(n), _p(out))
    return out


def sincos(phi):
    phi = np.ascontiguousarray(phi, np.float32).reshape(-1)
    s = np.zeros_like(phi)
    c = np.zeros_like(phi)
    fn = _lib().orc_sincos
    fn.restype = None
    fn(ctypes.c_int(len(phi)), _p(phi), _p(s), _p(c))
    return s, c


def particle_sdf(x, color, bbox_min, bake_size, grid_dim, inv_dx):
    """Returns (volume int64 [gx, gy, gz], sdf float32 [gx, gy, gz], rgb float32 [gx, gy, gz, 3])."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    color = np.ascontiguousarray(color, np.int32).reshape(-1)
    lo = np.ascontiguousarray(bbox_min, np.float32).reshape(3)
    grid = np.ascontiguousarray(grid_dim, np.int32).reshape(3)
    gx, gy, gz = (int(g) for g in grid)
    vol = np.full(gx * gy * gz, EMPTY_VOXEL, np.int64)
    sdf = np.zeros(gx * gy * gz, np.float32)
    rgb = np.zeros(gx * gy * gz * 3, np.float32)
    tmp = np.zeros(gx * gy * gz, np.float32)
    fn = _lib().orc_particle_sdf
    fn.restype = None
    fn(_p(vol), _p(x), _p(color), _p(lo), ctypes.c_int(int(bake_size)), _p(grid), ctypes.c_float(inv_dx), _p(sdf),
       _p(rgb), _p(tmp), ctypes.c_int(len(x)))
    return vol.reshape(gx, gy, gz), sdf.reshape(gx, gy, gz), rgb.reshape(gx, gy, gz, 3)


def render(sdf, box_min, box_max, rgb, body_pos, body_rot, body_tfsr, body_args, camera_rot, camera_pos,
           camera_intrinsic, image_hw, sdf_threshold=0.414, n_bodies=None, visualize_shape=True, max_ray_depth=2,
           spp=3, ground_height=0.0, seed=0, light_direction=(1.0, 1.0, 0.0), threads=None):
    """Returns (colour float32 [H, W, 3], depth float32 [H, W]); pixels are split over
    threads (each pixel's samples are traced in order by one thread)."""
    f = lambda a, n=None: np.ascontiguousarray(a, np.float32).reshape(-1 if n is None else n)  # noqa: E731
    sdf = f(sdf)
    grid = np.ascontiguousarray(np.shape(rgb)[:3] if np.ndim(rgb) == 4 else np.shape(sdf), np.int32)
    rgb = f(rgb)
    bpos, brot, tfsr, args = f(body_pos), f(body_rot), f(body_tfsr), f(body_args)
    nb = len(bpos) // 3 if n_bodies is None else int(n_bodies)
    H, W = (int(v) for v in image_hw)
    col = np.zeros(H * W * 3, np.float32)
    dep = np.zeros(H * W, np.float32)
    lo, hi = f(box_min, 3), f(box_max, 3)
    R, T, K, L = f(camera_rot, 9), f(camera_pos, 3), f(camera_intrinsic, 9), f(light_direction, 3)
    fn = _lib().orc_render
    fn.restype = None

    def run(rng):
        fn(_p(sdf), _p(lo), _p(hi), _p(rgb), _p(bpos), _p(brot), _p(tfsr), _p(args), _p(R), _p(T), _p(K), _p(col),
           _p(dep), ctypes.c_float(sdf_threshold), _p(grid), ctypes.c_int(nb), ctypes.c_int(int(bool(visualize_shape))),
           ctypes.c_int(H), ctypes.c_int(W), ctypes.c_int(max_ray_depth), ctypes.c_int(spp),
           ctypes.c_float(ground_height), ctypes.c_int(seed), _p(L), ctypes.c_int(rng[0]), ctypes.c_int(rng[1]))

    n = H * W
    k = threads or min(16, os.cpu_count() or 1)
    bounds = np.linspace(0, n, k + 1).astype(int)
    with ThreadPoolExecutor(k) as ex:
        list(ex.map(run, [(int(bounds[i]), int(bounds[i + 1])) for i in range(k)]))
    return col.reshape(H, W, 3), dep.reshape(H, W)
