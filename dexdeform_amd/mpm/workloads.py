"""Synthetic scenes for the benchmark configurations and the parity tests.

Particles are sampled with numpy's PCG64 generator and then ordered along a 4x4x4-cell
Morton-free spatial key (block index, then cell index) so that consecutive particles
are spatially close; the order is part of the synthetic input, the simulator accepts
any order. Units and conventions follow the reference simulator (gravity as configured;
MPMSimulator multiplies it by 30; capsules are (radius, segment length)).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .engine import SceneSpec, lame

CAPSULE, BOX = 1, 0
SOFTNESS = 666.0  # every primitive gets softness 666 (mpm/cuda_env.py:88)


def spatial_key(x: np.ndarray, dx: float) -> np.ndarray:
    """Sort key of each particle: its 4^3 cell block, then its cell in the block."""
    c = np.floor(np.asarray(x) / dx).astype(np.int64)
    b = c >> 2
    return ((b[:, 0] * 1024 + b[:, 1]) * 1024 + b[:, 2]) * 64 + ((c[:, 0] & 3) * 16 + (c[:, 1] & 3) * 4 + (c[:, 2] & 3))


def spatial_order(x: np.ndarray, dx: float) -> np.ndarray:
    """Permutation that groups particles by 4^3 cell block, then by cell. The particle
    kernels stage each chunk of 256 consecutive particles in an LDS tile around the
    chunk, so spatially ordered particles are several times faster than random order."""
    return np.argsort(spatial_key(x, dx), kind="stable")


def spatially_ordered(x: np.ndarray, dx: float) -> bool:
    k = spatial_key(x, dx)
    return bool(np.all(k[1:] >= k[:-1]))


def needs_sort(x: np.ndarray, dx: float, tolerance: float = 0.1) -> bool:
    """True when more than `tolerance` of neighbouring particles are out of block/cell
    order (e.g. randomly sampled); nearly ordered input (a few particles across a cell
    face) runs as fast as sorted input."""
    k = spatial_key(x, dx)
    return len(k) > 1 and float(np.mean(k[1:] < k[:-1])) > tolerance


def box_particles(n: int, lo, hi, seed: int, dx: float = 1 / 64) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.uniform(lo, hi, size=(n, 3))
    return x[spatial_order(x, dx)].astype(np.float32)


def ellipsoid_particles(n: int, center, radii, seed: int, dx: float = 1 / 64) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    need = n
    while need > 0:
        u = rng.uniform(-1, 1, size=(2 * need + 16, 3))
        u = u[(u * u).sum(1) <= 1.0]
        out.append(u[:need])
        need -= len(out[-1])
    x = np.concatenate(out)[:n] * np.asarray(radii) + np.asarray(center)
    return x[spatial_order(x, dx)].astype(np.float32)


def union_boxes_particles(n: int, boxes, seed: int, dx: float) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    vols = np.array([np.prod(np.subtract(hi, lo)) for lo, hi in boxes])
    counts = np.floor(n * vols / vols.sum()).astype(int)
    counts[0] += n - counts.sum()
    xs = [rng.uniform(lo, hi, size=(c, 3)) for (lo, hi), c in zip(boxes, counts)]
    x = np.concatenate(xs)
    return x[spatial_order(x, dx)].astype(np.float32)


@dataclass
class Scene:
    """A scene: kernel constants (spec), initial particle positions and body poses."""

    spec: SceneSpec
    x0: np.ndarray
    body_pos: np.ndarray
    body_rot: np.ndarray
    action_scales: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.x0)

    @property
    def nb(self):
        return len(self.body_pos)


def make_scene(x0: np.ndarray, grid: int, dx: float, dt: float, bodies_prim: np.ndarray, bodies_args: np.ndarray,
               body_pos: np.ndarray, body_rot: np.ndarray, E=5e3, nu=0.2, yield_stress=50.0, gravity=(0, -9.8, 0),
               vol=(1 / 128) ** 2, mass=(1 / 128) ** 2, ground_friction=0.0, adjoint_mode=0) -> Scene:
    n = len(x0)
    mu, lam = lame(E, nu)
    spec = SceneSpec(n_particles=n, grid_dim=(grid, grid, grid), dx=dx, dt=dt,
                     gravity=tuple(float(g) * 30.0 for g in gravity),
                     mass=np.full(n, mass, np.float32), vol=np.full(n, vol, np.float32),
                     mu_lam_yield=np.tile(np.float32([mu, lam, yield_stress]), (n, 1)),
                     body_prim=np.asarray(bodies_prim, np.float32).reshape(-1, 4),
                     body_args=np.asarray(bodies_args, np.float32).reshape(-1, 4),
                     ground_friction=ground_friction, adjoint_mode=adjoint_mode)
    return Scene(spec, np.asarray(x0, np.float32), np.asarray(body_pos, np.float32).reshape(-1, 3),
                 np.asarray(body_rot, np.float32).reshape(-1, 4))


def capsule_prims(n: int, friction=0.9):
    return np.tile(np.float32([CAPSULE, friction, SOFTNESS, 0.0]), (n, 1))


# ---------------------------------------------------------------- configurations
def plain_scene(n_particles=10000, grid=64, seed=0, dt=5e-5, adjoint_mode=0) -> Scene:
    """Plain single scene: 64^3 grid, particles in [0.4,0.6]x[0.05,0.25]x[0.4,0.6],
    E=5e3, nu=0.2, yield 50, two capsules r=0.03 with segment 0.10 (MuJoCo half-length
    0.05), friction 0.9, identity rotation, centres (0.45,0.30,0.50), (0.55,0.30,0.50)."""
    dx = 1.0 / grid
    x0 = box_particles(n_particles, [0.4, 0.05, 0.4], [0.6, 0.25, 0.6], seed, dx)
    args = np.float32([[0.03, 0.10, 0, 0], [0.03, 0.10, 0, 0]])
    pos = np.float32([[0.45, 0.30, 0.50], [0.55, 0.30, 0.50]])
    rot = np.float32([[1, 0, 0, 0], [1, 0, 0, 0]])
    sc = make_scene(x0, grid, dx, dt, capsule_prims(2), args, pos, rot, adjoint_mode=adjoint_mode)
    return sc


def plain_actions(n_env_steps=5, n_bodies=2, seed=1) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(-0.5, 0.5, size=(n_env_steps, n_bodies, 6)).astype(np.float32)


def plain_target(x0: np.ndarray, seed=2) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return (x0 + np.float32([0, 0.05, 0]) + rng.normal(0, 0.01, size=x0.shape)).astype(np.float32)


def random_capsules(rng, n_bodies: int, center, spread=0.12, r_range=(0.008, 0.03), h_range=(0.02, 0.10)):
    pos = np.asarray(center, np.float64) + rng.uniform(-spread, spread, size=(n_bodies, 3))
    axis = rng.normal(size=(n_bodies, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    ang = rng.uniform(0, np.pi, size=(n_bodies, 1))
    rot = np.concatenate([np.cos(ang / 2), axis * np.sin(ang / 2)], 1)
    args = np.zeros((n_bodies, 4))
    args[:, 0] = rng.uniform(*r_range, size=n_bodies)
    args[:, 1] = rng.uniform(*h_range, size=n_bodies)
    return pos.astype(np.float32), rot.astype(np.float32), args.astype(np.float32)


def batched_scenes(n_scenes=64, seed=7, n_range=(5000, 60000), grid=64, dt=5e-5, yield_range=(20.0, 100.0),
                   body_range=(2, 22)) -> List[Scene]:
    """Uneven scenes: particle counts U[n_range], yield stress U[yield_range],
    2-22 capsules each, boxes of random size and place."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 1.0 / grid
    out = []
    for s in range(n_scenes):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        half = rng.uniform(0.06, 0.14, size=3)
        c = np.array([rng.uniform(0.3, 0.7), rng.uniform(0.1 + half[1], 0.35), rng.uniform(0.3, 0.7)])
        lo, hi = c - half, c + half
        lo[1] = max(lo[1], 0.06)
        x0 = box_particles(n, lo, hi, int(rng.integers(1 << 30)), dx)
        nb = int(rng.integers(body_range[0], body_range[1] + 1))
        pos, rot, args = random_capsules(rng, nb, [c[0], hi[1] + 0.05, c[2]], spread=0.1)
        sc = make_scene(x0, grid, dx, dt, capsule_prims(nb), args, pos, rot,
                        yield_stress=float(rng.uniform(*yield_range)))
        sc.meta["seed"] = s
        out.append(sc)
    return out


def constant_velocity_poses(scenes, T: int, seed: int):
    """Body poses [T+1, nb, 3] / [T+1, nb, 4] per scene for capsules translating with
    constant velocities U(-0.5, 0.5) m/s (rotations fixed), drawn scene by scene from one
    generator."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for sc in scenes:
        vel = rng.uniform(-0.5, 0.5, size=(sc.nb, 3)).astype(np.float32)
        t = np.arange(T + 1, dtype=np.float32)[:, None, None] * np.float32(sc.spec.dt)
        pos = sc.body_pos[None] + t * vel[None]
        rot = np.broadcast_to(sc.body_rot[None], (T + 1, sc.nb, 4))
        out.append((pos.astype(np.float32), np.ascontiguousarray(rot, np.float32)))
    return out


def served_stream(rate: float, n_requests: int, seed=13, n_range=(5000, 100000), grids=(48, 64, 96),
                  horizon_range=(20, 200), checkpoint_every=10, dt=5e-5, ppc=6.0):
    """Rollout requests of a Poisson arrival stream (``rate`` requests/s): particle counts
    U[n_range] in a box sized for about ``ppc`` particles per cell, grid drawn from
    ``grids``, horizon U[horizon_range] in multiples of ``checkpoint_every`` (the
    checkpoint interval), 2-6 capsules with constant random velocities, yield stress
    U[20, 100], an L2 target near the initial shape. Returns a list of dicts with keys
    scene, pose_pos, pose_rot, target, arrival, checkpoint_every."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        grid = int(rng.choice(grids))
        dx = 1.0 / grid
        T = int(rng.integers(horizon_range[0] // checkpoint_every, horizon_range[1] // checkpoint_every + 1))
        T *= checkpoint_every
        side = min(0.6, dx * (n / ppc) ** (1.0 / 3.0))
        c = np.array([0.5, 0.06 + side / 2, 0.5])
        x0 = box_particles(n, c - side / 2, c + side / 2, int(rng.integers(1 << 30)), dx)
        nb = int(rng.integers(2, 7))
        pos, rot, args = random_capsules(rng, nb, [0.5, c[1] + side / 2 + 0.05, 0.5], spread=side / 2)
        sc = make_scene(x0, grid, dx, dt, capsule_prims(nb), args, pos, rot,
                        yield_stress=float(rng.uniform(20.0, 100.0)))
        vel = rng.uniform(-0.5, 0.5, size=(nb, 3)).astype(np.float32)
        steps = np.arange(T + 1, dtype=np.float32)[:, None, None] * np.float32(dt)
        pose_pos = (sc.body_pos[None] + steps * vel[None]).astype(np.float32)
        pose_rot = np.repeat(sc.body_rot[None], T + 1, 0)
        out.append(dict(scene=sc, pose_pos=pose_pos, pose_rot=pose_rot,
                        target=plain_target(x0, seed=int(rng.integers(1 << 30))), arrival=t,
                        checkpoint_every=checkpoint_every))
    return out


def hand_scene(n_particles=50000, seed=3, grid=64, dt=5e-5, friction=0.9):
    """Hand-contact scene: particles in a seeded ellipsoid (centre (0.5, 0.2, 0.5), radii
    (0.12, 0.08, 0.10)) with the 22-link synthetic capsule hand (tree seed 4) reaching
    into it from above. Returns (scene, hinge tree, initial joint angles); the scene's
    body poses are the hand's pose at those angles."""
    from .kinematics import synthetic_hand
    tree, args, q0 = synthetic_hand(seed=4)
    dx = 1.0 / grid
    x0 = ellipsoid_particles(n_particles, [0.5, 0.2, 0.5], [0.12, 0.08, 0.10], seed, dx)
    pos, rot = tree.fk(q0)
    sc = make_scene(x0, grid, dx, dt, capsule_prims(len(args), friction), args, pos.numpy(), rot.numpy())
    return sc, tree, q0


def hand_actions(n_env_steps: int, n_joints=22, seed=4) -> np.ndarray:
    """Seeded joint-velocity actions U(-1, 1) [n_env_steps, n_joints]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(-1.0, 1.0, size=(n_env_steps, n_joints)).astype(np.float32)


def large_scene(n_particles=1_000_000, seed=6, grid=128, dt=5e-5, friction=0.9):
    """Large single scene: a 128^3 grid, particles filling a union of 3 seeded boxes
    (half-extents U[0.06, 0.14] x U[0.05, 0.10] x U[0.06, 0.14], resting near the floor,
    centres within 0.15 of (0.5, 0.5)) with the 22-link capsule hand (palm at height 0.30)
    reaching into them. Returns (scene, hinge tree, initial joint angles)."""
    from .kinematics import synthetic_hand
    rng = np.random.Generator(np.random.PCG64(seed))
    boxes = []
    for _ in range(3):
        half = rng.uniform([0.06, 0.05, 0.06], [0.14, 0.10, 0.14])
        c = np.array([0.5 + rng.uniform(-0.15, 0.15), 0.04 + half[1], 0.5 + rng.uniform(-0.15, 0.15)])
        boxes.append((c - half, c + half))
    dx = 1.0 / grid
    x0 = union_boxes_particles(n_particles, boxes, int(rng.integers(1 << 30)), dx)
    tree, args, q0 = synthetic_hand(seed=4, palm=(0.5, 0.30, 0.5))
    pos, rot = tree.fk(q0)
    sc = make_scene(x0, grid, dx, dt, capsule_prims(len(args), friction), args, pos.numpy(), rot.numpy())
    sc.meta["boxes"] = boxes
    return sc, tree, q0
