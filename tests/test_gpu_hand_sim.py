"""The drop-in HandSimulator with the real Shadow hand (the lift-box configuration: right
hand, model scale 1.5, fixed wrist, hand friction 0.9, 19 capsule/box geoms, quality 1 so
40 substeps per environment step, ground friction 0.3, gravity (0, -2, 0)) on the GPU.

* ``step`` (the forward-only path the planner uses) over four environment steps of seeded
  joint-velocity actions against the FP32 oracle driven with the same geom poses:
  x, v, F, C within 1e-4 relative after 160 substeps;
* ``get_state`` / ``set_state`` round trip, with the geom poses following from the wrist
  frame and joint angles;
* GradModel over two environment steps, L2 loss to a seeded target, loss.backward():
  action gradients against the FP32 oracle's adjoint chained through the same kinematics
  on the CPU, and against the FP64 oracle's: within 1e-4 of the FP32 oracle and 1e-3 of the
  FP64 one.

The hand configuration comes from tests/golden/shadow_rh_scale1p5.npz (the MJCF files are
not needed on the GPU machine).
"""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import rel_err
from dexdeform_amd.mpm import hand as HM
from dexdeform_amd.mpm import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shadow_rh_scale1p5.npz")
N = 4000
ENV_STEPS = 2


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def _sim(max_steps):
    prims, cfg = HM.load_hand_config(GOLDEN)
    n_bodies, kw = HM.hand_bodies(prims)
    sim = HM.HandSimulator(n_bodies, cfg, cfg=dict(n_particles=N, E=5e3, nu=0.2, yield_stress=50.0,
                                                   ground_friction=0.3, quality=1, max_steps=max_steps,
                                                   gravity=(0.0, -2.0, 0.0), fixed_base=True), device=_dev())
    sim.init_bodies(**kw)
    x0 = W.box_particles(N, [0.445, 0.175, 0.405], [0.535, 0.265, 0.495], seed=5)
    base = HM.root_matrix([0.5, 0.2, 0.3], [0.0, 0.0, np.pi])
    z = np.zeros((N, 3), np.float32)
    state = (x0, z, np.tile(np.eye(3, dtype=np.float32), (N, 1, 1)), np.zeros((N, 3, 3), np.float32),
             base[None], np.zeros((1, 24)))
    sim.set_state(0, state)
    return sim, x0


def _actions(seed=6, steps=ENV_STEPS):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, size=(steps, 1, 26)).astype(np.float32)


def _oracle(sim, dtype):
    e = sim.engine
    sc = O.Scene(grid_dim=tuple(int(g) for g in sim.grid_dim_np), dx=float(np.float32(sim.dx)),
                 dt=float(np.float32(sim.dt)), gravity=tuple(np.float32(sim._spec.gravity)),
                 mass=e.mass.cpu().numpy(), vol=e.vol.cpu().numpy(), mly=e.mu_lam_yield.cpu().numpy(),
                 tfsr=e.body_prim.cpu().numpy(), args=e.body_args.cpu().numpy(),
                 ground_friction=float(sim.get_ground_friction()), ground_height=sim.get_ground_height())
    return O.OracleSim(sc, dtype)


def _initial(sim, x0, dtype):
    st = sim.states[0]
    pos = st.body_pos.download().astype(dtype)
    rot = st.body_rot.download().astype(dtype)
    return O.State(x0.astype(dtype), np.zeros((N, 3), dtype), np.tile(np.eye(3, dtype=dtype), (N, 1, 1)),
                   np.zeros((N, 3, 3), dtype), pos, rot)


STEP_ENV_STEPS = 4


def test_hand_step_matches_oracle_with_the_same_geom_poses():
    sim, x0 = _sim(max_steps=41)
    acts = _actions(steps=STEP_ENV_STEPS)
    s0 = _initial(sim, x0, np.float32)  # particles and geom poses before the steps
    P, R = [sim.states[0].body_pos.download()], [sim.states[0].body_rot.download()]
    for a in acts:
        pos, rot, _ = sim.JointVel_Fk(0, torch.tensor(a, device=_dev()))
        P.extend(pos.detach().cpu().numpy())
        R.extend(rot.detach().cpu().numpy())
        sim.step(torch.tensor(a, device=_dev()))
    torch.cuda.synchronize()
    got = sim.get_state(0)
    osim = _oracle(sim, np.float32)
    ref = osim.rollout(s0, np.stack(P), np.stack(R))[-1]
    moved = float(np.abs(ref.x - x0).max())
    print(f"max particle displacement {moved:.3e}")
    assert moved > 1e-4, "the hand did not touch the particles"
    errs = {name: rel_err(g, r) for name, g, r in zip("xvFC", got[:4], (ref.x, ref.v, ref.F, ref.C))}
    print(f"hand step x{STEP_ENV_STEPS} ({40 * STEP_ENV_STEPS} substeps) vs oracle: "
          + ", ".join(f"{k} {v:.3e}" for k, v in errs.items()))
    assert max(errs.values()) <= 1e-4, errs
    # the hand state advanced with the steps
    assert np.allclose(got[-2], sim.base_pose[0].cpu().numpy())
    np.testing.assert_allclose(sim.states[0].body_pos.download(), P[-1], atol=1e-6)


def test_hand_state_round_trip():
    sim, x0 = _sim(max_steps=41)
    rng = np.random.default_rng(2)
    st = list(sim.get_state(0))
    st[1] = rng.normal(size=(N, 3)).astype(np.float32) * 0.01
    st[-1] = rng.uniform(0.0, 0.3, size=(1, 24))
    st[-2] = HM.root_matrix([0.48, 0.22, 0.31], [0.1, 0.0, np.pi])[None]
    sim.set_state(0, st)
    back = sim.get_state(0)
    for a, b in zip(st[:4], back[:4]):
        np.testing.assert_allclose(np.asarray(a, np.float32).reshape(b.shape), b, atol=0)
    np.testing.assert_allclose(back[-2], st[-2].astype(np.float32), atol=1e-7)
    np.testing.assert_allclose(back[-1], st[-1].astype(np.float32), atol=1e-7)
    pos, rot = sim.hand_forward_kinematics(sim.base_pose[0][None], sim.joint_rot[0][None])
    pose = np.stack(back[4:-2])
    np.testing.assert_allclose(pose[:, :3], pos[0].cpu().numpy(), atol=1e-6)
    np.testing.assert_allclose(pose[:, 3:], rot[0].cpu().numpy(), atol=1e-6)


def _oracle_action_grad(sim, x0, target, dtype):
    """The oracle's rollout and adjoint (exact dtype) chained through the hand kinematics
    on the CPU in the same dtype."""
    prims, cfg = HM.load_hand_config(GOLDEN)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    kin = HM.HandKinematics(cfg, device="cpu", dtype=tdt)
    scale = torch.tensor([0.33 * 0.002] * 20 + [0.0] * 6, dtype=tdt)
    a = torch.tensor(_actions(), dtype=tdt, requires_grad=True)
    base = torch.tensor(HM.root_matrix([0.5, 0.2, 0.3], [0.0, 0.0, np.pi]), dtype=tdt)[None]
    q = torch.zeros(1, 24, dtype=tdt)
    p0, r0 = kin.forward_kinematics(base[None], q[None])
    Ps, Rs = [p0[0]], [r0[0]]
    for e in range(ENV_STEPS):
        nb, nq = kin.joint_velocity(base, q, a[e], scale, sim.substeps)
        p, r = kin.forward_kinematics(nb, nq)
        Ps.extend(list(p))
        Rs.extend(list(r))
        base, q = nb[-1], nq[-1]
    P, R = torch.stack(Ps), torch.stack(Rs)
    osim = _oracle(sim, dtype)
    s0 = _initial(sim, x0, dtype)
    s0.body_pos, s0.body_rot = P[0].detach().numpy(), R[0].detach().numpy()
    states = osim.rollout(s0, P.detach().numpy(), R.detach().numpy())
    gx = (2.0 / N) * (states[-1].x.astype(np.float64) - target)
    z = lambda *s: np.zeros(s, dtype)  # noqa: E731
    g_last = O.Grads(gx.astype(dtype), z(N, 3), z(N, 3, 3), z(N, 3, 3), z(19, 3), z(19, 4))
    _, pg, rg = osim.rollout_grad(states, g_last)
    ((P * torch.tensor(pg, dtype=tdt)).sum() + (R * torch.tensor(rg, dtype=tdt)).sum()).backward()
    return a.grad.double().numpy()


def test_hand_gradmodel_action_gradients_vs_oracle():
    from dexdeform_amd.mpm.torch_wrapper import GradModel
    sim, x0 = _sim(max_steps=ENV_STEPS * 40 + 1)
    target = W.plain_target(x0, seed=7)
    model = GradModel(sim, return_grid=())
    model.zero_grad()
    acts = torch.tensor(_actions(), device=_dev(), requires_grad=True)
    obs = model.get_obs(0, _dev())
    for s in range(ENV_STEPS):
        obs = model.forward(s, acts[s], *obs)
    loss = ((obs[0][:, :3] - torch.tensor(target, device=_dev())) ** 2).sum(-1).mean()
    loss.backward()
    g = acts.grad.double().cpu().numpy()
    g32 = _oracle_action_grad(sim, x0, target, np.float32)
    g64 = _oracle_action_grad(sim, x0, target, np.float64)
    e32, e64, n32 = rel_err(g, g32), rel_err(g, g64), rel_err(g32, g64)
    print(f"loss {loss.item():.6e}; action grads: GPU vs oracle FP32 {e32:.2e}, vs oracle FP64 {e64:.2e} "
          f"(oracle FP32 vs FP64 {n32:.2e}); |g| {np.linalg.norm(g):.3e}")
    assert np.isfinite(g).all() and np.linalg.norm(g) > 0
    # measured: 2.0e-5 from the FP32 oracle, 2.9e-4 from the FP64 one (the FP32 oracle itself
    # is 2.7e-4 from FP64: the FP32 state's floor over 80 substeps)
    assert e32 <= 1e-4 and e64 <= 1e-3
