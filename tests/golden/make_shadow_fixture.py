"""Writes the Shadow-hand configuration of the lift-box environment (right hand, model
scale 1.5, hand friction 0.9; mpm/assets/env_cfgs/lift_box.yml) as plain arrays, so that the
GPU tests can build the real hand without the MJCF files:

    python tests/golden/make_shadow_fixture.py [asset_dir]

asset_dir defaults to the reference checkout's mpm/assets/robots. The arrays are this
build's MJCF reader's output (dexdeform_amd.mpm.mjcf), pinned on the CPU against an
independent composition of the same files' transforms (tests/test_hand_model.py).
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dexdeform_amd.mpm import hand as HM  # noqa: E402

if __name__ == "__main__":
    assets = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/mpm/assets/robots"
    prims, cfg = HM.hand_config("rh", assets, scale=1.5, hand_friction=0.9)
    HM.save_hand_config(os.path.join(HERE, "shadow_rh_scale1p5.npz"), prims, cfg)
    print(f"{len(prims)} primitives written")
