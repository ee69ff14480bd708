"""Shadow-hand tables the hand simulator needs (mpm/robots/interface.py:5-145): joint and
actuator names, control ranges, coupled-joint tendons, default joint angles.

These are data of the robot model that DexDeform's hand controller is written against, so
they are kept identical: joint order, ranges and default angles decide the action layout,
the joint limits and the initial hand pose.
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np

FINGERTIP_SITE_NAMES = ["robot0:S_fftip", "robot0:S_mftip", "robot0:S_rftip", "robot0:S_lftip", "robot0:S_thtip"]

# (actuator, control range [rad], driven joints); the joints of one actuator are coupled
# by a tendon and share its range equally (mpm/robots/interface.py:73-139)
_ACTUATORS = [
    ("WRJ1", (-0.4887, 0.1396), ["WRJ1"]),
    ("WRJ0", (-0.6981, 0.4887), ["WRJ0"]),
    ("FFJ3", (-0.3491, 0.3491), ["FFJ3"]),
    ("FFJ2", (0.0, 1.5708), ["FFJ2"]),
    ("FFJ1", (0.0, 3.1416), ["FFJ1", "FFJ0"]),
    ("MFJ3", (-0.3491, 0.3491), ["MFJ3"]),
    ("MFJ2", (0.0, 1.5708), ["MFJ2"]),
    ("MFJ1", (0.0, 3.1416), ["MFJ1", "MFJ0"]),
    ("RFJ3", (-0.3491, 0.3491), ["RFJ3"]),
    ("RFJ2", (0.0, 1.5708), ["RFJ2"]),
    ("RFJ1", (0.0, 3.1416), ["RFJ1", "RFJ0"]),
    ("LFJ4", (0.0, 0.7854), ["LFJ4"]),
    ("LFJ3", (-0.3491, 0.3491), ["LFJ3"]),
    ("LFJ2", (0.0, 1.5708), ["LFJ2"]),
    ("LFJ1", (0.0, 3.1416), ["LFJ1", "LFJ0"]),
    ("THJ4", (-1.0472, 1.0472), ["THJ4"]),
    ("THJ3", (0.0, 1.2217), ["THJ3"]),
    ("THJ2", (-0.2094, 0.2094), ["THJ2"]),
    ("THJ1", (-0.5236, 0.5236), ["THJ1"]),
    ("THJ0", (-1.5708, 0.0), ["THJ0"]),
]
_P = "robot0:"

ACTUATOR_CTRLRANGE: Dict[str, List[float]] = {_P + "A_" + a: list(r) for a, r, _ in _ACTUATORS}
ACTUATOR_JOINT_MAPPING: Dict[str, List[str]] = {_P + "A_" + a: [_P + j for j in js] for a, _, js in _ACTUATORS}

# the 24 joints in simulator order (actuators in order, coupled joints distal last)
JOINTS: List[str] = [j for js in ACTUATOR_JOINT_MAPPING.values() for j in js]

JOINT_LIMITS: Dict[str, np.ndarray] = {}
for _a, _js in ACTUATOR_JOINT_MAPPING.items():
    for _j in _js:
        JOINT_LIMITS[_j] = np.array(ACTUATOR_CTRLRANGE[_a]) / len(_js)

_DEFAULT_Q = [
    -0.16514339750464327, -0.31973286565062153,
    0.14340512546557435, 0.32028208333591573, 0.7126053607727917, 0.6705281001412586,
    0.000246444303701037, 0.3152655251085491, 0.7659800313729842, 0.7323156897425923,
    0.00038520700007378114, 0.36743546201985233, 0.7119514095008576, 0.6699446327514138,
    0.0525442258033891, -0.13615534724474673, 0.39872030433433003, 0.7415570009679252, 0.704096378652974,
    0.003673823825070126, 0.5506291436028695, -0.014515151997119306, -0.0015229223564485414, -0.7894883021600622,
]
DEFAULT_INITIAL_QPOS: Dict[str, float] = dict(zip(JOINTS, _DEFAULT_Q))

# wrist orientations of the two named hand modes (euler angles, mpm/robots/interface.py:141-144)
MODES = {"face_up": [-np.pi / 2, 0, 0], "face_down": [np.pi / 2, 0, np.pi]}


def get_actuator_mapping() -> List[int]:
    """Actuator index of every joint, in JOINTS order (mpm/robots/interface.py:147-158)."""
    act = {j: i for i, js in enumerate(ACTUATOR_JOINT_MAPPING.values()) for j in js}
    return [act[j] for j in JOINTS]
