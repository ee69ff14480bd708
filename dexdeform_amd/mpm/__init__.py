"""MI355X drop-in for DexDeform's ``mpm`` simulator package (simulator + autograd bridge)."""
from .engine import Engine, SceneSpec, StateBuffers, lame  # noqa: F401
