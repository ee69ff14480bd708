"""dexdeform_amd — MI355X-native differentiable MLS-MPM substep (DexDeform hot path).

HIP kernels for gfx950 behind a C ABI (``include/dexdeform_mpm.h``), a Python mirror of
DexDeform's ``mpm/simulator.py`` + ``mpm/torch_wrapper.py`` (``dexdeform_amd.mpm``),
multi-scene batched rollouts and a continuous-batching rollout server.
"""
__version__ = "0.1.0"
