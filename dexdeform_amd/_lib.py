"""ctypes binding of libdexdeform_mpm.so (the C ABI declared in include/dexdeform_mpm.h).

The library is loaded after ``import torch`` so that its NEEDED ``libamdhip64.so.7``
resolves to the HIP runtime PyTorch-ROCm already loaded; streams and device pointers
from torch are therefore valid in the library. There is no fallback: if the library is
missing or fails to load, every simulator entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch  # noqa: F401  (must precede loading the HIP library)

LIB_PATH = os.environ.get("DDM_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib",
                                                     "libdexdeform_mpm.so")

c_int32 = ctypes.c_int32
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p

DDM_EINVAL = 10001
DDM_ETOOMANYBODIES = 10002
DDM_ESCRATCH = 10003
DDM_MAX_BODIES = 64
DDM_SCRATCH_BLOCKS = 8192

ERROR_NAMES = {DDM_EINVAL: "DDM_EINVAL (invalid argument)", DDM_ETOOMANYBODIES: "DDM_ETOOMANYBODIES",
               DDM_ESCRATCH: "DDM_ESCRATCH (more grid blocks than the fixed workspace holds)"}


class DdmScene(ctypes.Structure):
    _fields_ = [
        ("p_count", c_int32), ("blk_begin", c_int32), ("n_blocks", c_int32), ("dim", c_int32 * 3),
        ("n_bodies", c_int32), ("adjoint_mode", c_int32),
        ("dx", c_float), ("inv_dx", c_float), ("dt", c_float), ("ground_friction", c_float),
        ("ground_height", c_float), ("gravity", c_float * 3),
        ("mass", c_void_p), ("vol", c_void_p), ("mu_lam_yield", c_void_p), ("body_prim", c_void_p),
        ("body_args", c_void_p), ("tmp_qu", c_void_p), ("tmp_qv", c_void_p), ("tmp_sig", c_void_p),
        ("tape_mmv", c_void_p), ("tape_v", c_void_p), ("tape_list", c_void_p), ("tape_status", c_void_p),
        ("tape_capacity", c_int32), ("reserved", c_int32),
    ]


SCENE_DTYPE = np.dtype([
    ("p_count", "<i4"), ("blk_begin", "<i4"), ("n_blocks", "<i4"), ("dim", "<i4", (3,)),
    ("n_bodies", "<i4"), ("adjoint_mode", "<i4"),
    ("dx", "<f4"), ("inv_dx", "<f4"), ("dt", "<f4"), ("ground_friction", "<f4"),
    ("ground_height", "<f4"), ("gravity", "<f4", (3,)),
    ("mass", "<u8"), ("vol", "<u8"), ("mu_lam_yield", "<u8"), ("body_prim", "<u8"), ("body_args", "<u8"),
    ("tmp_qu", "<u8"), ("tmp_qv", "<u8"), ("tmp_sig", "<u8"),
    ("tape_mmv", "<u8"), ("tape_v", "<u8"), ("tape_list", "<u8"), ("tape_status", "<u8"),
    ("tape_capacity", "<i4"), ("reserved", "<i4"),
])
assert SCENE_DTYPE.itemsize == ctypes.sizeof(DdmScene) == 168


class DdmDet(ctypes.Structure):
    """Workspace of the deterministic sorted-reduction mode (device arrays)."""
    _fields_ = [
        ("cell_count", c_void_p), ("cell_start", c_void_p), ("order", c_void_p), ("order_p", c_void_p),
        ("key_slot", c_void_p), ("records", c_void_p), ("sorted", c_void_p), ("body_partial", c_void_p),
        ("range_count", c_void_p), ("counter", c_void_p),
    ]


assert ctypes.sizeof(DdmDet) == 80


class DdmTextureResources(ctypes.Structure):
    """(device volume, texture handle) returned by create_volume (the layout of types.py's
    texture_resources; the handle is 0 on gfx950, which has no texture objects)."""
    _fields_ = [("volume", c_void_p), ("texture", ctypes.c_ulonglong)]


class DdmBatch(ctypes.Structure):
    _fields_ = [
        ("n_scenes", c_int32), ("n_chunks", c_int32), ("n_blocks", c_int32), ("max_bodies", c_int32),
        ("scenes", c_void_p), ("chunks", c_void_p), ("grid_mmv", c_void_p), ("grid_v", c_void_p),
        ("grid_dv", c_void_p), ("grid_dmmv", c_void_p), ("block_flags", c_void_p), ("block_list", c_void_p),
        ("block_count", c_void_p), ("live", c_void_p), ("det", ctypes.POINTER(DdmDet)), ("options", c_int32),
        ("reserved", c_int32),
    ]


DDM_GRAD_OVERWRITE = 1
DDM_DETERMINISTIC = 2
DDM_TAPE_COMPLETE = 4


class DdmState(ctypes.Structure):
    _fields_ = [("x", c_void_p), ("v", c_void_p), ("F", c_void_p), ("C", c_void_p),
                ("body_pos", c_void_p), ("body_rot", c_void_p)]


STATE_DTYPE = np.dtype([("x", "<u8"), ("v", "<u8"), ("F", "<u8"), ("C", "<u8"), ("body_pos", "<u8"),
                        ("body_rot", "<u8")])
assert STATE_DTYPE.itemsize == ctypes.sizeof(DdmState) == 48


# Symbols declared in include/dexdeform_mpm.h
RUNTIME_SYMBOLS = ["cuda_alloc", "cuda_free", "print_memory_info", "cuda_stream_create", "cuda_stream_destroy",
                   "cuda_stream_sync", "cuda_upload", "cuda_download", "cuda_copy", "cuda_copy2d", "cuda_zero",
                   "cuda_zero_async", "cuda_upload_async", "cuda_download_async", "cuda_copy_async"]
REFERENCE_SYMBOLS = ["compute_grid_lower", "compute_svd", "compute_svd_grad", "p2g", "p2g_grad", "grid_op_v2", "grid_op_v2_grad", "g2p",
                     "g2p_grad", "compute_dist", "particle2mass"]
FUSED_SYMBOLS = ["ddm_substep_forward", "ddm_substep_backward", "ddm_substep_forward_timed",
                 "ddm_substep_backward_timed", "ddm_substep_forward_taped", "ddm_substep_backward_taped",
                 "ddm_substep_forward_taped_timed", "ddm_substep_backward_taped_timed", "ddm_batch_reset",
                 "ddm_substep_backward_taped_ahead", "ddm_substep_forward_ahead",
                 "ddm_version", "ddm_substep_forward_scratch"]
SCORE_SYMBOLS = ["ddm_pairwise_workspace_bytes", "ddm_softmin", "ddm_plan_grad", "ddm_nearest"]
RENDER_SYMBOLS = ["particle_sdf", "render", "create_volume", "destroy_volume"]
ALL_SYMBOLS = RUNTIME_SYMBOLS + REFERENCE_SYMBOLS + FUSED_SYMBOLS + SCORE_SYMBOLS + RENDER_SYMBOLS

_P = c_void_p
_I = ctypes.c_int
_F = c_float
_S = ctypes.c_size_t
_ARGTYPES = {
    "cuda_free": [_P],
    "cuda_stream_destroy": [_P],
    "cuda_stream_sync": [_P],
    "cuda_upload": [_P, _P, _S],
    "cuda_download": [_P, _P, _S],
    "cuda_copy": [_P, _P, _S],
    "cuda_copy2d": [_P, _S, _P, _S, _S, _S],
    "cuda_zero": [_P, _S],
    "cuda_zero_async": [_P, _S, _P],
    "cuda_upload_async": [_P, _P, _S, _P],
    "cuda_download_async": [_P, _P, _S, _P],
    "cuda_copy_async": [_P, _P, _S, _P],
    "compute_grid_lower": [_P, _F, _F, _P, _I, _P],
    "compute_svd": [_P, _P, _P, _P, _P, _P, _F, _I, _P],
    "compute_svd_grad": [_P] * 11 + [_F, _I, _P],
    "p2g": [_P] * 11 + [_P, _F, _F, _F, _P, _P, _P, _I, _P],
    "p2g_grad": [_P] * 11 + [_P, _F, _F, _F] + [_P] * 13 + [_I, _P],
    "grid_op_v2": [_P] * 11 + [_F, _F, _F, _F, _P, _P, _I, _P],
    "grid_op_v2_grad": [_P] * 17 + [_F, _F, _F, _F, _P, _P, _P, _I, _P],
    "g2p": [_P, _P, _P, _F, _F, _F, _P, _P, _F, _P, _P, _I, _P],
    "g2p_grad": [_P, _P, _P, _F, _F, _F, _P, _P, _F, _P, _P, _I, _P, _P, _P, _P, _P, _P],
    "compute_dist": [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _I, _I, _P],
    "particle2mass": [_P, _P, _P, _P, _F, _F, _P, _P, _P, _P, _I, _I, _I, _P],
    # state tables are device arrays of DdmState records: passed as raw device pointers
    "ddm_substep_forward": [ctypes.POINTER(DdmBatch), _P, _P, _P],
    "ddm_substep_backward": [ctypes.POINTER(DdmBatch), _P, _P, _P, _P, _P],
    "ddm_substep_forward_timed": [ctypes.POINTER(DdmBatch), _P, _P, _P, _P],
    "ddm_substep_backward_timed": [ctypes.POINTER(DdmBatch), _P, _P, _P, _P, _P, _P],
    "ddm_substep_forward_taped": [ctypes.POINTER(DdmBatch), _P, _P, _P],
    "ddm_substep_backward_taped": [ctypes.POINTER(DdmBatch), _P, _P, _P, _P, _P],
    "ddm_substep_backward_taped_ahead": [ctypes.POINTER(DdmBatch), _P, _P, _P, _P, ctypes.c_int, _P, _P],
    "ddm_substep_forward_ahead": [ctypes.POINTER(DdmBatch), _P, _P, _P, ctypes.c_int, ctypes.c_int, _P],
    "ddm_substep_forward_taped_timed": [ctypes.POINTER(DdmBatch), _P, _P, _P, _P],
    "ddm_substep_backward_taped_timed": [ctypes.POINTER(DdmBatch), _P, _P, _P, _P, _P, _P],
    "ddm_batch_reset": [ctypes.POINTER(DdmBatch), _P],
    "ddm_substep_forward_scratch": [ctypes.POINTER(DdmBatch), _P, _P, _P],
    "ddm_version": [],
    "ddm_softmin": [_P, _I, _P, _P, _I, _F, _I, _I, _P, _S, _P, _P],
    "ddm_plan_grad": [_P, _P, _I, _P, _P, _I, _F, _I, _I, _P, _S, _P, _P],
    "ddm_nearest": [_P, _I, _P, _I, _I, _P, _S, _P, _P, _P],
    # ivec3 arguments: pointers to 3 host ints (as types.py:284-286 passes POINTER(ivec3))
    "particle_sdf": [_P] * 5 + [_I, _P, _F] + [_P] * 3 + [_I, _P],
    "render": [_P] * 13 + [_F, _P, _I, ctypes.c_bool, _P, _I, _I, _F, _I, _P, _P],
    "create_volume": [_P, _I, _I, _I],
}

_lib = None


class ExtensionMissing(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load the HIP library (cached). Raises ExtensionMissing if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ExtensionMissing(
            f"{path} is not built; run `python -m dexdeform_amd.build` (the MI355X kernels have no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, args in _ARGTYPES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.cuda_alloc.argtypes = [_S]
    lib.cuda_alloc.restype = _P
    lib.cuda_stream_create.argtypes = []
    lib.cuda_stream_create.restype = _P
    lib.print_memory_info.argtypes = []
    lib.print_memory_info.restype = None
    lib.ddm_pairwise_workspace_bytes.argtypes = [_I, _I]
    lib.ddm_pairwise_workspace_bytes.restype = _S
    lib.create_volume.restype = DdmTextureResources
    lib.destroy_volume.argtypes = [DdmTextureResources]
    lib.destroy_volume.restype = ctypes.c_int
    _lib = lib
    return lib


def check(status: int, what: str = "kernel") -> None:
    if status != 0:
        name = ERROR_NAMES.get(status, f"hipError {status}")
        raise RuntimeError(f"dexdeform_mpm {what} failed: {name}")


def stream_ptr(stream=None) -> int:
    """Raw hipStream_t of a torch stream (default: the current stream)."""
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def ptr(t) -> int:
    """Device address of a tensor (None -> NULL)."""
    if t is None:
        return None
    return int(t.data_ptr())
