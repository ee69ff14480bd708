"""The C-ABI library loads on the host and exports every symbol the header declares
(no compute calls: there is no GPU in this check)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dexdeform_mpm.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|void\s*\*?|hipStream_t|ddm_texture_resources)\s*"
                                 r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt,
                                 flags=re.M)))


def test_header_declares_reference_launchers():
    fns = header_functions()
    for name in ["compute_svd", "compute_svd_grad", "p2g", "p2g_grad", "grid_op_v2", "grid_op_v2_grad", "g2p",
                 "g2p_grad", "compute_dist", "particle2mass"]:
        assert name in fns, name


def test_library_exports_every_declared_symbol():
    from dexdeform_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libdexdeform_mpm.so is not built (run __graft_entry__.build())")
    lib = _lib.load()
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    assert set(header_functions()) == set(_lib.ALL_SYMBOLS)
    assert lib.ddm_version() == 100


def test_struct_layouts_match_header():
    import ctypes
    from dexdeform_amd import _lib
    assert ctypes.sizeof(_lib.DdmScene) == 168
    assert ctypes.sizeof(_lib.DdmState) == 6 * 8
    assert ctypes.sizeof(_lib.DdmBatch) == 16 + 11 * 8 + 8
    assert ctypes.sizeof(_lib.DdmDet) == 10 * 8
    hdr = open(HEADER).read()
    assert "} ddm_scene;                   /* 168 bytes */" in hdr


# Every symbol the reference's ctypes layer binds (mpm/types.py:103-290) or calls
# (print_memory_info, mpm/simulator.py:226), the renderer's included.
REFERENCE_CTYPES_SYMBOLS = [
    "cuda_alloc", "cuda_free", "cuda_upload", "cuda_upload_async", "cuda_download", "cuda_download_async",
    "cuda_copy", "cuda_copy2d", "cuda_zero", "cuda_zero_async", "cuda_stream_create", "cuda_stream_destroy",
    "cuda_stream_sync", "compute_grid_lower", "compute_svd", "compute_svd_grad", "p2g", "p2g_grad", "grid_op_v2",
    "grid_op_v2_grad", "g2p", "g2p_grad", "compute_dist", "particle2mass", "print_memory_info", "render",
    "particle_sdf", "create_volume", "destroy_volume"]


def test_library_exports_every_symbol_the_reference_binds():
    """The reference's mpm/types.py can load this library: each symbol it binds is exported (checked with a plain ctypes.CDLL, as the reference loads it) and
    declared in the header."""
    import ctypes
    from dexdeform_amd import _lib
    import torch  # noqa: F401  (the HIP runtime first, as the reference's process has it)
    raw = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in REFERENCE_CTYPES_SYMBOLS if not hasattr(raw, s)]
    assert not missing, missing
    assert set(REFERENCE_CTYPES_SYMBOLS) <= set(header_functions())
