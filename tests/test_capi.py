"""The C ABI library loads and exports every entry point include/*.h declares (CPU only).

No compute call is made here: only argument validation paths that return before any
HIP call, mirroring the reference's error behaviour (unknown CollisionTemplate throws,
K17CompressibleNavierStokes.cu:139-141; BC functions need their node data).
"""
import ctypes
import glob
import os
import re

import pytest

from virtualfluids_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(vf_\w+)\s*\(", src))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("vf_k17_run", "vf_k17_run_on_indices", "vf_k17_run_box", "vf_NoSlipBounceBack",
                     "vf_VelocityBounceBack", "vf_PressureNonEquilibriumCompressible",
                     "vf_InitNavierStokesCompressible", "vf_CalculateMacroscopicQuantitiesCompressible"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.vf_abi_version() == capi.ABI_VERSION == 6


def test_signature_table_covers_header():
    # every declared entry point has a ctypes signature, so Python calls are type-checked
    for n in declared_functions():
        assert n in capi.SIGNATURES or n == "vf_get_last_error", n


def _param(**kw):
    p = capi.LBMParameter()
    p.distributions = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    p.numberOfNodes = 8
    p.omega = 1.5
    p.storage = capi.VF_STORAGE_FP32
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_null_arguments_are_rejected():
    lib = capi.load()
    assert lib.vf_k17_run(None, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_k17_run_box(None, 0, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_NoSlipBounceBack(None, None, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert b"null" in lib.vf_get_last_error()


def test_unknown_collision_template_is_an_error():
    lib = capi.load()
    p = _param(neighborX=ctypes.c_void_p(0x10), neighborY=ctypes.c_void_p(0x10), neighborZ=ctypes.c_void_p(0x10))
    rc = lib.vf_k17_run_on_indices(ctypes.byref(p), ctypes.c_void_p(0x10), 4, 99, None)
    assert rc == capi.VF_ERR_INVALID_ARGUMENT
    assert b"Invalid CollisionTemplate" in lib.vf_get_last_error()
    with pytest.raises(capi.VFError):
        capi.check(rc)


def test_box_path_validates_extents_and_omega():
    lib = capi.load()
    p = _param(nx=2, ny=2, nz=3)  # 12 != 8 nodes
    assert lib.vf_k17_run_box(ctypes.byref(p), 0, None) == capi.VF_ERR_INVALID_ARGUMENT
    p = _param(nx=2, ny=2, nz=2, omega=2.5)
    assert lib.vf_k17_run_box(ctypes.byref(p), 0, None) == capi.VF_ERR_INVALID_ARGUMENT
    p = _param(nx=2, ny=2, nz=2, storage=7)
    assert lib.vf_k17_run_box(ctypes.byref(p), 0, None) == capi.VF_ERR_INVALID_ARGUMENT


def test_velocity_bc_needs_wall_velocity():
    lib = capi.load()
    p = _param(nx=2, ny=2, nz=2)
    bc = capi.QBoundaryConditions()
    bc.k = ctypes.c_void_p(0x10)
    bc.q27 = ctypes.c_void_p(0x10)
    bc.numberOfBCnodes = 3
    assert lib.vf_VelocityBounceBack(ctypes.byref(p), ctypes.byref(bc), None) == capi.VF_ERR_INVALID_ARGUMENT
    bc.numberOfBCnodes = 0  # empty list: nothing to do
    assert lib.vf_VelocityBounceBack(ctypes.byref(p), ctypes.byref(bc), None) == capi.VF_OK


def test_stream_copy_validates_alignment():
    lib = capi.load()
    assert lib.vf_stream_copy(ctypes.c_void_p(0x10), ctypes.c_void_p(0x20), 24, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_stream_copy(ctypes.c_void_p(0x18), ctypes.c_void_p(0x20), 32, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_stream_copy(None, ctypes.c_void_p(0x20), 32, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_stream_copy(None, None, 0, None) == capi.VF_OK  # nothing to copy


def test_stream_inplace27_validates_sizes():
    lib = capi.load()
    assert lib.vf_stream_inplace27(None, 1024, 0, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_stream_inplace27(ctypes.c_void_p(0x100), 1000, 0, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_stream_inplace27(ctypes.c_void_p(0x100), 1 << 34, 0, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_stream_inplace27(ctypes.c_void_p(0x100), 1024, -1, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_stream_inplace27(ctypes.c_void_p(0x100), 1024, 33, None) == capi.VF_ERR_INVALID_ARGUMENT
    assert lib.vf_stream_inplace27(None, 0, 0, None) == capi.VF_OK  # nothing to do
