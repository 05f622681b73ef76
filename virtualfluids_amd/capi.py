"""ctypes binding of the C ABI in include/virtualfluids_amd.h (libvfamd.so).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvfamd.so")

VF_OK = 0
VF_ERR_INVALID_ARGUMENT = 1
VF_ERR_HIP = 2
VF_ERR_UNSUPPORTED = 3

VF_STORAGE_FP32 = 0
VF_STORAGE_FP16 = 1
VF_STORAGE_BF16 = 2

# CollisionTemplate (src/gpu/core/Calculation/Calculation.h:56-69)
COLLISION_DEFAULT = 0
COLLISION_WRITE_MACRO_VARS = 1
COLLISION_APPLY_BODY_FORCE = 2
COLLISION_ALL_FEATURES = 3
COLLISION_SUBDOMAIN_BORDER = 4

GEO_SOLID = 15
GEO_VOID = 16
GEO_FLUID = 19

_vp = ctypes.c_void_p


# vf_box_kernel, vf_row_addressing, vf_adjoint_kernel
BOX_KERNEL_AUTO, BOX_KERNEL_ROW, BOX_KERNEL_NODE = 0, 1, 2
ROW_ADDRESSING_AUTO, ROW_ADDRESSING_BUFFER, ROW_ADDRESSING_GLOBAL = 0, 1, 2
ADJOINT_KERNEL_AUTO, ADJOINT_KERNEL_ROW_WHOLE, ADJOINT_KERNEL_NODE, ADJOINT_KERNEL_ROW = 0, 1, 2, 3


class KernelChoice(ctypes.Structure):
    """vf_kernel_choice: pins a form of the dense-box collision and its adjoint (zero: automatic)."""

    _fields_ = [
        ("boxKernel", ctypes.c_int32),
        ("rowWidth", ctypes.c_int32),
        ("rowWorkgroupsPerCU", ctypes.c_int32),
        ("rowAddressing", ctypes.c_int32),
        ("rowXcdRanges", ctypes.c_int32),
        ("adjointKernel", ctypes.c_int32),
    ]


class LBMParameter(ctypes.Structure):
    """vf_lbm_parameter: one grid level (LBMSimulationParameter, Parameter.h:67-256)."""

    _fields_ = [
        ("distributions", _vp),
        ("storage", ctypes.c_int32),
        ("isEvenTimestep", ctypes.c_int32),
        ("numberOfNodes", ctypes.c_uint64),
        ("neighborX", _vp),
        ("neighborY", _vp),
        ("neighborZ", _vp),
        ("typeOfGridNode", _vp),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("fluidMask", _vp),
        ("omega", ctypes.c_float),
        ("quadricLimiters", ctypes.c_float * 3),
        ("forces", ctypes.c_float * 3),
        ("forceFactor", ctypes.c_float),
        ("bodyForceX", _vp),
        ("bodyForceY", _vp),
        ("bodyForceZ", _vp),
        ("rho", _vp),
        ("velocityX", _vp),
        ("velocityY", _vp),
        ("velocityZ", _vp),
        ("pressure", _vp),
        ("fluidNodeIndices", _vp),
        ("numberOfFluidNodes", ctypes.c_uint32),
        ("porosity", _vp),
        ("workingSetBytes", ctypes.c_uint64),
        ("kernelChoice", KernelChoice),
    ]


class QBoundaryConditions(ctypes.Structure):
    """vf_q_boundary_conditions (QforBoundaryConditions, Calculation.h:105-127)."""

    _fields_ = [
        ("k", _vp),
        ("kN", _vp),
        ("q27", _vp),
        ("numberOfBCnodes", ctypes.c_uint32),
        ("Vx", _vp),
        ("Vy", _vp),
        ("Vz", _vp),
        ("RhoBC", _vp),
        ("direction", ctypes.c_int32),
        ("postCollision", _vp),
        ("nodesAlongRows", ctypes.c_int32),
    ]


class PostBounceBack(ctypes.Structure):
    """vf_post_bounce_back: bounce-back lists the dense-box collision applies in its own launch."""

    _fields_ = [
        ("lists", ctypes.POINTER(QBoundaryConditions) * 3),
        ("velocity", ctypes.c_int32 * 3),
        ("nodeEntry", _vp),
        ("entryIndex", _vp),
        ("entryLinks", _vp),
        ("numberOfEntries", ctypes.c_uint32),
    ]


class InterpolationCells(ctypes.Structure):
    _fields_ = [("fineCellIndices", _vp), ("coarseCellIndices", _vp), ("numberOfCells", ctypes.c_uint32)]


class InterpolationCellNeighbor(ctypes.Structure):
    _fields_ = [("x", _vp), ("y", _vp), ("z", _vp)]


class InterpolationSources(ctypes.Structure):
    """vf_interpolation_sources: the distinct source nodes of interface cells and their lanes."""

    _fields_ = [("node", _vp), ("first", _vp), ("lanes", _vp), ("numberOfNodes", ctypes.c_uint32)]


# VF_ABI_VERSION of include/virtualfluids_amd.h that these structs mirror
ABI_VERSION = 6

_P = ctypes.POINTER(LBMParameter)
_Q = ctypes.POINTER(QBoundaryConditions)

# name -> argtypes; every entry point returns int (vf_status)
SIGNATURES = {
    "vf_abi_version": [],
    "vf_stream_copy": [_vp, _vp, ctypes.c_uint64, _vp],
    "vf_stream_inplace27": [_vp, ctypes.c_uint64, ctypes.c_int32, _vp],
    "vf_k17_run": [_P, _vp],
    "vf_k17_run_on_indices": [_P, _vp, ctypes.c_uint32, ctypes.c_int32, _vp],
    "vf_k17_run_box": [_P, ctypes.c_int32, _vp],
    "vf_k17_run_box_into": [_P, _vp, _vp],
    "vf_k17_run_box_bounce_back": [_P, ctypes.c_int32, ctypes.POINTER(PostBounceBack), _vp],
    "vf_NoSlipBounceBack": [_P, _Q, _vp],
    "vf_NoSlipInterpolatedCompressible": [_P, _Q, _vp],
    "vf_VelocityBounceBack": [_P, _Q, _vp],
    "vf_VelocityInterpolatedCompressible": [_P, _Q, _vp],
    "vf_VelocityWithPressureInterpolatedCompressible": [_P, _Q, _vp],
    "vf_PressureNonEquilibriumCompressible": [_P, _Q, _vp],
    "vf_InitNavierStokesCompressible": [_P, _vp],
    "vf_CalculateMacroscopicQuantitiesCompressible": [_P, _vp],
    "vf_ConservationSums": [_P, _vp, _vp],
    "vf_MomentSums": [_P, _vp, _vp],
    "vf_ScaleCoarseToFineCompressibleAdjoint": [_P, _P, _vp, _vp, _vp, _vp, _vp],
    "vf_ScaleFineToCoarseCompressibleAdjoint": [_P, _P, _vp, _vp, _vp, _vp, _vp],
    "vf_ScaleCoarseToFineCompressibleAdjointGather": [_P, _P, _vp, _vp, ctypes.POINTER(InterpolationSources), _vp,
                                                      _vp, _vp, _vp],
    "vf_ScaleCoarseToFineCompressibleAdjointGatherNoAlloc": [_P, _P, _vp, _vp, ctypes.POINTER(InterpolationSources),
                                                             _vp, _vp, _vp],
    "vf_DragLiftPostD27": [_P, _Q, _vp, _vp, _vp, _vp],
    "vf_DragLiftPreD27": [_P, _Q, _vp, _vp, _vp, _vp],
    "vf_DragLiftSum": [_P, _Q, _vp, _vp],
    "vf_KineticEnergy": [_P, _vp, _vp, _vp],
    "vf_ScaleCoarseToFineCompressible": [_P, _P, _vp, _vp, _vp],
    "vf_k17_adjoint_box": [_P, _vp, _vp, _vp, _vp],
    "vf_k17_adjoint": [_P, _vp, _vp, _vp, _vp],
    "vf_NoSlipBounceBackAdjoint": [_P, _Q, _vp, _vp, _vp],
    "vf_NoSlipInterpolatedCompressibleAdjoint": [_P, _Q, _vp, _vp, _vp],
    "vf_VelocityBounceBackAdjoint": [_P, _Q, _vp, _vp, _vp],
    "vf_VelocityInterpolatedCompressibleAdjoint": [_P, _Q, _vp, _vp, _vp],
    "vf_VelocityWithPressureInterpolatedCompressibleAdjoint": [_P, _Q, _vp, _vp, _vp],
    "vf_PressureNonEquilibriumCompressibleAdjoint": [_P, _Q, _vp, _vp, ctypes.c_int32, _vp],
    "vf_InitNavierStokesCompressibleNonEquilibrium": [_P, _vp, _vp],
    "vf_ScaleFineToCoarseCompressible": [_P, _P, _vp, _vp, _vp],
}

_lib = None


def load():
    """Load libvfamd.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run virtualfluids_amd/build.py (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        lib.vf_abi_version.restype = ctypes.c_int
        if lib.vf_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI {lib.vf_abi_version()}, this binding needs {ABI_VERSION}: rebuild it")
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        lib.vf_get_last_error.argtypes = []
        lib.vf_get_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


class VFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"virtualfluids_amd error {code}: {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != VF_OK:
        raise VFError(rc, load().vf_get_last_error().decode())


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
