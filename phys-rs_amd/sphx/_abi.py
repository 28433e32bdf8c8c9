"""ctypes declarations of include/sphx.h and the loader of libsphx.so.

The library is loaded from the package directory (``phys-rs_amd/libsphx.so``);
there is no fallback: a missing library raises at first use.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent.parent / "libsphx.so"

SPHX_OK = 0
ERRORS = {
    1: "SPHX_E_NODEVICE", 2: "SPHX_E_OOM", 3: "SPHX_E_CAPACITY", 4: "SPHX_E_INVALID",
    5: "SPHX_E_MULTISECTOR", 6: "SPHX_E_NONFINITE", 7: "SPHX_E_HIP", 8: "SPHX_E_STATE",
}
EOS_TAIT, EOS_POLYTROPIC, EOS_ISOTHERMAL, EOS_ADIABATIC = 0, 1, 2, 3
VISC_REFERENCE, VISC_MONAGHAN = 0, 1
BND_PENALTY_CUBE, BND_REFLECT, BND_PERIODIC = 1, 2, 4
MHD_INDUCTION, MHD_CLEANING, MHD_MAXWELL_STRESS, MHD_LORENTZ_CHARGE = 1, 2, 4, 8
PREC_FP64, PREC_FP32_MIXED = 0, 1
F_POS, F_VEL, F_RHO, F_P, F_ACC, F_B, F_PSI, F_U, F_ALL = 1, 2, 4, 8, 16, 32, 64, 128, 255
FORCES_FULL, FORCES_VISC_ONLY = 0, 1
TIMING_ALL = 1
(TIMING_SORT, TIMING_NEIGHBOUR, TIMING_DENSITY, TIMING_FORCE, TIMING_VISC_ONLY, TIMING_INTEGRATE,
 TIMING_MHD) = (1 << 1, 1 << 2, 1 << 3, 1 << 4, 1 << 5, 1 << 6, 1 << 7)

D3 = C.c_double * 3


class SphxParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("sector_size", C.c_double),
        ("rest_density", C.c_double), ("speed_of_sound", C.c_double), ("gamma", C.c_double),
        ("eos", C.c_int32), ("visc_convention", C.c_int32),
        ("poly_k", C.c_double), ("iso_cs2", C.c_double),
        ("visc_alpha", C.c_double), ("visc_beta", C.c_double),
        ("xsph_epsilon", C.c_double),
        ("half_extent", C.c_double), ("activation_distance", C.c_double), ("stiffness", C.c_double),
        ("boundary_flags", C.c_int32), ("mhd_flags", C.c_int32),
        ("box_lo", D3), ("box_hi", D3), ("gravity", D3), ("e_field", D3),
        ("resistivity", C.c_double), ("cleaning_speed", C.c_double), ("cleaning_damping", C.c_double),
        ("cell_size", C.c_double), ("grid_lo", D3), ("grid_hi", D3),
        ("precision", C.c_int32), ("max_neighbours", C.c_int32), ("neighbour_skin", C.c_double),
    ]


PD = C.POINTER(C.c_double)
PU = C.POINTER(C.c_uint32)


class SphxState(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("x", PD), ("y", PD), ("z", PD), ("vx", PD), ("vy", PD), ("vz", PD),
        ("m", PD), ("h", PD), ("rho", PD), ("p", PD), ("ax", PD), ("ay", PD), ("az", PD),
        ("bx", PD), ("by", PD), ("bz", PD), ("psi", PD), ("q", PD), ("id", PU), ("u", PD),
    ]


class SphxUniformData(C.Structure):
    """SphUniformData of the reference's WGSL path (src/gpu/buffers.rs:25-38)."""
    _fields_ = [
        ("num_particles", C.c_uint32), ("mass", C.c_float), ("h", C.c_float), ("dt", C.c_float),
        ("rest_density", C.c_float), ("speed_of_sound", C.c_float), ("gamma", C.c_float),
        ("alpha_visc", C.c_float), ("beta_visc", C.c_float), ("_padding", C.c_uint32),
    ]


class SphxTimes(C.Structure):
    _fields_ = [
        ("sort_ms", C.c_double), ("neighbour_ms", C.c_double), ("density_ms", C.c_double),
        ("force_ms", C.c_double), ("visc_only_ms", C.c_double), ("integrate_ms", C.c_double),
        ("mhd_ms", C.c_double), ("launches", C.c_uint64 * 7),
    ]


class SphxEnergy(C.Structure):
    _fields_ = [
        ("kinetic", C.c_double), ("potential", C.c_double), ("internal", C.c_double),
        ("magnetic", C.c_double), ("momentum", D3),
        ("max_neighbours", C.c_uint64), ("total_neighbours", C.c_uint64),
    ]


# name -> (restype, argtypes); every symbol include/sphx.h declares.
SIGNATURES = {
    "sphx_default_params": (None, [C.POINTER(SphxParams)]),
    "sphx_update": (C.c_int, [C.c_void_p, C.POINTER(SphxState), C.c_uint32]),
    "sphx_params_from_uniform": (C.c_int, [C.POINTER(SphxUniformData), C.POINTER(SphxParams)]),
    "sphx_state_from_vec4": (C.c_int, [C.POINTER(SphxUniformData), C.c_void_p, C.c_void_p,
                                       C.POINTER(SphxState)]),
    "sphx_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "sphx_create": (C.c_int, [C.c_int, C.c_uint32, C.POINTER(SphxParams), C.POINTER(C.c_void_p)]),
    "sphx_destroy": (None, [C.c_void_p]),
    "sphx_set_params": (C.c_int, [C.c_void_p, C.POINTER(SphxParams)]),
    "sphx_last_error": (C.c_char_p, [C.c_void_p]),
    "sphx_upload": (C.c_int, [C.c_void_p, C.POINTER(SphxState)]),
    "sphx_download": (C.c_int, [C.c_void_p, C.POINTER(SphxState), C.c_uint32]),
    "sphx_neighbours": (C.c_int, [C.c_void_p]),
    "sphx_density": (C.c_int, [C.c_void_p]),
    "sphx_forces": (C.c_int, [C.c_void_p, C.c_uint32]),
    "sphx_vv_position": (C.c_int, [C.c_void_p]),
    "sphx_vv_velocity": (C.c_int, [C.c_void_p]),
    "sphx_xsph": (C.c_int, [C.c_void_p]),
    "sphx_mhd_induction": (C.c_int, [C.c_void_p]),
    "sphx_div_clean": (C.c_int, [C.c_void_p]),
    "sphx_step": (C.c_int, [C.c_void_p, C.c_uint32]),
    "sphx_adjoint": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32] + [C.POINTER(C.c_double)] * 11),
    "sphx_synchronize": (C.c_int, [C.c_void_p]),
    "sphx_checkpoint_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "sphx_save": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "sphx_load": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "sphx_export_pairs": (C.c_int, [C.c_void_p, PU, C.c_uint64, C.POINTER(C.c_uint64)]),
    "sphx_export_pairs_noalloc": (C.c_int, [C.c_void_p, PU, C.c_uint64, C.POINTER(C.c_uint64)]),
    "sphx_export_candidates": (C.c_int, [C.c_void_p, PD, PU, C.c_uint64, C.POINTER(C.c_uint64)]),
    "sphx_enable_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sphx_kernel_times": (C.c_int, [C.c_void_p, C.POINTER(SphxTimes), C.c_int]),
    "sphx_energy": (C.c_int, [C.c_void_p, C.POINTER(SphxEnergy)]),
}

_lib = None


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Load libsphx.so and declare every entry point; raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise OSError(f"{p} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
