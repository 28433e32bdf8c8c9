"""sphx — Python binding of libsphx.so, the MI355X SPH particle step.

Mirrors the reference's ECS-facing API for the SPH step:
  * the configuration Resources (TaitEquationConfig, SphViscosityConfig,
    XsphConfig, SphBoundaryConfig, MhdConfig — src/physics/sph.rs:125-432,
    src/physics/mhd.rs:26-49) as dataclasses with the reference defaults;
  * one method per ECS system on :class:`Context` (density, forces, ...),
    each calling the C ABI entry point that replaces it.
Every call goes to the HIP library; there is no CPU code path here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from . import _abi
from ._abi import (  # noqa: F401  (re-exported constants)
    BND_PENALTY_CUBE, BND_PERIODIC, BND_REFLECT, EOS_ADIABATIC, EOS_ISOTHERMAL, EOS_POLYTROPIC, EOS_TAIT,
    F_ACC, F_ALL, F_B, F_P, F_POS, F_PSI, F_RHO, F_U, F_VEL, FORCES_FULL, FORCES_VISC_ONLY,
    MHD_CLEANING, MHD_INDUCTION, MHD_LORENTZ_CHARGE, MHD_MAXWELL_STRESS, PREC_FP32_MIXED,
    PREC_FP64, TIMING_ALL, TIMING_DENSITY, TIMING_FORCE, TIMING_INTEGRATE, TIMING_MHD, TIMING_NEIGHBOUR,
    TIMING_SORT, TIMING_VISC_ONLY, VISC_MONAGHAN, VISC_REFERENCE, SphxParams,
)


class SphxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_abi.ERRORS.get(code, code)}: {msg}")
        self.code = code


# ---- reference Resources ------------------------------------------------
@dataclass
class TaitEquationConfig:          # sph.rs:125-142
    reference_density: float = 1000.0
    speed_of_sound: float = 1500.0
    gamma: float = 7.0


@dataclass
class SphViscosityConfig:          # sph.rs:234-248
    alpha: float = 0.1
    beta: float = 0.2


@dataclass
class XsphConfig:                  # sph.rs:328-339
    epsilon: float = 0.5


@dataclass
class SphBoundaryConfig:           # sph.rs:415-432
    half_extent: float = 10.0
    activation_distance: float = 0.5
    stiffness: float = 1e6


@dataclass
class MhdConfig:                   # mhd.rs:26-49
    resistivity: float = 1e-4
    cleaning_speed: float = 100.0
    cleaning_damping: float = 0.5


def default_params() -> SphxParams:
    p = SphxParams()
    _abi.load().sphx_default_params(C.byref(p))
    return p


def make_params(dt: float,
                tait: Optional[TaitEquationConfig] = None,
                visc: Optional[SphViscosityConfig] = None,
                xsph: Optional[XsphConfig] = None,
                boundary: Optional[SphBoundaryConfig] = None,
                mhd: Optional[MhdConfig] = None,
                **extra) -> SphxParams:
    """SphxParams from reference-style configs plus the new fields by name."""
    p = SphxParams()
    _fill_defaults(p)
    p.dt = dt
    tait = tait or TaitEquationConfig()
    p.rest_density, p.speed_of_sound, p.gamma = tait.reference_density, tait.speed_of_sound, tait.gamma
    visc = visc or SphViscosityConfig()
    p.visc_alpha, p.visc_beta = visc.alpha, visc.beta
    p.xsph_epsilon = (xsph or XsphConfig(epsilon=0.0)).epsilon
    b = boundary or SphBoundaryConfig()
    p.half_extent, p.activation_distance, p.stiffness = b.half_extent, b.activation_distance, b.stiffness
    m = mhd or MhdConfig()
    p.resistivity, p.cleaning_speed, p.cleaning_damping = m.resistivity, m.cleaning_speed, m.cleaning_damping
    for k, v in extra.items():
        if not hasattr(p, k):
            raise AttributeError(f"SphxParams has no field {k!r}")
        if k in ("box_lo", "box_hi", "gravity", "e_field", "grid_lo", "grid_hi"):
            setattr(p, k, _abi.D3(*v))
        else:
            setattr(p, k, v)
    return p


def _fill_defaults(p: SphxParams) -> None:
    """Same values as sphx_default_params, without loading the library."""
    p.dt = 1.0 / 60.0
    p.sector_size = 1e12
    p.rest_density, p.speed_of_sound, p.gamma = 1000.0, 1500.0, 7.0
    p.eos, p.visc_convention = EOS_TAIT, VISC_REFERENCE
    p.poly_k, p.iso_cs2 = 1.0, 1.0
    p.visc_alpha, p.visc_beta = 0.1, 0.2
    p.xsph_epsilon = 0.5
    p.half_extent, p.activation_distance, p.stiffness = 10.0, 0.5, 1e6
    p.resistivity, p.cleaning_speed, p.cleaning_damping = 1e-4, 100.0, 0.5
    p.precision, p.max_neighbours = PREC_FP64, 0
    p.neighbour_skin = -1.0


# ---- particle state --------------------------------------------------------
STATE_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "h")
MHD_FIELDS = ("bx", "by", "bz", "psi")


def _ptr(a: Optional[np.ndarray], ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    """One device-resident particle set (a SphxCtx)."""

    def __init__(self, capacity: int, params: SphxParams, device: int = 0):
        self._lib = _abi.load()
        self._h = C.c_void_p()
        self.params = params
        rc = self._lib.sphx_create(device, capacity, C.byref(params), C.byref(self._h))
        if rc != 0:
            raise SphxError(rc, "sphx_create failed")
        self.n = 0
        self._mhd = False

    # -- plumbing
    def _check(self, rc: int) -> None:
        if rc != 0:
            raise SphxError(rc, self._lib.sphx_last_error(self._h).decode())

    def close(self) -> None:
        if self._h:
            self._lib.sphx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: SphxParams) -> None:
        self.params = params
        self._check(self._lib.sphx_set_params(self._h, C.byref(params)))

    # -- data
    def upload(self, state: Dict[str, np.ndarray]) -> None:
        arrs = {k: np.ascontiguousarray(state[k], dtype=np.float64) for k in STATE_FIELDS}
        n = len(arrs["x"])
        st = _abi.SphxState()
        st.n = n
        for k, a in arrs.items():
            setattr(st, k, _ptr(a))
        keep = [arrs]
        if "bx" in state:
            mh = {k: np.ascontiguousarray(state[k], dtype=np.float64) for k in ("bx", "by", "bz")}
            mh["psi"] = np.ascontiguousarray(state.get("psi", np.zeros(n)), dtype=np.float64)
            for k, a in mh.items():
                setattr(st, k, _ptr(a))
            keep.append(mh)
            self._mhd = True
        else:
            self._mhd = False
        if "q" in state:
            qa = np.ascontiguousarray(state["q"], dtype=np.float64)
            st.q = _ptr(qa)
            keep.append(qa)
        if "id" in state:
            ida = np.ascontiguousarray(state["id"], dtype=np.uint32)
            st.id = _ptr(ida, C.c_uint32)
            keep.append(ida)
        if "u" in state:
            ua = np.ascontiguousarray(state["u"], dtype=np.float64)
            st.u = _ptr(ua)
            keep.append(ua)
        self._check(self._lib.sphx_upload(self._h, C.byref(st)))
        self.n = n

    def download(self, fields: int = F_POS | F_VEL) -> Dict[str, np.ndarray]:
        n = self.n
        out: Dict[str, np.ndarray] = {}
        st = _abi.SphxState()
        st.n = n
        groups = [(F_POS, ("x", "y", "z")), (F_VEL, ("vx", "vy", "vz")), (F_RHO, ("rho",)),
                  (F_P, ("p",)), (F_ACC, ("ax", "ay", "az")), (F_B, ("bx", "by", "bz")),
                  (F_PSI, ("psi",)), (F_U, ("u",))]
        for bit, names in groups:
            if fields & bit:
                for k in names:
                    out[k] = np.empty(n, dtype=np.float64)
                    setattr(st, k, _ptr(out[k]))
        self._check(self._lib.sphx_download(self._h, C.byref(st), fields))
        return out

    def update(self, state: Dict[str, np.ndarray], fields: int) -> None:
        """sphx_update: new velocities (F_VEL), B (F_B), psi (F_PSI) and / or internal energy (F_U), upload order."""
        st = _abi.SphxState()
        st.n = self.n
        keep = []
        groups = [(F_VEL, ("vx", "vy", "vz")), (F_B, ("bx", "by", "bz")), (F_PSI, ("psi",)), (F_U, ("u",))]
        for bit, names in groups:
            if fields & bit:
                for k in names:
                    a = np.ascontiguousarray(state[k], dtype=np.float64)
                    keep.append(a)
                    setattr(st, k, _ptr(a))
        self._check(self._lib.sphx_update(self._h, C.byref(st), fields))

    # -- systems
    def neighbours(self) -> None:
        self._check(self._lib.sphx_neighbours(self._h))

    def density(self) -> None:
        self._check(self._lib.sphx_density(self._h))

    def forces(self, visc_only: bool = False) -> None:
        self._check(self._lib.sphx_forces(self._h, FORCES_VISC_ONLY if visc_only else FORCES_FULL))

    def vv_position(self) -> None:
        self._check(self._lib.sphx_vv_position(self._h))

    def vv_velocity(self) -> None:
        self._check(self._lib.sphx_vv_velocity(self._h))

    def xsph(self) -> None:
        self._check(self._lib.sphx_xsph(self._h))

    def mhd_induction(self) -> None:
        self._check(self._lib.sphx_mhd_induction(self._h))

    def div_clean(self) -> None:
        self._check(self._lib.sphx_div_clean(self._h))

    def step(self, nsteps: int = 1) -> None:
        self._check(self._lib.sphx_step(self._h, nsteps))

    def synchronize(self) -> None:
        self._check(self._lib.sphx_synchronize(self._h))

    # -- checkpoint / resume
    def save(self) -> bytes:
        """sphx_save: the checkpoint of the current state as bytes."""
        size = C.c_uint64(0)
        self._check(self._lib.sphx_checkpoint_size(self._h, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        written = C.c_uint64(0)
        self._check(self._lib.sphx_save(self._h, buf, size.value, C.byref(written)))
        return buf.raw[: written.value]

    def load(self, blob: bytes) -> None:
        """sphx_load: parameters and state from a checkpoint; the context continues bit for bit."""
        self._check(self._lib.sphx_load(self._h, blob, len(blob)))
        self.n = int.from_bytes(blob[20:24], "little")  # the header's particle count

    def adjoint(self, nsteps: int, target: np.ndarray, checkpoint_every: int = 0,
                with_grad_x: bool = False) -> Dict[str, object]:
        """Loss mean|x_N - target|^2 and its gradient in the initial velocities
        (n x 3), the viscosity alpha and (with_grad_x) the initial positions.
        target is n x 3 in upload order; the context ends at x_N."""
        n = self.n
        tgt = np.ascontiguousarray(np.asarray(target, dtype=np.float64).reshape(n, 3).T)
        gv = np.zeros((3, n))
        gx = np.zeros((3, n)) if with_grad_x else None
        loss = C.c_double(0.0)
        ga = C.c_double(0.0)
        d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._check(self._lib.sphx_adjoint(
            self._h, nsteps, checkpoint_every, d(tgt[0]), d(tgt[1]), d(tgt[2]), C.byref(loss),
            d(gx[0]) if gx is not None else None, d(gx[1]) if gx is not None else None,
            d(gx[2]) if gx is not None else None, d(gv[0]), d(gv[1]), d(gv[2]), C.byref(ga)))
        out = {"loss": loss.value, "grad_v": gv.T.copy(), "grad_alpha": ga.value}
        if gx is not None:
            out["grad_x"] = gx.T.copy()
        return out

    # -- measurement
    def export_pairs(self, noalloc: bool = False) -> np.ndarray:
        """Sorted (id_a, id_b) neighbour pairs; noalloc: through sphx_export_pairs_noalloc."""
        fn = self._lib.sphx_export_pairs_noalloc if noalloc else self._lib.sphx_export_pairs
        count = C.c_uint64(0)
        rc = fn(self._h, None, 0, C.byref(count))
        if rc not in (0, 3):  # SPHX_E_CAPACITY is expected with cap 0
            self._check(rc)
        out = np.empty((max(count.value, 1), 2), dtype=np.uint32)
        self._check(fn(self._h, _ptr(out, C.c_uint32), count.value, C.byref(count)))
        return out[: count.value]

    def export_candidates(self, radius: np.ndarray) -> np.ndarray:
        """sphx_export_candidates: broadphase pairs within 4 R of the smaller id (upload-order radii)."""
        r = np.ascontiguousarray(radius, dtype=np.float64)
        count = C.c_uint64(0)
        rc = self._lib.sphx_export_candidates(self._h, _ptr(r), None, 0, C.byref(count))
        if rc not in (0, 3):  # SPHX_E_CAPACITY is expected with cap 0
            self._check(rc)
        out = np.empty((max(count.value, 1), 2), dtype=np.uint32)
        self._check(self._lib.sphx_export_candidates(self._h, _ptr(r), _ptr(out, C.c_uint32), count.value,
                                                     C.byref(count)))
        return out[: count.value]

    def enable_timing(self, on=True) -> None:
        """True: every kernel; False: none; an int: the TIMING_* slots to time."""
        mask = (TIMING_ALL if on else 0) if isinstance(on, bool) else int(on)
        self._check(self._lib.sphx_enable_timing(self._h, mask))

    def kernel_times(self, reset: bool = False) -> Dict[str, object]:
        t = _abi.SphxTimes()
        self._check(self._lib.sphx_kernel_times(self._h, C.byref(t), 1 if reset else 0))
        names = ["sort", "neighbour", "density", "force", "visc_only", "integrate", "mhd"]
        ms = [t.sort_ms, t.neighbour_ms, t.density_ms, t.force_ms, t.visc_only_ms, t.integrate_ms, t.mhd_ms]
        return {k: {"ms": v, "launches": int(t.launches[i])} for i, (k, v) in enumerate(zip(names, ms))}

    def energy(self) -> Dict[str, object]:
        e = _abi.SphxEnergy()
        self._check(self._lib.sphx_energy(self._h, C.byref(e)))
        return {"kinetic": e.kinetic, "potential": e.potential, "internal": e.internal,
                "magnetic": e.magnetic, "momentum": list(e.momentum),
                "max_neighbours": int(e.max_neighbours), "total_neighbours": int(e.total_neighbours)}


def device_count() -> int:
    c = C.c_int(0)
    rc = _abi.load().sphx_device_count(C.byref(c))
    return c.value if rc == 0 else 0
