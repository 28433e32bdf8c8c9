"""ctypes wrapper of liboracle.so — the CPU f64 oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / CPU baseline; the product never does.
Particle index i of every array is the entity with ECS order id[i].
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Dict

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
PD = C.POINTER(C.c_double)
PU = C.POINTER(C.c_uint32)


class OrcState(C.Structure):
    _fields_ = [("n", C.c_uint32)] + [(k, PD) for k in (
        "x", "y", "z", "vx", "vy", "vz", "m", "h", "rho", "p", "ax", "ay", "az",
        "apx", "apy", "apz", "bx", "by", "bz", "psi", "q")] + [("id", PU)] + [
        (k, PD) for k in ("u", "du", "dup")]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OSError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        L.orc_kernel_w.restype = C.c_double
        L.orc_kernel_w.argtypes = [C.c_double, C.c_double]
        L.orc_kernel_dw.restype = C.c_double
        L.orc_kernel_dw.argtypes = [C.c_double, C.c_double]
        for name in ("orc_pairs_brute", "orc_pairs_cells"):
            f = getattr(L, name)
            f.restype = C.c_uint64
            f.argtypes = [C.POINTER(OrcState), C.c_void_p, PU, C.c_uint64]
        for name in ("orc_density", "orc_eval", "orc_xsph", "orc_mhd_induction", "orc_div_clean"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [C.POINTER(OrcState), C.c_void_p, PU, C.c_uint64]
        for name in ("orc_eos", "orc_eval_state", "orc_vv_position", "orc_vv_velocity"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [C.POINTER(OrcState), C.c_void_p]
        L.orc_step.restype = None
        L.orc_step.argtypes = [C.POINTER(OrcState), C.c_void_p, C.c_uint32]
        L.orc_energy.restype = None
        L.orc_energy.argtypes = [C.POINTER(OrcState), C.c_void_p, PD]
        L.orc_adjoint.restype = C.c_int
        L.orc_adjoint.argtypes = [C.POINTER(OrcState), C.c_void_p, C.c_uint32, PD, PD, PD, PD, PD]
        L.orc_loss.restype = C.c_double
        L.orc_loss.argtypes = [C.POINTER(OrcState), C.c_void_p, C.c_uint32, PD]
        _lib = L
    return _lib


def kernel_w(r: float, h: float) -> float:
    return lib().orc_kernel_w(r, h)


def kernel_dw(r: float, h: float) -> float:
    return lib().orc_kernel_dw(r, h)


class State:
    """Owns the numpy arrays behind an OrcState (copies the input dict)."""

    OPTIONAL = ("bx", "by", "bz", "psi", "q")

    def __init__(self, src: Dict[str, np.ndarray]):
        n = len(src["x"])
        self.n = n
        self.a: Dict[str, np.ndarray] = {}
        for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h"):
            self.a[k] = np.array(src[k], dtype=np.float64, copy=True)
        for k in ("rho", "p", "ax", "ay", "az", "apx", "apy", "apz"):
            self.a[k] = np.array(src[k], dtype=np.float64, copy=True) if k in src else np.zeros(n)
        if "bx" in src:
            for k in ("bx", "by", "bz"):
                self.a[k] = np.array(src[k], dtype=np.float64, copy=True)
            self.a["psi"] = np.array(src.get("psi", np.zeros(n)), dtype=np.float64, copy=True)
        if "q" in src:
            self.a["q"] = np.array(src["q"], dtype=np.float64, copy=True)
        if "u" in src:  # adiabatic EOS: specific internal energy and its rates
            self.a["u"] = np.array(src["u"], dtype=np.float64, copy=True)
            for k in ("du", "dup"):
                self.a[k] = np.array(src[k], dtype=np.float64, copy=True) if k in src else np.zeros(n)
        self.id = np.array(src.get("id", np.arange(n)), dtype=np.uint32, copy=True)
        self.c = OrcState()
        self.c.n = n
        for k, v in self.a.items():
            setattr(self.c, k, v.ctypes.data_as(PD))
        self.c.id = self.id.ctypes.data_as(PU)

    def __getitem__(self, k: str) -> np.ndarray:
        return self.a[k]

    def ref(self):
        return C.byref(self.c)


def _pp(params):
    return C.cast(C.byref(params), C.c_void_p)


def pairs(st: State, params, brute: bool = False) -> np.ndarray:
    """Undirected pairs as particle indices (a, b), id[a] < id[b], sorted by ids."""
    f = lib().orc_pairs_brute if brute else lib().orc_pairs_cells
    cap = max(st.n * 64, 1024)
    while True:
        buf = np.empty((cap, 2), dtype=np.uint32)
        np_ = f(st.ref(), _pp(params), buf.ctypes.data_as(PU), cap)
        if np_ <= cap:
            return buf[:np_].copy()
        cap = int(np_)


def id_pairs(st: State, idx_pairs: np.ndarray) -> np.ndarray:
    return st.id[idx_pairs]


def _pairs_args(pr: np.ndarray):
    pr = np.ascontiguousarray(pr, dtype=np.uint32)
    return pr, pr.ctypes.data_as(PU), len(pr)


def density(st: State, params, pr: np.ndarray) -> None:
    """sph_density_system + sph_eos_system on the given pair list."""
    pr, p, n = _pairs_args(pr)
    lib().orc_density(st.ref(), _pp(params), p, n)
    lib().orc_eos(st.ref(), _pp(params))


def evaluate(st: State, params, pr: np.ndarray | None = None) -> None:
    """One force evaluation of the schedule (rho, P, a)."""
    if pr is None:
        lib().orc_eval_state(st.ref(), _pp(params))
    else:
        pr, p, n = _pairs_args(pr)
        lib().orc_eval(st.ref(), _pp(params), p, n)


def vv_position(st: State, params) -> None:
    lib().orc_vv_position(st.ref(), _pp(params))


def vv_velocity(st: State, params) -> None:
    lib().orc_vv_velocity(st.ref(), _pp(params))


def xsph(st: State, params, pr: np.ndarray) -> None:
    pr, p, n = _pairs_args(pr)
    lib().orc_xsph(st.ref(), _pp(params), p, n)


def mhd_induction(st: State, params, pr: np.ndarray) -> None:
    pr, p, n = _pairs_args(pr)
    lib().orc_mhd_induction(st.ref(), _pp(params), p, n)


def div_clean(st: State, params, pr: np.ndarray) -> None:
    pr, p, n = _pairs_args(pr)
    lib().orc_div_clean(st.ref(), _pp(params), p, n)


def step(st: State, params, nsteps: int) -> None:
    lib().orc_step(st.ref(), _pp(params), nsteps)


def energy(st: State, params) -> Dict[str, float]:
    out = np.zeros(8)
    lib().orc_energy(st.ref(), _pp(params), out.ctypes.data_as(PD))
    return {"kinetic": out[0], "potential": out[1], "internal": out[2], "magnetic": out[3],
            "momentum": out[4:7].tolist()}


def adjoint(st: State, params, nsteps: int, target: np.ndarray, with_gx: bool = False):
    """Loss, dL/dv0 (n x 3) and dL/dalpha for L = mean |x_N - target|^2 (and
    dL/dx0 when with_gx); st ends at x_N."""
    tgt = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
    loss = C.c_double(0.0)
    ga = C.c_double(0.0)
    gv = np.zeros(3 * st.n)
    gx = np.zeros(3 * st.n)
    rc = lib().orc_adjoint(st.ref(), _pp(params), nsteps, tgt.ctypes.data_as(PD), C.byref(loss),
                           gx.ctypes.data_as(PD), gv.ctypes.data_as(PD), C.byref(ga))
    if rc != 0:
        raise ValueError("orc_adjoint: unsupported options")
    if with_gx:
        return loss.value, gx.reshape(-1, 3), gv.reshape(-1, 3), ga.value
    return loss.value, gv.reshape(-1, 3), ga.value


def loss(st: State, params, nsteps: int, target: np.ndarray) -> float:
    tgt = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
    return lib().orc_loss(st.ref(), _pp(params), nsteps, tgt.ctypes.data_as(PD))


# ---- the all-core gather form (sph_omp.c) -----------------------------------
def _omp_lib():
    L = lib()
    if not getattr(L, "_omp_ready", False):
        L.orc_omp_new.restype = C.c_void_p
        L.orc_omp_new.argtypes = []
        L.orc_omp_free.restype = None
        L.orc_omp_free.argtypes = [C.c_void_p]
        L.orc_omp_threads.restype = C.c_int
        L.orc_omp_threads.argtypes = []
        L.orc_omp_build.restype = C.c_uint64
        L.orc_omp_build.argtypes = [C.c_void_p, C.POINTER(OrcState), C.c_void_p]
        L.orc_omp_eval.restype = None
        L.orc_omp_eval.argtypes = [C.c_void_p, C.POINTER(OrcState), C.c_void_p]
        L.orc_omp_step.restype = C.c_int
        L.orc_omp_step.argtypes = [C.c_void_p, C.POINTER(OrcState), C.c_void_p, C.c_uint32]
        L._omp_ready = True
    return L


class Gather:
    """Neighbour lists and the OpenMP gather-form evaluation / tick."""

    def __init__(self):
        self.L = _omp_lib()
        self.h = self.L.orc_omp_new()

    def close(self):
        if self.h:
            self.L.orc_omp_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def threads() -> int:
        return _omp_lib().orc_omp_threads()

    def build(self, st: State, params) -> int:
        return int(self.L.orc_omp_build(self.h, st.ref(), _pp(params)))

    def evaluate(self, st: State, params) -> None:
        self.L.orc_omp_eval(self.h, st.ref(), _pp(params))

    def step(self, st: State, params, nsteps: int) -> None:
        if self.L.orc_omp_step(self.h, st.ref(), _pp(params), nsteps) != 0:
            raise ValueError("orc_omp_step: XSPH and the MHD passes are not in the gather form")
