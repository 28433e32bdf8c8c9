"""The C ABI without a GPU: the library loads, exports every symbol include/sphx.h
declares, and its host-side error paths return status codes instead of aborting."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import sphx
from sphx import _abi

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "sphx.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sphx_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_abi.SIGNATURES)


def test_header_constants_and_binding_agree():
    # every SPHX_<NAME> integer #define of the header is the binding's <NAME>
    text = (ROOT / "include" / "sphx.h").read_text()
    found = 0
    for name, expr in re.findall(r"#define\s+SPHX_([A-Z0-9_]+)\s+(.+)", text):
        expr = re.sub(r"/\*.*", "", expr).strip()
        if not re.fullmatch(r"[0-9()<> ]+", expr):
            continue
        assert getattr(_abi, name, None) == eval(expr), name
        found += 1
    assert found >= 8


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_abi.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing


def test_struct_layout_matches_header(tmp_path):
    # offsetof/sizeof as the C compiler lays out include/sphx.h, against ctypes
    structs = [_abi.SphxParams, _abi.SphxState, _abi.SphxTimes, _abi.SphxEnergy, _abi.SphxUniformData]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sphx.h"', "int main(void) {"]
    for s in structs:
        lines.append(f'printf("{s.__name__} sizeof %zu\\n", sizeof({s.__name__}));')
        for name, _ in s._fields_:
            lines.append(f'printf("{s.__name__} {name} %zu\\n", offsetof({s.__name__}, {name}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            s, f, v = line.split()
            got[(s, f)] = int(v)
    for s in structs:
        assert got[(s.__name__, "sizeof")] == C.sizeof(s)
        for name, _ in s._fields_:
            assert got[(s.__name__, name)] == getattr(s, name).offset, (s.__name__, name)


def test_default_params_match_reference_defaults():
    p = sphx.default_params()
    # TaitEquationConfig::default (sph.rs:134-142), SphViscosityConfig (241-248),
    # XsphConfig (333-339), SphBoundaryConfig (424-432), MhdConfig (mhd.rs:41-49)
    assert (p.rest_density, p.speed_of_sound, p.gamma) == (1000.0, 1500.0, 7.0)
    assert (p.visc_alpha, p.visc_beta) == (0.1, 0.2)
    assert p.xsph_epsilon == 0.5
    assert (p.half_extent, p.activation_distance, p.stiffness) == (10.0, 0.5, 1e6)
    assert (p.resistivity, p.cleaning_speed, p.cleaning_damping) == (1e-4, 100.0, 0.5)
    assert p.dt == 1.0 / 60.0
    q = sphx.SphxParams()
    sphx._fill_defaults(q)
    assert bytes(q) == bytes(p)


def test_no_device_is_an_error_code_not_a_crash():
    if sphx.device_count() > 0:
        pytest.skip("a device is visible")
    lib = _abi.load()
    p = sphx.default_params()
    h = C.c_void_p()
    rc = lib.sphx_create(0, 16, C.byref(p), C.byref(h))
    assert rc == 1  # SPHX_E_NODEVICE
    assert not h.value
    with pytest.raises(sphx.SphxError):
        sphx.Context(16, p)


def test_invalid_params_rejected_before_device():
    lib = _abi.load()
    p = sphx.default_params()
    p.dt = -1.0
    h = C.c_void_p()
    assert lib.sphx_create(0, 16, C.byref(p), C.byref(h)) == 4  # SPHX_E_INVALID
    p = sphx.default_params()
    p.boundary_flags = sphx.BND_PERIODIC | sphx.BND_REFLECT
    p.box_hi = _abi.D3(1, 1, 1)
    assert lib.sphx_create(0, 16, C.byref(p), C.byref(h)) == 4
    assert lib.sphx_create(0, 16, None, C.byref(h)) == 4
    assert lib.sphx_last_error(None) == b"null context"
    # null-context entry points return SPHX_E_INVALID
    assert lib.sphx_step(None, 1) == 4
    assert lib.sphx_density(None) == 4
    lib.sphx_destroy(None)


def test_scenarios_are_deterministic_and_shaped():
    from sphx import scenarios
    s1, p1 = scenarios.dam_break()
    s2, _ = scenarios.dam_break()
    assert len(s1["x"]) == 16384
    assert all(np.array_equal(s1[k], s2[k]) for k in s1)
    assert p1.dt == 1e-4 and p1.speed_of_sound == 20.0 and p1.visc_beta == 0.0
    assert np.abs(s1["x"] - np.round(s1["x"] / 0.01 - 0.5) * 0.01 - 0.005).max() <= 1e-4 + 1e-15
    m, pm = scenarios.mhd_periodic(16)
    assert len(m["x"]) == 4096 and (m["x"] >= 0).all() and (m["x"] < 1).all()
    assert pm.boundary_flags == sphx.BND_PERIODIC
    # the SplitMix64 stream: first output for seed 0 (published test vector)
    assert int(scenarios.SplitMix64(0).next_u64(1)[0]) == 0xE220A8397B1DCDAF


def test_uniform_block_is_the_wgsl_layout():
    # SphParams in src/gpu/sph.wgsl:1-12: ten 4-byte fields, no padding between
    assert C.sizeof(_abi.SphxUniformData) == 40
    assert [f for f, _ in _abi.SphxUniformData._fields_] == [
        "num_particles", "mass", "h", "dt", "rest_density", "speed_of_sound", "gamma", "alpha_visc",
        "beta_visc", "_padding"]


SPHX_E_INVALID = 4


def uniform(n=5, **kw):
    vals = dict(num_particles=n, mass=0.125, h=0.0265, dt=1e-4, rest_density=998.5, speed_of_sound=88.3,
                gamma=7.0, alpha_visc=0.35, beta_visc=0.7)
    vals.update(kw)
    return _abi.SphxUniformData(**vals)


def test_params_from_uniform_widens_and_keeps_the_rest():
    lib = _abi.load()
    u = uniform()
    p = sphx.default_params()
    p.stiffness, p.neighbour_skin = 123.0, 0.5
    assert lib.sphx_params_from_uniform(C.byref(u), C.byref(p)) == _abi.SPHX_OK
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    assert (p.dt, p.rest_density, p.speed_of_sound, p.gamma, p.visc_alpha, p.visc_beta) == (
        f32(1e-4), f32(998.5), f32(88.3), 7.0, f32(0.35), f32(0.7))
    assert (p.stiffness, p.neighbour_skin, p.eos) == (123.0, 0.5, _abi.EOS_TAIT)
    before = bytes(p)
    for bad in (dict(dt=0.0), dict(rest_density=-1.0), dict(speed_of_sound=float("nan")), dict(gamma=0.0)):
        assert lib.sphx_params_from_uniform(C.byref(uniform(**bad)), C.byref(p)) == SPHX_E_INVALID
        assert bytes(p) == before
    assert lib.sphx_params_from_uniform(None, C.byref(p)) == SPHX_E_INVALID


def test_state_from_vec4_fills_the_soa_arrays():
    lib = _abi.load()
    n = 7
    rng = np.random.default_rng(3)
    pos4 = rng.normal(size=(n, 4)).astype(np.float32)
    vel4 = rng.normal(size=(n, 4)).astype(np.float32)
    arrs = {k: np.full(n, -1.0) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    ids = np.full(n, 99, dtype=np.uint32)
    st = _abi.SphxState()
    for k, a in arrs.items():
        setattr(st, k, a.ctypes.data_as(_abi.PD))
    st.id = ids.ctypes.data_as(_abi.PU)
    u = uniform(n)
    rc = lib.sphx_state_from_vec4(C.byref(u), pos4.ctypes.data, vel4.ctypes.data, C.byref(st))
    assert rc == _abi.SPHX_OK and st.n == n
    for c, k in enumerate(("x", "y", "z")):
        assert np.array_equal(arrs[k], pos4[:, c].astype(np.float64))
        assert np.array_equal(arrs["v" + k], vel4[:, c].astype(np.float64))
    assert np.all(arrs["m"] == np.float32(0.125)) and np.all(arrs["h"] == np.float32(0.0265))
    assert np.array_equal(ids, np.arange(n, dtype=np.uint32))
    # without id the ids are left to the caller; bad inputs are status codes
    st.id = None
    assert lib.sphx_state_from_vec4(C.byref(u), pos4.ctypes.data, vel4.ctypes.data, C.byref(st)) == _abi.SPHX_OK
    assert lib.sphx_state_from_vec4(C.byref(u), None, vel4.ctypes.data, C.byref(st)) == SPHX_E_INVALID
    assert lib.sphx_state_from_vec4(C.byref(uniform(n, h=0.0)), pos4.ctypes.data, vel4.ctypes.data,
                                    C.byref(st)) == SPHX_E_INVALID
    st.m = None
    assert lib.sphx_state_from_vec4(C.byref(u), pos4.ctypes.data, vel4.ctypes.data, C.byref(st)) == SPHX_E_INVALID


def test_null_context_is_an_error_code():
    # every entry point taking a context answers a NULL one with SPHX_E_INVALID
    lib = _abi.load()
    st = _abi.SphxState()
    count = C.c_uint64(7)
    r = np.ones(4)
    assert lib.sphx_update(None, C.byref(st), 2) == SPHX_E_INVALID
    assert lib.sphx_export_candidates(None, r.ctypes.data_as(_abi.PD), None, 0, C.byref(count)) == SPHX_E_INVALID
    assert lib.sphx_export_pairs(None, None, 0, C.byref(count)) == SPHX_E_INVALID
    assert lib.sphx_download(None, C.byref(st), 1) == SPHX_E_INVALID
    assert lib.sphx_upload(None, C.byref(st)) == SPHX_E_INVALID
