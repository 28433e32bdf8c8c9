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
    return sorted(set(re.findall(r"\b(sphx_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_abi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_abi.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing


def test_struct_layout_matches_header(tmp_path):
    # offsetof/sizeof as the C compiler lays out include/sphx.h, against ctypes
    structs = [_abi.SphxParams, _abi.SphxState, _abi.SphxTimes, _abi.SphxEnergy]
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
