"""Checkpoint / resume (sphx_checkpoint_size, sphx_save, sphx_load) through
the C ABI: a context loaded from a checkpoint continues bit for bit as the
saved context does, for the dam break (open and reflective), the MHD box with
the adiabatic EOS, charge and shuffled ids, and the FP32-mixed form; the
loaded context keeps the caller's upload order; malformed checkpoints are
refused with a status code."""
import ctypes as C

import numpy as np
import pytest

import sphx
from sphx import scenarios

pytestmark = pytest.mark.gpu

ALL = sphx.F_POS | sphx.F_VEL


def fields(state):
    f = ALL | sphx.F_RHO | sphx.F_P | sphx.F_ACC
    if "bx" in state:
        f |= sphx.F_B | sphx.F_PSI
    if "u" in state:
        f |= sphx.F_U
    return f


def cases():
    st, p = scenarios.dam_break(16, 12, 10, v_sigma=0.3, seed=2)
    yield "dam_break_reflective", st, p
    st2 = dict(st)
    p2 = sphx.make_params(p.dt, tait=sphx.TaitEquationConfig(1000.0, 20.0, 7.0),
                          visc=sphx.SphViscosityConfig(0.3, 0.1), gravity=(0.0, 0.0, -9.81))
    yield "dam_break_open", st2, p2  # grid from the uploaded positions
    st3, p3 = scenarios.mhd_periodic(12)
    n = len(st3["x"])
    st3["id"] = np.argsort(scenarios.SplitMix64(5).uniform(n)).astype(np.uint32)
    st3["q"] = np.full(n, 1e-3)
    p3.mhd_flags |= sphx.MHD_LORENTZ_CHARGE
    p3.resistivity = 1e-3
    yield "mhd_adiabatic_charge", st3, p3
    st4, p4 = scenarios.dam_break(16, 12, 10, v_sigma=0.3, seed=3)
    p4.precision = sphx.PREC_FP32_MIXED
    yield "dam_break_mixed", st4, p4


@pytest.mark.parametrize("case", ["dam_break_reflective", "dam_break_open", "mhd_adiabatic_charge",
                                  "dam_break_mixed"])
def test_resume_is_bit_identical(require_gpu, case):
    name, st, p = next(c for c in cases() if c[0] == case)
    n = len(st["x"])
    f = fields(st)
    with sphx.Context(n, p) as a:
        a.upload(st)
        a.step(7)           # lists built some ticks ago: the resume must rebuild the same one
        blob = a.save()
        before = a.download(ALL | (sphx.F_B if "bx" in st else 0) | (sphx.F_U if "u" in st else 0))
        a.step(9)
        a.forces()
        ref = a.download(f)
    other = sphx.make_params(1e-3)  # the load replaces these
    with sphx.Context(n + 10, other) as b:
        b.load(blob)
        got0 = b.download(ALL | (sphx.F_B if "bx" in st else 0) | (sphx.F_U if "u" in st else 0))
        for k in before:
            assert np.array_equal(got0[k], before[k]), k
        b.step(9)
        b.forces()
        got = b.download(f)
        assert b.save()[:8] == b"SPHXCKP1"
    for k in ref:
        assert np.array_equal(got[k], ref[k]), (case, k)


def test_checkpoint_errors(require_gpu):
    st, p = scenarios.dam_break(8, 8, 8)
    n = len(st["x"])
    with sphx.Context(n, p) as a:
        a.upload(st)
        a.step(2)
        blob = a.save()
        lib = a._lib
        size = C.c_uint64(0)
        assert lib.sphx_checkpoint_size(a._h, C.byref(size)) == 0 and size.value == len(blob)
        written = C.c_uint64(0)
        small = C.create_string_buffer(16)
        assert lib.sphx_save(a._h, small, 16, C.byref(written)) == 3 and written.value == len(blob)
    with sphx.Context(n, p) as b:
        for bad in (blob[:100], b"XXXXXXXX" + blob[8:], blob[:-8]):
            with pytest.raises(sphx.SphxError) as e:
                b.load(bad)
            assert e.value.code == 4
        perm = bytearray(blob)
        perm[-8:-4] = perm[-4:]  # two equal upload slots
        with pytest.raises(sphx.SphxError) as e:
            b.load(bytes(perm))
        assert e.value.code == 4
        b.load(blob)  # a refused load leaves the context usable
        b.step(1)
    with sphx.Context(n - 1, p) as c:
        with pytest.raises(sphx.SphxError) as e:
            c.load(blob)
        assert e.value.code == 3
    q = sphx.make_params(p.dt, max_neighbours=256)
    with sphx.Context(n, q) as d:
        with pytest.raises(sphx.SphxError) as e:
            d.load(blob)
        assert e.value.code == 4
