"""The brick-staged pair passes against the unstaged ones.

The staged passes copy the neighbours' gather records of a brick of cells
into LDS and walk the same neighbour lists against LDS, reading from global
memory whatever is not staged (a periodic row that wraps in x, a halo larger
than LDS).  In the FP64 form the records, the order of every sum and the
arithmetic are the same, so densities, accelerations and whole trajectories
must be bit-identical to the unstaged passes.  The staged FP32-mixed density
and force passes hold positions relative to the brick's origin cell (the
unstaged ones add each pair's cell offset to cell-relative positions), so
their f32 pair vectors round differently: there the two must agree to well
inside the mixed form's own error against FP64.  SPHX_BRICKS is read when a
context is created: "off" runs the unstaged passes, "bx,by,bz" sets the
brick shape.
"""
import os

import numpy as np
import pytest

import sphx
from sphx import scenarios

pytestmark = pytest.mark.gpu

FIELDS = sphx.F_POS | sphx.F_VEL | sphx.F_ACC | sphx.F_RHO


def run(state, params, bricks, nsteps):
    # bricks None: the library's own choice of brick shape
    old = os.environ.pop("SPHX_BRICKS", None)
    if bricks is not None:
        os.environ["SPHX_BRICKS"] = bricks
    try:
        with sphx.Context(len(state["x"]), params) as ctx:
            ctx.upload(state)
            ctx.forces()
            first = ctx.download(sphx.F_ACC | sphx.F_RHO)
            ctx.step(nsteps)
            ctx.forces()
            last = ctx.download(FIELDS | (sphx.F_U if "u" in state else 0))
    finally:
        os.environ.pop("SPHX_BRICKS", None)
        if old is not None:
            os.environ["SPHX_BRICKS"] = old
    return first, last


def assert_identical(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# staged against unstaged FP32-mixed: relative density (each form within
# 1e-6 of FP64), acceleration relative to the largest (the bar the mixed form
# is held to against FP64), and the state after the ticks relative to its
# largest component
MIXED_TOL = {"rho": 2e-6, "ax": 1e-4, "ay": 1e-4, "az": 1e-4}


def assert_matches(got, ref, precision):
    if precision == sphx.PREC_FP64:
        assert_identical(got, ref)
        return
    for k in ref:
        a, b = got[k], ref[k]
        if k == "rho":
            err = np.abs(a / b - 1.0).max()
        else:
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        assert err <= MIXED_TOL.get(k, 1e-6), (k, err)


def random_box(n, seed, periodic, vsig=0.3):
    rng = scenarios.SplitMix64(seed)
    st = {k: rng.uniform(n) for k in ("x", "y", "z")}
    st.update(vx=vsig * rng.normal(n), vy=vsig * rng.normal(n), vz=vsig * rng.normal(n),
              m=np.full(n, 1.0 / n) * (1.0 + 0.2 * rng.uniform(n)), h=0.04 + 0.03 * rng.uniform(n),
              id=np.argsort(rng.uniform(n)).astype(np.uint32))
    kw = dict(boundary_flags=sphx.BND_PERIODIC, box_lo=(0, 0, 0), box_hi=(1, 1, 1)) if periodic else {}
    p = sphx.make_params(2e-5, tait=sphx.TaitEquationConfig(1.0, 5.0, 7.0),
                         visc=sphx.SphViscosityConfig(0.5, 0.3), **kw)
    return st, p


@pytest.mark.parametrize("shape", ["3,3,3", "1,1,1", "4,2,3", "8,4,4"])
def test_staged_dam_break_bit_identical(require_gpu, shape):
    # uniform h and m; "8,4,4" has a halo of ~6,500 records, more than LDS holds
    st, p = scenarios.dam_break(v_sigma=0.1, seed=3)
    ref = run(st, p, "off", 12)
    got = run(st, p, shape, 12)
    assert_identical(got[0], ref[0])
    assert_identical(got[1], ref[1])


def test_staged_dam_break_1M_bit_identical(require_gpu):
    # the 1,048,576-particle dam break on the library's own brick shape: the
    # force pass reads the slot lists the density pass wrote, also in the
    # bricks whose halo comes close to either pass's record capacity
    st, p = scenarios.dam_break(nx=128, ny=128, nz=64, spacing=0.01, box=(3.2, 1.6, 1.6), v_sigma=0.1, seed=7)
    ref = run(st, p, "off", 8)
    got = run(st, p, None, 8)
    assert_identical(got[0], ref[0])
    assert_identical(got[1], ref[1])


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("precision", [sphx.PREC_FP64, sphx.PREC_FP32_MIXED])
def test_staged_random_variable_h_matches_unstaged(require_gpu, periodic, precision):
    # per-particle h and m; in the periodic box the first and last bricks of
    # each row wrap in x, whose wrapped cell is read from global memory
    st, p = random_box(4000, 11, periodic)
    p.precision = precision
    ref = run(st, p, "off", 6)
    for shape in ("3,3,3", "2,4,1"):
        got = run(st, p, shape, 6)
        assert_matches(got[0], ref[0], precision)
        assert_matches(got[1], ref[1], precision)


def test_staged_mixed_dam_break_matches_unstaged(require_gpu):
    st, p = scenarios.dam_break(v_sigma=0.1, seed=5)
    p.precision = sphx.PREC_FP32_MIXED
    ref = run(st, p, "off", 12)
    got = run(st, p, "3,3,3", 12)
    assert_matches(got[0], ref[0], p.precision)
    assert_matches(got[1], ref[1], p.precision)


@pytest.mark.parametrize("precision", [sphx.PREC_FP64, sphx.PREC_FP32_MIXED])
@pytest.mark.parametrize("variable", [False, True])
def test_staged_adiabatic_gas_matches_unstaged(require_gpu, precision, variable):
    # the energy equation's sums in the staged force passes (u compared too)
    st, p = scenarios.ideal_gas_box(16, variable=variable)
    p.precision = precision
    ref = run(st, p, "off", 6)
    for shape in ("3,3,3", "2,2,2"):
        got = run(st, p, shape, 6)
        assert_matches(got[0], ref[0], precision)
        assert_matches(got[1], ref[1], precision)


@pytest.mark.parametrize("precision", [sphx.PREC_FP64, sphx.PREC_FP32_MIXED])
def test_staged_box_of_three_cells_matches_unstaged(require_gpu, precision):
    # a periodic box three cells wide (uniform h and m, so the slot lists are
    # walked): every halo row of a 3x3x3 or 2x2x2 brick holds the whole period
    # in x, and rows that wrap in y or z appear twice in the halo.  The FP64
    # passes stage such rows (absolute positions, minimum image per pair); the
    # brick-relative FP32-mixed records cannot (a cell would need two images),
    # so those entries come from global memory, converted by their cell codes.
    st, p = scenarios.ideal_gas_box(10)
    p.precision = precision
    # cells at least 2 h + skin wide (skin 0.04 h): three along each axis
    assert int(1.0 / (2.04 * st["h"][0])) == 3
    ref = run(st, p, "off", 4)
    for shape in ("3,3,3", "2,2,2"):
        got = run(st, p, shape, 4)
        assert_matches(got[0], ref[0], precision)
        assert_matches(got[1], ref[1], precision)


@pytest.mark.parametrize("precision", [sphx.PREC_FP64, sphx.PREC_FP32_MIXED])
def test_staged_mhd_periodic_matches_unstaged(require_gpu, precision):
    st, p = scenarios.mhd_periodic(16)
    p.precision = precision
    ref = run(st, p, "off", 4)
    got = run(st, p, "3,3,3", 4)
    assert_matches(got[0], ref[0], precision)
    assert_matches(got[1], ref[1], precision)


@pytest.mark.parametrize("precision", [sphx.PREC_FP64, sphx.PREC_FP32_MIXED])
def test_staged_mhd_variable_h_matches_unstaged(require_gpu, precision):
    # per-particle h and m: the non-uniform Maxwell-stress force passes, whose
    # staged records carry the B records beside the gather records
    st, p = scenarios.mhd_periodic(16)
    rng = scenarios.SplitMix64(7)
    n = len(st["x"])
    st["h"] = st["h"] * (1.0 + 0.1 * rng.uniform(n))
    st["m"] = st["m"] * (1.0 + 0.2 * rng.uniform(n))
    p.precision = precision
    ref = run(st, p, "off", 4)
    for shape in ("3,3,3", "2,2,2"):
        got = run(st, p, shape, 4)
        assert_matches(got[0], ref[0], precision)
        assert_matches(got[1], ref[1], precision)


def pass_sequence(state, params, bricks, nsteps, step):
    # nsteps ticks through sphx_step or through the separate passes in the
    # order of the tick (main.rs:103-157)
    old = os.environ.get("SPHX_BRICKS")
    os.environ["SPHX_BRICKS"] = bricks
    energy = params.eos == sphx.EOS_ADIABATIC and "u" in state
    mhd = "bx" in state
    try:
        with sphx.Context(len(state["x"]), params) as ctx:
            ctx.upload(state)
            if step:
                ctx.step(nsteps)
            else:
                for t in range(nsteps):
                    ctx.forces(visc_only=t > 0 and not energy)
                    ctx.vv_position()
                    ctx.density()
                    ctx.forces()
                    ctx.vv_velocity()
                    if params.xsph_epsilon != 0.0:
                        ctx.xsph()
                    if mhd and params.mhd_flags & sphx.MHD_INDUCTION:
                        ctx.mhd_induction()
                    if mhd and params.mhd_flags & sphx.MHD_CLEANING:
                        ctx.div_clean()
            f = sphx.F_POS | sphx.F_VEL | sphx.F_RHO
            f |= (sphx.F_U if energy else 0) | (sphx.F_B | sphx.F_PSI if mhd else 0)
            out = ctx.download(f)
            ctx.forces()  # a at the end state, from the state each path left
            out.update(ctx.download(sphx.F_ACC))
            return out
    finally:
        if old is None:
            del os.environ["SPHX_BRICKS"]
        else:
            os.environ["SPHX_BRICKS"] = old


@pytest.mark.parametrize("case", ["dam_break", "dam_break_xsph", "random_periodic", "mhd_adiabatic"])
@pytest.mark.parametrize("precision", [sphx.PREC_FP64, sphx.PREC_FP32_MIXED])
@pytest.mark.parametrize("bricks", ["off", "3,3,3"])
def test_step_equals_the_pass_sequence(require_gpu, case, precision, bricks):
    # sphx_step (passes enqueued before the moved flag is read, the kept
    # list, the reused pressure sum) against the separate passes, each of
    # which reads the moved flag before the next: the same bits
    if case.startswith("dam_break"):
        st, p = scenarios.dam_break(v_sigma=0.1, seed=5)
        if case == "dam_break_xsph":
            p.xsph_epsilon = 0.3
        n = 30
    elif case == "random_periodic":
        st, p = random_box(4000, 13, True)
        n = 8
    else:
        st, p = scenarios.mhd_periodic(16)
        n = 6
    p.precision = precision
    a = pass_sequence(st, p, bricks, n, step=True)
    b = pass_sequence(st, p, bricks, n, step=False)
    assert_identical(a, b)
