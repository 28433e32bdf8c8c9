"""Parity of the HIP path (libsphx.so on the GPU) against the CPU oracle.

Bars:
  * neighbour sets: bit-exact equality with the oracle's pair lists;
  * one f64 evaluation: rho within 1e-13 relative; accelerations within
    1e-11 of the largest |a| (gather vs. pair-scatter summation order);
  * trajectories: stated per test (chaotic growth of summation-order noise);
  * integrators given identical accelerations: bit-exact.
"""
import math

import os

import numpy as np
import pytest

import oracle as O
import sphx
from sphx import scenarios

pytestmark = pytest.mark.gpu


def make_ctx(state, params, cap=None):
    ctx = sphx.Context(cap or max(len(state["x"]), 1), params)
    ctx.upload(state)
    return ctx


def with_velocity(state, sigma, seed=7):
    rng = scenarios.SplitMix64(seed)
    n = len(state["x"])
    s = dict(state)
    s["vx"], s["vy"], s["vz"] = sigma * rng.normal(n), sigma * rng.normal(n), sigma * rng.normal(n)
    return s


def oracle_pairs_ids(state, params, brute=False):
    s = O.State(state)
    return O.id_pairs(s, O.pairs(s, params, brute=brute))


def shuffle_ids(state, seed=31):
    """ids in a seeded random order: the ECS rank then differs from the upload slot,
    so a pass that orients pairs by the slot instead of the id fails."""
    s = dict(state)
    s["id"] = np.argsort(scenarios.SplitMix64(seed).uniform(len(state["x"]))).astype(np.uint32)
    assert (s["id"] != np.arange(len(s["id"]))).mean() > 0.9
    return s


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ------------------------------------------------------------- neighbour sets
def test_pairs_bit_exact_dam_break(require_gpu):
    st, p = scenarios.dam_break()  # 16,384 particles
    with make_ctx(st, p) as ctx:
        got = ctx.export_pairs()
    ref = oracle_pairs_ids(st, p)
    assert len(ref) > 500_000
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("periodic", [False, True])
def test_pairs_bit_exact_random_variable_h(require_gpu, periodic):
    rng = scenarios.SplitMix64(21)
    n = 3000
    st = {k: rng.uniform(n) for k in ("x", "y", "z")}
    st.update(vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n), m=np.ones(n),
              h=0.02 + 0.03 * rng.uniform(n), id=np.argsort(rng.uniform(n)).astype(np.uint32))
    kw = dict(boundary_flags=sphx.BND_PERIODIC, box_lo=(0, 0, 0), box_hi=(1, 1, 1)) if periodic else {}
    p = sphx.make_params(1e-4, **kw)
    with make_ctx(st, p) as ctx:
        got = ctx.export_pairs()
    ref = oracle_pairs_ids(st, p, brute=True)
    assert len(ref) > 5000
    assert np.array_equal(got, ref)


def test_pairs_dense_periodic_chunked(require_gpu):
    # cell_size 0.25 -> 4 cells per axis, ~156 particles per cell: a 256-particle
    # tile spans 2-3 cells, so each neighbour row is the whole periodic grid row
    # (~625 candidates), more than one LDS chunk (384) of the neighbour build: staging runs in two
    # chunks.  h = 0.06 keeps ~72 neighbours per particle.
    rng = scenarios.SplitMix64(33)
    n = 10000
    st = {k: rng.uniform(n) for k in ("x", "y", "z")}
    st.update(vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n), m=np.full(n, 1.0 / n), h=np.full(n, 0.06))
    p = sphx.make_params(1e-4, boundary_flags=sphx.BND_PERIODIC, box_lo=(0, 0, 0), box_hi=(1, 1, 1),
                         cell_size=0.25, max_neighbours=160)
    with make_ctx(st, p) as ctx:
        got = ctx.export_pairs()
        ctx.density()
        rho = ctx.download(sphx.F_RHO)["rho"]
    s = O.State(st)
    pr = O.pairs(s, p, brute=True)
    assert np.array_equal(got, O.id_pairs(s, pr))
    O.density(s, p, pr)
    assert rel_err(rho, s["rho"]) < 1e-13


def test_timing_selected_slots(require_gpu):
    # only the chosen kernels are bracketed with events; unknown bits are an error
    st, p = scenarios.dam_break(nx=16, ny=16, nz=8, v_sigma=0.1, seed=2)
    with make_ctx(st, p) as ctx:
        ctx.enable_timing(sphx.TIMING_FORCE | sphx.TIMING_VISC_ONLY)
        ctx.step(3)
        t = ctx.kernel_times(reset=True)
        assert t["force"]["launches"] >= 3 and t["visc_only"]["launches"] >= 1
        assert all(t[k]["launches"] == 0 for k in ("sort", "neighbour", "density", "integrate", "mhd"))
        ctx.enable_timing(True)
        ctx.step(1)
        t = ctx.kernel_times(reset=True)
        assert t["density"]["launches"] >= 1 and t["integrate"]["launches"] == 2
        with pytest.raises(sphx.SphxError):
            ctx.enable_timing(1 << 9)


def test_skin_lists_keep_the_exact_neighbour_set(require_gpu):
    # a large skin keeps one list for many ticks of a moving fluid: at every
    # later state the exported set is still exactly {r < 2 h_avg}, and the
    # trajectory equals the one with a list rebuilt every tick
    st, p = scenarios.dam_break(nx=16, ny=16, nz=8, v_sigma=0.3, seed=7)
    n = len(st["x"])
    p.neighbour_skin = 0.3 * 0.013
    with make_ctx(st, p) as ctx:
        ctx.enable_timing(True)
        for _ in range(4):
            ctx.step(5)
            cur = ctx.download(sphx.F_POS)
            s = dict(st, x=cur["x"], y=cur["y"], z=cur["z"])
            assert np.array_equal(ctx.export_pairs(), oracle_pairs_ids(s, p, brute=True))
        builds = ctx.kernel_times()["neighbour"]["launches"]
        a = ctx.download(sphx.F_POS | sphx.F_VEL)
    assert 1 <= builds < 20  # the list outlived several ticks
    p0 = scenarios.dam_break(nx=16, ny=16, nz=8, v_sigma=0.3, seed=7)[1]
    p0.neighbour_skin = 0.0
    with make_ctx(st, p0) as ctx:
        ctx.step(20)
        b = ctx.download(sphx.F_POS | sphx.F_VEL)
    for k in ("x", "y", "z"):
        assert rel_err(a[k], b[k]) < 1e-12, k
    for k in ("vx", "vy", "vz"):
        assert rel_err(a[k], b[k]) < 1e-9, k


@pytest.mark.parametrize("precision", [sphx.PREC_FP64, sphx.PREC_FP32_MIXED])
def test_speculative_ticks_bit_identical(require_gpu, precision):
    # A tick may enqueue density and the second force pass before the host
    # has read the moved-skin/2 flag, and redo them on a rebuilt list when it
    # is set.  Forced on every tick and forced off, with a small skin (many
    # rebuilds, each one a redo when forced on), the trajectories are equal.
    st, p = scenarios.dam_break(nx=16, ny=16, nz=8, v_sigma=0.3, seed=7)
    p.neighbour_skin = 0.02 * 0.013
    p.precision = precision
    out = {}
    old = os.environ.get("SPHX_SPECULATE")
    try:
        for mode in ("always", "never"):
            os.environ["SPHX_SPECULATE"] = mode
            with make_ctx(st, p) as ctx:
                ctx.enable_timing(True)
                ctx.step(30)
                t = ctx.kernel_times()
                ctx.forces()
                out[mode] = (ctx.download(sphx.F_POS | sphx.F_VEL | sphx.F_ACC | sphx.F_RHO),
                             t["neighbour"]["launches"], t["density"]["launches"])
    finally:
        if old is None:
            del os.environ["SPHX_SPECULATE"]
        else:
            os.environ["SPHX_SPECULATE"] = old
    (a, builds_a, dens_a), (b, builds_b, dens_b) = out["always"], out["never"]
    assert builds_a == builds_b >= 4
    assert dens_a == dens_b + (builds_a - 1)  # every rebuild after the first was a redo
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_pairs_exactly_at_cutoff(require_gpu):
    x, y, z = scenarios.lattice(6, 6, 6, 1.0)
    n = x.size
    base = dict(x=x, y=y, z=z, vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n), m=np.ones(n))
    p = sphx.make_params(1e-3)
    for h, expect in ((0.5, 0), (0.5 * (1 + 2 ** -52), 3 * 6 * 6 * 5)):
        st = dict(base, h=np.full(n, h))
        with make_ctx(st, p) as ctx:
            got = ctx.export_pairs()
        assert len(got) == expect
        assert np.array_equal(got, oracle_pairs_ids(st, p, brute=True).reshape(-1, 2))


# ------------------------------------------------------------- one evaluation
def test_density_parity_dam_break(require_gpu):
    st, p = scenarios.dam_break()
    with make_ctx(st, p) as ctx:
        ctx.density()
        g = ctx.download(sphx.F_RHO | sphx.F_P)
    s = O.State(st)
    O.density(s, p, O.pairs(s, p))
    assert rel_err(g["rho"], s["rho"]) < 1e-13
    assert np.abs(g["p"] - s["p"]).max() <= 1e-11 * max(s["p"].max(), 1.0)


@pytest.mark.parametrize("conv", [sphx.VISC_REFERENCE, sphx.VISC_MONAGHAN])
def test_force_parity_dam_break(require_gpu, conv):
    st, p = scenarios.dam_break(nx=24, ny=24, nz=24)
    st = with_velocity(st, 0.3)
    # compress slightly so Tait pressures are non-zero everywhere inside
    st["m"] = st["m"] * 1.05
    p.visc_alpha, p.visc_beta, p.visc_convention = 0.5, 0.3, conv
    with make_ctx(st, p) as ctx:
        ctx.forces()
        g = ctx.download(sphx.F_ACC | sphx.F_RHO)
    s = O.State(st)
    O.evaluate(s, p)
    assert rel_err(g["rho"], s["rho"]) < 1e-13
    for k in ("ax", "ay", "az"):
        scale = np.abs(s[k]).max()
        assert scale > 1.0
        assert np.abs(g[k] - s[k]).max() < 1e-11 * scale, k


def test_visc_only_reuse_is_bit_identical(require_gpu):
    st, p = scenarios.dam_break()
    st = with_velocity(st, 0.2)
    with make_ctx(st, p) as ctx:
        ctx.forces()
        a1 = ctx.download(sphx.F_ACC)
        ctx.forces(visc_only=True)
        a2 = ctx.download(sphx.F_ACC)
    for k in a1:
        assert np.array_equal(a1[k], a2[k])


def test_penalty_walls_lorentz_parity(require_gpu):
    st, p = scenarios.dam_break(nx=16, ny=16, nz=16, spacing=0.05)
    n = len(st["x"])
    st = with_velocity(st, 0.5)
    rng = scenarios.SplitMix64(9)
    st.update(bx=rng.normal(n), by=rng.normal(n), bz=rng.normal(n), q=rng.normal(n) * 1e-3)
    p.boundary_flags = sphx.BND_PENALTY_CUBE
    p.half_extent, p.activation_distance, p.stiffness = 0.45, 0.1, 50.0
    p.mhd_flags = sphx.MHD_LORENTZ_CHARGE
    p.e_field = sphx._abi.D3(0.1, -0.2, 0.3)
    with make_ctx(st, p) as ctx:
        ctx.forces()
        g = ctx.download(sphx.F_ACC)
    s = O.State(st)
    O.evaluate(s, p)
    for k in ("ax", "ay", "az"):
        assert np.abs(g[k] - s[k]).max() < 1e-11 * np.abs(s[k]).max(), k


# ------------------------------------------------------------- MHD
def test_mhd_passes_parity(require_gpu):
    st, p = scenarios.mhd_periodic(16)
    st = shuffle_ids(st)
    p.resistivity = 1e-3  # exercise the resistive term too
    with make_ctx(st, p) as ctx:
        ctx.forces()
        a = ctx.download(sphx.F_ACC | sphx.F_RHO)
        ctx.mhd_induction()
        b1 = ctx.download(sphx.F_B)
        ctx.div_clean()
        b2 = ctx.download(sphx.F_B | sphx.F_PSI)
    s = O.State(st)
    pr = O.pairs(s, p)
    O.evaluate(s, p, pr)
    assert rel_err(a["rho"], s["rho"]) < 1e-13
    for k in ("ax", "ay", "az"):
        assert np.abs(a[k] - s[k]).max() < 1e-11 * np.abs(s[k]).max(), k
    b0 = {k: st[k].copy() for k in ("bx", "by", "bz")}
    O.mhd_induction(s, p, pr)
    for k in ("bx", "by", "bz"):
        d_ref = s[k] - b0[k]
        assert np.abs(b1[k] - s[k]).max() < 1e-12 * np.abs(d_ref).max() + 1e-15, k
    psi0 = s["psi"].copy()
    bb = {k: s[k].copy() for k in ("bx", "by", "bz")}
    O.div_clean(s, p, pr)
    for k in ("bx", "by", "bz"):
        assert np.abs(b2[k] - s[k]).max() < 1e-12 * max(np.abs(s[k] - bb[k]).max(), 1e-300) + 1e-15, k
    assert np.abs(b2["psi"] - s["psi"]).max() < 1e-12 * np.abs(s["psi"] - psi0).max()


def test_mixed_mhd_passes_error(require_gpu):
    # FP32-mixed induction and cleaning: f32 pair terms, f64 sums and update.
    # The change of B (and psi) is compared with the oracle's f64 change at
    # the same state, relative to its largest component.
    st, p = scenarios.mhd_periodic(16)
    p.resistivity = 1e-3
    p.precision = sphx.PREC_FP32_MIXED
    with make_ctx(st, p) as ctx:
        ctx.forces()
        ctx.mhd_induction()
        b1 = ctx.download(sphx.F_B)
        ctx.div_clean()
        b2 = ctx.download(sphx.F_B | sphx.F_PSI)
    p.precision = sphx.PREC_FP64
    s = O.State(st)
    pr = O.pairs(s, p)
    O.evaluate(s, p, pr)
    b0 = {k: st[k].copy() for k in ("bx", "by", "bz")}
    O.mhd_induction(s, p, pr)
    for k in ("bx", "by", "bz"):
        d_ref = s[k] - b0[k]
        assert np.abs(b1[k] - s[k]).max() < 2e-5 * np.abs(d_ref).max(), k
    psi0 = s["psi"].copy()
    bb = {k: s[k].copy() for k in ("bx", "by", "bz")}
    O.div_clean(s, p, pr)
    # cleaning's own change of B (zero here: psi starts at 0) and of psi
    for k in ("bx", "by", "bz"):
        d_ref = s[k] - bb[k]
        assert np.abs((b2[k] - b1[k]) - d_ref).max() <= 2e-5 * np.abs(d_ref).max(), k
    d_psi = s["psi"] - psi0
    assert np.abs(b2["psi"] - s["psi"]).max() < 1e-4 * np.abs(d_psi).max()


def test_xsph_parity(require_gpu):
    st, p = scenarios.dam_break(nx=16, ny=16, nz=16)
    st = with_velocity(st, 0.4)
    p.xsph_epsilon = 0.5
    with make_ctx(st, p) as ctx:
        ctx.density()
        ctx.xsph()
        g = ctx.download(sphx.F_VEL)
    s = O.State(st)
    pr = O.pairs(s, p)
    O.density(s, p, pr)
    v0 = {k: s[k].copy() for k in ("vx", "vy", "vz")}
    O.xsph(s, p, pr)
    for k in ("vx", "vy", "vz"):
        assert np.abs(g[k] - s[k]).max() < 1e-12 * np.abs(s[k] - v0[k]).max(), k


# ------------------------------------------------------------- trajectories
def test_step_parity_dam_break(require_gpu):
    st, p = scenarios.dam_break()
    nsteps = 20
    with make_ctx(st, p) as ctx:
        ctx.step(nsteps)
        g = ctx.download(sphx.F_POS | sphx.F_VEL)
    s = O.State(st)
    O.step(s, p, nsteps)
    for k in ("x", "y", "z"):
        assert np.abs(g[k] - s[k]).max() < 1e-12, k
    for k in ("vx", "vy", "vz"):
        assert np.abs(g[k] - s[k]).max() < 1e-9 * max(np.abs(s[k]).max(), 1e-3), k


def test_c0_dam_break_200_steps(require_gpu):
    # the full 16,384-particle dam break: 200 velocity-Verlet ticks against the
    # oracle (about 15 s of CPU); differences stay at round-off level
    st, p = scenarios.dam_break()
    nsteps = 200
    with make_ctx(st, p) as ctx:
        ctx.step(nsteps)
        g = ctx.download(sphx.F_POS | sphx.F_VEL)
        e_gpu = ctx.energy()
    s = O.State(st)
    O.step(s, p, nsteps)
    O.evaluate(s, p)
    e_orc = O.energy(s, p)
    for k in ("x", "y", "z"):
        assert np.abs(g[k] - s[k]).max() < 1e-10, k
    for k in ("vx", "vy", "vz"):
        assert np.abs(g[k] - s[k]).max() < 1e-7 * max(np.abs(s[k]).max(), 1e-3), k
    for k in ("kinetic", "potential", "internal"):
        assert e_gpu[k] == pytest.approx(e_orc[k], rel=1e-9, abs=1e-12), k


def test_step_parity_mhd(require_gpu):
    # the 262,144-particle workload's physics (adiabatic ideal gas, induction,
    # cleaning, Maxwell stress) on a 16^3 lattice, ids shuffled
    st, p = scenarios.mhd_periodic(16)
    st = shuffle_ids(st, seed=32)
    nsteps = 5
    with make_ctx(st, p) as ctx:
        ctx.step(nsteps)
        g = ctx.download(sphx.F_POS | sphx.F_VEL | sphx.F_B | sphx.F_PSI | sphx.F_U)
    s = O.State(st)
    O.step(s, p, nsteps)
    for k in ("x", "y", "z"):
        d = np.abs(g[k] - s[k])
        d = np.minimum(d, 1.0 - d)  # periodic wrap
        assert d.max() < 1e-12, k
    for k in ("vx", "vy", "vz", "bx", "by", "bz"):
        assert np.abs(g[k] - s[k]).max() < 1e-10 * np.abs(s[k]).max(), k
    assert np.abs(g["psi"] - s["psi"]).max() < 1e-8 * max(np.abs(s["psi"]).max(), 1e-12)
    du = np.abs(s["u"] - st["u"]).max()
    assert du > 0
    assert np.abs(g["u"] - s["u"]).max() < 1e-9 * du


# ------------------------------------------------------------- adiabatic ideal gas
@pytest.mark.parametrize("monaghan", [True, False])
@pytest.mark.parametrize("variable", [False, True])
def test_adiabatic_step_parity(require_gpu, monaghan, variable):
    # P = (gamma - 1) rho u and du/dt from the full force passes; uniform and
    # per-particle h and m, both viscosity conventions, shuffled ids
    st, p = scenarios.ideal_gas_box(16, monaghan=monaghan, variable=variable)
    st = shuffle_ids(st, seed=33)
    nsteps = 10
    with make_ctx(st, p) as ctx:
        ctx.forces()
        first = ctx.download(sphx.F_RHO | sphx.F_P | sphx.F_ACC)
        with pytest.raises(sphx.SphxError) as e:
            ctx.forces(visc_only=True)  # the energy equation needs the full pass
        assert e.value.code == 8
        ctx.step(nsteps)
        g = ctx.download(sphx.F_POS | sphx.F_VEL | sphx.F_U | sphx.F_RHO | sphx.F_P)
        eg = ctx.energy()
    s0 = O.State(st)
    O.evaluate(s0, p)
    assert rel_err(first["rho"], s0["rho"]) < 1e-13
    assert rel_err(first["p"], s0["p"]) < 1e-13
    for k in ("ax", "ay", "az"):
        assert np.abs(first[k] - s0[k]).max() < 1e-11 * np.abs(s0[k]).max(), k
    s = O.State(st)
    O.step(s, p, nsteps)
    O.lib().orc_eos(s.ref(), O._pp(p))  # P at the new u (the download recomputes it)
    for k in ("x", "y", "z"):
        d = np.abs(g[k] - s[k])
        assert np.minimum(d, 1.0 - d).max() < 1e-12, k
    for k in ("vx", "vy", "vz"):
        assert np.abs(g[k] - s[k]).max() < 1e-10 * np.abs(s[k]).max(), k
    du = np.abs(s["u"] - st["u"]).max()
    assert du > 1e-4
    assert np.abs(g["u"] - s["u"]).max() < 1e-9 * du
    assert rel_err(g["rho"], s["rho"]) < 1e-12
    assert rel_err(g["p"], s["p"]) < 1e-11
    assert eg["internal"] == pytest.approx(float(np.sum(st["m"] * s["u"])), rel=1e-12)


def test_mixed_adiabatic_error(require_gpu):
    # FP32-mixed form of the energy equation: f32 pair terms, f64 sums and u
    st, p = scenarios.ideal_gas_box(16)
    out = {}
    for prec in (sphx.PREC_FP64, sphx.PREC_FP32_MIXED):
        p.precision = prec
        with make_ctx(st, p) as ctx:
            ctx.step(10)
            out[prec] = ctx.download(sphx.F_U | sphx.F_VEL)
    a, b = out[sphx.PREC_FP32_MIXED], out[sphx.PREC_FP64]
    du = np.abs(b["u"] - st["u"]).max()
    err = np.abs(a["u"] - b["u"]).max() / du
    print(f"mixed adiabatic: max |u32 - u64| / max|du64| = {err:.3e}")
    assert err < 1e-4


def test_free_fall_matches_oracle_bitwise(require_gpu):
    # isolated particles: no pair terms, so the integrator path is bit-exact
    n = 64
    rng = scenarios.SplitMix64(4)
    st = dict(x=rng.uniform(n) * 10, y=rng.uniform(n) * 10, z=rng.uniform(n) * 10, vx=rng.normal(n),
              vy=rng.normal(n), vz=rng.normal(n), m=np.ones(n), h=np.full(n, 1e-3))
    p = sphx.make_params(1e-3, gravity=(0.1, -0.2, -9.81), boundary_flags=sphx.BND_REFLECT,
                         box_lo=(0, 0, 0), box_hi=(10, 10, 10))
    with make_ctx(st, p) as ctx:
        ctx.step(500)
        g = ctx.download(sphx.F_POS | sphx.F_VEL)
    s = O.State(st)
    O.step(s, p, 500)
    for k in g:
        assert np.array_equal(g[k], s[k]), k


# ------------------------------------------------------------- edge cases
def test_empty_and_single_particle(require_gpu):
    p = sphx.make_params(1e-4, gravity=(0, 0, -9.81))
    empty = {k: np.zeros(0) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    with make_ctx(empty, p, cap=8) as ctx:
        ctx.step(3)
        assert len(ctx.export_pairs()) == 0
        assert ctx.download(sphx.F_POS)["x"].size == 0
    one = {k: np.array([0.5]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "h")}
    with make_ctx(one, p) as ctx:
        ctx.step(10)
        g = ctx.download(sphx.F_POS)
    s = O.State(one)
    O.step(s, p, 10)
    assert np.array_equal(g["z"], s["z"])


def test_error_paths(require_gpu):
    st, p = scenarios.dam_break(nx=8, ny=8, nz=8)
    with sphx.Context(100, p) as ctx:
        with pytest.raises(sphx.SphxError) as e:
            ctx.upload(st)  # 512 > capacity 100
        assert e.value.code == 3
        bad = {k: v[:50].copy() for k, v in st.items()}
        bad["x"][7] = np.nan
        with pytest.raises(sphx.SphxError) as e:
            ctx.upload(bad)
        assert e.value.code == 6
        with pytest.raises(sphx.SphxError) as e:
            ctx.forces(visc_only=True)
        assert e.value.code == 8
    p.max_neighbours = 8  # lattice interior has 80
    with sphx.Context(len(st["x"]), p) as ctx:
        ctx.upload(st)
        with pytest.raises(sphx.SphxError) as e:
            ctx.step(1)
        assert e.value.code == 3


def test_set_params_failure_leaves_context_unchanged(require_gpu):
    # max_neighbours is rounded up to a multiple of 4 at create (190 -> 192):
    # passing the create-time value back is accepted; a different capacity is
    # SPHX_E_INVALID and the context keeps its old parameters and state
    st, p = scenarios.dam_break(nx=12, ny=12, nz=12)
    st = with_velocity(st, 0.3)
    p.max_neighbours = 190
    with make_ctx(st, p) as ctx:
        ctx.set_params(p)
        ctx.forces()
        a0 = ctx.download(sphx.F_ACC)
        bad = sphx.make_params(p.dt, visc=sphx.SphViscosityConfig(2.0, 0.0), max_neighbours=300)
        with pytest.raises(sphx.SphxError) as e:
            ctx.set_params(bad)
        assert e.value.code == 4
        a1 = ctx.download(sphx.F_ACC)  # derived data still valid: nothing was reset
        ctx.forces()
        a2 = ctx.download(sphx.F_ACC)  # recomputed with the old parameters
        p.max_neighbours = 0           # keeps the capacity
        ctx.set_params(p)
        ctx.forces()
        a3 = ctx.download(sphx.F_ACC)
    for k in ("ax", "ay", "az"):
        assert np.array_equal(a0[k], a1[k]) and np.array_equal(a0[k], a2[k]) and np.array_equal(a0[k], a3[k]), k


def test_determinism(require_gpu):
    st, p = scenarios.dam_break()
    out = []
    for _ in range(2):
        with make_ctx(st, p) as ctx:
            ctx.step(10)
            out.append(ctx.download(sphx.F_POS | sphx.F_VEL))
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k])


# ------------------------------------------------------------- full size
def test_full_size_lattice_properties(require_gpu):
    # 1,048,576 particles without jitter: every interior particle has 80
    # neighbours and rho = 0.99726183 m / dx^3; pair lists are symmetric.
    dx = 0.01
    st, p = scenarios.dam_break(nx=128, ny=128, nz=64, spacing=dx, box=(3.2, 1.6, 1.6), jitter=0.0)
    with make_ctx(st, p) as ctx:
        ctx.density()
        g = ctx.download(sphx.F_RHO | sphx.F_POS)
        e = ctx.energy()
        pairs = ctx.export_pairs()
    ix = np.rint(g["x"] / dx - 0.5).astype(int)
    iy = np.rint(g["y"] / dx - 0.5).astype(int)
    iz = np.rint(g["z"] / dx - 0.5).astype(int)
    interior = (ix >= 2) & (ix < 126) & (iy >= 2) & (iy < 126) & (iz >= 2) & (iz < 62)
    m = st["m"][0]
    assert np.abs(g["rho"][interior] / (0.99726183 * m / dx ** 3) - 1).max() < 1e-7
    deg = np.bincount(pairs.ravel(), minlength=len(st["x"]))
    assert (deg[interior] == 80).all()
    assert e["total_neighbours"] == 2 * len(pairs)
    assert e["max_neighbours"] == 80


# ------------------------------------------------------------- FP32-mixed form
def mixed(params):
    params.precision = sphx.PREC_FP32_MIXED
    return params


def test_mixed_density_and_force_error(require_gpu):
    # f32 pair terms on cell-relative coordinates, f64 sums: relative density
    # error <= 1e-6 against the f64 oracle; accelerations within 1e-4 of max|a|.
    st, p = scenarios.dam_break(nx=24, ny=24, nz=24)
    st = with_velocity(st, 0.3)
    st["m"] = st["m"] * 1.05
    p.visc_alpha, p.visc_beta = 0.5, 0.3
    with make_ctx(st, mixed(p)) as ctx:
        ctx.forces()
        g = ctx.download(sphx.F_ACC | sphx.F_RHO)
    p.precision = sphx.PREC_FP64
    s = O.State(st)
    O.evaluate(s, p)
    err_rho = np.abs(g["rho"] / s["rho"] - 1).max()
    assert err_rho < 1e-6, err_rho
    for k in ("ax", "ay", "az"):
        assert np.abs(g[k] - s[k]).max() < 1e-4 * np.abs(s[k]).max(), k


def test_mixed_visc_only_reuse_bit_identical(require_gpu):
    st, p = scenarios.dam_break()
    st = with_velocity(st, 0.2)
    with make_ctx(st, mixed(p)) as ctx:
        ctx.forces()
        a1 = ctx.download(sphx.F_ACC)
        ctx.forces(visc_only=True)
        a2 = ctx.download(sphx.F_ACC)
    for k in a1:
        assert np.array_equal(a1[k], a2[k])


def test_mixed_mhd_periodic_step(require_gpu):
    st, p = scenarios.mhd_periodic(16)
    with make_ctx(st, mixed(p)) as ctx:
        ctx.step(5)
        g = ctx.download(sphx.F_POS | sphx.F_VEL | sphx.F_B)
    p.precision = sphx.PREC_FP64
    s = O.State(st)
    O.step(s, p, 5)
    for k in ("vx", "vy", "vz", "bx", "by", "bz"):
        assert np.abs(g[k] - s[k]).max() < 1e-5 * np.abs(s[k]).max(), k


def test_update_keeps_residency_and_matches_upload(require_gpu):
    # sphx_update writes new velocities, B / psi and internal energies into
    # the resident sorted state.  Writing back the values the context already
    # holds changes no bit of the trajectory; new values agree with a fresh
    # upload of the same state (a different sort and list order:
    # rounding-level differences).
    st, p = scenarios.mhd_periodic(16)
    fields = sphx.F_POS | sphx.F_VEL | sphx.F_B | sphx.F_PSI | sphx.F_U
    with sphx.Context(len(st["x"]), p) as a, sphx.Context(len(st["x"]), p) as b:
        a.upload(st)
        b.upload(st)
        a.step(3)
        b.step(3)
        now = a.download(fields)
        a.update(now, sphx.F_VEL | sphx.F_B | sphx.F_PSI | sphx.F_U)
        a.step(3)
        b.step(3)
        ra, rb = a.download(fields), b.download(fields)
        for k in ra:
            assert np.array_equal(ra[k], rb[k]), k
        # new velocities and field: against a fresh upload of the same state
        new = dict(ra)
        rng = scenarios.SplitMix64(5)
        n = len(st["x"])
        for k in ("vx", "vy", "vz"):
            new[k] = ra[k] + 0.01 * rng.normal(n)
        new["bz"] = ra["bz"] + 1e-3
        new["u"] = ra["u"] * (1.0 + 0.01 * rng.uniform(n))
        a.update(new, sphx.F_VEL | sphx.F_B | sphx.F_U)
        fresh = dict(st)
        fresh.update({k: new[k] for k in ("x", "y", "z", "vx", "vy", "vz", "bx", "by", "bz", "psi", "u")})
        with sphx.Context(n, p) as c:
            c.upload(fresh)
            a.step(2)
            c.step(2)
            ga, gc = a.download(fields), c.download(fields)
    for k in ga:
        scale = max(np.abs(gc[k]).max(), 1e-30)
        assert np.abs(ga[k] - gc[k]).max() <= 1e-9 * scale, k


def brute_candidates(st, radius, periodic, L=1.0):
    # broadphase_query_system (broadphase.rs:62-96): (a, b) with a < b and
    # ((dx*dx + dy*dy) + dz*dz) < (4 R_a)^2, R_a the querying (smaller) id's radius
    ids = st["id"].astype(np.int64)
    x = np.stack([st["x"], st["y"], st["z"]], axis=1)
    out = []
    for i in range(len(ids)):
        d = x - x[i]
        if periodic:  # the minimum image as the kernels form it
            d = np.where(d > 0.5 * L, d - L, np.where(d < -0.5 * L, d + L, d))
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        r4 = 4.0 * radius[i]
        sel = (ids[i] < ids) & (d2 < r4 * r4)
        out.extend((int(ids[i]), int(b)) for b in ids[sel])
    return np.array(sorted(out), dtype=np.uint32).reshape(-1, 2)


@pytest.mark.parametrize("periodic", [False, True])
def test_export_candidates_match_brute_force(require_gpu, periodic):
    # CandidatePairs for non-SPH consumers: per-particle bounding radii, the
    # query radius 4 R of the smaller id; before and after ticks that moved the
    # particles without a list rebuild
    rng = scenarios.SplitMix64(31)
    n = 3000
    st = {k: rng.uniform(n) for k in ("x", "y", "z")}
    st.update(vx=0.3 * rng.normal(n), vy=0.3 * rng.normal(n), vz=0.3 * rng.normal(n), m=np.full(n, 1.0 / n),
              h=0.03 + 0.02 * rng.uniform(n), id=np.argsort(rng.uniform(n)).astype(np.uint32))
    radius = 0.45 * st["h"].max() * (0.5 + 0.5 * rng.uniform(n))
    kw = dict(boundary_flags=sphx.BND_PERIODIC, box_lo=(0, 0, 0), box_hi=(1, 1, 1)) if periodic else {}
    p = sphx.make_params(2e-5, tait=sphx.TaitEquationConfig(1.0, 5.0, 7.0),
                         visc=sphx.SphViscosityConfig(0.5, 0.3), **kw)
    with make_ctx(st, p) as ctx:
        got0 = ctx.export_candidates(radius)
        ctx.step(5)
        cur = ctx.download(sphx.F_POS)
        got1 = ctx.export_candidates(radius)
        # radii beyond the cell width: three cells per axis and more
        got3 = ctx.export_candidates(radius * 3.0)
        with pytest.raises(sphx.SphxError) as e:
            ctx.export_candidates(-radius)
        assert e.value.code == 4
    ref0 = brute_candidates(st, radius, periodic)
    moved = dict(st)
    moved.update(cur)
    ref1 = brute_candidates(moved, radius, periodic)
    ref3 = brute_candidates(moved, radius * 3.0, periodic)
    assert len(ref0) > 3000 and len(ref3) > 20 * len(ref1)
    assert np.array_equal(got0, ref0)
    assert np.array_equal(got1, ref1)
    assert np.array_equal(got3, ref3)


@pytest.mark.parametrize("periodic", [False, True])
def test_export_candidates_radius_beyond_half_the_box(require_gpu, periodic):
    # 4 R above half the box: in a periodic box every cell is searched once
    # (minimum image), in an open one the search is clamped to the grid
    rng = scenarios.SplitMix64(32)
    n = 600
    st = {k: rng.uniform(n) for k in ("x", "y", "z")}
    st.update(vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n), m=np.full(n, 1.0 / n), h=np.full(n, 0.04),
              id=np.argsort(rng.uniform(n)).astype(np.uint32))
    radius = 0.15 * (0.5 + rng.uniform(n))
    kw = dict(boundary_flags=sphx.BND_PERIODIC, box_lo=(0, 0, 0), box_hi=(1, 1, 1)) if periodic else {}
    p = sphx.make_params(1e-5, tait=sphx.TaitEquationConfig(1.0, 5.0, 7.0), **kw)
    with make_ctx(st, p) as ctx:
        got = ctx.export_candidates(radius)
    ref = brute_candidates(st, radius, periodic)
    assert len(ref) > 0.4 * n * (n - 1) / 2
    assert np.array_equal(got, ref)
