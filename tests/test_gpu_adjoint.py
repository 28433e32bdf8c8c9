"""GPU parity of sphx_adjoint (the differentiable SPH step) against the oracle.

The HIP adjoint is checked
  * against the oracle's reverse pass (oracle/sph_adjoint.c, itself pinned by
    finite differences and closed forms) on a compressed blob in a tight
    reflecting box, for both viscosity conventions: loss, dL/dv0, dL/dx0 and
    dL/dalpha within 1e-9 relative (the forward trajectories differ in
    summation order only, ~1e-13);
  * for checkpoint invariance: recomputing the tape from checkpoints every
    1, 2, 3 ticks gives the same gradient as storing the whole tape;
  * against central finite differences of its own forward (eps = 1e-6) on a
    dam break with seeded initial velocities, checkpoint every 10 ticks;
  * for the state it leaves behind: bitwise the state sphx_step reaches.
"""
import numpy as np
import pytest

import oracle as O
import sphx
from sphx import scenarios

pytestmark = pytest.mark.gpu


def blob_case(conv=sphx.VISC_REFERENCE, seed=5, dt=2e-4):
    rng = scenarios.SplitMix64(seed)
    sp = 0.05
    x, y, z = scenarios.lattice(5, 5, 4, sp)
    n = x.size
    x = x + (2 * rng.uniform(n) - 1) * 0.1 * sp
    y = y + (2 * rng.uniform(n) - 1) * 0.1 * sp
    z = z + (2 * rng.uniform(n) - 1) * 0.1 * sp
    st = {"x": x, "y": y, "z": z, "vx": 0.4 * rng.normal(n), "vy": 0.4 * rng.normal(n),
          "vz": 0.4 * rng.normal(n), "m": np.full(n, 1000.0 * sp ** 3 * 1.15), "h": np.full(n, 0.07)}
    lo = (x.min() - 2e-4, y.min() - 2e-4, z.min() - 2e-4)
    hi = (x.max() + 2e-4, y.max() + 2e-4, z.max() + 2e-4)
    prm = sphx.make_params(dt, tait=sphx.TaitEquationConfig(1000.0, 20.0, 7.0),
                           visc=sphx.SphViscosityConfig(0.3, 0.2), boundary_flags=sphx.BND_REFLECT,
                           box_lo=lo, box_hi=hi, gravity=(0.0, 0.0, -9.81), visc_convention=conv)
    target = np.stack([x, y, z], axis=1) + 0.01 * rng.normal(3 * n).reshape(n, 3)
    return st, prm, target


def gpu_adjoint(st, prm, nsteps, target, every=0):
    with sphx.Context(len(st["x"]), prm) as ctx:
        ctx.upload(st)
        out = ctx.adjoint(nsteps, target, checkpoint_every=every, with_grad_x=True)
        fin = ctx.download(sphx.F_POS | sphx.F_VEL)
    return out, fin


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("conv", [sphx.VISC_REFERENCE, sphx.VISC_MONAGHAN])
def test_adjoint_matches_oracle(require_gpu, conv):
    nsteps = 6
    st, prm, target = blob_case(conv=conv)
    out, fin = gpu_adjoint(st, prm, nsteps, target)
    s = O.State(st)
    loss, gx, gv, ga = O.adjoint(s, prm, nsteps, target, with_gx=True)
    assert out["loss"] == pytest.approx(loss, rel=1e-11)
    assert rel(out["grad_v"], gv) < 1e-9
    assert rel(out["grad_x"], gx) < 1e-9
    assert out["grad_alpha"] == pytest.approx(ga, rel=1e-9)
    # the context is left at x_N
    assert rel(fin["x"], s["x"]) < 1e-12


def test_adjoint_checkpoint_invariance(require_gpu):
    st, prm, target = blob_case()
    ref, _ = gpu_adjoint(st, prm, 7, target, every=0)
    for every in (1, 2, 3):
        out, _ = gpu_adjoint(st, prm, 7, target, every=every)
        assert out["loss"] == ref["loss"]
        assert rel(out["grad_v"], ref["grad_v"]) < 1e-13
        assert rel(out["grad_x"], ref["grad_x"]) < 1e-13
        assert out["grad_alpha"] == pytest.approx(ref["grad_alpha"], rel=1e-13)


def test_adjoint_leaves_the_step_state(require_gpu):
    st, prm = scenarios.dam_break(nx=12, ny=8, nz=8, v_sigma=0.05, seed=4)
    n = len(st["x"])
    target = np.stack([st["x"], st["y"], st["z"]], axis=1)
    with sphx.Context(n, prm) as ctx:
        ctx.upload(st)
        ctx.adjoint(9, target, checkpoint_every=4)
        a = ctx.download(sphx.F_POS | sphx.F_VEL)
    with sphx.Context(n, prm) as ctx:
        ctx.upload(st)
        ctx.step(9)
        b = ctx.download(sphx.F_POS | sphx.F_VEL)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(a[k], b[k]), k


def test_adjoint_dam_break_finite_differences(require_gpu):
    # dam break with seeded initial velocities (at rest every pair sits on
    # the viscosity switch, where the step has no derivative); the target is
    # the same start run with alpha = 0.15
    nsteps, every = 30, 10
    st, prm = scenarios.dam_break(nx=16, ny=12, nz=10, v_sigma=0.05, seed=2)
    n = len(st["x"])
    prm15 = sphx.make_params(prm.dt, tait=sphx.TaitEquationConfig(1000.0, 20.0, 7.0),
                             visc=sphx.SphViscosityConfig(0.15, 0.0), boundary_flags=sphx.BND_REFLECT,
                             box_lo=tuple(prm.box_lo), box_hi=tuple(prm.box_hi), gravity=(0.0, 0.0, -9.81))
    with sphx.Context(n, prm15) as ctx:
        ctx.upload(st)
        ctx.step(nsteps)
        f = ctx.download(sphx.F_POS)
    target = np.stack([f["x"], f["y"], f["z"]], axis=1)
    with sphx.Context(n, prm) as ctx:
        ctx.upload(st)
        out = ctx.adjoint(nsteps, target, checkpoint_every=every)

        def loss(v3=None, dalpha=0.0):
            s = dict(st)
            if v3 is not None:
                s["vx"], s["vy"], s["vz"] = st["vx"] + v3[:, 0], st["vy"] + v3[:, 1], st["vz"] + v3[:, 2]
            p = prm
            if dalpha:
                p = sphx.make_params(prm.dt, tait=sphx.TaitEquationConfig(1000.0, 20.0, 7.0),
                                     visc=sphx.SphViscosityConfig(prm.visc_alpha + dalpha, 0.0),
                                     boundary_flags=sphx.BND_REFLECT, box_lo=tuple(prm.box_lo),
                                     box_hi=tuple(prm.box_hi), gravity=(0.0, 0.0, -9.81))
            ctx.set_params(p)
            ctx.upload(s)
            ctx.step(nsteps)
            r = ctx.download(sphx.F_POS)
            x = np.stack([r["x"], r["y"], r["z"]], axis=1)
            return float(np.sum((x - target) ** 2) / (3 * n))

        eps = 1e-6
        rng = np.random.default_rng(3)
        for _ in range(4):
            d = rng.standard_normal((n, 3))
            d /= np.linalg.norm(d)
            fd = (loss(eps * d) - loss(-eps * d)) / (2 * eps)
            an = float(np.sum(out["grad_v"] * d))
            assert abs(fd - an) <= 1e-4 * abs(an), (fd, an)
        fda = (loss(dalpha=eps) - loss(dalpha=-eps)) / (2 * eps)
        assert abs(fda - out["grad_alpha"]) <= 1e-4 * abs(out["grad_alpha"]), (fda, out["grad_alpha"])


def test_adjoint_rejects_unsupported(require_gpu):
    st, prm, target = blob_case()
    prm.precision = sphx.PREC_FP32_MIXED
    with sphx.Context(len(st["x"]), prm) as ctx:
        ctx.upload(st)
        with pytest.raises(sphx.SphxError) as e:
            ctx.adjoint(2, target)
        assert e.value.code == 4
    st2, p2 = scenarios.mhd_periodic(n_side=8)
    with sphx.Context(len(st2["x"]), p2) as ctx:
        ctx.upload(st2)
        with pytest.raises(sphx.SphxError) as e:
            ctx.adjoint(2, np.zeros((len(st2["x"]), 3)))
        assert e.value.code == 4


def test_adjoint_65k_dam_break_32_directions(require_gpu):
    # the differentiable workload at full size: 64 x 32 x 32 dam break, seeded
    # initial velocities, 100 steps, checkpoint every 10, target from the same
    # start with alpha = 0.15; central differences (eps = 1e-6) on 32 seeded
    # unit directions of the initial velocities, and on alpha
    nsteps, every, eps = 100, 10, 1e-6
    st, prm = scenarios.dam_break(nx=64, ny=32, nz=32, v_sigma=0.05, seed=0)
    p15 = scenarios.dam_break(nx=64, ny=32, nz=32, v_sigma=0.05, seed=0, alpha=0.15)[1]
    n = len(st["x"])
    with sphx.Context(n, prm) as ctx:
        ctx.set_params(p15)
        ctx.upload(st)
        ctx.step(nsteps)
        f = ctx.download(sphx.F_POS)
        target = np.stack([f["x"], f["y"], f["z"]], axis=1)
        ctx.set_params(prm)
        ctx.upload(st)
        out = ctx.adjoint(nsteps, target, checkpoint_every=every)

        def loss(dv=None, p=prm):
            s = dict(st)
            if dv is not None:
                s["vx"], s["vy"], s["vz"] = st["vx"] + dv[:, 0], st["vy"] + dv[:, 1], st["vz"] + dv[:, 2]
            ctx.set_params(p)
            ctx.upload(s)
            ctx.step(nsteps)
            r = ctx.download(sphx.F_POS)
            x = np.stack([r["x"], r["y"], r["z"]], axis=1)
            return float(np.sum((x - target) ** 2) / (3 * n))

        assert loss() == pytest.approx(out["loss"], rel=1e-12)
        rng = np.random.default_rng(2024)
        errs = []
        for _ in range(32):
            d = rng.standard_normal((n, 3))
            d /= np.linalg.norm(d)
            fd = (loss(eps * d) - loss(-eps * d)) / (2 * eps)
            an = float(np.sum(out["grad_v"] * d))
            errs.append(abs(fd - an) / abs(an))
        p_hi = scenarios.dam_break(nx=64, ny=32, nz=32, v_sigma=0.05, seed=0, alpha=0.1 + eps)[1]
        p_lo = scenarios.dam_break(nx=64, ny=32, nz=32, v_sigma=0.05, seed=0, alpha=0.1 - eps)[1]
        fda = (loss(p=p_hi) - loss(p=p_lo)) / (2 * eps)
    print("65k adjoint: max rel err over 32 directions %.3g, median %.3g; alpha %.3g vs %.3g"
          % (max(errs), float(np.median(errs)), out["grad_alpha"], fda))
    assert max(errs) < 1e-3, errs
    assert float(np.median(errs)) < 1e-5, errs
    assert abs(fda - out["grad_alpha"]) <= 1e-3 * abs(out["grad_alpha"])
