"""Pins of the oracle's reverse-mode gradient (oracle/sph_adjoint.c).

orc_adjoint returns L = mean |x_N - target|^2 and its gradient in the initial
velocities and in the viscosity alpha.  It is pinned by
  * the loss it reports == orc_loss (the plain forward, orc_step) bit for bit,
  * central finite differences of orc_loss along seeded random directions
    (velocities and alpha), for Tait / polytropic / isothermal EOS and both
    viscosity conventions, on a compressed blob in a tight reflecting box,
  * closed forms for one particle (free fall: dL/dv0 = 2 e T / 3; a wall
    reflection flips the sign of the normal component).
The neighbour sets and the viscosity switch are fixed per tick; the
finite-difference steps (3e-6, 1e-6 and 3e-7 m/s on velocities of ~0.4 m/s; 1e-4 and
1e-5 on alpha) sit between the round-off floor and the kinks of the viscosity switch
(Pi is continuous but not differentiable where a pair's approach speed
crosses 0).  A step that straddles such a kink shows as a one-sided error, so
the best step of each set must agree to 3e-6; any dropped term of
the VJP (density path, dG/dr, the viscosity's velocity dependence) is off by
1e-3 or more.
"""
import numpy as np
import pytest

import oracle as O
import sphx
from sphx import scenarios


def blob_case(eos="tait", conv=sphx.VISC_REFERENCE, seed=5, dt=2e-4):
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
    kw = dict(tait=sphx.TaitEquationConfig(1000.0, 20.0, 7.0), visc=sphx.SphViscosityConfig(0.3, 0.2),
              boundary_flags=sphx.BND_REFLECT, box_lo=lo, box_hi=hi, gravity=(0.0, 0.0, -9.81),
              visc_convention=conv)
    prm = sphx.make_params(dt, **kw)
    if eos == "polytropic":
        prm.eos = sphx.EOS_POLYTROPIC
        prm.poly_k = 40.0
        prm.gamma = 5.0 / 3.0
    elif eos == "isothermal":
        prm.eos = sphx.EOS_ISOTHERMAL
        prm.iso_cs2 = 30.0
    target = np.stack([x, y, z], axis=1) + 0.01 * rng.normal(3 * n).reshape(n, 3)
    return st, prm, target


def loss_at(st, prm, nsteps, target, dv=None, dalpha=0.0):
    s = dict(st)
    if dv is not None:
        s["vx"] = st["vx"] + dv[:, 0]
        s["vy"] = st["vy"] + dv[:, 1]
        s["vz"] = st["vz"] + dv[:, 2]
    alpha = prm.visc_alpha
    prm.visc_alpha = alpha + dalpha
    try:
        return O.loss(O.State(s), prm, nsteps, target)
    finally:
        prm.visc_alpha = alpha


def reflections(st, prm, nsteps):
    """Number of velocity components whose sign differs from a run without walls."""
    a = O.State(st)
    O.step(a, prm, nsteps)
    free = sphx.make_params(prm.dt, tait=sphx.TaitEquationConfig(prm.rest_density, prm.speed_of_sound,
                                                                  prm.gamma),
                            visc=sphx.SphViscosityConfig(prm.visc_alpha, prm.visc_beta),
                            visc_convention=prm.visc_convention, gravity=tuple(prm.gravity))
    free.eos, free.poly_k, free.iso_cs2 = prm.eos, prm.poly_k, prm.iso_cs2
    b = O.State(st)
    O.step(b, free, nsteps)
    va = np.stack([a["vx"], a["vy"], a["vz"]])
    vb = np.stack([b["vx"], b["vy"], b["vz"]])
    return int(np.sum(np.sign(va) != np.sign(vb)))


def test_adjoint_loss_is_forward_loss():
    st, prm, target = blob_case()
    loss, _, _ = O.adjoint(O.State(st), prm, 6, target)
    assert loss == loss_at(st, prm, 6, target)


@pytest.mark.parametrize("eos", ["tait", "polytropic", "isothermal"])
@pytest.mark.parametrize("conv", [sphx.VISC_REFERENCE, sphx.VISC_MONAGHAN])
def test_adjoint_matches_finite_differences(eos, conv):
    nsteps = 6
    st, prm, target = blob_case(eos=eos, conv=conv)
    assert reflections(st, prm, nsteps) > 0  # the wall map is exercised
    _, gv, ga = O.adjoint(O.State(st), prm, nsteps, target)
    n = len(st["x"])
    rng = np.random.default_rng(11)
    for trial in range(2):
        d = rng.standard_normal((n, 3))
        an = float(np.sum(gv * d))
        errs = []
        for eps in (3e-6, 1e-6, 3e-7):
            fd = (loss_at(st, prm, nsteps, target, dv=eps * d)
                  - loss_at(st, prm, nsteps, target, dv=-eps * d)) / (2 * eps)
            errs.append(abs(fd - an) / abs(an))
        assert min(errs) <= 3e-6, (trial, errs)
    assert ga != 0.0
    errs = []
    for ea in (1e-4, 1e-5):
        fda = (loss_at(st, prm, nsteps, target, dalpha=ea)
               - loss_at(st, prm, nsteps, target, dalpha=-ea)) / (2 * ea)
        errs.append(abs(fda - ga) / abs(ga))
    assert min(errs) <= 3e-6, errs


@pytest.mark.parametrize("eos", ["tait", "polytropic"])
def test_adjoint_position_gradient_matches_finite_differences(eos):
    nsteps = 6
    st, prm, target = blob_case(eos=eos)
    _, gx, _, _ = O.adjoint(O.State(st), prm, nsteps, target, with_gx=True)
    n = len(st["x"])
    d = np.random.default_rng(12).standard_normal((n, 3))
    an = float(np.sum(gx * d))

    def loss_dx(eps):
        s = dict(st)
        s["x"] = st["x"] + eps * d[:, 0]
        s["y"] = st["y"] + eps * d[:, 1]
        s["z"] = st["z"] + eps * d[:, 2]
        return O.loss(O.State(s), prm, nsteps, target)

    errs = []
    for eps in (1e-7, 3e-8, 1e-8):
        errs.append(abs((loss_dx(eps) - loss_dx(-eps)) / (2 * eps) - an) / abs(an))
    assert min(errs) <= 3e-6, errs


def test_adjoint_single_particle_free_fall():
    dt, nsteps = 1e-3, 7
    st = {"x": np.array([0.1]), "y": np.array([0.2]), "z": np.array([0.3]),
          "vx": np.array([0.5]), "vy": np.array([-0.2]), "vz": np.array([0.1]),
          "m": np.array([1.0]), "h": np.array([0.05])}
    prm = sphx.make_params(dt, tait=sphx.TaitEquationConfig(1000.0, 20.0, 7.0),
                           visc=sphx.SphViscosityConfig(0.1, 0.0), gravity=(0.0, 0.0, -9.81))
    target = np.array([[0.0, 0.0, 0.0]])
    s = O.State(st)
    loss, gv, ga = O.adjoint(s, prm, nsteps, target)
    T = nsteps * dt
    g = np.array(prm.gravity[:])
    xN = np.array([0.1, 0.2, 0.3]) + np.array([0.5, -0.2, 0.1]) * T + 0.5 * g * T * T
    e = xN - target[0]
    np.testing.assert_allclose([s["x"][0], s["y"][0], s["z"][0]], xN, rtol=0, atol=1e-15)
    np.testing.assert_allclose(gv[0], 2 * e * T / 3, rtol=1e-12)
    assert ga == 0.0
    assert loss == pytest.approx(np.sum(e * e) / 3, rel=1e-14)


def test_adjoint_single_particle_wall_reflection():
    # x0 = 8e-4 from the wall at 0, moving at -1 m/s, no gravity: the first
    # tick overshoots to -2e-4 and is mirrored, so x_N = -(x0 + v0 T) and
    # dL/dvx0 = 2 e (-T) / 3, while y and z keep +T.
    dt, nsteps = 1e-3, 4
    st = {"x": np.array([1e-3 - 2e-4]), "y": np.array([0.5]), "z": np.array([0.5]),
          "vx": np.array([-1.0]), "vy": np.array([0.3]), "vz": np.array([0.0]),
          "m": np.array([1.0]), "h": np.array([0.05])}
    prm = sphx.make_params(dt, tait=sphx.TaitEquationConfig(1000.0, 20.0, 7.0),
                           visc=sphx.SphViscosityConfig(0.1, 0.0),
                           boundary_flags=sphx.BND_REFLECT, box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0))
    target = np.array([[0.01, 0.4, 0.45]])
    s = O.State(st)
    _, gv, _ = O.adjoint(s, prm, nsteps, target)
    T = nsteps * dt
    xN = np.array([-(st["x"][0] + st["vx"][0] * T), 0.5 + 0.3 * T, 0.5])
    np.testing.assert_allclose([s["x"][0], s["y"][0], s["z"][0]], xN, rtol=0, atol=1e-15)
    e = xN - target[0]
    np.testing.assert_allclose(gv[0], 2 * e * np.array([-T, T, T]) / 3, rtol=1e-12)


def test_adjoint_rejects_unsupported_options():
    st, prm, target = blob_case()
    prm.xsph_epsilon = 0.5
    with pytest.raises(ValueError):
        O.adjoint(O.State(st), prm, 2, target)
