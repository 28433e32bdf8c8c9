"""Seeded synthetic particle sets for the five workloads (host-side numpy).

One SplitMix64 stream per scenario (seed -> uniform doubles; Gaussians by
Box-Muller) so that the GPU and the oracle receive identical bytes.
"""
from __future__ import annotations

import math
from typing import Dict, Tuple

import numpy as np

from . import (BND_PERIODIC, BND_REFLECT, EOS_ADIABATIC, EOS_POLYTROPIC, EOS_TAIT, MHD_CLEANING, MHD_INDUCTION,
               MHD_MAXWELL_STRESS, VISC_MONAGHAN, VISC_REFERENCE, MhdConfig, SphViscosityConfig, TaitEquationConfig,
               make_params)

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64: output k is mix(seed + (k+1) * golden)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
            z = self.state + k * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.count += n
        return z

    def uniform(self, n: int) -> np.ndarray:
        """Doubles in [0, 1) from the top 53 bits."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def lattice(nx: int, ny: int, nz: int, spacing: float, origin=(0.0, 0.0, 0.0)) -> Tuple[np.ndarray, ...]:
    """Cell-centred lattice, x fastest: particle (i,j,k) at origin + (idx + 0.5) * spacing."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = origin[0] + (i.ravel() + 0.5) * spacing
    y = origin[1] + (j.ravel() + 0.5) * spacing
    z = origin[2] + (k.ravel() + 0.5) * spacing
    return x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)


def dam_break(nx: int = 32, ny: int = 32, nz: int = 16, spacing: float = 0.01,
              box=(0.8, 0.4, 0.4), jitter: float = 0.01, seed: int = 0, alpha: float = 0.1,
              dt: float = 1e-4, v_sigma: float = 0.0) -> Tuple[Dict[str, np.ndarray], object]:
    """Dam-break block in a reflective box (the 16,384 / 65,536 / 1,048,576 workloads).

    Lattice nx*ny*nz at `spacing` from the box corner, uniform jitter of
    +-jitter*spacing per axis, h = 1.3*spacing, m = rho0*spacing^3, Tait
    rho0 = 1000, gamma = 7, c0 = 20, Monaghan alpha (beta = 0), g = -9.81 z.
    Particles start at rest unless v_sigma > 0 (seeded N(0, v_sigma^2) per
    component; the differentiable form uses it: at rest every pair sits on
    the viscosity switch (v_ij . x_ij = 0), where the step is not
    differentiable).
    """
    rng = SplitMix64(seed)
    x, y, z = lattice(nx, ny, nz, spacing)
    n = x.size
    x = x + (2.0 * rng.uniform(n) - 1.0) * jitter * spacing
    y = y + (2.0 * rng.uniform(n) - 1.0) * jitter * spacing
    z = z + (2.0 * rng.uniform(n) - 1.0) * jitter * spacing
    if v_sigma > 0.0:
        vx, vy, vz = (v_sigma * rng.normal(n) for _ in range(3))
    else:
        vx, vy, vz = np.zeros(n), np.zeros(n), np.zeros(n)
    rho0 = 1000.0
    state = {
        "x": x, "y": y, "z": z,
        "vx": vx, "vy": vy, "vz": vz,
        "m": np.full(n, rho0 * spacing ** 3), "h": np.full(n, 1.3 * spacing),
        "id": np.arange(n, dtype=np.uint32),
    }
    params = make_params(
        dt,
        tait=TaitEquationConfig(rho0, 20.0, 7.0),
        visc=SphViscosityConfig(alpha, 0.0),
        eos=EOS_TAIT, visc_convention=VISC_REFERENCE,
        boundary_flags=BND_REFLECT, box_lo=(0.0, 0.0, 0.0), box_hi=tuple(box),
        gravity=(0.0, 0.0, -9.81),
    )
    return state, params


def mhd_periodic(n_side: int = 64, jitter: float = 0.02, seed: int = 1, h_factor: float = 1.06,
                 dt: float = 1e-4, random_fill: int = 0,
                 eos: str = "adiabatic") -> Tuple[Dict[str, np.ndarray], object]:
    """Ideal-MHD periodic unit cube (the 262,144 and 4,194,304 workloads).

    Lattice n_side^3 with +-jitter*spacing (or, with random_fill = N, N
    seeded uniform positions and h from <N_nb> = 40), ideal gas gamma = 5/3
    with rho = 1, P = 1: the adiabatic EOS P = (gamma - 1) rho u with u = 1.5
    evolved by the energy equation (eos="polytropic": P = K rho^gamma, K = 1,
    at fixed entropy instead), v ~ N(0, 0.1^2), B = (0.1, 0, 0) plus
    0.01 * (0, sin(2 pi x + phi), cos(2 pi x + phi)) with a seeded phase phi,
    Monaghan alpha = 1, c = sqrt(gamma P / rho), ideal induction (eta = 0),
    Dedner cleaning at c_h = fast speed, Maxwell-stress force on.
    """
    rng = SplitMix64(seed)
    gamma = 5.0 / 3.0
    if random_fill:
        n = int(random_fill)
        x, y, z = rng.uniform(n), rng.uniform(n), rng.uniform(n)
        # <N_nb> = n * 4/3 pi (2h)^3 = 40
        h = 0.5 * (40.0 / (n * 4.0 / 3.0 * math.pi)) ** (1.0 / 3.0)
        spacing = n ** (-1.0 / 3.0)
    else:
        spacing = 1.0 / n_side
        x, y, z = lattice(n_side, n_side, n_side, spacing)
        n = x.size
        x = np.mod(x + (2.0 * rng.uniform(n) - 1.0) * jitter * spacing, 1.0)
        y = np.mod(y + (2.0 * rng.uniform(n) - 1.0) * jitter * spacing, 1.0)
        z = np.mod(z + (2.0 * rng.uniform(n) - 1.0) * jitter * spacing, 1.0)
        h = h_factor * spacing
    vx, vy, vz = 0.1 * rng.normal(n), 0.1 * rng.normal(n), 0.1 * rng.normal(n)
    phi = 2.0 * math.pi * float(rng.uniform(1)[0])
    state = {
        "x": x, "y": y, "z": z, "vx": vx, "vy": vy, "vz": vz,
        "m": np.full(n, 1.0 / n), "h": np.full(n, h),
        "bx": np.full(n, 0.1),
        "by": 0.01 * np.sin(2.0 * math.pi * x + phi),
        "bz": 0.01 * np.cos(2.0 * math.pi * x + phi),
        "psi": np.zeros(n),
        "id": np.arange(n, dtype=np.uint32),
    }
    if eos == "adiabatic":
        state["u"] = np.full(n, 1.0 / ((gamma - 1.0) * 1.0))  # P = (gamma - 1) rho u = 1 at rho = 1
    elif eos != "polytropic":
        raise ValueError(f"eos must be 'adiabatic' or 'polytropic', not {eos!r}")
    cs = math.sqrt(gamma * 1.0 / 1.0)
    c_fast = math.sqrt(cs * cs + 0.1 ** 2 + 0.01 ** 2)
    params = make_params(
        dt,
        tait=TaitEquationConfig(1.0, cs, gamma),
        visc=SphViscosityConfig(1.0, 0.0),
        mhd=MhdConfig(0.0, c_fast, 0.5),
        eos=EOS_ADIABATIC if eos == "adiabatic" else EOS_POLYTROPIC, poly_k=1.0, visc_convention=VISC_REFERENCE,
        boundary_flags=BND_PERIODIC, box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0),
        mhd_flags=MHD_INDUCTION | MHD_CLEANING | MHD_MAXWELL_STRESS,
    )
    return state, params


def ideal_gas_box(n_side: int = 16, jitter: float = 0.2, seed: int = 41, v_sigma: float = 0.3,
                  alpha: float = 1.0, dt: float = 2e-4, monaghan: bool = True,
                  variable: bool = False) -> Tuple[Dict[str, np.ndarray], object]:
    """Adiabatic ideal gas (gamma = 5/3, rho ~ 1, u ~ 1.5 + 0.3 U(0,1)) in the
    periodic unit cube: jittered n_side^3 lattice, h = 1.3 spacing, seeded
    Gaussian velocities, Monaghan viscosity alpha (approaching pairs unless
    monaghan=False).  variable: per-particle h and m spread by 10 % / 20 %."""
    rng = SplitMix64(seed)
    sp = 1.0 / n_side
    x, y, z = lattice(n_side, n_side, n_side, sp)
    n = x.size
    x = np.mod(x + (2.0 * rng.uniform(n) - 1.0) * jitter * sp, 1.0)
    y = np.mod(y + (2.0 * rng.uniform(n) - 1.0) * jitter * sp, 1.0)
    z = np.mod(z + (2.0 * rng.uniform(n) - 1.0) * jitter * sp, 1.0)
    vx, vy, vz = (v_sigma * rng.normal(n) for _ in range(3))
    u = 1.5 + 0.3 * rng.uniform(n)
    h = np.full(n, 1.3 * sp)
    m = np.full(n, 1.0 / n)
    if variable:
        h = h * (1.0 + 0.1 * rng.uniform(n))
        m = m * (1.0 + 0.2 * rng.uniform(n))
    gamma = 5.0 / 3.0
    state = {"x": x, "y": y, "z": z, "vx": vx, "vy": vy, "vz": vz, "m": m, "h": h, "u": u,
             "id": np.arange(n, dtype=np.uint32)}
    params = make_params(
        dt,
        tait=TaitEquationConfig(1.0, math.sqrt(gamma * (gamma - 1.0) * 1.5), gamma),
        visc=SphViscosityConfig(alpha, 0.0),
        eos=EOS_ADIABATIC, visc_convention=VISC_MONAGHAN if monaghan else VISC_REFERENCE,
        boundary_flags=BND_PERIODIC, box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0),
    )
    return state, params
