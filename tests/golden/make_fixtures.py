"""Writes the golden fixtures of the full-size workloads from the CPU oracle.

Run in a container that has the oracle built (make -C oracle):
    python tests/golden/make_fixtures.py [c3] [c4]

c3_mhd_4M.npz  -- the 4,194,304-particle random-fill MHD box (adiabatic ideal
   gas, induction, cleaning, Maxwell stress): one f64 force evaluation and the
   two MHD passes at the initial state, from the oracle's pair form
   (sph_oracle.c).  Stored: a hash of the generated input, the pair count and
   three checksums of the sorted (id_a, id_b) list, the momentum sum of the
   forces, and rho, P, a, the neighbour count, the sum of neighbour ids, B after
   induction and B, psi after cleaning at 2,048 seeded sample particles.
c4_adjoint_65k.npz -- the 65,536-particle differentiable dam break: loss,
   dL/dalpha, dL/dv0 at 1,024 seeded sample particles and dL/dv0 along 32
   seeded unit directions, from the oracle's reverse pass (sph_adjoint.c), with
   the target from the oracle's forward at alpha = 0.15.

The fixtures are data (inputs are regenerated from the seeded scenarios and
checked against the stored hash); the test reading them is
tests/test_gpu_configs.py.
"""
from __future__ import annotations

import hashlib
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
for p in (ROOT / "phys-rs_amd", ROOT / "oracle"):
    sys.path.insert(0, str(p))

import oracle as O  # noqa: E402
from sphx import scenarios  # noqa: E402

HERE = Path(__file__).resolve().parent
C3_KEYS = ("x", "y", "z", "vx", "vy", "vz", "m", "h", "bx", "by", "bz", "psi", "u", "id")
C4_KEYS = ("x", "y", "z", "vx", "vy", "vz", "m", "h", "id")


def input_hash(state, keys) -> str:
    h = hashlib.sha256()
    for k in keys:
        h.update(np.ascontiguousarray(state[k]).tobytes())
    return h.hexdigest()


def pair_checksums(ids: np.ndarray, n: int):
    """Three wrapping uint64 checksums of an (id_a, id_b) pair list."""
    key = ids[:, 0].astype(np.uint64) * np.uint64(n) + ids[:, 1].astype(np.uint64)
    with np.errstate(over="ignore"):
        s1 = key.sum(dtype=np.uint64)
        s2 = (key * key).sum(dtype=np.uint64)
        s3 = np.bitwise_xor.reduce(key * np.uint64(0x9E3779B97F4A7C15))
    return np.array([s1, s2, s3], dtype=np.uint64)


def c3_inputs():
    return scenarios.mhd_periodic(random_fill=4194304)


def c3_samples(n: int) -> np.ndarray:
    return np.sort(np.unique((scenarios.SplitMix64(2024).uniform(2048) * n).astype(np.int64)))


def c4_inputs():
    st, p = scenarios.dam_break(nx=64, ny=32, nz=32, v_sigma=0.05, seed=0)
    p15 = scenarios.dam_break(nx=64, ny=32, nz=32, v_sigma=0.05, seed=0, alpha=0.15)[1]
    return st, p, p15


def c4_directions(n: int, k: int = 32) -> np.ndarray:
    rng = scenarios.SplitMix64(99)
    d = rng.normal(k * n * 3).reshape(k, n, 3)
    return d / np.linalg.norm(d.reshape(k, -1), axis=1)[:, None, None]


def c4_samples(n: int) -> np.ndarray:
    return np.sort(np.unique((scenarios.SplitMix64(77).uniform(1024) * n).astype(np.int64)))


def make_c3() -> None:
    t0 = time.time()
    st, p = c3_inputs()
    n = len(st["x"])
    s = O.State(st)
    pr = O.pairs(s, p)
    print(f"c3: {len(pr)} pairs ({time.time() - t0:.0f} s)", flush=True)
    ids = O.id_pairs(s, pr)  # ids equal indices here (id = arange)
    deg = np.bincount(pr.ravel(), minlength=n)
    nsum = (np.bincount(pr[:, 0], weights=pr[:, 1].astype(np.float64), minlength=n)
            + np.bincount(pr[:, 1], weights=pr[:, 0].astype(np.float64), minlength=n))
    O.evaluate(s, p, pr)
    m = st["m"]
    F = np.stack([m * s["ax"], m * s["ay"], m * s["az"]], 1)
    print(f"c3: evaluated ({time.time() - t0:.0f} s)", flush=True)
    smp = c3_samples(n)
    out = dict(input_hash=np.array(input_hash(st, C3_KEYS)), npairs=np.array(len(pr)),
               pair_sums=pair_checksums(ids, n), momentum=F.sum(0), force_abs_sum=np.abs(F).sum(0),
               acc_max=np.array([np.abs(s[k]).max() for k in ("ax", "ay", "az")]),
               sample=smp, degree=deg[smp], nbr_id_sum=nsum[smp],
               rho=s["rho"][smp], p=s["p"][smp], ax=s["ax"][smp], ay=s["ay"][smp], az=s["az"][smp])
    O.mhd_induction(s, p, pr)
    for k in ("bx", "by", "bz"):
        out["ind_" + k] = s[k][smp]
    O.div_clean(s, p, pr)
    for k in ("bx", "by", "bz", "psi"):
        out["cln_" + k] = s[k][smp]
    np.savez_compressed(HERE / "c3_mhd_4M.npz", **out)
    print(f"c3: written ({time.time() - t0:.0f} s)", flush=True)


def make_c4() -> None:
    t0 = time.time()
    st, p, p15 = c4_inputs()
    n = len(st["x"])
    s = O.State(st)
    O.step(s, p15, 100)
    target = np.stack([s["x"], s["y"], s["z"]], 1)
    print(f"c4: target ({time.time() - t0:.0f} s)", flush=True)
    s = O.State(st)
    loss, gv, ga = O.adjoint(s, p, 100, target)
    print(f"c4: adjoint ({time.time() - t0:.0f} s)", flush=True)
    smp = c4_samples(n)
    d = c4_directions(n)
    out = dict(input_hash=np.array(input_hash(st, C4_KEYS)), target_hash=np.array(
        hashlib.sha256(np.ascontiguousarray(target).tobytes()).hexdigest()), loss=np.array(loss),
        grad_alpha=np.array(ga), sample=smp, grad_v=gv[smp], directional=np.einsum("knc,nc->k", d, gv),
        grad_v_norm=np.array(np.linalg.norm(gv)))
    np.savez_compressed(HERE / "c4_adjoint_65k.npz", **out)
    print(f"c4: written ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c3", "c4"]
    if "c3" in which:
        make_c3()
    if "c4" in which:
        make_c4()
