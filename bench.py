#!/usr/bin/env python3
"""SPH particle-steps/s on one MI355X: the 1,048,576-particle dam break.

Workload: 128 x 128 x 64 lattice, spacing 0.01 m, +-1 % seeded jitter, in a
3.2 x 1.6 x 1.6 m reflective box; cubic spline h = 1.3 dx, Tait rho0 = 1000,
gamma = 7, c0 = 20, Monaghan alpha = 0.1 (beta = 0), g = -9.81 z, dt = 1e-4.
A step is one velocity-Verlet tick of the reference schedule (main.rs:103-157)
over all particles, state resident in HBM (no host copies inside the timed
region).  The neighbour list carries a skin of 0.04 h and is rebuilt once
some particle has moved half of it (kernels.neighbour.launches counts the
builds in the timed steps); every pass still uses exactly r < 2 h_avg.
The FP64 drop-in is `value`; beside it the line reports
  * "per_tick_rebuild": the same FP64 run with the list rebuilt every tick;
  * "mixed": the FP32-mixed form on the same workload (ms per step, speed-up)
    and its error against FP64 at the same positions (the state the FP64 run
    ends in): max and RMS relative error of density and acceleration;
  * "adjoint": the differentiable step on the 65,536-particle dam break
    (64 x 32 x 32, seeded initial velocities sigma 0.05 m/s, 100 steps,
    checkpoint every 10, target = the same start run with alpha = 0.15):
    forward and adjoint (recompute + reverse) milliseconds per step.
  * "c2_500": the whole 500-step run of the same workload from the upload,
    FP64 and FP32-mixed, mean ms per tick over every tick and every
    neighbour-list build (the timed steps above are a quiet stretch of it);
  * "cpu_baseline": the CPU path on the host's cores (all cores, gather form;
    and one thread, the reference's pair structure).
--precision-report runs the 500-step FP64 / FP32-mixed comparison (errors at
the FP64 end state, total-energy drift of each form) and prints it instead.
--mhd-report runs the 4,194,304-particle periodic MHD box (seeded uniform
random fill, <N_nb> = 40, induction + cleaning + Maxwell stress) in both
forms and prints the per-kernel breakdown instead (--mhd-particles to resize;
--mhd-lattice 64 --steps 100 for the 262,144-particle jittered lattice).

Prints one JSON line (rank 0).  --gpus N > 1 is not supported: this build runs
on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "phys-rs_amd"))

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (packed)

# Algorithmic work per unit (f64 SoA, state touched once per kernel, neighbour
# reads on-chip): bytes per particle per launch, FLOPs per directed neighbour
# pair (add/sub/mul/div/sqrt = 1, fma = 2) from the reference formulas.
ALG = {
    "sort":      {"bytes": 224, "flops_pair": 0.0},   # hash 28 + counting sort 12 + reorder 184
    # x, y, z, h read and one 4-byte list entry written per neighbour; candidates are not work
    "neighbour": {"bytes": 32, "bytes_pair": 4, "flops_pair": 0.0},
    "density":   {"bytes": 56, "flops_pair": 26.0},   # x, m, h -> rho, P
    "force":     {"bytes": 128, "flops_pair": 58.0},  # x, v, m, h, rho, P -> F_p, a*
    "visc_only": {"bytes": 120, "flops_pair": 45.0},  # x, v, m, h, rho, F_p -> a_n
    "integrate": {"bytes": 96, "flops_pair": 0.0},    # vv_position / vv_velocity, each 96
    # induction (73 flops a pair: mhd.rs:68-152) and cleaning (38: mhd.rs:172-259),
    # one launch each, so 55.5 per launch on average
    "mhd":       {"bytes": 200, "flops_pair": 55.5},
}
STEP_BYTES_PER_PARTICLE = 720.0          # sum over one tick (no MHD)

# HBM bytes per launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a
# 256 MiB copy) of the FP64 kernels on this workload, from the rocprofv3
# counter runs summarised in profiles/dam_break_1M_fp64_mixed_kernels.md.
MEASURED_TRAFFIC = {
    "force": (301.6 + 50.7) * 1e6,
    "visc_only": (325.3 + 25.4) * 1e6,
    "density": (287.8 + 92.6) * 1e6,
}
STEP_FLOPS_PER_PAIR = 26.0 + 58.0 + 45.0


END_STATE = {}


def sync_all(ctx) -> None:
    ctx.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def run_gpu(precision: str, steps: int, warmup: int, skin=None):
    import sphx
    from sphx import scenarios

    st, p = scenarios.dam_break(nx=128, ny=128, nz=64, spacing=0.01, box=(3.2, 1.6, 1.6))
    p.precision = sphx.PREC_FP64 if precision == "fp64" else sphx.PREC_FP32_MIXED
    if skin is not None:
        p.neighbour_skin = skin
    n = len(st["x"])
    ctx = sphx.Context(n, p)
    ctx.upload(st)
    ctx.step(warmup)
    sync_all(ctx)
    # the timed region brackets only the dominant kernel (the full force pass)
    # with HIP events: an event pair costs a few microseconds, and events
    # around all five kernels of a tick add about 7 % to the step
    ctx.enable_timing(sphx.TIMING_FORCE)
    ctx.kernel_times(reset=True)
    sync_all(ctx)
    t0 = time.perf_counter()
    ctx.step(steps)
    sync_all(ctx)
    t1 = time.perf_counter()
    times = ctx.kernel_times(reset=True)
    # per-kernel breakdown: the next `steps` steps with every kernel bracketed
    ctx.enable_timing(True)
    ctx.step(steps)
    sync_all(ctx)
    BREAKDOWN[precision] = ctx.kernel_times(reset=True)
    ctx.enable_timing(False)
    e = ctx.energy()
    end = download_state(ctx)
    end.update(m=st["m"], h=st["h"], id=st["id"])
    END_STATE[precision] = end
    ctx.close()
    return n, t1 - t0, times, e


BREAKDOWN = {}


def eval_errors(st):
    """Density and acceleration of the FP32-mixed form against FP64 at one state."""
    import numpy as np
    import sphx
    from sphx import scenarios

    _, p = scenarios.dam_break(nx=128, ny=128, nz=64, spacing=0.01, box=(3.2, 1.6, 1.6))
    out = {}
    for prec in ("fp64", "mixed"):
        p.precision = sphx.PREC_FP64 if prec == "fp64" else sphx.PREC_FP32_MIXED
        with sphx.Context(len(st["x"]), p) as ctx:
            ctx.upload(st)
            ctx.forces()
            out[prec] = ctx.download(sphx.F_RHO | sphx.F_ACC)
    a, b = out["mixed"], out["fp64"]
    drho = np.abs(a["rho"] - b["rho"]) / np.abs(b["rho"])
    acc64 = np.stack([b["ax"], b["ay"], b["az"]], axis=1)
    acc32 = np.stack([a["ax"], a["ay"], a["az"]], axis=1)
    # acceleration error relative to the largest acceleration in the field
    # (componentwise relative error is meaningless where a ~ 0)
    dacc = np.linalg.norm(acc32 - acc64, axis=1) / np.max(np.linalg.norm(acc64, axis=1))
    return {"density_rel_max": float(drho.max()), "density_rel_rms": float(np.sqrt(np.mean(drho ** 2))),
            "accel_rel_max": float(dacc.max()), "accel_rel_rms": float(np.sqrt(np.mean(dacc ** 2)))}


def download_state(ctx):
    import sphx
    s = ctx.download(sphx.F_POS | sphx.F_VEL)
    return s


def run_adjoint(steps: int = 100, every: int = 10, reps: int = 3):
    """Forward vs adjoint time on the 65,536-particle differentiable dam break."""
    import numpy as np
    import sphx
    from sphx import scenarios

    st, p = scenarios.dam_break(nx=64, ny=32, nz=32, v_sigma=0.05, seed=0)
    n = len(st["x"])
    p15 = scenarios.dam_break(nx=64, ny=32, nz=32, v_sigma=0.05, seed=0, alpha=0.15)[1]
    with sphx.Context(n, p15) as ctx:
        ctx.upload(st)
        ctx.step(steps)
        f = ctx.download(sphx.F_POS)
    target = np.stack([f["x"], f["y"], f["z"]], axis=1)
    with sphx.Context(n, p) as ctx:
        ctx.upload(st)
        ctx.adjoint(steps, target, checkpoint_every=every)  # warm-up (allocates the tape)
        fwd, adj = [], []
        for _ in range(reps):
            ctx.upload(st)
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.step(steps)
            ctx.synchronize()
            fwd.append(time.perf_counter() - t0)
            ctx.upload(st)
            ctx.synchronize()
            t0 = time.perf_counter()
            out = ctx.adjoint(steps, target, checkpoint_every=every)
            adj.append(time.perf_counter() - t0)
    tf, ta = min(fwd), min(adj)
    return {"workload": "dam_break_65k_adjoint", "particles": n, "steps": steps, "checkpoint_every": every,
            "forward_ms_per_step": round(tf / steps * 1e3, 4),
            "forward_plus_adjoint_ms_per_step": round(ta / steps * 1e3, 4),
            "adjoint_over_forward": round((ta - tf) / tf, 3),
            "loss": out["loss"], "grad_alpha": out["grad_alpha"],
            "note": "adjoint = checkpoint recompute + reverse pass (the call also runs the forward it "
                    "differentiates, subtracted here); host-timed, state resident"}


def mhd_eval_errors(st, p):
    """FP32-mixed against FP64 at one MHD state: density and acceleration of a
    force evaluation, the change of B by induction and of psi by cleaning."""
    import copy

    import numpy as np
    import sphx

    res = {}
    for prec in ("fp64", "mixed"):
        q = copy.copy(p)
        q.precision = sphx.PREC_FP64 if prec == "fp64" else sphx.PREC_FP32_MIXED
        with sphx.Context(len(st["x"]), q) as ctx:
            ctx.upload(st)
            ctx.forces()
            a = ctx.download(sphx.F_RHO | sphx.F_ACC)
            ctx.mhd_induction()
            b1 = ctx.download(sphx.F_B)
            ctx.div_clean()
            b2 = ctx.download(sphx.F_PSI)
        res[prec] = (a, b1, b2)
    (a64, b164, b264), (a32, b132, b232) = res["fp64"], res["mixed"]
    drho = np.abs(a32["rho"] / a64["rho"] - 1)
    acc64 = np.stack([a64[k] for k in ("ax", "ay", "az")], 1)
    acc32 = np.stack([a32[k] for k in ("ax", "ay", "az")], 1)
    dacc = np.linalg.norm(acc32 - acc64, axis=1) / np.linalg.norm(acc64, axis=1).max()
    dB = max(float(np.abs(b132[k] - b164[k]).max() / np.abs(b164[k] - st[k]).max()) for k in ("bx", "by", "bz"))
    dpsi = float(np.abs(b232["psi"] - b264["psi"]).max() / np.abs(b264["psi"] - st["psi"]).max())
    return {"density_rel_max": float(drho.max()), "density_rel_rms": float(np.sqrt(np.mean(drho ** 2))),
            "accel_rel_max": float(dacc.max()), "accel_rel_rms": float(np.sqrt(np.mean(dacc ** 2))),
            "induction_dB_rel_max": dB, "cleaning_dpsi_rel_max": dpsi,
            "note": "at the FP64 run's end state; acceleration relative to the largest |a|, dB and dpsi relative "
                    "to the largest FP64 change"}


# List skin (in units of h) of each form on the random-fill MHD box, the
# optimum of the measured sweep (bench.py --mhd-report --mhd-skin S on the
# 4,194,304-particle box, where a 0.04 h list lives two ticks): the FP32-mixed
# passes cost less per list entry, so the longer lists of a larger skin cost
# them less than the builds it saves.  The lattice box keeps 0.04 h (its lists
# live about twelve ticks; 0.12 h measured slower there in both forms).
MHD_SKIN_H = {"fp64": 0.04, "mixed": 0.12}


def mhd_report(n_fill: int, steps: int, warmup: int, lattice: int = 0, skin_h: float = -1.0):
    """Per-kernel breakdown of the periodic MHD box in the FP64 and FP32-mixed forms
    (seeded random fill of n_fill particles, or the lattice^3 jittered lattice).
    `steps` ticks are timed with only the sort and build bracketed (their count);
    the per-kernel breakdown comes from the next `steps` ticks, every kernel
    bracketed.  skin_h < 0: each form's skin from MHD_SKIN_H (random fill) or
    0.04 h (lattice)."""
    import sphx
    from sphx import scenarios

    if lattice:
        st, p = scenarios.mhd_periodic(n_side=lattice)
    else:
        st, p = scenarios.mhd_periodic(random_fill=n_fill)
    n = len(st["x"])
    kind = f"lattice_{lattice}^3" if lattice else "random"
    out = {"report": "mhd", "workload": f"mhd_periodic_{kind}_{n}", "particles": n, "steps": steps,
           "h": float(st["h"][0])}
    for prec in ("fp64", "mixed"):
        p.precision = sphx.PREC_FP64 if prec == "fp64" else sphx.PREC_FP32_MIXED
        sk = skin_h if skin_h >= 0 else (0.04 if lattice else MHD_SKIN_H[prec])
        p.neighbour_skin = sk * float(st["h"].max())
        with sphx.Context(n, p) as ctx:
            ctx.upload(st)
            ctx.step(warmup)
            sync_all(ctx)
            ctx.enable_timing(sphx.TIMING_SORT | sphx.TIMING_NEIGHBOUR)
            ctx.kernel_times(reset=True)
            t0 = time.perf_counter()
            ctx.step(steps)
            sync_all(ctx)
            secs = time.perf_counter() - t0
            builds = ctx.kernel_times(reset=True)["neighbour"]["launches"]
            ctx.enable_timing(True)
            ctx.step(steps)
            sync_all(ctx)
            times = ctx.kernel_times(reset=True)
            e = ctx.energy()
            if prec == "fp64":
                end = ctx.download(sphx.F_POS | sphx.F_VEL | sphx.F_B | sphx.F_PSI | sphx.F_U)
                end.update(m=st["m"], h=st["h"], id=st["id"])
        pairs = e["total_neighbours"]
        kern = {}
        for k, v in times.items():
            if v["launches"] == 0:
                continue
            avg_s = v["ms"] / v["launches"] * 1e-3
            gbs = (ALG[k]["bytes"] * n + ALG[k].get("bytes_pair", 0) * pairs) / avg_s / 1e9
            tflops = ALG[k]["flops_pair"] * pairs / avg_s / 1e12
            # the mixed form's pair math is f32 (against the f32 vector peak)
            f32 = prec == "mixed" and k in ("density", "force", "visc_only", "mhd")
            kern[k] = {"ms_per_step": round(v["ms"] / steps, 4), "avg_ms": round(avg_s * 1e3, 4),
                       "launches": v["launches"], "alg_GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                       "alg_TFLOPs": round(tflops, 3),
                       "flop_frac": round(tflops / (FP32_PEAK_TFLOPS if f32 else FP64_PEAK_TFLOPS), 4),
                       "flop_peak": "fp32 157.3" if f32 else "fp64 78.6"}
        out[prec] = {"ms_per_step": round(secs / steps * 1e3, 4), "particle_steps_per_s": round(n * steps / secs, 1),
                     "skin_h": sk, "neighbour_builds": builds,
                     "directed_neighbour_pairs": pairs, "max_neighbours": e["max_neighbours"], "kernels": kern,
                     "kernels_note": "HIP events around every kernel over the next `steps` ticks"}
    out["mixed_speedup"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    out["mixed_errors"] = mhd_eval_errors(end, p)
    return out


def precision_report(steps: int = 500):
    import numpy as np
    import sphx
    from sphx import scenarios

    st, p = scenarios.dam_break(nx=128, ny=128, nz=64, spacing=0.01, box=(3.2, 1.6, 1.6))
    n = len(st["x"])
    res = {}
    end = None
    for prec in ("fp64", "mixed"):
        p.precision = sphx.PREC_FP64 if prec == "fp64" else sphx.PREC_FP32_MIXED
        with sphx.Context(n, p) as ctx:
            ctx.upload(st)
            ctx.forces()
            e0 = ctx.energy()
            ctx.step(steps)
            ctx.forces()
            e1 = ctx.energy()
            tot = lambda e: e["kinetic"] + e["potential"] + e["internal"]
            res[prec] = {"energy_start": tot(e0), "energy_end": tot(e1),
                         "energy_drift_rel": (tot(e1) - tot(e0)) / abs(tot(e0))}
            if prec == "fp64":
                end = download_state(ctx)
                end.update(m=st["m"], h=st["h"], id=st["id"])
    res["errors_at_fp64_end_state"] = eval_errors(end)
    return {"report": "precision", "workload": "dam_break_1M", "steps": steps, **res}


def roofline(n: int, times: dict, pairs: int, steps: int, timed: dict = None):
    """Per-kernel rates from `times`; the dominant kernel's achieved rate from
    `timed` (the timed region's events) when it holds that kernel."""
    kern = {}
    for k, v in times.items():
        if v["launches"] == 0:
            continue
        avg_s = v["ms"] / v["launches"] * 1e-3
        b = ALG[k]["bytes"] * n + ALG[k].get("bytes_pair", 0) * pairs
        f = ALG[k]["flops_pair"] * pairs
        kern[k] = {
            "avg_ms": round(v["ms"] / v["launches"], 5),
            "launches": v["launches"],
            "ms_per_step": round(v["ms"] / steps, 5),
            "alg_GBs": round(b / avg_s / 1e9, 1),
            "alg_TFLOPs": round(f / avg_s / 1e12, 3),
        }
    dom = max(kern, key=lambda k: kern[k]["ms_per_step"])
    v = timed[dom] if timed and timed.get(dom, {}).get("launches") else times[dom]
    avg_s = v["ms"] / v["launches"] * 1e-3
    b = ALG[dom]["bytes"] * n + ALG[dom].get("bytes_pair", 0) * pairs
    f = ALG[dom]["flops_pair"] * pairs
    ridge = FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
    if f > 0 and f / b > ridge:
        # FP64 vector ALU bound; MI355X's FP64 vector and matrix peaks coincide
        # and no MFMA instruction is used.
        ach = f / avg_s / 1e12
        rf = {"kernel": dom, "bound": "mfma", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
              "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": MEASURED_TRAFFIC.get(dom),
              "note": "FP64 vector ALU (no MFMA); peak = MI355X FP64 78.6 TFLOP/s",
              "hbm_GBs": round(b / avg_s / 1e9, 1)}
    else:
        ach = b / avg_s / 1e9
        rf = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
              "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": MEASURED_TRAFFIC.get(dom)}
    return rf, kern


def c2_full_run(steps: int = 500):
    """The whole 500-step 1M dam break from the initial upload, in both forms:
    mean ms per step over every tick, every neighbour-list build included."""
    import sphx
    from sphx import scenarios

    st, p = scenarios.dam_break(nx=128, ny=128, nz=64, spacing=0.01, box=(3.2, 1.6, 1.6))
    n = len(st["x"])
    out = {"workload": "dam_break_1M", "steps": steps}
    for prec in ("fp64", "mixed"):
        p.precision = sphx.PREC_FP64 if prec == "fp64" else sphx.PREC_FP32_MIXED
        with sphx.Context(n, p) as ctx:
            ctx.upload(st)
            ctx.enable_timing(sphx.TIMING_SORT | sphx.TIMING_NEIGHBOUR)  # the builds only
            ctx.kernel_times(reset=True)
            sync_all(ctx)
            t0 = time.perf_counter()
            ctx.step(steps)
            sync_all(ctx)
            secs = time.perf_counter() - t0
            kt = ctx.kernel_times(reset=True)
        out[prec] = {"ms_per_step": round(secs / steps * 1e3, 4), "particle_steps_per_s": round(n * steps / secs, 1),
                     "neighbour_builds": kt["neighbour"]["launches"],
                     "build_ms_per_step": round((kt["neighbour"]["ms"] + kt["sort"]["ms"]) / steps, 4)}
    out["mixed_speedup"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    return out


def host_cpu():
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(steps: int):
    """The CPU path on this host's cores, on the same 1M dam break:
      * all cores (the top-level value): the oracle's gather form (oracle/sph_omp.c,
        OpenMP over particles, lists rebuilt for each evaluation as CandidatePairs
        is each tick), full velocity-Verlet ticks for at least 10 s; beside it, shaped like
        the reference's bench_sph_cpu (tests/gpu_bench.rs:220-254), one warm-up
        and 20 timed force evaluations on a fixed list (particles/s = n / mean);
      * one thread: the oracle's pair form (the reference's structure: an
        undirected pair list, two-sided accumulation per system), `steps` ticks."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from sphx import scenarios

    st, p = scenarios.dam_break(nx=128, ny=128, nz=64, spacing=0.01, box=(3.2, 1.6, 1.6))
    n = len(st["x"])
    model, nproc = host_cpu()
    g = O.Gather()
    threads = O.Gather.threads()
    s = O.State(st)
    g.build(s, p)
    g.evaluate(s, p)  # warm-up
    t0 = time.perf_counter()
    for _ in range(20):
        g.evaluate(s, p)
    t_eval = (time.perf_counter() - t0) / 20
    # all cores: ticks in pairs until at least 10 s of work
    s = O.State(st)
    ticks = 0
    t0 = time.perf_counter()
    while True:
        g.step(s, p, 2)
        ticks += 2
        t_all = time.perf_counter() - t0
        if t_all >= 10.0:
            break
    g.close()
    s = O.State(st)
    t0 = time.perf_counter()
    O.step(s, p, steps)
    t_one = time.perf_counter() - t0
    return {"value": round(n * ticks / t_all, 1), "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle gather form (C, OpenMP, {threads} threads) on the same 1,048,576-particle dam "
                      f"break, {ticks} ticks, {t_all:.1f} s",
            "host": {"cpu_model": model, "nproc": nproc, "omp_threads": threads},
            "evaluation_like_bench_sph_cpu": {"particles_per_s": round(n / t_eval, 1), "ms_per_evaluation":
                                              round(t_eval * 1e3, 2), "threads": threads, "evaluations": 20},
            "single_thread": {"value": round(n * steps / t_one, 1), "unit": "particle-steps/s", "cores": 1,
                              "sample": f"oracle pair form (the reference's structure), 1 thread, {steps} "
                                        f"ticks, {t_one:.1f} s"}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the mixed and adjoint side runs")
    ap.add_argument("--precision-report", action="store_true")
    ap.add_argument("--mhd-report", action="store_true")
    ap.add_argument("--adjoint-report", action="store_true", help="only the 65,536-particle adjoint/forward run")
    ap.add_argument("--mhd-particles", type=int, default=4194304)
    ap.add_argument("--mhd-skin", type=float, default=-1.0, help="list skin in units of h (default 0.04)")
    ap.add_argument("--mhd-lattice", type=int, default=0,
                    help="use the n^3 jittered lattice (64: the 262,144-particle box) instead of the random fill")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device-to-device copy (FETCH/WRITE_SIZE calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 25, dtype=torch.float64, device="cuda")  # 256 MiB
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8, "copies": 5}))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU domain decomposition is a design, not part of this build)", file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0
    if args.precision_report:
        print(json.dumps(precision_report()), flush=True)
        return 0
    if args.adjoint_report:
        print(json.dumps(run_adjoint()), flush=True)
        return 0
    if args.mhd_report:
        steps = args.steps if args.mhd_lattice else min(args.steps, 20)
        print(json.dumps(mhd_report(args.mhd_particles, steps, min(args.warmup, 2), args.mhd_lattice,
                                    args.mhd_skin)), flush=True)
        return 0

    n, secs, times, e = run_gpu(args.precision, args.steps, args.warmup)
    pairs = e["total_neighbours"]
    value = n * args.steps / secs
    ms_step = secs / args.steps * 1e3
    rf, kern = roofline(n, BREAKDOWN[args.precision], pairs, args.steps, timed=times)
    rf["timed_launches"] = times["force"]["launches"]
    step_s = secs / args.steps
    out = {
        "metric": "SPH particle-steps/s @1M (FP64 vs FP32-mixed, err<=1e-6); adjoint/fwd time",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-pair/f64-accumulate",
        "data": "synthetic",
        "config": {"workload": "dam_break_1M", "particles": n, "lattice": "128x128x64", "spacing_m": 0.01,
                   "h_m": 0.013, "box_m": [3.2, 1.6, 1.6], "dt_s": 1e-4, "precision": args.precision,
                   "directed_neighbour_pairs": pairs, "max_neighbours": e["max_neighbours"]},
        "roofline": rf,
        "step_roofline": {
            "hbm_frac": round(n * STEP_BYTES_PER_PARTICLE / (HBM_PEAK_GBS * 1e9) / step_s, 4),
            "fp64_frac": round(pairs * STEP_FLOPS_PER_PAIR / (FP64_PEAK_TFLOPS * 1e12) / step_s, 4),
            "alg_bytes_per_particle": STEP_BYTES_PER_PARTICLE,
            "alg_flops_per_pair": STEP_FLOPS_PER_PAIR,
        },
        "kernels": kern,
        "kernels_note": "per-kernel HIP-event times of the next `steps` steps with every kernel bracketed "
                        "(the timed region brackets only the force pass, whose times give roofline.achieved)",
    }
    if not args.no_extras and args.precision == "fp64":
        n2, secs2, times2, _ = run_gpu("mixed", args.steps, args.warmup)
        mixed = {"ms_per_step": round(secs2 / args.steps * 1e3, 4),
                 "value": round(n2 * args.steps / secs2, 1),
                 "speedup_vs_fp64": round(secs / secs2, 3)}
        mixed.update(eval_errors(END_STATE["fp64"]))
        mixed["kernels"] = {k: {"avg_ms": round(v["ms"] / v["launches"], 5), "launches": v["launches"],
                                "ms_per_step": round(v["ms"] / args.steps, 5)}
                            for k, v in BREAKDOWN["mixed"].items() if v["launches"]}
        out["mixed"] = mixed
        # the same FP64 run with the neighbour list rebuilt at every tick
        n3, secs3, times3, _ = run_gpu("fp64", args.steps, args.warmup, skin=0.0)
        out["per_tick_rebuild"] = {"ms_per_step": round(secs3 / args.steps * 1e3, 4),
                                   "value": round(n3 * args.steps / secs3, 1),
                                   "neighbour_builds_per_step": round(
                                       BREAKDOWN["fp64"]["neighbour"]["launches"] / args.steps, 3)}
        out["adjoint"] = run_adjoint()
        out["c2_500"] = c2_full_run()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
