"""Benchmark of the per-timestep periodic fluid update on one MI355X.

Workload (the headline configuration): forced isotropic turbulence, periodic 2pi box at
256^3 cells (32^3 blocks of 8^3), seeded solenoidal initial condition with the
E(k) ~ k^4 exp(-2 (k/4)^2) spectrum scaled to u_rms = 1, band-limited linear forcing
holding the kinetic energy at its initial value, nu = 0.005, CFL 0.3, no ramp. A "step"
is one full timestep of the path: CFL dt from the previous step's max velocity,
advection-diffusion, forcing, pressure RHS, FFT Poisson solve, projection.

The FP32 form (fused decomposition, FP64 reductions and spectral division) is the
measured line; the FP64 plain form (the reference's operator decomposition) is timed
beside it on the same GPU, and the FP64 CPU oracle (OpenMP) on a bounded sample of the
same workload on the host cores.

    python bench.py [--gpus 1] [--steps 90] [--warmup 10]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Algorithmic bytes per cell and step, FP32 form (4-byte fields): every field materialised
# between pipeline stages read once and written once, each 3-D FFT one read plus one write:
# advection 24, pressure RHS 16, FFT r2c + division + c2r 16, gradient (+ reductions) 28.
ALGO_BYTES_PER_CELL_FP32 = 84.0
# The third-order upwind + midpoint RK2 form: stage 1 reads u, writes u_half (24), stage 2
# reads u_half and u, writes u* (36), so advection moves 60 instead of 24.
ALGO_BYTES_PER_CELL_FP32_RK2 = 120.0
# Per-launch algorithmic bytes per cell of each stage of the fused FP32 step.
STAGE_BYTES_PER_CELL = {"advdiff": 24.0, "xfwd": 16.0, "yfwd": 8.0, "zpass": 8.0, "yinv": 8.0, "xinv": 8.0,
                        "gradp": 28.0}
# Stages timed for the report but carrying no per-cell traffic (small reductions).
AUX_STAGES = ("forcing_reduce",)
SEED = 20260101
NU = 0.005
CFL = 0.3


def make_grid(cup, n, precision, decomposition):
    G = cup.Grid((n // 8,) * 3, precision=precision, decomposition=decomposition)
    G.ic_hit(4.0, 1.0, SEED)
    return G


def target_tke(G):
    return G.diagnostics(NU)[18] / float(G.extent.prod())


def time_steps(G, steps, tke, **kw):
    G.synchronize()
    t0 = time.perf_counter()
    G.run_steps(steps, CFL=CFL, nu=NU, rampup=0, spectral_forcing=True, tke_target=tke, **kw)
    G.synchronize()
    return time.perf_counter() - t0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, warm=2, timed=5):
    """FP64 oracle (OpenMP over blocks, the reference CPU path's structure: AoS blocks,
    per-block labs, its own FFT) on the same workload: `warm` + `timed` steps at n^3 on the
    host cores this process is given (OMP_NUM_THREADS, else every core), threads bound close."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import oracle as O
    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    O.lib().orc_set_threads(threads)
    s = O.OracleSim((n // 8,) * 3)
    s.ic_hit(4.0, 1.0, SEED)
    u, v, w = s.velocity()
    tke = 0.5 * float((u * u + v * v + w * w).mean())
    del u, v, w
    step = 0
    for _ in range(warm):
        dt, _ = s.calc_max_timestep(CFL, NU, (0, 0, 0), step, 0)
        s.timestep(dt, NU, spectral_forcing=True, tke_target=tke)
        step += 1
    t0 = time.perf_counter()
    for _ in range(timed):
        dt, _ = s.calc_max_timestep(CFL, NU, (0, 0, 0), step, 0)
        s.timestep(dt, NU, spectral_forcing=True, tke_target=tke)
        step += 1
    el = time.perf_counter() - t0
    return {"value": round(n ** 3 * timed / el / 1e6, 2), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n}^3 forced isotropic turbulence, FP64 oracle (OpenMP, {threads} threads, {cpu_model()}), "
                      f"{warm} warm-up + {timed} timed steps ({el:.1f} s)"}


def c1_rows(cup, with_cpu):
    """C1, the reference's CPU-runnable case: the 64^3 Taylor-Green vortex (the TGV IC with
    uMax = -1), FP64 plain, nu = 0.01, fixed dt = 1e-3, no forcing, all 20 steps timed, on the
    GPU in the FP64 plain and FP32 fused forms and, with the CPU baseline, on the host cores (the
    oracle)."""
    n, steps, nu, dt = 64, 20, 0.01, 1e-3
    rows = {"workload": "Taylor-Green vortex, periodic 2pi box 64^3, nu=0.01, dt=1e-3, 20 steps", "cells": n ** 3,
            "steps": steps}
    for key, prec, dec in (("gpu_fp64_plain", 8, "plain"), ("gpu_fp32_fused", 4, "fused")):
        G = cup.Grid((n // 8,) * 3, precision=prec, decomposition=dec)
        G.ic_taylor_green(-1.0)
        G.run_steps(2, nu=nu, fixed_dt=dt)  # warm-up: first launches
        G.ic_taylor_green(-1.0)
        G.synchronize()
        t0 = time.perf_counter()
        G.run_steps(steps, nu=nu, fixed_dt=dt)
        G.synchronize()
        el = time.perf_counter() - t0
        rows[key] = {"value": round(n ** 3 * steps / el / 1e6, 1), "unit": "Mcell-steps/s",
                     "ms_per_step": round(el / steps * 1e3, 4)}
        del G
    if with_cpu:
        from oracle import oracle as O
        O.build()
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        O.lib().orc_set_threads(threads)
        s = O.OracleSim((n // 8,) * 3)
        s.ic_taylor_green(-1.0)
        t0 = time.perf_counter()
        for _ in range(steps):
            s.timestep(dt, nu)
        el = time.perf_counter() - t0
        rows["cpu_fp64_oracle"] = {"value": round(n ** 3 * steps / el / 1e6, 2), "unit": "Mcell-steps/s",
                                   "ms_per_step": round(el / steps * 1e3, 3), "cores": threads, "kind": "port",
                                   "cpu": cpu_model()}
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=90)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256, help="cells per side")
    ap.add_argument("--plain-steps", type=int, default=20, help="timed steps of the FP64 plain form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-c1", action="store_true", help="skip the 64^3 Taylor-Green rows")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is described, not built)", file=sys.stderr)
        sys.exit(2)

    import cubismup_3d_amd as cup
    n = args.size
    ncell = n ** 3

    # FP32 form, fused decomposition
    G = make_grid(cup, n, 4, "fused")
    tke = target_tke(G)
    time_steps(G, args.warmup, tke)
    # per-stage launch durations over a few steps (HIP events on the context's stream)
    stage_ms = {}
    for st in (*STAGE_BYTES_PER_CELL, *AUX_STAGES):
        G.stage_timing_begin(st, 3)
        time_steps(G, 3, tke)
        stage_ms[st] = G.stage_timing_end()[0]
    dominant = max(STAGE_BYTES_PER_CELL, key=lambda k: stage_ms[k])
    # timed region: K steps, events around the dominant kernel in every one of them after the
    # first (whose events follow the idle stream of the opening synchronize)
    G.stage_timing_begin(dominant, args.steps)
    el = time_steps(G, args.steps, tke)
    dom_ms, dom_n = G.stage_timing_end()
    state = G.step_state()
    ms = el / args.steps * 1e3
    value = ncell * args.steps / el / 1e6
    step_frac = ALGO_BYTES_PER_CELL_FP32 * ncell / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9)
    achieved = STAGE_BYTES_PER_CELL[dominant] * ncell / (dom_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{dominant}_{n}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    del G

    # FP32 fused form with third-order upwind advection and midpoint RK2, same workload
    R = make_grid(cup, n, 4, "fused")
    tke_r = target_tke(R)
    time_steps(R, 5, tke_r, scheme=2)
    rk_steps = max(10, args.steps // 3)
    el_rk = time_steps(R, rk_steps, tke_r, scheme=2)
    rk_ms = el_rk / rk_steps * 1e3
    del R

    # FP64 plain form (reference decomposition), same workload
    P = make_grid(cup, n, 8, "plain")
    tke64 = target_tke(P)
    time_steps(P, 3, tke64)
    el64 = time_steps(P, args.plain_steps, tke64)
    v64 = ncell * args.plain_steps / el64 / 1e6
    del P

    out = {
        "metric": "Mcell-steps/s at 256^3 HIT, FP32 form vs FP64 plain on GPU; % HBM roofline",
        "value": round(value, 1),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"forced isotropic turbulence, periodic 2pi box {n}^3 ({n // 8}^3 blocks of 8^3), "
                               f"seeded art-spectrum IC (k0=4, u_rms=1), linear forcing to the initial energy, "
                               f"nu={NU}, CFL {CFL}, central advection + Euler, FFT Poisson",
                   "cells": ncell, "decomposition": "fused", "precision": "fp32 fields, fp64 reductions"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": STAGE_BYTES_PER_CELL[dominant] * ncell,
                     "avg_launch_ms": round(dom_ms, 5), "launches_timed": dom_n},
        "step_roofline": {"algorithmic_bytes_per_cell": ALGO_BYTES_PER_CELL_FP32,
                          "achieved_GBs": round(ALGO_BYTES_PER_CELL_FP32 * ncell / (ms * 1e-3) / 1e9, 1),
                          "frac": round(step_frac, 4)},
        "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "fp32_upwind3_rk2": {"value": round(ncell / (rk_ms * 1e-3) / 1e6, 1), "unit": "Mcell-steps/s",
                             "steps": rk_steps, "ms_per_step": round(rk_ms, 4),
                             "algorithmic_bytes_per_cell": ALGO_BYTES_PER_CELL_FP32_RK2,
                             "step_roofline_frac": round(ALGO_BYTES_PER_CELL_FP32_RK2 * ncell / (rk_ms * 1e-3)
                                                         / (HBM_PEAK_GBS * 1e9), 4)},
        "fp64_plain": {"value": round(v64, 1), "unit": "Mcell-steps/s", "steps": args.plain_steps,
                       "ms_per_step": round(el64 / args.plain_steps * 1e3, 4)},
        "speedup_vs_fp64_plain": round(value / v64, 3),
        "final_state": {"step": state["step"], "time": state["time"], "umax": state["umax"]},
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(n, 2, args.cpu_steps)
        out["cpu_baseline"] = cb
    if not args.no_c1:
        out["c1_tgv64"] = c1_rows(cup, not args.no_cpu_baseline)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
