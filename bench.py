"""Throughput of the He-4 TDDFT + Bohmian-impurity time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]

A step is one Hamming predictor-corrector step of the reference time loop
(steppc + the potential evaluation that follows it, DFT4He3d.f90:855-863)
on the 128^3, N=500, 10^4-particle workload, with every array resident in
HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "4hetddft-isotropic-bohmian_amd"))


def model_bytes(n, n_par, precision, ironx=3):
    """Algorithmic HBM bytes of one PC step (stage model of the path).

    Veval = 22(R+S) + 12k + (44+2 ironx)R; psi side = 16(16N) + 4(8N);
    particles = N_par(2*268 + 288).  R = field bytes, S = 2*(bytes per
    element)*H, k = (bytes per element)*H, H = (n/2+1) n^2.
    """
    N = n ** 3
    H = (n // 2 + 1) * n * n
    b = 8 if precision == 0 else 4
    R = b * N
    S = 2 * b * H
    k = b * H
    veval = 22 * (R + S) + 12 * k + (44 + 2 * ironx) * R
    psi_side = 16 * (16 * N) + 4 * (8 * N)
    parts = n_par * (2 * 268 + 288)
    return 2 * veval + psi_side + parts


def run_gpu(wl, precision, steps, warmup):
    import hetddft
    import inputs
    p = inputs.params_for(wl, precision=precision)
    k = hetddft.make_kernels(p)
    psi, r, v = inputs.initial_state(wl)
    g = hetddft.Instance(p)
    g.set_kernels(k)
    g.set_state(psi, r, v)
    g.start()
    for _ in range(3):                      # RK-Gill start-up (DFT4He3d.f90:851)
        g.iterate(wl.deltat, want_errors=False)
    for _ in range(warmup):
        g.iterate(wl.deltat, want_errors=False)
    g.sync()
    e0 = g.energy()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.iterate(wl.deltat, want_errors=False)
    g.sync()
    t1 = time.perf_counter()
    errs = g.last_errors()
    e1 = g.energy()
    out = dict(seconds=t1 - t0, steps_per_s=steps / (t1 - t0), energy0=e0, energy1=e1, errs=errs,
               device_bytes=g.device_bytes)
    g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="time one precision only (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this path has no multi-GPU mode "
              "(an instance is one GPU's share; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import inputs
    wl = inputs.WORKLOADS[args.workload]
    if args.only:
        res = run_gpu(wl, 0 if args.only == "fp64" else 1, args.steps, args.warmup)
        print(json.dumps({"only": args.only, "steps_per_s": res["steps_per_s"],
                          "ms_per_step": 1000.0 / res["steps_per_s"]}))
        return
    res64 = run_gpu(wl, 0, args.steps, args.warmup)
    res32 = run_gpu(wl, 1, args.steps, args.warmup)
    e64 = res64["energy1"]["etot"]
    e32 = res32["energy1"]["etot"]
    rel_e = abs(e32 - e64) / abs(e64)
    use_mixed = rel_e <= 1e-6
    best = res32 if use_mixed else res64
    prec = 1 if use_mixed else 0
    ms = 1000.0 / best["steps_per_s"]
    line = {
        "metric": "PC steps/s at 128^3 N=500 (mixed if rel energy err <= 1e-6 vs FP64)",
        "value": best["steps_per_s"],
        "unit": "PC steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if use_mixed else "f64",
        "data": "synthetic (seeded Woods-Saxon droplet, Gaussian particle cloud)",
        "config": {"workload": f"{wl.name}: {wl.n}^3 grid dx={wl.h} A, N={wl.n_he} He, "
                               f"N_par={wl.n_par}, dt={wl.dt_ps*1000:.1f} fs, OTC, ironx={wl.ironx}",
                   "parallelism": "single GPU"},
        "fp64_steps_per_s": res64["steps_per_s"],
        "mixed_steps_per_s": res32["steps_per_s"],
        "mixed_vs_fp64_speedup": res32["steps_per_s"] / res64["steps_per_s"],
        "mixed_rel_energy_err": rel_e,
        "model_gb_per_step_fp64": model_bytes(wl.n, wl.n_par, 0, wl.ironx) / 1e9,
        "model_gb_per_step_mixed": model_bytes(wl.n, wl.n_par, 1, wl.ironx) / 1e9,
        "model_gbs_fp64": model_bytes(wl.n, wl.n_par, 0, wl.ironx) * res64["steps_per_s"] / 1e9,
        "model_gbs_mixed": model_bytes(wl.n, wl.n_par, 1, wl.ironx) * res32["steps_per_s"] / 1e9,
        "device_bytes": res64["device_bytes"],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
