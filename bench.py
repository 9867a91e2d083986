"""Throughput of the He-4 TDDFT + Bohmian-impurity time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one Hamming predictor-corrector step of the reference time loop
(steppc, which holds one potential evaluation, plus the potential evaluation
the driver runs after it: DFT4He3d.f90:855-863) on the 128^3, N=500,
10^4-particle workload, with every array resident in HBM.  The FP64 path and
the mixed path (FP32 convolutions and potential fields, FP64 psi, particles
and reductions) are both timed; the headline value is the mixed rate when
its total energy stays within 1e-6 (relative) of the FP64 run, else the FP64
rate.  Prints one JSON line.
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

METRIC = "PC steps/s at 128^3 N=500, mixed vs FP64, at ≤1e-6 rel energy err; HBM GB/s"
HBM_PEAK_GBS = 8000.0


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


def _load_tuning(path, precision):
    """A tuning saved by --save-tuning for this precision, as a hetddft.Tuning, or None."""
    import hetddft
    if not path:
        return None
    rec = json.loads(Path(path).read_text()).get("fp64" if precision == 0 else "mixed")
    if not rec:
        return None
    t = hetddft.Tuning()
    for k in range(5):
        t.tile[k], t.zc[k], t.prefetch[k] = rec["tile"][k], rec["zc"][k], rec["prefetch"][k]
    t.fft_z_tile, t.fft_z_wgs = rec["fft_z_tile"], rec["fft_z_wgs"]
    return t


def _tuning_record(t):
    return {"tile": list(t.tile), "zc": list(t.zc), "prefetch": list(t.prefetch),
            "fft_z_tile": t.fft_z_tile, "fft_z_wgs": t.fft_z_wgs}


def run_gpu(wl, precision, steps, warmup, prof_steps=0, tune=True, tuning_file=None, dropin_steps=0):
    """Time `steps` iterations of the time loop after the RK start-up.

    Resident loop (the headline): hetddft_iterate with err = NULL, so the
    error estimates stay on the device; the closing hetddft_sync reduces and
    checks the last step's estimates (HETDDFT_E_NONFINITE on divergence).
    Drop-in loop (dropin_steps > 0): the reference's loop as a Fortran host
    drives it, with the three error estimates and the nine particle means
    fetched to the host after every step (DFT4He3d.f90:855-878)."""
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
    tuning = None
    saved = _load_tuning(tuning_file, precision)
    if saved is not None:                   # a tuning of an earlier run (no trial launches)
        g.set_tuning(saved)
        tuning = g.get_tuning().as_dict()
        tuning["from"] = str(tuning_file)
    elif tune:                              # launch shapes picked by the autotuner (state restored)
        ta = time.perf_counter()
        tuning = g.autotune(wl.deltat, steps=8).as_dict()
        tuning["seconds"] = time.perf_counter() - ta
    tuning_rec = _tuning_record(g.get_tuning())
    for _ in range(warmup):
        g.iterate(wl.deltat, want_errors=False)
    g.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.iterate(wl.deltat, want_errors=False)
    g.sync()
    t1 = time.perf_counter()
    errs = g.last_errors()
    dropin = None
    if dropin_steps:
        t2 = time.perf_counter()
        for _ in range(dropin_steps):
            g.iterate(wl.deltat, want_errors=True)
            g.particle_means()
        t3 = time.perf_counter()
        dropin = dropin_steps / (t3 - t2)
    energy = g.energy()
    den = g.get_field("den")
    prof = None
    if prof_steps:
        g.profile(True)
        for _ in range(prof_steps):
            g.iterate(wl.deltat, want_errors=False)
        prof = g.profile_read()
        g.profile(False)
    out = dict(seconds=t1 - t0, steps_per_s=steps / (t1 - t0), energy=energy, errs=errs, den=den,
               device_bytes=g.device_bytes, prof=prof, tuning=tuning, tuning_rec=tuning_rec,
               dropin_steps_per_s=dropin)
    g.close()
    return out


def run_ensemble(wls, precision, steps, warmup, threads=8, tune=True, executor="library"):
    """Every instance of the ensemble advances one PC step per round; the
    rounds are enqueued without host synchronisation (each instance has its
    own two streams), and the timed region ends when every instance is done.
    Launching is host-bound for instances this small, so `threads` host
    threads each drive a share of the instances: by default the library's
    batch executor (hetddft_batch_iterate, C++ threads, largest grids spread
    first); executor="python" drives them from Python threads instead."""
    import threading
    import hetddft
    import inputs
    kcache = {}
    insts = []
    for wl in wls:
        p = inputs.params_for(wl, precision=precision)
        if wl.n not in kcache:
            kcache[wl.n] = hetddft.make_kernels(p)
        g = hetddft.Instance(p)
        g.set_kernels(kcache[wl.n])
        psi, r, v = inputs.initial_state(wl)
        g.set_state(psi, r, v)
        g.start()
        insts.append(g)
    dts = [wl.deltat for wl in wls]
    for _ in range(3):
        for g, dt in zip(insts, dts):
            g.iterate(dt, want_errors=False)
    ta = time.perf_counter()
    if tune:                                # one autotuner run per grid size (cached by the library)
        for g, dt in zip(insts, dts):
            g.autotune(dt, steps=2)
    tune_s = time.perf_counter() - ta
    for g in insts:
        g.sync()
    if executor == "library":
        b = hetddft.Batch(insts, dts, nthreads=threads)
        b.iterate(warmup)                   # the tuned shapes, steady state
        t0 = time.perf_counter()
        b.iterate(steps)
        t1 = time.perf_counter()
        b.close()
        full = [g.energy() for g in insts]
        energies = [e["etot"] for e in full]
        dev = sum(g.device_bytes for g in insts)
        for g in insts:
            g.close()
        return dict(seconds=t1 - t0, instance_steps_per_s=len(insts) * steps / (t1 - t0), energies=energies,
                    terms=[_term_scale(e) for e in full], device_bytes=dev, tune_seconds=tune_s)
    for _ in range(warmup):
        for g, dt in zip(insts, dts):
            g.iterate(dt, want_errors=False)
    for g in insts:
        g.sync()
    groups = [list(range(t, len(insts), threads)) for t in range(min(threads, len(insts)))]
    errors = []

    def drive(idx):
        try:
            for _ in range(steps):
                for i in idx:
                    insts[i].iterate(dts[i], want_errors=False)
            for i in idx:
                insts[i].sync()
        except Exception as e:  # reported after the join
            errors.append(e)

    workers = [threading.Thread(target=drive, args=(idx,)) for idx in groups]
    t0 = time.perf_counter()
    for w in workers:
        w.start()
    for w in workers:
        w.join()
    t1 = time.perf_counter()
    if errors:
        raise errors[0]
    full = [g.energy() for g in insts]
    energies = [e["etot"] for e in full]
    dev = sum(g.device_bytes for g in insts)
    for g in insts:
        g.close()
    return dict(seconds=t1 - t0, instance_steps_per_s=len(insts) * steps / (t1 - t0), energies=energies,
                terms=[_term_scale(e) for e in full], device_bytes=dev, tune_seconds=tune_s)


def _term_scale(e):
    """Sum of the magnitudes of the energy terms whose sum is etot
    (energy.f90:62-124): the scale against which a difference of totals that
    nearly cancel is measured."""
    return sum(abs(e[k]) for k in ("ekin4", "elj4", "ealphas", "esolid", "ecor4", "ekinx", "ekinqx", "ehex"))


def ensemble_main(args):
    import inputs
    wls = inputs.ensemble_c4(n_pc=args.steps)
    r64 = run_ensemble(wls, 0, args.steps, args.warmup, args.threads, tune=not args.no_tune, executor=args.executor)
    r32 = run_ensemble(wls, 1, args.steps, args.warmup, args.threads, tune=not args.no_tune, executor=args.executor)
    rels = [abs(a - b) / abs(a) for a, b in zip(r64["energies"], r32["energies"])]
    rel = max(rels)
    worst = rels.index(rel)
    abs_err = max(abs(a - b) for a, b in zip(r64["energies"], r32["energies"]))
    rel_terms = max(abs(a - b) / s for a, b, s in zip(r64["energies"], r32["energies"], r64["terms"]))
    sizes = {}
    for wl in wls:
        sizes[f"{wl.n}^3"] = sizes.get(f"{wl.n}^3", 0) + 1
    mb = sum(model_bytes(wl.n, wl.n_par, 1) for wl in wls) / len(wls)
    print(json.dumps({
        "metric": "instance-steps/s, ensemble of 32 uneven instances (mixed vs FP64)",
        "value": r32["instance_steps_per_s"], "unit": "instance-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
        "data": "synthetic (seeded Woods-Saxon droplets, Gaussian particle clouds)",
        "config": {"workload": f"c4: 32 instances {sizes}, dx=0.45 A, N in [17, 300], N_par=2000, "
                               f"dt=0.5 fs, OTC, ironx=3", "precision": "mixed"},
        "fp64_instance_steps_per_s": r64["instance_steps_per_s"],
        "mixed_instance_steps_per_s": r32["instance_steps_per_s"],
        "mixed_vs_fp64_speedup": r32["instance_steps_per_s"] / r64["instance_steps_per_s"],
        "mixed_max_rel_energy_err": rel,
        # the relative error is largest where the total energy nearly cancels
        "mixed_max_rel_energy_err_instance": {"name": wls[worst].name, "n_he": wls[worst].n_he,
                                              "etot_fp64": r64["energies"][worst]},
        "mixed_max_abs_energy_err_K": abs_err,
        # the same differences relative to the sum of the terms' magnitudes
        "mixed_max_energy_err_vs_terms": rel_terms,
        "model_gb_per_instance_step_mixed": mb / 1e9,
        "model_hbm_gbs": mb * r32["instance_steps_per_s"] / 1e9,
        "device_bytes_fp64": r64["device_bytes"], "device_bytes_mixed": r32["device_bytes"],
        "host_threads": args.threads,
        "executor": "hetddft_batch_iterate" if args.executor == "library" else "python threads",
        "autotune_seconds": {"fp64": r64["tune_seconds"], "mixed": r32["tune_seconds"]},
    }))


TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic.json"


def measured_traffic(workload, precision, cls):
    """HBM bytes per launch of a kernel class from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), or None."""
    try:
        rec = json.loads(TRAFFIC_FILE.read_text())
        ent = rec["entries"][f"{workload}/{precision}/{cls}"]
        return float(ent["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def roofline(prof, workload=None, precision=None):
    """Dominant kernel class: algorithmic bytes per launch / mean event time."""
    name, (ms, n, b) = max(prof.items(), key=lambda kv: kv[1][0])
    sec = ms / 1e3 / n
    achieved = b / n / sec / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(workload, precision, name),
            "avg_us": sec * 1e6, "algorithmic_bytes_per_launch": b / n, "launches": n}


def copy_calibration(nbytes=1 << 30, reps=5):
    """Plain device-to-device copy of nbytes, reps times: the reference kernel for
    scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
            "GB_per_s": 2 * nbytes * reps / (t1 - t0) / 1e9}


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workloads, samples=5):
    """The oracle on the host cores: PC steps/s as the median of `samples`
    timed PC steps (steppc + the driver's potential evaluation, as the GPU
    loop counts them) after start() and the 3 RK steps, per workload."""
    import hetddft
    import inputs
    from oracle import hetddft_oracle as orc
    sys.path.insert(0, str(ROOT))
    cores = len(os.sched_getaffinity(0))
    out = {}
    from oracle.threaded import ThreadedOracle
    with ThreadedOracle(orc, cores):
        for wl in workloads:
            p = inputs.params_for(wl)
            k = hetddft.make_kernels(p)
            psi, r, v = inputs.initial_state(wl)
            prm = orc.Params(nx=wl.n, ny=wl.n, nz=wl.n, xmax=wl.xmax, ymax=wl.xmax, zmax=wl.xmax,
                             n_par=wl.n_par, mimpur=wl.mimpur, ironx=wl.ironx)
            o = orc.Oracle(prm, k)
            o.set_state(psi, r, v)
            o.start()
            for it in range(3):
                o.iterate(it, wl.deltat)
            times = []
            for it in range(samples):
                t0 = time.perf_counter()
                o.iterate(3 + it, wl.deltat)
                times.append(time.perf_counter() - t0)
            out[wl.name] = {"steps_per_s": 1.0 / float(np.median(times)),
                            "seconds": [round(t, 4) for t in times]}
    main = workloads[0].name
    return {"value": out[main]["steps_per_s"], "unit": "PC steps/s", "cores": cores, "kind": "port",
            "cpu": _cpu_model(),
            "sample": f"numpy oracle on {cores} threads (transforms and stencils threaded, whole-array "
                      f"expressions single-threaded): median of {samples} PC steps after start + 3 RK "
                      f"steps; value is {main}",
            "per_workload": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes (no autotuner)")
    ap.add_argument("--executor", choices=["library", "python"], default="library",
                    help="ensemble (c4) driver: the library's batch executor or Python threads")
    ap.add_argument("--threads", type=int, default=8, help="host threads driving the ensemble (c4)")
    ap.add_argument("--only", choices=["fp64", "mixed", "copy"], default=None,
                    help="time one precision only, or run the copy calibration (profiling runs)")
    ap.add_argument("--tuning", default=None, help="apply a tuning saved by --save-tuning (no autotuner)")
    ap.add_argument("--save-tuning", default=None, help="write the launch shapes used to this JSON file")
    ap.add_argument("--dropin-steps", type=int, default=100,
                    help="steps of the drop-in loop (errors and particle means fetched every step)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the path has no multi-GPU mode "
              "(an instance is one GPU's share of the work)", file=sys.stderr)
        sys.exit(2)
    import inputs
    if args.workload == "c4":
        ensemble_main(args)
        return
    wl = inputs.WORKLOADS[args.workload]
    if args.only == "copy":
        print(json.dumps(copy_calibration()))
        return
    if args.only:
        res = run_gpu(wl, 0 if args.only == "fp64" else 1, args.steps, args.warmup, prof_steps=5,
                      tune=not args.no_tune, tuning_file=args.tuning)
        kernels = {k: {"avg_us": round(ms * 1e3 / n, 2), "launches": n,
                       "GB_per_s": round(b / (ms * 1e-3) / 1e9, 1)}
                   for k, (ms, n, b) in res["prof"].items() if n}
        print(json.dumps({"only": args.only, "steps_per_s": res["steps_per_s"],
                          "ms_per_step": 1000.0 / res["steps_per_s"], "kernels": kernels}))
        return
    res64 = run_gpu(wl, 0, args.steps, args.warmup, prof_steps=20, tune=not args.no_tune,
                    tuning_file=args.tuning, dropin_steps=args.dropin_steps)
    res32 = run_gpu(wl, 1, args.steps, args.warmup, prof_steps=20, tune=not args.no_tune,
                    tuning_file=args.tuning, dropin_steps=args.dropin_steps)
    if args.save_tuning:
        Path(args.save_tuning).write_text(json.dumps({"fp64": res64["tuning_rec"], "mixed": res32["tuning_rec"]}))
    e64 = res64["energy"]["etot"]
    e32 = res32["energy"]["etot"]
    rel_e = abs(e32 - e64) / abs(e64)
    d64, d32 = res64["den"], res32["den"]
    den_l2 = float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64))
    use_mixed = rel_e <= 1e-6
    best = res32 if use_mixed else res64
    prec = 1 if use_mixed else 0
    gb_step = model_bytes(wl.n, wl.n_par, prec, wl.ironx) / 1e9
    line = {
        "metric": METRIC,
        "value": best["steps_per_s"],
        "unit": "PC steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 / best["steps_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if use_mixed else "f64",
        "data": "synthetic (seeded Woods-Saxon droplet, Gaussian particle cloud)",
        "config": {"workload": f"{wl.name}: {wl.n}^3 grid dx={wl.h} A, N={wl.n_he} He Woods-Saxon droplet "
                               f"with a {wl.cavity} A cavity at the impurity, {wl.phase_noise:g} rad phase "
                               f"noise, N_par={wl.n_par} (sigma {wl.sigma} A), v0={wl.v0 * 100:g} m/s, "
                               f"Morse D=20 K, dt={wl.dt_ps * 1000:.1f} fs, OTC, ironx={wl.ironx}",
                   "precision": "mixed" if use_mixed else "fp64",
                   "loop": "resident: iterate with the error estimates on the device, checked at the end "
                           "of the timed region (hetddft_sync)"},
        "roofline": roofline(best["prof"], wl.name, "mixed" if use_mixed else "fp64"),
        "fp64_steps_per_s": res64["steps_per_s"],
        "mixed_steps_per_s": res32["steps_per_s"],
        # the reference's loop: errors and particle means to the host every step
        "dropin_loop_steps_per_s": {"fp64": res64["dropin_steps_per_s"], "mixed": res32["dropin_steps_per_s"],
                                    "steps": args.dropin_steps},
        "mixed_vs_fp64_speedup": res32["steps_per_s"] / res64["steps_per_s"],
        "mixed_rel_energy_err": rel_e,
        "mixed_density_rel_l2_err": den_l2,
        "model_gb_per_step": gb_step,
        "model_hbm_gbs": gb_step * best["steps_per_s"],
        "model_hbm_frac": gb_step * best["steps_per_s"] / HBM_PEAK_GBS,
        "fp64_model_hbm_frac": model_bytes(wl.n, wl.n_par, 0, wl.ironx) / 1e9 * res64["steps_per_s"] / HBM_PEAK_GBS,
        "fp64_roofline": roofline(res64["prof"], wl.name, "fp64"),
        "tuning": {"mixed": res32["tuning"], "fp64": res64["tuning"]},
        "device_bytes": best["device_bytes"],
    }
    if not args.no_cpu:
        others = [inputs.WORKLOADS[n] for n in ("c1", "c2") if n != wl.name]
        line["cpu_baseline"] = cpu_baseline([wl] + others)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
