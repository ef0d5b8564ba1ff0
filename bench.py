"""Benchmark: cell-updates/s of the Fluid2d timestep path on one MI355X.

Workload (default): Euler (vorticity-streamfunction), 1024x1024 closed
free-slip box, FP64, UP5 parabolic advection, RK3_SSP, fixed dt at CFL 0.8
of the initial maxspeed, multigrid Poisson (2 V-cycles per inverted stage +
the end-of-step F-cycle solve, maxite 4, tol 1e-11), k^-3 turbulence init.
A "step" is one RK3_SSP step of that grid including every multigrid cycle
and the per-step diagnostics; inputs are resident in HBM before timing.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="euler-closed-1024")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="oracle steps for the CPU baseline (0: auto, ~15 s)")
    return ap.parse_args()


def words_per_cell_update(model, n_f):
    """Algorithmic FP64 words per interior cell per step (each fused pass
    reads its inputs and writes its outputs once; multigrid operator
    matrix-free; masks and halos ignored): Euler RK3_SSP 109 + 28 n_F,
    Boussinesq RK3_SSP 162."""
    return 109 + 28 * n_f if model == "euler" else 162


def cpu_baseline(name, steps_hint):
    from fluid2d_amd import workloads
    from oracle import oracle as orc

    kw, _, _ = workloads.CONFIGS[name]
    g = workloads.build_grid(name)
    o = orc.OracleModel(kw["modelname"], g.nx, g.ny, g.Lx, g.Ly, geometry=g.geometry,
                        diffusion=kw.get("diffusion", False), Kdiff=kw.get("Kdiff", 0.0),
                        gravity=kw.get("gravity", 1.0))
    for k, v in workloads.initial_fields(name, g).items():
        o.get(k)[...] = v
    o.set_psi_from_vorticity()
    d = o.diagnostics()
    dt = workloads.time_step(name, g, d["maxspeed"])
    steps = steps_hint
    if steps <= 0:
        t0 = time.perf_counter()
        o.run(1, dt)
        one = time.perf_counter() - t0
        steps = max(1, min(200, int(15.0 / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.run(steps, dt)
    el = time.perf_counter() - t0
    return {"value": g.nx * g.ny * steps / el, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": "%s, %d RK3_SSP steps of the C oracle (FP64, 1 thread, -O2, "
                      "no FMA) after the initial solve, %.1f s" % (name, steps, el)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import fluid2d_amd as F
    from fluid2d_amd import workloads

    if F.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    name = a.workload
    kw, _, _ = workloads.CONFIGS[name]
    p = workloads.make_param(name)
    g = F.Grid(p)
    f = F.Fluid2d(p, g)
    for k, v in workloads.initial_fields(name, g).items():
        f.model.var.get(k)[...] = v
    f.model.set_psi_from_vorticity()
    d = f.model.diagnostics(f.model.var, 0.0)
    dt = workloads.time_step(name, g, d["maxspeed"])
    F.load().f2d_synchronize()
    if a.warmup:
        f.model.run(a.warmup, dt)
    F.load().f2d_synchronize()
    t0 = time.perf_counter()
    nlog = f.model.run(a.steps, dt)
    F.load().f2d_synchronize()
    el = time.perf_counter() - t0
    cells = g.nx * g.ny
    value = cells * a.steps / el
    n_f = float(np.mean(nlog)) if kw["modelname"] == "euler" else 0.0
    wpc = words_per_cell_update(kw["modelname"], n_f)
    step_bytes = wpc * 8.0 * cells
    achieved = step_bytes / (el / a.steps) / 1e9
    out = {
        "metric": "cell-updates/s incl. MG Poisson, batched vs plain @ rel-L2<=1e-6; p99 latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded k^-3 turbulence init)",
        "config": {"workload": name, "path": "plain-fp64", "nx": g.nx, "ny": g.ny,
                   "model": kw["modelname"], "geometry": g.geometry, "scheme": "RK3_SSP",
                   "advection": "UP5 parabolic", "mg": "2 V-cycles/stage + F-cycle solve",
                   "mean_F_cycles": n_f},
        "roofline": {"bound": "hbm", "kernel": "whole step (algorithmic words %d/cell)" % wpc,
                     "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None},
        "cpu_baseline": None,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(name, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
