"""Benchmark: cell-updates/s of the Fluid2d timestep path on one MI355X.

Headline (``value``): the plain FP64 path on the Euler 1024x1024 closed
free-slip box (UP5 parabolic advection, RK3_SSP, fixed dt at CFL 0.8 of the
initial maxspeed, multigrid Poisson: 2 V-cycles per inverted stage + the
end-of-step F-cycle solve with maxite 4 / tol 1e-11, k^-3 turbulence init).
A "step" is one RK3_SSP step of that grid including every multigrid cycle
and the per-step diagnostics, replayed from a hipGraph; inputs are resident
in HBM before timing.

Further legs, reported as extra keys of the same JSON line:

* ``roofline``  the kernel with the largest share of the plain step, its
  algorithmic bytes per launch / mean launch time (HIP events on the
  engine's stream) against 8 TB/s;
* ``batched``   256-instance ensemble (sizes 64..512, periodic/closed) in
  mixed precision, one graph per step for all instances;
* ``precision`` rel-L2 of mixed precision against FP64 (the FP64 engine is
  bit-identical to the oracle) on the 1024^2 box, the 2048x1024 Boussinesq
  Rayleigh-Taylor run and the ensemble;
* ``serving``   Poisson arrivals (50/s) of Euler/Boussinesq runs of uneven
  size, continuous batching, p50/p99 time-to-result;
* ``cpu_baseline`` the C oracle (FP64) with its row loops under OpenMP on
  the host's cores, and ``cpu_baseline_1thread`` the same on one thread,
  each on a bounded sample of the headline workload.

--gpus N > 1 is refused: the library runs on one GPU.
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
HEADLINE = "euler-closed-1024"
BOUSSINESQ = "boussinesq-rt-2048x1024"
KERNEL_NAMES = {"smooth": "k_smooth", "resid_restrict": "k_resid_restrict", "rhs": "k_rhs",
                "coarse": "k_coarse", "persist": "k_persist", "smooth_rr": "k_smooth_rr",
                "up2": "k_up2"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--legs", default="plain,roofline,batched,seqplain,precision,serving,cpu",
                    help="comma list of legs to run")
    ap.add_argument("--ensemble", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--bous-steps", type=int, default=1000)
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch configuration instead of the autotuned one")
    ap.add_argument("--tuning", default=None,
                    help="JSON launch configuration to use instead of running the autotuner")
    ap.add_argument("--workload", default=HEADLINE,
                    help="workload of the plain leg; other than %s only with --legs plain or copy,plain "
                         "(kernel traces of the other configurations)" % HEADLINE)
    ap.add_argument("--copy-mib", type=int, default=256,
                    help="array size of the copy leg's calibration passes (MiB)")
    ap.add_argument("--eager", action="store_true",
                    help="launch the step kernel by kernel instead of replaying a hipGraph "
                         "(for kernel-trace profilers)")
    return ap.parse_args()


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def max_rel(sa, sb):
    return max(rel(sa[k], sb[k]) for k in range(sa.shape[0]))


USE_GRAPH = True


FIXED_TUNING = None


def tuned(pg, precision, enabled):
    """Launch configuration from the autotuner (timed on a scratch batch,
    outside every timed region), or the defaults."""
    if not enabled:
        return None
    if FIXED_TUNING is not None:
        return FIXED_TUNING
    from fluid2d_amd import autotune
    best, _ = autotune.tune(pg, precision, use_cache=False)
    return best


def config_batch(name, precision, dt=None, tune=False, fast_solve=None):
    from fluid2d_amd import workloads
    from fluid2d_amd.batch import Batch
    from fluid2d_amd.grid import Grid

    p = workloads.make_param(name)
    g = Grid(p)
    b = Batch([(p, g)], precision=precision, use_graph=USE_GRAPH,
              tuning=tuned([(p, g)], precision, tune))
    if fast_solve:
        b.set_fast_solve(*fast_solve)
    for k, v in workloads.initial_fields(name, g).items():
        b.set_field(0, k, v)
    b.set_psi_from_vorticity()
    if dt is None:
        dt = workloads.time_step(name, g, b.diagnostics(0)["maxspeed"])
    b.set_dt(dt)
    return b, g, dt


def ensemble_batch(n, precision, dts=None, tune=False, perturb=0.0):
    """The ensemble as one batch; `perturb` > 0 multiplies every initial
    vorticity value by (1 + perturb * r), r seeded uniform in [-1, 1]."""
    from fluid2d_amd import workloads
    from fluid2d_amd.batch import Batch
    from fluid2d_amd.grid import Grid
    from fluid2d_amd.param import Param

    specs = workloads.ensemble_specs(n)
    pg = []
    for kw, _ in specs:
        p = Param(**kw)
        pg.append((p, Grid(p)))
    b = Batch(pg, precision=precision, use_graph=USE_GRAPH, tuning=tuned(pg, precision, tune))
    rng = np.random.default_rng(12345)
    for i, (_, seed) in enumerate(specs):
        w = workloads.turbulence(b.grids[i], seed)
        if perturb > 0:
            w = w * (1.0 + perturb * rng.uniform(-1.0, 1.0, w.shape))
        b.set_field(i, "vorticity", w)
    b.set_psi_from_vorticity()
    if dts is None:
        dts = [0.8 * min(g.dx, g.dy) / b.diagnostics(i)["maxspeed"] for i, g in enumerate(b.grids)]
    for i, dt in enumerate(dts):
        b.set_dt(dt, i)
    cells = sum(g.nx * g.ny for g in b.grids)
    return b, dts, cells


def timed(b, steps):
    import fluid2d_amd as F
    b.synchronize()
    F.load().f2d_synchronize()
    t0 = time.perf_counter()
    b.step_async(steps)
    b.synchronize()
    F.load().f2d_synchronize()
    return time.perf_counter() - t0


def roofline(b, nsteps=2):
    """Dominant HBM-streaming kernel of one step: of the tiled kinds
    (smoother, residual/restriction, advection), the one with the largest
    event-timed share; its finest-level launches give the roofline figure.
    The one-workgroup coarse kernel is latency-bound, not HBM-bound: its
    share is reported beside the others."""
    shares = {}
    for kind in ("smooth", "resid_restrict", "rhs", "coarse", "persist", "smooth_rr", "up2"):
        pr = b.profile(kind, -1, nsteps)
        shares[kind] = pr["avg_ms"] * pr["launches"] / nsteps
    dom = max(("smooth", "resid_restrict", "rhs"), key=shares.get)
    pr = b.profile(dom, 0, nsteps)
    achieved = pr["alg_bytes"] / (pr["avg_ms"] * 1e-3) / 1e9 if pr["avg_ms"] > 0 else 0.0
    return dom, shares, pr, achieved


def cpu_baseline(name, seconds, variant="serial"):
    """The C oracle on the host cores: "serial" (1 thread) or "omp" (its row
    loops under OpenMP with OMP_NUM_THREADS threads, the same results)."""
    from fluid2d_amd import workloads
    from oracle import oracle as orc

    orc.use(variant)
    cores = 1
    if variant == "omp":
        cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(cores)
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
    t0 = time.perf_counter()
    o.run(1, dt)
    one = time.perf_counter() - t0
    steps = max(1, min(200, int(seconds / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.run(steps, dt)
    el = time.perf_counter() - t0
    orc.use("serial")
    return {"value": g.nx * g.ny * steps / el, "unit": "cell-updates/s", "cores": cores,
            "kind": "port", "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": "%s: %d RK3_SSP steps (after 1 untimed) of the C oracle, FP64, %s, "
                      "gcc -O3 -ffp-contract=off, %.1f s" % (
                          name, steps, "1 thread" if variant == "serial" else
                          "%d OpenMP threads over grid rows" % cores, el)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def finest_tile_height(nx, ny, min_tiles):
    """Tile height of the finest-level smoother launches: the tallest of
    16, 8, 4 rows (64 columns) giving at least min_tiles tiles."""
    if min_tiles <= 0:
        return 16
    for ty in (16, 8):
        if nx * ny / (64.0 * ty) >= min_tiles:
            return ty
    return 4


def load_traffic(kernel, workload, ty):
    """HBM bytes per launch of the finest-level launches of `kernel` with
    tile height `ty`, from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(REPO, "profiles", "traffic_%s_%s.json" % (kernel, workload))
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    # the kernel's name as the trace shows it: k_smooth<T, rows>,
    # k_resid_restrict<T, R, coarse rows>
    if kernel == "k_resid_restrict":
        key = "%s<double, double, %d>" % (kernel, ty // 2)
    else:
        key = "%s<double, %d>" % (kernel, ty)
    if key not in t:
        return None, os.path.relpath(path, REPO) + " (no entry for %s)" % key
    return t[key].get("bytes_per_launch"), os.path.relpath(path, REPO) + " [%s]" % key


def sequential_plain(n, tune, steps=20):
    """The ensemble run the plain way: one FP64 instance at a time.  Each
    (size, geometry) class is timed once (graph-replayed steps) and the
    ensemble's time is the sum over its instances."""
    from fluid2d_amd import workloads
    from fluid2d_amd.batch import Batch
    from fluid2d_amd.grid import Grid
    from fluid2d_amd.param import Param

    specs = workloads.ensemble_specs(n)
    classes = {}
    for kw, seed in specs:
        classes.setdefault((kw["nx"], kw["geometry"]), []).append((kw, seed))
    total_t, total_cells = 0.0, 0
    per = {}
    for key, members in sorted(classes.items()):
        kw, seed = members[0]
        p = Param(**kw)
        g = Grid(p)
        b = Batch([(p, g)], precision="fp64", use_graph=USE_GRAPH,
                  tuning=tuned([(p, g)], "fp64", tune))
        b.set_field(0, "vorticity", workloads.turbulence(g, seed))
        b.set_psi_from_vorticity()
        b.set_dt(0.8 * min(g.dx, g.dy) / b.diagnostics(0)["maxspeed"])
        b.step(2)
        el = timed(b, steps) / steps
        per["%dx%d-%s" % (key[0], key[0], key[1])] = el * 1e3
        total_t += el * len(members)
        total_cells += g.nx * g.ny * len(members)
        del b
    return total_cells / total_t, per


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: the library runs on one GPU; --gpus %d is not supported "
              "(instances shard across GPUs with no exchange; not built here)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import fluid2d_amd as F

    if F.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    legs = set(a.legs.split(","))
    if "copy" in legs:
        # calibration for the HBM byte counters: a plain device copy of a
        # known size (the runtime's copy kernel), timed with the same clock
        from fluid2d_amd.device import DeviceArray
        nbytes = a.copy_mib << 20
        src = DeviceArray((nbytes // 8,), np.float64)
        dst = DeviceArray((nbytes // 8,), np.float64)
        for _ in range(5):
            dst.copy_from(src)
        F.load().f2d_synchronize()
        log("copy: 5 x %d bytes" % nbytes)
        del src, dst
        # the same volume through a streaming stencil kernel of this library
        # (celltocorner: each cell read once from HBM, its neighbours from
        # cache, one value written), whose loads are plain global loads like
        # the solver's
        from fluid2d_amd import kernels
        m, n = 4096, nbytes // 8 // 4096
        xs = DeviceArray((m, n), np.float64)
        xd = DeviceArray((m, n), np.float64)
        for _ in range(5):
            kernels.celltocorner(xs, xd)
        F.load().f2d_synchronize()
        log("stream copy (celltocorner): 5 x %d bytes" % nbytes)
        del xs, xd
    global USE_GRAPH, FIXED_TUNING
    USE_GRAPH = not a.eager
    if a.tuning:
        FIXED_TUNING = json.loads(a.tuning)
    from fluid2d_amd import workloads
    name = a.workload
    if name not in workloads.CONFIGS:
        raise SystemExit("bench.py: unknown workload %s (one of %s)"
                         % (name, ", ".join(sorted(workloads.CONFIGS))))
    if name != HEADLINE and not legs <= {"plain", "copy"}:
        raise SystemExit("bench.py: --workload %s runs with --legs plain (and copy) only" % name)
    wkw = workloads.CONFIGS[name][0]
    out = {
        "metric": "cell-updates/s incl. MG Poisson, batched vs plain @ rel-L2<=1e-6; p99 latency",
        "value": None, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded k^-3 turbulence, vortex pair, Rayleigh-Taylor inits)",
        "config": {"workload": name, "path": "plain FP64, one instance, graph-replayed step",
                   "nx": wkw["nx"], "ny": wkw["ny"], "model": wkw["modelname"],
                   "geometry": wkw["geometry"],
                   "scheme": "RK3_SSP", "advection": "UP5 parabolic",
                   "mg": ("2 V-cycles per inverted stage + F-cycle solve (maxite 4, tol 1e-11)"
                          if wkw["modelname"] == "euler" else "2 V-cycles per inverted stage")},
        "roofline": None, "cpu_baseline": None,
    }
    plain_state = None
    b = None
    if legs & {"plain", "roofline", "precision"}:
        log("plain FP64 %s: warmup %d, %d steps" % (name, a.warmup, a.steps))
        b, g, dt = config_batch(name, "fp64", tune=not a.no_tune)
        out["config"]["tuning"] = b.tuning()
        out["config"]["launches_per_step"] = b.info(0)["launches_per_step"]
        b.step(a.warmup)
        el = timed(b, a.steps)
        cells = g.nx * g.ny
        out["value"] = cells * a.steps / el
        out["ms_per_step"] = el / a.steps * 1e3
        if wkw["modelname"] == "euler":
            # words per cell-update: 109 + 28 per F-cycle of the full solve
            nite = b.solve_info(0)[0]
            out["config"]["F_cycles_last_step"] = nite
            wpc = 109 + 28 * float(nite)
        else:
            wpc = 162.0  # Boussinesq: fast inversions only, no full solve
        out["step_roofline"] = {
            "words_per_cell_update": wpc,
            "achieved": out["value"] * wpc * 8 / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": out["value"] * wpc * 8 / 1e9 / PEAK_HBM_GBS}
        plain_state = b.state(0)
        log("plain: %.3f ms/step, %.3e cell-updates/s" % (out["ms_per_step"], out["value"]))
    if "roofline" in legs and b is not None:
        dom, shares, pr, achieved = roofline(b)
        kname = KERNEL_NAMES[dom] + "<double>"
        tun = out["config"].get("tuning") or {}
        ty = finest_tile_height(out["config"]["nx"], out["config"]["ny"],
                                int(tun.get("min_tiles", 0)))
        traffic, src = load_traffic(KERNEL_NAMES[dom], name, ty)
        out["roofline"] = {
            "bound": "hbm",
            "kernel": kname + " finest-level launches",
            "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
            "alg_bytes_per_launch": pr["alg_bytes"], "avg_launch_ms": pr["avg_ms"],
            "launches_timed": pr["launches"],
            "step_share_ms": {k: round(v, 4) for k, v in shares.items()}}
        if src:
            out["roofline"]["traffic_source"] = src
        log("roofline: %s %.1f GB/s (%.1f%%)" % (kname, achieved, 100 * achieved / PEAK_HBM_GBS))
    del b
    if "batched" in legs or "precision" in legs:
        log("ensemble %d instances mixed, %d steps" % (a.ensemble, a.ensemble_steps))
        eb, dts, ecells = ensemble_batch(a.ensemble, "mixed", tune=not a.no_tune)
        eb.step(2)
        el = timed(eb, a.ensemble_steps - 2)
        bval = ecells * (a.ensemble_steps - 2) / el
        out["batched"] = {"value": bval, "unit": "cell-updates/s", "dtype": "f32+f64",
                          "workload": "ensemble-%d-euler-64..512" % a.ensemble,
                          "instances": a.ensemble, "cells": ecells,
                          "steps": a.ensemble_steps, "ms_per_step": el / (a.ensemble_steps - 2) * 1e3,
                          "speedup_vs_plain_fp64": (bval / out["value"]) if out["value"] else None,
                          "speedup_note": "ensemble of 64^2..512^2 grids against one 1024^2 grid: "
                                          "a batching gain, not a precision gain; the like-for-like "
                                          "ratios are speedup_vs_batched_fp64 and "
                                          "speedup_vs_plain_fp64_same_workload",
                          "tuning": eb.tuning()}
        log("batched mixed: %.3e cell-updates/s" % bval)
        if "seqplain" in legs:
            sval, per = sequential_plain(a.ensemble, not a.no_tune)
            out["batched"]["plain_fp64_same_workload"] = {
                "value": sval, "unit": "cell-updates/s", "ms_per_step_by_class": per,
                "method": "each (size, geometry) class timed alone in FP64, summed over the "
                          "ensemble's instances"}
            out["batched"]["speedup_vs_plain_fp64_same_workload"] = bval / sval
            log("plain FP64, same ensemble one instance at a time: %.3e cell-updates/s" % sval)
        if "precision" in legs:
            mixed_states = [eb.state(i) for i in range(eb.n)]
            del eb
            log("ensemble FP64, %d steps" % a.ensemble_steps)
            fb, _, _ = ensemble_batch(a.ensemble, "fp64", dts, tune=not a.no_tune)
            fb.step(2)
            el64 = timed(fb, a.ensemble_steps - 2)
            out["batched"]["fp64_value"] = ecells * (a.ensemble_steps - 2) / el64
            out["batched"]["speedup_vs_batched_fp64"] = bval / out["batched"]["fp64_value"]
            f64_states = [fb.state(i) for i in range(fb.n)]
            sizes = [g.nx for g in fb.grids]
            del fb
            errs = [max_rel(mixed_states[i], f64_states[i]) for i in range(len(sizes))]
            out.setdefault("precision", {})["ensemble_max_rel_l2"] = max(errs)
            out["precision"]["ensemble_steps"] = a.ensemble_steps
            out["precision"]["ensemble_max_rel_l2_by_size"] = {
                str(nx): max(e for e, s_ in zip(errs, sizes) if s_ == nx) for nx in sorted(set(sizes))}
            out["precision"]["ensemble_meets_1e-6_by_size"] = {
                k: v <= 1e-6 for k, v in out["precision"]["ensemble_max_rel_l2_by_size"].items()}
            out["batched"]["meets_1e-6"] = max(errs) <= 1e-6
            # the flows' own sensitivity: FP64 from initial vorticity perturbed
            # by half an FP32 ulp (relative 2^-24), against unperturbed FP64
            log("ensemble FP64 from initial vorticity perturbed by 2^-24, %d steps" %
                a.ensemble_steps)
            pb, _, _ = ensemble_batch(a.ensemble, "fp64", dts, tune=not a.no_tune,
                                      perturb=2.0 ** -24)
            pb.step(a.ensemble_steps)
            perrs = [max_rel(pb.state(i), f64_states[i]) for i in range(len(sizes))]
            del pb
            out["precision"]["ensemble_fp64_perturbed_2^-24_max_rel_l2_by_size"] = {
                str(nx): max(e for e, s_ in zip(perrs, sizes) if s_ == nx) for nx in sorted(set(sizes))}
        else:
            del eb
    if "precision" in legs and plain_state is not None:
        # the configuration's own step count (1000), FP64 and mixed from the
        # same start with the same dt
        nprec = workloads.CONFIGS[name][2]
        log("%s FP64 + mixed, %d steps" % (name, nprec))
        fb, g, dt = config_batch(name, "fp64", tune=not a.no_tune)
        el64 = timed(fb, nprec)
        s64 = fb.state(0)
        del fb
        mb, g, _ = config_batch(name, "mixed", dt, tune=not a.no_tune)
        el = timed(mb, nprec)
        out.setdefault("precision", {})
        err = max_rel(mb.state(0), s64)
        out["precision"]["euler_closed_1024_rel_l2"] = err
        out["precision"]["euler_closed_1024_steps"] = nprec
        out["precision"]["euler_closed_1024_fp64_value"] = g.nx * g.ny * nprec / el64
        out["precision"]["euler_closed_1024_mixed_value"] = g.nx * g.ny * nprec / el
        out["precision"]["euler_closed_1024_mixed_vs_fp64"] = el64 / el
        out["precision"]["euler_closed_1024_meets_1e-6"] = err <= 1e-6
        del mb
        log("Boussinesq %s fp64 + mixed, %d steps" % (BOUSSINESQ, a.bous_steps))
        bb, g, dt = config_batch(BOUSSINESQ, "fp64", tune=not a.no_tune)
        el64 = timed(bb, a.bous_steps)
        s64 = bb.state(0)
        del bb
        bm, _, _ = config_batch(BOUSSINESQ, "mixed", dt, tune=not a.no_tune)
        elm = timed(bm, a.bous_steps)
        out["precision"]["boussinesq_rt_rel_l2"] = max_rel(bm.state(0), s64)
        out["precision"]["boussinesq_rt_steps"] = a.bous_steps
        out["precision"]["boussinesq_rt_fp64_value"] = g.nx * g.ny * a.bous_steps / el64
        out["precision"]["boussinesq_rt_mixed_value"] = g.nx * g.ny * a.bous_steps / elm
        out["precision"]["boussinesq_rt_meets_1e-6"] = out["precision"]["boussinesq_rt_rel_l2"] <= 1e-6
        del bm
        # FP64 residual refinement of the per-stage inversions (gmg.solve to
        # 1e-10, at most 6 F-cycles, instead of twoVcycle): FP32 F-cycles
        # between FP64 residuals in mixed precision, against FP64 in the
        # same mode
        log("Boussinesq %s refined inversions (tol 1e-10), fp64 + mixed, %d steps"
            % (BOUSSINESQ, a.bous_steps))
        ref = {}
        for prec in ("fp64", "mixed"):
            rb, g, _ = config_batch(BOUSSINESQ, prec, dt, tune=not a.no_tune, fast_solve=(6, 1e-10))
            ref[prec] = (timed(rb, a.bous_steps), rb.state(0), rb.solve_info(0))
            del rb
        e = max_rel(ref["mixed"][1], ref["fp64"][1])
        out["precision"]["boussinesq_rt_refined"] = {
            "steps": a.bous_steps, "refinement": "gmg.solve per stage, tol 1e-10, maxite 6",
            "rel_l2": e, "meets_1e-6": e <= 1e-6,
            "fp64_value": g.nx * g.ny * a.bous_steps / ref["fp64"][0],
            "mixed_value": g.nx * g.ny * a.bous_steps / ref["mixed"][0],
            "mixed_vs_fp64": ref["fp64"][0] / ref["mixed"][0],
            "last_solve_fp64": {"F_cycles": ref["fp64"][2][0], "res": ref["fp64"][2][1]},
            "last_solve_mixed": {"F_cycles": ref["mixed"][2][0], "res": ref["mixed"][2][1]}}
        out["precision"]["reference"] = "FP64 engine (bit-identical to the oracle)"
    if "serving" in legs:
        from fluid2d_amd import serving
        reqs = serving.stream_requests(a.serve_requests)
        total = sum(r.est_bytes("mixed") for r in reqs)
        # a budget of about half the stream's footprint binds: requests wait
        # for slots and memory; the same stream with room for everything
        # resident at once is the baseline of the latencies
        for key, budget in (("serving", max(total // 2, max(r.est_bytes("mixed") for r in reqs))),
                            ("serving_unbounded", 64 << 30)):
            log("%s: %d requests, budget %.2f GB (stream footprint %.2f GB)"
                % (key, a.serve_requests, budget / 1e9, total / 1e9))
            srv = serving.Server(hbm_budget=budget, precision="mixed", chunk=20)
            summary, _ = srv.run(serving.stream_requests(a.serve_requests))
            summary["arrival_rate_per_s"] = 50.0
            summary["stream_footprint_bytes"] = total
            out[key] = summary
            log("%s: p50 %.3f s p99 %.3f s (queue p50 %.3f / p99 %.3f, service p50 %.3f / p99 %.3f),"
                " peak %.2f GB, %d slot re-uses"
                % (key, summary["p50_s"], summary["p99_s"], summary["queue_p50_s"],
                   summary["queue_p99_s"], summary["service_p50_s"], summary["service_p99_s"],
                   summary["peak_resident_bytes"] / 1e9, summary["slot_reuses"]))
    if "cpu" in legs:
        log("cpu baseline, all cores (~%.0f s)" % a.cpu_seconds)
        out["cpu_baseline"] = cpu_baseline(name, a.cpu_seconds, "omp")
        log("cpu baseline, 1 thread (~%.0f s)" % a.cpu_seconds)
        out["cpu_baseline_1thread"] = cpu_baseline(name, a.cpu_seconds, "serial")
        # the other two single-grid configurations, shorter samples
        other = {}
        for cname in ("euler-perio-128", BOUSSINESQ):
            other[cname] = {}
            for variant in ("omp", "serial"):
                log("cpu baseline %s, %s (~%.0f s)" % (cname, variant, a.cpu_seconds / 3))
                other[cname][variant] = cpu_baseline(cname, a.cpu_seconds / 3, variant)
        out["cpu_baseline_other_configs"] = other
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
