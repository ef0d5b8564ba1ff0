"""Serving loop: continuous batching of Fluid2d runs on one GPU.

Requests (model, grid, step count, initial condition) arrive over time.

* Admission: pending requests (FIFO) are admitted as long as the estimated
  HBM footprint of all resident requests stays within the budget; the
  requests admitted together form a cohort, one engine :class:`Batch` with
  its own stream, so cohorts run concurrently on the device.
* Each request keeps its own step budget on the device (a gate kernel at
  the top of every step), so requests of one cohort finish independently.
* Time-to-result = arrival -> final state on the host.  A cohort's memory is
  released when all its requests have finished, which admits more work.

Every request performs exactly the operations it would perform alone (the
engine's instances are independent), so its result is the same alone and in
the mix; ``tests/test_gpu_serving.py`` checks that bit for bit.
"""
import collections
import time

import numpy as np

from . import workloads
from .batch import Batch
from .grid import Grid
from .param import Param


class Request:
    def __init__(self, rid, arrival, modelname, n, steps, seed, geometry=None):
        self.id = rid
        self.arrival = float(arrival)
        self.modelname = modelname
        self.n = int(n)
        self.steps = int(steps)
        self.seed = int(seed)
        self.geometry = geometry or ("perio" if modelname == "euler" else "xchannel")

    def param(self):
        kw = dict(modelname=self.modelname, nx=self.n, ny=self.n, Lx=1.0, Ly=1.0,
                  geometry=self.geometry, adaptable_dt=False, order=5,
                  flux_splitting_method="parabolic", timestepping="RK3_SSP")
        if self.modelname == "boussinesq":
            kw.update(diffusion=True, Kdiff=1e-5, gravity=1.0)
        return Param(**kw)

    def est_bytes(self, precision):
        """Device memory of one instance (state, tendencies, multigrid)."""
        cells = (self.n + 6) ** 2
        nvar = 5 if self.modelname == "euler" else 6
        tz = 8 if precision == "fp64" else 4
        pw = nvar + 8 + (0 if precision == "fp64" else 3)  # FP64 pool words per cell
        return int(1.15 * cells * (8 * pw + tz * (3 * nvar + 3 + 11) + 5))


def stream_requests(count, rate=50.0, seed=5, sizes_log2=(7, 11), steps=(50, 2000)):
    """Poisson arrivals (rate per second), model 50/50 Euler/Boussinesq,
    N = 2^U{lo..hi}, steps U{steps}, seeded vortex-pair (Euler) or
    Rayleigh-Taylor (Boussinesq) initial conditions."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        model = "euler" if rng.random() < 0.5 else "boussinesq"
        n = 2 ** int(rng.integers(sizes_log2[0], sizes_log2[1] + 1))
        st = int(rng.integers(steps[0], steps[1] + 1))
        out.append(Request(i, t, model, n, st, seed=100 + i))
    return out


def initial_fields(req, grid):
    if req.modelname == "euler":
        return {"vorticity": workloads.vortex_pair(grid, req.seed)}
    return {"buoyancy": workloads.rayleigh_taylor(grid, req.seed)}


def time_step(req, grid, maxspeed0):
    u = maxspeed0 if req.modelname == "euler" else 1.0
    return 0.8 * min(grid.dx, grid.dy) / u


class Cohort:
    def __init__(self, reqs, precision, use_graph=True, tuning=None):
        self.reqs = list(reqs)
        pg = []
        for r in self.reqs:
            p = r.param()
            pg.append((p, Grid(p)))
        self.grids = [g for _, g in pg]
        self.batch = Batch(pg, precision=precision, use_graph=use_graph, tuning=tuning)
        for i, r in enumerate(self.reqs):
            for k, v in initial_fields(r, self.grids[i]).items():
                self.batch.set_field(i, k, v)
        self.batch.set_psi_from_vorticity()
        for i, r in enumerate(self.reqs):
            ms = self.batch.diagnostics(i)["maxspeed"]
            self.batch.set_dt(time_step(r, self.grids[i], ms), i)
            self.batch.set_steps(r.steps, i)
        self.finished = set()
        self.bytes = sum(r.est_bytes(precision) for r in self.reqs)
        self.max_steps = max(r.steps for r in self.reqs)
        self.enqueued = 0

    def advance(self, chunk):
        k = min(chunk, self.max_steps - self.enqueued)
        if k > 0:
            self.batch.step_async(k)
            self.enqueued += k

    def collect(self, now, results, keep_state):
        for i, r in enumerate(self.reqs):
            if i in self.finished:
                continue
            left, done = self.batch.progress(i)
            if left == 0:
                res = {"id": r.id, "arrival": r.arrival, "finish": now,
                       "latency": now - r.arrival, "cells": r.n * r.n, "steps": r.steps,
                       "model": r.modelname, "n": r.n, "diag": self.batch.diagnostics(i)}
                if keep_state:
                    res["state"] = self.batch.state(i)
                results[r.id] = res
                self.finished.add(i)
        return len(self.finished) == len(self.reqs)


class Server:
    """Continuous-batching request loop on one GPU."""

    def __init__(self, hbm_budget=64 << 30, precision="mixed", chunk=20, max_cohort=64,
                 use_graph=True, keep_state=False, tuning=None):
        self.tuning = tuning  # launch shapes of every cohort's batch (None: defaults)
        self.hbm_budget = int(hbm_budget)
        self.precision = precision
        self.chunk = int(chunk)
        self.max_cohort = int(max_cohort)
        self.use_graph = use_graph
        self.keep_state = keep_state

    def run(self, requests):
        todo = collections.deque(sorted(requests, key=lambda r: r.arrival))
        pending = collections.deque()
        cohorts = []
        used = 0
        results = {}
        peak = 0
        t0 = time.perf_counter()
        while todo or pending or cohorts:
            now = time.perf_counter() - t0
            while todo and todo[0].arrival <= now:
                pending.append(todo.popleft())
            admit = []
            while pending and len(admit) < self.max_cohort:
                need = pending[0].est_bytes(self.precision)
                if used + need > self.hbm_budget and (cohorts or admit):
                    break  # wait until a cohort releases memory
                admit.append(pending.popleft())
                used += need
            if admit:
                c = Cohort(admit, self.precision, self.use_graph, self.tuning)
                cohorts.append(c)
                c.advance(self.chunk)
                peak = max(peak, used)
            busy = False
            for c in list(cohorts):
                if not c.batch.done():
                    busy = True
                    continue
                now = time.perf_counter() - t0
                if c.collect(now, results, self.keep_state) or c.enqueued >= c.max_steps:
                    c.collect(now, results, self.keep_state)
                    cohorts.remove(c)
                    used -= c.bytes
                    del c
                else:
                    c.advance(self.chunk)
                    busy = True
            if not admit and not busy:
                if todo and not pending and not cohorts:
                    time.sleep(max(0.0, min(0.002, todo[0].arrival - (time.perf_counter() - t0))))
                else:
                    time.sleep(0.0002)
            elif busy:
                time.sleep(0.0002)
        wall = time.perf_counter() - t0
        lat = np.array([results[k]["latency"] for k in sorted(results)])
        work = sum(r["cells"] * r["steps"] for r in results.values())
        summary = {"requests": len(results), "wall_s": wall,
                   "p50_s": float(np.percentile(lat, 50)) if len(lat) else 0.0,
                   "p99_s": float(np.percentile(lat, 99)) if len(lat) else 0.0,
                   "cell_updates_per_s": work / wall if wall > 0 else 0.0,
                   "cell_updates": work, "peak_resident_bytes": peak,
                   "hbm_budget": self.hbm_budget, "precision": self.precision}
        return summary, results
