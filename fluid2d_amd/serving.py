"""Serving loop: continuous batching of Fluid2d runs on one GPU.

Requests (model, grid, step count, initial condition) arrive over time.

* Requests are grouped by class (model, grid size).  A class's requests run
  in pools: engine batches of slots (:class:`Batch`), each on its own
  stream, so pools run concurrently on the device.
* Admission into a running pool: when a slot's request finishes, a waiting
  request of the same class takes the slot between two chunks of steps
  (the instance is reset, initialised and given its own dt and step count
  on the device) while the pool's other requests keep running.
* A new pool is created for a class with waiting requests and no free slot
  while the estimated HBM footprint of all pools fits the budget; a pool
  with no resident request and nothing waiting is destroyed, which returns
  its memory.
* Each request keeps its own step budget on the device (a gate kernel at
  the top of every step), so requests of one pool finish independently.
* Time-to-result = arrival -> final state on the host, split into queueing
  (arrival -> slot) and service (slot -> result).

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
        self.admit = None          # time it took a slot
        self.slot_reused = False   # the slot had served an earlier request

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


class Pool:
    """One engine :class:`Batch` of ``nslots`` slots of one request class
    (model, grid size, hence geometry and mask), on its own stream.  A slot
    holds one request at a time; when its request finishes, a waiting
    request of the same class takes the slot between two chunks while the
    other slots keep their state."""

    def __init__(self, cls, nslots, precision, use_graph=True, tuning=None):
        self.cls = cls
        modelname, n = cls
        proto = Request(-1, 0.0, modelname, n, 1, 0)
        pg = []
        for _ in range(nslots):
            p = proto.param()
            pg.append((p, Grid(p)))
        self.grids = [g for _, g in pg]
        self.batch = Batch(pg, precision=precision, use_graph=use_graph, tuning=tuning)
        self.batch.set_steps(0)  # every slot idle until a request takes it
        self.slot_bytes = proto.est_bytes(precision)
        self.bytes = nslots * self.slot_bytes
        self.req = [None] * nslots
        self.fresh = [True] * nslots  # slot never used since the batch was created
        self.enqueued = 0  # steps enqueued since the last admission
        self.left = 0      # steps the longest resident request still needs

    def free_slots(self):
        return [i for i, r in enumerate(self.req) if r is None]

    def occupied(self):
        return any(r is not None for r in self.req)

    def admit(self, i, req):
        """Request ``req`` into free slot i: a fresh start of the instance,
        its initial fields, its initial inversion, dt at CFL 0.8 of its
        initial maxspeed, and its own step budget."""
        b = self.batch
        if not self.fresh[i]:
            b.reset_instance(i)
        self.fresh[i] = False
        for k, v in initial_fields(req, self.grids[i]).items():
            b.set_field(i, k, v)
        b.init_instance(i)
        ms = b.diagnostics(i)["maxspeed"]  # waits for the initial inversion
        b.set_dt(time_step(req, self.grids[i], ms), i)
        b.set_steps(req.steps, i)
        self.req[i] = req
        self.left = max(self.left - self.enqueued, req.steps)
        self.enqueued = 0

    def advance(self, chunk):
        k = min(chunk, self.left - self.enqueued)
        if k > 0:
            self.batch.step_async(k)
            self.enqueued += k

    def collect(self, now, results, keep_state):
        """Results of the requests that have run their steps; their slots
        become free."""
        for i, r in enumerate(self.req):
            if r is None:
                continue
            left, done = self.batch.progress(i)
            if left == 0:
                res = {"id": r.id, "arrival": r.arrival, "admit": r.admit, "finish": now,
                       "latency": now - r.arrival, "queue": r.admit - r.arrival,
                       "service": now - r.admit, "cells": r.n * r.n, "steps": r.steps,
                       "model": r.modelname, "n": r.n, "diag": self.batch.diagnostics(i),
                       "slot_reused": r.slot_reused}
                if keep_state:
                    res["state"] = self.batch.state(i)
                results[r.id] = res
                self.req[i] = None


class Server:
    """Continuous-batching request loop on one GPU.

    Requests are grouped by class (model, grid size).  Each class has pools
    (engine batches of slots, one stream each); a pool's slots take waiting
    requests of their class as they free up, between chunks of steps, while
    the pool's other requests keep running.  A new pool is created when a
    class has waiting requests and no free slot and the HBM budget allows;
    a pool whose slots are all free with nothing waiting for its class is
    destroyed, which returns its memory to the budget."""

    def __init__(self, hbm_budget=64 << 30, precision="mixed", chunk=20, max_slots=16,
                 use_graph=True, keep_state=False, tuning=None):
        self.tuning = tuning  # launch shapes of every pool's batch (None: defaults)
        self.hbm_budget = int(hbm_budget)
        self.precision = precision
        self.chunk = int(chunk)
        self.max_slots = int(max_slots)
        self.use_graph = use_graph
        self.keep_state = keep_state

    def run(self, requests):
        todo = collections.deque(sorted(requests, key=lambda r: r.arrival))
        pending = collections.OrderedDict()  # class -> deque of waiting requests
        pools = []
        used = 0
        results = {}
        peak = 0
        reused = 0
        t0 = time.perf_counter()

        def clock():
            return time.perf_counter() - t0

        def npending():
            return sum(len(q) for q in pending.values())

        while todo or npending() or pools:
            now = clock()
            while todo and todo[0].arrival <= now:
                r = todo.popleft()
                pending.setdefault((r.modelname, r.n), collections.deque()).append(r)
            busy = False
            # pools between chunks: results, slot re-use, next chunk
            for pool in list(pools):
                if not pool.batch.done():
                    busy = True
                    continue
                pool.collect(clock(), results, self.keep_state)
                q = pending.get(pool.cls)
                for i in pool.free_slots():
                    if not q:
                        break
                    r = q.popleft()
                    r.admit = clock()
                    r.slot_reused = not pool.fresh[i]
                    reused += r.slot_reused
                    pool.admit(i, r)
                if not pool.occupied():
                    pools.remove(pool)  # nothing resident, nothing waiting: release
                    used -= pool.bytes
                    del pool
                    continue
                pool.advance(self.chunk)
                busy = True
            # classes still waiting: a new pool when the budget allows
            for cls, q in pending.items():
                if not q:
                    continue
                if any(p.cls == cls and p.free_slots() for p in pools):
                    continue  # a slot of this class frees up at its next chunk
                est = q[0].est_bytes(self.precision)
                nslots = min(len(q), self.max_slots, max(1, (self.hbm_budget - used) // est))
                if used + nslots * est > self.hbm_budget and (pools or used > 0):
                    continue  # wait until memory is released
                pool = Pool(cls, nslots, self.precision, self.use_graph, self.tuning)
                used += pool.bytes
                peak = max(peak, used)
                pools.append(pool)
                for i in range(nslots):
                    r = q.popleft()
                    r.admit = clock()
                    r.slot_reused = False
                    pool.admit(i, r)
                pool.advance(self.chunk)
                busy = True
            if not busy:
                if todo and not npending() and not pools:
                    time.sleep(max(0.0, min(0.002, todo[0].arrival - clock())))
                else:
                    time.sleep(0.0002)
            else:
                time.sleep(0.0002)
        wall = clock()
        ids = sorted(results)
        lat = np.array([results[k]["latency"] for k in ids])
        que = np.array([results[k]["queue"] for k in ids])
        srv = np.array([results[k]["service"] for k in ids])
        work = sum(r["cells"] * r["steps"] for r in results.values())

        def pct(a, q):
            return float(np.percentile(a, q)) if len(a) else 0.0

        summary = {"requests": len(results), "wall_s": wall,
                   "p50_s": pct(lat, 50), "p99_s": pct(lat, 99),
                   "queue_p50_s": pct(que, 50), "queue_p99_s": pct(que, 99),
                   "service_p50_s": pct(srv, 50), "service_p99_s": pct(srv, 99),
                   "cell_updates_per_s": work / wall if wall > 0 else 0.0,
                   "cell_updates": work, "peak_resident_bytes": peak,
                   "hbm_budget": self.hbm_budget, "precision": self.precision,
                   "slot_reuses": int(reused)}
        return summary, results
