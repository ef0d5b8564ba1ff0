"""Launch-configuration autotuner for the batched engine.

Two knobs shape the launches of a batch (``f2d_batch_tuning``):

* ``coarse_lds_budget``: how many of the coarse multigrid levels the
  one-workgroup coarse kernel takes over (its operands must fit in LDS);
  more levels there means fewer, but serial, phases per V-cycle.
* ``min_tiles``: the fine phases use 64 x 16, 64 x 8 or 64 x 4 tiles, the
  tallest that still gives this many workgroups, so that small levels
  spread over more compute units.
* ``coarse_wave_cells``: inside the coarse kernel, levels up to this size
  run on a single wavefront (no workgroup barriers) instead of the whole
  workgroup.
* ``persist``: runs of consecutive smoother and residual/restriction phases
  below the finest level run as one persistent launch with device-wide
  barriers (at most this many workgroups per compute unit), or one launch
  per phase (0).  Not searched by default: on MI355X a device-wide barrier
  costs more than a dependent launch of a replayed graph (about 1.7 us)
  as soon as more than a few workgroups take part (about 4 us at 64 and
  12 us at 256 workgroups), so one launch per phase wins.  ``tune(...,
  persist=(1, 2))`` tries it on the best of the other three.

:func:`tune` builds a scratch batch with the same instance specs (so the
caller's state is never touched), starts it from seeded turbulence,
replays each candidate's captured step a few times and keeps the fastest.
Results are cached per batch signature in ``$FLUID2D_AMD_CACHE``
(default ``~/.cache/fluid2d_amd/tune.json``).
"""
import json
import os
import time

from . import workloads
from .batch import Batch
from .grid import Grid
from .param import Param

BUDGETS = (0, 98304, 49152, 24576)
MIN_TILES = (0, 1024, 4096, 1 << 24)
WAVE_CELLS = (256, 0)
PERSIST = ()


def _pg(specs):
    out = []
    for s in specs:
        if isinstance(s, tuple):
            out.append(s)
        else:
            p = Param(**s)
            out.append((p, Grid(p)))
    return out


def signature(pg, precision):
    keys = {}
    for p, g in pg:
        k = "%s:%dx%d:%s" % (p.modelname, g.nx, g.ny, "perio" if g.geometry == "perio" else "wall")
        keys[k] = keys.get(k, 0) + 1
    return "v2|" + precision + "|" + ",".join("%s*%d" % (k, keys[k]) for k in sorted(keys))


def _cache_path():
    d = os.environ.get("FLUID2D_AMD_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "fluid2d_amd"))
    return os.path.join(d, "tune.json")


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(c):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _init(b, seed=11):
    for i, (p, g) in enumerate(zip(b.params, b.grids)):
        b.set_field(i, "vorticity", workloads.turbulence(g, seed + i))
        if p.modelname == "boussinesq":
            b.set_field(i, "buoyancy", workloads.rayleigh_taylor(g, seed + i))
    b.set_psi_from_vorticity()
    for i, (p, g) in enumerate(zip(b.params, b.grids)):
        ms = b.diagnostics(i)["maxspeed"]
        u = ms if (p.modelname == "euler" and ms > 0) else 1.0
        b.set_dt(0.8 * min(g.dx, g.dy) / u, i)


def candidates():
    return [{"coarse_lds_budget": cb, "min_tiles": mt, "coarse_wave_cells": wc}
            for cb in BUDGETS for mt in MIN_TILES for wc in WAVE_CELLS]


def tune(specs, precision="fp64", steps=3, cands=None, use_cache=True, verbose=False,
         persist=PERSIST):
    """Best tuning for a batch of ``specs``: returns (tuning dict, table of
    (tuning, ms per step)) with the default configuration first.  The
    persistent multigrid launches (``persist`` workgroups per compute unit)
    are tried on the best configuration found without them."""
    pg = _pg(specs)
    sig = signature(pg, precision)
    cache = _load_cache() if use_cache else {}
    if use_cache and sig in cache:
        return cache[sig]["best"], cache[sig]["table"]
    cands = cands or candidates()
    b = Batch(pg, precision=precision)
    _init(b)
    table = []

    def trial(c):
        try:
            b.retune(c)
            b.step(1)  # capture + warm
            t0 = time.perf_counter()
            b.step(steps)
        except Exception as e:  # a budget too small for the coarsest level
            if verbose:
                print("autotune: skip %s (%s)" % (c, e))
            return
        ms = (time.perf_counter() - t0) / steps * 1e3
        table.append((c, ms))
        if verbose:
            print("autotune: %s %.3f ms/step" % (c, ms), flush=True)

    for c in cands:
        trial(dict(c, persist=0))
    best = min(table, key=lambda r: r[1])[0]
    for p in persist:
        trial(dict(best, persist=p))
    del b
    best = min(table, key=lambda r: r[1])[0]
    if use_cache:
        cache[sig] = {"best": best, "table": table}
        _store_cache(cache)
    return best, table
