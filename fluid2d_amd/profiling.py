"""Summaries of rocprofv3 result databases (``--kernel-trace`` and
``--pmc`` runs written as ``*_results.db``), used to produce the text
summaries kept with the benchmark and the HBM-traffic figure of the bench's
roofline.

    python -m fluid2d_amd.profiling kernels DB [PATTERN]
    python -m fluid2d_amd.profiling pmc DB COUNTER [PATTERN]
    python -m fluid2d_amd.profiling traffic FETCH_DB WRITE_DB KERNEL GRID_X OUT_JSON [TUNING_JSON]
    python -m fluid2d_amd.profiling sq DB [ROWS]
    python -m fluid2d_amd.profiling stamps [WORKLOAD]
    python -m fluid2d_amd.profiling coarse [WORKLOAD]
    python -m fluid2d_amd.profiling rr [WORKLOAD]

``traffic`` scales the per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of the
kernel launches with grid size GRID_X by what the same counters read for
the plain copies of known size in the same run (``bench.py --legs
copy,...``: 5 passes of the library's celltocorner kernel over 256 MiB
arrays, a streaming kernel with plain global loads; the runtime's copy
kernel is timed beside it).  The entry for KERNEL is added to OUT_JSON
(a dict keyed by kernel name), keeping the entries of other kernels.

``stamps`` loads the profiling build of the library (``make -C
fluid2d_amd/csrc stamps``: per-workgroup phase timestamps in the 64 x 16
smoother tiles), runs three plain FP64 steps of WORKLOAD (default
euler-closed-1024) and prints the phase timeline of the last finest-level
smoother launch.  ``coarse`` does the same for the coarse kernel: the
shader-clock cycles of each level visit of its last launch of each mode
(a V-cycle's tail and an F-cycle), and ``rr`` for the fused down-leg
kernel on each level.
"""
import json
import sqlite3
import sys


def _db(path):
    return sqlite3.connect(path)


def kernel_table(path, pattern="%"):
    """(name, grid_x, grid_y, calls, avg_us, total_ms) per kernel and grid."""
    cur = _db(path).cursor()
    q = ("select name, grid_x, grid_y, count(*), avg(end-start)/1e3, sum(end-start)/1e6 "
         "from kernels where name like ? group by name, grid_x, grid_y "
         "order by sum(end-start) desc")
    return [(r[0].split("(")[0], r[1], r[2], r[3], r[4], r[5]) for r in cur.execute(q, (pattern,))]


def pmc_table(path, counter, pattern="%"):
    """(name, grid_size_x, dispatches, mean value per dispatch) of one
    counter, values summed over the hardware units of each dispatch."""
    cur = _db(path).cursor()
    q = ("select kernel_name, grid_size_x, count(*), avg(v) from "
         "(select dispatch_id, kernel_name, grid_size_x, sum(value) as v from counters_collection "
         " where counter_name = ? and kernel_name like ? group by dispatch_id) "
         "group by kernel_name, grid_size_x order by avg(v) * count(*) desc")
    return [(r[0].split("(")[0], r[1], r[2], r[3]) for r in cur.execute(q, (counter, pattern))]


def _copy_kib(path, counter, pattern="%copyBuffer%"):
    """Total counter value (KiB) over the dispatches of a calibration copy
    (the kernel matching `pattern` at its largest grid; one copy may be
    split over several dispatches)."""
    rows = pmc_table(path, counter, pattern)
    if not rows:
        return None
    big = max(r[1] for r in rows)
    return sum(r[3] * r[2] for r in rows if r[1] == big)


def traffic(fetch_db, write_db, kernel, grid_x, copy_bytes=5 * (256 << 20)):
    """HBM bytes per launch of `kernel` launches with grid size `grid_x`:
    FETCH_SIZE and WRITE_SIZE scaled so that the five streaming calibration
    passes (celltocorner over 256 MiB arrays; `copy_bytes` read and written
    in total) read and write exactly `copy_bytes`.  The scale from the
    runtime's copy kernel is reported beside it; that one also counts any
    other copies of the same grid size in the run."""
    out = {"kernel": kernel, "grid_x": grid_x, "copy_bytes": copy_bytes,
           "calibration": "k_celltocorner<double>, 5 passes over 256 MiB"}
    tot = 0.0
    for name, path in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        rows = [r for r in pmc_table(path, name, "%" + kernel + "%") if r[1] == grid_x]
        if not rows:
            raise ValueError("no %s rows for %s grid %d" % (name, kernel, grid_x))
        kib = sum(r[3] * r[2] for r in rows) / sum(r[2] for r in rows)
        skib = _copy_kib(path, name, "%k_celltocorner%")
        ckib = _copy_kib(path, name)
        factor = copy_bytes / (skib * 1024.0) if skib else 1.0
        cfactor = copy_bytes / (ckib * 1024.0) if ckib else None
        out[name] = {"kib_per_dispatch": kib, "calibration_kib_total": skib, "scale": factor,
                     "runtime_copy_kib_total": ckib, "runtime_copy_scale": cfactor,
                     "bytes_per_launch": kib * 1024.0 * factor}
        tot += kib * 1024.0 * factor
    out["bytes_per_launch"] = tot
    return out


def sq_table(path, wave_cycle_rate=None):
    """Per (kernel, grid) mean per dispatch of the SQ counters of a --pmc
    run (values summed over the hardware units), with VALU instructions per
    wavefront, ordered by total VALU instructions."""
    cur = _db(path).cursor()
    q = ("select kernel_name, grid_size_x, counter_name, count(*), avg(v) from "
         "(select dispatch_id, kernel_name, grid_size_x, counter_name, sum(value) as v "
         " from counters_collection group by dispatch_id, counter_name) "
         "group by kernel_name, grid_size_x, counter_name")
    d = {}
    for k, g, c, n, v in cur.execute(q):
        e = d.setdefault((k.split("(")[0], g), {"calls": 0})
        e[c] = v
        e["calls"] = max(e["calls"], n)
    rows = []
    for (k, g), e in d.items():
        waves = e.get("SQ_WAVES", 0.0)
        valu = e.get("SQ_INSTS_VALU", 0.0)
        rows.append((k, g, e["calls"], waves, valu, valu / waves if waves else 0.0, e))
    rows.sort(key=lambda r: -r[4] * r[2])
    return rows


def stamps(workload="euler-closed-1024"):
    """Phase timeline of the last 64 x 16 smoother launch (profiling build):
    per-phase durations (us, 100 MHz clock) over the workgroups, the end
    times, and wall tiles against interior tiles."""
    import ctypes as C
    import os

    import numpy as np

    from . import _lib
    _lib.LIB_PATH = os.path.join(os.path.dirname(_lib.LIB_PATH), "libf2dgpu_stamps.so")
    from . import workloads
    from .batch import Batch
    from .grid import Grid
    p = workloads.make_param(workload)
    g = Grid(p)
    tun = {"coarse_lds_budget": 159744, "min_tiles": 1024, "coarse_wave_cells": 0, "persist": 0}
    b = Batch([(p, g)], precision="fp64", tuning=tun)
    for k, v in workloads.initial_fields(workload, g).items():
        b.set_field(0, k, v)
    b.set_psi_from_vorticity()
    b.set_dt(workloads.time_step(workload, g, b.diagnostics(0)["maxspeed"]))
    b.step(3)
    lib = _lib.load()
    buf = (C.c_ulonglong * (2048 * 8))()
    if lib.f2d_debug_smooth_stamps(buf) != 0:
        raise RuntimeError("f2d_debug_smooth_stamps failed")
    a = np.frombuffer(buf, dtype=np.uint64).reshape(2048, 8).astype(np.int64)
    a = a[a[:, 0] > 0]
    t = (a[:, :6] - a[:, 0].min()) * 0.01
    info = a[:, 7]
    y0, x0 = (info >> 4) & 0xFFF, (info >> 16) & 0xFFFF
    m, n = g.nyl, g.nxl
    wall = (y0 <= 3) | (y0 + 16 >= m - 3) | (x0 <= 3) | (x0 + 64 >= n - 3)
    print("workgroups %d on %d compute units, mode %s" % (len(a), len(np.unique(info >> 32)),
                                                          np.unique(info & 15).tolist()))
    names = ["start", "tile loads issued, gate read", "tile in LDS", "prolongation",
             "first sweep", "second sweep, stores"]
    print("%-30s %7s %7s %7s %7s %8s %8s" % ("phase (us)", "mean", "p10", "p50", "p90", "wall",
                                              "interior"))
    for k, nm in enumerate(names):
        v = t[:, 0] if k == 0 else t[:, k] - t[:, k - 1]
        print("%-30s %7.2f %7.2f %7.2f %7.2f %8.2f %8.2f" %
              (nm, v.mean(), *np.percentile(v, [10, 50, 90]), v[wall].mean(), v[~wall].mean()))
    e = t[:, 5]
    print("workgroup end (us): p10 %.2f p50 %.2f p90 %.2f max %.2f" %
          (*np.percentile(e, [10, 50, 90]), e.max()))
    print("wall tiles %d: mean end %.2f us; interior tiles: %.2f us" %
          (wall.sum(), e[wall].mean(), e[~wall].mean()))
    cu = info >> 32
    last = np.argsort(-e)[:8]
    print("last to end: " + ", ".join("(y0 %d, x0 %d, cu %x, %.2f us, %d on its cu, %d wall)" % (
        y0[q], x0[q], cu[q], e[q], (cu == cu[q]).sum(), wall[cu == cu[q]].sum()) for q in last))


def coarse_stamps(workload="euler-closed-1024"):
    """Phase breakdown of the last coarse-kernel launch of each mode
    (profiling build): the shader-clock cycles of each level visit, with the
    launch's real-time length for the clock rate."""
    import ctypes as C
    import os

    import numpy as np

    from . import _lib
    _lib.LIB_PATH = os.path.join(os.path.dirname(_lib.LIB_PATH), "libf2dgpu_stamps.so")
    from . import workloads
    from .batch import Batch
    from .grid import Grid
    p = workloads.make_param(workload)
    g = Grid(p)
    tun = {"coarse_lds_budget": 159744, "min_tiles": 1024, "coarse_wave_cells": 0, "persist": 0}
    b = Batch([(p, g)], precision="fp64", tuning=tun)
    for k, v in workloads.initial_fields(workload, g).items():
        b.set_field(0, k, v)
    b.set_psi_from_vorticity()
    b.set_dt(workloads.time_step(workload, g, b.diagnostics(0)["maxspeed"]))
    b.step(3)
    lib = _lib.load()
    buf = (C.c_ulonglong * 512)()
    if lib.f2d_debug_coarse_stamps(buf) != 0:
        raise RuntimeError("f2d_debug_coarse_stamps failed")
    a = np.frombuffer(buf, dtype=np.uint64).reshape(2, 256).astype(np.int64)
    names = {0: "start", 1: "operators in LDS, rhs widened", 2: "end (x stored)", 20: "coarsest solve"}
    for mode in (0, 1):
        r = a[mode]
        nl = int(r[255] - 100)
        wall_us = (r[252] - r[254]) * 0.01
        pairs = [(int(r[2 * k]), int(r[2 * k + 1])) for k in range(126) if r[2 * k] > 0]
        if not pairs:
            continue
        t0 = pairs[0][0]
        cyc = pairs[-1][0] - t0
        mhz = cyc / wall_us if wall_us > 0 else 0.0
        print("mode %d (%s): %d levels, %.2f us, %d cycles (%.0f MHz)" % (
            mode, "V-cycle tail" if mode == 0 else "F-cycle", nl, wall_us, cyc, mhz))
        prev = t0
        tot = {}
        for c, code in pairs[1:]:
            if code in names:
                nm = names[code]
            elif 10 <= code < 20:
                nm = "down, level +%d (smooth, residual, restrict)" % (code - 10)
            elif 30 <= code < 40:
                nm = "up, level +%d (interpolate, smooth)" % (code - 30)
            elif 40 <= code < 50:
                nm = "restrict b, level +%d" % (code - 40)
            else:
                nm = "interpolate onto level +%d" % (code - 50)
            tot[nm] = tot.get(nm, 0) + (c - prev)
            if mode == 0:
                print("  %-46s %7d cycles %6.2f us" % (nm, c - prev, (c - prev) / mhz if mhz else 0))
            prev = c
        if mode == 1:
            for nm, v in tot.items():
                print("  %-46s %7d cycles %6.2f us (sum)" % (nm, v, v / mhz if mhz else 0))


def rr_stamps(workload="euler-closed-1024"):
    """Phase timeline of the fused down-leg kernel (k_smooth_rr) per level
    (profiling build): mean phase durations over the first 256 workgroups of
    the last launch on each level (us, 100 MHz clock)."""
    import ctypes as C
    import os

    import numpy as np

    from . import _lib
    _lib.LIB_PATH = os.path.join(os.path.dirname(_lib.LIB_PATH), "libf2dgpu_stamps.so")
    from . import workloads
    from .batch import Batch
    from .grid import Grid
    p = workloads.make_param(workload)
    g = Grid(p)
    tun = {"coarse_lds_budget": 159744, "min_tiles": 1024, "coarse_wave_cells": 0, "persist": 0}
    b = Batch([(p, g)], precision="fp64", tuning=tun)
    for k, v in workloads.initial_fields(workload, g).items():
        b.set_field(0, k, v)
    b.set_psi_from_vorticity()
    b.set_dt(workloads.time_step(workload, g, b.diagnostics(0)["maxspeed"]))
    b.step(3)
    lib = _lib.load()
    buf = (C.c_ulonglong * (8 * 256 * 8))()
    if lib.f2d_debug_rr_stamps(buf) != 0:
        raise RuntimeError("f2d_debug_rr_stamps failed")
    a = np.frombuffer(buf, dtype=np.uint64).reshape(8, 256, 8).astype(np.int64)
    names = ["task read", "loads issued, gate read", "first sweep", "second sweep",
             "residuals, x stored", "restriction, end"]
    print("%-6s %5s %7s" % ("level", "WGs", "start") + "".join(" %11s" % n[:11] for n in names)
          + " %7s %7s" % ("end p50", "max"))
    for lev in range(8):
        r = a[lev]
        r = r[r[:, 0] > 0]
        if len(r) == 0:
            continue
        ng = int(r[0, 7] & 0xFFFFFFFF)
        t = (r[:, :7] - r[:, 0].min()) * 0.01
        ph = [np.mean(t[:, k + 1] - t[:, k]) for k in range(6)]
        print("%-6d %5d %7.2f" % (lev, ng, np.mean(t[:, 0])) + "".join(" %11.2f" % v for v in ph)
              + " %7.2f %7.2f" % (np.median(t[:, 6]), t[:, 6].max()))


def main(argv):
    if argv and argv[0] == "stamps":
        stamps(*argv[1:2])
        return 0
    if argv and argv[0] == "rr":
        rr_stamps(*argv[1:2])
        return 0
    if argv and argv[0] == "coarse":
        coarse_stamps(*argv[1:2])
        return 0
    if len(argv) < 2:
        print(__doc__)
        return 2
    cmd = argv[0]
    if cmd == "kernels":
        pat = argv[2] if len(argv) > 2 else "%"
        print("%-48s %9s %4s %7s %10s %10s" % ("kernel", "grid_x", "gy", "calls", "avg_us",
                                               "total_ms"))
        for r in kernel_table(argv[1], pat):
            print("%-48s %9d %4d %7d %10.2f %10.3f" % (r[0][:48], r[1], r[2], r[3], r[4], r[5]))
    elif cmd == "pmc":
        pat = argv[3] if len(argv) > 3 else "%"
        print("%-48s %9s %6s %14s" % ("kernel", "grid_x", "calls", argv[2]))
        for r in pmc_table(argv[1], argv[2], pat):
            print("%-48s %9d %6d %14.1f" % (r[0][:48], r[1], r[2], r[3]))
    elif cmd == "sq":
        cs = ["SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_BUSY_CYCLES", "SQ_WAIT_INST_ANY"]
        print("%-48s %8s %6s %9s %12s %9s" % ("kernel", "grid_x", "calls", "waves", "VALU",
                                              "VALU/wave") + "".join(" %14s" % c for c in cs))
        for k, g, n, w, v, vw, e in sq_table(argv[1])[:int(argv[2]) if len(argv) > 2 else 30]:
            print("%-48s %8d %6d %9.0f %12.4g %9.1f" % (k[:48], g, n, w, v, vw)
                  + "".join(" %14.4g" % e.get(c, 0.0) for c in cs))
    elif cmd == "traffic":
        t = traffic(argv[1], argv[2], argv[3], int(argv[4]))
        if len(argv) > 6:
            t["tuning"] = json.loads(argv[6])
        try:
            with open(argv[5]) as f:
                allk = json.load(f)
        except (OSError, ValueError):
            allk = {}
        if "kernel" in allk:  # a single-kernel file: keep it under its name
            allk = {allk["kernel"]: allk}
        allk[argv[3]] = t
        with open(argv[5], "w") as f:
            json.dump(allk, f, indent=1, sort_keys=True)
        print(json.dumps(t))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
