"""Every benchmark configuration, as its workload names it, on the device.

* euler-perio-128: 128x128 doubly periodic, Gaussian vortex pair + seeded
  noise, fixed dt at CFL 0.8, 200 RK3_SSP steps; the FP64 engine and the
  ``Fluid2d`` drop-in against the oracle, rel-L2 <= 1e-12 per field.
* euler-closed-1024 and boussinesq-rt-2048x1024 in mixed precision: 1000
  steps against the FP64 engine (bit-identical to the oracle at these sizes,
  ``tests/test_gpu_fullsize.py``); rel-L2 <= 1e-6 per field.
* the 256-instance ensemble (sizes 64..512, periodic and closed, k^-3
  inits), FP64, a few steps: one instance of every (size, geometry) class
  against the oracle, and the smallest and largest instance alone == in
  the mix (bit for bit).
* the serving stream (32 Poisson arrivals at 50/s, Euler/Boussinesq,
  128^2..2048^2, 50..2000 steps) under a memory budget that binds: every
  request runs its own step count, admission stays within the budget, and a
  sample of requests equals the same request run alone, bit for bit.
* the per-step diagnostic sums (core/euler.py:185-221,
  core/boussinesq.py:123-153) against the oracle, and the engine's
  on-device adaptive dt (core/fluid2d.py:376-397) against the oracle
  stepped with the host rule.
"""
import numpy as np
import pytest

import fluid2d_amd as F
from fluid2d_amd import serving, workloads
from fluid2d_amd.batch import Batch
from fluid2d_amd.grid import Grid
from fluid2d_amd.param import Param
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12
TOL_MIXED = 1e-6


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def oracle_like(p, g):
    return orc.OracleModel(p.modelname, g.nx, g.ny, g.Lx, g.Ly, geometry=g.geometry,
                           order=p.order, flux_splitting_method=p.flux_splitting_method,
                           aparab=p.aparab, timestepping=p.timestepping,
                           diffusion=p.diffusion, Kdiff=p.Kdiff, gravity=p.gravity)


def config_engine(name, precision, dt=None):
    p = workloads.make_param(name)
    g = Grid(p)
    b = Batch([(p, g)], precision=precision)
    fields = workloads.initial_fields(name, g)
    for k, v in fields.items():
        b.set_field(0, k, v)
    b.set_psi_from_vorticity()
    if dt is None:
        dt = workloads.time_step(name, g, b.diagnostics(0)["maxspeed"])
    b.set_dt(dt)
    return b, p, g, fields, dt


def test_c1_vortex_pair_200_steps_engine_vs_oracle():
    name = "euler-perio-128"
    print("case", name, "engine", flush=True)
    b, p, g, fields, dt = config_engine(name, "fp64")
    o = oracle_like(p, g)
    for k, v in fields.items():
        o.get(k)[...] = v
    o.set_psi_from_vorticity()
    d0 = o.diagnostics()
    assert b.diagnostics(0)["maxspeed"] == d0["maxspeed"]
    assert dt == workloads.time_step(name, g, d0["maxspeed"])
    nsteps = workloads.CONFIGS[name][2]
    assert nsteps == 200
    b.step(nsteps)
    log = o.run(nsteps, dt)
    s = b.state(0)
    for k, vname in enumerate(o.varname_list):
        r = rel(s[k], o.state[k])
        assert r <= TOL, "%s rel-L2 %.3e" % (vname, r)
    nite, _, div = b.solve_info(0)
    assert nite == log[-1] and not div


def test_c1_vortex_pair_fluid2d_dropin_vs_oracle():
    name = "euler-perio-128"
    print("case", name, "Fluid2d", flush=True)
    p = workloads.make_param(name)
    g = F.Grid(p)
    f = F.Fluid2d(p, g)
    o = oracle_like(p, g)
    w = workloads.initial_fields(name, g)["vorticity"]
    f.model.var.get("vorticity")[...] = w
    o.get("vorticity")[...] = w
    f.model.set_psi_from_vorticity()
    o.set_psi_from_vorticity()
    d = f.model.diagnostics(f.model.var, 0.0)
    dt = workloads.time_step(name, g, d["maxspeed"])
    assert d["maxspeed"] == o.diagnostics()["maxspeed"]
    f.model.run(200, dt)
    o.run(200, dt)
    s = f.model.var.state
    for k, vname in enumerate(f.model.var.varname_list):
        r = rel(s[k], o.state[k])
        assert r <= TOL, "%s rel-L2 %.3e" % (vname, r)


@pytest.mark.parametrize("name", ["euler-closed-1024", "boussinesq-rt-2048x1024"])
def test_mixed_precision_1000_steps_vs_fp64(name):
    """The configuration's 1000 steps in mixed precision against the FP64
    engine (same dt), every field within rel-L2 1e-6."""
    print("case mixed", name, flush=True)
    nsteps = workloads.CONFIGS[name][2]
    assert nsteps == 1000
    b64, _, _, _, dt = config_engine(name, "fp64")
    b64.step(nsteps)
    s64 = b64.state(0)
    del b64
    bm, _, _, _, _ = config_engine(name, "mixed", dt)
    bm.step(nsteps)
    sm = bm.state(0)
    names = F.models.VARNAMES[workloads.CONFIGS[name][0]["modelname"]]
    for k in range(s64.shape[0]):
        if not np.any(s64[k]):
            continue  # Euler's source plane stays zero
        r = rel(sm[k], s64[k])
        print("  %s rel-L2 %.3e" % (names[k], r), flush=True)
        assert r <= TOL_MIXED, "%s rel-L2 %.3e after %d steps" % (names[k], r, nsteps)


def _ensemble():
    specs = workloads.ensemble_specs(256)
    pg = []
    for kw, _ in specs:
        p = Param(**kw)
        pg.append((p, Grid(p)))
    return specs, pg


def test_c4_ensemble_fp64_sample_vs_oracle_and_alone():
    nsteps = 3
    specs, pg = _ensemble()
    print("case ensemble 256 fp64", flush=True)
    b = Batch(pg, precision="fp64")
    fields = []
    for i, (_, seed) in enumerate(specs):
        w = workloads.turbulence(b.grids[i], seed)
        b.set_field(i, "vorticity", w)
        fields.append(w)
    b.set_psi_from_vorticity()
    dts = []
    for i, g in enumerate(b.grids):
        dts.append(0.8 * min(g.dx, g.dy) / b.diagnostics(i)["maxspeed"])
        b.set_dt(dts[i], i)
    b.step(nsteps)
    # one instance of every (size, geometry) class
    first = {}
    for i, (p, g) in enumerate(pg):
        first.setdefault((g.nx, g.geometry), i)
    assert len(first) == 8
    for (nx, geo), i in sorted(first.items()):
        print("  oracle instance", i, nx, geo, flush=True)
        p, g = pg[i]
        o = oracle_like(p, g)
        o.get("vorticity")[...] = fields[i]
        o.set_psi_from_vorticity()
        o.diagnostics()
        o.run(nsteps, dts[i])
        s = b.state(i)
        for k, vname in enumerate(o.varname_list):
            r = rel(s[k], o.state[k])
            assert r <= TOL, "instance %d (%d %s) %s rel-L2 %.3e" % (i, nx, geo, vname, r)
    sizes = [g.nx for _, g in pg]
    for i in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        print("  alone instance", i, flush=True)
        a = Batch([pg[i]], precision="fp64")
        a.set_field(0, "vorticity", fields[i])
        a.set_psi_from_vorticity()
        a.set_dt(dts[i])
        a.step(nsteps)
        assert (a.state(0) == b.state(i)).all(), "instance %d differs alone vs in the ensemble" % i


def run_alone(req, precision):
    p = req.param()
    g = Grid(p)
    b = Batch([(p, g)], precision=precision)
    for k, v in serving.initial_fields(req, g).items():
        b.set_field(0, k, v)
    b.set_psi_from_vorticity()
    b.set_dt(serving.time_step(req, g, b.diagnostics(0)["maxspeed"]))
    b.step(req.steps)
    return b.state(0)


def test_c5_full_stream_alone_equals_in_mix():
    reqs = serving.stream_requests(32)
    assert min(r.n for r in reqs) >= 128 and max(r.n for r in reqs) <= 2048
    # a budget of a third of the whole stream's footprint: admission binds
    total = sum(r.est_bytes("mixed") for r in reqs)
    budget = max(total // 3, max(r.est_bytes("mixed") for r in reqs))
    print("case serving stream 32 requests, budget %.2f GB of %.2f" % (budget / 1e9, total / 1e9),
          flush=True)
    srv = serving.Server(hbm_budget=budget, precision="mixed", chunk=20, keep_state=True)
    summary, res = srv.run(reqs)
    assert summary["requests"] == len(reqs)
    assert summary["peak_resident_bytes"] <= budget
    for r in reqs:
        assert res[r.id]["steps"] == r.steps
    # sample: the smallest and the largest grid of each model
    sample = set()
    for model in ("euler", "boussinesq"):
        rs = [r for r in reqs if r.modelname == model]
        sample.add(min(rs, key=lambda r: (r.n, r.steps)).id)
        sample.add(max(rs, key=lambda r: (r.n, -r.steps)).id)
    for r in reqs:
        if r.id not in sample:
            continue
        print("  alone request", r.id, r.modelname, r.n, r.steps, flush=True)
        a = run_alone(r, "mixed")
        assert (a == res[r.id]["state"]).all(), "request %d differs alone vs in the stream" % r.id


def abs_sums(o, kind, g):
    """Sums of the magnitudes of the terms of each diagnostic sum
    (core/fortran_diag.f90: masks msk == 1 or msk != 0 as the routine
    takes them), over the interior."""
    s = np.s_[3:-3, 3:-3]
    st = o.state
    m1 = (o.msk[s] == 1)
    m0 = (o.msk[s] != 0)
    xr, yr = o.grid["xr"][s], o.grid["yr"][s]
    w = st[o.varname_list.index("vorticity")][s]
    u, v = st[o.varname_list.index("u")], st[o.varname_list.index("v")]
    ke = (u[s] ** 2 + u[3:-3, 2:-4] ** 2 + v[s] ** 2 + v[2:-4, 3:-3] ** 2)
    out = np.zeros(8)
    out[1] = np.sum(ke * m1)
    out[2] = np.sum(np.abs(w) * m1)
    out[3] = np.sum(w ** 2 * m1)
    if kind == "euler":
        psi = st[o.varname_list.index("psi")][s]
        src = st[o.varname_list.index("source")][s]
        out[4] = np.sum(np.abs(w * xr) * m0)
        out[5] = np.sum(np.abs(w * yr) * m0)
        out[6] = np.sum(np.abs(psi) * m0)
        out[7] = np.sum(np.abs(w * src) * m0)
    else:
        bu = st[o.varname_list.index("buoyancy")][s]
        out[4] = np.sum(np.abs(bu * yr) * m0)
        out[5] = np.sum(np.abs(bu) * m1)
        out[6] = np.sum(bu ** 2 * m1)
    return np.maximum(out, 1e-300)


@pytest.mark.parametrize("case", [
    ("euler", 128, 128, 1.0, 1.0, "closed"),
    ("euler", 128, 128, 1.0, 1.0, "perio"),
    ("boussinesq", 256, 128, 2.0, 1.0, "xchannel"),
])
def test_diagnostic_sums_vs_oracle(case):
    kind, nx, ny, Lx, Ly, geo = case
    print("case diagnostics", case, flush=True)
    kw = dict(modelname=kind, nx=nx, ny=ny, Lx=Lx, Ly=Ly, geometry=geo, adaptable_dt=False)
    if kind == "boussinesq":
        kw.update(diffusion=True, Kdiff=1e-5)
    p = Param(**kw)
    g = Grid(p)
    b = Batch([(p, g)])
    o = oracle_like(p, g)
    w = workloads.turbulence(g, 21)
    b.set_field(0, "vorticity", w)
    o.get("vorticity")[...] = w
    if kind == "boussinesq":
        bb = workloads.rayleigh_taylor(g, 22) + 0.1 * workloads.turbulence(g, 23)
        b.set_field(0, "buoyancy", bb)
        o.get("buoyancy")[...] = bb
    b.set_psi_from_vorticity()
    o.set_psi_from_vorticity()
    dt = 0.5 * g.dx / b.diagnostics(0)["maxspeed"]
    for when in ("initial", "after 4 steps"):
        if when != "initial":
            b.set_dt(dt)
            b.step(4)
            o.run(4, dt)
        got = b.raw_diagnostics(0)
        ref = o.diagnostics()["raw"]
        # maxspeed is a max: exact.  The sums are tree-ordered on the device
        # and serial in the reference: they agree to rounding, i.e. to a
        # small multiple of eps times the sum of the terms' magnitudes (a sum
        # that cancels, such as the total vorticity, is compared at that
        # scale, not at its own)
        assert got[0] == ref[0], when
        scale = abs_sums(o, kind, g)
        for k in range(1, 8):
            assert abs(got[k] - ref[k]) <= 1e-13 * scale[k], (when, k, got[k], ref[k], scale[k])


@pytest.mark.parametrize("cfl,cap", [(0.5, None), (0.8, 0.9)])
def test_engine_adaptive_dt_vs_oracle(cfl, cap):
    """dt = cfl * min(dx, dy) / maxspeed of the previous diagnostics, capped
    at dtmax (core/fluid2d.py:376-397), evaluated on the device every step."""
    print("case adaptive dt", cfl, cap, flush=True)
    p = Param(modelname="euler", nx=128, ny=128, geometry="closed", adaptable_dt=True)
    g = Grid(p)
    o = oracle_like(p, g)
    w = workloads.turbulence(g, 31)
    o.get("vorticity")[...] = w
    o.set_psi_from_vorticity()
    d = o.diagnostics()
    dt_cfl = cfl * min(g.dx, g.dy) / d["maxspeed"]
    dtmax = 1e9 if cap is None else cap * dt_cfl
    b = Batch([(p, g)])
    b.set_adaptive(True, cfl=cfl, dtmax=dtmax)
    b.set_field(0, "vorticity", w)
    b.set_psi_from_vorticity()
    nsteps = 12
    b.step(nsteps)
    capped = 0
    for _ in range(nsteps):
        dt = cfl * min(g.dx, g.dy) / o.diags["maxspeed"]
        if dt > dtmax:
            dt = dtmax
            capped += 1
        o.step(dt)
    if cap is not None:
        assert capped > 0
    s = b.state(0)
    for k, vname in enumerate(o.varname_list):
        r = rel(s[k], o.state[k])
        assert r <= TOL, "%s rel-L2 %.3e" % (vname, r)
