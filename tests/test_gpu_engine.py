"""The batched, graph-replayed engine against the oracle.

FP64 results must match the oracle within rel-L2 1e-12 per field (the fused
kernels keep the reference's operation order, so they are expected to be
bit-identical), for a single instance and for the same instance inside a
mixed batch ("alone == in the mix", bit for bit).
"""
import numpy as np
import pytest

from fluid2d_amd import workloads
from fluid2d_amd.batch import Batch
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def spec(kind, nx, ny, Lx, Ly, geom, **kw):
    d = dict(modelname=kind, nx=nx, ny=ny, Lx=Lx, Ly=Ly, geometry=geom, adaptable_dt=False)
    d.update(kw)
    return d


def init_fields(b, i, seed):
    g = b.grids[i]
    w = workloads.turbulence(g, seed)
    b.set_field(i, "vorticity", w)
    out = {"vorticity": w}
    if b.params[i].modelname == "boussinesq":
        bb = workloads.rayleigh_taylor(g, seed + 1)
        b.set_field(i, "buoyancy", bb)
        out["buoyancy"] = bb
    return out


def oracle_for(b, i, fields):
    p, g = b.params[i], b.grids[i]
    o = orc.OracleModel(p.modelname, g.nx, g.ny, g.Lx, g.Ly, geometry=g.geometry,
                        diffusion=p.diffusion, Kdiff=p.Kdiff, gravity=p.gravity)
    for k, v in fields.items():
        o.get(k)[...] = v
    o.set_psi_from_vorticity()
    o.diagnostics()
    return o


def dt_for(b, i, maxspeed):
    g = b.grids[i]
    u = maxspeed if b.params[i].modelname == "euler" else 1.0
    return 0.8 * min(g.dx, g.dy) / u


@pytest.mark.parametrize("case", [
    ("euler", 128, 128, 1.0, 1.0, "perio", 30),
    ("euler", 256, 256, 1.0, 1.0, "closed", 20),
    ("boussinesq", 256, 128, 2.0, 1.0, "xchannel", 20),
])
def test_engine_single_vs_oracle(case):
    kind, nx, ny, Lx, Ly, geom, nsteps = case
    kw = dict(diffusion=True, Kdiff=1e-5) if kind == "boussinesq" else {}
    b = Batch([spec(kind, nx, ny, Lx, Ly, geom, **kw)])
    f = init_fields(b, 0, 3)
    b.set_psi_from_vorticity()
    o = oracle_for(b, 0, f)
    s0 = b.state(0)
    for k in range(s0.shape[0]):
        assert rel(s0[k], o.state[k]) <= TOL, ("init", k)
    d = b.diagnostics(0)
    assert d["maxspeed"] == o.diags["maxspeed"]
    dt = dt_for(b, 0, d["maxspeed"])
    b.set_dt(dt)
    b.step(nsteps)
    o.run(nsteps, dt)
    s = b.state(0)
    for k in range(s.shape[0]):
        r = rel(s[k], o.state[k])
        assert r <= TOL, "field %d rel-L2 %.3e" % (k, r)
    if kind == "euler":
        nite, res, div = b.solve_info(0)
        assert nite == o.last_solve()[0] and not div


def test_alone_equals_in_mix():
    specs = [
        spec("euler", 64, 64, 1.0, 1.0, "closed"),
        spec("euler", 128, 128, 1.0, 1.0, "perio"),
        spec("boussinesq", 128, 64, 2.0, 1.0, "xchannel", diffusion=True, Kdiff=1e-5),
        spec("euler", 256, 256, 1.0, 1.0, "closed"),
    ]
    mix = Batch(specs)
    fields = [init_fields(mix, i, 10 + i) for i in range(len(specs))]
    mix.set_psi_from_vorticity()
    dts = []
    for i in range(len(specs)):
        dts.append(dt_for(mix, i, mix.diagnostics(i)["maxspeed"]))
        mix.set_dt(dts[i], i)
    mix.step(12)
    for i in range(len(specs)):
        alone = Batch([specs[i]])
        for k, v in fields[i].items():
            alone.set_field(0, k, v)
        alone.set_psi_from_vorticity()
        alone.set_dt(dts[i])
        alone.step(12)
        a, m = alone.state(0), mix.state(i)
        assert (a == m).all(), "instance %d differs alone vs in the mix" % i


def test_graph_equals_eager():
    s = spec("euler", 128, 128, 1.0, 1.0, "closed")
    out = []
    for use_graph in (True, False):
        b = Batch([s], use_graph=use_graph)
        init_fields(b, 0, 4)
        b.set_psi_from_vorticity()
        b.set_dt(dt_for(b, 0, b.diagnostics(0)["maxspeed"]))
        b.step(3)
        b.step(2)  # replays after the first capture
        out.append(b.state(0))
    assert (out[0] == out[1]).all()


@pytest.mark.parametrize("case", [
    ("euler", 128, 128, 1.0, 1.0, "perio", 50),
    ("euler", 256, 256, 1.0, 1.0, "closed", 50),
    ("boussinesq", 256, 128, 2.0, 1.0, "xchannel", 50),
])
def test_mixed_precision_error(case):
    """Mixed mode (FP32 advection/V-cycles, FP64 state and refinement) stays
    within rel-L2 1e-6 of the FP64 path (itself bit-identical to the oracle)."""
    kind, nx, ny, Lx, Ly, geom, nsteps = case
    kw = dict(diffusion=True, Kdiff=1e-5) if kind == "boussinesq" else {}
    out = {}
    for prec in ("fp64", "mixed"):
        b = Batch([spec(kind, nx, ny, Lx, Ly, geom, **kw)], precision=prec)
        init_fields(b, 0, 3)
        b.set_psi_from_vorticity()
        if prec == "fp64":
            dt = dt_for(b, 0, b.diagnostics(0)["maxspeed"])
        b.set_dt(dt)
        b.step(nsteps)
        out[prec] = b.state(0)
    for k in range(out["fp64"].shape[0]):
        r = rel(out["mixed"][k], out["fp64"][k])
        assert r <= 1e-6, "field %d rel-L2 %.3e" % (k, r)


@pytest.mark.parametrize("tuning", [
    {"coarse_lds_budget": 24576, "min_tiles": 0},
    {"coarse_lds_budget": 0, "min_tiles": 1024},
    {"coarse_lds_budget": 49152, "min_tiles": 256},
])
def test_tuning_does_not_change_results(tuning):
    """Tile heights and the coarse-kernel split change only how the work is
    launched: FP64 results stay bit-identical (and so equal to the oracle)."""
    specs = [spec("euler", 256, 256, 1.0, 1.0, "closed"),
             spec("euler", 64, 64, 1.0, 1.0, "perio"),
             spec("boussinesq", 128, 64, 2.0, 1.0, "xchannel", diffusion=True, Kdiff=1e-5)]
    out = []
    for t in (None, tuning):
        b = Batch(specs, tuning=t)
        for i in range(len(specs)):
            init_fields(b, i, 20 + i)
        b.set_psi_from_vorticity()
        for i in range(len(specs)):
            b.set_dt(dt_for(b, i, b.diagnostics(i)["maxspeed"]), i)
        b.step(6)
        out.append([b.state(i) for i in range(len(specs))])
        if t is not None:
            got = b.tuning()
            assert got["min_tiles"] == t["min_tiles"]
    for i in range(len(specs)):
        assert (out[0][i] == out[1][i]).all(), "instance %d" % i


def test_retune_between_steps():
    s = [spec("euler", 128, 128, 1.0, 1.0, "closed")]
    ref = Batch(s)
    init_fields(ref, 0, 5)
    ref.set_psi_from_vorticity()
    dt = dt_for(ref, 0, ref.diagnostics(0)["maxspeed"])
    ref.set_dt(dt)
    ref.step(6)
    b = Batch(s)
    init_fields(b, 0, 5)
    b.set_psi_from_vorticity()
    b.set_dt(dt)
    b.step(3)
    b.retune({"coarse_lds_budget": 24576, "min_tiles": 512})
    b.step(3)
    assert (b.state(0) == ref.state(0)).all()


@pytest.mark.parametrize("geom", ["perio", "closed", "xchannel", "disc"])
def test_finest_operator_from_stencil_bits(geom):
    """The finest operator (a constant stencil times masks) is applied from
    one byte of stencil bits per cell, not from five FP64 planes: the
    finest-level smoother's implemented bytes stay near its algorithmic
    bytes (explicit planes add 40 bytes per cell to about 26)."""
    b = Batch([spec("euler", 128, 128, 1.0, 1.0, geom)])
    init_fields(b, 0, 3)
    b.set_psi_from_vorticity()
    b.set_dt(dt_for(b, 0, b.diagnostics(0)["maxspeed"]))
    pr = b.profile("smooth", 0, nsteps=1)
    assert pr["launches"] > 0
    assert pr["impl_bytes"] < 1.5 * pr["alg_bytes"], pr


@pytest.mark.parametrize("order,method,scheme,geom,kind", [
    (1, "parabolic", "RK3_SSP", "closed", "euler"),
    (2, "parabolic", "RK3_SSP", "perio", "euler"),
    (3, "minmax", "RK3", "closed", "euler"),
    (4, "parabolic", "RK3_SSP", "closed", "euler"),
    (5, "minmax", "RK3_SSP", "perio", "euler"),
    (6, "parabolic", "RK3", "perio", "euler"),
    (4, "parabolic", "RK3_SSP", "xchannel", "boussinesq"),
    (3, "parabolic", "RK3", "xchannel", "boussinesq"),
])
def test_engine_schemes_vs_oracle(order, method, scheme, geom, kind):
    """Every advection order and flux splitting and both time schemes on the
    engine (each edge flux once per tile) against the oracle, 12 steps,
    bitwise."""
    nx, Lx = (128, 2.0) if kind == "boussinesq" else (64, 1.0)
    kw = dict(order=order, flux_splitting_method=method, timestepping=scheme)
    if kind == "boussinesq":
        kw.update(diffusion=True, Kdiff=1e-4, gravity=1.0)
    b = Batch([spec(kind, nx, 64, Lx, 1.0, geom, **kw)], precision="fp64")
    fields = init_fields(b, 0, 31)
    b.set_psi_from_vorticity()
    p, g = b.params[0], b.grids[0]
    o = orc.OracleModel(kind, nx, 64, Lx, 1.0, geometry=geom, order=order,
                        flux_splitting_method=method, timestepping=scheme,
                        diffusion=p.diffusion, Kdiff=p.Kdiff, gravity=p.gravity)
    for k, v in fields.items():
        o.get(k)[...] = v
    o.set_psi_from_vorticity()
    d = o.diagnostics()
    assert b.diagnostics(0)["maxspeed"] == d["maxspeed"]
    dt = dt_for(b, 0, d["maxspeed"])
    b.set_dt(dt, 0)
    b.step(12)
    o.run(12, dt)
    s = b.state(0)
    for k, name in enumerate(b.varnames[0]):
        assert (s[k] == o.state[k]).all(), "%s differs (rel %.3e)" % (name, rel(s[k], o.state[k]))


@pytest.mark.parametrize("order,method,scheme", [(4, "parabolic", "RK3_SSP"), (3, "minmax", "RK3"),
                                                 (6, "parabolic", "RK3_SSP")])
def test_mixed_precision_other_schemes(order, method, scheme):
    """Mixed mode with centred and third-order upwind fluxes (the unit-field
    flux of each edge in FP64 for every order) stays within rel-L2 1e-6 of
    the FP64 path after 30 steps."""
    kw = dict(order=order, flux_splitting_method=method, timestepping=scheme)
    out = {}
    for prec in ("fp64", "mixed"):
        b = Batch([spec("euler", 128, 128, 1.0, 1.0, "closed", **kw)], precision=prec)
        init_fields(b, 0, 5)
        b.set_psi_from_vorticity()
        if prec == "fp64":
            dt = dt_for(b, 0, b.diagnostics(0)["maxspeed"])
        b.set_dt(dt)
        b.step(30)
        out[prec] = b.state(0)
    for k in range(out["fp64"].shape[0]):
        r = rel(out["mixed"][k], out["fp64"][k])
        assert r <= 1e-6, "field %d rel-L2 %.3e" % (k, r)


def test_profiled_steps_equal_plain_steps():
    """Kernel timing (Batch.profile) issues every timed launch 8 times back
    to back, so each timed launch must leave the state as one launch does:
    steps run under the timing of every launch kind equal plain steps bit
    for bit (the finest-level smoothing that also adds its result into psi,
    x0 += x, is never repeated)."""
    sp = [spec("euler", 256, 256, 1.0, 1.0, "closed")]
    kinds = ["smooth", "resid_restrict", "rhs", "coarse", "smooth_rr", "up2"]
    out = []
    for profiled in (False, True):
        b = Batch(sp, precision="fp64", tuning={"min_tiles": 1024})
        init_fields(b, 0, 77)
        b.set_psi_from_vorticity()
        b.set_dt(dt_for(b, 0, b.diagnostics(0)["maxspeed"]))
        b.step(1)
        if profiled:
            for k in kinds:
                pr = b.profile(k, -1, 1)
                assert pr["launches"] > 0, k
        else:
            b.step(len(kinds))
        out.append(b.state(0))
        del b
    assert (out[0] == out[1]).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_persistent_multigrid_equals_launch_per_phase(precision):
    """Runs of mid-level smoother and residual/restriction phases as one
    persistent launch with device-wide barriers: bit for bit the
    launch-per-phase program, single instance and mixed batch, with the
    reference schedule and with the refined per-stage solves."""
    specs = [
        spec("euler", 256, 256, 1.0, 1.0, "closed"),
        spec("euler", 128, 128, 1.0, 1.0, "perio"),
        spec("boussinesq", 256, 128, 2.0, 1.0, "xchannel", diffusion=True, Kdiff=1e-5),
    ]
    for fast in (0, 4):
        for group in ([specs[0]], specs):
            print("case persistent", precision, fast, len(group), flush=True)
            out = []
            for persist in (0, 2):
                b = Batch(group, precision=precision,
                          tuning={"min_tiles": 1024, "persist": persist})
                if fast:
                    b.set_fast_solve(fast, 1e-10)
                for i in range(len(group)):
                    init_fields(b, i, 50 + i)
                b.set_psi_from_vorticity()
                for i in range(len(group)):
                    b.set_dt(dt_for(b, i, b.diagnostics(i)["maxspeed"]), i)
                b.step(8)
                out.append((b.info(0)["launches_per_step"], [b.state(i) for i in range(len(group))],
                            b.profile("persist", -1, 1)["launches"]))
                assert b.tuning()["persist"] == persist
                del b
            # the persistent program runs its merged phases in k_persist, in
            # at most the per-phase program's launches (whose F-cycle descent
            # of b already takes two levels a launch)
            assert out[1][2] > 0 and out[0][2] == 0, "persistent launches %d / %d" % (
                out[1][2], out[0][2])
            assert out[1][0] <= out[0][0], "persistent program has %d launches, per-phase %d" % (
                out[1][0], out[0][0])
            for a, m in zip(out[0][1], out[1][1]):
                assert (a == m).all()
