"""No-slip walls, islands and the loop's momentum enforcement on the batched
engine (core/operators.py:157-185, :245-290, :421-498; core/island.py;
core/fluid2d.py:352-358): Fluid2d's engine-backed model against the
operator-level model (itself matched to the oracle in test_gpu_model.py),
bit for bit where both sum in the same order; instances with walls alone
and in a mixed batch; mixed precision against FP64."""
import numpy as np
import pytest

import fluid2d_amd as F
from fluid2d_amd import workloads
from fluid2d_amd.batch import Batch

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    d = np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel())
    nb = np.linalg.norm(np.asarray(b).ravel())
    return d / nb if nb > 0 else d


def karman(p, g):
    """A disc island in a channel whose walls carry psi = +-psi0/2."""
    nh = 3
    r = np.sqrt((g.xr - 0.5) ** 2 + (g.yr - 0.5) ** 2)
    idx = np.where(r <= 0.12)
    g.msk[idx] = 0
    g.msknoslip[idx] = 0
    g.island.add(idx, 0.0)
    low = np.zeros_like(g.msk)
    low[:nh, :] = 1
    g.island.add(np.where(low == 1), 0.1)
    high = np.zeros_like(g.msk)
    high[-nh:-1, :] = 1
    g.island.add(np.where(high == 1), -0.1)
    g.msknoslip[:nh, :] = 1
    g.msknoslip[-nh:, :] = 1


def run_fluid2d(engine, kind, nx, ny, Lx, geom, seed, nsteps, islands=None, loop=False, **kw):
    p = F.Param(modelname=kind, nx=nx, ny=ny, Lx=Lx, Ly=1.0, geometry=geom, engine=engine, **kw)
    g = F.Grid(p)
    if islands is not None:
        islands(p, g)
    f = F.Fluid2d(p, g)
    assert isinstance(f.model, F.EngineModel) == engine
    if islands is None:
        f.model.var.get("vorticity")[...] = workloads.turbulence(g, seed) * g.msk
    if kind == "boussinesq":
        rng = np.random.default_rng(seed + 1)
        b = (rng.standard_normal(g.msk.shape) * 0.1) * g.msk
        g.fill_halo(b)
        f.model.var.get("buoyancy")[...] = b
    f.model.set_psi_from_vorticity()
    d0 = f.model.diagnostics(f.model.var, 0)
    if loop:
        f.loop(nsteps=nsteps)
        log = []
    else:
        log = list(f.model.run(nsteps, 0.5 * g.dx / max(d0["maxspeed"], 0.1)))
    return f.model.var.state.copy(), log, f


@pytest.mark.parametrize("kind,geom", [("euler", "closed"), ("euler", "disc"),
                                       ("boussinesq", "closed"), ("boussinesq", "xchannel")])
def test_noslip_engine_equals_operator_model(kind, geom):
    """No-slip walls after each step: the engine-backed model and the
    operator-level model take 15 steps to the same bits."""
    nx = 128 if kind == "boussinesq" else 64
    out = [run_fluid2d(e, kind, nx, 64, nx / 64, geom, 31, 15, noslip=True) for e in (True, False)]
    (a, la, fa), (b, lb, _) = out
    print("case noslip", kind, geom, flush=True)
    assert la == lb
    if kind == "euler":
        src = a[fa.model.var.varname_list.index("source")]
        assert np.abs(src[3:-3, 3:-3]).max() > 0
    assert (a == b).all()


def test_noslip_enforce_engine_close_to_operator_model():
    """With enforce_momentum the source's two momentum sums are folded in
    another order than the operator-level kernels fold them: the states
    agree to rounding (rel-L2 <= 1e-12 per field after 15 steps)."""
    out = [run_fluid2d(e, "euler", 64, 64, 1.0, "closed", 32, 15, noslip=True,
                       enforce_momentum=True) for e in (True, False)]
    (a, la, f), (b, lb, _) = out
    assert la == lb
    for k, name in enumerate(f.model.var.varname_list):
        assert rel_l2(a[k], b[k]) <= 1e-12, name


def test_islands_engine_equals_operator_model():
    """Islands with no-slip walls (the flow past a disc between walls at
    +-psi0/2, from rest): engine and operator-level model, 20 steps,
    bit for bit, psi on the islands held at their values."""
    out = [run_fluid2d(e, "euler", 128, 64, 2.0, "xchannel", 33, 20, islands=karman, noslip=True,
                       isisland=True) for e in (True, False)]
    (a, la, f), (b, lb, _) = out
    assert la == lb
    assert abs(a[1][3:-3, 3:-3]).max() > 0.05  # the imposed through-flow
    assert (a == b).all()


def test_momentum_loop_engine_close_to_operator_model():
    """The loop's momentum enforcement (core/fluid2d.py:270-272, 352-358)
    in the engine: the same loop as on the operator-level model within
    rounding (the diagnostics that give px, py are summed per tile), and the
    momentum stays at rounding level."""
    out = [run_fluid2d(e, "euler", 64, 64, 1.0, "closed", 34, 4, loop=True, adaptable_dt=True,
                       cfl=0.8, tend=1e9, enforce_momentum=True) for e in (True, False)]
    (a, _, f), (b, _, _) = out
    for k, name in enumerate(f.model.var.varname_list):
        assert rel_l2(a[k], b[k]) <= 1e-11, name
    d = f.model.diags
    w0 = workloads.turbulence(f.grid, 34) * f.grid.msk
    p0 = max(abs(f.grid.domain_integration(w0 * f.grid.xr0)),
             abs(f.grid.domain_integration(w0 * f.grid.yr0))) / f.grid.area
    assert abs(d["px"]) < 1e-10 * p0 and abs(d["py"]) < 1e-10 * p0


def _specs():
    specs = []
    for kind, geom, ns in (("euler", "closed", True), ("euler", "closed", False),
                           ("boussinesq", "closed", True)):
        nx = 128 if kind == "boussinesq" else 64
        p = F.Param(modelname=kind, nx=nx, ny=64, Lx=nx / 64, geometry=geom, noslip=ns)
        specs.append((p, F.Grid(p)))
    return specs


def _init(b, specs, seed):
    for i, (p, g) in enumerate(specs):
        b.set_field(i, "vorticity", workloads.turbulence(g, seed + i) * g.msk)
        if p.modelname == "boussinesq":
            rng = np.random.default_rng(seed + 10 + i)
            bb = (rng.standard_normal(g.msk.shape) * 0.1) * g.msk
            g.fill_halo(bb)
            b.set_field(i, "buoyancy", bb)
    b.set_psi_from_vorticity()
    for i, (p, g) in enumerate(specs):
        b.set_dt(0.5 * g.dx / b.diagnostics(i)["maxspeed"], i)


def test_batch_walls_alone_equals_in_mix():
    """A batch of an Euler instance with walls, one without and a Boussinesq
    instance with walls: each equals itself run alone, bit for bit."""
    specs = _specs()
    b = Batch(specs, precision="fp64")
    _init(b, specs, 40)
    b.step(12)
    for i, (p, g) in enumerate(specs):
        a = Batch([(p, g)], precision="fp64")
        a.set_field(0, "vorticity", workloads.turbulence(g, 40 + i) * g.msk)
        if p.modelname == "boussinesq":
            rng = np.random.default_rng(40 + 10 + i)
            bb = (rng.standard_normal(g.msk.shape) * 0.1) * g.msk
            g.fill_halo(bb)
            a.set_field(0, "buoyancy", bb)
        a.set_psi_from_vorticity()
        a.set_dt(0.5 * g.dx / a.diagnostics(0)["maxspeed"], 0)
        a.step(12)
        print("case batch walls instance", i, flush=True)
        assert (a.state(0) == b.state(i)).all()


def test_batch_walls_mixed_precision():
    """Mixed precision with walls (FP32 V-cycles in the correction form of
    the state's fast inversion): within 1e-6 of the FP64 batch after 10
    steps."""
    specs = _specs()
    out = []
    for prec in ("fp64", "mixed"):
        b = Batch(specs, precision=prec)
        _init(b, specs, 50)
        b.step(10)
        out.append([b.state(i) for i in range(len(specs))])
    for i in range(len(specs)):
        for k in (0, 1):
            r = rel_l2(out[1][i][k], out[0][i][k])
            print("case mixed walls", i, k, r, flush=True)
            assert r <= 1e-6
