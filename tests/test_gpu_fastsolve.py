"""Per-stage inversions by gmg.solve (the alternative core/operators.py:452-455
keeps commented out) on the batched engine.

* FP64: equal to the oracle in the same mode (rel-L2 <= 1e-12 per field;
  the kernels keep the reference's operation order), Euler and Boussinesq.
* Mixed precision is FP64 iterative refinement of FP32 F-cycles: within
  rel-L2 1e-8 of the FP64 run in the same mode.
* maxite = 0 restores the reference's twoVcycle path bit for bit.
* boussinesq-rt-2048x1024 (FP32 advection and V-cycles, FP64 residual
  refinement to 1e-10), the configuration's 1000 steps: mixed against FP64
  in the same mode, every field within rel-L2 1e-6.
"""
import numpy as np
import pytest

from fluid2d_amd import workloads
from fluid2d_amd.batch import Batch
from fluid2d_amd.grid import Grid
from fluid2d_amd.param import Param
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
MAXITE, TOL = 6, 1e-10


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


CASES = [
    dict(modelname="boussinesq", nx=256, ny=128, Lx=2.0, Ly=1.0, geometry="xchannel",
         diffusion=True, Kdiff=1e-5),
    dict(modelname="euler", nx=128, ny=128, Lx=1.0, Ly=1.0, geometry="closed"),
]


def start(kw, precision, seed=41, fast=True):
    p = Param(adaptable_dt=False, **kw)
    g = Grid(p)
    b = Batch([(p, g)], precision=precision)
    fields = {"vorticity": workloads.turbulence(g, seed)}
    if p.modelname == "boussinesq":
        fields["buoyancy"] = workloads.rayleigh_taylor(g, seed + 1)
    for k, v in fields.items():
        b.set_field(0, k, v)
    if fast:
        b.set_fast_solve(MAXITE, TOL)
    b.set_psi_from_vorticity()
    return p, g, b, fields


@pytest.mark.parametrize("kw", CASES, ids=["boussinesq", "euler"])
def test_fast_solve_fp64_vs_oracle(kw):
    print("case fast solve fp64", kw["modelname"], flush=True)
    p, g, b, fields = start(kw, "fp64")
    o = orc.OracleModel(p.modelname, g.nx, g.ny, g.Lx, g.Ly, geometry=g.geometry,
                        diffusion=p.diffusion, Kdiff=p.Kdiff, gravity=p.gravity)
    for k, v in fields.items():
        o.get(k)[...] = v
    o.set_fast_solve(MAXITE, TOL)
    o.set_psi_from_vorticity()
    d = o.diagnostics()
    assert b.diagnostics(0)["maxspeed"] == d["maxspeed"]
    dt = 0.5 * g.dx / max(d["maxspeed"], 1.0)
    b.set_dt(dt)
    b.step(10)
    o.run(10, dt)
    s = b.state(0)
    for k, name in enumerate(o.varname_list):
        r = rel(s[k], o.state[k])
        assert r <= 1e-12, "%s rel-L2 %.3e" % (name, r)


@pytest.mark.parametrize("kw", CASES, ids=["boussinesq", "euler"])
def test_fast_solve_mixed_vs_fp64(kw):
    print("case fast solve mixed", kw["modelname"], flush=True)
    _, g, b64, _ = start(kw, "fp64")
    dt = 0.5 * g.dx / max(b64.diagnostics(0)["maxspeed"], 1.0)
    b64.set_dt(dt)
    b64.step(20)
    _, _, bm, _ = start(kw, "mixed")
    bm.set_dt(dt)
    bm.step(20)
    s64, sm = b64.state(0), bm.state(0)
    for k in range(s64.shape[0]):
        if np.any(s64[k]):
            r = rel(sm[k], s64[k])
            assert r <= 1e-8, "field %d rel-L2 %.3e" % (k, r)


def test_fast_solve_off_is_the_reference_path():
    kw = CASES[0]
    _, g, ref, _ = start(kw, "fp64", fast=False)
    _, _, b, _ = start(kw, "fp64")
    b.set_fast_solve(0)
    for x in (ref, b):
        x.set_dt(0.3 * g.dx)
        x.step(5)
    assert (ref.state(0) == b.state(0)).all()


def test_boussinesq_rt_refined_1000_steps_mixed_vs_fp64():
    name = "boussinesq-rt-2048x1024"
    print("case", name, "refined 1000 steps", flush=True)
    out = {}
    dt = None
    for prec in ("fp64", "mixed"):
        p = workloads.make_param(name)
        g = Grid(p)
        b = Batch([(p, g)], precision=prec)
        for k, v in workloads.initial_fields(name, g).items():
            b.set_field(0, k, v)
        b.set_fast_solve(MAXITE, TOL)
        b.set_psi_from_vorticity()
        if dt is None:
            dt = workloads.time_step(name, g, b.diagnostics(0)["maxspeed"])
        b.set_dt(dt)
        b.step(1000)
        out[prec] = b.state(0)
        nite, res, div = b.solve_info(0)
        print("  %s: last solve %d F-cycles, res %.2e" % (prec, nite, res), flush=True)
        assert not div
        del b
    for k in range(out["fp64"].shape[0]):
        if np.any(out["fp64"][k]):
            r = rel(out["mixed"][k], out["fp64"][k])
            print("  field %d rel-L2 %.3e" % (k, r), flush=True)
            assert r <= 1e-6, "field %d rel-L2 %.3e" % (k, r)
