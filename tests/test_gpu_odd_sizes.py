"""Grids whose sizes are not multiples of the 64-column tiles: the clipped
(general) forms of the smoother, the residual/restriction, the fused
down-leg and the two-level up-leg run on partial tiles, against the oracle
at the FP64 bar (rel-L2 <= 1e-12 per field; 96 x 96 closed, 160 x 96
periodic, 384 x 192 channel with Boussinesq diffusion)."""
import numpy as np
import pytest

from fluid2d_amd.batch import Batch
from tests.test_gpu_engine import TOL, dt_for, init_fields, oracle_for, rel, spec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [
    ("euler", 96, 96, 1.0, 1.0, "closed", 20),
    ("euler", 160, 96, 1.0, 0.6, "perio", 20),
    ("boussinesq", 384, 192, 2.0, 1.0, "xchannel", 12),
])
def test_odd_sizes_vs_oracle(case):
    kind, nx, ny, Lx, Ly, geom, nsteps = case
    kw = dict(diffusion=True, Kdiff=1e-5) if kind == "boussinesq" else {}
    b = Batch([spec(kind, nx, ny, Lx, Ly, geom, **kw)])
    f = init_fields(b, 0, 5)
    b.set_psi_from_vorticity()
    o = oracle_for(b, 0, f)
    print("case odd size", kind, nx, ny, geom, b.info(0), flush=True)
    s0 = b.state(0)
    for k in range(s0.shape[0]):
        assert rel(s0[k], o.state[k]) <= TOL, ("init", k)
    dt = dt_for(b, 0, b.diagnostics(0)["maxspeed"])
    b.set_dt(dt)
    b.step(nsteps)
    o.run(nsteps, dt)
    s = b.state(0)
    for k in range(s.shape[0]):
        r = rel(s[k], o.state[k])
        print("field", k, "rel-L2", r, "bitwise", bool((s[k] == o.state[k]).all()), flush=True)
        assert r <= TOL, "field %d rel-L2 %.3e after %d steps" % (k, r, nsteps)
    if kind == "euler":
        assert b.solve_info(0)[0] == o.last_solve()[0]


def test_odd_sizes_alone_equals_in_mix():
    """The three odd grids and a 64^2 box in one batch, in mixed precision:
    each equals itself run alone."""
    specs = [spec("euler", 96, 96, 1.0, 1.0, "closed"), spec("euler", 160, 96, 1.0, 0.6, "perio"),
             spec("boussinesq", 384, 192, 2.0, 1.0, "xchannel", diffusion=True, Kdiff=1e-5),
             spec("euler", 64, 64, 1.0, 1.0, "closed")]
    mix = Batch(specs, precision="mixed")
    fields = [init_fields(mix, i, 20 + i) for i in range(len(specs))]
    mix.set_psi_from_vorticity()
    dts = [dt_for(mix, i, mix.diagnostics(i)["maxspeed"]) for i in range(len(specs))]
    for i, dt in enumerate(dts):
        mix.set_dt(dt, i)
    mix.step(8)
    for i in range(len(specs)):
        alone = Batch([specs[i]], precision="mixed")
        for k, v in fields[i].items():
            alone.set_field(0, k, v)
        alone.set_psi_from_vorticity()
        alone.set_dt(dts[i])
        alone.step(8)
        assert (alone.state(0) == mix.state(i)).all(), "instance %d" % i
