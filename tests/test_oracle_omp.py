"""The OpenMP build of the oracle (the multi-core CPU baseline) computes the
same bits as the serial build: only its row loops run in parallel."""
import os

import numpy as np

import fluid2d_amd as F
from fluid2d_amd import workloads
from oracle import oracle as orc


def run(variant, kind, geom, nx, ny, Lx, steps):
    orc.use(variant)
    try:
        o = orc.OracleModel(kind, nx, ny, Lx, 1.0, geometry=geom)
        g = F.Grid(F.Param(modelname=kind, nx=nx, ny=ny, Lx=Lx, geometry=geom))
        o.get("vorticity")[...] = workloads.turbulence(g, 3)
        o.set_psi_from_vorticity()
        d = o.diagnostics()
        log = o.run(steps, 0.5 * g.dx / d["maxspeed"])
        return o.state.copy(), list(log)
    finally:
        orc.use("serial")


def test_omp_build_is_bitwise_serial():
    orc.build(target="liboracle_omp.so")
    os.environ.setdefault("OMP_NUM_THREADS", "4")
    for kind, geom, nx, Lx in (("euler", "closed", 64, 1.0), ("boussinesq", "xchannel", 128, 2.0)):
        a, la = run("serial", kind, geom, nx, 64, Lx, 3)
        b, lb = run("omp", kind, geom, nx, 64, Lx, 3)
        assert la == lb
        assert (a == b).all(), kind
