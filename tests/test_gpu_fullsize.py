"""The device path at the full benchmark sizes.

- euler-closed-1024 (1024 x 1024 closed box, k^-3 turbulence) and
  boussinesq-rt-2048x1024 (2048 x 1024 channel, Rayleigh-Taylor start): the
  FP64 batched engine steps twice from the workload's start and matches the
  oracle stepped with the same dt, every field bit for bit (and so within
  rel-L2 1e-12; the oracle finishes these two steps in a few seconds); the 1024 x 1024 case
  also with the benchmark's launch shapes (tiles for at least 1024
  workgroups per finest-level launch, the coarse kernel's LDS budget).
- The converged multigrid solve on the 1024 x 1024 closed box equals the
  DST-I solution of the 9-point Laplacian (psi = 0 on the boundary corners)
  within 1e-10 of its maximum: a size-independent check that needs no
  oracle run.
"""
import numpy as np
import pytest
from scipy.fft import dstn, idstn

import fluid2d_amd as F
from fluid2d_amd import workloads
from fluid2d_amd.batch import Batch
from fluid2d_amd.device import DeviceArray
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
NH = 3


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


BENCH_TUNING = {"coarse_lds_budget": 159744, "min_tiles": 1024, "coarse_wave_cells": 0,
                "persist": 0}


@pytest.mark.parametrize("name,tuning", [("euler-closed-1024", None),
                                         ("euler-closed-1024", BENCH_TUNING),
                                         ("boussinesq-rt-2048x1024", None)],
                         ids=["euler-closed-1024", "euler-closed-1024-bench-tiles",
                              "boussinesq-rt-2048x1024"])
def test_engine_full_size_vs_oracle(name, tuning):
    print("case", name, tuning, flush=True)
    p = workloads.make_param(name)
    g = F.Grid(p)
    b = Batch([(p, g)], tuning=tuning)
    if tuning is not None:
        assert b.tuning()["min_tiles"] == tuning["min_tiles"]
    o = orc.OracleModel(p.modelname, g.nx, g.ny, g.Lx, g.Ly, geometry=g.geometry,
                        diffusion=p.diffusion, Kdiff=p.Kdiff, gravity=p.gravity)
    for k, v in workloads.initial_fields(name, g).items():
        b.set_field(0, k, v)
        o.get(k)[...] = v
    b.set_psi_from_vorticity()
    o.set_psi_from_vorticity()
    d = b.diagnostics(0)
    assert d["maxspeed"] == o.diagnostics()["maxspeed"]
    dt = workloads.time_step(name, g, d["maxspeed"])
    b.set_dt(dt)
    b.step(2)
    o.run(2, dt)
    s = b.state(0)
    for k, vname in enumerate(o.varname_list):
        assert rel(s[k], o.state[k]) <= 1e-12, vname
        # the FP64 engine reproduces the oracle's operations in order
        assert np.array_equal(s[k], o.state[k]), vname
    if p.modelname == "euler":
        assert b.diagnostics(0)["maxspeed"] == pytest.approx(o.diagnostics()["maxspeed"],
                                                             rel=1e-12)


def test_converged_solve_full_size_is_the_dst_solution():
    n = 1024
    dx = 1.0 / n
    print("case gmg solve closed", n, flush=True)
    mskp = orc.OracleModel("euler", n, n, geometry="closed").mskp
    s = np.s_[NH:NH + n - 1, NH:NH + n - 1]
    assert (mskp[s] == 1).all() and mskp.sum() == (n - 1) ** 2
    gm = F.Gmg({"n": n, "m": n, "nh": NH, "dx": dx, "dy": dx}, mskp)
    rng = np.random.default_rng(9)
    bi = rng.standard_normal((n - 1, n - 1))
    b = np.zeros(mskp.shape)
    b[s] = bi
    xd = DeviceArray.from_host(np.zeros_like(b))
    nite, res = gm.solve(xd, DeviceArray.from_host(b), {"maxite": 30, "tol": 1e-12})
    assert res < 1e-12 and nite < 30
    a = np.pi * np.arange(1, n) / n
    lam = (-3 + np.cos(a)[None, :] + np.cos(a)[:, None]
           + np.cos(a)[None, :] * np.cos(a)[:, None]) / (dx * dx)
    xf = idstn(dstn(bi, type=1) / lam, type=1)
    x = xd.get()
    assert np.abs(x[s] - xf).max() <= 1e-10 * np.abs(xf).max()
