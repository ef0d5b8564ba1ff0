"""Restart drop-in (core/restart.py) and the history files of a device run:
two jobs of Fluid2d.loop under Restart against the oracle, which replays
each job's dt sequence from the same start state (the second from the
restart file), bitwise."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import fluid2d_amd as F
from fluid2d_amd import workloads
from fluid2d_amd.ncio import Restart, read_restart
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
N = 64


def job(tmp_path, w0=None):
    p = F.Param(modelname="euler", nx=N, ny=N, geometry="closed", adaptable_dt=True, cfl=0.8,
                tend=0.02, freq_his=0.01, freq_diag=0.005, datadir=str(tmp_path), expname="rs",
                var_to_save=["vorticity", "psi"])
    g = F.Grid(p)
    f = F.Fluid2d(p, g)
    if w0 is not None:
        f.model.var.get("vorticity")[...] = w0
        f.model.set_psi_from_vorticity()
    dts = []
    step = f.model.step
    f.model.step = lambda t, dt: (dts.append(dt), step(t, dt))
    Restart(p, g, f)
    return f, dts


def oracle_replay(start, dts):
    o = orc.OracleModel("euler", N, N, geometry="closed")
    o.state[...] = start
    o.diagnostics()
    for dt in dts:
        o.step(dt)
    return o.state.copy()


def test_two_restart_jobs_match_oracle(tmp_path):
    g = F.Grid(F.Param(nx=N, ny=N, geometry="closed"))
    w0 = workloads.turbulence(g, 14)
    # job 0, from scratch
    f0, dts0 = job(tmp_path, w0)
    o = orc.OracleModel("euler", N, N, geometry="closed")
    o.get("vorticity")[...] = w0
    o.set_psi_from_vorticity()
    start0 = o.state.copy()
    names = f0.model.var.varname_list
    r0 = np.zeros_like(start0)
    hdr0 = read_restart(str(tmp_path / "rs" / "rs_00_restart_000.nc"), names, r0)
    assert abs(hdr0[1] - 0.02) < 1e-14 and hdr0[3] == len(dts0)
    assert (r0 == oracle_replay(start0, dts0)).all()
    # job 1 resumes from restart 0 (fresh tendencies, as in the reference)
    f1, dts1 = job(tmp_path)
    assert abs(f1.t - 0.04) < 1e-14 and f1.kt == len(dts0) + len(dts1)
    r1 = np.zeros_like(start0)
    read_restart(str(tmp_path / "rs" / "rs_01_restart_000.nc"), names, r1)
    assert (r1 == oracle_replay(r0, dts1)).all()
    assert (f1.model.var.state == r1).all()
    # history files of both jobs
    with netcdf_file(str(tmp_path / "rs" / "rs_00_his_000.nc"), "r", mmap=False) as nc:
        assert np.allclose(nc.variables["t"][:], [0.0, 0.01, 0.02], atol=1e-7)
        assert (nc.variables["vorticity"][-1] == r0[0, 3:-3, 3:-3].astype(np.float32)).all()
    with netcdf_file(str(tmp_path / "rs" / "rs_01_his_000.nc"), "r", mmap=False) as nc:
        assert np.allclose(nc.variables["t"][:], [0.03, 0.04], atol=1e-7)
        assert (nc.variables["psi"][-1] == r1[1, 3:-3, 3:-3].astype(np.float32)).all()
    with netcdf_file(str(tmp_path / "rs" / "rs_01_diag.nc"), "r", mmap=False) as nc:
        assert "ke" in nc.variables and "maxspeed" in nc.variables
        kt = nc.variables["kt"][:]
        assert len(kt) > 0 and kt[0] > len(dts0) and kt[-1] <= f1.kt
