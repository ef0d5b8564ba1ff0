"""History, diagnostic and restart files (core/output.py, core/restart.py)
written by the Output schedule and fluid2d_amd.ncio, read back with
scipy's NetCDF reader.  Host only: the state is a NumPy array."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import fluid2d_amd as F
from fluid2d_amd.fluid2d import Output
from fluid2d_amd.ncio import Restart, read_restart, write_restart

NAMES = ["vorticity", "psi", "u", "v", "source"]


def test_history_and_diag_files(tmp_path):
    p = F.Param(nx=12, ny=8, Lx=1.5, geometry="closed", datadir=str(tmp_path), expname="exp",
                var_to_save=["vorticity", "psi"], freq_his=0.1, freq_diag=0.05)
    g = F.Grid(p)
    out = Output(p, g, NAMES)
    rng = np.random.default_rng(1)
    states, diags = [], []
    t, kt = 0.0, 0
    for k in range(8):
        x = rng.standard_normal((5, 14, 18))
        d = {"maxspeed": float(k), "ke": 0.5 * k, "dkedt": -1.0}
        states.append(x)
        diags.append(d)
        out.do({"diag": d, "his": (lambda x=x: x)}, t, kt)
        t += 0.04
        kt += 1
    his_t = [r[0] for r in out.his]
    with netcdf_file(str(tmp_path / "exp" / "exp_his.nc"), "r", mmap=False) as nc:
        assert nc.dimensions["x"] == 12 and nc.dimensions["y"] == 8
        assert np.allclose(nc.variables["t"][:], his_t)
        kts = nc.variables["kt"][:]
        for r, k in enumerate(kts):
            for name in ("vorticity", "psi"):
                want = states[k][NAMES.index(name)][3:-3, 3:-3].astype(np.float32)
                assert (nc.variables[name][r] == want).all()
        assert "u" not in nc.variables
        assert (nc.variables["msk"][:] == g.msk[3:-3, 3:-3]).all()
        assert np.allclose(nc.variables["x"][:], g.xr[0, 3:-3])
        assert nc.nx == 12 and nc.adaptable_dt == 1 and nc.geometry == b"closed"
    with netcdf_file(str(tmp_path / "exp" / "exp_diag.nc"), "r", mmap=False) as nc:
        kt_d = list(nc.variables["kt"][:])
        assert len(kt_d) == len(out.diag)
        for r, k in enumerate(kt_d):
            assert nc.variables["ke"][r] == np.float32(diags[k]["ke"])
        assert set(["maxspeed", "ke", "dkedt"]) <= set(nc.variables)


def test_var_to_save_is_checked(tmp_path):
    p = F.Param(nx=8, ny=8, datadir=str(tmp_path), var_to_save=["vorticity", "pressure"])
    with pytest.raises(ValueError, match="pressure"):
        Output(p, F.Grid(p), NAMES)


def test_restart_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((5, 14, 18))
    f = str(tmp_path / "exp_03_restart_000.nc")
    write_restart(f, 10.0, 3.25, 0.01, 325, 3.3, 3.5, NAMES, x)
    y = np.zeros_like(x)
    got = read_restart(f, NAMES, y)
    assert got == (10.0, 3.25, 0.01, 325, 3.3, 3.5)
    assert (y == x).all()


def test_last_restart_index(tmp_path):
    for k in (0, 1, 7):
        (tmp_path / ("exp_%02i_restart_000.nc" % k)).write_bytes(b"")
    r = Restart.__new__(Restart)
    r.expdir, r.expname = str(tmp_path), "exp"
    r.get_lastrestart()
    assert r.lastrestart == 7
