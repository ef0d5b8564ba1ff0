"""The Helmholtz operator of the device multigrid (core/gmg/hierarchy.py:80-84,
``Gmg(param)`` with 'qgoperator' and 'Rd') against the oracle: every level's
operator bit for bit, twoVcycle and solve bit for bit (periodic and closed),
and at 1024 x 1024 the converged periodic solve against the FFT solution of
the shifted 9-point symbol (a size-independent property)."""
import numpy as np
import pytest

import fluid2d_amd as F
from fluid2d_amd.device import DeviceArray
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
NH = 3


def lam(a, b, dxdy):
    return (-3 + np.cos(a) + np.cos(b) + np.cos(a) * np.cos(b)) / dxdy


@pytest.mark.parametrize("geom", ["perio", "closed"])
def test_helmholtz_levels_cycles_and_solve_match_oracle(geom):
    nx, ny, Lx, Rd = 64, 32, 2.0, 0.1
    print("case helmholtz", geom, flush=True)
    dx, dy = Lx / nx, 1.0 / ny
    o = orc.OracleModel("euler", nx, ny, Lx, 1.0, geometry=geom)
    og = orc.OracleGmg(nx, ny, dx, dy, o.mskp, Rd=Rd)
    g = F.Gmg({"n": nx, "m": ny, "nh": 3, "dx": dx, "dy": dy, "qgoperator": True, "Rd": Rd},
              o.mskp)
    assert g.nlevs == og.nlevs
    for lev in range(og.nlevs):
        assert (g.level(lev, "A").get() == og.A(lev)).all(), lev
    rng = np.random.default_rng(27)
    shape = (ny + 6, nx + 6)
    b = rng.standard_normal(shape)
    x = rng.standard_normal(shape) * 1e-2
    xo, xd = x.copy(), DeviceArray.from_host(x)
    og.twoVcycle(xo, b)
    g.twoVcycle(xd, DeviceArray.from_host(b))
    assert (xd.get() == xo).all()
    x0 = np.zeros(shape)
    xo, xd = x0.copy(), DeviceArray.from_host(x0)
    no, ro = og.solve(xo, b, maxite=6, tol=1e-11)
    nd, rd = g.solve(xd, DeviceArray.from_host(b), {"maxite": 6, "tol": 1e-11})
    assert no == nd
    assert abs(rd - ro) <= 1e-10 * ro
    assert (xd.get() == xo).all()


def test_helmholtz_full_size_solve_is_the_fft_solution():
    n, Rd = 1024, 0.02
    dx = 1.0 / n
    print("case helmholtz fft", n, flush=True)
    mskp = orc.OracleModel("euler", n, n, geometry="perio").mskp
    assert (mskp[NH:-NH, NH:-NH] == 1).all()
    g = F.Gmg({"n": n, "m": n, "nh": 3, "dx": dx, "dy": dx, "qgoperator": True, "Rd": Rd}, mskp)
    rng = np.random.default_rng(28)
    qi = rng.standard_normal((n, n))
    q = np.zeros((n + 2 * NH, n + 2 * NH))
    q[NH:-NH, NH:-NH] = qi
    orc.fillhalo(q, NH)
    xd = DeviceArray.from_host(np.zeros_like(q))
    nite, res = g.solve(xd, DeviceArray.from_host(q), {"maxite": 40, "tol": 1e-12})
    assert res < 1e-12 and nite < 40
    a = 2 * np.pi * np.fft.fftfreq(n)
    L = lam(a[None, :], a[:, None], dx * dx) - 1.0 / Rd ** 2
    xf = np.real(np.fft.ifft2(np.fft.fft2(qi) / L))
    x = xd.get()[NH:-NH, NH:-NH]
    assert np.abs(x - xf).max() <= 1e-10 * np.abs(xf).max()
