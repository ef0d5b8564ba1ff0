"""The oracle's Helmholtz operator (core/gmg/hierarchy.py:80-84, the
quasi-geostrophic ``qgoperator``): after the hierarchy is built, every
level's centre coefficient becomes A5 - 1/Rd^2 wherever A5 is nonzero.

Pinned by what the mathematics fixes:

- the converged solve of (A - 1/Rd^2) psi = q equals the FFT solution with
  the symbol (-3 + cos a + cos b + cos a cos b)/(dx dy) - 1/Rd^2 on a
  periodic domain (the shifted operator is not singular: no mean is
  removed) and the DST-I solution in a closed box;
- a dropped shift, a shift with the wrong sign, or one scaled by the grid
  spacing fails those comparisons by orders of magnitude (Rd is chosen so
  that 1/Rd^2 is comparable to the operator's small eigenvalues);
- the shift lands on every level, halo included, exactly where the
  unshifted centre is nonzero, and nowhere else (compared against the same
  hierarchy built without it).  The coarse levels' shift is the reference's
  definition (the same constant on every level); it changes the cycles'
  iterates but not the converged solution, so beyond this placement check it
  is pinned by the reference text only.
"""
import numpy as np
import pytest
from scipy.fft import dstn, idstn

from oracle import oracle as orc

NH = 3


def lam(a, b, dxdy):
    return (-3 + np.cos(a) + np.cos(b) + np.cos(a) * np.cos(b)) / dxdy


def hierarchy(geom, nx, ny, Lx, Rd):
    o = orc.OracleModel("euler", nx, ny, Lx=Lx, geometry=geom)
    dx, dy = Lx / nx, 1.0 / ny
    return o, orc.OracleGmg(nx, ny, dx, dy, o.mskp, Rd=Rd), dx, dy


@pytest.mark.parametrize("nx,ny,Lx,Rd", [(32, 32, 1.0, 0.1), (64, 32, 2.0, 0.05)])
def test_periodic_helmholtz_solve_is_the_fft_solution(nx, ny, Lx, Rd):
    o, g, dx, dy = hierarchy("perio", nx, ny, Lx, Rd)
    rng = np.random.default_rng(17)
    qi = rng.standard_normal((ny, nx))  # no zero-mean condition: the operator is invertible
    q = np.zeros(o.mskp.shape)
    q[NH:-NH, NH:-NH] = qi
    orc.fillhalo(q, NH)
    x = np.zeros_like(q)
    nite, res = g.solve(x, q, maxite=30, tol=1e-12)
    assert res < 1e-12 and nite < 30
    a = 2 * np.pi * np.fft.fftfreq(nx)
    c = 2 * np.pi * np.fft.fftfreq(ny)
    L = lam(a[None, :], c[:, None], dx * dy) - 1.0 / Rd ** 2
    xf = np.real(np.fft.ifft2(np.fft.fft2(qi) / L))
    xi = x[NH:-NH, NH:-NH]
    assert np.abs(xi - xf).max() <= 1e-11 * np.abs(xf).max()
    # the mean mode is -Rd^2 times the mean of q (a constant is no longer free)
    assert xi.mean() == pytest.approx(-(Rd ** 2) * qi.mean(), rel=1e-9, abs=1e-13)


def test_closed_box_helmholtz_solve_is_the_dst_solution():
    nx, ny, Lx, Rd = 32, 32, 1.0, 0.08
    o, g, dx, dy = hierarchy("closed", nx, ny, Lx, Rd)
    s = np.s_[NH:NH + ny - 1, NH:NH + nx - 1]
    rng = np.random.default_rng(18)
    qi = rng.standard_normal((ny - 1, nx - 1))
    q = np.zeros(o.mskp.shape)
    q[s] = qi
    x = np.zeros_like(q)
    nite, res = g.solve(x, q, maxite=30, tol=1e-12)
    assert res < 1e-12 and nite < 30
    a = np.pi * np.arange(1, nx) / nx
    c = np.pi * np.arange(1, ny) / ny
    L = lam(a[None, :], c[:, None], dx * dy) - 1.0 / Rd ** 2
    xf = idstn(dstn(qi, type=1) / L, type=1)
    assert np.abs(x[s] - xf).max() <= 1e-11 * np.abs(xf).max()


@pytest.mark.parametrize("geom", ["perio", "closed"])
def test_shift_on_every_level_where_the_centre_is_nonzero(geom):
    nx = ny = 64
    Rd = 0.2
    o, g, dx, dy = hierarchy(geom, nx, ny, 1.0, Rd)
    plain = orc.OracleGmg(nx, ny, dx, dy, o.mskp)
    assert g.nlevs == plain.nlevs >= 5
    for lev in range(g.nlevs):
        a_s, a_p = g.A(lev), plain.A(lev)
        # off-centre couplings untouched
        assert (a_s[:4] == a_p[:4]).all(), lev
        on = a_p[4] != 0.0
        assert on.any() and (~on).any() if geom == "closed" else on.any()
        assert (a_s[4][on] == a_p[4][on] - 1.0 / Rd ** 2).all(), lev
        assert (a_s[4][~on] == 0.0).all(), lev


def test_without_the_shift_the_solution_differs():
    """Guard: the periodic comparison above cannot pass without the shift
    (the unshifted solve converges to the Poisson solution instead)."""
    nx = ny = 32
    Rd = 0.1
    o = orc.OracleModel("euler", nx, ny, geometry="perio")
    dx = dy = 1.0 / nx
    g = orc.OracleGmg(nx, ny, dx, dy, o.mskp)
    rng = np.random.default_rng(17)
    qi = rng.standard_normal((ny, nx))
    qi -= qi.mean()
    q = np.zeros(o.mskp.shape)
    q[NH:-NH, NH:-NH] = qi
    orc.fillhalo(q, NH)
    x = np.zeros_like(q)
    g.solve(x, q, maxite=30, tol=1e-12)
    a = 2 * np.pi * np.fft.fftfreq(nx)
    L = lam(a[None, :], a[:, None], dx * dy) - 1.0 / Rd ** 2
    xf = np.real(np.fft.ifft2(np.fft.fft2(qi) / L))
    xi = x[NH:-NH, NH:-NH] - x[NH:-NH, NH:-NH].mean()
    assert np.abs(xi - xf).max() > 1e-3 * np.abs(xf).max()
