"""The oracle against what the mathematics fixes, independently of the
reference's own Fortran:

- the finest multigrid operator (core/gmg/level.py:227-335 with
  core/gmg/fortran_multigrid.f90:706-811 for dx == dy) is the 9-point
  Laplacian whose periodic symbol is
  (-3 + cos a + cos b + cos a cos b) / (dx dy): every Fourier mode is an
  eigenvector with that eigenvalue;
- the converged solve (core/gmg/hierarchy.py:153-218, F-cycles to a
  tolerance) equals the FFT solution on a periodic domain (zero-mean rhs,
  solution up to a constant) and the DST-I solution in a closed box (psi = 0
  on the boundary corners);
- the advection tendency (core/fortran_advection.f90:2-284) is a flux
  divergence, so its interior sum vanishes on a periodic domain and in a
  closed box whose wall-normal velocities are zero;
- a constant tracer on a discretely divergence-free flow (u, v from
  core/fortran_operators.f90:2-39) has a zero tendency up to rounding: the
  edge values are c times the sum of the interpolation weights.

A dropped term, a wrong weight, a sign or an index slip in the operator, the
cycle logic, the flux form or the edge interpolation fails one of these.
"""
import numpy as np
import pytest
from scipy.fft import dstn, idstn

from oracle import oracle as orc

NH = 3


def lam(a, b, dxdy):
    return (-3 + np.cos(a) + np.cos(b) + np.cos(a) * np.cos(b)) / dxdy


def gmg_for(geom, nx, ny, Lx):
    o = orc.OracleModel("euler", nx, ny, Lx=Lx, geometry=geom)
    dx, dy = Lx / nx, 1.0 / ny
    return o, orc.OracleGmg(nx, ny, dx, dy, o.mskp), dx, dy


@pytest.mark.parametrize("kx,ky", [(1, 0), (3, 5), (7, 2), (16, 16)])
def test_operator_is_the_9point_symbol(kx, ky):
    nx = ny = 32
    o, g, dx, dy = gmg_for("perio", nx, ny, 1.0)
    m, n = o.mskp.shape
    J, I = np.meshgrid(np.arange(m) - NH, np.arange(n) - NH, indexing="ij")
    x = np.cos(2 * np.pi * (kx * I / nx + ky * J / ny)) + 0.5 * np.sin(2 * np.pi * (kx * I / nx + ky * J / ny))
    b, r = np.zeros_like(x), np.zeros_like(x)
    orc.residual(g.msk(0), g.A(0), x, b, r)
    want = lam(2 * np.pi * kx / nx, 2 * np.pi * ky / ny, dx * dy) * x
    s = np.s_[NH:-NH, NH:-NH]
    assert np.abs(-r[s] - want[s]).max() <= 1e-11 * np.abs(want[s]).max()


@pytest.mark.parametrize("nx,ny,Lx", [(32, 32, 1.0), (64, 32, 2.0)])
def test_periodic_solve_is_the_fft_solution(nx, ny, Lx):
    o, g, dx, dy = gmg_for("perio", nx, ny, Lx)
    rng = np.random.default_rng(7)
    bi = rng.standard_normal((ny, nx))
    bi -= bi.mean()
    b = np.zeros(o.mskp.shape)
    b[NH:-NH, NH:-NH] = bi
    orc.fillhalo(b, NH)
    x = np.zeros_like(b)
    nite, res = g.solve(x, b, maxite=20, tol=1e-12)
    assert res < 1e-12 and nite < 20
    a = 2 * np.pi * np.fft.fftfreq(nx)
    c = 2 * np.pi * np.fft.fftfreq(ny)
    L = lam(a[None, :], c[:, None], dx * dy)
    L[0, 0] = 1.0
    xh = np.fft.fft2(bi) / L
    xh[0, 0] = 0.0
    xf = np.real(np.fft.ifft2(xh))
    xi = x[NH:-NH, NH:-NH] - x[NH:-NH, NH:-NH].mean()
    assert np.abs(xi - xf).max() <= 1e-11 * np.abs(xf).max()


@pytest.mark.parametrize("nx,ny,Lx", [(32, 32, 1.0), (64, 32, 2.0)])
def test_closed_box_solve_is_the_dst_solution(nx, ny, Lx):
    o, g, dx, dy = gmg_for("closed", nx, ny, Lx)
    # the unknowns are the nx-1 by ny-1 inner corners
    s = np.s_[NH:NH + ny - 1, NH:NH + nx - 1]
    assert (o.mskp[s] == 1).all() and o.mskp.sum() == (nx - 1) * (ny - 1)
    rng = np.random.default_rng(8)
    bi = rng.standard_normal((ny - 1, nx - 1))
    b = np.zeros(o.mskp.shape)
    b[s] = bi
    x = np.zeros_like(b)
    nite, res = g.solve(x, b, maxite=20, tol=1e-12)
    assert res < 1e-12 and nite < 20
    a = np.pi * np.arange(1, nx) / nx
    c = np.pi * np.arange(1, ny) / ny
    xf = idstn(dstn(bi, type=1) / lam(a[None, :], c[:, None], dx * dy), type=1)
    assert np.abs(x[s] - xf).max() <= 1e-11 * np.abs(xf).max()


def flow(geom, n=32, seed=3):
    """A cell mask and edge velocities from a random streamfunction, zero on
    the boundary corners of a closed box as the model's psi is (psi *=
    mskp), so that the walls are streamlines and the flow is discretely
    divergence-free in every fluid cell."""
    grid = orc.make_grid(n, n, geometry=geom)
    msk = grid["msk"]
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal(msk.shape)
    if geom == "perio":
        orc.fillhalo(psi, NH)
    else:
        psi *= orc.OracleModel("euler", n, n, geometry=geom).mskp
    u, v = np.zeros_like(psi), np.zeros_like(psi)
    orc.orthogradient(msk, psi, grid["dx"], grid["dy"], u, v)
    if geom == "perio":
        orc.fillhalo(u, NH)
        orc.fillhalo(v, NH)
    cst = np.array([grid["dx"], grid["dy"], 0.05, max(np.abs(u).max(), np.abs(v).max()), 0.05])
    return msk, u, v, cst, grid


SCHEMES = [("up", o, m) for o in (1, 3, 5) for m in (0, 1)] + [("c", o, 0) for o in (2, 4, 6)]


def tendency(kind, order, method, msk, q, u, v, cst):
    y = np.zeros_like(q)
    if kind == "up":
        orc.adv_upwind(msk, q, y, u, v, cst, NH, method, order)
    else:
        orc.adv_centered(msk, q, y, u, v, cst, NH, order)
    return y


@pytest.mark.parametrize("geom", ["perio", "closed"])
@pytest.mark.parametrize("kind,order,method", SCHEMES)
def test_advection_conserves_the_tracer(geom, kind, order, method):
    msk, u, v, cst, grid = flow(geom)
    rng = np.random.default_rng(5)
    q = rng.standard_normal(msk.shape) * msk
    if geom == "perio":
        orc.fillhalo(q, NH)
    y = tendency(kind, order, method, msk, q, u, v, cst)
    s = np.s_[NH:-NH, NH:-NH]
    assert np.abs(y[s]).max() > 0
    assert abs(np.sum(y[s])) <= 1e-12 * np.sum(np.abs(y[s]))
    # no tendency outside the fluid
    assert (y[s][msk[s] == 0] == 0).all()


@pytest.mark.parametrize("geom", ["perio", "closed"])
@pytest.mark.parametrize("kind,order,method", SCHEMES)
def test_constant_tracer_on_divergence_free_flow(geom, kind, order, method):
    msk, u, v, cst, grid = flow(geom)
    c = 2.5
    q = np.full(msk.shape, c)
    y = tendency(kind, order, method, msk, q, u, v, cst)
    s = np.s_[NH:-NH, NH:-NH]
    # the scale of one edge flux difference
    scale = c * max(np.abs(u).max() / grid["dx"], np.abs(v).max() / grid["dy"])
    # Edge value = c * S, S the sum of the edge's float32-rounded weights.
    # On a periodic domain every edge has the same stencil, hence the same S,
    # and the tendency is -c S div(u, v) = 0 up to FP64 rounding.  Near walls
    # the stencil's order drops edge by edge, and S differs between a cell's
    # edges by a few float32 ulps.
    tol = 1e-13 if geom == "perio" else 1e-6
    assert np.abs(y[s]).max() <= tol * scale
