"""The C oracle against golden vectors written by the reference's own Fortran
kernels (tests/golden/make_golden.py; oracle/build_ref.sh).

Kernel fixtures: every routine on the path, one call each on random inputs
with periodic, closed-box and irregular masks -- compared bit for bit.

Composed fixtures: the reference's Python driver logic (multigrid set-up and
cycles, inversion, RK3_SSP, Euler/Boussinesq) restated in NumPy around the
reference Fortran kernels, run for three steps -- compared bit for bit with
the oracle's own (independent) restatement in C.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NH = 3


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def same(a, b):
    """Bitwise-equal values (NaN-free); +0/-0 count as equal."""
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    bad = np.argwhere(a != b)
    assert bad.size == 0, "first mismatch at %s: %r vs %r (n=%d)" % (
        tuple(bad[0]), a[tuple(bad[0])], b[tuple(bad[0])], len(bad))


KFILES = ["kernels_16x16.npz", "kernels_24x16.npz"]
MASKS = ["perio", "closed", "islands"]


@pytest.fixture(scope="module", params=KFILES)
def kf(request):
    return load(request.param)


@pytest.mark.parametrize("mname", MASKS)
def test_upwind(kf, mname):
    p = mname + "_"
    msk, x, u, v, y0, cst = (kf[p + k] for k in ("msk", "x", "u", "v", "y0", "cst"))
    for order in (1, 3, 5):
        for method in (0, 1):
            y = y0.copy()
            orc.adv_upwind(msk, x, y, u, v, cst, NH, method, order)
            same(y, kf[p + "up_o%d_m%d" % (order, method)])
        y, xf, yf = y0.copy(), np.zeros_like(x), np.zeros_like(x)
        orc.adv_upwind(msk, x, y, u, v, cst, NH, 1, order, xf, yf)
        same(y, kf[p + "upflx_o%d" % order])
        same(xf, kf[p + "upflx_o%d_xf" % order])
        same(yf, kf[p + "upflx_o%d_yf" % order])


@pytest.mark.parametrize("mname", MASKS)
def test_centered(kf, mname):
    p = mname + "_"
    msk, x, u, v, y0, cst = (kf[p + k] for k in ("msk", "x", "u", "v", "y0", "cst"))
    for order in (2, 4, 6):
        y = y0.copy()
        orc.adv_centered(msk, x, y, u, v, cst, NH, order)
        same(y, kf[p + "ce_o%d" % order])
        y, xf, yf = y0.copy(), np.zeros_like(x), np.zeros_like(x)
        orc.adv_centered(msk, x, y, u, v, cst, NH, order, xf, yf)
        same(y, kf[p + "ceflx_o%d" % order])
        same(xf, kf[p + "ceflx_o%d_xf" % order])
        same(yf, kf[p + "ceflx_o%d_yf" % order])


@pytest.mark.parametrize("mname", MASKS)
def test_operators(kf, mname):
    p = mname + "_"
    msk, x, y0, cst = kf[p + "msk"], kf[p + "x"], kf[p + "y0"], kf[p + "cst"]
    dx, dy = cst[0], cst[1]
    xp = y0.copy()
    orc.celltocorner(x, xp)
    same(xp, kf[p + "corner"])
    u, v = y0.copy(), (y0 * 0.5).copy()
    orc.orthogradient(msk, x, dx, dy, u, v)
    same(u, kf[p + "ortho_u"])
    same(v, kf[p + "ortho_v"])
    d = y0.copy()
    orc.add_diffusion(msk, x, dx, 1.7e-3, d)
    same(d, kf[p + "diffusion"])
    w = y0.copy()
    orc.add_torque(msk, x, dx, NH, 9.81, w)
    same(w, kf[p + "torque"])
    h = x.copy()
    orc.fillhalo(h, NH)
    same(h, kf[p + "fillhalo"])


@pytest.mark.parametrize("mname", MASKS)
def test_diagnostics(kf, mname):
    p = mname + "_"
    msk, u, v, x, y0 = (kf[p + k] for k in ("msk", "u", "v", "x", "y0"))
    ke, mx = orc.kemaxu(msk, u, v, NH)
    s, s2 = orc.sumandnorm(msk, x, NH)
    z = orc.dotprod(msk, x, y0, NH)
    sm = orc.sum_masked(msk, x, NH)
    same(np.array([ke, mx, s, s2, z, sm]), kf[p + "diag"])


@pytest.mark.parametrize("mname", MASKS)
def test_multigrid_kernels(kf, mname):
    p = mname + "_"
    msk, x, y0, A5, b = (kf[p + k] for k in ("msk", "x", "y0", "A5", "b"))
    xs = x.copy()
    orc.smoothtwice(msk, A5, xs, b, 8.0 / 9.0)
    same(xs, kf[p + "smooth"])
    r = y0.copy()
    orc.residual(msk, A5, x, b, r)
    same(r, kf[p + "residual"])
    same(orc.norm2(msk, x, NH), kf[p + "norm"])
    msk2, xc = kf[p + "msk2"], kf[p + "xc"]
    # restrict writes coarse rows/cols [nh-1, m2-nh-1] only; the fixture
    # started from random contents, so compare that window
    x2 = np.zeros_like(xc)
    orc.restrict(msk2, x, NH, x2)
    g = kf[p + "restrict"]
    same(x2[NH - 1:-NH, NH - 1:-NH], g[NH - 1:-NH, NH - 1:-NH])
    x1 = y0.copy()
    orc.interpolate(msk, msk2, xc, NH, x1)
    same(x1, kf[p + "interpolate"])
    # coarsenmatrix likewise writes only [nh-1, m2-nh-1]
    Ac = orc.coarsenmatrix(kf[p + "A9"], msk, kf[p + "mskc"], NH)
    g = kf[p + "coarsen"]
    same(Ac[:, NH - 1:-NH, NH - 1:-NH], g[:, NH - 1:-NH, NH - 1:-NH])


CFILES = {
    "composed_euler_perio_16.npz": ("euler", 16, 16, 1.0, 1.0, "perio", "RK3_SSP"),
    "composed_euler_closed_32x16.npz": ("euler", 32, 16, 2.0, 1.0, "closed", "RK3_SSP"),
    "composed_bous_xchannel_32x16.npz": ("boussinesq", 32, 16, 2.0, 1.0, "xchannel", "RK3_SSP"),
    # core/timescheme.py:152-164 (dt/3, dt/2, dt; kforcing 0, dtcoef 1)
    "composed_euler_closed_rk3_32x16.npz": ("euler", 32, 16, 2.0, 1.0, "closed", "RK3"),
    "composed_bous_xchannel_rk3_32x16.npz": ("boussinesq", 32, 16, 2.0, 1.0, "xchannel", "RK3"),
}


@pytest.mark.parametrize("fname", sorted(CFILES))
def test_composed(fname):
    g = load(fname)
    kind, nx, ny, Lx, Ly, geom, scheme = CFILES[fname]
    M = orc.OracleModel(kind, nx, ny, Lx, Ly, geometry=geom, diffusion=(kind == "boussinesq"),
                        Kdiff=1e-3, msk=g["msk"], timestepping=scheme)
    gmg = orc.OracleGmg(nx, ny, M.grid["dx"], M.grid["dy"], M.mskp)
    assert gmg.nlevs == len(g["levels"])
    for lev in range(gmg.nlevs):
        same(gmg.msk(lev), g["gmg_msk_%d" % lev])
        same(gmg.A(lev), g["gmg_A_%d" % lev])
    M.state[...] = g["state0"]
    M.set_psi_from_vorticity()
    same(M.get("psi"), g["psi_init"])
    same(M.get("u"), g["u_init"])
    same(M.get("v"), g["v_init"])
    # fast inversion of a tendency through the model's multigrid
    t = g["tend_in"].copy()
    psi = t[1].copy()
    work = np.zeros_like(psi)
    orc.celltocorner(t[0], work)
    gmg.twoVcycle(psi, work)
    psi = psi * M.mskp
    same(psi, g["tend_out"][1])
    d = M.diagnostics()
    dt = float(g["dt"])
    assert dt == 0.8 * min(M.grid["dx"], M.grid["dy"]) / d["maxspeed"]
    nites = []
    for k in range(3):
        M.step(dt)
        if kind == "euler":
            nites.append(M.last_solve()[0])
        same(M.state, g["state_%d" % (k + 1)])
    if kind == "euler":
        assert nites == list(g["nite"][1:])
