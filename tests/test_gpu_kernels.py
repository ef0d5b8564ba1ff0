"""GPU kernels (libf2dgpu.so through the C ABI) against the reference's own
Fortran outputs (tests/golden/kernels_*.npz) and the oracle.

Bar: bit-identical values for every element-wise/stencil kernel (FP64, no
fused multiply-add, same operation order).  Reductions (norm, kinetic
energy) sum in a different order: relative 1e-13; the maximum speed is
order-independent and must be exact.
"""
import os

import numpy as np
import pytest

from fluid2d_amd import kernels as K
from fluid2d_amd.device import DeviceArray
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NH = 3
MASKS = ["perio", "closed", "islands"]


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def D(a):
    return DeviceArray.from_host(np.ascontiguousarray(a))


def same(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = np.argwhere(a != b)
    assert bad.size == 0, "%s: first mismatch at %s: %r vs %r (%d cells)" % (
        what, tuple(bad[0]), a[tuple(bad[0])], b[tuple(bad[0])], len(bad))


def filled(a):
    a = np.array(a, copy=True)
    orc.fillhalo(a, NH)
    return a


@pytest.fixture(scope="module", params=["kernels_16x16.npz", "kernels_24x16.npz"])
def kf(request):
    return load(request.param)


@pytest.mark.parametrize("mname", MASKS)
def test_advection_vs_fortran(kf, mname):
    p = mname + "_"
    msk, x, u, v, y0, cst = (kf[p + k] for k in ("msk", "x", "u", "v", "y0", "cst"))
    dm, dx_, du, dv = D(msk), D(x), D(u), D(v)
    for order in (1, 3, 5):
        for method in (0, 1):
            y = D(y0)
            K.adv_upwind(dm, dx_, y, du, dv, cst, NH, method, order)
            same(y.get(), kf[p + "up_o%d_m%d" % (order, method)], "upwind o%d m%d" % (order, method))
        y, xf, yf = D(y0), D(np.zeros_like(x)), D(np.zeros_like(x))
        K.adv_upwind(dm, dx_, y, du, dv, cst, NH, 1, order, xf, yf)
        same(y.get(), kf[p + "upflx_o%d" % order])
        same(xf.get(), kf[p + "upflx_o%d_xf" % order])
        same(yf.get(), kf[p + "upflx_o%d_yf" % order])
    for order in (2, 4, 6):
        y = D(y0)
        K.adv_centered(dm, dx_, y, du, dv, cst, NH, 0, order)
        same(y.get(), kf[p + "ce_o%d" % order], "centred o%d" % order)
        y, xf, yf = D(y0), D(np.zeros_like(x)), D(np.zeros_like(x))
        K.adv_centered(dm, dx_, y, du, dv, cst, NH, 0, order, xf, yf)
        same(y.get(), kf[p + "ceflx_o%d" % order])
        same(xf.get(), kf[p + "ceflx_o%d_xf" % order])
        same(yf.get(), kf[p + "ceflx_o%d_yf" % order])


@pytest.mark.parametrize("mname", MASKS)
def test_operators_vs_fortran(kf, mname):
    p = mname + "_"
    msk, x, y0, cst = kf[p + "msk"], kf[p + "x"], kf[p + "y0"], kf[p + "cst"]
    dx, dy = float(cst[0]), float(cst[1])
    dm, dX = D(msk), D(x)
    xp = D(y0)
    K.celltocorner(dX, xp)
    same(xp.get(), kf[p + "corner"], "celltocorner")
    u, v = D(y0), D(y0 * 0.5)
    K.computeorthogradient(dm, dX, dx, dy, NH, u, v)
    same(u.get(), kf[p + "ortho_u"], "u")
    same(v.get(), kf[p + "ortho_v"], "v")
    d = D(y0)
    K.add_diffusion_fill(dm, dX, dx, NH, 1.7e-3, d)
    same(d.get(), filled(kf[p + "diffusion"]), "diffusion")
    # torque as the reference calls it: y *= msk, add_torque, fill
    w = D(y0)
    K.add_torque_fill(dm, dX, dx, NH, 9.81, w)
    ref = y0 * msk
    orc.add_torque(msk, x, dx, NH, 9.81, ref)
    same(w.get(), filled(ref), "torque")
    h = D(x)
    K.fillhalo(h, NH)
    same(h.get(), kf[p + "fillhalo"], "fillhalo")


@pytest.mark.parametrize("mname", MASKS)
def test_multigrid_kernels_vs_fortran(kf, mname):
    p = mname + "_"
    msk, x, y0, A5, b = (kf[p + k] for k in ("msk", "x", "y0", "A5", "b"))
    dm, dA, db = D(msk), D(A5), D(b)
    xs, work = D(x), D(np.zeros_like(x))
    K.smoothtwicewitha_fill(dm, dA, xs, db, 8.0 / 9.0, work)
    same(xs.get(), filled(kf[p + "smooth"]), "smooth")
    r = D(y0)
    K.computeresidualwitha_fill(dm, dA, D(x), db, r)
    same(r.get(), filled(kf[p + "residual"]), "residual")
    n2 = K.computenorm(dm, D(x), NH)
    assert abs(n2 - float(kf[p + "norm"])) <= 1e-13 * abs(float(kf[p + "norm"]))
    msk2, xc = kf[p + "msk2"], kf[p + "xc"]
    x2 = D(np.zeros_like(xc))
    K.restrict_fill(D(msk2), D(x), NH, x2)
    g = kf[p + "restrict"].copy()
    same(x2.get()[NH:-NH, NH:-NH], g[NH:-NH, NH:-NH], "restrict interior")
    same(x2.get(), filled(g), "restrict + fill")
    x1 = D(y0)
    K.interpolate(dm, D(msk2), D(xc), NH, x1)
    same(x1.get(), kf[p + "interpolate"], "interpolate")
    x1 = D(y0)
    K.interpolate(dm, D(msk2), D(xc), NH, x1, accumulate=True)
    same(x1.get(), y0 + kf[p + "interpolate"], "interpolate + x")
    mskc = kf[p + "mskc"]
    Ac = D(np.zeros((9,) + mskc.shape))
    K.coarsenmatrix(D(kf[p + "A9"]), dm, D(mskc), NH, Ac)
    g = kf[p + "coarsen"].copy()
    for k in range(9):
        same(Ac.get()[k], filled(g[k]), "coarsen plane %d" % k)


@pytest.mark.parametrize("mname", MASKS)
def test_diag_vs_fortran(kf, mname):
    p = mname + "_"
    msk, u, v = kf[p + "msk"], kf[p + "u"], kf[p + "v"]
    ke, mx = K.computekemaxu(D(msk), D(u), D(v), NH)
    gke, gmx = kf[p + "diag"][0], kf[p + "diag"][1]
    assert mx == gmx
    assert abs(ke - gke) <= 1e-13 * abs(gke)
    # computesumandnorm / computedotprod / computesum: the reference sums
    # serially, the GPU in a tree; the two differ by rounding only, bounded
    # by 1e-13 of the sum of the terms' magnitudes
    x, y0 = kf[p + "x"], kf[p + "y0"]
    inner = (slice(NH, -NH), slice(NH, -NH))
    fl1 = msk[inner] == 1
    fl = msk[inner] != 0
    xi, yi = x[inner], y0[inner]
    s, s2 = K.computesumandnorm(D(msk), D(x), NH)
    g = kf[p + "diag"]
    assert abs(s - g[2]) <= 1e-13 * np.abs(xi[fl1]).sum()
    assert abs(s2 - g[3]) <= 1e-13 * (xi[fl1] ** 2).sum()
    z = K.computedotprod(D(msk), D(x), D(y0), NH)
    assert abs(z - g[4]) <= 1e-13 * np.abs(xi[fl] * yi[fl]).sum()
    sm = K.computesum(D(msk), D(x), NH)
    assert abs(sm - g[5]) <= 1e-13 * np.abs(xi[fl]).sum()


@pytest.mark.parametrize("shape", [(22, 22), (70, 70), (134, 134), (262, 518), (1030, 1030)])
def test_domain_integration_numpy_order(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape) * np.exp(rng.standard_normal(shape) * 4)
    got = K.domain_integration(D(x))
    assert got == np.sum(x[NH:-NH, NH:-NH]) * 1.0
    assert got == orc.np_sum_interior(np.ascontiguousarray(x), NH)
