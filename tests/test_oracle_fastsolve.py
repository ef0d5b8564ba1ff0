"""The oracle's fast inversions by ``gmg.solve`` (the alternative that
core/operators.py:452-455 keeps commented out, the basis of the refined
per-stage inversions of the mixed-precision path).

Pinned by what the mathematics fixes: with a tight tolerance every stage's
psi tendency is the converged solution of the 9-point Laplacian with the
stage's corner vorticity tendency as right-hand side -- in a closed box the
DST-I solution, masked by the corner mask -- and (u, v) tendencies are its
orthogonal gradient.  With maxite <= 0 the mode is off and a step equals the
reference's twoVcycle step bit for bit.
"""
import numpy as np
from scipy.fft import dstn, idstn

from oracle import oracle as orc

NH = 3


def lam(a, b, dxdy):
    return (-3 + np.cos(a) + np.cos(b) + np.cos(a) * np.cos(b)) / dxdy


def model(seed=5, n=32):
    o = orc.OracleModel("boussinesq", n, n, geometry="closed", diffusion=True, Kdiff=1e-4)
    rng = np.random.default_rng(seed)
    msk = o.msk
    w = np.zeros(o.state.shape[1:])
    w[NH:-NH, NH:-NH] = rng.standard_normal((n, n))
    b = np.zeros_like(w)
    b[NH:-NH, NH:-NH] = 0.3 * rng.standard_normal((n, n))
    o.get("vorticity")[...] = w * msk
    o.get("buoyancy")[...] = b * msk
    orc.fillhalo(o.get("vorticity"), NH)
    orc.fillhalo(o.get("buoyancy"), NH)
    o.set_psi_from_vorticity()
    o.diagnostics()
    return o


def test_stage_psi_tendencies_are_converged_solutions():
    n = 32
    o = model(n=n)
    o.set_fast_solve(40, 1e-13)
    dx = 1.0 / n
    o.step(0.2 * dx / max(o.diags["maxspeed"], 1e-3))
    s = np.s_[NH:NH + n - 1, NH:NH + n - 1]
    a = np.pi * np.arange(1, n) / n
    L = lam(a[None, :], a[:, None], dx * dx)
    iw, ip, iu, iv = (o.varname_list.index(k) for k in ("vorticity", "psi", "u", "v"))
    for k in range(3):
        t = o.tendency(k)
        corner = np.zeros_like(t[iw])
        orc.celltocorner(t[iw], corner)
        want = idstn(dstn(corner[s], type=1) / L, type=1)
        got = t[ip][s]
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max(), k
        assert (t[ip] * (1 - o.mskp) == 0).all(), k
        u = np.zeros_like(t[iu])
        v = np.zeros_like(t[iv])
        orc.orthogradient(o.msk, t[ip], dx, dx, u, v)
        assert (u == t[iu]).all() and (v == t[iv]).all(), k


def test_mode_off_is_the_twovcycle_step():
    n = 32
    a, b = model(n=n), model(n=n)
    b.set_fast_solve(6, 1e-10)
    b.set_fast_solve(0, 0.0)
    dt = 0.2 / n
    a.run(3, dt)
    b.run(3, dt)
    assert (a.state == b.state).all()
    c = model(n=n)
    c.set_fast_solve(6, 1e-10)
    c.run(3, dt)
    assert not (a.state == c.state).all()
