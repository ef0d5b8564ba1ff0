"""Pins of the oracle's time schemes (oracle/f2d_oracle.c rk_forward,
core/timescheme.py:152-180) and of where the models take their
single-stage terms (kforcing, dtcoef; core/euler.py:149-169,
core/boussinesq.py:102-114), independently of the oracle's own text:

* bit for bit against the reference's own ``Timescheme`` (imported in the
  build container by tests/golden/make_timescheme_golden.py) on
  dx/dt = lam * x, which fixes NumPy's operation order, the stage times and
  the kforcing / dtcoef values;
* the stability polynomial that the order conditions of a three-stage,
  third-order method fix on a linear problem, R(z) = 1 + z + z^2/2 + z^3/6;
* the temporal order of the full model with every inversion solved to
  convergence: 3 for RK3_SSP (Euler and Boussinesq), 2 for the reference's
  "RK3" on the nonlinear Euler problem (its weights (0, 0, 1) at stage
  times (0, 1/3, 1/2) satisfy sum b c^2 = 1/4, not 1/3);
* diffusion taken at stage kforcing only: with RK3_SSP (dtcoef 1.5 at the
  last stage, weight 4/6) a pure diffusion problem converges at order 1 to
  the exact exponential of the discrete operator; with RK3 (kforcing 0,
  weights (0, 0, 1)) it leaves the field unchanged, bit for bit.

A plausible slip fails one of them: dt/2 for dt/3 in RK3 changes the z^3
coefficient of R; 0.5 for 0.25 in RK3_SSP's second stage changes R and
drops the order; kforcing or dtcoef off by a stage or a factor changes the
diffusion limit.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCHEMES = ("RK3_SSP", "RK3")


@pytest.fixture(scope="module")
def ts():
    return dict(np.load(os.path.join(GOLD, "timescheme_rk3.npz"), allow_pickle=False))


@pytest.mark.parametrize("scheme", SCHEMES)
def test_stage_algebra_bitwise_vs_reference_timescheme(ts, scheme):
    x = ts["x0"].copy()
    for step in range(3):
        orc.rk_linear(scheme, ts["lam"], x, float(ts["dt"]))
        ref = ts["%s_x%d" % (scheme, step + 1)]
        assert np.array_equal(x, ref), (scheme, step, np.abs(x - ref).max())


def test_stage_times_and_forcing_stage_of_reference(ts):
    dt, t0 = float(ts["dt"]), float(ts["t0"])
    calls = ts["RK3_SSP_calls"]
    assert calls[:3, 0].tolist() == [0, 1, 2]
    np.testing.assert_allclose(calls[:3, 1], [t0, t0 + dt, t0 + 0.5 * dt], rtol=0, atol=1e-15)
    calls = ts["RK3_calls"]
    np.testing.assert_allclose(calls[:3, 1], [t0, t0 + dt / 3, t0 + 0.5 * dt], rtol=0, atol=1e-15)
    assert int(ts["RK3_SSP_kforcing"]) == 2 and float(ts["RK3_SSP_dtcoef"]) == 1.5
    assert int(ts["RK3_kforcing"]) == 0 and float(ts["RK3_dtcoef"]) == 1.0


@pytest.mark.parametrize("scheme", SCHEMES)
def test_stability_polynomial(scheme):
    z = np.linspace(-2.5, 1.0, 57)
    lam = z / 0.3
    x = np.ones_like(z)
    orc.rk_linear(scheme, lam, x, 0.3)
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6
    np.testing.assert_allclose(x, R, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_euler_inverts_except_at_kforcing(ts, scheme):
    """Euler's stage tendency of psi is written by the fast inversion at
    every stage but kforcing (core/euler.py:156-169); the tendency buffers
    start at zero, so after a step the skipped stage's psi tendency is
    still zero and the others are not.  Boussinesq inverts at every stage."""
    kf = int(ts["%s_kforcing" % scheme])
    for model in ("euler", "boussinesq"):
        o = orc.OracleModel(model, 16, 16, geometry="perio", timestepping=scheme)
        rng = np.random.default_rng(3)
        w = rng.standard_normal(o.get("vorticity").shape)
        w -= w[3:-3, 3:-3].mean()
        orc.fillhalo(w, 3)
        o.get("vorticity")[...] = w
        o.set_psi_from_vorticity()
        o.diagnostics()
        o.step(1e-3)
        for k in range(3):
            zero = not np.any(o.tendency(k)[1])
            assert zero == (model == "euler" and k == kf), (model, scheme, k)


def _vortex_pair(o, r=0.1):
    xr, yr = o.grid["xr"], o.grid["yr"]
    return (np.exp(-((xr - 0.4) ** 2 + (yr - 0.5) ** 2) / r ** 2)
            - np.exp(-((xr - 0.6) ** 2 + (yr - 0.52) ** 2) / r ** 2))


def _run(model, scheme, geom, dt, nsteps, nx=32):
    """Centred order-4 advection (a smooth right-hand side), fixed umax,
    every inversion converged."""
    o = orc.OracleModel(model, nx, nx, 1.0, 1.0, geometry=geom, order=4, timestepping=scheme)
    o.set_exact_inversions(True)
    w = _vortex_pair(o) * o.msk
    if geom == "perio":
        w = w - np.sum(w[3:-3, 3:-3]) / o.area
    w = w * o.msk
    orc.fillhalo(w, 3)
    o.get("vorticity")[...] = w
    if model == "boussinesq":
        b = 0.3 * np.tanh((o.grid["yr"] - 0.5) / 0.1) * o.msk
        orc.fillhalo(b, 3)
        o.get("buoyancy")[...] = b
    o.set_psi_from_vorticity()
    for _ in range(nsteps):
        o.diags["maxspeed"] = 1.0
        o.step(dt)
    return o.state[0].copy()


@pytest.mark.parametrize("model,scheme,geom,lo,hi", [
    ("euler", "RK3_SSP", "perio", 7.0, 9.0),
    ("euler", "RK3_SSP", "closed", 7.0, 9.0),
    ("boussinesq", "RK3_SSP", "xchannel", 7.0, 9.0),
    ("euler", "RK3", "perio", 3.6, 4.6),
])
def test_temporal_order_with_converged_inversions(model, scheme, geom, lo, hi):
    T = 0.08
    sol = {k: _run(model, scheme, geom, T / k, k) for k in (4, 8, 16, 64)}
    e = [np.linalg.norm(sol[k] - sol[64]) for k in (4, 8, 16)]
    ratios = [e[0] / e[1], e[1] / e[2]]
    assert all(lo <= r <= hi for r in ratios), (model, scheme, geom, e, ratios)


def _diffusion_only(scheme, dt, nsteps, K=2e-3, nx=32, ny=16):
    """Boussinesq channel (periodic x, walls in y) with b = b(y), no flow:
    the torque vanishes, omega stays 0, and b only diffuses.  Centred
    advection, whose fluxes vanish with the velocity (the parabolic upwind
    splitting keeps U = u1/2 at u = 0, a numerical diffusion of its own)."""
    o = orc.OracleModel("boussinesq", nx, ny, 2.0, 1.0, geometry="xchannel", timestepping=scheme,
                        diffusion=True, Kdiff=K, order=4)
    yr = o.grid["yr"]
    b = (np.cos(np.pi * yr) + 0.4 * np.cos(3 * np.pi * yr)) * o.msk
    orc.fillhalo(b, 3)
    o.get("buoyancy")[...] = b
    o.set_psi_from_vorticity()
    b0 = o.get("buoyancy").copy()
    for _ in range(nsteps):
        o.diags["maxspeed"] = 1.0
        o.step(dt)
    return o, b0


def test_diffusion_only_at_kforcing_with_dtcoef():
    K, T, ny = 2e-3, 2.0, 16
    o, b0 = _diffusion_only("RK3_SSP", T / 8, 8, K)
    dx = o.grid["dx"]
    # 1D discrete Neumann Laplacian of add_diffusion (core/fortran_operators.f90:125-156)
    L = np.zeros((ny, ny))
    for j in range(ny):
        for jj in (j - 1, j + 1):
            if 0 <= jj < ny:
                L[j, jj] += 1.0
                L[j, j] -= 1.0
    lam, V = np.linalg.eigh(L * K / dx ** 2)
    col = b0[3:-3, 3 + 5]
    exact = V @ (np.exp(lam * T) * (V.T @ col))
    errs, cols = [], []
    for nsteps in (8, 16, 32):
        o, _ = _diffusion_only("RK3_SSP", T / nsteps, nsteps, K)
        got = o.get("buoyancy")[3:-3, 3:-3]
        assert np.allclose(got, got[:, :1], rtol=0, atol=1e-14)  # still b(y)
        cols.append(got[:, 5])
        errs.append(np.abs(got[:, 5] - exact).max())
    # first order towards exp(T K L) b0: a wrong dtcoef or stage would
    # converge to another limit.  The Richardson extrapolation of the two
    # finest runs cancels the first-order term and lands on the limit.
    assert 1.7 <= errs[0] / errs[1] <= 2.3 and 1.7 <= errs[1] / errs[2] <= 2.3, errs
    extrap = 2 * cols[2] - cols[1]
    assert np.abs(extrap - exact).max() < 0.05 * errs[2], (np.abs(extrap - exact).max(), errs)
    # RK3: diffusion at stage 0 only, and stage 0 has weight 0
    o, b0 = _diffusion_only("RK3", T / 8, 8, K)
    assert np.array_equal(o.get("buoyancy"), b0)
