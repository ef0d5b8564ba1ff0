"""Host logic of Fluid2d.loop (core/fluid2d.py:188-358, core/output.py:62-96)
on a stand-in model: the time-step schedule, the output clocks and the
momentum removal.  No GPU: the model's step and diagnostics are replaced by
a stub whose maxspeed varies from step to step."""
import types

import numpy as np
import pytest

import fluid2d_amd as F
from fluid2d_amd.fluid2d import Fluid2d, Output


class StubModel:
    def __init__(self, speeds, ens=1.0, ke=None):
        self.speeds = speeds
        self.k = 0
        self.diags = {"maxspeed": speeds(0), "ke": 1.0, "enstrophy": ens}
        self.ke = ke
        self.ope = types.SimpleNamespace(cst=np.zeros(5))
        st = np.zeros((1, 2, 2))
        self.var = types.SimpleNamespace(state=st, sync_to_device=lambda: None,
                                         dev=types.SimpleNamespace(get=st.copy))
        self.steps = []

    def step(self, t, dt):
        self.steps.append((t, dt, self.ope.cst[3]))
        self.k += 1

    def diagnostics(self, var, t):
        self.diags["maxspeed"] = self.speeds(self.k)
        if self.ke is not None:
            self.diags["ke"] = self.ke(self.k)
        return self.diags


def make(model, **kw):
    p = F.Param(**kw)
    f = Fluid2d.__new__(Fluid2d)
    for k in ("modelname", "tend", "adaptable_dt", "dt", "cfl", "dtmax", "geometry", "nx",
              "ny", "npx", "npy", "decay", "nprint", "exacthistime", "rescaledtime",
              "enforce_momentum", "diag_fluxes"):
        setattr(f, k, getattr(p, k))
    f.dt0 = f.dt
    f.dx, f.dy = p.Lx / p.nx, p.Ly / p.ny
    f.model = model
    f.enstrophyname = "enstrophy"
    f.output = Output(p)
    f.t, f.kt, f.stop, f.walltime, f.ke_increases = 0.0, 0, False, 0.0, 0
    return f


def speeds(k):
    return 1.0 + 0.4 * np.sin(0.7 * k)


def test_exacthistime_lands_on_every_history_time():
    m = StubModel(speeds)
    f = make(m, nx=64, ny=64, cfl=0.5, tend=0.5, freq_his=0.1, freq_diag=0.05,
             adaptable_dt=True, exacthistime=True)
    f.loop()
    his_t = [t for t, _, _ in f.output.his]
    assert len(his_t) == 6
    assert np.allclose(his_t, np.arange(6) * 0.1, rtol=0, atol=1e-13)
    # the steps stay positive; the step that lands on a history time ends
    # exactly there (dt = tnexthis - t)
    assert all(dt > 0 for _, dt, _ in m.steps)
    ends = [t + dt for t, dt, _ in m.steps]
    for th in his_t[1:]:
        assert th in ends
    # umax handed to the advection is the maxspeed that set the step
    assert all(ms == speeds(k) for k, (_, _, ms) in enumerate(m.steps))
    diag_t = [t for t, _, _ in f.output.diag]
    assert diag_t[0] == 0.0 and all(b > a for a, b in zip(diag_t, diag_t[1:]))


def test_without_exacthistime_history_follows_the_step():
    m = StubModel(speeds)
    f = make(m, nx=64, ny=64, cfl=0.5, tend=0.3, freq_his=0.1, adaptable_dt=True,
             exacthistime=False)
    f.loop()
    dx = 1.0 / 64
    for k, (t, dt, ms) in enumerate(m.steps):
        assert dt == 0.5 * dx / speeds(k)  # plain CFL step, core/fluid2d.py:378-390
    his_t = [t for t, _, _ in f.output.his]
    # a record at the first t >= each multiple of freq_his
    for j, t in enumerate(his_t[1:], 1):
        assert t >= 0.1 * j - 1e-12
    assert not np.allclose(his_t[1:], 0.1 * np.arange(1, len(his_t)), atol=1e-9)


def test_dtmax_and_fixed_dt():
    m = StubModel(lambda k: 1e-6)
    f = make(m, nx=64, ny=64, cfl=0.5, dtmax=0.01, tend=0.05, adaptable_dt=True,
             exacthistime=False)
    f.loop()
    assert all(dt == 0.01 for _, dt, _ in m.steps)
    m = StubModel(speeds)
    f = make(m, nx=64, ny=64, dt=0.02, tend=0.1, adaptable_dt=False)
    f.loop()
    assert all(dt == 0.02 for _, dt, _ in m.steps) and len(m.steps) == 5


def test_rescaled_time_and_energy_rates():
    m = StubModel(lambda k: 1.0, ens=4.0, ke=lambda k: 1.0 + 0.1 * k)
    f = make(m, nx=64, ny=64, dt=0.01, tend=0.1, adaptable_dt=False, rescaledtime="enstrophy")
    n = f.loop()
    assert n == 5  # t advances by dt * sqrt(enstrophy) = 0.02 per step
    assert f.model.diags["dkedt"] == pytest.approx(0.1 / 0.01, rel=1e-12)
    assert f.model.diags["dvdt"] == 0.0
    assert f.ke_increases == n  # ke rises at every step of a decaying run


def test_enforce_zero_momentum_removes_px_py():
    """omega -= xr0 px/x2 + yr0 py/y2 zeroes the momentum px = <omega x>,
    py = <omega y> of a symmetric closed domain (core/fluid2d.py:351-358)."""
    p = F.Param(nx=48, ny=48, geometry="closed")
    g = F.Grid(p)
    rng = np.random.default_rng(3)
    w = rng.standard_normal(g.msk.shape) * g.msk

    def mom(v):
        return (g.domain_integration(v * g.xr) / g.area, g.domain_integration(v * g.yr) / g.area)

    state = w[None].copy()
    calls = []
    model = types.SimpleNamespace(
        diags=dict(zip(("px", "py"), mom(w))),
        var=types.SimpleNamespace(get=lambda name: state[0], dev="device-state"),
        ope=types.SimpleNamespace(invert_vorticity=lambda x, flag: calls.append((x, flag))))
    f = Fluid2d.__new__(Fluid2d)
    f.enforce_momentum, f.grid, f.model = True, g, model
    f.enforce_zero_momentum()
    px, py = mom(state[0])
    p0 = max(abs(v) for v in mom(w))
    assert abs(px) < 1e-12 * p0 and abs(py) < 1e-12 * p0
    assert calls == [("device-state", "fast")]
    # only the fluid cells change
    assert (state[0][g.msk == 0] == 0).all()


def test_boundary_mask_closed_box():
    """Operators.set_boundary_msk (core/operators.py:157-185): in a closed
    n x n box the no-slip boundary cells are the ring of fluid cells along
    the walls, 4n - 4 of them, with second moments sum(x0^2) over the ring."""
    n = 16
    p = F.Param(nx=n, ny=n, geometry="closed")
    g = F.Grid(p)
    mskbc, bcarea, x2bc, y2bc = g.boundary_mask()
    assert bcarea == 4 * n - 4
    ring = np.zeros((n, n), bool)
    ring[0, :] = ring[-1, :] = ring[:, 0] = ring[:, -1] = True
    assert (mskbc[3:-3, 3:-3] == ring).all()
    xc = (np.arange(n) + 0.5) / n - 0.5
    X = np.broadcast_to(xc, (n, n))
    assert x2bc == pytest.approx(np.sum(X[ring] ** 2), rel=1e-13)
    assert y2bc == pytest.approx(np.sum(X.T[ring] ** 2), rel=1e-13)


def test_island_set_up_by_brute_force():
    """Island.finalize (core/island.py:21-43): psi0 on every cell corner that
    touches an island cell (and on the last row and column); rhsp = (number of such corners among the four
    neighbour corners) * psi0 / (dx dy) wherever that number is positive."""
    p = F.Param(nx=24, ny=16, Lx=1.5, geometry="closed", isisland=True)
    g = F.Grid(p)
    blk = np.zeros_like(g.msk, bool)
    blk[8:11, 9:12] = True
    g.msk[blk] = 0
    g.island.add(np.where(blk), 0.3)
    g.island.finalize()
    m, n = g.msk.shape
    # corner (j, i): cells (j..j+1, i..i+1); celltocorner leaves the last
    # row and column at 0, which the set-up reads as island corners too
    corner = np.ones((m, n), bool)
    for j in range(m - 1):
        for i in range(n - 1):
            corner[j, i] = blk[j:j + 2, i:i + 2].any()
    assert (g.island.psi == np.where(corner, 0.3, 0.0)).all()
    cnt = np.zeros((m, n), int)
    for j in range(m):
        for i in range(n):
            cnt[j, i] = sum(corner[(j + a) % m, (i + b) % n] for a, b in ((1, 0), (-1, 0), (0, 1), (0, -1)))
    want = np.where(cnt > 0, cnt * 0.3 / (g.dx * g.dy), 0.0)
    assert np.allclose(g.island.rhsp, want, rtol=1e-15, atol=0)

