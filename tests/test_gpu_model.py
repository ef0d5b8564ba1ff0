"""Multigrid and full-model parity of the GPU path against the oracle.

The plain FP64 path must match the reference algorithm within rel-L2 1e-12
per field after N steps; with identical operation order it is expected to be
bit-identical except where the solve's stopping test (a norm summed in a
different order) flips, which the F-cycle counts would show.
"""
import numpy as np
import pytest

import fluid2d_amd as F
from fluid2d_amd import workloads
from fluid2d_amd.device import DeviceArray
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
NH = 3
TOL = 1e-12


def rel_l2(a, b, msk=None):
    a, b = np.asarray(a), np.asarray(b)
    if msk is not None:
        a, b = a * msk, b * msk
    d = np.linalg.norm((a - b).ravel())
    nb = np.linalg.norm(b.ravel())
    return d / nb if nb > 0 else d


def setup_pair(kind, nx, ny, Lx, Ly, geom, seed, **kw):
    p = F.Param(modelname=kind, nx=nx, ny=ny, Lx=Lx, Ly=Ly, geometry=geom, **kw)
    g = F.Grid(p)
    f = F.Fluid2d(p, g)
    o = orc.OracleModel(kind, nx, ny, Lx, Ly, geometry=geom,
                        diffusion=kw.get("diffusion", False), Kdiff=kw.get("Kdiff", 0.0),
                        gravity=kw.get("gravity", 1.0), order=kw.get("order", 5))
    msk = g.msk
    w = workloads.turbulence(g, seed)
    f.model.var.get("vorticity")[...] = w
    o.get("vorticity")[...] = w
    if kind == "boussinesq":
        rng = np.random.default_rng(seed + 1)
        b = np.where(g.yr < 0.5 * Ly + 1e-3 * np.sin(2 * np.pi * g.xr / Lx), 0.5, -0.5)
        b = (b + 1e-3 * rng.uniform(-1, 1, b.shape)) * msk
        g.fill_halo(b)
        f.model.var.get("buoyancy")[...] = b
        o.get("buoyancy")[...] = b
    f.model.set_psi_from_vorticity()
    o.set_psi_from_vorticity()
    return f, o


def compare_states(f, o, what):
    s = f.model.var.state
    msk = f.model.msk
    for k, name in enumerate(f.model.var.varname_list):
        r = rel_l2(s[k], o.state[k])
        assert r <= TOL, "%s: %s rel-L2 %.3e" % (what, name, r)


def test_gmg_setup_matches_oracle():
    nx = ny = 64
    o = orc.OracleModel("euler", nx, ny, geometry="closed")
    p = F.Param(modelname="euler", nx=nx, ny=ny, geometry="closed", engine=False)
    f = F.Fluid2d(p, F.Grid(p))
    og = orc.OracleGmg(nx, ny, 1.0 / nx, 1.0 / ny, o.mskp)
    assert (f.model.mskp == o.mskp).all()
    assert f.model.gmg.nlevs == og.nlevs
    for lev in range(og.nlevs):
        assert (f.model.gmg.level(lev, "msk").get() == og.msk(lev)).all()
        assert (f.model.gmg.level(lev, "A").get() == og.A(lev)).all(), lev


@pytest.mark.parametrize("geom", ["perio", "closed"])
def test_twovcycle_and_solve_bitwise(geom):
    nx, ny = 64, 32
    o = orc.OracleModel("euler", nx, ny, 2.0, 1.0, geometry=geom)
    og = orc.OracleGmg(nx, ny, 2.0 / nx, 1.0 / ny, o.mskp)
    g = F.Gmg({"n": nx, "m": ny, "nh": 3, "dx": 2.0 / nx, "dy": 1.0 / ny}, o.mskp)
    rng = np.random.default_rng(7)
    shape = (ny + 6, nx + 6)
    b = rng.standard_normal(shape)
    if geom == "perio":
        b -= b[NH:-NH, NH:-NH].mean()
    x = rng.standard_normal(shape) * 1e-2
    xo, xd = x.copy(), DeviceArray.from_host(x)
    og.twoVcycle(xo, b)
    g.twoVcycle(xd, DeviceArray.from_host(b))
    assert (xd.get() == xo).all()
    x0 = np.zeros(shape)
    xo, xd = x0.copy(), DeviceArray.from_host(x0)
    no, ro = og.solve(xo, b)
    nd, rd = g.solve(xd, DeviceArray.from_host(b))
    assert no == nd
    assert abs(rd - ro) <= 1e-10 * ro
    assert (xd.get() == xo).all()


@pytest.mark.parametrize("engine", [True, False])
def test_euler_perio_128_200_steps(engine):
    """Config: Euler 128x128 doubly periodic, RK3_SSP, UP5 parabolic, 200
    steps at CFL 0.8 of the initial maxspeed (k^-3 initial field), on the
    engine-backed model and on the operator-level one."""
    f, o = setup_pair("euler", 128, 128, 1.0, 1.0, "perio", 1, engine=engine)
    assert isinstance(f.model, F.EngineModel) == engine
    compare_states(f, o, "init")
    d0 = f.model.diagnostics(f.model.var, 0)
    do = o.diagnostics()
    assert d0["maxspeed"] == do["maxspeed"]
    dt = 0.8 * f.model.grid.dx / d0["maxspeed"]
    nd = f.model.run(200, dt)
    no = o.run(200, dt)
    assert list(nd) == list(no)
    compare_states(f, o, "step 200")


def test_euler_closed_256_50_steps():
    f, o = setup_pair("euler", 256, 256, 1.0, 1.0, "closed", 2)
    compare_states(f, o, "init")
    d0 = f.model.diagnostics(f.model.var, 0)
    dt = 0.8 * f.model.grid.dx / d0["maxspeed"]
    o.diagnostics()
    nd = f.model.run(50, dt)
    no = o.run(50, dt)
    assert list(nd) == list(no)
    compare_states(f, o, "step 50")


def test_boussinesq_xchannel_256x128_30_steps():
    f, o = setup_pair("boussinesq", 256, 128, 2.0, 1.0, "xchannel", 3, diffusion=True,
                      Kdiff=1e-5, gravity=1.0)
    compare_states(f, o, "init")
    d0 = f.model.diagnostics(f.model.var, 0)
    o.diagnostics()
    # zero initial velocity: dt from the buoyancy velocity scale sqrt(g db Ly) = 1
    dt = 0.8 * f.model.grid.dx / 1.0
    f.model.run(30, dt)
    o.run(30, dt)
    compare_states(f, o, "step 30")


def test_fluid2d_loop_drop_in():
    """The reference's driver flow: Param -> Grid -> Fluid2d -> set fields ->
    set_psi_from_vorticity -> loop (adaptive dt from maxspeed)."""
    p = F.Param(modelname="euler", nx=64, ny=64, geometry="closed", adaptable_dt=True, cfl=0.8,
                tend=0.05)
    g = F.Grid(p)
    f = F.Fluid2d(p, g)
    f.model.var.get("vorticity")[...] = workloads.turbulence(g, 5)
    f.model.set_psi_from_vorticity()
    n = f.loop()
    assert n > 0 and abs(f.t - 0.05) < f.dt + 1e-12
    ke = f.model.diags["ke"]
    assert np.isfinite(ke) and ke > 0


def test_fluid2d_loop_exacthistime_matches_oracle():
    """Fluid2d.loop with adaptive dt shortened onto the history times
    (core/fluid2d.py:239-255): the history records land on multiples of
    freq_his, and the state equals the oracle stepped with the same dt
    sequence (each step's umax is the oracle's own maxspeed)."""
    p = F.Param(modelname="euler", nx=64, ny=64, geometry="closed", adaptable_dt=True, cfl=0.8,
                tend=0.03, freq_his=0.01, exacthistime=True)
    g = F.Grid(p)
    f = F.Fluid2d(p, g)
    o = orc.OracleModel("euler", 64, 64, geometry="closed")
    w = workloads.turbulence(g, 5)
    f.model.var.get("vorticity")[...] = w
    o.get("vorticity")[...] = w
    f.model.set_psi_from_vorticity()
    o.set_psi_from_vorticity()
    dts = []
    step = f.model.step
    f.model.step = lambda t, dt: (dts.append(dt), step(t, dt))
    f.loop()
    his_t = [t for t, _, _ in f.output.his]
    assert np.allclose(his_t, [0.0, 0.01, 0.02, 0.03], rtol=0, atol=1e-14)
    o.diagnostics()
    states = {}
    for k, dt in enumerate(dts, 1):
        o.step(dt)
        states[k] = o.state.copy()
    compare_states(f, o, "loop end")
    for t, kt, s in f.output.his[1:]:
        for v, name in enumerate(f.model.var.varname_list):
            r = rel_l2(s[v], states[kt][v])
            assert r <= TOL, "history t=%g %s rel-L2 %.3e" % (t, name, r)


def test_fluid2d_loop_enforce_momentum():
    """enforce_momentum (core/fluid2d.py:270-272, 351-358): after each step
    the vorticity's linear profile is removed and psi, u, v re-inverted, so
    the momentum diagnostics px, py stay at rounding level."""
    p = F.Param(modelname="euler", nx=64, ny=64, geometry="closed", adaptable_dt=True, cfl=0.8,
                tend=1e9, enforce_momentum=True)
    g = F.Grid(p)
    f = F.Fluid2d(p, g)
    w = workloads.turbulence(g, 6)
    f.model.var.get("vorticity")[...] = w
    f.model.set_psi_from_vorticity()
    d0 = dict(f.model.diagnostics(f.model.var, 0))
    p0 = max(abs(d0["px"]), abs(d0["py"]))
    assert p0 > 1e-6
    f.loop(nsteps=4)
    d = f.model.diags
    assert abs(d["px"]) < 1e-10 * p0 and abs(d["py"]) < 1e-10 * p0
    s = f.model.var.state
    assert np.isfinite(s).all()
    # (u, v) is the orthogonal gradient of the stored psi (operators.py:487-497)
    u, v = np.zeros_like(s[2]), np.zeros_like(s[3])
    orc.orthogradient(f.model.msk, np.ascontiguousarray(s[1]), g.dx, g.dy, u, v)
    assert (u == s[2]).all() and (v == s[3]).all()
    # and psi was re-inverted from the corrected vorticity: a fresh full
    # solve moves it by no more than two V-cycles' error
    psi = s[1].copy()
    f.model.set_psi_from_vorticity()
    assert rel_l2(f.model.var.get("psi"), psi) < 1e-3


@pytest.mark.parametrize("kind,geom,order", [("euler", "closed", 5), ("euler", "perio", 4),
                                             ("boussinesq", "xchannel", 5)])
def test_diag_fluxes_matches_oracle(kind, geom, order):
    """Fluxes.diag_fluxes (core/fluxes.py:100-170) on the GPU against the
    oracle, plane by plane, on two calls in a row (the second starts its
    fast inversions from the tendencies the first left)."""
    from fluid2d_amd.fluxes import Fluxes

    nx, ny, Lx = (128, 64, 2.0) if kind == "boussinesq" else (64, 64, 1.0)
    f, o = setup_pair(kind, nx, ny, Lx, 1.0, geom, 8, order=order, engine=False)
    d0 = f.model.diagnostics(f.model.var, 0)
    o.diagnostics()
    dt = 0.5 * f.model.grid.dx / max(d0["maxspeed"], 1.0)
    f.model.run(2, dt)
    o.run(2, dt)
    compare_states(f, o, "before the flux diagnostics")
    umax = f.model.diags["maxspeed"]
    assert umax == o.diags["maxspeed"]
    f.model.ope.cst[3] = umax
    orc.lib().orc_model_set_umax(o._h, umax)
    fl = Fluxes(f.param, f.model.grid, f.model.ope)
    assert fl.fullflx_list == o.flux_names()
    for call in range(2):
        fl.diag_fluxes(f.model.var.dev, 0.0, dt)
        got = fl.flx.get()
        ref = o.diag_fluxes(dt)
        for k, name in enumerate(fl.fullflx_list):
            r = rel_l2(got[k], ref[k])
            assert r <= TOL, "call %d %s rel-L2 %.3e" % (call, name, r)
    compare_states(f, o, "after the flux diagnostics")


def test_fluid2d_loop_diag_fluxes():
    """diag_fluxes = True: the loop records the fluxes with each history
    record (core/fluid2d.py:203-210, 293-295)."""
    p = F.Param(modelname="euler", nx=64, ny=64, geometry="closed", adaptable_dt=True, cfl=0.8,
                tend=0.02, freq_his=0.01, diag_fluxes=True)
    g = F.Grid(p)
    f = F.Fluid2d(p, g)
    f.model.var.get("vorticity")[...] = workloads.turbulence(g, 9)
    f.model.set_psi_from_vorticity()
    f.loop()
    assert [t for t, _, _ in f.output.flx] == [t for t, _, _ in f.output.his]
    assert len(f.output.flx) == 3
    for t, kt, flx in f.output.flx:
        assert flx.shape == (12, 70, 70) and np.isfinite(flx).all()
        assert np.abs(flx[0]).max() > 0


def noslip_pair(kind, nx, ny, Lx, geom, seed, enforce=False, islands=None, **kw):
    """GPU Fluid2d and oracle with no-slip walls (and islands), from the same
    host set-up (grid masks, boundary mask, island arrays)."""
    p = F.Param(modelname=kind, nx=nx, ny=ny, Lx=Lx, Ly=1.0, geometry=geom, noslip=True,
                enforce_momentum=enforce, isisland=islands is not None, **kw)
    g = F.Grid(p)
    if islands is not None:
        islands(p, g)
    f = F.Fluid2d(p, g)
    o = orc.OracleModel(kind, nx, ny, Lx, 1.0, geometry=geom, msk=g.msk,
                        diffusion=kw.get("diffusion", False), Kdiff=kw.get("Kdiff", 0.0))
    mskbc, bcarea, x2bc, y2bc = g.boundary_mask()
    o.set_noslip(g.msknoslip, mskbc, g.xr0, g.yr0, bcarea, x2bc, y2bc, enforce)
    if islands is not None:
        o.set_islands(g.island.rhsp, g.island.psi)
    assert (f.model.mskp == o.mskp).all()
    return f, o, g


@pytest.mark.parametrize("kind,geom,enforce", [("euler", "closed", False),
                                               ("euler", "closed", True),
                                               ("boussinesq", "xchannel", False)])
def test_noslip_matches_oracle(kind, geom, enforce):
    """No-slip walls (core/operators.py:245-290 after each step): 20 steps
    against the oracle.  Bitwise expected except with enforce_momentum,
    whose two dot products are summed in a different order."""
    nx, Lx = (128, 2.0) if kind == "boussinesq" else (64, 1.0)
    f, o, g = noslip_pair(kind, nx, 64, Lx, geom, 11, enforce=enforce)
    w = workloads.turbulence(g, 11) * g.msk
    f.model.var.get("vorticity")[...] = w
    o.get("vorticity")[...] = w
    if kind == "boussinesq":
        rng = np.random.default_rng(12)
        b = (rng.standard_normal(g.msk.shape) * 0.1) * g.msk
        g.fill_halo(b)
        f.model.var.get("buoyancy")[...] = b
        o.get("buoyancy")[...] = b
    f.model.set_psi_from_vorticity()
    o.set_psi_from_vorticity()
    d0 = f.model.diagnostics(f.model.var, 0)
    o.diagnostics()
    dt = 0.5 * g.dx / d0["maxspeed"]
    f.model.run(20, dt)
    o.run(20, dt)
    compare_states(f, o, "noslip 20 steps")
    s = f.model.var.state
    if kind == "euler":
        src = s[f.model.var.varname_list.index("source")]
        assert np.abs(src[3:-3, 3:-3]).max() > 0
        if not enforce:
            assert (s == o.state).all()


def karman(p, g):
    """An island disc in a channel whose walls carry psi = +-psi0/2 (the
    pattern of experiments/VonKarman/karman_street.py), no-slip off along
    the outer walls."""
    nh = 3
    r = np.sqrt((g.xr - 0.5) ** 2 + (g.yr - 0.5) ** 2)
    idx = np.where(r <= 0.12)
    g.msk[idx] = 0
    g.msknoslip[idx] = 0
    g.island.add(idx, 0.0)
    psi0 = 0.2
    low = np.zeros_like(g.msk)
    low[:nh, :] = 1
    g.island.add(np.where(low == 1), psi0 * 0.5)
    high = np.zeros_like(g.msk)
    high[-nh:-1, :] = 1
    g.island.add(np.where(high == 1), -psi0 * 0.5)
    g.msknoslip[:nh, :] = 1
    g.msknoslip[-nh:, :] = 1


def test_islands_and_noslip_match_oracle():
    """Islands (core/island.py, core/operators.py:421-498 island branch) with
    no-slip walls: the flow past a disc between two walls held at +-psi0/2,
    from rest, 30 steps against the oracle, bitwise."""
    f, o, g = noslip_pair("euler", 128, 64, 2.0, "xchannel", 13, islands=karman)
    f.model.set_psi_from_vorticity()
    o.set_psi_from_vorticity()
    compare_states(f, o, "island psi")
    psi = f.model.var.get("psi")
    assert abs(psi[3:-3, 3:-3]).max() > 0.05  # the imposed through-flow
    d0 = f.model.diagnostics(f.model.var, 0)
    o.diagnostics()
    dt = 0.5 * g.dx / d0["maxspeed"]
    f.model.run(30, dt)
    o.run(30, dt)
    compare_states(f, o, "islands 30 steps")
    assert (f.model.var.state == o.state).all()


@pytest.mark.parametrize("kind,geom", [("euler", "disc"), ("euler", "ychannel"),
                                       ("euler", "xperio"), ("boussinesq", "closed")])
def test_engine_model_equals_operator_model(kind, geom):
    """Fluid2d's engine-backed model and the operator-level model take the
    same 10 steps to the same bits, for the geometries not covered above."""
    states = []
    for engine in (True, False):
        p = F.Param(modelname=kind, nx=64, ny=64, geometry=geom, engine=engine)
        g = F.Grid(p)
        f = F.Fluid2d(p, g)
        assert isinstance(f.model, F.EngineModel) == engine
        f.model.var.get("vorticity")[...] = workloads.turbulence(g, 21) * g.msk
        f.model.set_psi_from_vorticity()
        d = f.model.diagnostics(f.model.var, 0)
        log = f.model.run(10, 0.5 * g.dx / d["maxspeed"])
        states.append((f.model.var.state.copy(), list(log)))
    (a, la), (b, lb) = states
    assert la == lb
    assert (a == b).all()
