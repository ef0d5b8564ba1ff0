"""Euler and Boussinesq models on the device, with the reference's interface.

``Euler(param, grid)`` / ``Boussinesq(param, grid)`` mirror core/euler.py and
core/boussinesq.py: ``model.var`` (packed state, ``get(name)``),
``model.ope`` (Operators: ``rhs_adv``, ``invert_vorticity``, ``fill_halo``),
``model.step(t, dt)``, ``model.diagnostics(var, t)``,
``model.set_psi_from_vorticity()``, ``model.diags``.  The state, the RK
tendencies (core/timescheme.py:35-41) and the multigrid live on the GPU; the
whole step is one call into libf2dgpu.so.

Host access: ``var.state`` / ``var.get(name)`` return a host mirror,
downloaded on first access after the device changed; writes into it are
uploaded before the next device operation.  Re-fetch with ``get`` after a
step (a view held across a step is not refreshed).
"""
import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray, ptr
from .gmg import Gmg

MODELS = {"euler": 0, "boussinesq": 1}
VARNAMES = {"euler": ["vorticity", "psi", "u", "v", "source"],
            "boussinesq": ["vorticity", "psi", "u", "v", "buoyancy", "banom"]}
TRACERS = {"euler": ["vorticity"], "boussinesq": ["vorticity", "buoyancy"]}


class Var:
    """Packed model state (core/variables.py:5-30) mirrored on the host."""

    def __init__(self, model, varname_list, dev):
        self._model = model
        self.varname_list = list(varname_list)
        self.nvar = len(self.varname_list)
        self.dev = dev  # DeviceArray (nvar, m, n)
        self._host = np.zeros(dev.shape)
        self._host_valid = True
        self._dirty = False

    @property
    def state(self):
        if not self._host_valid:
            self.pull()
        self._dirty = True
        return self._host

    def get(self, name):
        return self.state[self.varname_list.index(name)]

    def pull(self):
        self._host[...] = self.dev.get()
        self._host_valid = True
        self._dirty = False

    def push(self):
        self.dev.upload(self._host)
        self._dirty = False

    def sync_to_device(self):
        if self._dirty:
            self.push()

    def device_changed(self):
        self._host_valid = False
        self._dirty = False


class Operators:
    """core/operators.py Operators on the model's device buffers."""

    def __init__(self, model):
        self._m = model
        self.gmg = model.gmg
        self.cst = np.array([model.grid.dx, model.grid.dy, 0.05, 0.0,
                             model.aparab if model.order % 2 else 0.0])

    def rhs_adv(self, x, t, dxdt):
        """dxdt[tracer] = -div(u tracer) + periodic halo (operators.py:214-236).
        x, dxdt: device arrays shaped like the state."""
        self._m.var.sync_to_device()
        _lib.call("f2d_model_set_umax", self._m._h, float(self.cst[3]))
        _lib.call("f2d_model_rhs_adv", self._m._h, ptr(x), ptr(dxdt))

    def invert_vorticity(self, x, flag="full", island=False):
        """operators.py:421-498 on a device state-shaped array (island: the
        islands' rhs correction and boundary psi, when the model has them)."""
        self._m.var.sync_to_device()
        _lib.call("f2d_model_invert_vorticity", self._m._h, ptr(x),
                  (1 if flag == "full" else 0) | (2 if island else 0))
        if ptr(x) == self._m.var.dev.ptr:
            self._m.var.device_changed()

    def fill_halo(self, x):
        """Periodic halo fill (core/gmg/halo.py:136-141) of a host array."""
        self._m.grid.fill_halo(x)


class _Model:
    kind = None

    def __init__(self, param, grid):
        self.grid = grid
        self.order = int(param.order)
        self.aparab = float(param.aparab)
        self.timestepping = param.timestepping
        self.diffusion = bool(param.diffusion)
        kd = param.Kdiff
        if not isinstance(kd, dict):
            kd = {t: kd for t in TRACERS[self.kind]}
        self.Kdiff = kd
        self.gravity = float(param.gravity)
        self.nh = grid.nh
        self.msk = grid.msk
        self.area = grid.area
        p = _lib.f2d_model_params()
        p.model = MODELS[self.kind]
        p.nx, p.ny = grid.nx, grid.ny
        p.Lx, p.Ly = grid.Lx, grid.Ly
        p.order = self.order
        p.flux_splitting = ["minmax", "parabolic"].index(param.flux_splitting_method)
        p.aparab = self.aparab
        p.timestepping = ["RK3_SSP", "RK3"].index(self.timestepping)
        p.perio = 1 if grid.geometry == "perio" else 0
        p.diffusion = 1 if self.diffusion else 0
        p.kdiff_vort = float(kd.get("vorticity", 0.0))
        p.kdiff_buoy = float(kd.get("buoyancy", 0.0))
        p.gravity = self.gravity
        p.hydroepsilon = float(param.hydroepsilon)
        h = C.c_void_p()
        msk = np.ascontiguousarray(grid.msk, dtype=np.int8)
        _lib.call("f2d_model_create", C.byref(p), msk.ctypes.data_as(C.c_void_p), C.byref(h))
        self._h = h.value
        sp, nv, m, n = C.c_void_p(), C.c_int(), C.c_int(), C.c_int()
        _lib.call("f2d_model_state", self._h, C.byref(sp), C.byref(nv), C.byref(m), C.byref(n))
        self.var = Var(self, VARNAMES[self.kind],
                       DeviceArray((nv.value, m.value, n.value), np.float64, ptr=sp.value, owner=self))
        gp = C.c_void_p()
        _lib.call("f2d_model_gmg", self._h, C.byref(gp))
        self.gmg = Gmg(None, None, handle=gp.value)
        self.ope = Operators(self)
        self.diags = {"maxspeed": 0.0}
        self._coords_sent = False
        self.last_nite, self.last_res = 0, 0.0
        self.noslip = bool(getattr(param, "noslip", False))
        self.isisland = bool(getattr(param, "isisland", False))
        if self.noslip:  # Operators.set_boundary_msk (core/operators.py:97, 157-185)
            mskbc, bcarea, x2bc, y2bc = grid.boundary_mask()
            a = [np.ascontiguousarray(v) for v in
                 (grid.msknoslip.astype(np.int8), mskbc.astype(np.int8), grid.xr0, grid.yr0)]
            _lib.call("f2d_model_set_noslip", self._h, *[v.ctypes.data_as(C.c_void_p) for v in a],
                      float(bcarea), float(x2bc), float(y2bc),
                      1 if getattr(param, "enforce_momentum", False) else 0)
            self.mskbc, self.bcarea, self.x2bc, self.y2bc = mskbc, bcarea, x2bc, y2bc

    def set_islands(self, rhsp, psi):
        """core/fluid2d.py:110-113: the islands' rhs correction and boundary
        psi (core/island.py:21-43) for the state's inversions."""
        r = np.ascontiguousarray(rhsp, dtype=np.float64)
        p = np.ascontiguousarray(psi, dtype=np.float64)
        _lib.call("f2d_model_set_islands", self._h, r.ctypes.data_as(C.c_void_p),
                  p.ctypes.data_as(C.c_void_p))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().f2d_model_destroy(self._h)
            except Exception:
                pass
            self._h = None

    # -- tendencies (core/timescheme.py:35-41) on the device
    def tendency(self, k):
        p = C.c_void_p()
        _lib.call("f2d_model_tendency", self._h, k, C.byref(p))
        return DeviceArray(self.var.dev.shape, np.float64, ptr=p.value, owner=self)

    @property
    def mskp(self):
        p = C.c_void_p()
        _lib.call("f2d_model_corner_mask", self._h, C.byref(p))
        return DeviceArray(self.var.dev.shape[1:], np.int8, ptr=p.value, owner=self).get()

    def set_psi_from_vorticity(self):
        self.var.sync_to_device()
        _lib.call("f2d_model_set_psi_from_vorticity", self._h)
        self._read_solve()
        self.var.device_changed()

    def _read_solve(self):
        nite, res = C.c_int(), C.c_double()
        _lib.call("f2d_model_last_solve", self._h, C.byref(nite), C.byref(res))
        self.last_nite, self.last_res = nite.value, res.value

    def step(self, t, dt):
        """One RK step (core/euler.py:103-147, core/boussinesq.py:87-100);
        uses ope.cst[3] as umax like the reference."""
        self.var.sync_to_device()
        _lib.call("f2d_model_set_umax", self._h, float(self.ope.cst[3]))
        _lib.call("f2d_model_step", self._h, float(dt))
        self._read_solve()
        self.var.device_changed()

    def run(self, nsteps, dt):
        """``nsteps`` iterations of the Fluid2d.loop body with a fixed dt, on
        the device without host round trips beyond the solve norms and the
        per-step maxspeed; returns the F-cycle count of each step."""
        self.var.sync_to_device()
        self._raw_diag()  # the loop starts with diagnostics (core/fluid2d.py:199)
        log = np.zeros(nsteps, dtype=np.int32)
        _lib.call("f2d_model_run", self._h, int(nsteps), float(dt), log.ctypes.data_as(C.c_void_p))
        self._read_solve()
        self.var.device_changed()
        self.diagnostics(self.var, None)
        return log

    def _raw_diag(self):
        self.var.sync_to_device()
        xr, yr = (self.grid.xr, self.grid.yr)
        out = (C.c_double * 8)()
        if self._coords_sent:
            _lib.call("f2d_model_diagnostics", self._h, None, None, out)
        else:
            _lib.call("f2d_model_diagnostics", self._h, xr.ctypes.data_as(C.c_void_p),
                      yr.ctypes.data_as(C.c_void_p), out)
            self._coords_sent = True
        return np.array(out[:])


class Euler(_Model):
    """core/euler.py (no forcing, sponge or extra tracers)."""
    kind = "euler"

    def diagnostics(self, var, t):
        """core/euler.py:185-221"""
        o = self._raw_diag()
        a = self.area
        self.diags.update(maxspeed=o[0], ke=o[1] / a, vorticity=o[2] / a, enstrophy=0.5 * o[3] / a,
                          px=o[4] / a, py=o[5] / a, angmom=o[6] / a, source=o[7] / a)
        return self.diags


class Boussinesq(_Model):
    """core/boussinesq.py (no forcing)."""
    kind = "boussinesq"

    def diagnostics(self, var, t):
        """core/boussinesq.py:123-153"""
        o = self._raw_diag()
        a = self.area
        self.diags.update(maxspeed=o[0], ke=o[1] / a, pe=o[4] / a, energy=(o[1] + o[4]) / a,
                          vorticity=o[2] / a, enstrophy=0.5 * o[3] / a, buoyancy=o[5] / a,
                          brms=float(np.sqrt(max(o[6] / a - (o[5] / a) ** 2, 0.0))))
        return self.diags


class _EngineOps:
    """The Operators of an engine-backed model: cst (cst[3] = umax is kept by
    the engine itself, from the last diagnostics) and the periodic fill."""

    def __init__(self, model):
        self._m = model
        self.cst = np.array([model.grid.dx, model.grid.dy, 0.05, 0.0,
                             model.aparab if model.order % 2 else 0.0])

    def fill_halo(self, x):
        self._m.grid.fill_halo(x)

    def rhs_adv(self, x, t, dxdt):
        raise NotImplementedError("Fluid2d(param, grid) with param.engine = False gives the "
                                  "operator-level model")

    invert_vorticity = rhs_adv


class EngineModel:
    """Euler or Boussinesq (core/euler.py, core/boussinesq.py) on the batched
    engine: one FP64 instance whose step is a replayed hipGraph, bit-identical
    to :class:`Euler` / :class:`Boussinesq` and about three times faster at
    1024^2.  Same ``var``, ``step``, ``run``, ``diagnostics``,
    ``set_psi_from_vorticity`` and ``diags``; the engine computes the
    diagnostics at the end of each step and inversion, and again after a
    host write of the state.  No-slip walls, islands and the loop's
    momentum enforcement run in the engine; the flux diagnostics use the
    operator-level models."""

    def __init__(self, param, grid):
        from .batch import Batch, _call as bcall
        self.kind = param.modelname
        self.grid = grid
        self.nh = grid.nh
        self.msk = grid.msk
        self.area = grid.area
        self.order = int(param.order)
        self.aparab = float(param.aparab)
        # launch shapes picked by the autotuner (timed once per grid and
        # model on a scratch batch, then cached)
        from . import autotune
        cands = [{"coarse_lds_budget": 0, "min_tiles": mt, "coarse_wave_cells": wc}
                 for mt in (0, 1024, 4096) for wc in (256, 0)]
        tuning = None
        if getattr(param, "autotune", True):
            tuning, _ = autotune.tune([(param, grid)], "fp64", cands=cands)
        self.batch = Batch([(param, grid)], precision="fp64", tuning=tuning)
        p = C.c_void_p()
        bcall("f2d_batch_state_ptr", self.batch._h, 0, C.byref(p))
        nv = len(VARNAMES[self.kind])
        self.var = Var(self, VARNAMES[self.kind],
                       DeviceArray((nv, grid.nyl, grid.nxl), np.float64, ptr=p.value, owner=self))
        self.ope = _EngineOps(self)
        self.diags = {"maxspeed": 0.0}
        self._diag_fresh = False  # the engine's diagnostics match the state

    @property
    def last_nite(self):
        return self.batch.solve_info(0)[0]

    @property
    def last_res(self):
        return self.batch.solve_info(0)[1]

    def _sync(self):
        if self.var._dirty:
            self._diag_fresh = False
        self.var.sync_to_device()

    def set_psi_from_vorticity(self):
        self._sync()
        self.batch.set_psi_from_vorticity()
        self._diag_fresh = True
        self.var.device_changed()

    def step(self, t, dt):
        """One RK step with the engine's graph (core/euler.py:103-147,
        core/boussinesq.py:87-100)."""
        self._sync()
        if not self._diag_fresh:  # this step's umax comes from the diagnostics
            self.batch.refresh_diagnostics()
        self.batch.set_dt(dt, 0)
        self.batch.step(1)
        self._diag_fresh = True
        self.var.device_changed()

    def run(self, nsteps, dt):
        """``nsteps`` steps with a fixed dt; returns each step's F-cycle count."""
        self._sync()
        if not self._diag_fresh:
            self.batch.refresh_diagnostics()
        self._diag_fresh = True
        self.batch.set_dt(dt, 0)
        log = np.zeros(nsteps, dtype=np.int32)
        for k in range(nsteps):
            self.batch.step(1)
            log[k] = self.batch.solve_info(0)[0]
        self.var.device_changed()
        self.diagnostics(self.var, None)
        return log

    @property
    def mskp(self):
        return self.batch.corner_mask(0)

    def set_islands(self, rhsp, psi):
        """The batch took the grid's islands at construction (the same
        ``rhsp`` and ``psi``: core/island.py:21-43 is deterministic)."""

    def enforce_zero_momentum(self, px, py):
        """core/fluid2d.py:352-358 in the engine: vorticity -= xr0 px + yr0 py,
        the fast inversion of the state, fresh diagnostics."""
        self._sync()
        self.batch.enforce_momentum(0, px, py)
        self._diag_fresh = True
        self.var.device_changed()

    def diagnostics(self, var, t):
        """core/euler.py:185-221, core/boussinesq.py:123-153 (the engine's
        values of the current state, recomputed after a host write)."""
        self._sync()
        if not self._diag_fresh:
            self.batch.refresh_diagnostics()
            self._diag_fresh = True
        o = self.batch.raw_diagnostics(0)
        a = self.area
        if self.kind == "euler":
            self.diags.update(maxspeed=o[0], ke=o[1] / a, vorticity=o[2] / a,
                              enstrophy=0.5 * o[3] / a, px=o[4] / a, py=o[5] / a,
                              angmom=o[6] / a, source=o[7] / a)
        else:
            self.diags.update(maxspeed=o[0], ke=o[1] / a, pe=o[4] / a, energy=(o[1] + o[4]) / a,
                              vorticity=o[2] / a, enstrophy=0.5 * o[3] / a, buoyancy=o[5] / a,
                              brms=float(np.sqrt(max(o[6] / a - (o[5] / a) ** 2, 0.0))))
        return self.diags
