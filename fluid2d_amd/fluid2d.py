"""Fluid2d driver on the device: core/fluid2d.py ``Fluid2d`` (model factory,
``loop``, ``set_dt``, ``enforce_zero_momentum``) and core/output.py's
history/diagnostic output, without plotting or logging."""
import os
import time

import numpy as np

from .models import Boussinesq, EngineModel, Euler


class Output:
    """The history/diagnostic schedule of core/output.py:56-96 (``Output.do``).

    A diagnostic record is taken whenever ``t >= tnextdiag`` (then
    ``tnextdiag += freq_diag``) and a history record whenever
    ``t >= tnexthis`` (then ``tnexthis += freq_his``); both clocks start at 0.

    With ``param.datadir`` set, the records go to NetCDF files in
    ``<datadir>/<expname>/`` (``<expname>_his.nc`` with ``var_to_save``,
    ``<expname>_diag.nc`` with ``list_diag``, ``<expname>_flx.nc``), laid out
    as the reference's (see :mod:`fluid2d_amd.ncio`).  Without it they stay
    in host memory: ``diag`` holds ``(t, kt, diags)``, ``his`` holds ``(t,
    kt, state)`` (a host copy of the packed state) and ``flx`` holds ``(t,
    kt, fluxes)`` with flux diagnostics on.
    """

    def __init__(self, param, grid=None, varname_list=None, flx_list=None):
        self.param, self.grid = param, grid
        self.freq_his = float(param.freq_his)
        self.freq_diag = float(param.freq_diag)
        self.tnexthis = 0.0
        self.tnextdiag = 0.0
        self.his = []
        self.diag = []
        self.flx = []
        self.expdir = None
        self.flx_list = flx_list
        self.first = True
        datadir = getattr(param, "datadir", None)
        self.keep_his = not datadir
        if datadir and grid is not None:
            datadir = datadir.replace("~", os.getenv("HOME", "~"))
            self.expdir = os.path.join(datadir, param.expname)
            os.makedirs(self.expdir, exist_ok=True)
            vs = param.var_to_save
            vs = list(varname_list) if vs == "all" else ([vs] if isinstance(vs, str) else list(vs))
            for v in vs:  # core/output.py:36-39
                if v not in varname_list:
                    raise ValueError("%s is not a model variable => modify param.var_to_save" % v)
            self.var_to_save, self.varname_list = vs, list(varname_list)
            self.rename_for_job(None)

    def rename_for_job(self, job):
        """File names of a run (job None) or of restart job ``job``
        (core/restart.py:74-89)."""
        if self.expdir is None:
            return
        e = self.param.expname
        if job is None:
            his, diag, flx = "%s_his.nc" % e, "%s_diag.nc" % e, "%s_flx.nc" % e
        else:
            his, flx = "%s_%02i_his_000.nc" % (e, job), "%s_%02i_flx_000.nc" % (e, job)
            diag = "%s_%02i_diag.nc" % (e, job)
        self.hisfile, self.diagfile, self.flxfile = (os.path.join(self.expdir, f)
                                                     for f in (his, diag, flx))

    def _create(self, data):
        from .ncio import DiagFile, HisFile
        ld = self.param.list_diag
        ld = list(data["diag"].keys()) if ld == "all" else ([ld] if isinstance(ld, str) else ld)
        self.nchis = HisFile(self.hisfile, self.param, self.grid, self.var_to_save,
                             self.varname_list)
        self.nchis.create()
        self.ncdiag = DiagFile(self.diagfile, ld)
        self.ncdiag.create()
        self.ncflx = None
        if "flx" in data:
            self.ncflx = HisFile(self.flxfile, self.param, self.grid, self.flx_list, self.flx_list)
            self.ncflx.create()

    def do(self, data, t, kt):
        """``data``: {'diag': dict, 'his': callable returning the state
        [, 'flx': callable returning the fluxes]}."""
        files = self.expdir is not None
        if files and self.first:
            self._create(data)
        self.first = False
        if t >= self.tnextdiag:
            self.tnextdiag += self.freq_diag
            self.diag.append((t, kt, dict(data["diag"])))
            if files:
                self.ncdiag.write(data["diag"], t, kt)
        if t >= self.tnexthis:
            self.tnexthis += self.freq_his
            x = data["his"]() if (files or self.keep_his) else None
            self.his.append((t, kt, np.array(x) if self.keep_his else None))
            if files:
                self.nchis.write(x, t, kt)
            if "flx" in data:
                fl = data["flx"]()
                self.flx.append((t, kt, fl if self.keep_his else None))
                if files:
                    self.ncflx.write(fl, t, kt)


class Fluid2d:
    def __init__(self, param, grid):
        param.checkall()
        self.param = param
        for k in ("modelname", "tend", "adaptable_dt", "dt", "cfl", "dtmax", "geometry",
                  "nx", "ny", "npx", "npy", "decay", "nprint", "exacthistime", "rescaledtime",
                  "enforce_momentum", "diag_fluxes"):
            setattr(self, k, getattr(param, k))
        self.dt0 = self.dt
        self.dx, self.dy = grid.dx, grid.dy
        self.grid = grid
        grid.finalize_msk()  # core/fluid2d.py:64: the user may have edited msk
        # core/fluid2d.py:67-77: momentum is enforced for Euler in closed
        # domains only
        if param.modelname == "euler":
            if self.geometry not in ("closed", "disc"):
                self.enforce_momentum = False
        elif param.modelname == "boussinesq":
            self.enforce_momentum = False
        else:
            raise ValueError("model %s is outside the device path" % param.modelname)
        # flux diagnostics re-run the time step term by term on the
        # operator-level models
        if getattr(param, "engine", True) and not self.diag_fluxes:
            self.model = EngineModel(param, grid)
            if self.enforce_momentum:
                self.model.batch.set_momentum(0)
        elif param.modelname == "euler":
            self.model = Euler(param, grid)
        else:
            self.model = Boussinesq(param, grid)
        if getattr(param, "isisland", False):  # core/fluid2d.py:110-113
            grid.island.finalize(self.model.mskp)
            self.model.set_islands(grid.island.rhsp, grid.island.psi)
        self.enstrophyname = "enstrophy"  # core/fluid2d.py:102-108
        if self.diag_fluxes:  # core/fluid2d.py:116-118
            from .fluxes import Fluxes
            self.flx = Fluxes(param, grid, self.model.ope)
        self.output = Output(param, grid, self.model.var.varname_list,
                             self.flx.fullflx_list if self.diag_fluxes else None)
        self.t = 0.0
        self.kt = 0
        self.stop = False
        self.walltime = 0.0
        self.ke_increases = 0

    def set_dt(self, kt=0):
        """core/fluid2d.py:376-397"""
        ms = self.model.diags["maxspeed"]
        if self.adaptable_dt and ms != 0:
            self.dt = self.cfl * min(self.dx, self.dy) / ms
            if self.dt > self.dtmax:
                self.dt = self.dtmax
        else:
            self.dt = self.dt0
        self.model.ope.cst[3] = ms

    def _exact_his_dt(self):
        """core/fluid2d.py:239-255: with exacthistime and an adaptive dt, the
        step is shortened over the 8 steps before the next history time so
        that one step lands on it."""
        tnext = self.output.tnexthis
        if self.t + 8 * self.dt > tnext and self._reduce == 0:
            self._reduce = 8
        if self._reduce > 0:
            self.dt = (tnext - self.t) / (self._reduce * 0.95)
            self._reduce -= 1
        if self.t + self.dt > tnext:
            self._reduce = 0
            self.dt = tnext - self.t

    def _state(self):
        # the device state, read without marking the host mirror as written
        self.model.var.sync_to_device()
        return self.model.var.dev.get()

    def _data(self):
        d = {"diag": self.model.diags, "his": self._state}
        if self.diag_fluxes:
            d["flx"] = self.flx.flx.get
        return d

    def loop(self, nsteps=None, joinhis=True, keepplotalive=False):
        """core/fluid2d.py:188-349: until t >= tend (or ``nsteps`` steps):
        set_dt (shortened to hit history times), step, advance t (by dt, or
        by dt * sqrt(enstrophy) with rescaledtime = 'enstrophy'),
        diagnostics, momentum enforcement, dke/dt and denstrophy/dt, the
        flux diagnostics before each history record, the output schedule,
        blow-up check."""
        self.model.diagnostics(self.model.var, self.t)
        self.model.diags["dkedt"] = 0.0
        self.model.diags["dvdt"] = 0.0
        if self.diag_fluxes:  # core/fluid2d.py:203-210: dt first
            self.set_dt(self.kt)
            self.flx.diag_fluxes(self.model.var.dev, self.t, self.dt)
        self.output.do(self._data(), self.t, self.kt)
        self._reduce = 0
        t0 = time.perf_counter()
        k = 0
        while (self.t < self.tend if nsteps is None else k < nsteps) and not self.stop:
            self.set_dt(self.kt)
            if self.exacthistime and self.adaptable_dt:
                self._exact_his_dt()
            self.model.step(self.t, self.dt)
            if self.rescaledtime == "enstrophy":
                self.t += self.dt * np.sqrt(self.model.diags["enstrophy"])
            else:
                self.t += self.dt
            self.kt += 1
            k += 1
            ke_old = self.model.diags["ke"]
            ens_old = self.model.diags[self.enstrophyname]
            self.model.diagnostics(self.model.var, self.t)
            if self.enforce_momentum:
                self.enforce_zero_momentum()
                self.model.diagnostics(self.model.var, self.t)
            ke = self.model.diags["ke"]
            ens = self.model.diags[self.enstrophyname]
            self.model.diags["dkedt"] = (ke - ke_old) / self.dt
            self.model.diags["dvdt"] = (ens - ens_old) / self.dt
            if ke > ke_old and self.decay and self.modelname == "euler":
                self.ke_increases += 1
            if self.diag_fluxes and self.t >= self.output.tnexthis:
                # core/fluid2d.py:293-295: only before a history record
                self.flx.diag_fluxes(self.model.var.dev, self.t, self.dt)
            self.output.do(self._data(), self.t, self.kt)
            if self.model.diags["maxspeed"] > 1e3:
                self.stop = True
                print("max|u| > 1000, blow-up detected, stopping")
        self.walltime = time.perf_counter() - t0
        return k

    def enforce_zero_momentum(self):
        """core/fluid2d.py:351-358: remove the linear vorticity profile that
        carries the domain's momentum, omega -= xr0 px / x2 + yr0 py / y2,
        then the fast inversion of the state."""
        if self.enforce_momentum:
            g = self.grid
            px = self.model.diags["px"] / g.x2
            py = self.model.diags["py"] / g.y2
            if isinstance(self.model, EngineModel):
                self.model.enforce_zero_momentum(px, py)
                return
            vor = self.model.var.get("vorticity")
            vor[:] -= g.xr0 * px + g.yr0 * py
            self.model.ope.invert_vorticity(self.model.var.dev, flag="fast")
