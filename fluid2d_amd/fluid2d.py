"""Fluid2d driver on the device: core/fluid2d.py ``Fluid2d`` (model factory,
``loop``, ``set_dt``) without output, plotting or logging."""
import time

import numpy as np

from .models import Boussinesq, Euler


class Fluid2d:
    def __init__(self, param, grid):
        param.checkall()
        self.param = param
        for k in ("modelname", "tend", "adaptable_dt", "dt", "cfl", "dtmax", "geometry",
                  "nx", "ny", "npx", "npy", "decay", "nprint"):
            setattr(self, k, getattr(param, k))
        self.dt0 = self.dt
        self.dx, self.dy = grid.dx, grid.dy
        if param.modelname == "euler":
            self.model = Euler(param, grid)
        elif param.modelname == "boussinesq":
            self.model = Boussinesq(param, grid)
        else:
            raise ValueError("model %s is outside the device path" % param.modelname)
        self.t = 0.0
        self.kt = 0
        self.stop = False
        self.walltime = 0.0

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

    def loop(self, nsteps=None):
        """core/fluid2d.py:188-349: until t >= tend (or ``nsteps`` steps),
        set_dt, step, t += dt, diagnostics, blow-up check."""
        self.model.diagnostics(self.model.var, self.t)
        t0 = time.perf_counter()
        k = 0
        while (self.t < self.tend if nsteps is None else k < nsteps) and not self.stop:
            self.set_dt(self.kt)
            self.model.step(self.t, self.dt)
            self.t += self.dt
            self.kt += 1
            k += 1
            self.model.diagnostics(self.model.var, self.t)
            if self.model.diags["maxspeed"] > 1e3:
                self.stop = True
                print("max|u| > 1000, blow-up detected, stopping")
        self.walltime = time.perf_counter() - t0
        return k
