"""Flux diagnostics on the device: core/fluxes.py ``Fluxes``.

``Fluxes(param, grid, ope).diag_fluxes(x, t, dt)`` splits the advective
edge fluxes of vorticity [, buoyancy] and of the cell-centred velocities
(uc, vc) over one step dt into a reversible and an irreversible part: the
extended state (the model variables, uc, vc, zeroed flux planes) is stepped
forward by dt with the flux-recording advection (core/fortran_fluxes.f90),
then, from the time-reversed state, by -dt; the halves of the sum and the
difference of the two time-integrated fluxes, divided by dt, are the two
parts (core/fluxes.py:100-170).  The whole computation runs on the GPU in
one call into libf2dgpu.so; ``flx`` is a device array of
``len(fullflx_list)`` planes, ``flx.get()`` brings it to the host.
"""
import numpy as np

from . import _lib
from .device import DeviceArray, ptr


class Fluxes:
    def __init__(self, param, grid, ope):
        self._model = ope._m
        if not hasattr(self._model, "_h"):
            raise ValueError("flux diagnostics run on the operator-level model: build Fluid2d "
                             "with param.diag_fluxes = True or param.engine = False")
        kind = self._model.kind
        self.tracer_list = ["vorticity"] + (["buoyancy"] if kind == "boussinesq" else [])
        ext = self.tracer_list + ["uc", "vc"]
        self.flx_list = ["flx_%s_%s" % (v, d) for v in ext for d in ("x", "y")]
        self.fullflx_list = ["%s_%s_%s" % (r, d, v) for r in ("rev", "irr") for v in ext
                             for d in ("x", "y")]
        m, n = self._model.var.dev.shape[1:]
        self.flx = DeviceArray((len(self.fullflx_list), m, n), np.float64)

    def diag_fluxes(self, x, t, dt):
        """core/fluxes.py:100-170.  ``x``: the model state (its host mirror
        ``model.var.state`` or a state-shaped device array)."""
        mod = self._model
        mod.var.sync_to_device()
        xd = x if isinstance(x, DeviceArray) else mod.var.dev
        _lib.call("f2d_model_set_umax", mod._h, float(mod.ope.cst[3]))
        _lib.call("f2d_model_diag_fluxes", mod._h, ptr(xd), float(dt), ptr(self.flx))
