"""Geometric multigrid on the device: drop-in for core/gmg/hierarchy.py Gmg.

``Gmg(param, mskp)`` builds the level hierarchy, masks and operators on the
GPU (core/gmg/hierarchy.py:23-84, core/gmg/level.py:24-335) and exposes the
reference's cycle methods.  Arrays passed to ``solve`` / ``twoVcycle`` are
device arrays of the finest level's shape.
"""
import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray, ptr


class Gmg:
    """param: the dict core/operators.py:35-41 builds ('n', 'm', 'nh', 'dx',
    'dy', 'omega', 'hydroepsilon', ...; with 'qgoperator' and 'Rd' the
    operator is the Helmholtz A - 1/Rd^2, core/operators.py:74-77);
    mskp: level-0 corner mask (host)."""

    def __init__(self, param, mskp, handle=None):
        self._own = handle is None
        if handle is None:
            h = C.c_void_p()
            mk = np.ascontiguousarray(mskp, dtype=np.int8)
            _lib.call("f2d_gmg_create", int(param["n"]), int(param["m"]), int(param.get("nh", 3)),
                      float(param["dx"]), float(param["dy"]), float(param.get("hydroepsilon", 1.0)),
                      float(param.get("omega", 8.0 / 9.0)), mk.ctypes.data_as(C.c_void_p),
                      C.byref(h))
            handle = h.value
        self._h = handle
        if self._own and "qgoperator" in param:
            # core/gmg/hierarchy.py:80-84: the Helmholtz operator A - 1/Rd^2
            _lib.call("f2d_gmg_set_helmholtz", self._h, float(param["Rd"]))
        nl = C.c_int()
        _lib.call("f2d_gmg_nlevels", self._h, C.byref(nl))
        self.nlevs = nl.value
        self.npre = self.npost = 1
        self.ndeepest = 16

    def __del__(self):
        if getattr(self, "_own", False) and getattr(self, "_h", None):
            try:
                _lib.load().f2d_gmg_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def level(self, lev, which):
        """Device view of level ``lev``: which in msk, A, x, b, r."""
        code = {"msk": 0, "A": 1, "x": 2, "b": 3, "r": 4}[which]
        p, m, n = C.c_void_p(), C.c_int(), C.c_int()
        _lib.call("f2d_gmg_level", self._h, lev, code, C.byref(p), C.byref(m), C.byref(n))
        if which == "msk":
            return DeviceArray((m.value, n.value), np.int8, ptr=p.value, owner=self)
        if which == "A":
            return DeviceArray((5, m.value, n.value), np.float64, ptr=p.value, owner=self)
        return DeviceArray((m.value, n.value), np.float64, ptr=p.value, owner=self)

    def Vcycle(self, lev1):
        _lib.call("f2d_gmg_vcycle", self._h, lev1)

    def Fcycle(self, lev1):
        _lib.call("f2d_gmg_fcycle", self._h, lev1)

    def solve(self, x, b, param=None):
        """core/gmg/hierarchy.py:154-192; raises DivergenceError where the
        reference calls exit(0)."""
        param = param or {}
        nite, res = C.c_int(), C.c_double()
        _lib.call("f2d_gmg_solve", self._h, ptr(x), ptr(b), int(param.get("maxite", 4)),
                  float(param.get("tol", 1e-11)), C.byref(nite), C.byref(res))
        return nite.value, res.value

    def twoVcycle(self, x, b, param=None):
        _lib.call("f2d_gmg_twovcycle", self._h, ptr(x), ptr(b))
        return 1, 0.0
