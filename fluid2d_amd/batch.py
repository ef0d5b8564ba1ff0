"""Batched engine: many independent Fluid2d instances advanced together.

``Batch([spec, ...])`` builds one engine for a list of instance specs (model,
grid, geometry, physics; one time scheme per batch).  Every kernel launch
covers all instances taking part in that phase of the step; the step runs
from a hipGraph with the multigrid stopping test and dt on the device.  An
instance's FP64 result is bit-identical whether it runs alone or in a mix.

A spec is a dict of :class:`fluid2d_amd.Param` keywords, or a ``(Param,
Grid)`` pair.  ``precision="mixed"`` keeps the state, the RK combination,
the full-solve residual and correction and the coarse-level solve in FP64
and runs advection, tendencies and the fine-level V-cycles in FP32.
"""
import ctypes as C

import numpy as np

from . import _lib
from .grid import Grid
from .models import VARNAMES
from .param import Param

_p = C.c_void_p
_i = C.c_int
_d = C.c_double

_SIG = {
    "f2d_batch_last_error": ([], C.c_char_p),
    "f2d_batch_create": ([_i, C.POINTER(_lib.f2d_model_params), C.POINTER(C.c_void_p), _i, _p,
                          C.POINTER(C.c_void_p)], _i),
    "f2d_batch_retune": ([_p, _p], _i),
    "f2d_batch_tuning_info": ([_p, _p], _i),
    "f2d_batch_destroy": ([_p], _i),
    "f2d_batch_info": ([_p, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(C.c_longlong)], _i),
    "f2d_batch_set_field": ([_p, _i, _i, _p], _i),
    "f2d_batch_get_state": ([_p, _i, _p], _i),
    "f2d_batch_state_ptr": ([_p, _i, C.POINTER(C.c_void_p)], _i),
    "f2d_batch_set_dt": ([_p, _i, _d], _i),
    "f2d_batch_set_adaptive": ([_p, _i, _d, _d], _i),
    "f2d_batch_set_fast_solve": ([_p, _i, _d], _i),
    "f2d_batch_set_noslip": ([_p, _i, _p, _p, _p, _p, _d, _d, _d, _i], _i),
    "f2d_batch_set_islands": ([_p, _i, _p, _p], _i),
    "f2d_batch_corner_mask": ([_p, _i, _p], _i),
    "f2d_batch_set_momentum": ([_p, _i, _p, _p], _i),
    "f2d_batch_enforce_momentum": ([_p, _i, _d, _d], _i),
    "f2d_batch_set_steps": ([_p, _i, _i], _i),
    "f2d_batch_progress": ([_p, _i, C.POINTER(_i), C.POINTER(C.c_longlong)], _i),
    "f2d_batch_use_graph": ([_p, _i], _i),
    "f2d_batch_set_psi_from_vorticity": ([_p], _i),
    "f2d_batch_reset_instance": ([_p, _i], _i),
    "f2d_batch_init_instance": ([_p, _i], _i),
    "f2d_batch_refresh_diagnostics": ([_p], _i),
    "f2d_batch_step": ([_p, _i], _i),
    "f2d_batch_step_async": ([_p, _i], _i),
    "f2d_batch_synchronize": ([_p], _i),
    "f2d_batch_query": ([_p, C.POINTER(_i)], _i),
    "f2d_batch_profile": ([_p, _i, _i, _i, C.POINTER(_d), C.POINTER(C.c_longlong), C.POINTER(_d),
                           C.POINTER(_d)], _i),
    "f2d_batch_diagnostics": ([_p, _i, C.POINTER(_d)], _i),
    "f2d_batch_solve_info": ([_p, _i, C.POINTER(_i), C.POINTER(_d), C.POINTER(_i)], _i),
}
_bound = False


def _L():
    global _bound
    lib = _lib.load()
    if not _bound:
        for name, (args, res) in _SIG.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _bound = True
    return lib


def _call(name, *args):
    rc = getattr(_L(), name)(*args)
    if rc != 0:
        msg = _L().f2d_batch_last_error().decode()
        if rc == _lib.F2D_E_DIVERGED:
            raise _lib.DivergenceError(rc, msg)
        raise _lib.F2DError(rc, msg)


class Tuning(C.Structure):
    """f2d_batch_tuning: coarse-kernel LDS budget (bytes, 0 = device
    maximum), the minimum workgroup count of a fine phase (0 = always
    16-row tiles), the largest coarse level (cells) run on one wavefront
    (-1 = default 256) and the persistent multigrid launches (workgroups per
    compute unit, 0 = one launch per phase)."""
    _fields_ = [("coarse_lds_budget", C.c_int), ("min_tiles", C.c_int),
                ("coarse_wave_cells", C.c_int), ("persist", C.c_int)]

    def as_dict(self):
        return {"coarse_lds_budget": self.coarse_lds_budget, "min_tiles": self.min_tiles,
                "coarse_wave_cells": self.coarse_wave_cells, "persist": self.persist}


def _tuning(t):
    if t is None:
        return None
    if isinstance(t, dict):
        t = Tuning(int(t.get("coarse_lds_budget", 0)), int(t.get("min_tiles", 0)),
                   int(t.get("coarse_wave_cells", -1)), int(t.get("persist", 0)))
    return C.byref(t)


def params_struct(param, grid):
    p = _lib.f2d_model_params()
    p.model = {"euler": 0, "boussinesq": 1}[param.modelname]
    p.nx, p.ny = grid.nx, grid.ny
    p.Lx, p.Ly = grid.Lx, grid.Ly
    p.order = int(param.order)
    p.flux_splitting = ["minmax", "parabolic"].index(param.flux_splitting_method)
    p.aparab = float(param.aparab)
    p.timestepping = ["RK3_SSP", "RK3"].index(param.timestepping)
    p.perio = 1 if grid.geometry == "perio" else 0
    p.diffusion = 1 if param.diffusion else 0
    kd = param.Kdiff if isinstance(param.Kdiff, dict) else {"vorticity": param.Kdiff,
                                                             "buoyancy": param.Kdiff}
    p.kdiff_vort = float(kd.get("vorticity", 0.0))
    p.kdiff_buoy = float(kd.get("buoyancy", 0.0))
    p.gravity = float(param.gravity)
    p.hydroepsilon = float(param.hydroepsilon)
    return p


class Batch:
    def __init__(self, specs, precision="fp64", tuning=None, use_graph=True):
        self.params, self.grids = [], []
        for s in specs:
            if isinstance(s, tuple):
                prm, g = s
            else:
                prm = Param(**s)
                g = Grid(prm)
            prm.checkall()
            self.params.append(prm)
            self.grids.append(g)
        n = len(self.params)
        arr = (_lib.f2d_model_params * n)(*[params_struct(p, g) for p, g in zip(self.params, self.grids)])
        self._msks = [np.ascontiguousarray(g.msk, dtype=np.int8) for g in self.grids]
        mp = (C.c_void_p * n)(*[m.ctypes.data for m in self._msks])
        h = C.c_void_p()
        self.precision = precision
        prec = {"fp64": 0, "mixed": 1}[precision]
        _call("f2d_batch_create", n, arr, mp, prec, _tuning(tuning), C.byref(h))
        self._h = h.value
        self.n = n
        _call("f2d_batch_use_graph", self._h, 1 if use_graph else 0)
        self.varnames = [VARNAMES[p.modelname] for p in self.params]
        for i, (prm, g) in enumerate(zip(self.params, self.grids)):
            if getattr(prm, "noslip", False):
                # Operators copies enforce_momentum from the parameters as
                # given (core/operators.py:17), not as Fluid2d narrows it
                self.set_noslip(i, enforce=bool(getattr(prm, "enforce_momentum", False)))
            if getattr(prm, "isisland", False):
                g.island.finalize()
                self.set_islands(i, g.island.rhsp, g.island.psi)

    def set_noslip(self, inst, enforce=False):
        """No-slip walls on instance ``inst`` (core/operators.py:157-185,
        :245-290) from its grid's ``msknoslip`` and boundary mask; with
        ``enforce`` the source's momentum is removed as well."""
        g = self.grids[inst]
        mskbc, bcarea, x2bc, y2bc = g.boundary_mask()
        a = [np.ascontiguousarray(v) for v in (g.msknoslip.astype(np.int8), mskbc.astype(np.int8),
                                              np.asarray(g.xr0, np.float64),
                                              np.asarray(g.yr0, np.float64))]
        _call("f2d_batch_set_noslip", self._h, inst, *[v.ctypes.data for v in a],
              float(bcarea), float(x2bc), float(y2bc), 1 if enforce else 0)

    def set_islands(self, inst, rhsp, psi):
        """Islands on instance ``inst``: the rhs correction ``rhsp`` and the
        boundary psi of core/island.py:21-43."""
        r = np.ascontiguousarray(rhsp, dtype=np.float64)
        p = np.ascontiguousarray(psi, dtype=np.float64)
        _call("f2d_batch_set_islands", self._h, inst, r.ctypes.data, p.ctypes.data)

    def corner_mask(self, inst):
        """The instance's corner mask mskp (int8, the state's plane shape)."""
        g = self.grids[inst]
        out = np.zeros((g.nyl, g.nxl), dtype=np.int8)
        _call("f2d_batch_corner_mask", self._h, inst, out.ctypes.data)
        return out

    def set_momentum(self, inst):
        """Enable the loop's momentum enforcement on instance ``inst``."""
        g = self.grids[inst]
        x = np.ascontiguousarray(g.xr0, dtype=np.float64)
        y = np.ascontiguousarray(g.yr0, dtype=np.float64)
        _call("f2d_batch_set_momentum", self._h, inst, x.ctypes.data, y.ctypes.data)

    def enforce_momentum(self, inst, px, py):
        """Fluid2d.enforce_zero_momentum (core/fluid2d.py:352-358) on instance
        ``inst``: vorticity -= xr0 px + yr0 py, fast inversion, diagnostics."""
        _call("f2d_batch_enforce_momentum", self._h, inst, float(px), float(py))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _L().f2d_batch_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return self.n

    def retune(self, tuning):
        """Rebuild the launch program for a new tuning (dict or Tuning)."""
        _call("f2d_batch_retune", self._h, _tuning(tuning))

    def tuning(self):
        t = Tuning()
        _call("f2d_batch_tuning_info", self._h, C.byref(t))
        return t.as_dict()

    def info(self, i):
        nl, nf, nlaunch = C.c_int(), C.c_int(), C.c_longlong()
        _call("f2d_batch_info", self._h, i, C.byref(nl), C.byref(nf), C.byref(nlaunch))
        return {"nlev": nl.value, "nfine": nf.value, "launches_per_step": nlaunch.value}

    def set_field(self, i, name, a):
        k = self.varnames[i].index(name)
        a = np.ascontiguousarray(a, dtype=np.float64)
        g = self.grids[i]
        if a.shape != (g.nyl, g.nxl):
            raise ValueError("field shape %s != %s" % (a.shape, (g.nyl, g.nxl)))
        _call("f2d_batch_set_field", self._h, i, k, a.ctypes.data_as(C.c_void_p))

    def state(self, i):
        g = self.grids[i]
        out = np.empty((len(self.varnames[i]), g.nyl, g.nxl))
        _call("f2d_batch_get_state", self._h, i, out.ctypes.data_as(C.c_void_p))
        return out

    def set_dt(self, dt, i=-1):
        _call("f2d_batch_set_dt", self._h, int(i), float(dt))

    def set_steps(self, nsteps, i=-1):
        """Step budget of instance i (all if i < 0; nsteps < 0: unlimited)."""
        _call("f2d_batch_set_steps", self._h, int(i), int(nsteps))

    def progress(self, i):
        left, done = C.c_int(), C.c_longlong()
        _call("f2d_batch_progress", self._h, i, C.byref(left), C.byref(done))
        return left.value, done.value

    def set_adaptive(self, on=True, cfl=0.5, dtmax=5.0):
        _call("f2d_batch_set_adaptive", self._h, 1 if on else 0, float(cfl), float(dtmax))

    def set_fast_solve(self, maxite, tol=1e-10):
        """Per-stage inversions by gmg.solve(maxite, tol) -- the alternative
        core/operators.py:452-455 keeps commented out -- instead of
        twoVcycle; in mixed precision this is FP64 iterative refinement of
        FP32 F-cycles.  maxite = 0 restores the reference's twoVcycle."""
        _call("f2d_batch_set_fast_solve", self._h, int(maxite), float(tol))

    def set_psi_from_vorticity(self):
        _call("f2d_batch_set_psi_from_vorticity", self._h)

    def reset_instance(self, i):
        """Instance i becomes a fresh slot (its carried buffers zeroed, its
        device record reset); the other instances are untouched."""
        _call("f2d_batch_reset_instance", self._h, int(i))

    def init_instance(self, i):
        """set_psi_from_vorticity of instance i alone (enqueued)."""
        _call("f2d_batch_init_instance", self._h, int(i))

    def refresh_diagnostics(self):
        """Recompute every instance's diagnostics (and next umax) from its
        current state."""
        _call("f2d_batch_refresh_diagnostics", self._h)

    def step(self, nsteps=1):
        _call("f2d_batch_step", self._h, int(nsteps))

    def step_async(self, nsteps=1):
        _call("f2d_batch_step_async", self._h, int(nsteps))

    def synchronize(self):
        _call("f2d_batch_synchronize", self._h)

    def done(self):
        """True when all enqueued work on this batch's stream has finished."""
        d = C.c_int()
        _call("f2d_batch_query", self._h, C.byref(d))
        return bool(d.value)

    KINDS = {"smooth": 0, "resid_restrict": 1, "rhs": 2, "coarse": 3, "persist": 4,
             "smooth_rr": 5, "up2": 6}

    def profile(self, kind, level=-1, nsteps=2):
        """Time every launch of ``kind`` at ``level`` over ``nsteps`` eager
        steps with HIP events; returns mean ms and bytes per launch."""
        ms, nl, ab, ib = C.c_double(), C.c_longlong(), C.c_double(), C.c_double()
        _call("f2d_batch_profile", self._h, self.KINDS[kind], int(level), int(nsteps),
              C.byref(ms), C.byref(nl), C.byref(ab), C.byref(ib))
        return {"avg_ms": ms.value, "launches": nl.value, "alg_bytes": ab.value,
                "impl_bytes": ib.value}

    def raw_diagnostics(self, i):
        out = (C.c_double * 8)()
        _call("f2d_batch_diagnostics", self._h, i, out)
        return np.array(out[:])

    def diagnostics(self, i):
        o = self.raw_diagnostics(i)
        a = self.grids[i].area
        d = {"maxspeed": o[0], "ke": o[1] / a, "vorticity": o[2] / a, "enstrophy": 0.5 * o[3] / a}
        if self.params[i].modelname == "euler":
            d.update(px=o[4] / a, py=o[5] / a, angmom=o[6] / a, source=o[7] / a)
        else:
            d.update(pe=o[4] / a, energy=(o[1] + o[4]) / a, buoyancy=o[5] / a)
        return d

    def solve_info(self, i):
        nite, res, div = C.c_int(), C.c_double(), C.c_int()
        _call("f2d_batch_solve_info", self._h, i, C.byref(nite), C.byref(res), C.byref(div))
        return nite.value, res.value, bool(div.value)
