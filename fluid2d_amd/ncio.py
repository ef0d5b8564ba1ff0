"""History, diagnostic and restart files: core/output.py (Output, NcfileIO)
and core/restart.py (Restart) on NetCDF files.

The reference writes NETCDF4 (HDF5) files through netCDF4-python.  This
package writes the same dimensions, variables, types, attributes and file
names in the NetCDF classic format (64-bit offset) through
``scipy.io.netcdf_file``; netCDF4-python, xarray and ncdump read both
formats alike.

* ``<expdir>/<expname>_his.nc``: dims t (unlimited), x, y (interior);
  variables x, y (float32 cell centres), msk (int32), t (float32), kt
  (int32) and each saved variable (float32, t x y x interior); every
  parameter as a global attribute (booleans as 0/1, lists and dicts
  skipped) -- core/output.py:185-246.
* ``<expdir>/<expname>_diag.nc``: t, kt and each diagnostic (float32) per
  record -- core/output.py:98-155.
* ``<expdir>/<expname>_flx.nc``: the flux diagnostics, history layout.
* restart ``<expdir>/<expname>_<job:02d>_restart_000.nc``: global
  attributes tend, t, dt, kt, tnextdiag, tnexthis and every state variable
  in FP64 with its halo -- core/restart.py:140-177.
"""
import glob
import os

import numpy as np
from scipy.io import netcdf_file

_RESERVED = set(dir(netcdf_file)) | {"fp", "filename", "mode", "use_mmap", "version_byte",
                                     "maskandscale", "dimensions", "variables", "_dims",
                                     "_recs", "_recsize", "_attributes", "_mm", "_mm_buf"}


def _set_param_attributes(nc, param):
    """core/output.py:187-205: each parameter as a global attribute."""
    for k, v in vars(param).items():
        if k.startswith("_") or k in _RESERVED or v is None:
            continue
        if isinstance(v, (list, tuple, dict)):
            continue
        if isinstance(v, (bool, np.bool_)):
            v = 1 if v else 0
        if isinstance(v, str):
            setattr(nc, k, v)
        elif isinstance(v, (int, np.integer)) and -2**31 <= int(v) < 2**31:
            setattr(nc, k, np.int32(v))
        elif isinstance(v, (int, float, np.integer, np.floating)):
            setattr(nc, k, np.float64(v))  # a Python float would be stored as float32


class HisFile:
    """NcfileIO (core/output.py:167-246): a file of 2D snapshots."""

    def __init__(self, path, param, grid, var_to_save, varname_list):
        self.path = path
        self.var_to_save = list(var_to_save)
        self.varname_list = list(varname_list)
        self.param, self.grid = param, grid
        self.khis = 0

    def create(self):
        g, nh = self.grid, self.grid.nh
        with netcdf_file(self.path, "w", version=2) as nc:
            _set_param_attributes(nc, self.param)
            nc.createDimension("t", None)
            nc.createDimension("x", g.nxl - 2 * nh)
            nc.createDimension("y", g.nyl - 2 * nh)
            d = nc.createVariable("x", "f", ("x",))
            d.long_name = "coordinate in the x-direction (at cell center)"
            d[:] = g.xr[0, nh:-nh]
            d = nc.createVariable("y", "f", ("y",))
            d.long_name = "coordinate in the y-direction (at cell center)"
            d[:] = g.yr[nh:-nh, 0]
            d = nc.createVariable("msk", "i", ("y", "x"))
            d.long_name = "mask at cell center"
            d[:, :] = g.msk[nh:-nh, nh:-nh]
            d = nc.createVariable("t", "f", ("t",))
            d.long_name = "model time"
            d = nc.createVariable("kt", "i", ("t",))
            d.long_name = "model iteration"
            for v in self.var_to_save:
                d = nc.createVariable(v, "f", ("t", "y", "x"))
                d.long_name = v
        self.khis = 0

    def write(self, x, t, kt):
        nh = self.grid.nh
        with netcdf_file(self.path, "a", version=2) as nc:
            nc.variables["t"][self.khis] = t
            nc.variables["kt"][self.khis] = kt
            for v in self.var_to_save:
                iv = self.varname_list.index(v)
                nc.variables[v][self.khis] = x[iv][nh:-nh, nh:-nh]
        self.khis += 1


class DiagFile:
    """The diagnostics file (core/output.py:98-155), written record by
    record (the reference buffers 10 records between writes)."""

    def __init__(self, path, list_diag):
        self.path = path
        self.list_diag = list(list_diag)
        self.kdiag = 0

    def create(self):
        with netcdf_file(self.path, "w", version=2) as nc:
            nc.createDimension("t", None)
            d = nc.createVariable("t", "f", ("t",))
            d.long_name = "model time"
            d = nc.createVariable("kt", "i", ("t",))
            d.long_name = "model iteration"
            for v in self.list_diag:
                d = nc.createVariable(v, "f", ("t",))
                d.long_name = v
        self.kdiag = 0

    def write(self, diag, t, kt):
        with netcdf_file(self.path, "a", version=2) as nc:
            nc.variables["t"][self.kdiag] = t
            nc.variables["kt"][self.kdiag] = kt
            for v in self.list_diag:
                nc.variables[v][self.kdiag] = diag[v]
        self.kdiag += 1


def write_restart(path, tend, t, dt, kt, tnextdiag, tnexthis, varname_list, state):
    """Restart.write (core/restart.py:140-160): FP64 fields with halo."""
    m, n = state.shape[-2:]
    with netcdf_file(path, "w", version=2) as nc:
        f64 = np.float64
        nc.tend, nc.t, nc.dt, nc.kt = f64(tend), f64(t), f64(dt), np.int32(kt)
        nc.tnextdiag, nc.tnexthis = f64(tnextdiag), f64(tnexthis)
        nc.createDimension("x", n)
        nc.createDimension("y", m)
        for k, v in enumerate(varname_list):
            d = nc.createVariable(v, "d", ("y", "x"))
            d[:, :] = state[k]


def read_restart(path, varname_list, state):
    """Restart.read (core/restart.py:162-177): fills ``state`` in place and
    returns (tend, t, dt, kt, tnextdiag, tnexthis)."""
    with netcdf_file(path, "r", mmap=False) as nc:
        out = (float(nc.tend), float(nc.t), float(nc.dt), int(nc.kt), float(nc.tnextdiag),
               float(nc.tnexthis))
        for k, v in enumerate(varname_list):
            state[k][...] = nc.variables[v][:, :]
    return out


class Restart:
    """core/restart.py Restart: ``Restart(param, grid, f2d)`` runs
    ``f2d.loop`` in ``param.ninterrestart`` slices of ``param.tend``,
    writing a restart file after each, and first resumes from the last
    restart found in the experiment folder (file index = job number)."""

    def __init__(self, param, grid, f2d, launch=True):
        self.expname = param.expname
        self.expdir = f2d.output.expdir
        if self.expdir is None:
            raise ValueError("restarts need param.datadir (the experiment folder)")
        self.ninterrestart = int(getattr(param, "ninterrestart", 1))
        self.varname_list = f2d.model.var.varname_list
        self.template = os.path.join(self.expdir, "%s_%02i_restart_%03i.nc")
        self.get_lastrestart()
        self.timelength = f2d.tend
        if self.lastrestart is not None:
            self.restart_file = self.template % (self.expname, self.lastrestart, 0)
            tend, t, dt, kt, tnextdiag, tnexthis = read_restart(
                self.restart_file, self.varname_list, f2d.model.var.state)
            f2d.tend += round(t)  # core/restart.py:61
            f2d.t, f2d.dt, f2d.kt = t, dt, kt
            f2d.output.tnextdiag, f2d.output.tnexthis = tnextdiag, tnexthis
            # the host copy is now the state; the model uploads it (and
            # recomputes its diagnostics) before its next device operation
            self.nextrestart = self.lastrestart + 1
        else:
            self.nextrestart = 0
        f2d.output.rename_for_job(self.nextrestart)
        self.lengthsubint = self.timelength / self.ninterrestart
        if launch:
            self.launchf2d(f2d)

    def launchf2d(self, f2d):
        """core/restart.py:97-122"""
        for kres in range(self.ninterrestart):
            f2d.tend = f2d.t + self.lengthsubint
            f2d.loop()
            self.restart_file = self.template % (self.expname, self.nextrestart, 0)
            write_restart(self.restart_file, f2d.tend, f2d.t, f2d.dt, f2d.kt,
                          f2d.output.tnextdiag, f2d.output.tnexthis, self.varname_list,
                          f2d.model.var.dev.get())

    def get_lastrestart(self):
        """core/restart.py:124-138"""
        files = glob.glob(os.path.join(self.expdir, "%s_*_restart_000.nc" % self.expname))
        self.lastrestart = None
        for f in files:
            b = os.path.basename(f)
            pos = b.rfind("restart")
            k = int(b[pos - 3:pos - 1])
            self.lastrestart = k if self.lastrestart is None else max(self.lastrestart, k)
