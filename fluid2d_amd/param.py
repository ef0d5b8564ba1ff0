"""Parameters of the timestep path, with the reference's names and defaults.

Restates the entries of core/defaults.json that the path reads (and the
``checkall`` validation of core/param.py:78-86 for them).  Unknown names are
accepted and ignored, like the reference's Param, so experiment scripts can
set output/plotting options freely.
"""

DEFAULTS = {
    # general
    "modelname": "euler",
    "expname": "myexp",
    # numerics (core/defaults.json:24-60)
    "timestepping": "RK3_SSP",
    "order": 5,
    "aparab": 0.05,
    "flux_splitting_method": "parabolic",
    "relaxation": "default",
    "nh": 3,
    # time
    "adaptable_dt": True,
    "exacthistime": True,
    "rescaledtime": "none",
    "dt": 0.1,
    "cfl": 0.5,
    "dtmax": 5.0,
    "tend": 10.0,
    # domain
    "nx": 128,
    "ny": 128,
    "Lx": 1.0,
    "Ly": 1.0,
    "geometry": "disc",
    "npx": 1,
    "npy": 1,
    "myrank": 0,
    # physics
    "hydroepsilon": 1.0,
    "diffusion": False,
    "Kdiff": 0.0,
    "gravity": 1.0,
    "noslip": False,
    "forcing": False,
    "isisland": False,
    "enforce_momentum": False,
    "decay": True,
    # output (core/defaults.json:229-273, 103-107); files are written only
    # when datadir is set (the reference's default is ~/data/fluid2d)
    "datadir": None,
    "var_to_save": "vorticity",
    "list_diag": "all",
    "ninterrestart": 1,
    "freq_his": 1.0,
    "diag_fluxes": False,
    "freq_diag": 1.0,
    "nprint": 20,
    # this package: Fluid2d runs the step on the batched engine (one
    # instance, graph-replayed) unless an option needs the operator-level
    # model (noslip, isisland, enforce_momentum, diag_fluxes)
    "engine": True,
    "autotune": True,  # the engine's launch shapes from the autotuner (cached)
    "plot_interactive": False,
}

AVAIL = {
    "modelname": ["euler", "boussinesq"],
    # the reference's driver accepts only these; 'RK3' is reachable only by
    # building Timescheme directly (core/timescheme.py:152-164), offered here
    "timestepping": ["RK3_SSP", "RK3"],
    "flux_splitting_method": ["minmax", "parabolic"],
    "geometry": ["disc", "perio", "closed", "ychannel", "xchannel", "xperio", "yperio"],
    "relaxation": ["default"],
    "rescaledtime": ["none", "enstrophy"],
}

UNSUPPORTED_TRUE = ["forcing"]


class Param:
    def __init__(self, **kw):
        for k, v in DEFAULTS.items():
            setattr(self, k, v)
        for k, v in kw.items():
            setattr(self, k, v)

    def checkall(self):
        """Validate the options the device path reads (core/param.py:78-86)."""
        for k, avail in AVAIL.items():
            v = getattr(self, k)
            if v not in avail:
                raise ValueError("%s = %r not in %s" % (k, v, avail))
        if self.nh != 3:
            raise ValueError("nh = 3 is compulsory (core/fortran_advection.f90:30-34)")
        if not 1 <= int(self.order) <= 6:
            raise ValueError("order must be 1..6")
        if self.npx * self.npy != 1:
            raise ValueError("one subdomain per GPU instance (npx = npy = 1)")
        for k in UNSUPPORTED_TRUE:
            if getattr(self, k):
                raise ValueError("%s is outside the device path" % k)

    def copy(self, obj, list_param):
        for k in list_param:
            setattr(obj, k, getattr(self, k))
