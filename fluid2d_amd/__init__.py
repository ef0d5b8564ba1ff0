"""fluid2d_amd -- the Fluid2d timestep hot path (advection, psi/u,v
operators, geometric multigrid, RK3) on AMD MI355X (gfx950) HIP kernels.

Drop-in surface (reference names): Param, Grid, Fluid2d, Euler, Boussinesq,
Gmg, and the kernel-level module ``fluid2d_amd.kernels``.
"""
from ._lib import DivergenceError, F2DError, device_count, load  # noqa: F401
from .device import DeviceArray  # noqa: F401
from .fluid2d import Fluid2d  # noqa: F401
from .gmg import Gmg  # noqa: F401
from .grid import Grid  # noqa: F401
from .models import Boussinesq, EngineModel, Euler  # noqa: F401
from .param import Param  # noqa: F401

__version__ = "0.1.0"
