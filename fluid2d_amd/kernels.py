"""Kernel-level drop-in for the reference's f2py modules on this path.

Same function names and argument meaning as the Fortran routines the
reference's Python calls (``fa.adv_upwind``, ``fo.celltocorner``,
``fortmg.smoothtwicewitha`` ...), on device arrays (:class:`DeviceArray` or
any object with ``data_ptr()``).  Every call enqueues HIP kernels from
libf2dgpu.so on the null stream; there is no CPU path.

Where the reference always follows a routine with ``fill_halo`` at its call
site, the GPU kernel writes the halo itself and the function name says so
(``smoothtwicewitha_fill`` etc.).

================================  =============================================
function                          reference
================================  =============================================
adv_upwind, adv_centered          core/fortran_advection.f90:2-284 (and
                                  core/fortran_fluxes.f90 with xflx/yflx)
celltocorner                      core/fortran_operators.f90:44-64
computeorthogradient              core/fortran_operators.f90:2-39
add_diffusion_fill                core/fortran_operators.f90:125-156 + fill
add_torque_fill                   y *= msk; core/fortran_operators.f90:330-381 + fill
fillhalo                          core/gmg/fortran_multigrid.f90:365-412
smoothtwicewitha_fill             core/gmg/fortran_multigrid.f90:2-127 + fill
computeresidualwitha_fill         core/gmg/fortran_multigrid.f90:320-362 + fill
restrict_fill                     core/gmg/fortran_multigrid.f90:501-546 + fill
interpolate                       core/gmg/fortran_multigrid.f90:415-498
coarsenmatrix                     core/gmg/fortran_multigrid.f90:706-811 + fill
computenorm                       core/gmg/fortran_multigrid.f90:813-839
computekemaxu                     core/fortran_diag.f90:155-195
computedotprod                    core/fortran_diag.f90:3-30
computesum                        core/fortran_diag.f90:63-90
computesumandnorm                 core/fortran_diag.f90:93-120
domain_integration                core/grid.py:132-140 (NumPy summation order)
================================  =============================================
"""
import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray, ptr

NH = 3


def _shape(a):
    return a.shape[-2], a.shape[-1]


def _cst(cst):
    c = (C.c_double * 5)(*[float(v) for v in np.asarray(cst).ravel()[:5]])
    return c


def adv_upwind(msk, x, y, u, v, cst, nh, method, order, xflx=None, yflx=None, stream=None):
    m, n = _shape(x)
    _lib.call("f2d_adv_upwind_f64", ptr(msk), ptr(x), ptr(y), ptr(u), ptr(v), ptr(xflx), ptr(yflx),
              _cst(cst), nh, method, order, m, n, stream)


def adv_centered(msk, x, y, u, v, cst, nh, method, order, xflx=None, yflx=None, stream=None):
    m, n = _shape(x)
    _lib.call("f2d_adv_centered_f64", ptr(msk), ptr(x), ptr(y), ptr(u), ptr(v), ptr(xflx),
              ptr(yflx), _cst(cst), nh, order, m, n, stream)


def celltocorner(xr, xp, stream=None):
    m, n = _shape(xr)
    _lib.call("f2d_celltocorner_f64", ptr(xr), ptr(xp), m, n, stream)


def computeorthogradient(msk, psi, dx, dy, nh, u, v, stream=None):
    m, n = _shape(psi)
    _lib.call("f2d_orthogradient_f64", ptr(msk), ptr(psi), dx, dy, nh, ptr(u), ptr(v), m, n, stream)


def add_diffusion_fill(msk, trac, dx, nh, Kdiff, dtrac, stream=None):
    m, n = _shape(trac)
    _lib.call("f2d_add_diffusion_f64", ptr(msk), ptr(trac), dx, nh, Kdiff, ptr(dtrac), m, n, stream)


def add_torque_fill(msk, buoy, dx, nh, gravity, domega, stream=None):
    m, n = _shape(buoy)
    _lib.call("f2d_add_torque_f64", ptr(msk), ptr(buoy), dx, nh, gravity, ptr(domega), m, n, stream)


def fillhalo(x, nh, stream=None):
    m, n = _shape(x)
    _lib.call("f2d_fillhalo_f64", ptr(x), nh, m, n, stream)


def smoothtwicewitha_fill(msk, A, x, b, omega, work, stream=None):
    m, n = _shape(x)
    _lib.call("f2d_smooth_f64", ptr(msk), ptr(A), ptr(x), ptr(b), omega, ptr(work), m, n, stream)


def computeresidualwitha_fill(msk, A, x, b, r, stream=None):
    m, n = _shape(x)
    _lib.call("f2d_residual_f64", ptr(msk), ptr(A), ptr(x), ptr(b), ptr(r), m, n, stream)


def restrict_fill(msk2, x1, nh, x2, stream=None):
    m2, n2 = _shape(x2)
    m1, n1 = _shape(x1)
    _lib.call("f2d_restrict_f64", ptr(msk2), ptr(x1), nh, ptr(x2), m2, n2, m1, n1, stream)


def interpolate(msk1, msk2, x2, nh, x1, accumulate=False, stream=None):
    m2, n2 = _shape(x2)
    m1, n1 = _shape(x1)
    _lib.call("f2d_interpolate_f64", ptr(msk1), ptr(msk2), ptr(x2), nh, ptr(x1), m2, n2, m1, n1,
              1 if accumulate else 0, stream)


def coarsenmatrix(Afine, msk1, msk2, nh, Acoarse, stream=None):
    m1, n1 = _shape(msk1)
    m2, n2 = _shape(msk2)
    _lib.call("f2d_coarsenmatrix_f64", ptr(Afine), ptr(msk1), ptr(msk2), nh, ptr(Acoarse),
              m1, n1, m2, n2, stream)


def _scalar_call(name, nout, *args):
    out = DeviceArray((nout,), np.float64)
    _lib.call(name, *args, out.ptr, None)
    return out.get()


def computenorm(msk, x, nh):
    """Sum of squares over the fluid interior (the reference's computenorm)."""
    m, n = _shape(x)
    return float(_scalar_call("f2d_norm2_f64", 1, ptr(msk), ptr(x), nh, m, n)[0])


def computekemaxu(msk, u, v, nh):
    m, n = _shape(u)
    r = _scalar_call("f2d_kemaxu_f64", 2, ptr(msk), ptr(u), ptr(v), nh, m, n)
    return float(r[0]), float(r[1])


def computedotprod(msk, x, y, nh):
    """Sum of x*y over the interior where msk != 0."""
    m, n = _shape(x)
    return float(_scalar_call("f2d_dotprod_f64", 1, ptr(msk), ptr(x), ptr(y), nh, m, n)[0])


def computesum(msk, x, nh):
    """Sum of x over the interior where msk != 0."""
    m, n = _shape(x)
    return float(_scalar_call("f2d_sum_f64", 1, ptr(msk), ptr(x), nh, m, n)[0])


def computesumandnorm(msk, x, nh):
    """(sum of x, sum of x^2) over the interior where msk == 1."""
    m, n = _shape(x)
    r = _scalar_call("f2d_sumandnorm_f64", 2, ptr(msk), ptr(x), nh, m, n)
    return float(r[0]), float(r[1])


def domain_integration(x, nh=NH):
    m, n = _shape(x)
    return float(_scalar_call("f2d_domain_sum_f64", 1, ptr(x), nh, m, n)[0])
