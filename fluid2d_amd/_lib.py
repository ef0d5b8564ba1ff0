"""ctypes binding of libf2dgpu.so (C ABI declared in include/fluid2d_amd.h).

The library is built in-tree (``fluid2d_amd/libf2dgpu.so``) by
``__graft_entry__.build()`` / ``make -C fluid2d_amd/csrc``.  There is no
fallback: if the library is missing or no HIP device is visible, the calls
that need it raise.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libf2dgpu.so")

F2D_OK, F2D_E_ARG, F2D_E_HIP, F2D_E_NOMEM, F2D_E_DIVERGED, F2D_E_UNSUPPORTED = range(6)


class F2DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("f2d error %d: %s" % (code, msg))
        self.code = code


class DivergenceError(F2DError):
    """The multigrid solve stopped converging (the reference exits the process)."""


class f2d_model_params(C.Structure):
    _fields_ = [("model", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double),
                ("order", C.c_int), ("flux_splitting", C.c_int), ("aparab", C.c_double),
                ("timestepping", C.c_int), ("perio", C.c_int), ("diffusion", C.c_int),
                ("kdiff_vort", C.c_double), ("kdiff_buoy", C.c_double),
                ("gravity", C.c_double), ("hydroepsilon", C.c_double)]


_p = C.c_void_p
_i = C.c_int
_d = C.c_double
_z = C.c_size_t
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_pp = C.POINTER(C.c_void_p)

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "f2d_last_error": [],
    "f2d_version": [],
    "f2d_device_count": [_ip],
    "f2d_malloc": [_pp, _z],
    "f2d_free": [_p],
    "f2d_memcpy_h2d": [_p, _p, _z],
    "f2d_memcpy_d2h": [_p, _p, _z],
    "f2d_memcpy_d2d": [_p, _p, _z],
    "f2d_synchronize": [],
    "f2d_adv_upwind_f64": [_p, _p, _p, _p, _p, _p, _p, _dp, _i, _i, _i, _i, _i, _p],
    "f2d_adv_centered_f64": [_p, _p, _p, _p, _p, _p, _p, _dp, _i, _i, _i, _i, _p],
    "f2d_celltocorner_f64": [_p, _p, _i, _i, _p],
    "f2d_orthogradient_f64": [_p, _p, _d, _d, _i, _p, _p, _i, _i, _p],
    "f2d_add_diffusion_f64": [_p, _p, _d, _i, _d, _p, _i, _i, _p],
    "f2d_add_torque_f64": [_p, _p, _d, _i, _d, _p, _i, _i, _p],
    "f2d_fillhalo_f64": [_p, _i, _i, _i, _p],
    "f2d_smooth_f64": [_p, _p, _p, _p, _d, _p, _i, _i, _p],
    "f2d_residual_f64": [_p, _p, _p, _p, _p, _i, _i, _p],
    "f2d_restrict_f64": [_p, _p, _i, _p, _i, _i, _i, _i, _p],
    "f2d_interpolate_f64": [_p, _p, _p, _i, _p, _i, _i, _i, _i, _i, _p],
    "f2d_coarsenmatrix_f64": [_p, _p, _p, _i, _p, _i, _i, _i, _i, _p],
    "f2d_norm2_f64": [_p, _p, _i, _i, _i, _p, _p],
    "f2d_kemaxu_f64": [_p, _p, _p, _i, _i, _i, _p, _p],
    "f2d_dotprod_f64": [_p, _p, _p, _i, _i, _i, _p, _p],
    "f2d_sum_f64": [_p, _p, _i, _i, _i, _p, _p],
    "f2d_sumandnorm_f64": [_p, _p, _i, _i, _i, _p, _p],
    "f2d_domain_sum_f64": [_p, _i, _i, _i, _p, _p],
    "f2d_domain_sum_f64_noalloc": [_p, _i, _i, _i, _p, _p],
    "f2d_gmg_create": [_i, _i, _i, _d, _d, _d, _d, _p, _pp],
    "f2d_gmg_destroy": [_p],
    "f2d_gmg_set_helmholtz": [_p, _d],
    "f2d_gmg_nlevels": [_p, _ip],
    "f2d_gmg_level": [_p, _i, _i, _pp, _ip, _ip],
    "f2d_gmg_vcycle": [_p, _i],
    "f2d_gmg_fcycle": [_p, _i],
    "f2d_gmg_twovcycle": [_p, _p, _p],
    "f2d_gmg_solve": [_p, _p, _p, _i, _d, _ip, _dp],
    "f2d_model_create": [C.POINTER(f2d_model_params), _p, _pp],
    "f2d_model_destroy": [_p],
    "f2d_model_state": [_p, _pp, _ip, _ip, _ip],
    "f2d_model_tendency": [_p, _i, _pp],
    "f2d_model_corner_mask": [_p, _pp],
    "f2d_model_gmg": [_p, _pp],
    "f2d_model_set_psi_from_vorticity": [_p],
    "f2d_model_invert_vorticity": [_p, _p, _i],
    "f2d_model_rhs_adv": [_p, _p, _p],
    "f2d_model_set_umax": [_p, _d],
    "f2d_model_step": [_p, _d],
    "f2d_model_diagnostics": [_p, _p, _p, _dp],
    "f2d_model_run": [_p, _i, _d, _p],
    "f2d_model_diag_fluxes": [_p, _p, _d, _p],
    "f2d_model_set_noslip": [_p, _p, _p, _p, _p, _d, _d, _d, _i],
    "f2d_model_set_islands": [_p, _p, _p],
    "f2d_model_last_solve": [_p, _ip, _dp],
}
_RESTYPE = {"f2d_last_error": C.c_char_p}

_lib = None


def load():
    """Load libf2dgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libf2dgpu.so is not built: run `python -c 'import __graft_entry__ as g; g.build()'`"
                           " or `make -C fluid2d_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def last_error():
    return load().f2d_last_error().decode()


def check(rc):
    if rc != F2D_OK:
        msg = last_error()
        if rc == F2D_E_DIVERGED:
            raise DivergenceError(rc, msg)
        raise F2DError(rc, msg)
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))


def device_count():
    n = C.c_int(0)
    rc = load().f2d_device_count(C.byref(n))
    return n.value if rc == F2D_OK else 0
