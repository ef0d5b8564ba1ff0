"""CPU checks of the C ABI: the library loads, exports every function that
include/fluid2d_amd.h declares, the ctypes table matches the header, and
argument validation fails loudly (no device needed for those paths)."""
import os
import re

import pytest

import fluid2d_amd
from fluid2d_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "fluid2d_amd.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(f2d_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_parses():
    names = header_functions()
    assert "f2d_adv_upwind_f64" in names and "f2d_model_step" in names
    assert len(names) >= 40


def test_library_exports_every_header_symbol():
    lib = fluid2d_amd.load()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_covers_header():
    from fluid2d_amd import batch
    bound = set(_lib.SIGNATURES) | set(batch._SIG)
    assert sorted(bound) == header_functions()


def test_version_and_errors():
    lib = fluid2d_amd.load()
    assert lib.f2d_version() >= 100
    # nh != 3 is rejected before any device work (the reference aborts)
    rc = lib.f2d_fillhalo_f64(None, 2, 10, 10, None)
    assert rc == _lib.F2D_E_ARG
    assert "nh" in _lib.last_error()
    rc = lib.f2d_adv_upwind_f64(None, None, None, None, None, None, None, None, 3, 0, 4, 12, 12,
                                None)
    assert rc == _lib.F2D_E_ARG
    with pytest.raises(_lib.F2DError):
        _lib.call("f2d_restrict_f64", None, None, 3, None, 10, 10, 15, 15, None)


def test_param_validation():
    p = fluid2d_amd.Param(modelname="euler", geometry="perio", timestepping="AB2")
    with pytest.raises(ValueError):
        p.checkall()
    p = fluid2d_amd.Param(modelname="euler", geometry="perio", nh=2)
    with pytest.raises(ValueError):
        p.checkall()
    fluid2d_amd.Param(modelname="boussinesq", geometry="xchannel").checkall()


def test_grid_mask_matches_oracle_grid():
    from oracle import oracle as orc
    for geom in ("perio", "closed", "xchannel", "disc"):
        p = fluid2d_amd.Param(nx=32, ny=16, Lx=2.0, Ly=1.0, geometry=geom)
        g = fluid2d_amd.Grid(p)
        o = orc.make_grid(32, 16, 2.0, 1.0, geom)
        assert (g.msk == o["msk"]).all()
        assert g.area == o["area"]
        assert (g.xr == o["xr"]).all() and (g.yr == o["yr"]).all()


def _params(nx, ny, model=0):
    from fluid2d_amd import batch
    p = _lib.f2d_model_params()
    p.model, p.nx, p.ny, p.Lx, p.Ly = model, nx, ny, 1.0, 1.0
    p.order, p.flux_splitting, p.aparab = 5, 1, 0.05
    p.hydroepsilon = 1.0
    assert batch  # the engine's ctypes table is bound on first use
    return p


@pytest.mark.parametrize("nx,ny,model,prec,ok", [
    # FP64 level-0 operator, 5 planes of (ny+6)(nx+6) doubles: 2^32 bytes at
    # about 10,360 cells a side; the state (5 or 6 planes) comes just below
    (16384, 16384, 0, 0, False),
    (10368, 10368, 0, 1, False),
    (16384, 8, 1, 0, True),       # long and thin: small arrays, passes the guard
    (9472, 9472, 1, 0, False),    # Boussinesq state: 6 planes of FP64
])
def test_batch_rejects_arrays_beyond_32bit_offsets(nx, ny, model, prec, ok):
    """Kernels address one instance's arrays through 32-bit byte offsets;
    f2d_batch_create rejects any instance whose state, tendency or operator
    array reaches 2^32 bytes, before any device call (no GPU here)."""
    import ctypes as C
    from fluid2d_amd import batch
    lib = batch._L()
    F = (ny + 6) * (nx + 6)
    largest = max((5 if model == 0 else 6) * F * 8, 5 * F * 8)
    assert (largest < 2 ** 32) == ok
    if ok:
        return  # an accepted batch would go on to allocate on the device
    arr = (_lib.f2d_model_params * 1)(_params(nx, ny, model))
    msks = (C.c_void_p * 1)(None)  # never read: the guard runs first
    h = C.c_void_p()
    rc = lib.f2d_batch_create(1, arr, msks, prec, None, C.byref(h))
    assert rc == _lib.F2D_E_ARG
    msg = lib.f2d_batch_last_error().decode()
    assert "2^32" in msg and "instance 0" in msg
    assert h.value is None


def test_noalloc_sum_rejects_oversize_without_device_work():
    """The allocation-free NumPy-order sum accepts interiors of at most
    8192 * 8192 cells (its chunk partials live in a fixed 64 KiB device
    array); a larger one is refused before any device call."""
    lib = fluid2d_amd.load()
    rc = lib.f2d_domain_sum_f64_noalloc(None, 3, 8193 + 6, 8192 + 6, None, None)
    assert rc == _lib.F2D_E_ARG
    assert "at most 8192" in _lib.last_error()
