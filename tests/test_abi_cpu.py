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
