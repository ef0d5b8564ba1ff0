"""The oracle's no-slip source term against the reference Fortran
(computenoslipsourceterm, core/fortran_operators.f90:221-268, built with
amdflang by oracle/build_ref.sh; vectors in tests/golden/noslip_24x16.npz,
written by tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden", "noslip_24x16.npz")


@pytest.mark.parametrize("mname", ["perio", "closed", "islands"])
def test_noslip_source_vs_fortran(mname):
    g = np.load(GOLD)
    p = mname + "_"
    msk, x, y0 = g[p + "msk"], g[p + "x"], g[p + "y0"]
    dx, dy = g[p + "dxdy"]
    y = y0.copy()
    tot = orc.noslip_source(msk, x, y, float(dx), float(dy), 3)
    assert (y == g[p + "noslip"]).all()
    assert tot == float(g[p + "total"])
    if mname != "perio":  # walls: a nonzero source
        assert np.abs(y[3:-3, 3:-3]).max() > 0
