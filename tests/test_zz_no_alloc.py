"""The allocation-free NumPy-order sum (f2d_domain_sum_f64_noalloc) against
the allocating entry point (f2d_domain_sum_f64): the same kernels, so the
same bits, on two calls in a row, the second on a smaller field (the shared
scratch keeps the first call's partials, and the second must not read
them).  Both are called through the C ABI, on the null stream."""
import ctypes as C

import numpy as np
import pytest

import fluid2d_amd as F
from fluid2d_amd import _lib
from fluid2d_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
NH = 3


def _sum(name, x, out):
    m, n = x.shape
    rc = getattr(F.load(), name)(C.c_void_p(x.ptr), NH, m, n, C.c_void_p(out.ptr), None)
    assert rc == 0, _lib.last_error()
    assert F.load().f2d_synchronize() == 0
    return out.get().copy()


def test_noalloc_sum_equals_allocating_sum_twice():
    rng = np.random.default_rng(77)
    # interiors of 1100 x 900 (121 chunks of 8192) then 300 x 170 (7 chunks)
    for (ny, nx) in ((1100, 900), (300, 170)):
        print("case noalloc sum", ny, nx, flush=True)
        a = np.zeros((ny + 2 * NH, nx + 2 * NH))
        a[NH:-NH, NH:-NH] = rng.standard_normal((ny, nx))
        x = DeviceArray.from_host(a)
        o1 = DeviceArray.from_host(np.full(1, np.nan))
        o2 = DeviceArray.from_host(np.full(1, np.nan))
        r_alloc = _sum("f2d_domain_sum_f64", x, o1)
        r_free = _sum("f2d_domain_sum_f64_noalloc", x, o2)
        assert r_alloc[0] == r_free[0], (ny, nx, r_alloc[0], r_free[0])
        # NumPy's own sum over the same interior (the order both restate)
        assert r_free[0] == np.sum(a[NH:-NH, NH:-NH]), (ny, nx)


def test_noalloc_sum_rejects_oversize_before_device_work():
    # 8192 * 8192 interior cells is the largest accepted; one row more is not
    rc = F.load().f2d_domain_sum_f64_noalloc(None, NH, 8193 + 2 * NH, 8192 + 2 * NH, None, None)
    assert rc == _lib.F2D_E_ARG
    assert "at most 8192" in _lib.last_error()
