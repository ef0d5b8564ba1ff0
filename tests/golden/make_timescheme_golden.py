"""Golden vectors of the reference's own RK3_SSP and RK3 stage algebra.

Runs in the build container only, where /root/reference exists.  It imports
the reference's ``core/timescheme.py`` (which needs only NumPy and the
reference's ``core/param.py`` / ``core/variables.py``) and drives
``Timescheme.forward`` with a right-hand side dx/dt = lam * x (elementwise),
recording for each scheme:

* the state after each of three steps (NumPy's own operation order of
  ``core/timescheme.py:152-180``, bit for bit);
* the stage index ``kstage`` and time ``t`` of every right-hand-side call;
* ``kforcing`` and ``dtcoef`` (``core/timescheme.py:30-33, 61-75``), which
  decide at which stage Euler skips its inversion and Boussinesq adds
  diffusion.

``Param()`` itself imports the MPI grid module (absent here), so the
parameter object is made with ``Param.set_parameters`` from the reference's
own ``core/defaults.json``; ``Timescheme`` reads only ``timestepping`` from
it.  Output: ``timescheme_rk3.npz`` next to this script.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/core"


def main():
    if not os.path.isdir(REF):
        sys.exit("%s not present: the fixture is made in the build container only" % REF)
    sys.path.insert(0, REF)
    from param import Param
    from timescheme import Timescheme

    with open(os.path.join(REF, "defaults.json")) as f:
        namelist = json.load(f)
    out = {}
    rng = np.random.default_rng(2024)
    shape = (2, 16, 24)
    lam = rng.uniform(-4.0, 1.5, shape)
    x0 = rng.standard_normal(shape)
    dt, t0 = 0.137, 0.25
    out["lam"], out["x0"], out["dt"], out["t0"] = lam, x0, np.array(dt), np.array(t0)
    for scheme in ("RK3_SSP", "RK3"):
        p = Param.__new__(Param)
        p.set_parameters(namelist)
        p.timestepping = scheme
        x = x0.copy()
        ts = Timescheme(p, x)
        calls = []

        def rhs(xs, t, dxdt, ts=ts, calls=calls):
            calls.append((ts.kstage, t))
            dxdt[...] = lam * xs

        ts.set(rhs, scheme)
        t = t0
        for step in range(3):
            ts.forward(x, t, dt)
            t += dt
            out["%s_x%d" % (scheme, step + 1)] = x.copy()
        out["%s_calls" % scheme] = np.array(calls, dtype=np.float64)
        out["%s_kforcing" % scheme] = np.array(ts.kforcing)
        out["%s_dtcoef" % scheme] = np.array(ts.dtcoef)
    path = os.path.join(HERE, "timescheme_rk3.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
