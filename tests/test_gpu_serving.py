"""Continuous-batching serving loop: every request's result is the same
alone and in the mix (bit for bit), every request runs exactly its own step
count, and the admission never exceeds the memory budget."""
import numpy as np
import pytest

from fluid2d_amd import serving
from fluid2d_amd.batch import Batch

pytestmark = pytest.mark.gpu


def run_alone(req, precision):
    p = req.param()
    from fluid2d_amd.grid import Grid
    g = Grid(p)
    b = Batch([(p, g)], precision=precision)
    for k, v in serving.initial_fields(req, g).items():
        b.set_field(0, k, v)
    b.set_psi_from_vorticity()
    b.set_dt(serving.time_step(req, g, b.diagnostics(0)["maxspeed"]))
    b.step(req.steps)
    return b.state(0)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_serving_alone_equals_in_mix(precision):
    reqs = serving.stream_requests(10, rate=200.0, seed=7, sizes_log2=(6, 8), steps=(5, 40))
    # a budget that holds about three requests at a time forces several cohorts
    budget = 3 * max(r.est_bytes(precision) for r in reqs)
    srv = serving.Server(hbm_budget=budget, precision=precision, chunk=7, keep_state=True)
    summary, res = srv.run(reqs)
    assert summary["requests"] == len(reqs)
    assert summary["peak_resident_bytes"] <= budget
    assert summary["p99_s"] >= summary["p50_s"] > 0
    for r in reqs:
        assert res[r.id]["steps"] == r.steps
    for r in reqs[:5]:
        a = run_alone(r, precision)
        m = res[r.id]["state"]
        assert a.shape == m.shape
        assert (a == m).all(), "request %d differs alone vs in the serving mix" % r.id


def test_step_budget_gate():
    """Instances with different step budgets stop at their own count while
    the batch keeps stepping (device-side gate)."""
    reqs = serving.stream_requests(3, seed=9, sizes_log2=(6, 6), steps=(5, 5))
    c = serving.Cohort(reqs, "fp64")
    budgets = [3, 8, 5]
    for i, n in enumerate(budgets):
        c.batch.set_steps(n, i)
    c.batch.step(10)
    for i, n in enumerate(budgets):
        left, done = c.batch.progress(i)
        assert left == 0 and done == n
    for i, r in enumerate(reqs):
        r.steps = budgets[i]
        assert (run_alone(r, "fp64") == c.batch.state(i)).all()
