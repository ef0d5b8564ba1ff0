"""Continuous-batching serving loop: every request's result is the same
alone and in the mix (bit for bit), also when it runs in the slot of an
earlier request inside a running pool; every request runs exactly its own
step count, and the admission never exceeds the memory budget."""
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
    from fluid2d_amd.grid import Grid
    pg = [(r.param(), None) for r in reqs]
    pg = [(p, Grid(p)) for p, _ in pg]
    b = Batch(pg, precision="fp64")
    for i, r in enumerate(reqs):
        for k, v in serving.initial_fields(r, pg[i][1]).items():
            b.set_field(i, k, v)
    b.set_psi_from_vorticity()
    for i, r in enumerate(reqs):
        b.set_dt(serving.time_step(r, pg[i][1], b.diagnostics(i)["maxspeed"]), i)
    budgets = [3, 8, 5]
    for i, n in enumerate(budgets):
        b.set_steps(n, i)
    b.step(10)
    for i, n in enumerate(budgets):
        left, done = b.progress(i)
        assert left == 0 and done == n
    for i, r in enumerate(reqs):
        r.steps = budgets[i]
        assert (run_alone(r, "fp64") == b.state(i)).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_slot_reuse_alone_equals_in_mix(precision):
    """A budget of four slots for eight requests of two classes: later
    requests take the slots of finished ones in running pools (reset,
    initialised and stepped between chunks while the other slots keep
    running).  Every request equals the same request run alone, bit for
    bit, and the resident estimate never exceeds the budget."""
    reqs = []
    for i in range(8):
        model = "euler" if i % 2 == 0 else "boussinesq"
        reqs.append(serving.Request(i, 0.001 * i, model, 64, 6 + 5 * (i % 3), seed=300 + i))
    est = {m: serving.Request(-1, 0, m, 64, 1, 0).est_bytes(precision)
           for m in ("euler", "boussinesq")}
    budget = 2 * est["euler"] + 2 * est["boussinesq"]
    srv = serving.Server(hbm_budget=budget, precision=precision, chunk=4, keep_state=True)
    summary, res = srv.run(reqs)
    print("case slot reuse", precision, summary["slot_reuses"], flush=True)
    assert summary["requests"] == len(reqs)
    assert summary["peak_resident_bytes"] <= budget
    assert summary["slot_reuses"] > 0
    assert any(res[r.id]["slot_reused"] for r in reqs)
    assert summary["queue_p99_s"] >= summary["queue_p50_s"] >= 0
    for r in reqs:
        assert res[r.id]["steps"] == r.steps
        a = run_alone(r, precision)
        assert (a == res[r.id]["state"]).all(), \
            "request %d (slot re-used: %s) differs alone vs in the mix" % (r.id, res[r.id]["slot_reused"])
