"""Served requests equal the same requests solved alone (plain calls)."""
import numpy as np
import pytest

from xgrid_amd import workloads as W

pytestmark = pytest.mark.gpu


def _trace(seed, count=7):
    """A long first request, then short ones arriving while it runs."""
    rng = np.random.default_rng(seed)
    out = [W.InstanceSpec("cavity", 120, 0.05, W.stable_dt("cavity", 120, 0.05), 60, 0.0, 0)]
    t = 0.0
    for i in range(1, count):
        t += 0.001
        n = int(rng.integers(36, 120))
        nu = float(rng.uniform(0.001, 0.1))
        out.append(W.InstanceSpec("cavity", n, nu, W.stable_dt("cavity", n, nu), int(rng.integers(2, 14)), t, i))
    return out


@pytest.mark.parametrize("precision,budget_requests", [("float", 100), ("float", 2), ("double", 3)])
def test_served_results_equal_alone(precision, budget_requests):
    from xgrid_amd.serving import Server, footprint, serve_plain
    trace = _trace(11)
    elem = 4 if precision == "float" else 8
    budget = budget_requests * max(footprint(s, elem) for s in trace)
    srv = Server(precision, hbm_budget=budget, steps_per_tick=3)
    got, summ = srv.run(trace)
    ref, _ = serve_plain(trace, precision, solver=srv.ms)
    assert sorted(got) == sorted(ref) == list(range(len(trace)))
    for rid in got:
        for c in "bpuv":
            a, b = got[rid].fields[c], ref[rid].fields[c]
            assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), (rid, c)
    assert summ["latency_p99_s"] >= summ["latency_p50_s"] > 0
    assert srv.peak_used <= budget
    if budget_requests > 3:
        assert srv.peak_active >= 2                     # arrivals joined while the first request ran
        assert got[0].done > max(got[r].admitted for r in got if r)
    srv.close()
