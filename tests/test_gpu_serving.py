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


def test_served_cavities_at_configured_sizes_equal_oracle_and_alone():
    """Serving at the sizes the serving workload draws from (sides 128..2048,
    FP32 fields): four overlapping cavity requests (128^2 arriving first and
    running longest, 300^2, 700^2 and 2048^2 joining while it runs) are served
    through the continuous-batching server.  Every response equals, bit for bit,
    the same request solved alone with plain calls and the FP32 oracle (the
    restated reference CPU path) run for the request's steps from the same zero
    fields.  All four spacings 2/(n-1) are not powers of two, so the uniform
    divisions divide."""
    from oracle.cpu_path import CpuKernel, HostGrid
    from xgrid_amd.precision import precision_scope
    from xgrid_amd.serving import Server, footprint, serve_plain
    specs = [(128, 0.02, 24, 0.0), (300, 0.05, 10, 0.001), (700, 0.004, 6, 0.002), (2048, 0.1, 3, 0.003)]
    trace = [W.InstanceSpec("cavity", n, nu, W.stable_dt("cavity", n, nu), steps, t, i)
             for i, (n, nu, steps, t) in enumerate(specs)]
    budget = sum(footprint(s, 4) for s in trace)
    srv = Server("float", hbm_budget=budget, steps_per_tick=3)
    got, summ = srv.run(trace)
    alone, _ = serve_plain(trace, "float", solver=srv.ms)
    assert srv.peak_active >= 2
    with precision_scope("float"):
        ck = CpuKernel(W.make_cavity())
        for s in trace:
            print(f"serving oracle: request {s.rid} side {s.n} steps {s.steps}", flush=True)
            grids, scalars = W.instance_grids(s, HostGrid)
            for _ in range(s.steps):
                ck(*grids, *scalars)
            for c, g in zip("bpuv", grids):
                a, b, h = got[s.rid].fields[c], alone[s.rid].fields[c], g.now
                assert a.dtype == b.dtype == h.dtype == np.float32
                assert np.array_equal(a.view(np.uint32), h.view(np.uint32)), (s.rid, c, "served vs oracle")
                assert np.array_equal(b.view(np.uint32), h.view(np.uint32)), (s.rid, c, "alone vs oracle")
            assert np.abs(grids[1].now).max() > 0                 # live pressure
    assert summ["requests"] == len(trace) and summ["latency_p99_s"] >= summ["latency_p50_s"] > 0
    srv.close()
