"""Serving loop logic on the CPU (a stand-in solver counts steps): admission by
HBM budget in arrival order, exact step counts, latency bookkeeping."""
import pytest

from xgrid_amd import workloads as W
from xgrid_amd.serving import Server, footprint


class FakeSolver:
    """Counts the steps each admitted request receives."""

    def __init__(self):
        self.members = {}
        self.steps = {}
        self.order = []
        self._next = 1

    def add(self, spec):
        h = self._next
        self._next += 1
        self.members[h] = spec
        self.steps[spec.rid] = 0
        self.order.append(spec.rid)
        return h

    def step(self, k):
        for s in self.members.values():
            self.steps[s.rid] += k

    def remove(self, h):
        self.members.pop(h)
        return []

    def close(self):
        pass


class FakeTimer:
    def begin(self):
        pass

    def end(self):
        pass

    def sync(self):
        pass

    def busy_seconds(self):
        return 0.0


def _trace(n=12):
    out = []
    for i in range(n):
        out.append(W.InstanceSpec("cavity", 64 + 16 * (i % 5), 0.05, 1e-5, steps=3 + 5 * (i % 4),
                                  arrival=0.0005 * i, rid=i))
    return out


def test_every_request_gets_exactly_its_steps_in_fifo_order():
    trace = _trace()
    fake = FakeSolver()
    srv = Server("float", hbm_budget=1 << 30, steps_per_tick=4, solver=fake, timer=FakeTimer(), keep_fields=False)
    results, summ = srv.run(trace)
    assert sorted(results) == [s.rid for s in trace]
    for s in trace:
        assert fake.steps[s.rid] == s.steps
        r = results[s.rid]
        assert r.done >= r.admitted >= r.arrival - 1e-9
    assert fake.order == sorted(fake.order)            # admitted in arrival order
    assert summ["requests"] == len(trace) and summ["latency_p99_s"] >= summ["latency_p50_s"] >= 0


def test_budget_bounds_the_resident_set():
    trace = _trace(10)
    one = max(footprint(s, 4) for s in trace)
    fake = FakeSolver()
    srv = Server("float", hbm_budget=2 * one, steps_per_tick=2, solver=fake, timer=FakeTimer(), keep_fields=False)
    results, _ = srv.run(trace)
    assert len(results) == 10
    assert srv.peak_used <= 2 * one
    assert srv.peak_active <= 2


def test_cell_cap_bounds_the_resident_work():
    """With a resident-cell cap, admitted requests never hold more cells than the
    cap (one request alone may exceed it), requests are still admitted in arrival
    order, and every one gets exactly its steps."""
    from xgrid_amd.serving import cells
    trace = _trace(12)
    cap = 2 * max(cells(s) for s in trace)
    fake = FakeSolver()
    srv = Server("float", hbm_budget=1 << 30, steps_per_tick=2, solver=fake, timer=FakeTimer(), keep_fields=False,
                 max_cells=cap)
    resident = []
    orig_add, orig_remove = fake.add, fake.remove
    live = {}

    def add(spec):
        h = orig_add(spec)
        live[h] = cells(spec)
        resident.append(sum(live.values()))
        return h

    def remove(h):
        live.pop(h)
        return orig_remove(h)

    fake.add, fake.remove = add, remove
    results, _ = srv.run(trace)
    assert len(results) == 12 and max(resident) <= cap
    assert fake.order == sorted(fake.order)
    for s in trace:
        assert fake.steps[s.rid] == s.steps
    big = [W.InstanceSpec("cavity", 300, 0.05, 1e-5, steps=2, arrival=0.0, rid=0)]
    results, _ = Server("float", hbm_budget=1 << 30, solver=FakeSolver(), timer=FakeTimer(), keep_fields=False,
                        max_cells=10).run(big)
    assert len(results) == 1


def test_request_larger_than_budget_is_refused():
    trace = [W.InstanceSpec("cavity", 2048, 0.01, 1e-6, steps=5, arrival=0.0, rid=0)]
    srv = Server("float", hbm_budget=1 << 20, solver=FakeSolver(), timer=FakeTimer(), keep_fields=False)
    with pytest.raises(ValueError):
        srv.run(trace)


def test_footprint_counts_levels_scratch_and_mask():
    s = W.InstanceSpec("cavity", 100, 0.1, 1e-5)
    assert footprint(s, 4) == 4 * 100 * 100 * (3 * 4 + 4)
    d = W.InstanceSpec("diffusion", 100, 0.1, 1e-5)
    assert footprint(d, 8) == 100 * 100 * (3 * 8 + 4)


def test_serving_trace_matches_configured_distribution():
    t = W.serving_specs(400, seed=2)
    arr = [s.arrival for s in t]
    assert arr == sorted(arr) and 1.5 < arr[-1] < 2.5          # about 400 / 200 s
    assert all(128 <= s.n <= 2048 and 100 <= s.steps <= 2000 and 0.001 <= s.nu <= 0.1 for s in t)
    assert all(s.kind == "cavity" for s in t)
