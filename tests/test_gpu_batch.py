"""Batched execution: every instance of a batch ends bit-exact with the same
instance run alone through plain kernel calls (and with the oracle)."""
import numpy as np
import pytest

from oracle import restate as R

pytestmark = pytest.mark.gpu


def _cavity(xg, n, dt=2e-6, nu=0.1):
    from xgrid_amd import workloads as W
    masks = W.cavity_masks(n)
    grids = []
    for c in "bpuv":
        g = xg.Grid((n, n), float)
        g.boundary[...] = masks[c]
        grids.append(g)
    return grids, W.cavity_config(n, dt=dt, nu=nu)


def _assert_same(a, b):
    for x, y in zip(a, b):
        ax, by = x.now, y.now
        assert ax.dtype == by.dtype and ax.shape == by.shape
        if not np.array_equal(ax.view(np.uint8), by.view(np.uint8)):
            d = np.argwhere(ax != by)
            raise AssertionError(f"{len(d)} cells differ, first {d[:3].tolist()}")


def _run_cavities(xg, sizes, steps, graphs=True, max_blocks=0, nus=None):
    from xgrid_amd import workloads as W
    from xgrid_amd.batch import Batch
    kb, kp = W.make_cavity(), W.make_cavity()
    batch = Batch(kb, graphs=graphs, max_blocks=max_blocks)
    members = []
    for i, n in enumerate(sizes):
        nu = 0.1 if nus is None else nus[i]
        gb, cfg = _cavity(xg, n, nu=nu)
        ga, _ = _cavity(xg, n, nu=nu)
        batch.add(*gb, cfg)
        members.append((gb, ga, cfg))
    batch.step(steps)
    for gb, ga, cfg in members:
        for _ in range(steps):
            kp(*ga, cfg)
    return batch, members


@pytest.mark.parametrize("graphs,max_blocks", [(True, 0), (False, 0), (True, 64)])
def test_batched_cavities_equal_alone(double, graphs, max_blocks):
    sizes = [33, 64, 97, 130, 41, 200]
    nus = [0.1, 0.05, 0.1, 0.02, 0.08, 0.1]
    steps = 14
    batch, members = _run_cavities(double, sizes, steps, graphs, max_blocks, nus)
    for gb, ga, cfg in members:
        _assert_same(gb, ga)
    # every member also equals the numpy restatement of the reference's cavity
    for (gb, ga, cfg), n in zip(members, sizes):
        ref = R.cavity(n, steps, cfg.rho, cfg.nu, cfg.dt, cfg.dx, cfg.dy)
        for g, r in zip(gb, ref):
            assert np.array_equal(g.now, r), (n, "batched member vs restatement")
    c = batch.counters()
    assert c["steps"] == steps and c["validations"] >= 1
    if graphs:
        assert c["graph_captures"] >= 1 and c["graph_replays"] >= 1
    else:
        assert c["graph_replays"] == 0
    batch.close()


def test_batched_cavities_fp32_equal_alone(single):
    from oracle.cpu_path import CpuKernel, HostGrid
    from xgrid_amd import workloads as W
    sizes, steps = [48, 80, 33], 9
    batch, members = _run_cavities(single, sizes, steps)
    ck = CpuKernel(W.make_cavity())
    for (gb, ga, cfg), n in zip(members, sizes):
        assert gb[1].now.dtype == np.float32
        _assert_same(gb, ga)
        # and the FP32 oracle (the restated reference CPU path in FP32)
        masks = W.cavity_masks(n)
        hs = []
        for c in "bpuv":
            h = HostGrid((n, n), float)
            h.boundary[...] = masks[c]
            hs.append(h)
        for _ in range(steps):
            ck(*hs, cfg)
        for g, h in zip(gb, hs):
            assert np.array_equal(g.now.view(np.uint32), h.now.view(np.uint32)), (n, "batched member vs oracle")
    batch.close()


def test_batched_diffusion_2d_equal_alone(double):
    from xgrid_amd import workloads as W
    from xgrid_amd.batch import Batch
    kb, kp = W.make_diffusion_2d(), W.make_diffusion_2d()
    batch = Batch(kb)
    rng = np.random.default_rng(4)
    pairs = []
    for n in (64, 100, 257, 31):
        u0 = rng.random((n, n))
        nu = float(rng.uniform(0.01, 0.1))
        dx = 2.0 / (n - 1)
        dt = 0.2 * dx * dx / nu
        gs = []
        for _ in range(2):
            g = double.Grid((n, n), float)
            g.now[...] = u0
            g.boundary[...] = W.box_walls((n, n))
            gs.append(g)
        batch.add(gs[0], nu, dt, dx)
        pairs.append((gs, (nu, dt, dx)))
    batch.step(23)
    for gs, a in pairs:
        for _ in range(23):
            kp(gs[1], *a)
        _assert_same([gs[0]], [gs[1]])
    batch.close()


@pytest.mark.parametrize("prec,max_blocks", [("double", 0), ("float", 0), ("double", 16), ("float", 16)])
def test_batched_heat_3d_equal_alone(prec, max_blocks, request):
    """Batched 3-D statements (staged loads: 128-column tiles in FP64, 256 in
    FP32) on non-cubic grids, one narrower than a 64-column tile and one with two
    full tiles of either width, equal the same instances run alone and the numpy
    restatement of the reference's C.  With a block cap the launch is persistent:
    a workgroup stages tile after tile through the same LDS."""
    xg = request.getfixturevalue("double" if prec == "double" else "single")
    from xgrid_amd import workloads as W
    from xgrid_amd.batch import Batch
    kb, kp = W.make_heat_3d(), W.make_heat_3d()
    batch = Batch(kb, max_blocks=max_blocks)
    rng = np.random.default_rng(5)
    dt_np = np.float64 if prec == "double" else np.float32
    pairs = []
    for shape in ((34, 34, 34), (20, 45, 150), (66, 9, 130), (20, 30, 40), (7, 9, 600)):
        u0 = rng.random(shape).astype(dt_np)
        nu, dx = 0.1, 1.0 / (max(shape) - 1)
        dt = 0.1 * dx * dx / nu
        gs = []
        for _ in range(2):
            g = xg.Grid(shape, float)
            g.now[...] = u0
            g.boundary[...] = W.box_walls(shape)
            gs.append(g)
        batch.add(gs[0], nu, dt, dx)
        pairs.append((gs, (nu, dt, dx), u0))
    batch.step(7)
    for gs, a, u0 in pairs:
        for _ in range(7):
            kp(gs[1], *a)
        _assert_same([gs[0]], [gs[1]])
        assert np.array_equal(gs[0].now, R.heat_3d(u0, *a, 7, dt_np))
    batch.close()


def test_batch_partitions_instances_by_control_flow(double):
    """Instances whose calls take different paths (here: loop trip counts) are
    walked in separate groups and still match their plain runs."""
    from xgrid_amd import workloads as W
    from xgrid_amd.batch import Batch
    kb, kp = W.make_cavity(sweeps=7), W.make_cavity(sweeps=7)
    batch = Batch(kb)
    members = []
    for n, nit in ((40, 3), (64, 12), (33, 3), (90, 0)):
        gb, cfg = _cavity(double, n)
        ga, _ = _cavity(double, n)
        batch.add(*gb, cfg, nit)
        members.append((gb, ga, cfg, nit))
    batch.step(5)
    for gb, ga, cfg, nit in members:
        for _ in range(5):
            kp(*ga, cfg, nit)
        _assert_same(gb, ga)
    batch.close()


def test_batch_membership_changes_between_steps(double):
    from xgrid_amd import workloads as W
    from xgrid_amd.batch import Batch
    kb, kp = W.make_cavity(), W.make_cavity()
    batch = Batch(kb)
    a, cfg_a = _cavity(double, 50)
    b, cfg_b = _cavity(double, 70)
    c, cfg_c = _cavity(double, 36)
    ia = batch.add(*a, cfg_a)
    batch.add(*b, cfg_b)
    batch.step(4)
    batch.remove(ia)
    batch.add(*c, cfg_c)
    batch.step(3)
    assert len(batch) == 2
    for grids, cfg, steps in ((a, cfg_a, 4), (b, cfg_b, 7), (c, cfg_c, 3)):
        alone, _ = _cavity(double, grids[0].shape[0])
        for _ in range(steps):
            kp(*alone, cfg)
        _assert_same(grids, alone)
    with pytest.raises(Exception):
        batch.add(*b, cfg_b)                 # a grid may belong to one instance only
    batch.close()


def test_batch_mixed_with_plain_calls_on_other_grids(double):
    """A batch and plain calls of the same operator interleave without sharing state."""
    from xgrid_amd import workloads as W
    from xgrid_amd.batch import Batch
    k = W.make_cavity()
    batch = Batch(k)
    gb, cfg = _cavity(double, 45)
    gp, _ = _cavity(double, 45)
    batch.add(*gb, cfg)
    for _ in range(3):
        batch.step(1)
        k(*gp, cfg)
    _assert_same(gb, gp)
    batch.close()


@pytest.mark.parametrize("precision", ["double", "float"])
def test_multisolver_mixed_kinds_equal_plain(precision):
    from xgrid_amd import workloads as W
    from xgrid_amd.multi import MultiSolver
    specs = W.batch_specs(8, seed=5, lo=40, hi=160)
    assert {s.kind for s in specs} == {"cavity", "diffusion"}
    ms = MultiSolver(precision)
    handles = [ms.add(s) for s in specs]
    ms.step(6)
    for h, s in zip(handles, specs):
        grids, scalars = ms.grids_for(s)
        for _ in range(6):
            ms.plain_call(s, grids, scalars)
        _assert_same(ms.grids(h), grids)
    ms.close()


def test_autotuned_batch_equals_plain(double):
    """The autotuner times every candidate shape on copies of the instances
    (leaving them untouched) and the tuned batch still equals plain calls."""
    from xgrid_amd import workloads as W
    from xgrid_amd.multi import MultiSolver
    specs = W.batch_specs(10, seed=9, lo=48, hi=400)
    ms = MultiSolver("double", size_classes=(192,))
    tuning = ms.autotune(specs, limit=6, use_cache=False)
    assert tuning and all(k.startswith("cavity/") for k in tuning)  # diffusion has no blocked loop
    assert len({ms.key(s) for s in specs if s.kind == "cavity"}) >= 2   # sizes span two classes
    for t in tuning.values():
        assert all(v is not None and v > 0 for v in t["timings"].values())
    handles = [ms.add(s) for s in specs]
    for key, t in tuning.items():
        assert ms.shapes[key] == t["shape"]
    ms.step(5)
    for h, s in zip(handles, specs):
        grids, scalars = ms.grids_for(s)
        for _ in range(5):
            ms.plain_call(s, grids, scalars)
        _assert_same(ms.grids(h), grids)
    ms.close()


def test_batched_register_layout_equals_plain(double):
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import HipKernel
    from xgrid_amd.batch import Batch
    kb, kp = W.make_cavity(), W.make_cavity()
    batch = Batch(kb, native=HipKernel(kb, (40, 64, 512, 6, "reg")))
    members = []
    for n in (65, 130, 41, 200):
        gb, cfg = _cavity(double, n)
        ga, _ = _cavity(double, n)
        batch.add(*gb, cfg)
        members.append((gb, ga, cfg))
    batch.step(9)
    for gb, ga, cfg in members:
        for _ in range(9):
            kp(*ga, cfg)
        _assert_same(gb, ga)
    batch.close()


def test_batch_mixes_validated_and_declined_blocked_plans(double):
    """One batch, two kinds of masks: on one the blocked plan validates, on the
    other boundary data creeps faster than the halo and the plan is declined;
    the batched loop runs blocked for the first kind and statement by statement
    for the second, and every instance still equals its plain run."""
    import xgrid_amd as xgrid
    from xgrid_amd.batch import Batch
    f2 = xgrid.grid[float, 2]

    def make():
        @xgrid.kernel()
        def creep(u: f2, v: f2) -> None:
            for _ in range(0, 12):
                u[0, 0] = (u[0, 1][0] + u[0, -1][0] + u[1, 0][0] + u[-1, 0][0]) * 0.25 + v[0, 0]
                with xgrid.boundary(1):
                    u[0, 0] = u[0, -1][0] + 1.0
                with xgrid.boundary(2):
                    u[0, 0] = u[0, -1][0] * 0.5
                with xgrid.boundary(3):
                    u[0, 0] = u[0, -1][0] - 0.25
        return creep

    kb, kp = make(), make()
    batch = Batch(kb)
    rng = np.random.default_rng(21)
    members = []
    for i, n in enumerate((150, 120, 90, 140)):
        mask = np.zeros((n, n), np.int32)
        if i % 2 == 0:
            mask[0, :] = 1 + np.arange(n) % 3           # creeping along row 0: declined
        else:
            mask[0, :] = 1                               # one class on row 0: blocked
            mask[:, 0] = 2
        mask[-1, :] = mask[:, -1] = 3
        u0 = rng.random((n, n))
        v0 = rng.random((n, n)) * 1e-3
        pair = []
        for _ in range(2):
            u = xgrid.Grid((n, n), float)
            v = xgrid.Grid((n, n), float)
            u.now[...] = u0
            u.boundary[...] = mask
            v.now[...] = v0
            pair.append((u, v))
        batch.add(*pair[0])
        members.append(pair)
    batch.step(3)
    for (ub, vb), (ua, va) in members:
        for _ in range(3):
            kp(ua, va)
        _assert_same([ub, vb], [ua, va])
    c = batch.counters()
    assert c["validations"] >= 1
    batch.close()


def test_batch_runs_two_phase_statements(double):
    """A statement with a shifted store runs compute-then-store; batched, the
    two phases run for all instances and results equal plain calls."""
    import xgrid_amd as xgrid
    from xgrid_amd.batch import Batch
    f2 = xgrid.grid[float, 2]

    def make():
        @xgrid.kernel()
        def shift(u: f2, w: f2, a: float) -> None:
            u[0, 1] = u[0, 0][0] * a + w[0, 0]       # reads the store grid now: compute, then store
            with xgrid.boundary(1):
                w[0, 0] = w[0, 0][0] + u[0, 0][0]
        return shift

    kb, kp = make(), make()
    batch = Batch(kb)
    rng = np.random.default_rng(5)
    members = []
    for n, a in ((40, 0.5), (77, 1.5), (64, -0.25)):
        mask = np.zeros((n, n), np.int32)
        mask[:, -1] = 1
        pair = []
        for _ in range(2):
            u = xgrid.Grid((n, n), float)
            w = xgrid.Grid((n, n), float)
            u.now[...] = rng.random((n, n)) if not pair else pair[0][0].now
            w.now[...] = rng.random((n, n)) if not pair else pair[0][1].now
            u.boundary[...] = mask
            w.boundary[...] = mask
            pair.append((u, w))
        batch.add(*pair[0], a)
        members.append((pair, a))
    assert kb.native.lowered.stmts[0].stmt.mode == "twophase"
    batch.step(4)
    for ((ub, wb), (ua, wa)), a in members:
        for _ in range(4):
            kp(ua, wa, a)
        _assert_same([ub, wb], [ua, wa])
    batch.close()
