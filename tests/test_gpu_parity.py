"""Device path vs the oracle (restated reference CPU path): bit-exact."""
import numpy as np
import pytest

from oracle import restate as R
from oracle.cpu_path import CpuKernel, HostGrid
from tests.helpers import assert_bitexact, grid_pair, run_pair

pytestmark = pytest.mark.gpu


def test_diffusion_1d_c1_full(double):
    from xgrid_amd import workloads as W
    k = W.make_diffusion_1d()
    u0, mask, dx = W.diffusion_1d_setup(4096)
    nu = 0.3
    dt = 0.2 * dx * dx / nu
    d, h = grid_pair(double, (4096,), float, mask, u0)
    run_pair(k, (d, nu, dt, dx), (h, nu, dt, dx), steps=500)
    assert_bitexact(d, h)
    assert np.array_equal(d.now, R.diffusion_1d(u0, nu, dt, dx, 500))


@pytest.mark.parametrize("n,steps", [(33, 5), (65, 20), (130, 3)])
def test_cavity_small(double, n, steps):
    from xgrid_amd import workloads as W
    k = W.make_cavity()
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    pairs = {name: grid_pair(double, (n, n), float, masks[name]) for name in "bpuv"}
    dev = [pairs[c][0] for c in "bpuv"] + [cfg]
    host = [pairs[c][1] for c in "bpuv"] + [cfg]
    run_pair(k, dev, host, steps=steps)
    for c in "bpuv":
        assert_bitexact(*pairs[c])
    ref = R.cavity(n, steps, cfg.rho, cfg.nu, cfg.dt, cfg.dx, cfg.dy)
    for c, r in zip("bpuv", ref):
        assert np.array_equal(pairs[c][0].now, r)


def test_cavity_full_size_few_steps(double):
    from xgrid_amd import workloads as W
    n, steps = 2049, 2
    k = W.make_cavity()
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    pairs = {name: grid_pair(double, (n, n), float, masks[name]) for name in "bpuv"}
    run_pair(k, [pairs[c][0] for c in "bpuv"] + [cfg], [pairs[c][1] for c in "bpuv"] + [cfg], steps=steps)
    for c in "bpuv":
        assert_bitexact(*pairs[c], levels=[0])


def _cavity_live(xg, n, steps, seed):
    """The cavity on the device and on the oracle from seeded non-zero interior
    velocities, so that `b` and `p` are non-zero from the first call and every
    blocked pressure launch works on live data."""
    from xgrid_amd import workloads as W
    k = W.make_cavity()
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    rng = np.random.default_rng(seed)
    init = {"b": None, "p": None,
            "u": rng.uniform(-0.5, 0.5, (n, n)), "v": rng.uniform(-0.5, 0.5, (n, n))}
    pairs = {c: grid_pair(xg, (n, n), float, masks[c], init[c]) for c in "bpuv"}
    run_pair(k, [pairs[c][0] for c in "bpuv"] + [cfg], [pairs[c][1] for c in "bpuv"] + [cfg], steps=steps)
    return k, pairs


@pytest.mark.parametrize("n", [2049, 1501])
def test_cavity_headline_tile_on_live_data(double, n):
    """The 2049^2 FP64 cavity the benchmark times, on the blocked-loop tile it
    times (48 x 40 core, 1024 work-items, 10 iterations per launch; plan 0), from
    non-zero velocities: b and p are live in every blocked launch.  2049^2 has
    power-of-two spacings (dx = dy = 2^-10), so its uniform divisions take the
    reciprocal form; 1501^2 (dx = 2/1500) keeps the divisions."""
    k, pairs = _cavity_live(double, n, 3, seed=7)
    L = k.native.lowered.loops[0]
    assert (L.core, L.threads, L.tmax, L.variant) == ((48, 40), 1024, 10, "lds")
    assert k.native.loop_plan(0, n * n) == 0
    info = k.native.loop_info(0)
    assert info["state"] == 1 and info["iterations"] == 50 * 3
    assert np.abs(pairs["p"][1].now).max() > 0 and np.abs(pairs["b"][1].now).max() > 0
    for c in "bpuv":
        assert_bitexact(*pairs[c], levels=[0])


def test_cavity_fp32_headline_tile_on_live_data(single):
    """The 2049^2 cavity with FP32 fields on its blocked tile (72 x 40 core, 1024
    work-items; plan 0, not the small-grid tile) from non-zero velocities, bit for
    bit against the oracle in FP32."""
    n = 2049
    k, pairs = _cavity_live(single, n, 3, seed=8)
    L = k.native.lowered.loops[0]
    assert (L.core, L.threads, L.tmax, L.variant) == ((72, 40), 1024, 10, "lds")
    assert k.native.loop_plan(0, n * n) == 0
    info = k.native.loop_info(0)
    assert info["state"] == 1 and info["iterations"] == 50 * 3
    assert pairs["p"][0].now.dtype == np.float32
    for c in "bpuv":
        assert_bitexact(*pairs[c], levels=[0])


@pytest.mark.parametrize("n", [34, 130, 150])
@pytest.mark.parametrize("prec", ["double", "float"])
def test_heat_3d(prec, n, request):
    """130^3: several column tiles (the last partial) and, in FP32, two groups
    of 16 row bands; 150^3: 37 bands, a partial last group in both precisions
    (FP64's 128-column tiles group 32 bands)."""
    xg = request.getfixturevalue("double" if prec == "double" else "single")
    from xgrid_amd import workloads as W
    steps = 25 if n < 100 else 6
    dt_np = np.float64 if prec == "double" else np.float32
    u0, mask, dx = W.heat_3d_setup(n, dt_np, seed=3)
    nu = 0.1
    dt = 0.1 * dx * dx / nu
    k = W.make_heat_3d()
    d, h = grid_pair(xg, (n, n, n), float, mask, u0)
    run_pair(k, (d, nu, dt, dx), (h, nu, dt, dx), steps=steps)
    assert_bitexact(d, h)
    assert np.array_equal(d.now, R.heat_3d(u0, nu, dt, dx, steps, dt_np))


@pytest.mark.parametrize("prec", ["float", "double"])
def test_heat_3d_full_size(prec, request):
    """The 3-D heat workload the bench times, at its own size (512^3, seed 0,
    nu = 0.1, dt = 0.1 dx^2 / nu): the staged kernels (256-column tiles in FP32,
    128 in FP64, 16-byte loads and stores) for 3 steps, bit for bit against the
    numpy restatement of the reference's C."""
    xg = request.getfixturevalue("double" if prec == "double" else "single")
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    n, steps, nu = 512, 3, 0.1
    dt_np = np.float64 if prec == "double" else np.float32
    u0, mask, dx = W.heat_3d_setup(n, dt_np, seed=0)
    dt = 0.1 * dx * dx / nu
    k = W.make_heat_3d()
    low = lower_kernel(k)[0]
    assert low.stmts[0].tile3_cols == (256 if prec == "float" else 128)
    g = xg.Grid((n, n, n), float)
    g.now[...] = u0
    g.boundary[...] = mask
    for _ in range(steps):
        k(g, nu, dt, dx)
    got = g.now
    want = R.heat_3d(u0, nu, dt, dx, steps, dt_np)
    assert got.dtype == want.dtype
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_heat_3d_unstaged_above_cache_size(double, monkeypatch):
    """A grid larger than the 256 MB last-level cache makes 3-D box tiles store
    non-temporally.  The staged tiles are covered at 512^3 above; here staging
    is off (XGRID_STAGE3=0), so the plain tile's non-temporal variant runs:
    330^3 FP64 (287 MB), 2 steps, bit for bit against the numpy restatement."""
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    monkeypatch.setenv("XGRID_STAGE3", "0")
    n, steps, nu = 330, 2, 0.1
    assert n ** 3 * 8 > 256 << 20
    u0, mask, dx = W.heat_3d_setup(n, np.float64, seed=5)
    dt = 0.1 * dx * dx / nu
    k = W.make_heat_3d()
    assert lower_kernel(k)[0].stmts[0].tile3_cols == 64          # unstaged
    g = double.Grid((n, n, n), float)
    g.now[...] = u0
    g.boundary[...] = mask
    for _ in range(steps):
        k(g, nu, dt, dx)
    want = R.heat_3d(u0, nu, dt, dx, steps, np.float64)
    assert np.array_equal(g.now.view(np.uint8), want.view(np.uint8))


def test_convection_and_diffusion_2d(double):
    from xgrid_amd import workloads as W
    nx = 41
    dx = 2 / (nx - 1)
    u0 = np.ones(nx)
    u0[int(.5 / dx):int(1 / dx + 1)] = 2
    mask = np.zeros(nx, np.int32)
    mask[0] = 1
    for maker, args, nonlin in [(W.make_convection_1d, (1.0, .025, dx), False),
                                (W.make_convection_1d_nonlinear, (.025, dx), True)]:
        k = maker()
        d, h = grid_pair(double, (nx,), float, mask, u0)
        run_pair(k, (d, *args), (h, *args), steps=25)
        assert_bitexact(d, h)
        ref = R.convection_1d(u0, 1.0, .025, dx, 25, nonlinear=nonlin)
        assert np.array_equal(d.now, ref)
    n = 101
    dx = 2 / (n - 1)
    dt = .5 * dx
    u0 = np.ones((n, n))
    u0[int(.5 / dx):int(1 / dx) + 1, int(.5 / dx):int(1 / dx) + 1] = 2
    mask = np.zeros((n, n), np.int32)
    mask[0, :] = mask[:, 0] = 1
    k = W.make_convection_2d()
    d, h = grid_pair(double, (n, n), float, mask, u0)
    run_pair(k, (d, 1.0, dt, dx, dx), (h, 1.0, dt, dx, dx), steps=int(.7 / dt))
    assert_bitexact(d, h)
    assert np.array_equal(d.now, R.convection_2d(u0, 1.0, dt, dx, dx, int(.7 / dt)))
    k = W.make_diffusion_2d()
    rng = np.random.default_rng(5)
    u0 = rng.random((70, 70))
    d, h = grid_pair(double, (70, 70), float, W.box_walls((70, 70)), u0)
    run_pair(k, (d, 0.05, 1e-4, 0.02), (h, 0.05, 1e-4, 0.02), steps=30)
    assert_bitexact(d, h)


def test_grids_dropped_behind_queued_kernels(double):
    """Grid memory is released behind the kernels already queued and taken on a
    copy stream without waiting for them.  Grids dropped while their calls may
    still be queued, and new grids of the same size made and uploaded at once
    (likely reusing the released blocks and addresses), each give the numpy
    restatement's result bit for bit, and so does a grid kept across the churn
    (its calls are replayed graphs, keyed by grid identity, not address)."""
    import gc
    from xgrid_amd import workloads as W
    n, nu = 768, 0.1
    dx = 2.0 / (n - 1)
    dt = 0.2 * dx * dx / nu
    k = W.make_diffusion_2d()
    mask = np.ones((n, n), np.int32)
    mask[1:-1, 1:-1] = 0
    rng = np.random.default_rng(7)

    def fresh(u0):
        g = double.Grid((n, n), float)
        g.now[...] = u0
        g.boundary[...] = mask
        return g

    u_keep = rng.random((n, n))
    keep = fresh(u_keep)
    kept = 0
    for _ in range(4):
        a = fresh(rng.random((n, n)))
        for _ in range(30):
            k(a, nu, dt, dx)
        k(keep, nu, dt, dx)
        kept += 1
        del a                        # its calls may still be queued
        gc.collect()
        u1 = rng.random((n, n))
        b = fresh(u1)
        for _ in range(7):
            k(b, nu, dt, dx)
        assert np.array_equal(b.now, R.diffusion_2d(u1, nu, dt, dx, 7))
        del b
    assert np.array_equal(keep.now, R.diffusion_2d(u_keep, nu, dt, dx, kept))
