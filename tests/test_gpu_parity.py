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


@pytest.mark.parametrize("n", [34, 130])
@pytest.mark.parametrize("prec", ["double", "float"])
def test_heat_3d(prec, n, request):
    """130^3: three column tiles (the last partial) and two groups of row bands."""
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
