"""The oracle (restated reference CPU path) pinned against independent checks.

* numpy restatements of every workload (bit-exact: same association, IEEE
  double/float without FMA);
* the reference's own known answers (test.py:123,140,165,178-180);
* closed forms / invariants: the reference's linear index equals row-major for
  1-D and square 2-D grids; Jacobi with zero source and zero boundary stays
  zero; mass conservation of periodic diffusion; a constant field is a fixed
  point of diffusion;
* mutations planted in the oracle are caught by those checks.
"""
import dataclasses
import random

import numpy as np
import pytest

import xgrid_amd as xgrid
from xgrid_amd import workloads as W
from oracle import restate as R
from oracle.cpu_path import CpuKernel, CSource, HostGrid
from tests.helpers import Cplx, struct_kernel, struct_restated, struct_setup


def _host(shape, dtype, init=None, mask=None):
    g = HostGrid(shape, dtype)
    if init is not None:
        g.now[...] = init
    if mask is not None:
        g.boundary[...] = mask
    return g


def test_diffusion_1d_matches_restatement(double):
    u0, mask, dx = W.diffusion_1d_setup(4096)
    nu = 0.3
    dt = 0.2 * dx * dx / nu
    g = _host((4096,), float, u0, mask)
    ck = CpuKernel(W.make_diffusion_1d())
    for _ in range(500):
        ck(g, nu, dt, dx)
    assert np.array_equal(g.now, R.diffusion_1d(u0, nu, dt, dx, 500))


@pytest.mark.parametrize("n,steps", [(17, 3), (41, 10)])
def test_cavity_matches_restatement(double, n, steps):
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    grids = [_host((n, n), float, mask=masks[c]) for c in "bpuv"]
    ck = CpuKernel(W.make_cavity())
    for _ in range(steps):
        ck(*grids, cfg)
    ref = R.cavity(n, steps, cfg.rho, cfg.nu, cfg.dt, cfg.dx, cfg.dy)
    for g, r in zip(grids, ref):
        assert np.array_equal(g.now, r)


def test_cavity_sweep_count_is_51(double):
    """examples/cavity.py runs one sweep before the loop and 50 in it."""
    n = 17
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    grids = [_host((n, n), float, mask=masks[c]) for c in "bpuv"]
    ck = CpuKernel(W.make_cavity())
    for _ in range(3):
        ck(*grids, cfg)
    r50 = R.cavity(n, 3, cfg.rho, cfg.nu, cfg.dt, cfg.dx, cfg.dy, sweeps=50)
    r49 = R.cavity(n, 3, cfg.rho, cfg.nu, cfg.dt, cfg.dx, cfg.dy, sweeps=49)
    assert np.array_equal(grids[1].now, r50[1])
    assert not np.array_equal(grids[1].now, r49[1])


def test_convection_matches_restatement(double):
    nx = 41
    dx = 2 / (nx - 1)
    u0 = np.ones(nx)
    u0[int(.5 / dx):int(1 / dx + 1)] = 2
    mask = np.zeros(nx, np.int32)
    mask[0] = 1
    g = _host((nx,), float, u0, mask)
    ck = CpuKernel(W.make_convection_1d())
    for _ in range(25):
        ck(g, 1.0, .025, dx)
    assert np.array_equal(g.now, R.convection_1d(u0, 1.0, .025, dx, 25))
    g = _host((nx,), float, u0, mask)
    ck = CpuKernel(W.make_convection_1d_nonlinear())
    for _ in range(10):
        ck(g, .025, dx)
    assert np.array_equal(g.now, R.convection_1d(u0, 1.0, .025, dx, 10, nonlinear=True))
    n = 31
    dx = 2 / (n - 1)
    u0 = np.ones((n, n))
    u0[7:16, 7:16] = 2
    m = np.zeros((n, n), np.int32)
    m[0, :] = m[:, 0] = 1
    g = _host((n, n), float, u0, m)
    ck = CpuKernel(W.make_convection_2d())
    for _ in range(12):
        ck(g, 1.0, .5 * dx, dx, dx)
    assert np.array_equal(g.now, R.convection_2d(u0, 1.0, .5 * dx, dx, dx, 12))


@pytest.mark.parametrize("prec,npdt", [("double", np.float64), ("float", np.float32)])
def test_heat_3d_matches_restatement(prec, npdt, request):
    request.getfixturevalue("double" if prec == "double" else "single")
    n = 12
    u0, mask, dx = W.heat_3d_setup(n, npdt, seed=1)
    nu = 0.1
    dt = 0.1 * dx * dx / nu
    g = _host((n, n, n), float, u0, mask)
    ck = CpuKernel(W.make_heat_3d())
    for _ in range(7):
        ck(g, nu, dt, dx)
    assert np.array_equal(g.now, R.heat_3d(u0, nu, dt, dx, 7, npdt))


def test_diffusion_2d_matches_restatement(double):
    rng = np.random.default_rng(0)
    u0 = rng.random((20, 20))
    g = _host((20, 20), float, u0, W.box_walls((20, 20)))
    ck = CpuKernel(W.make_diffusion_2d())
    for _ in range(9):
        ck(g, 0.05, 1e-4, 0.02)
    assert np.array_equal(g.now, R.diffusion_2d(u0, 0.05, 1e-4, 0.02, 9))


# ----------------------------------------------------------- known answers (test.py)
TEMP = 10


@dataclasses.dataclass
class Vector3i:
    x: int
    y: int
    z: int

    @xgrid.function(method=True)
    def dot(self, b: "Vector3i") -> int:
        return self.x * b.x + self.y * b.y + self.z * b.z


def test_known_answers(double):
    @xgrid.kernel()
    def aux(a: int, b: int) -> int:
        return a + b + TEMP

    ck = CpuKernel(aux)
    for _ in range(5):
        a, b = random.randint(0, 1000), random.randint(0, 1000)
        assert ck(a, b) == a + b + TEMP                      # test.py:140

    @xgrid.kernel()
    def aux2(a: Vector3i, b: Vector3i) -> int:
        return a.dot(b)

    ck = CpuKernel(aux2)
    a, b = Vector3i(3, 4, 5), Vector3i(7, 8, 9)
    assert ck(a, b) == a.dot(b) == 3 * 7 + 4 * 8 + 5 * 9     # test.py:165

    @xgrid.kernel()
    def aux3(a: xgrid.grid[int, 2]) -> None:
        a[0, 0] = 4

    g = HostGrid((10, 10), int)
    CpuKernel(aux3)(g)
    assert (g.now == 4).all() and g.now.dtype == np.int32   # test.py:178-180


def test_float32_add_known_answer(single):
    """test.py:123: universe(1.2f, 2.3f) == 3.5 with float arithmetic."""
    @xgrid.kernel()
    def universe(a: float, b: float) -> float:
        return a + b

    assert CpuKernel(universe)(1.2, 2.3) == 3.5


# ----------------------------------------------------------- closed forms / invariants
@pytest.mark.parametrize("shape", [(40,), (23, 23)])
def test_reference_index_equals_rowmajor_1d_square(double, shape):
    rng = np.random.default_rng(2)
    u0 = rng.random(shape)
    mask = W.box_walls(shape) if len(shape) == 2 else np.pad(np.zeros(shape[0] - 2, np.int32), 1, constant_values=1)
    k = W.make_diffusion_2d() if len(shape) == 2 else W.make_diffusion_1d()
    args = (0.05, 1e-4, 0.02)
    a, b = _host(shape, float, u0, mask), _host(shape, float, u0, mask)
    ca, cb = CpuKernel(k, index="reference"), CpuKernel(k, index="rowmajor")
    for _ in range(6):
        ca(a, *args)
        cb(b, *args)
    assert np.array_equal(a.now, b.now)


def test_reference_index_differs_on_nonsquare(double):
    """F1: the reference's formula is not row-major for a non-square grid."""
    cs = CSource(W.make_diffusion_2d(), index="reference")
    assert "s1 * 1 + s0 * grid.shape[0]" in cs.source().replace("(", "").replace(")", "")


def test_jacobi_zero_source_stays_zero(double):
    n = 33
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    grids = [_host((n, n), float, mask=masks[c]) for c in "bpuv"]
    # u, v zero and no lid motion: b = 0, so p must stay exactly 0
    grids[2].boundary[...] = 1
    ck = CpuKernel(W.make_cavity())
    for _ in range(3):
        ck(*grids, cfg)
    assert not grids[1].now.any()


def test_periodic_diffusion_conserves_mass(double):
    xgrid.init(precision="double", overstep="wrap")

    @xgrid.kernel()
    def periodic(u: xgrid.grid[float, 1], c: float) -> None:
        u[0] = u[0] + c * (u[1] - 2.0 * u[0] + u[-1])

    rng = np.random.default_rng(4)
    u0 = rng.random(64)
    g = _host((64,), float, u0)
    ck = CpuKernel(periodic, overstep="wrap")
    for _ in range(50):
        ck(g, 0.25)
    assert abs(g.now.sum() - u0.sum()) < 1e-12 * u0.sum()
    assert g.now.std() < u0.std()


def test_constant_field_is_fixed_point(double):
    g = _host((30, 30), float, np.full((30, 30), 1.0), W.box_walls((30, 30)))
    ck = CpuKernel(W.make_diffusion_2d())
    for _ in range(5):
        ck(g, 0.05, 1e-4, 0.02)
    assert (g.now == 1.0).all()


# ----------------------------------------------------------- planted mistakes are caught
def test_planted_mistakes_fail_the_checks(double, monkeypatch):
    """A sign flip, a swapped index formula and a dropped two-phase step in the
    oracle's C must each break a restatement comparison."""
    import oracle.cpu_path as cp
    n, steps = 17, 5          # b (and so p) first becomes non-zero at step 3
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    ref = R.cavity(n, steps, cfg.rho, cfg.nu, cfg.dt, cfg.dx, cfg.dy)

    def run_with(patch):
        monkeypatch.setattr(cp.CSource, "source", patch)
        grids = [_host((n, n), float, mask=masks[c]) for c in "bpuv"]
        ck = cp.CpuKernel(W.make_cavity())
        for _ in range(steps):
            ck(*grids, cfg)
        monkeypatch.undo()
        return all(np.array_equal(g.now, r) for g, r in zip(grids, ref))

    orig = cp.CSource.source
    assert run_with(orig)
    # a wrong sign of the pressure stencil's vertical offset
    assert not run_with(lambda self: orig(self).replace("(p, _d0 + -1", "(p, _d0 + 1"))
    # the implicit statement run in place (Gauss-Seidel) instead of two-phase (Jacobi)
    assert not run_with(lambda self: orig(self).replace(
        "_value[(_d0) * p.shape[1] + _d1] = ", "*__Grid2d_d_at(p, _d0 + 0, _d1 + 0, 0) = "))
    # loads of the previous level read the current one
    assert not run_with(lambda self: orig(self).replace("_d1 + 1, 1))", "_d1 + 1, 0))"))


# ----------------------------------------------------------- DSL features outside the workloads
def _c_div(a, d):
    """C integer division and remainder (truncation toward zero), from numpy floor ops."""
    q = np.abs(a) // np.abs(d) * np.sign(a) * np.sign(d)
    return q, a - q * d


@pytest.mark.parametrize("mode", ["limit", "wrap"])
def test_overstep_2d_matches_padding_and_roll(mode):
    """generator.py:172-177: `limit` clamps each coordinate (numpy pad mode "edge"),
    `wrap` takes it modulo the extent (numpy roll); a square grid, where the
    reference's index formula is row-major."""
    xgrid.init(precision="double", overstep=mode)

    @xgrid.kernel()
    def edge(u: xgrid.grid[float, 2]) -> None:
        u[0, 0] = (u[0, 1] + u[2, 0]) * 0.5 - u[-1, -1] * 0.25 + u[0, -2]

    rng = np.random.default_rng(41)
    u0 = rng.random((23, 23))
    g = _host((23, 23), float, u0)
    CpuKernel(edge, overstep=mode)(g)
    if mode == "limit":
        P = np.pad(u0, 2, mode="edge")
        at = lambda o0, o1: P[2 + o0:2 + o0 + 23, 2 + o1:2 + o1 + 23]    # noqa: E731
    else:
        at = lambda o0, o1: np.roll(u0, (-o0, -o1), axis=(0, 1))         # noqa: E731
    want = (at(0, 1) + at(2, 0)) * 0.5 - at(-1, -1) * 0.25 + at(0, -2)
    assert np.array_equal(g.now, want)


def test_three_time_levels_leapfrog():
    """parser.py:512-525 and xgrid/xgrid/__init__.py:43-54: loads at times -1 and -2
    read the two previous calls' levels (the ring rotates the oldest level to the
    front, and new levels start at zero), checked against a numpy leapfrog."""
    xgrid.init(precision="double")

    @xgrid.kernel()
    def wave(u: xgrid.grid[float, 1], c: float) -> None:
        u[0] = 2.0 * u[0][-1] - u[0][-2] + c * (u[1][-1] - 2.0 * u[0][-1] + u[-1][-1])
        with xgrid.boundary(1):
            u[0] = 0.0

    n = 50
    rng = np.random.default_rng(42)
    u0 = rng.random(n)
    mask = np.zeros(n, np.int32)
    mask[0] = mask[-1] = 1
    g = _host((n,), float, u0, mask)
    ck = CpuKernel(wave)
    assert ck.depth == 3
    prev, prev2 = u0.copy(), np.zeros(n)
    for _ in range(7):
        ck(g, 0.3)
        new = np.zeros(n)
        new[1:-1] = 2.0 * prev[1:-1] - prev2[1:-1] + 0.3 * (prev[2:] - 2.0 * prev[1:-1] + prev[:-2])
        prev2, prev = prev, new
        assert np.array_equal(g.now, new)
        assert np.array_equal(g._data[1], prev2)


def test_shifted_two_phase_store():
    """generator.py:312-352: an implicit statement with a store offset computes every
    class cell into a temporary first, then stores it at the shifted cell: here
    p[i+1] = p[i-1] for the class-0 cells 1..n-2, all reads before any write."""
    xgrid.init(precision="double")

    @xgrid.kernel()
    def shift(p: xgrid.grid[float, 1]) -> None:
        p[1] = p[-1][0] * 1.5

    n = 31
    rng = np.random.default_rng(43)
    p0 = rng.random(n)
    mask = np.zeros(n, np.int32)
    mask[0] = mask[-1] = 1
    g = _host((n,), float, p0, mask)
    ck = CpuKernel(shift)
    want = p0.copy()
    for _ in range(3):
        ck(g, )
        nxt = want.copy()
        nxt[2:] = want[:-2] * 1.5
        want = nxt
        assert np.array_equal(g.now, want)


def test_integer_division_and_remainder_truncate():
    """C integer '/' and '%' (generator.py:377-385 emits them as written) truncate
    toward zero, unlike Python's floor division, on negative operands too."""
    xgrid.init(precision="double")
    i1 = xgrid.grid[int, 1]

    @xgrid.kernel()
    def idiv(q: i1, r: i1, a: i1, d: int) -> None:
        q[0] = a[0] / d
        r[0] = a[0] % d

    a0 = np.arange(-40, 41, dtype=np.int32) * 7 + 3
    for d in (3, -5, 7, -1):
        q, r, a = _host((81,), int), _host((81,), int), _host((81,), int, a0)
        CpuKernel(idiv)(q, r, a, d)
        wq, wr = _c_div(a0.astype(np.int64), d)
        assert np.array_equal(q.now, wq) and np.array_equal(r.now, wr), d
        assert (np.abs(r.now) < abs(d)).all() and ((r.now == 0) | (np.sign(r.now) == np.sign(a0))).all()


def test_struct_grid_matches_restatement():
    xgrid.init(precision="double")
    shape = (12, 17)
    re, im, tag = struct_setup(shape, 44)
    g = _host(shape, Cplx, mask=W.box_walls(shape))
    assert g.now.dtype.itemsize == 24 and g.now.dtype.fields["tag"][1] == 16     # C layout, aligned
    g.now["re"], g.now["im"], g.now["tag"] = re, im, tag
    ck = CpuKernel(struct_kernel())
    for _ in range(4):
        ck(g, 0.75)
    wre, wim, wtag = struct_restated(re, im, tag, 0.75, 4)
    assert np.array_equal(g.now["re"], wre) and np.array_equal(g.now["im"], wim)
    assert np.array_equal(g.now["tag"], wtag)
