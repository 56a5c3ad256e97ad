"""Device semantics vs the oracle: time levels, ticking, masks, overstep, the
implicit / in-place rules, aliasing, host writes between calls, graph replay
and the reference's known answers (test.py)."""
import ctypes
import dataclasses
import random
import re
import typing

import numpy as np
import pytest

import xgrid_amd as xgrid
from oracle.cpu_path import CpuKernel, HostGrid
from tests.helpers import assert_bitexact, grid_pair, run_pair
from xgrid_amd import workloads as W

pytestmark = pytest.mark.gpu

f1 = xgrid.grid[float, 1]
f2 = xgrid.grid[float, 2]
f3 = xgrid.grid[float, 3]
i2 = xgrid.grid[int, 2]
TEMP = 10


@dataclasses.dataclass
class Vector3i:
    x: int
    y: int
    z: int

    @xgrid.function(method=True)
    def dot(self, b: "Vector3i") -> int:
        return self.x * b.x + self.y * b.y + self.z * b.z


def test_scalar_and_struct_kernels(double):
    @xgrid.kernel()
    def aux(a: int, b: int) -> int:
        return a + b + TEMP

    @xgrid.kernel()
    def aux2(a: Vector3i, b: Vector3i) -> int:
        return a.dot(b)

    for _ in range(5):
        a, b = random.randint(0, 1000), random.randint(0, 1000)
        assert aux(a, b) == a + b + TEMP
        va = Vector3i(*(random.randint(0, 1000) for _ in range(3)))
        vb = Vector3i(*(random.randint(0, 1000) for _ in range(3)))
        assert aux2(va, vb) == va.dot(vb)
    with pytest.raises(TypeError):
        aux(1)


def test_float_add_known_answer(single):
    @xgrid.kernel()
    def universe(a: float, b: float) -> float:
        return a + b

    assert universe(1.2, 2.3) == 3.5


def test_int_grid_fill(double):
    @xgrid.kernel()
    def aux(a: i2) -> None:
        a[0, 0] = 4

    g = xgrid.Grid((10, 10), dtype=int)
    aux(g)
    assert g.now.dtype == np.int32 and (g.now == 4).all()


def test_index_guard_runs_and_matches_inside_the_array(double):
    @xgrid.kernel()
    def aux(a: i2) -> None:
        a[0, 0] = a[-1, -1][-1]

    rng = np.random.default_rng(1)
    init = rng.integers(0, 100, (10, 10), dtype=np.int32)
    d, h = grid_pair(xgrid, (10, 10), int, None, init)
    run_pair(aux, (d,), (h,), steps=3)
    # cells 0..10 read before element 0 (undefined in the reference); each step
    # carries those values 11 cells further
    dv, hv = d.now.reshape(-1), h.now.reshape(-1)
    assert np.array_equal(dv[33:], hv[33:])


def test_readme_elementwise_ticks_every_grid(double):
    @xgrid.kernel()
    def elementwise_mul(result: f1, a: f1, b: f1) -> None:
        result[0] = a[0] * b[0]

    rng = np.random.default_rng(0)
    av, bv = rng.random(1000), rng.random(1000)
    r = xgrid.Grid((1000,), float)
    a = xgrid.Grid((1000,), float)
    b = xgrid.Grid((1000,), float)
    a.now[:] = av
    b.now[:] = bv
    elementwise_mul(r, a, b)
    assert np.array_equal(r.now, av * bv)
    assert not a.now.any() and not b.now.any()
    elementwise_mul(r, a, b)
    assert not r.now.any()


def test_three_time_levels(double):
    @xgrid.kernel()
    def leapfrog(u: f1, c: float) -> None:
        u[0] = u[0][-2] + c * (u[1] - u[-1])
        with xgrid.boundary(1):
            u[0] = 0.0

    rng = np.random.default_rng(2)
    mask = np.zeros(64, np.int32)
    mask[0] = mask[-1] = 1
    d, h = grid_pair(xgrid, (64,), float, mask, rng.random(64))
    run_pair(leapfrog, (d, 0.3), (h, 0.3), steps=9)
    assert d.time_levels == 3
    assert_bitexact(d, h)


@pytest.mark.parametrize("shape", [(37, 53), (8, 9, 11)])
def test_non_square_grids_rowmajor(double, shape):
    from xgrid_amd import workloads as W
    rng = np.random.default_rng(3)
    if len(shape) == 2:
        k = W.make_diffusion_2d()
    else:
        k = W.make_heat_3d()
    d, h = grid_pair(xgrid, shape, float, W.box_walls(shape), rng.random(shape))
    run_pair(k, (d, 0.05, 1e-4, 0.02), (h, 0.05, 1e-4, 0.02), steps=5)
    assert_bitexact(d, h)


@pytest.mark.parametrize("mode", ["wrap", "limit"])
def test_overstep_modes(mode):
    xgrid.init(precision="double", overstep=mode)

    @xgrid.kernel()
    def smooth(u: f2, c: float) -> None:
        u[0, 0] = u[0, 0] + c * (u[1, 0] + u[-1, 0] + u[0, 1] + u[0, -2] - 4.0 * u[0, 0])

    rng = np.random.default_rng(4)
    d, h = grid_pair(xgrid, (24, 24), float, None, rng.random((24, 24)))
    run_pair(smooth, (d, 0.1), (h, 0.1), steps=6)
    assert_bitexact(d, h)
    if mode == "wrap":
        assert abs(d.now.sum() - h.now.sum()) < 1e-9


@pytest.mark.parametrize("mode", ["wrap", "limit"])
def test_overstep_modes_3d_staged(mode):
    """FP64 3-D statements with every cell in class 0: tiles away from the faces
    take the staged loads, tiles at the faces the overstep rule (wrap or clamp)."""
    xgrid.init(precision="double", overstep=mode)

    @xgrid.kernel()
    def smooth3(u: f3, c: float) -> None:
        u[0, 0, 0] = u[0, 0, 0] + c * (u[1, 0, 0] + u[-1, 0, 0] + u[0, 1, 0] + u[0, -1, 0] +
                                       u[0, 0, 1] + u[0, 0, -1] - 6.0 * u[0, 0, 0])

    rng = np.random.default_rng(6)
    shape = (12, 13, 150)
    d, h = grid_pair(xgrid, shape, float, None, rng.random(shape))
    run_pair(smooth3, (d, 0.1), (h, 0.1), steps=4)
    assert_bitexact(d, h)


def test_3d_staged_neighbours_of_other_planes(double):
    """Staged FP64 loads of row and column neighbours on the planes above and
    below (LDS copies and lane shifts of planes other than the cell's own),
    with a diagonal load left global, equal the oracle."""
    @xgrid.kernel()
    def k(u: f3, v: f3) -> None:
        u[0, 0, 0] = (v[1, 1, 0] + v[-1, 0, 1] * 0.5 + v[1, -1, 0] * 0.25 - v[-1, 0, -1] + v[0, 0, 0] +
                      v[0, 1, 0] * 0.125 - v[1, 1, 1])
        with xgrid.boundary(1):
            u[0, 0, 0] = v[0, 0, 0] * 2.0

    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    src = lower_kernel(k)[0].device_src
    cpl = int(re.search(r"static constexpr int CPL = (\d)", src).group(1))
    # two plane offsets staged in LDS: 4 planes x 6 rows x (64 x CPL) columns of 8 B each
    assert "ST3 = true" in src and f"ST3_LDS = {2 * 4 * 6 * 64 * cpl * 8}" in src
    rng = np.random.default_rng(12)
    shape = (14, 19, 140)
    du, hu = grid_pair(xgrid, shape, float, W.box_walls(shape), rng.random(shape))
    dv, hv = grid_pair(xgrid, shape, float, None, rng.random(shape))
    run_pair(k, (du, dv), (hu, hv), steps=3)
    assert_bitexact(du, hu)
    assert_bitexact(dv, hv)


@pytest.mark.parametrize("prec", ["float", "double"])
def test_3d_staged_mixed_element_sizes(prec, request):
    """A 3-D statement reading neighbours of a float grid and of an int grid:
    with FP32 floats both arrays are 4 bytes and both are staged (the int array
    through the same lane shifts and LDS copy); with FP64 floats only the 8-byte
    array is, and the int array keeps global loads.  Bit for bit against the
    oracle on rows with two full 256-column tiles and a partial one."""
    xg = request.getfixturevalue("double" if prec == "double" else "single")
    i3 = xgrid.grid[int, 3]

    @xgrid.kernel()
    def mixed(u: f3, v: f3, w: i3) -> None:
        u[0, 0, 0] = (v[0, 0, 1] * 0.5 + v[0, -1, 0] + typing.cast(float, w[0, 0, -1] - w[0, 1, 0]) * 0.25 +
                          v[1, 0, 0] * typing.cast(float, w[0, 0, 0]))

    from xgrid_amd.backend import lower_kernel
    src = lower_kernel(mixed)[0].device_src
    st = src[src.index("struct xg_s0_t "):src.index("extern", src.index("struct xg_s0_t "))]
    assert "ST3 = true" in st
    if prec == "float":
        assert "CPL = 4" in st and "int32_t q" in st          # the int array staged too
    else:
        assert "CPL = 2" in st and "int32_t q" not in st      # 4-byte array: global loads
    rng = np.random.default_rng(31)
    shape = (6, 9, 600)
    du, hu = grid_pair(xg, shape, float, W.box_walls(shape), rng.random(shape))
    dv, hv = grid_pair(xg, shape, float, None, rng.random(shape))
    dw, hw = grid_pair(xg, shape, int, None, rng.integers(-50, 50, shape).astype(np.int32))
    run_pair(mixed, (du, dv, dw), (hu, hv, hw), steps=2)
    assert_bitexact(du, hu)


def test_twophase_store_offset(double):
    @xgrid.kernel()
    def shift(u: f1) -> None:
        u[1] = u[-1][0] * 2.0

    rng = np.random.default_rng(5)
    mask = np.zeros(40, np.int32)
    mask[0] = mask[-1] = 7              # cell 0 would read before the array
    d, h = grid_pair(xgrid, (40,), float, mask, rng.random(40))
    run_pair(shift, (d,), (h,), steps=4)
    # cells 1..38 are stored from cells 0..37 of the same level (reads before writes)
    assert_bitexact(d, h)


def test_same_class_reads_run_compute_then_store(double):
    """A boundary statement reading its own class at time 0 is a race in the
    reference's parallel C; the device runs it reading the old values."""
    @xgrid.kernel()
    def chain(u: f2) -> None:
        with xgrid.boundary(1):
            u[0, 0] = u[0, -1][0] + 1.0

    mask = np.zeros((6, 10), np.int32)
    mask[2, :] = 1
    init = np.arange(60, dtype=np.float64).reshape(6, 10)
    g = xgrid.Grid((6, 10), float)
    g.now[...] = init
    g.boundary[...] = mask
    chain(g)                            # depth 1: level 0 is the initial data
    exp = init.copy()
    flat_old = init.reshape(-1)
    for j in range(10):
        exp[2, j] = flat_old[2 * 10 + j - 1] + 1.0
    assert np.array_equal(g.now, exp)


def test_aliased_grid_argument(double):
    @xgrid.kernel()
    def two(a: f1, b: f1) -> None:
        a[0] = b[1] + 1.0

    rng = np.random.default_rng(6)
    init = rng.random(30)
    d = xgrid.Grid((30,), float)
    h = HostGrid((30,), float)
    d.now[:] = init
    h.now[:] = init
    d.boundary[-1] = h.boundary[-1] = 1
    ck = CpuKernel(two)
    for _ in range(3):
        two(d, d)
        ck(h, h)
    assert_bitexact(d, h)


def test_host_writes_and_mask_changes_between_calls(double):
    from xgrid_amd import workloads as W
    k = W.make_diffusion_2d()
    rng = np.random.default_rng(7)
    d, h = grid_pair(xgrid, (33, 33), float, W.box_walls((33, 33)), rng.random((33, 33)))
    ck = CpuKernel(k)
    for step in range(8):
        if step == 3:
            v = rng.random((33, 33))
            d.now[...] = v
            h.now[...] = v
        if step == 5:
            m = W.box_walls((33, 33))
            m[10:14, 10:20] = 1        # an obstacle
            d.boundary[...] = m
            h.boundary[...] = m
        k(d, 0.05, 1e-4, 0.02)
        ck(h, 0.05, 1e-4, 0.02)
    assert_bitexact(d, h)


def test_list_and_masked_classes(double):
    @xgrid.kernel()
    def k(u: f2, v: f2) -> None:
        u[0, 0] = v[0, 1] + v[-1, 0] * 0.5
        with xgrid.boundary(3):
            u[0, 0] = v[0, 0] * 2.0 - 1.0
        with xgrid.boundary(5):
            u[0, 0] = u[0, 0][0] + v[1, 1]

    rng = np.random.default_rng(8)
    n = 50
    mask = np.where(rng.random((n, n)) < 0.3, 0, 9).astype(np.int32)  # class 0: irregular (masked)
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = 9           # keep reads inside the array
    inner = [i * n + j for i in range(1, n - 1) for j in range(1, n - 1)]
    idx = rng.choice(inner, 40, replace=False)
    mask.reshape(-1)[idx[:20]] = 3                                     # sparse: list
    mask.reshape(-1)[idx[20:]] = 5
    du, hu = grid_pair(xgrid, (n, n), float, mask, rng.random((n, n)))
    dv, hv = grid_pair(xgrid, (n, n), float, None, rng.random((n, n)))
    run_pair(k, (du, dv), (hu, hv), steps=4)
    assert du.class_info(3)["kind"] == "list" and du.class_info(0)["kind"] in ("masked", "list")
    assert_bitexact(du, hu)
    assert_bitexact(dv, hv)


@pytest.mark.parametrize("shape", [(9, 11, 150), (6, 7, 64), (5, 40, 200), (7, 9, 259), (6, 10, 300)])
def test_3d_classes_at_any_column(double, shape):
    _classes_at_any_column(shape)


@pytest.mark.parametrize("shape", [(5, 9, 150), (6, 10, 600), (7, 9, 518)])
def test_3d_classes_at_any_column_fp32(single, shape):
    """The same classes over FP32 arrays, staged four columns per lane on
    256-column tiles: rows of 600 columns (a multiple of 4: 16-byte loads and
    stores) and 518 (one element at a time), each with two full tiles and a
    partial one."""
    _classes_at_any_column(shape)


def _classes_at_any_column(shape):
    """3-D column tiles start at multiples of 64 columns.  A box class and a box
    with holes that start and end mid-row, and the remaining interior, are each
    updated exactly once and only inside their class; the class reading its own
    cells at time 0 runs compute-then-store.  Rows of 259 and 300 columns hold
    two full 128-column tiles and a partial one: an odd row pitch reads the
    staged columns one element at a time, an even one as 16-byte pairs.

    The oracle runs sequentially here.  Class 4 reads u[0, 0, 1] of its own class
    at time 0, and the reference's OpenMP nest (parallel for collapse(3), in
    place) races where one thread's chunk of cells ends: that thread reads the
    next cell after or before its owner rewrote it, depending on timing and the
    thread count.  Sequential order reads the cell before it is rewritten, which
    is what compute-then-store gives on the device."""
    @xgrid.kernel()
    def k(u: f3, v: f3) -> None:
        u[0, 0, 0] = v[0, 0, 1] + v[-1, 0, 0] * 0.5 + v[0, 1, 0] * 0.25
        with xgrid.boundary(2):
            u[0, 0, 0] = v[0, 0, -1] * 2.0 - v[1, 0, 0]
        with xgrid.boundary(4):
            u[0, 0, 0] = u[0, 0, 1][0] + v[0, -1, 1]

    rng = np.random.default_rng(11)
    n0, n1, n2 = shape
    mask = np.full(shape, 9, np.int32)
    mask[1:-1, 1:-1, 1:-1] = 0
    h1 = n1 // 2
    mask[1:-1, 1:h1, 5:n2 - 3] = 2                                     # a box from column 5
    holes = rng.random((n0 - 2, n1 - 1 - h1, n2 - 1 - (n2 // 2 + 3))) < 0.3
    sub = mask[1:-1, h1:-1, n2 // 2 + 3:-1]
    sub[...] = np.where(holes, 0, 4)                                   # mid-row, with holes
    du, hu = grid_pair(xgrid, shape, float, mask, rng.random(shape))
    dv, hv = grid_pair(xgrid, shape, float, None, rng.random(shape))
    run_pair(k, (du, dv), (hu, hv), steps=3, parallel=False)
    assert du.class_info(2)["kind"] == "box" and du.class_info(2)["lo"][2] == 5
    assert du.class_info(4)["kind"] == "masked"
    assert_bitexact(du, hu)
    assert_bitexact(dv, hv)


def _pair_kernel():
    @xgrid.kernel()
    def two(u: f2, v: f2, w: f2) -> None:
        u[0, 0] = w[0, 1] + w[-1, 0] * 0.5 + v[0, 0]
        v[0, 0] = w[1, 0] - w[0, -1] * 0.25 + u[0, 1]
        with xgrid.boundary(1):
            u[0, 0] = 0.5
    return two


def _forward_kernel():
    @xgrid.kernel()
    def fwd(u: f2, v: f2, w: f2) -> None:
        u[0, 0] = w[0, 1] * 0.5 - w[-1, 0]
        v[0, 0] = v[0, 1] + u[0, 0][0] * 3.0 + w[0, 0][0]    # u at time 0 at its own cell: forwarded
        with xgrid.boundary(1):
            u[0, 0] = 0.25
    return fwd


@pytest.mark.parametrize("case", ["paired", "aliased", "different_boxes", "forwarded", "forward_aliased"])
def test_statement_pairs(double, case):
    """Two consecutive in-place statements over the same box run as one launch
    (the second reads the first's store grid only at time -1, or at time 0 at
    its own cell, where the first's value is forwarded).  When the same grid is
    passed for both stores, another argument aliases the first's store, or the
    classes are different boxes, they run one after the other.  Every case
    equals the oracle bit for bit."""
    from xgrid_amd import workloads as W
    k = _forward_kernel() if case.startswith("forward") else _pair_kernel()
    rng = np.random.default_rng(21)
    n = (90, 130)                                # interiors above the chaining cap of 4096 cells
    walls = W.box_walls(n)
    du, hu = grid_pair(xgrid, n, float, walls, rng.random(n))
    vmask = walls.copy()
    if case == "different_boxes":
        vmask[1:-1, -2] = 1                      # class 0 of v is a narrower box
    dv, hv = grid_pair(xgrid, n, float, vmask, rng.random(n))
    dw, hw = grid_pair(xgrid, n, float, None, rng.random(n))
    if case == "aliased":
        dargs, hargs = (du, du, dw), (hu, hu, hw)
    elif case == "forward_aliased":
        dargs, hargs = (du, dv, du), (hu, hv, hu)        # w[0, 0][0] then reads u's new values
    else:
        dargs, hargs = (du, dv, dw), (hu, hv, hw)
    run_pair(k, dargs, hargs, steps=5)
    for d, h in zip(dargs, hargs):
        assert_bitexact(d, h, levels=[0])
    # a pair counts 2 fused statements per run (graph capture attempts run a call twice)
    fused = k.native.counters()["fused"]
    assert (fused > 0) == (case in ("paired", "forwarded")), fused


def test_float_precision_mixed_literals(single):
    """precision='float': literals are C doubles, so arithmetic is mixed (F11)."""
    @xgrid.kernel()
    def k(u: f1, c: float) -> None:
        u[0] = u[0] * 0.1 + c * u[1] / 3.0

    rng = np.random.default_rng(9)
    mask = np.zeros(100, np.int32)
    mask[-1] = 1
    d, h = grid_pair(xgrid, (100,), float, mask, rng.random(100).astype(np.float32))
    run_pair(k, (d, np.float32(0.3).item()), (h, np.float32(0.3).item()), steps=5)
    assert_bitexact(d, h)


def test_graphs_and_fusion_do_not_change_results(double):
    from xgrid_amd import workloads as W
    n = 65
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    out = []
    for graphs, fuse in [(True, True), (False, False), (True, False)]:
        xgrid.init(precision="double", graphs=graphs, fuse=fuse)
        k = W.make_cavity()
        gs = [xgrid.Grid((n, n), float) for _ in "bpuv"]
        for g, c in zip(gs, "bpuv"):
            g.boundary[...] = masks[c]
        # the buffer states of a call repeat with a period of up to 6 calls (ticks of
        # 2-level grids, 3-buffer rotations of the Jacobi grid); the third sighting
        # of a state replays its graph
        for _ in range(14):
            k(*gs, cfg)
        out.append([g.now.copy() for g in gs])
        if graphs:
            assert k.native.counters()["graph_replays"] >= 2
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)


def test_graph_cache_eviction_keeps_results(double):
    """More live grid sets than the kernel's graph cache holds (16 entries): each
    set is called 14 times in a row (its buffer states repeat within 6 calls and a
    state's graph is captured on its second sighting, so the last calls replay),
    and entries of earlier sets are evicted while recent replays may still run.
    Every set ends equal to the same calls without graphs."""
    from xgrid_amd import workloads as W
    n, sets, rounds, calls = 33, 20, 2, 14
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    res = {}
    for graphs in (True, False):
        xgrid.init(precision="double", graphs=graphs)
        k = W.make_cavity()
        gsets = []
        for i in range(sets):
            gs = [xgrid.Grid((n, n), float) for _ in "bpuv"]
            for g, c in zip(gs, "bpuv"):
                g.boundary[...] = masks[c]
            gs[2].now[n // 2, 1 + i % (n - 2)] = 1.0      # a different start per set
            gsets.append(gs)
        for _ in range(rounds):
            for gs in gsets:
                for _ in range(calls):
                    k(*gs, cfg)
        res[graphs] = [[g.now.copy() for g in gs] for gs in gsets]
        if graphs:
            assert k.native.counters()["graph_replays"] > 0
    for a_set, b_set in zip(res[True], res[False]):
        for a, b in zip(a_set, b_set):
            assert np.array_equal(a, b)


def test_division_by_scalar_is_ieee_for_every_double(double):
    """A device division must give the bits of IEEE a / c (the reference's C
    divides with the CPU's correctly rounded division): random bit patterns (all
    exponents, subnormals, infinities, NaNs), ordinary values and signed zeros,
    over divisors of every size and sign, and zero."""
    @xgrid.kernel()
    def divk(u: f1, a: f1, c: float) -> None:
        u[0] = a[0][0] / c

    rng = np.random.default_rng(21)
    n = 1 << 20
    vals = rng.integers(0, 2**63, n, dtype=np.int64).view(np.float64).copy()
    vals[: n // 2] = rng.random(n // 2) * 10.0 ** rng.integers(-30, 30, n // 2)
    vals[::3] = np.negative(vals[::3])             # a sign flip (NaNs included)
    vals[:8] = [0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 1.7976931348623157e308, 2.0 ** -900]
    # numerators near 2^+-500
    vals[8:16] = [1.5 * 2.0 ** 500, 1.5 * 2.0 ** 501, 1.5 * 2.0 ** -500, 1.5 * 2.0 ** -501,
                  -(2.0 ** 500), -(2.0 ** -500), np.nextafter(2.0 ** 501, 0), np.nextafter(2.0 ** -500, 0)]
    a = xgrid.Grid((n,), float)
    a.now[:] = vals
    u = xgrid.Grid((n,), float)
    # powers of two take the reciprocal form (the largest, 2^1023, has the subnormal
    # reciprocal 2^-1023); the subnormal 2^-1023 itself and the other divisors divide
    # (all-ones and near-one significands, very large and very small ones among them)
    for c in (3.0, 0.1, -7.25, 1e-5, 6.02e23, 2.0 ** -120, 2.0 ** 150, -1e300, 0.0,
              1.0, -0.5, 2.0 ** -10, 2.0 ** -18, 2.0 ** -1022, -(2.0 ** 1023), 2.0 ** 1000, 2.0 ** -1023, np.inf,
              np.nextafter(2.0, 0), np.nextafter(1.0, 2), 1.5 * 2.0 ** 400, 1.5 * 2.0 ** 401,
              1.5 * 2.0 ** -400, 1.5 * 2.0 ** -401, 3e-300, 7e250, 2.0 / 2048 ** 2 * 3.0):
        divk(u, a, c)
        with np.errstate(all="ignore"):
            want = vals / c
        got = u.now
        nan = np.isnan(want)
        assert np.array_equal(np.isnan(got), nan), c
        assert np.array_equal(got[~nan].view(np.uint64), want[~nan].view(np.uint64)), c


def test_division_by_scalar_is_ieee_for_every_float(single):
    """The FP32 counterpart: float grids divided by a float scalar, including the
    power-of-two divisors of the reciprocal form (2^127: subnormal reciprocal
    2^-127) and subnormal results, against numpy's IEEE float32 division."""
    @xgrid.kernel()
    def divf(u: f1, a: f1, c: float) -> None:
        u[0] = a[0][0] / c

    rng = np.random.default_rng(22)
    n = 1 << 20
    vals = rng.integers(0, 2**31, n, dtype=np.int64).astype(np.uint32).view(np.float32).copy()
    vals[: n // 2] = (rng.random(n // 2) * 10.0 ** rng.integers(-20, 20, n // 2)).astype(np.float32)
    vals[::3] = np.negative(vals[::3])
    vals[:8] = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, 3.4028235e38, 2.0 ** -100], np.float32)
    # numerators near 2^+-60
    vals[8:14] = np.array([1.5 * 2.0 ** 60, 1.5 * 2.0 ** 61, 1.5 * 2.0 ** -60, 1.5 * 2.0 ** -61,
                           -(2.0 ** 60), -(2.0 ** -60)], np.float32)
    a = xgrid.Grid((n,), float)
    a.now[:] = vals
    u = xgrid.Grid((n,), float)
    assert a.now.dtype == np.float32
    for c in (3.0, 0.1, -7.25, 2.0 ** -3, 2.0 ** 20, -(2.0 ** 127), 2.0 ** -126, 2.0 ** -127, 1.0, 0.0,
              float(np.nextafter(np.float32(2), np.float32(0))), float(np.nextafter(np.float32(1), np.float32(2))),
              1.5 * 2.0 ** 40, 1.5 * 2.0 ** 41, 1.5 * 2.0 ** -40, 1.5 * 2.0 ** -41, 1e-30, 7e30, 1.1):
        cf = np.float32(c)
        divf(u, a, float(cf))
        with np.errstate(all="ignore"):
            want = vals / cf
        got = u.now
        assert got.dtype == np.float32
        nan = np.isnan(want)
        assert np.array_equal(np.isnan(got), nan), c
        assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32)), c


def test_blocked_loop_revalidates_longer_launches(double):
    """A blocked plan proven on a 3-iteration launch is not trusted for 10: the
    boundary row below moves data 3 cells per iteration (classes 1, 2, 3 each
    read the cell the previous statement just wrote), which a 12-cell halo holds
    for 3 iterations but not for 10.  The first call (trip 3) validates and runs
    blocked; the second (trip 12) must check again, decline, and run unblocked,
    with the oracle's bits in both."""
    @xgrid.kernel()
    def creep_n(u: f2, v: f2, trip: int) -> None:
        for _ in range(0, trip):
            u[0, 0] = (u[0, 1][0] + u[0, -1][0] + u[1, 0][0] + u[-1, 0][0]) * 0.25 + v[0, 0]
            with xgrid.boundary(1):
                u[0, 0] = u[0, -1][0] + 1.0
            with xgrid.boundary(2):
                u[0, 0] = u[0, -1][0] * 0.5
            with xgrid.boundary(3):
                u[0, 0] = u[0, -1][0] - 0.25

    n = 150
    mask = np.zeros((n, n), np.int32)
    mask[0, :] = 1 + np.arange(n) % 3
    mask[-1, :] = mask[:, 0] = mask[:, -1] = 4
    rng = np.random.default_rng(13)
    du, hu = grid_pair(xgrid, (n, n), float, mask, rng.random((n, n)))
    dv, hv = grid_pair(xgrid, (n, n), float, None, rng.random((n, n)) * 1e-3)
    run_pair(creep_n, (du, dv, 3), (hu, hv, 3), steps=1)
    assert creep_n.native.lowered.loops[0].tmax == 10
    info = creep_n.native.loop_info(creep_n.native.loop_plan(0, n * n))
    assert info["state"] == 1 and info["iterations"] == 3
    assert_bitexact(du, hu)
    run_pair(creep_n, (du, dv, 12), (hu, hv, 12), steps=1)
    info = creep_n.native.loop_info(creep_n.native.loop_plan(0, n * n))
    assert info["state"] == -1 and info["iterations"] == 3
    assert_bitexact(du, hu)


def test_pointer_argument(double):
    @xgrid.kernel()
    def k(u: f1, s: xgrid.ptr[float]) -> None:
        u[0] = u[1] * s

    rng = np.random.default_rng(10)
    mask = np.zeros(20, np.int32)
    mask[-1] = 1
    d, h = grid_pair(xgrid, (20,), float, mask, rng.random(20))
    s1, s2 = ctypes.c_double(1.5), ctypes.c_double(1.5)
    run_pair(k, (d, s1), (h, s2), steps=2)
    assert_bitexact(d, h)


def test_control_flow_and_shape_intrinsics(double):
    @xgrid.kernel()
    def k(u: f2, n: int) -> int:
        total = 0
        for i in range(0, n):
            if i % 2 == 0:
                u[0, 0] = u[0, 0][0] + 1.0
            else:
                u[0, 0] = u[0, 0][0] * 0.5
            total = total + xgrid.shape(u, 1)
        while total > 100:
            total = total - 7
        return total + xgrid.dimension(u)

    rng = np.random.default_rng(11)
    d, h = grid_pair(xgrid, (5, 12), float, None, rng.random((5, 12)))
    run_pair(k, (d, 5), (h, 5), steps=2)
    assert_bitexact(d, h)


def test_blocked_loop_matches_unblocked_and_oracle(double):
    from xgrid_amd import workloads as W
    n, steps = 97, 4
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    res = []
    for blocked in (True, False):
        xgrid.init(precision="double", blocked=blocked)
        k = W.make_cavity()
        pairs = {c: grid_pair(xgrid, (n, n), float, masks[c]) for c in "bpuv"}
        run_pair(k, [pairs[c][0] for c in "bpuv"] + [cfg], [pairs[c][1] for c in "bpuv"] + [cfg], steps=steps)
        for c in "bpuv":
            assert_bitexact(*pairs[c], levels=[0])
        info = k.native.loop_info(k.native.loop_plan(0, n * n))
        if blocked:
            assert info["state"] == 1 and info["iterations"] == 50 * steps
        else:
            assert info["iterations"] == 0
        res.append([pairs[c][0].now.copy() for c in "bpuv"])
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def test_blocked_loop_declined_when_halo_too_small(double):
    """Three hops along a boundary row per iteration (classes 1, 2, 3 cycling
    along row 0, each statement reading the class before it) outrun the halo
    (one cell per iteration): validation must decline the blocked plan and the
    loop falls back to per-statement launches with identical results."""
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

    n = 150
    mask = np.zeros((n, n), np.int32)
    mask[0, :] = 1 + np.arange(n) % 3
    mask[-1, :] = mask[:, 0] = mask[:, -1] = 4
    rng = np.random.default_rng(12)
    du, hu = grid_pair(xgrid, (n, n), float, mask, rng.random((n, n)))
    dv, hv = grid_pair(xgrid, (n, n), float, None, rng.random((n, n)) * 1e-3)
    run_pair(creep, (du, dv), (hu, hv), steps=2)
    assert creep.native.loop_info(creep.native.loop_plan(0, n * n))["state"] == -1
    assert_bitexact(du, hu)


@pytest.mark.parametrize("shape", [(40, 64, 512, 6, "reg"), (24, 64, 512, 4, "reg"), (48, 64, 768, 10, "reg"),
                                   (32, 64, 512, 8, "lds"), (24, 64, 256, 6, "lds"), (72, 64, 1024, 10, "lds"),
                                   (48, 64, 768, 10, "lds"), (48, 64, 1024, 10, "lds")])
def test_blocked_loop_shapes_and_layouts_match_oracle(double, shape):
    """Every tile shape and both layouts (region in LDS / columns in registers)
    of the blocked loop give the oracle's bits, on grids whose edges cut tiles."""
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import build_kernel
    for n, steps in ((97, 3), (130, 2)):
        cfg = W.cavity_config(n, dt=2e-6)
        masks = W.cavity_masks(n)
        k = W.make_cavity()
        k.native, k.depth = build_kernel(k, shape)
        assert k.native.lowered.loops and k.native.lowered.loops[0].variant == shape[4]
        pairs = {c: grid_pair(xgrid, (n, n), float, masks[c]) for c in "bpuv"}
        run_pair(k, [pairs[c][0] for c in "bpuv"] + [cfg], [pairs[c][1] for c in "bpuv"] + [cfg], steps=steps)
        for c in "bpuv":
            assert_bitexact(*pairs[c], levels=[0])
        info = k.native.loop_info(0)
        assert info["state"] == 1 and info["iterations"] == 50 * steps


def test_fp32_default_tiles_by_grid_size(single):
    """The default build holds two tiles for the blocked loop: plain calls on grids
    of at most 800^2 FP32 cells run the small-grid tile (40 x 40 core, loop plan 1),
    larger grids the 72 x 40 one (plan 0).  Both give the oracle's bits."""
    from xgrid_amd import workloads as W
    k = W.make_cavity()
    for n, plan, steps in ((97, 1, 2), (810, 0, 1)):
        cfg = W.cavity_config(n, dt=2e-6)
        masks = W.cavity_masks(n)
        pairs = {c: grid_pair(xgrid, (n, n), float, masks[c]) for c in "bpuv"}
        before = k.native.loop_info(plan)["iterations"] if k.native is not None else 0
        run_pair(k, [pairs[c][0] for c in "bpuv"] + [cfg], [pairs[c][1] for c in "bpuv"] + [cfg], steps=steps)
        for c in "bpuv":
            assert_bitexact(*pairs[c], levels=[0])
        L = k.native.lowered.loops[0]
        assert L.core == (72, 40) and L.alt is not None and L.alt.core == (40, 40)
        assert k.native.loop_plan(0, n * n) == plan
        info = k.native.loop_info(plan)
        assert info["state"] == 1 and info["iterations"] - before == 50 * steps


def test_register_layout_declines_when_halo_too_small(double):
    """The register layout's validation declines the same too-fast boundary
    creep the LDS layout declines, and the fallback keeps the oracle's bits."""
    from xgrid_amd.backend import build_kernel

    @xgrid.kernel()
    def creep2(u: f2, v: f2) -> None:
        for _ in range(0, 12):
            u[0, 0] = (u[0, 1][0] + u[0, -1][0] + u[1, 0][0] + u[-1, 0][0]) * 0.25 + v[0, 0]
            with xgrid.boundary(1):
                u[0, 0] = u[0, -1][0] + 1.0
            with xgrid.boundary(2):
                u[0, 0] = u[0, -1][0] * 0.5
            with xgrid.boundary(3):
                u[0, 0] = u[0, -1][0] - 0.25

    creep2.native, creep2.depth = build_kernel(creep2, (40, 64, 512, 6, "reg"))
    n = 150
    mask = np.zeros((n, n), np.int32)
    mask[0, :] = 1 + np.arange(n) % 3
    mask[-1, :] = mask[:, 0] = mask[:, -1] = 4
    rng = np.random.default_rng(12)
    du, hu = grid_pair(xgrid, (n, n), float, mask, rng.random((n, n)))
    dv, hv = grid_pair(xgrid, (n, n), float, None, rng.random((n, n)) * 1e-3)
    run_pair(creep2, (du, dv), (hu, hv), steps=2)
    assert creep2.native.loop_info(0)["state"] == -1
    assert_bitexact(du, hu)


def test_inline_c_and_external_function_in_host_code(double):
    """`with xgrid.c()` text and an external C function run in the host stub,
    between launches, as in the reference's generated C; the stencil statement
    still runs on the device."""
    from xgrid_amd.types import Integer

    @xgrid.external(name="abs", includes=["stdlib.h"], typecheck_override=lambda args: Integer(4))
    def c_abs(x: int) -> int:
        ...

    @xgrid.kernel()
    def k(u: f1, a: int) -> int:
        b = c_abs(a)
        with xgrid.c():
            "b = b * 2 + 1;"
        u[0] = u[0] + 1.0
        return b

    d, h = grid_pair(xgrid, (33,), float, None, np.arange(33.0))
    ck = CpuKernel(k)
    for a in (-5, 7, 0):
        assert k(d, a) == ck(h, a) == abs(a) * 2 + 1
    assert_bitexact(d, h)
    assert k.native.lowered.graphs_ok is False        # host side effects: no graph replay


def test_bool_grid_statements(double):
    """A bool grid written from comparisons and read back as a condition."""
    b2 = xgrid.grid[bool, 2]

    @xgrid.kernel()
    def k(u: f2, flag: b2, lim: float) -> None:
        flag[0, 0] = u[0, 0] > lim and u[0, 1] <= lim
        u[0, 0] = u[0, 0] * 0.5 if flag[0, 0][0] else u[0, 0] + 1.0

    rng = np.random.default_rng(3)
    u0 = rng.random((20, 31))
    mask = np.zeros((20, 31), np.int32)
    mask[:, -1] = 1
    du, hu = grid_pair(xgrid, (20, 31), float, mask, u0)
    df, hf = grid_pair(xgrid, (20, 31), bool, mask)
    run_pair(k, (du, df, 0.5), (hu, hf, 0.5), steps=3)
    assert df.now.dtype == np.bool_
    assert_bitexact(du, hu)
    assert np.array_equal(df.now[:, :-1], hf.now[:, :-1])


def test_struct_element_grid_matches_oracle(double):
    """A grid of structures (xgrid/xgrid/__init__.py:10-18: a structured numpy dtype
    in C layout) on the device: fields of neighbouring cells read, a constructor
    stored, walls reset; bit for bit against the oracle and its numpy restatement."""
    from tests.helpers import Cplx, struct_kernel, struct_restated, struct_setup
    k = struct_kernel()
    shape = (70, 131)
    re, im, tag = struct_setup(shape, 45)
    d, h = grid_pair(xgrid, shape, Cplx, W.box_walls(shape))
    for g in (d, h):
        g.now["re"], g.now["im"], g.now["tag"] = re, im, tag
    run_pair(k, (d, 0.75), (h, 0.75), steps=4)
    assert d.now.dtype == h.now.dtype and d.now.dtype.itemsize == 24
    assert_bitexact(d, h)
    wre, wim, wtag = struct_restated(re, im, tag, 0.75, 4)
    assert np.array_equal(d.now["re"], wre) and np.array_equal(d.now["im"], wim)
    assert np.array_equal(d.now["tag"], wtag)


def _ulps(a, b):
    """Distance in units in the last place between two float64 arrays (finite values)."""
    ia, ib = a.view(np.int64).astype(object), b.view(np.int64).astype(object)
    ia = [x if x >= 0 else -(x & 0x7fffffffffffffff) for x in ia]
    ib = [x if x >= 0 else -(x & 0x7fffffffffffffff) for x in ib]
    return max(abs(x - y) for x, y in zip(ia, ib))


def test_math_externals_inside_stencil_statements(double):
    """@xgrid.external <math.h> functions called inside stencil statements run on
    the device (the reference's generated C calls them inside its loop nests,
    xgrid/lang/generator.py:204-212).  Correctly rounded ones (sqrt, fabs, floor,
    fmin, fmax, fmod, copysign, fma) give the oracle's bits; the transcendental
    ones come from the device's math library instead of glibc and are held to
    4 ulp of the oracle."""
    from xgrid_amd.types import Floating
    dbl = lambda args: Floating(8)      # noqa: E731

    @xgrid.external(name="sqrt", includes=["math.h"], typecheck_override=dbl)
    def c_sqrt(x: float) -> float: ...

    @xgrid.external(name="fabs", includes=["math.h"], typecheck_override=dbl)
    def c_fabs(x: float) -> float: ...

    @xgrid.external(name="floor", includes=["math.h"], typecheck_override=dbl)
    def c_floor(x: float) -> float: ...

    @xgrid.external(name="fmin", includes=["math.h"], typecheck_override=dbl)
    def c_fmin(x: float, y: float) -> float: ...

    @xgrid.external(name="fmod", includes=["math.h"], typecheck_override=dbl)
    def c_fmod(x: float, y: float) -> float: ...

    @xgrid.external(name="fma", includes=["math.h"], typecheck_override=dbl)
    def c_fma(x: float, y: float, z: float) -> float: ...

    @xgrid.external(name="sin", includes=["math.h"], typecheck_override=dbl)
    def c_sin(x: float) -> float: ...

    @xgrid.external(name="exp", includes=["math.h"], typecheck_override=dbl)
    def c_exp(x: float) -> float: ...

    @xgrid.external(name="atan2", includes=["math.h"], typecheck_override=dbl)
    def c_atan2(x: float, y: float) -> float: ...

    @xgrid.kernel()
    def exact(u: f2, v: f2) -> None:
        u[0, 0] = c_sqrt(c_fabs(v[0, 0] - v[1, 0])) + c_floor(v[0, 1] * 7.5) - c_fmin(v[0, -1], 0.5) + \
            c_fmod(v[-1, 0] * 10.0, 0.3) + c_fma(v[0, 0], v[0, 1], -v[1, 0])

    # one function per output grid, so that a last-place difference is not magnified
    # by cancellation in a sum
    @xgrid.kernel()
    def transcendental(s: f2, e: f2, t: f2, v: f2) -> None:
        s[0, 0] = c_sin(v[0, 1] * 3.0)
        e[0, 0] = c_exp(v[0, 0])
        t[0, 0] = c_atan2(v[1, 0] - 0.5, v[-1, 0] + 0.1)

    rng = np.random.default_rng(46)
    v0 = rng.random((40, 70)) * 2.0 - 0.5
    walls = W.box_walls((40, 70))
    du, hu = grid_pair(xgrid, (40, 70), float, walls)
    dv, hv = grid_pair(xgrid, (40, 70), float, None, v0)
    run_pair(exact, (du, dv), (hu, hv), steps=1)
    assert_bitexact(du, hu)
    outs = [grid_pair(xgrid, (40, 70), float, walls) for _ in range(3)]
    dv, hv = grid_pair(xgrid, (40, 70), float, None, v0)
    run_pair(transcendental, [d for d, _ in outs] + [dv], [h for _, h in outs] + [hv], steps=1)
    for d, h in outs:
        a, b = d.now[1:-1, 1:-1], h.now[1:-1, 1:-1]
        assert np.isfinite(a).all() and _ulps(a.ravel(), b.ravel()) <= 4
    for k in (exact, transcendental):
        assert k.native.lowered.graphs_ok                   # pure device functions: graphs stay on
