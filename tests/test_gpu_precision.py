"""FP64 reductions on the device (difference norms, Poisson residual) and FP32
storage against FP64: closed forms, numpy in FP64, and the oracle."""
import numpy as np
import pytest

from tests.helpers import assert_bitexact, grid_pair, run_pair

pytestmark = pytest.mark.gpu

REL = 1e-12   # FP64 sums in another order: relative tolerance of an L2 norm


def _grid(xg, arr, mask=None):
    g = xg.Grid(arr.shape, float)
    g.now[...] = arr
    if mask is not None:
        g.boundary[...] = mask
    return g


def test_diff_norm_closed_forms(double):
    from xgrid_amd import norms
    from xgrid_amd.workloads import box_walls
    n = 257
    mask = box_walls((n, n))
    a = _grid(double, np.full((n, n), 3.0), mask)
    r = norms.diff_norm(a, a)
    assert r["l2"] == 0.0 and r["linf"] == 0.0 and r["cells"] == (n - 2) ** 2
    r = norms.diff_norm(a)                          # constant 3 over (n-2)^2 cells
    assert r["linf"] == 3.0
    assert abs(r["l2"] - 3.0 * (n - 2)) <= REL * r["l2"]
    r = norms.diff_norm(a, klass=1)                 # the walls: 4(n-1) cells
    assert r["cells"] == 4 * (n - 1) and r["linf"] == 3.0
    assert abs(r["l2"] - 3.0 * np.sqrt(4 * (n - 1))) <= REL * r["l2"]


@pytest.mark.parametrize("shape", [(300, 517), (37, 41, 45), (5000,)])
def test_diff_norm_matches_numpy_fp32_vs_fp64(shape):
    import os
    import xgrid_amd
    from xgrid_amd import norms
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape)
    mask = np.zeros(shape, np.int32)
    mask[rng.random(shape) < 0.05] = 2              # sparse class: a device list
    mask[tuple(slice(0, s // 2) for s in shape)][rng.random(tuple(s // 2 for s in shape)) < 0.6] = 1  # dense in a box
    xgrid_amd.init(precision="float", cacheroot=os.path.join(root, ".xgrid"))
    a32 = xgrid_amd.Grid(shape, float)
    a32.now[...] = x.astype(np.float32)
    a32.boundary[...] = mask
    xgrid_amd.init(precision="double", cacheroot=os.path.join(root, ".xgrid"))
    b64 = xgrid_amd.Grid(shape, float)
    b64.now[...] = x
    for k in (0, 1, 2):
        d = x.astype(np.float32).astype(np.float64) - x
        sel = mask == k
        ref_l2 = np.sqrt(np.sum(d[sel] ** 2))
        r = norms.diff_norm(a32, b64, klass=k)
        assert r["cells"] == sel.sum()
        assert abs(r["l2"] - ref_l2) <= REL * ref_l2
        assert r["linf"] == np.abs(d[sel]).max()
    kinds = {k: a32.class_info(k)["kind"] for k in (0, 1, 2)}
    assert kinds[2] == "list"


def test_diff_norm_propagates_nan(double):
    from xgrid_amd import norms
    x = np.ones((64, 64))
    x[10, 20] = np.nan
    a = _grid(double, x)
    r = norms.diff_norm(a)
    assert np.isnan(r["l2"]) and np.isnan(r["linf"])


def test_poisson_residual_closed_form_2d_3d(double):
    from xgrid_amd import norms
    from xgrid_amd.workloads import box_walls
    for shape in [(65, 130), (20, 24, 28)]:
        idx = np.indices(shape).astype(np.float64)
        p = sum(i * i for i in idx)                  # second difference of i^2 is exactly 2 per axis
        D = len(shape)
        mask = box_walls(shape)
        pg = _grid(double, p, mask)
        b = _grid(double, np.full(shape, 2.0 * D - 1.0), mask)
        r = norms.poisson_residual(pg, b, [1.0] * D)  # residual 2D - (2D - 1) = 1 at every interior cell
        cells = np.prod([s - 2 for s in shape])
        assert r["linf"] == 1.0
        assert abs(r["l2"] - np.sqrt(cells)) <= REL * np.sqrt(cells)


def test_poisson_residual_matches_numpy_and_rejects_edge_classes(double):
    from xgrid_amd import _lib, norms
    from xgrid_amd.workloads import cavity_masks
    n = 129
    rng = np.random.default_rng(11)
    p = rng.standard_normal((n, n))
    b = rng.standard_normal((n, n))
    masks = cavity_masks(n)
    pg = _grid(double, p, masks["p"])
    bg = _grid(double, b, masks["b"])
    cy, cx = 1.0 / 0.03 ** 2, 1.0 / 0.02 ** 2
    lap = cy * ((p[2:, 1:-1] - 2.0 * p[1:-1, 1:-1]) + p[:-2, 1:-1]) + \
        cx * ((p[1:-1, 2:] - 2.0 * p[1:-1, 1:-1]) + p[1:-1, :-2])
    res = lap - b[1:-1, 1:-1]
    r = norms.poisson_residual(pg, bg, [cy, cx])
    assert abs(r["l2"] - np.sqrt(np.sum(res ** 2))) <= REL * r["l2"]
    assert r["linf"] == np.abs(res).max()           # same association per cell: exact
    with pytest.raises(_lib.BackendError, match="edge"):
        norms.poisson_residual(pg, bg, [cy, cx], klass=1)     # the right wall: the stencil would leave the grid


def test_cavity_fp32_matches_oracle(single):
    from xgrid_amd import workloads as W
    n, steps = 65, 6
    k = W.make_cavity()
    cfg = W.cavity_config(n, dt=2e-6)
    masks = W.cavity_masks(n)
    pairs = {c: grid_pair(single, (n, n), float, masks[c]) for c in "bpuv"}
    run_pair(k, [pairs[c][0] for c in "bpuv"] + [cfg], [pairs[c][1] for c in "bpuv"] + [cfg], steps=steps)
    for c in "bpuv":
        assert pairs[c][0].now.dtype == np.float32
        assert_bitexact(*pairs[c], levels=[0])


def test_cavity_fp32_storage_error_against_fp64():
    """C3 at a test size: the FP32-stored cavity against its FP64 twin, L2/Linf
    measured on the device; the FP64 residual of each field agrees with numpy."""
    from xgrid_amd import workloads as W
    from xgrid_amd.precision import CavityTwin
    twin = CavityTwin(97, dt=2e-6)
    twin.run(10)
    rep = twin.report()
    for c in "puv":
        e = rep["error"][c]
        assert e["rel_l2"] < 1e-4, (c, e)
    # residual of the FP32 field, in FP64, against numpy on the host copy
    p32 = twin.f32["p"].now.astype(np.float64)
    b32 = twin.f32["b"].now.astype(np.float64)
    cfg = twin.cfg
    cy, cx = 1.0 / cfg.dy ** 2, 1.0 / cfg.dx ** 2
    lap = cy * ((p32[2:, 1:-1] - 2.0 * p32[1:-1, 1:-1]) + p32[:-2, 1:-1]) + \
        cx * ((p32[1:-1, 2:] - 2.0 * p32[1:-1, 1:-1]) + p32[1:-1, :-2])
    res = lap - b32[1:-1, 1:-1]
    ref = np.sqrt(np.sum(res ** 2))
    assert abs(rep["residual"]["f32"]["l2"] - ref) <= REL * ref
    assert rep["residual"]["f32"]["linf"] == np.abs(res).max()
