"""Shared helpers: the same kernel on the device and on the oracle."""
import dataclasses

import numpy as np

import xgrid_amd as xgrid

from oracle.cpu_path import CpuKernel, HostGrid


def grid_pair(xg, shape, dtype, mask=None, init=None):
    """A device Grid and an oracle HostGrid with identical level 0 and mask."""
    d = xg.Grid(shape, dtype)
    h = HostGrid(shape, dtype)
    if init is not None:
        d.now[...] = init
        h.now[...] = init
    if mask is not None:
        d.boundary[...] = mask
        h.boundary[...] = mask
    return d, h


def run_pair(op, dev_args, host_args, steps=1, **oracle_kw):
    """Call the kernel `steps` times on both paths; return the oracle kernel."""
    from xgrid_amd.config import get_config
    ck = CpuKernel(op, overstep=get_config().overstep, **oracle_kw)
    for _ in range(steps):
        r1 = op(*dev_args)
        r2 = ck(*host_args)
        assert r1 == r2 or (r1 is None and r2 is None)
    return ck


def assert_bitexact(dev, host, levels=None):
    n = len(host._data)
    assert dev.time_levels == n
    for t in (range(n) if levels is None else levels):
        a = dev.level(t)
        b = host._data[t]
        assert a.dtype == b.dtype and a.shape == b.shape
        if not np.array_equal(a.view(np.uint8), b.view(np.uint8)):
            diff = np.argwhere(a != b)
            raise AssertionError(f"level {t}: {len(diff)} cells differ, first at {diff[:5].tolist()}: "
                                 f"dev={a[tuple(diff[0])]!r} host={b[tuple(diff[0])]!r}")


@dataclasses.dataclass
class Cplx:
    re: float
    im: float
    tag: int


def struct_kernel():
    """A grid of structures (xgrid/xgrid/__init__.py:10-18): fields read at
    neighbouring cells, a constructor stored at the cell, walls reset."""
    @xgrid.kernel()
    def rotate(z: xgrid.grid[Cplx, 2], w: float) -> None:
        z[0, 0] = Cplx(z[0, 0].re * w - z[0, 1].im, z[1, 0].im * w + z[0, -1].re, z[0, 0].tag + 1)
        with xgrid.boundary(1):
            z[0, 0] = Cplx(0.0, 0.0, z[0, 0].tag)
    return rotate


def struct_setup(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.random(shape), rng.random(shape), rng.integers(-5, 5, shape).astype(np.int32)


def struct_restated(re, im, tag, w, steps):
    """numpy restatement of `struct_kernel` over `steps` calls."""
    for _ in range(steps):
        nre, nim, ntag = np.zeros_like(re), np.zeros_like(im), tag.copy()
        nre[1:-1, 1:-1] = re[1:-1, 1:-1] * w - im[1:-1, 2:]
        nim[1:-1, 1:-1] = im[2:, 1:-1] * w + re[1:-1, :-2]
        ntag[1:-1, 1:-1] = tag[1:-1, 1:-1] + 1
        re, im, tag = nre, nim, ntag
    return re, im, tag
