"""Shared helpers: the same kernel on the device and on the oracle."""
import numpy as np

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
