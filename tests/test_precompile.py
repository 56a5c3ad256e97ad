"""Background compiles of kernels (no GPU): started when a kernel is decorated,
isolated from the configuration and the operators' parses of the main thread."""
import os

import numpy as np
import pytest

import xgrid_amd as xgrid
from xgrid_amd.compilers import compile_device
from xgrid_amd.frontend import DSLSyntaxError
from xgrid_amd.types import Floating

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCALE = 2.0


def test_decorating_starts_a_background_compile_into_the_cache(tmp_path):
    xgrid.init(precision="double", cacheroot=str(tmp_path))

    @xgrid.kernel()
    def scale(u: xgrid.grid[float, 1], a: float) -> None:
        u[0] = u[0] * a + SCALE

    pre = scale._precompiled
    assert pre is not None and pre.wait(120) and pre.error is None and pre.seconds > 0
    # the main thread's own build text is what the background compiled: a cache hit
    from xgrid_amd.backend import lower_kernel
    low, _ = lower_kernel(scale)
    assert any(p.endswith(".hsaco") for p in os.listdir(tmp_path))
    assert os.path.exists(compile_device(low.device_src, str(tmp_path)))
    assert len([p for p in os.listdir(tmp_path) if p.endswith(".hsaco")]) == 1


def test_background_parse_does_not_decide_the_first_calls_types(tmp_path):
    """The reference parses at the first call, under the configuration then in
    force: a kernel decorated under precision="double" and first parsed after
    init(precision="float") has FP32 floats."""
    xgrid.init(precision="double", cacheroot=str(tmp_path))

    @xgrid.kernel()
    def k(u: xgrid.grid[float, 1]) -> None:
        u[0] = u[1] * 0.5

    k._precompiled.wait(120)
    xgrid.init(precision="float", cacheroot=str(tmp_path))
    assert k.ir.args[0][1].element == Floating(4)


def test_precompile_errors_surface_at_the_first_parse(tmp_path):
    xgrid.init(precision="double", cacheroot=str(tmp_path))

    @xgrid.kernel()
    def bad(a: int, b: float) -> float:
        return a + b

    assert bad._precompiled.wait(120) and isinstance(bad._precompiled.error, DSLSyntaxError)
    with pytest.raises(DSLSyntaxError):
        bad.ir


def test_precompile_can_be_switched_off_and_requested(tmp_path):
    xgrid.init(precision="double", cacheroot=str(tmp_path), precompile=False)

    @xgrid.kernel()
    def k(u: xgrid.grid[float, 1]) -> None:
        u[0] = u[0] + 1.0

    assert k._precompiled is None
    (h,) = xgrid.precompile(k, wait=True)
    assert h.done() and h.error is None and k._precompiled is h


def test_concurrent_stub_compiles_of_one_source_all_load(tmp_path):
    """Threads of one process compiling the same host stub at once (a background
    precompile and a batch build, say) each get a complete library: every
    compile writes files of its own and renames them into place."""
    import ctypes
    import threading
    from xgrid_amd.compilers import compile_host
    src = "".join(f"double xg_f{i}(double x) {{ return x * {i}.5 + {i}; }}\n" for i in range(400))
    errors, paths = [], []
    start = threading.Barrier(8)

    def work():
        try:
            start.wait()
            p = compile_host(src, str(tmp_path), ["gcc", "clang"])
            ctypes.CDLL(p)
            paths.append(p)
        except Exception as e:         # collected and reported below
            errors.append(repr(e))

    ts = [threading.Thread(target=work) for _ in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    assert len(set(paths)) == 1
    lib = ctypes.CDLL(paths[0])
    lib.xg_f3.restype = ctypes.c_double
    lib.xg_f3.argtypes = [ctypes.c_double]
    assert lib.xg_f3(2.0) == 2.0 * 3.5 + 3
    assert not [f for f in os.listdir(os.path.dirname(paths[0])) if ".tmp" in f or f.startswith("tmp")]
