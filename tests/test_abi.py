"""The C-ABI library (no GPU): it loads, exports every symbol include/xgrid_hip.h
declares, its structs have one layout in C and in ctypes, generated kernels
compile for gfx950 with hipRTC, and the device path fails loudly without a GPU."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from xgrid_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xgrid_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*\b(xg_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), f"missing export {n}"
    assert set(names) == set(_lib.EXPORTED), set(names) ^ set(_lib.EXPORTED)


def test_struct_layouts_match_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "xgrid_hip.h"
#include "xgrid_abi.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(xg_stmt_desc), offsetof(xg_stmt_desc, param_bytes),
         offsetof(xg_stmt_desc, fn_main), sizeof(xg_group_desc), sizeof(xg_api), sizeof(xg_dom),
         sizeof(xg_gview), sizeof(xg_launch_hdr), offsetof(xg_dom, count), offsetof(xg_stmt_desc, pair_fwd),
         offsetof(xg_stmt_desc, fn_pair), sizeof(xg_loop_desc), offsetof(xg_loop_desc, alt_max_cells));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", c, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "xgrid_amd", "csrc"),
                        "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    exp = [ctypes.sizeof(_lib.StmtDesc), _lib.StmtDesc.param_bytes.offset, _lib.StmtDesc.fn_main.offset,
           ctypes.sizeof(_lib.GroupDesc), ctypes.sizeof(_lib.Api), 72, 24, 120, 32,
           _lib.StmtDesc.pair_fwd.offset, _lib.StmtDesc.fn_pair.offset,
           ctypes.sizeof(_lib.LoopDesc), _lib.LoopDesc.alt_max_cells.offset]
    assert got == exp


def test_device_path_fails_loudly_without_gpu(double):
    if _lib.load().xg_initialized():
        pytest.skip("a device is initialised")
    lib = _lib.load()
    n = ctypes.c_int()
    if lib.xg_init(0) == 0:
        pytest.skip("a GPU is present")
    from xgrid_amd import workloads as W
    k = W.make_diffusion_1d()
    g = __import__("xgrid_amd").Grid((16,), float)
    with pytest.raises(_lib.BackendError):
        k(g, 0.1, 0.1, 0.1)


def test_calls_before_init_are_refused():
    lib = _lib.load()
    if lib.xg_initialized():
        pytest.skip("device initialised")
    h = ctypes.c_void_p()
    shape = (ctypes.c_int32 * 3)(4, 1, 1)
    assert lib.xg_grid_create(1, shape, 8, ctypes.byref(h)) == -5
    assert "xg_init" in _lib.last_error()


@pytest.mark.parametrize("prec", ["double", "float"])
def test_generated_kernels_compile_for_gfx950(prec, request, tmp_path):
    request.getfixturevalue(prec if prec == "double" else "single")
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    from xgrid_amd.compilers import compile_device, compile_host
    from xgrid_amd.backend import packed_args_type
    from xgrid_amd.compilers import load_host
    for make in (W.make_diffusion_1d, W.make_cavity, W.make_heat_3d, W.make_convection_2d):
        op = make()
        low, _ = lower_kernel(op)
        assert os.path.getsize(compile_device(low.device_src, str(tmp_path))) > 1000
        assert os.path.getsize(compile_device(low.batch_src, str(tmp_path))) > 1000
        stub = compile_host(low.stub_src, str(tmp_path), ["gcc"])
        assert os.path.exists(stub)
        # batches call the stub with every argument packed in one struct: ctypes must agree with C
        lib = load_host(stub)
        lib.xg_args_size.restype = ctypes.c_size_t
        assert lib.xg_args_size() == ctypes.sizeof(packed_args_type(list(op.ir.args)))
        assert hasattr(lib, "xg_entry_packed")


def test_generated_device_source_has_batched_variants(double):
    """Every launch has a batched variant, in the second source (compiled when a
    batch is made); the plain source defines none of them, and both sources hold
    every plain launch's statement structs."""
    import re
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    low, _ = lower_kernel(W.make_cavity())
    src, bsrc = low.device_src, low.batch_src
    for m in low.stmts:
        assert f" {m.kernel}(" in src and f"{m.kernel}_b(" in bsrc and f"{m.kernel}_b(" not in src
        if m.kernel_pp:
            assert f"{m.kernel_pp}b(" in bsrc
        assert f"struct xg_s{m.stmt.sid}_t " in src and f"struct xg_s{m.stmt.sid}_t " in bsrc
    for g in low.groups:
        assert f"{g.kernel}_b(" in bsrc
    for L in low.loops:
        assert f"{L.kernel}_b(" in bsrc and f"{L.kernel}_chkb(" in bsrc
    defined = lambda s: set(re.findall(r'extern "C" __global__[^\n]*?\b(xg_\w+)\(', s))   # noqa: E731
    assert not defined(src) & defined(bsrc)
    assert all(re.search(r"_(b|chkb|ppb)$", n) for n in defined(bsrc))


def test_rtc_compile_error_is_reported(tmp_path):
    from xgrid_amd.compilers import CompileError, compile_device
    with pytest.raises(CompileError) as e:
        compile_device('extern "C" __global__ void k() { undeclared_thing(); }', str(tmp_path))
    assert "undeclared_thing" in str(e.value)


def test_host_compile_error_is_reported(tmp_path):
    from xgrid_amd.compilers import CompileError, compile_host
    with pytest.raises(CompileError):
        compile_host("innt main() {}", str(tmp_path), ["gcc"])
