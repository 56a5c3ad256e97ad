"""Frontend semantics (no GPU): the rules of xgrid/lang/parser.py and the
stencil analysis of xgrid/lang/generator.py:50-76."""
import dataclasses
import os

import numpy as np
import pytest

import xgrid_amd as xgrid
from xgrid_amd import ir
from xgrid_amd.analysis import StencilError, analyze
from xgrid_amd.frontend import DSLSyntaxError
from xgrid_amd.types import Boolean, Floating, GridType, Integer, Pointer, Structure, parse_annotation

f1 = xgrid.grid[float, 1]
f2 = xgrid.grid[float, 2]
CONST = 7


def test_annotations(double):
    assert parse_annotation(int) == Integer(4)
    assert parse_annotation(float) == Floating(8)
    assert parse_annotation(bool) == Boolean()
    assert parse_annotation(xgrid.grid[float, 3]) == GridType(Floating(8), 3)
    assert parse_annotation(xgrid.ptr[int]) == Pointer(Integer(4))
    xgrid.init(precision="float")
    assert parse_annotation(float) == Floating(4)


def test_time_offsets_default_load_minus1_store_0(double):
    @xgrid.kernel()
    def k(u: f1, v: f1) -> None:
        u[0] = v[1] + v[-1][0] + u[0][-2]

    f = k.ir
    st = f.body[0]
    assert isinstance(st.target, ir.Cell) and st.target.time == 0 and st.target.ctx == "store"
    cells = [n for n in ir.walk(st.value) if isinstance(n, ir.Cell)]
    assert [(c.grid.name, c.offsets, c.time) for c in cells] == [("v", (1,), -1), ("v", (-1,), 0), ("u", (0,), -2)]
    info = analyze(f)
    assert info.depth == 3                      # max |t| + 1
    assert info.stmts[0].refs == [("v", 1), ("v", 0), ("u", 2)]


def test_implicit_rule(double):
    @xgrid.kernel()
    def k(p: f2, b: f2) -> None:
        p[0, 0] = p[0, 1][0] + b[0, 0]          # self read at time 0, class 0: implicit
        p[0, 0] = p[0, 1] + b[0, 0][0]          # time -1 self read: explicit
        with xgrid.boundary(2):
            p[0, 0] = p[1, 0][0]                # class 2: in place (not implicit)
        p[0, 0] = p[0, 0][0] * 2.0              # only the cell itself: implicit but in place

    info = analyze(k.ir)
    assert [s.implicit for s in info.stmts] == [True, False, False, True]
    assert [s.mode for s in info.stmts] == ["pingpong", "inplace", "inplace", "inplace"]
    assert [s.klass for s in info.stmts] == [0, 0, 2, 0]
    assert info.stmts[2].hazard_offsets == [(1, 0)]


def test_store_offset_implicit_is_twophase(double):
    @xgrid.kernel()
    def k(p: f1) -> None:
        p[1] = p[-1][0]

    assert analyze(k.ir).stmts[0].mode == "twophase"


def test_chained_compare_compares_left_with_each(double):
    @xgrid.kernel()
    def k(a: int, b: int, c: int) -> bool:
        return a < b < c

    ret = k.ir.body[0].value
    assert ret.op == "&&"
    assert ret.left.left.var.name == "a" and ret.right.left.var.name == "a"   # a<b && a<c


def test_pow_typing(double):
    @xgrid.kernel()
    def k(x: float, i: int) -> float:
        y = x ** 2.0
        z = i ** 2
        return y

    body = k.ir.body
    assert body[0].value.type == Floating(8)
    assert body[1].value.type == Floating(8)        # int ** int -> configured float
    xgrid.init(precision="float")

    @xgrid.kernel()
    def k2(x: float) -> float:
        return x ** 2.0

    assert k2.ir.body[0].value.type == Floating(4)


def test_globals_are_folded(double):
    @xgrid.kernel()
    def k(a: int) -> int:
        return a + CONST

    c = k.ir.body[0].value.right
    assert isinstance(c, ir.Const) and c.value == 7


@pytest.mark.parametrize("src_case", ["mixed", "unknown", "load_outside", "two_arg_boundary", "bad_offset",
                                      "grid_value", "pow_aug"])
def test_syntax_errors(double, src_case):
    if src_case == "mixed":
        @xgrid.kernel()
        def k(a: int, b: float) -> float:
            return a + b
    elif src_case == "unknown":
        @xgrid.kernel()
        def k(a: int) -> int:
            return a + not_defined_anywhere  # noqa: F821
    elif src_case == "load_outside":
        @xgrid.kernel()
        def k(u: f1) -> float:
            return u[0]
    elif src_case == "two_arg_boundary":
        @xgrid.kernel()
        def k(u: f1) -> None:
            with xgrid.boundary(u, 1):
                u[0] = 1.0
    elif src_case == "bad_offset":
        @xgrid.kernel()
        def k(u: f1, i: int) -> None:
            u[i] = 1.0
    elif src_case == "grid_value":
        @xgrid.kernel()
        def k(u: f1) -> None:
            x = u
    else:
        @xgrid.kernel()
        def k(x: float) -> float:
            x **= 2.0
            return x
    with pytest.raises((DSLSyntaxError, StencilError)):
        analyze(k.ir)


def test_struct_method_and_constructor(double):
    @dataclasses.dataclass
    class V2:
        x: float
        y: float

        @xgrid.function(method=True)
        def norm2(self) -> float:
            return self.x * self.x + self.y * self.y

    @xgrid.kernel()
    def k(a: V2) -> float:
        b = V2(a.y, a.x)
        return b.norm2()

    body = k.ir.body
    assert isinstance(body[0].value, ir.Call) and isinstance(body[0].value.ctor, Structure)
    assert body[1].value.callee.c_name == "V2_norm2"


def test_readme_elementwise_semantics_on_host_grids(double):
    """README.md:26-57: every Grid argument is ticked, so after one call the
    inputs' `now` is a fresh zero level and `result` holds a*b."""
    from oracle.cpu_path import CpuKernel, HostGrid

    @xgrid.kernel()
    def elementwise_mul(result: f1, a: f1, b: f1) -> None:
        result[0] = a[0] * b[0]

    rng = np.random.default_rng(0)
    r, a, b = HostGrid((100,), float), HostGrid((100,), float), HostGrid((100,), float)
    av, bv = rng.random(100), rng.random(100)
    a.now[:] = av
    b.now[:] = bv
    ck = CpuKernel(elementwise_mul)
    ck(r, a, b)
    assert np.array_equal(r.now, av * bv)
    assert not a.now.any() and not b.now.any()
    ck(r, a, b)                                  # second call reads the zero levels
    assert not r.now.any()


def test_fill_default_time_truncates_and_raises(double):
    """xgrid/xgrid/__init__.py:74-80: fill(data) with time=0 empties the level list."""
    g = xgrid.Grid((4,), float)
    with pytest.raises(IndexError):
        g.fill(np.ones(4))
    g2 = xgrid.Grid((4,), float)
    g2.fill(np.full(4, 3.0), time=-1)
    assert (g2.now == 3.0).all()


def _stmt_struct(src: str, sid: int) -> str:
    i = src.index(f"struct xg_s{sid}_t ")
    return src[i:src.index("extern", i)]


def test_staged_3d_plan(double, monkeypatch):
    """3-D statements whose loads include column or row neighbours are staged:
    plane registers, DPP column shifts, an LDS copy of the tile's rows (only for
    row neighbours), 16 bytes of a row per lane (2 columns of FP64, 4 of FP32);
    diagonal loads stay global loads; 2-D statements, XGRID_STAGE3=0 and, for
    4-byte arrays, XGRID_STAGE3_F32=0 keep plain loads."""
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    f3 = xgrid.grid[float, 3]

    low = lower_kernel(W.make_heat_3d())[0]
    src = low.device_src
    s0 = _stmt_struct(src, 0)
    assert "ST3 = true" in s0 and "ST3_LDS = 24576" in s0           # 4 planes x 6 rows x 128 columns x 8 B
    assert "CPL = 2" in s0 and low.stmts[0].tile3_cols == 128        # two columns per lane, 128-column tiles
    assert "xg::ld<" not in s0[s0.index("value_st3"):]              # every load of the 7-point update staged
    assert "amdgpu_waves_per_eu(6)))" in src[src.index("xg_k0(") - 80:src.index("xg_k0(")]
    assert "CPL = 1" in _stmt_struct(src, 1) and low.stmts[1].tile3_cols == 64     # the walls: plain loads
    monkeypatch.setenv("XGRID_STAGE3_COLS", "1")                      # one column per lane, as before
    low1 = lower_kernel(W.make_heat_3d())[0]
    assert "ST3_LDS = 12288" in _stmt_struct(low1.device_src, 0) and low1.stmts[0].tile3_cols == 64
    monkeypatch.delenv("XGRID_STAGE3_COLS")

    @xgrid.kernel()
    def diag(u: f3, v: f3) -> None:
        u[0, 0, 0] = v[0, 1, 1] + v[1, -1, 1]
        with xgrid.boundary(1):
            u[0, 0, 0] = v[0, 0, 1] * 2.0 - v[1, 0, 0]

    src = lower_kernel(diag)[0].device_src
    assert "ST3 = false" in _stmt_struct(src, 0)
    s1 = _stmt_struct(src, 1)
    assert "ST3 = true" in s1 and "ST3_LDS = 16;" in s1 and "__syncthreads" not in s1

    assert "ST3 = false" in _stmt_struct(lower_kernel(W.make_diffusion_2d())[0].device_src, 0)
    monkeypatch.setenv("XGRID_STAGE3", "0")
    assert "ST3 = true" not in lower_kernel(W.make_heat_3d())[0].device_src
    monkeypatch.delenv("XGRID_STAGE3")
    xgrid.init(precision="float")                                     # 4-byte arrays: 4 columns per lane
    low = lower_kernel(W.make_heat_3d())[0]
    s0 = _stmt_struct(low.device_src, 0)
    assert "CPL = 4" in s0 and "ST3_LDS = 24576" in s0 and low.stmts[0].tile3_cols == 256
    assert "xg::vecn<float, 4>" in s0
    monkeypatch.setenv("XGRID_STAGE3_F32", "0")
    assert "ST3 = true" not in lower_kernel(W.make_heat_3d())[0].device_src
    monkeypatch.delenv("XGRID_STAGE3_F32")


def _dpp_in_masked_regions(asm: str) -> dict:
    """Per kernel of an AMD assembly listing: (DPP moves, DPP moves executed
    inside an exec-masked region).  The kernel is split into basic blocks
    (labels, and the ends of branches).  The open exec-masked regions are a stack
    of the registers their s_and_saveexec saved the mask into; an
    `s_or_b64 exec, exec, sN` closes every region opened since sN was saved (the
    compiler closes an inner region and the one around it with one restore of
    the outer mask).  The stack flows along the branch and fall-through edges
    from empty at entry, the deeper one winning where paths join.  Following
    edges rather than layout order matters: the compiler lays a region's blocks
    out before or after the block that opens it."""
    import re
    kernels, cur = {}, None
    for line in asm.splitlines():
        m = re.match(r"^([A-Za-z_]\w*):", line)
        if m:
            cur = m.group(1)
            kernels[cur] = []
            continue
        if cur is None:
            continue
        t = line.split(";")[0].strip()
        if t.startswith(".Lfunc_end"):
            cur = None
            continue
        if t:
            kernels[cur].append(t)

    def step(stack, t):
        if t.startswith("s_and_saveexec"):
            return stack + (t.split()[1].rstrip(","),)
        if t.startswith("s_or_b64 exec, exec,"):
            reg = t.split(",")[-1].strip()
            return stack[:stack.index(reg)] if reg in stack else stack[:-1]
        return stack

    out = {}
    for name, lines in kernels.items():
        blocks, labels, b = [], {}, []
        for t in lines:
            lm = re.match(r"^(\.LBB\w+):", t)
            if lm:
                if b:
                    blocks.append(b)
                b = []
                labels[lm.group(1)] = len(blocks)
                continue
            if t.startswith("."):
                continue
            b.append(t)
            if t.startswith(("s_branch", "s_cbranch", "s_endpgm", "s_setpc")):
                blocks.append(b)
                b = []
        if b:
            blocks.append(b)
        if not blocks:
            continue
        succ = []
        for i, blk in enumerate(blocks):
            last = blk[-1] if blk else ""
            tgt = last.split()[1] if last.startswith(("s_branch", "s_cbranch")) and len(last.split()) > 1 else None
            nxt = [i + 1] if i + 1 < len(blocks) else []
            if last.startswith("s_branch"):
                succ.append([labels[tgt]] if tgt in labels else [])
            elif last.startswith("s_cbranch"):
                succ.append(([labels[tgt]] if tgt in labels else []) + nxt)
            elif last.startswith(("s_endpgm", "s_setpc")):
                succ.append([])
            else:
                succ.append(nxt)
        stack_in = [None] * len(blocks)
        stack_in[0] = ()
        work = [0]
        while work:
            i = work.pop()
            st = stack_in[i]
            for t in blocks[i]:
                st = step(st, t)
            for j in succ[i]:
                if stack_in[j] is None or (len(st) > len(stack_in[j]) and len(st) <= 64):
                    stack_in[j] = st
                    work.append(j)
        n = bad = 0
        for i, blk in enumerate(blocks):
            st = stack_in[i] if stack_in[i] is not None else ()
            for t in blk:
                st = step(st, t)
                if "_dpp" in t:
                    n += 1
                    bad += len(st) > 0
        if n:
            out[name] = (n, bad)
    return out


def _hipcc():
    import shutil
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _need_hipcc():
    if not os.path.exists(_hipcc()):
        pytest.skip("hipcc not available")


def _assembly(text):
    """gfx950 assembly of a generated device source, built with the run-time options."""
    import subprocess
    import tempfile
    from xgrid_amd import compilers
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "k.hip")
        with open(p, "w") as f:
            f.write(text)
        r = subprocess.run([_hipcc(), "-x", "hip", "--cuda-device-only", "-S", *compilers.device_options(),
                            "-I", compilers.CSRC, p, "-o", os.path.join(d, "k.s")],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        return open(os.path.join(d, "k.s")).read()


def _kernel_body(asm, name):
    import re
    m = re.search(rf"^{name}:[^\n]*\n(.*?)^\s*s_endpgm", asm, re.S | re.M)
    assert m, name
    return m.group(1)


def test_nontemporal_stores_only_in_3d_box_kernels(double):
    """An in-place statement over a grid larger than the last-level cache is
    launched with XG_F_STREAM, and its 3-D box tiles then store with the
    non-temporal hint.  The two store forms are separate template instances:
    written as a run-time branch, the compiler merges them into one plain store
    and the hint is lost.  So a 3-D kernel's code holds both forms (16-byte
    staged stores included), and a 2-D kernel's only plain stores."""
    import re
    _need_hipcc()
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    f2 = xgrid.grid[float, 2]

    @xgrid.kernel()
    def smooth(u: f2, v: f2) -> None:
        u[0, 0] = (v[0, 1] + v[0, -1] + v[1, 0] + v[-1, 0]) * 0.25

    def stores(body):
        forms = re.findall(r"^\s*(global_store_\w+)\s.*?(\bnt)?\s*$", body, re.M)
        return {(op, bool(nt)) for op, nt in forms}

    heat = stores(_kernel_body(_assembly(lower_kernel(W.make_heat_3d())[0].device_src), "xg_k0"))
    assert ("global_store_dwordx4", True) in heat and ("global_store_dwordx4", False) in heat, heat
    assert ("global_store_dwordx2", True) in heat and ("global_store_dwordx2", False) in heat, heat
    flat = stores(_kernel_body(_assembly(lower_kernel(smooth)[0].device_src), "xg_k0"))
    assert flat and not any(nt for _, nt in flat), flat


def _lds_loads_unwaited(lines):
    """Indexes of barriers reached (in code order) after a load into LDS with
    no wait for the vector-memory counter to reach 0 in between."""
    bad, pending = [], False
    for i, ln in enumerate(lines):
        if ln.startswith("global_load_lds_"):
            pending = True
        elif ln.startswith("s_waitcnt") and "vmcnt(0)" in ln:
            pending = False
        elif ln.startswith("s_barrier") and pending:
            bad.append(i)
    return bad


def test_staged_halo_rows_load_into_lds_before_the_barrier(double):
    """The staged 3-D tiles load their aligned halo rows straight into LDS.
    Those loads are counted by the vector-memory counter, not the LDS one, so
    after them the wave waits for that counter to reach 0 before the
    workgroup's next barrier; only then may other waves read the rows.  The
    check is pinned by the same code with those waits taken out."""
    _need_hipcc()
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    body = _kernel_body(_assembly(lower_kernel(W.make_heat_3d())[0].device_src), "xg_k0")
    lines = [ln.strip() for ln in body.splitlines()
             if ln.strip() and not ln.strip().startswith((";", ".")) and not ln.rstrip().endswith(":")]
    assert sum(ln.startswith("global_load_lds_dwordx4") for ln in lines) >= 2
    assert not _lds_loads_unwaited(lines)
    planted = [ln for ln in lines if not (ln.startswith("s_waitcnt") and "vmcnt(0)" in ln)]
    assert _lds_loads_unwaited(planted)


def test_staged_lane_shifts_run_on_every_lane(double):
    """A lane shift reads another lane's register; run by a subset of lanes it
    reads 0 from the disabled ones.  In the generated source every shift is a
    declaration of its own, never an operand of a conditional (which would be
    evaluated only where the condition selects it).  In the compiled gfx950 code
    (same options as at run time) no DPP move, in plain or batched kernels, is
    reached inside an exec-masked region.  The check itself is pinned by a
    planted shift inside the edge lane's conditional, which it must flag."""
    import re
    _need_hipcc()
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    f3 = xgrid.grid[float, 3]

    @xgrid.kernel()
    def mixed(u: f3, v: f3) -> None:
        u[0, 0, 0] = v[0, 0, 1] + v[-1, 0, 0] * 0.5 + v[0, 1, 0] * 0.25
        with xgrid.boundary(4):
            u[0, 0, 0] = u[0, 0, 1][0] + v[0, -1, 1] - v[0, 0, -1]

    assembly = _assembly

    for op in (W.make_heat_3d(), mixed):
        low = lower_kernel(op)[0]
        src = low.device_src
        shifts = [ln.strip() for ln in src.splitlines() if "lane_shift<" in ln]
        assert shifts, "no staged statement"
        assert all(re.fullmatch(r"const \w+ [ab] = xg::lane_shift<-?1>\(c[fl]\);", ln) for ln in shifts), shifts
        found = _dpp_in_masked_regions(assembly(src))
        found_b = _dpp_in_masked_regions(assembly(low.batch_src))
        assert found and found_b, "no staged kernel with lane shifts"
        assert any(k.endswith("_b") for k in found_b)
        assert all(bad == 0 for _, bad in list(found.values()) + list(found_b.values())), (found, found_b)
        if op is not mixed:
            continue
        # the planted mistake: a shift evaluated only where the conditional selects it
        planted = re.sub(r"const (\w+) a = (xg::lane_shift<1>\(cf\));\n(\s*)(st\.xp\d+) = (st\.lane == 63 \? [^:]+ : )a;",
                         r"\3\4 = \5\2;", src, count=1)
        assert planted != src
        bad = _dpp_in_masked_regions(assembly(planted))
        assert any(b > 0 for _, b in bad.values()), bad


def test_statement_pairs_marked(double):
    """Consecutive in-place statements of one group storing to different grids
    get a pair kernel when the first does not load the second's store grid at
    time 0 and the second loads the first's at time 0 only at its own cell
    (then forwarded)."""
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    low = lower_kernel(W.make_cavity())[0]
    pairs = {m.stmt.sid: (m.pair, m.pair_fwd) for m in low.stmts if m.pair}
    assert pairs.get(11) == ("xg_f11", False)              # u and v updates
    assert pairs.get(0) == ("xg_f0", True)                 # b, then the first sweep reading b at its cell
    assert low.stmts[1].fwd_from == "b"
    assert "run_pair<xg_s11, xg_s12, XG_OVERSTEP, false>" in low.device_src
    assert "run_pair<xg_s0, xg_s1, XG_OVERSTEP, true>" in low.device_src
    f2 = xgrid.grid[float, 2]

    @xgrid.kernel()
    def chain(u: f2, v: f2) -> None:
        u[0, 0] = v[0, 1]
        v[0, 0] = u[0, 1][0] * 2.0                         # reads u at time 0 at another cell: no pair

    assert not any(m.pair for m in lower_kernel(chain)[0].stmts)
