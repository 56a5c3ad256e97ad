"""Lowering of a parsed kernel to (1) HIP device source and (2) a C host stub.

Device source (compiled by hipRTC for gfx950): one struct per stencil
statement exposing ``value(refs, params, i0, i1, i2)`` - the cell expression,
with every grid access turned into an overstep-aware load of one of the
statement's (grid, level) views - plus ``extern "C"`` kernels that plug it into
the traversal templates of ``csrc/xgrid_device.h``.

Host stub (compiled by the system C compiler): the kernel's scalar code and
control flow, as the reference runs it (``xgrid/lang/generator.py:251-283,
362-374``), with every stencil statement replaced by one ``stmt()`` call that
hands the statement's captured scalars to the runtime.

Arithmetic keeps the reference's C semantics: the same operator text, fully
parenthesised (``generator.py:377-422``), the same constant spelling
(``repr(v).lower()``, ``generator.py:393-394``) and the same ``pow``/``powf``
choice (``generator.py:377-383``).  gcc folds ``pow(x, 2.0)`` to ``x*x``,
``pow(x, 1.0)`` to ``x`` and ``pow(x, -1.0)`` to ``1/x`` at -O1 and above; the
device code applies exactly those three folds and calls the device ``pow``
otherwise.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

from . import ir
from .analysis import KernelInfo, StencilStmt, captured_value_type, structures_in
from .types import (Boolean, Floating, GridType, Ignore, Integer, Pointer, Structure, Value, Void,
                    element_nbytes)


class EmitError(RuntimeError):
    pass


OVERSTEP_CODE = {"none": 0, "limit": 1, "wrap": 2}


def ctype_name(t) -> str:
    if isinstance(t, Void):
        return "void"
    if isinstance(t, (Boolean, Integer, Floating, Structure)):
        return t.cname
    if isinstance(t, Pointer):
        return ctype_name(t.element) + "*"
    raise EmitError(f"type {t!r} has no C spelling")


def device_type_name(t) -> str:
    if isinstance(t, Structure):
        return t.name
    return ctype_name(t)


def const_text(c: ir.Const) -> str:
    return repr(c.value).lower()


def _pow_fold(right) -> str | None:
    """Exponents gcc folds at -O1+: 2 -> x*x, 1 -> x, -1 -> 1/x."""
    if isinstance(right, ir.Const) and type(right.value) in (int, float):
        v = float(right.value)
        if v == 2.0:
            return "sq"
        if v == 1.0:
            return "id"
        if v == -1.0:
            return "inv"
    if isinstance(right, ir.Un) and right.op == "-" and isinstance(right.operand, ir.Const) \
            and type(right.operand.value) in (int, float) and float(right.operand.value) == 1.0:
        return "inv"
    return None


@dataclass
class StmtMeta:
    stmt: StencilStmt
    refs: list                    # [(grid name, level)]; level -1 = shape only
    params: list                  # [ir.Var]
    kernel: str
    kernel_pp: str | None
    phase2: str
    group: int = -1
    group_pos: int = 0
    pair: str | None = None       # kernel running this statement and the next in one launch
    pair_fwd: bool = False        # ... with the next one's loads of this store (own cell, time 0) forwarded
    fwd_from: str | None = None   # (second of a forwarding pair) the grid whose time-0 loads are forwarded
    hoisted: list = field(default_factory=list)   # [Hoist]: launch-uniform subexpressions computed by the host
    tile3_cols: int = 64          # columns of a 3-D tile of this statement's run_main kernels


@dataclass
class Hoist:
    """A launch-uniform subexpression of a statement (no grid load): the host stub
    evaluates it once per call and passes the value in the parameter block, field
    `name`, so the device does not evaluate it for every cell.  The host's C
    evaluates the same operations on the same values in the same order (IEEE,
    no contraction), so the value is the one the device would have computed."""
    nodes: list                   # the IR nodes with this text
    ctype: str                    # "double" | "float"
    name: str                     # parameter field
    host: str                     # C text the stub evaluates


@dataclass
class GroupMeta:
    first: int
    count: int
    kernel: str


@dataclass
class LoopMeta:
    """A `for` loop that may run temporally blocked (several iterations per launch)."""
    lid: int
    node: object                  # ir.For
    stmts: list                   # StmtMeta of the body, in order
    grid: str                     # the one grid the body stores to
    core: tuple                   # (rows, columns) of a workgroup's core tile
    halo: int
    tmax: int
    kernel: str
    threads: int = 512
    reach: int = 1                # largest |offset| of the body's reads of the loop grid
    variant: str = "lds"          # "lds": region in LDS; "reg": one column per lane in registers
    alt: object = None            # LoopMeta of the same loop with the tile used on small grids
    alt_max_cells: int = 0        # ... on grids of at most this many cells
    batched: bool = True          # emit the batched variants (the small-grid tile has none)


@dataclass
class Lowered:
    device_src: str               # the kernels plain calls launch
    stub_src: str
    stmts: list
    groups: list
    grid_args: list               # [(arg index, name, GridType)]
    scalar_args: list             # [(arg index, name, type)]
    graphs_ok: bool
    loops: list = field(default_factory=list)
    block_shape: tuple | None = None   # shape of the blocked loops (None: the kernel has none)
    batch_src: str = ""           # the batched variants (many instances per launch), compiled on demand


_BATCHED_KERNEL = re.compile(r"_(b|chkb|ppb)$")


def split_batched(src: str) -> tuple:
    """(plain source, batched source) of one emitted module.  Both keep every line
    outside the kernel definitions (the statement structs, templates and device
    functions: templates cost nothing until a kernel instantiates them); each kernel
    definition goes to one of them, by its name (batched variants end in _b, _chkb or
    _ppb).  Plain calls then compile only what they launch: the batched kernels were
    about 40 % of the compile time of the 2049^2 cavity module."""
    lines = src.split("\n")
    plain, batch = [], []
    i = 0
    while i < len(lines):
        line = lines[i]
        if not line.startswith('extern "C" __global__'):
            plain.append(line)
            batch.append(line)
            i += 1
            continue
        name = re.search(r"\b(xg_\w+)\(", line).group(1)
        block = [line]
        depth = line.count("{") - line.count("}")
        i += 1
        while depth > 0:
            block.append(lines[i])
            depth += lines[i].count("{") - lines[i].count("}")
            i += 1
        (batch if _BATCHED_KERNEL.search(name) else plain).extend(block)
    return "\n".join(plain), "\n".join(batch)


class _ExprC:
    """Expression text with C semantics, shared by host and device lowering."""

    def __init__(self, device: bool, cell=None, name=None, grid_info=None, udivs: list | None = None,
                 hoisted: dict | None = None, fold_pow: bool = False):
        self.device = device
        self.cell = cell              # callback for ir.Cell (device statements only)
        self.name = name              # callback for ir.Name
        self.grid_info = grid_info    # callback for ir.GridInfo
        # device statements: divisor texts of the per-cell divisions by a launch-uniform
        # value, which are emitted as xg::udiv<DV> (see _uniform_division)
        self.udivs = udivs
        self.hoisted = hoisted or {}  # id(node) -> text of its host-computed value
        self.fold_pow = fold_pow      # host text with the device's pow folds (hoisted values)

    def __call__(self, e) -> str:
        h = self.hoisted.get(id(e))
        if h is not None:
            return h
        m = getattr(self, "e_" + type(e).__name__, None)
        if m is None:
            raise EmitError(f"cannot lower expression {type(e).__name__}")
        return m(e)

    def e_Const(self, e):
        return const_text(e)

    def e_Name(self, e):
        return self.name(e)

    def e_Field(self, e):
        return f"({self(e.value)}).{e.attr}"

    def e_Cell(self, e):
        if self.cell is None:
            raise EmitError(f"{e.loc}: Unable to perform load operation to grid '{e.grid.name}' without stencil context")
        return self.cell(e)

    def e_Bin(self, e):
        if e.op == "**":
            double = isinstance(e.type, Floating) and e.type.width_bytes == 8
            fp = "double" if double else "float"
            if not self.device and not self.fold_pow:
                return f"{'pow' if double else 'powf'}({self(e.left)}, {self(e.right)})"
            fold = _pow_fold(e.right)
            x = f"(({fp})({self(e.left)}))"
            if fold == "sq":
                return f"xg::sq{x}" if self.device else f"({x} * {x})"
            if fold == "id":
                return x
            if fold == "inv":
                return f"(({'1.0' if double else '1.0f'}) / {x})"
            return f"{'pow' if double else 'powf'}({x}, (({fp})({self(e.right)})))"
        if self.udivs is not None and _uniform_division(e):
            d = self(e.right)
            if d not in self.udivs:
                self.udivs.append(d)
            return f"xg::udiv<DV>({self(e.left)}, {d})"
        return f"({self(e.left)} {e.op} {self(e.right)})"

    def e_Un(self, e):
        return f"({e.op} {self(e.operand)})"

    def e_Select(self, e):
        return f"({self(e.cond)} ? {self(e.then)} : {self(e.orelse)})"

    def e_Cast(self, e):
        tn = device_type_name(e.type) if self.device else ctype_name(e.type)
        return f"(({tn})({self(e.value)}))"

    def e_Call(self, e):
        args = [self(a) for a in e.args]
        if e.ctor is not None:
            if self.device:
                return f"{e.ctor.name}{{{', '.join(args)}}}"
            return f"(({ctype_name(e.ctor)}) {{{', '.join(args)}}})"
        if self.device and is_device_math(e.callee):
            ptypes = DEVICE_MATH[e.callee.name][1]
            if len(ptypes) != len(args):
                raise EmitError(f"{e.loc}: '{e.callee.name}' takes {len(ptypes)} argument(s)")
            args = [f"(({t})({a}))" for t, a in zip(ptypes, args)]
        return f"{function_symbol(e.callee)}({', '.join(args)})"

    def e_GridInfo(self, e):
        if e.info == "dimension":
            return str(e.grid.type.dimension)
        return self.grid_info(e)


_FP_RANK = {"float": 1, "double": 2}


def _c_type(e):
    """The C type of an expression's text as this module writes it (C's usual
    arithmetic conversions; an unsuffixed float literal is a double, as in the
    reference's generated C), or None where it is not a plain scalar."""
    if isinstance(e, ir.Const):
        if type(e.value) is float:
            return "double"
        if type(e.value) is bool:
            return "int"
        return "int" if abs(e.value) < 2 ** 31 else None
    if isinstance(e, (ir.Name, ir.Field)):
        t = e.type
        if isinstance(t, Floating):
            return "double" if t.width_bytes == 8 else "float"
        return "int" if isinstance(t, (Integer, Boolean)) else None
    if isinstance(e, ir.GridInfo):
        return "int"
    if isinstance(e, ir.Cast):
        t = e.type
        if isinstance(t, Floating):
            return "double" if t.width_bytes == 8 else "float"
        return "int" if isinstance(t, (Integer, Boolean)) else None
    if isinstance(e, ir.Un):
        return _c_type(e.operand)
    if isinstance(e, ir.Bin):
        if e.op == "**":
            return "double" if isinstance(e.type, Floating) and e.type.width_bytes == 8 else "float"
        if e.op in ("==", "!=", "<", "<=", ">", ">=", "&&", "||"):
            return "int"
        a, b = _c_type(e.left), _c_type(e.right)
        if a is None or b is None:
            return None
        ra, rb = _FP_RANK.get(a, 0), _FP_RANK.get(b, 0)
        return a if ra >= rb else b
    if isinstance(e, ir.Select):
        a, b = _c_type(e.then), _c_type(e.orelse)
        if a is None or b is None:
            return None
        return a if _FP_RANK.get(a, 0) >= _FP_RANK.get(b, 0) else b
    return None


def _host_exact(e) -> bool:
    """Can the host stub compute this subexpression, bit for bit as the device
    would, once per call: no grid load, no call (device and host functions may
    differ), no integer division or remainder (division by zero would stop the
    host), no conversion of a floating value to an integer (undefined out of
    range), and `**` only where both sides fold it (x*x, x, 1/x)."""
    if isinstance(e, ir.Const):
        return _c_type(e) is not None
    if isinstance(e, ir.Name):
        return not isinstance(e.type, Structure)
    if isinstance(e, ir.Field):
        return _host_exact(e.value) if not isinstance(e.value, ir.Name) else True
    if isinstance(e, ir.GridInfo):
        return e.dim is None or isinstance(e.dim, ir.Const)
    if isinstance(e, ir.Cast):
        src = _c_type(e.value)
        if src in _FP_RANK and not isinstance(e.type, Floating):
            return False
        return src is not None and _host_exact(e.value)
    if isinstance(e, ir.Un):
        return _host_exact(e.operand)
    if isinstance(e, ir.Select):
        return _host_exact(e.cond) and _host_exact(e.then) and _host_exact(e.orelse)
    if isinstance(e, ir.Bin):
        if e.op == "**" and _pow_fold(e.right) is None:
            return False
        if e.op in ("/", "%") and not isinstance(e.type, Floating):
            return False
        return _host_exact(e.left) and (e.op == "**" or _host_exact(e.right))
    return False


def param_fields(m, type_name) -> str:
    """Fields of a statement's parameter block: its captured scalars, then its
    hoisted values (the host stub, the device code and ctypes share this layout)."""
    parts = [f"{type_name(captured_value_type(v))} {v.name};" for v in m.params]
    parts += [f"{h.ctype} {h.name};" for h in m.hoisted]
    return " ".join(parts) if parts else "int32_t xg_unused;"


def _is_leaf(e) -> bool:
    return isinstance(e, (ir.Const, ir.Name, ir.GridInfo)) or \
        (isinstance(e, ir.Field) and _is_leaf(e.value))


def _uniform_division(e) -> bool:
    """A floating-point division whose dividend loads a grid (a value per cell) and
    whose divisor loads none (one value per launch: parameters, constants, shapes).
    Such a divisor is tested once per workgroup; when it is a power of two whose
    reciprocal is a normal or exact subnormal number, the kernel multiplies by
    that reciprocal instead, which is exactly IEEE division there (see
    xg::udiv)."""
    if e.op != "/" or not isinstance(e.type, Floating):
        return False
    return any(isinstance(n, ir.Cell) for n in ir.walk(e.left)) and \
        not any(isinstance(n, ir.Cell) for n in ir.walk(e.right))


def _math_table() -> dict:
    """C99 <math.h> functions the device provides under the same name: name ->
    (return type, parameter types).  An @xgrid.external of one of these may be
    called inside a stencil statement; its arguments are converted to the C
    prototype's types first, so the device evaluates what the reference's C
    evaluates (a float argument of `sin` is widened to double, as in C)."""
    one = ["sqrt", "cbrt", "fabs", "floor", "ceil", "trunc", "round", "exp", "exp2", "expm1", "log", "log2",
           "log10", "log1p", "sin", "cos", "tan", "asin", "acos", "atan", "sinh", "cosh", "tanh", "asinh",
           "acosh", "atanh", "erf", "erfc", "tgamma", "lgamma"]
    two = ["pow", "atan2", "fmod", "fmin", "fmax", "hypot", "copysign", "fdim"]
    out = {}
    for names, k in ((one, 1), (two, 2), (["fma"], 3)):
        for n in names:
            out[n] = ("double", ["double"] * k)
            out[n + "f"] = ("float", ["float"] * k)
    return out


DEVICE_MATH = _math_table()


def is_device_math(op) -> bool:
    return op.mode == "external" and op.name in DEVICE_MATH


def function_symbol(op) -> str:
    """C name of a user function: an external keeps its own name (its header
    declares it, as in the reference's generated C); others get a prefix."""
    if op.mode == "external":
        return op.name
    return "xgf_" + op.c_name


class _St3Plan:
    """Staged loads of a 3-D statement over one tile (4 planes x 4 rows x 64 * CPL
    columns, work-item = (row wave, column lane), CPL columns of 4 planes per
    work-item), for the arrays whose elements have the statement's size (4 or
    8 bytes).

    Per array j (a distinct (grid, level) of the statement) the columns of the
    planes the statement reads are held in registers `q{j}[u][k]`, plane
    p0 + pmin + u, the lane's k-th column.  A load at offset (o0, 0, 0) reads
    q{j}[n + o0 - pmin][k]; at (o0, 0, +-1) the lane's next or previous column,
    or the neighbouring lane's, moved with a DPP lane shift (lanes 0 and 63 read
    their outer neighbour from memory); at (o0, +-1, 0) a row of an LDS copy of
    the tile's 4 rows plus one halo row on each side.  Other loads (diagonals,
    larger offsets, lower-dimensional grids, arrays of another element size)
    stay global loads.  Only used in tiles whose cells and reach lie inside the
    grid, so every value a stored cell uses is exact; padding lanes and rows
    read clamped addresses and are not stored.

    A lane holds 16 bytes of each row (CPL = 2 columns of 8-byte elements, 4 of
    4-byte ones), so a wave reads 1 KB of a row at a time: the 512^3 FP64 7-point
    update measured 376 us against 534 us with one column per lane in a probe.
    Where a tile lies inside the row and the array is 16-byte aligned with a row
    pitch that is a multiple of CPL, a lane's columns are one 16-byte load;
    otherwise one element at a time.  XGRID_STAGE3_COLS=1 selects one column per
    lane, XGRID_STAGE3_F32=0 leaves 4-byte arrays unstaged."""

    NP, ROWS = 4, 4
    MAX_WINDOW = 8          # most planes held in registers per array
    MAX_LDS_GROUPS = 3      # most (array, plane offset) pairs staged in LDS

    def __init__(self, refs, types, sizes, esize: int = 8):
        self.refs = refs            # [(grid, level)]
        self.types = types          # device element type per ref index
        self.sizes = sizes          # element bytes per ref index
        self.esize = esize          # bytes of the statement's elements: arrays of this size are staged
        self.planes = {}            # j -> set of o0 read by staged loads
        self.xg = []                # [(j, o0)] with a column-neighbour load
        self.xside = {}             # (j, o0) -> set of +-1
        self.yg = []                # [(j, o0)] with a row-neighbour load
        # 16 bytes of a row per lane: 2 columns of 8-byte elements, 4 of 4-byte ones
        self.cpl = 1 if os.environ.get("XGRID_STAGE3_COLS", "") == "1" else 16 // esize
        self.COLS = 64 * self.cpl   # columns of a tile

    def add(self, j, off) -> bool:
        o0, o1, o2 = off
        if self._size(j) != self.esize:
            return False            # an array of another element size keeps global loads
        if self.esize == 4 and os.environ.get("XGRID_STAGE3_F32", "1") == "0":
            return False
        if o1 == 0 and o2 == 0:
            self.planes.setdefault(j, set()).add(o0)
        elif o1 == 0 and abs(o2) == 1:
            self.planes.setdefault(j, set()).add(o0)
            if (j, o0) not in self.xg:
                self.xg.append((j, o0))
            self.xside.setdefault((j, o0), set()).add(o2)
        elif abs(o1) == 1 and o2 == 0:
            self.planes.setdefault(j, set()).add(o0)
            if (j, o0) not in self.yg:
                self.yg.append((j, o0))
        else:
            return False
        return True

    def usable(self) -> bool:
        if not self.xg and not self.yg:
            return False            # nothing a plain load would not do as well
        if len(self.yg) > self.MAX_LDS_GROUPS:
            return False
        return all(max(p) - min(p) + self.NP <= self.MAX_WINDOW for p in self.planes.values())

    def pmin(self, j) -> int:
        return min(self.planes[j])

    def window(self, j) -> int:
        return max(self.planes[j]) - self.pmin(j) + self.NP

    def access(self, j, off):
        """The staged value of cell n's load at `off` (n = plane * CPL + column)."""
        if j not in self.planes:
            return None
        o0, o1, o2 = off
        pl = "n" if self.cpl == 1 else f"(n / {self.cpl})"
        k = "0" if self.cpl == 1 else f"(n % {self.cpl})"
        q = f"st.q{j}[{pl} + {o0 - self.pmin(j)}]"
        if o1 == 0 and o2 == 0:
            return f"{q}[{k}]"
        if o1 == 0 and abs(o2) == 1 and (j, o0) in self.xg:
            g = self.xg.index((j, o0))
            if self.cpl == 1:
                return f"st.x{'p' if o2 > 0 else 'n'}{g}"
            # the lane's other column, or the neighbouring lane's
            if o2 > 0:
                return f"({k} == {self.cpl - 1} ? st.xp{g} : {q}[{k} + 1])"
            return f"({k} == 0 ? st.xn{g} : {q}[{k} - 1])"
        if abs(o1) == 1 and o2 == 0 and (j, o0) in self.yg:
            return f"st.y{'p' if o1 > 0 else 'n'}{self.yg.index((j, o0))}[{k}]"
        return None

    def lds_bytes(self) -> int:
        rows = self.ROWS + 2
        return sum(self.NP * rows * self.COLS * self._size(j) for j, _ in self.yg)

    def _size(self, j) -> int:
        return self.sizes[j]

    def members(self, expr: str) -> list:
        """st3_load: the plane registers, the edge lanes' outer columns and the
        LDS copy (ends with a barrier); st3_plane_n: the neighbours of plane
        n, formed just before its cells so that only one plane's are live;
        value_st3: the statement's expression over them."""
        NP, R2, C, K = self.NP, self.ROWS + 2, self.COLS, self.cpl
        out = ["  static constexpr bool ST3 = true;", f"  static constexpr int ST3_LDS = {max(16, self.lds_bytes())};",
               f"  static constexpr int CPL = {K};"]
        fields = []
        for j in sorted(self.planes):
            fields.append(f"{self.types[j]} q{j}[{self.window(j)}][{K}];")
        for g, (j, o0) in enumerate(self.xg):
            fields.append(f"{self.types[j]} e{g}[{NP}];")
            for sd in sorted(self.xside[(j, o0)]):
                fields.append(f"{self.types[j]} x{'p' if sd > 0 else 'n'}{g};")
        for g, (j, o0) in enumerate(self.yg):
            fields.append(f"{self.types[j]} yp{g}[{K}]; {self.types[j]} yn{g}[{K}];")
        fields.append("const unsigned char* lds; int lane, wr;")
        out.append(f"  struct St3 {{ {' '.join(fields)} }};")
        out.append("  __device__ static __forceinline__ void st3_load(St3& st, unsigned char* lds, const Refs& R, "
                   "const int32_t* sh, int p0, int r0, int c0) {")
        # In a tile whose columns all lie inside the row (`full`: every tile but a
        # partial last one) a lane's columns need no clamp; where their address is a
        # multiple of 16 bytes they are read as one 16-byte load.  (Marking the wave's
        # row uniform with readfirstlane made the FP64 heat kernel spill.)
        # 8-byte tiles load at a raised wave priority, so a wave that starts its
        # tile issues its loads ahead of other waves' arithmetic (512^3 FP64
        # update 385.3-385.6 against 389.4-390.4 us; FP32 slower, 205-207 against 204)
        prio = self.esize == 8
        if prio:
            out.append("    __builtin_amdgcn_s_setprio(1);")
        out.append("    const int lane = threadIdx.x & 63, wr = threadIdx.x >> 6;")
        out.append("    st.lds = lds; st.lane = lane; st.wr = wr;")
        out.append("    const int64_t P1 = sh[2], P0 = (int64_t)sh[1] * sh[2];")
        out.append("    const int64_t row = min(r0 + wr, sh[1] - 1);")
        out.append(f"    const int cl = c0 + lane * {K};")
        out.append(f"    const bool full = c0 + {C} <= sh[2];")
        out.append("    (void)lds; (void)P1; (void)P0; (void)cl; (void)full;")
        for j in sorted(self.planes):
            t = self.types[j]
            out.append(f"    const xg::gptr<const {t}> g{j} = xg::gp<const {t}>(R.v[{j}].ptr);")
            if K * self._size(j) == 16:
                # a lane's columns at a 16-byte aligned address: a 16-byte aligned base,
                # a row pitch that is a multiple of K (then every plane pitch is) and
                # column multiples of K
                out.append(f"    if (full && ((uint64_t)R.v[{j}].ptr & 15) == 0 && (P1 % {K}) == 0) {{")
                out.append("#pragma unroll")
                out.append(f"      for (int u = 0; u < {self.window(j)}; ++u) {{")
                out.append(f"        const xg::gptr<const {t}> rp = g{j} + (xg::st3_plane(p0 + u + {self.pmin(j)}, sh) * P0 + row * P1);")
                out.append(f"        const xg::vecn<{t}, {K}> v = *(xg::gptr<const xg::vecn<{t}, {K}>>)(rp + cl);")
                out.append("#pragma unroll")
                out.append(f"        for (int k = 0; k < {K}; ++k) st.q{j}[u][k] = v[k];")
                out.append("      }")
                out.append("    } else {")
            else:
                out.append("    {")
            out.append("#pragma unroll")
            out.append(f"      for (int u = 0; u < {self.window(j)}; ++u) {{")
            out.append(f"        const xg::gptr<const {t}> rp = g{j} + (xg::st3_plane(p0 + u + {self.pmin(j)}, sh) * P0 + row * P1);")
            out.append("#pragma unroll")
            out.append(f"        for (int k = 0; k < {K}; ++k) st.q{j}[u][k] = rp[min(cl + k, sh[2] - 1)];")
            out.append("      }")
            out.append("    }")
        if self.xg:
            out.append("    const bool edge = lane == 0 || lane == 63;")
            out.append(f"    const int64_t ecol = lane == 0 ? max(c0 - 1, 0) : min(c0 + {C}, sh[2] - 1);")
        for g, (j, o0) in enumerate(self.xg):
            t = self.types[j]
            out.append("#pragma unroll")
            out.append(f"    for (int n = 0; n < {NP}; ++n) st.e{g}[n] = edge ? "
                       f"g{j}[xg::st3_plane(p0 + n + {o0}, sh) * P0 + row * P1 + ecol] : ({t})0;")
        off = 0
        lds_dma = False
        self._offs = []
        for g, (j, o0) in enumerate(self.yg):
            t = self.types[j]
            self._offs.append(off)
            out.append(f"    {{ {t}* L = ({t}*)(lds + {off});")
            out.append("#pragma unroll")
            out.append(f"      for (int n = 0; n < {NP}; ++n)")
            out.append("#pragma unroll")
            out.append(f"        for (int k = 0; k < {K}; ++k) L[(n * {R2} + wr + 1) * {C} + lane * {K} + k] = "
                       f"st.q{j}[n + {o0 - self.pmin(j)}][k];")
            # the halo rows above and below the tile: 2 x NP rows of C columns over 256 work-items
            if K * self._size(j) == 16:
                # K columns per work-item, one 16-byte load, where aligned as above,
                # straight into LDS: a wave loads one halo row (64 lanes x 16 bytes =
                # C columns), so its LDS row is the same for the whole wave
                out.append(f"      if (full && ((uint64_t)R.v[{j}].ptr & 15) == 0 && (P1 % {K}) == 0) {{")
                out.append("#pragma unroll")
                out.append(f"        for (int h = 0; h < {2 * NP * 64 // 256}; ++h) {{")
                out.append(f"          const int e = wr * 64 + h * 256, cc = lane * {K}, "
                           f"pl = (e / 64) % {NP}, hr = e / {64 * NP};")
                out.append(f"          const int64_t gr = min(max(hr ? r0 + {self.ROWS} : r0 - 1, 0), sh[1] - 1);")
                out.append(f"          xg::load_lds16(g{j} + (xg::st3_plane(p0 + pl + {o0}, sh) * P0 + gr * P1 + c0 + cc), "
                           f"L + (pl * {R2} + (hr ? {R2 - 1} : 0)) * {C});")
                out.append("        }")
                lds_dma = True
                out.append("      } else {")
            else:
                out.append("      {")
            out.append("#pragma unroll")
            out.append(f"      for (int h = 0; h < {2 * NP * C // 256}; ++h) {{")
            out.append(f"        const int e = threadIdx.x + h * 256, cc = e % {C}, pl = (e / {C}) % {NP}, "
                       f"hr = e / {C * NP};")
            out.append(f"        const int64_t gr = min(max(hr ? r0 + {self.ROWS} : r0 - 1, 0), sh[1] - 1);")
            out.append(f"        L[(pl * {R2} + (hr ? {R2 - 1} : 0)) * {C} + cc] = "
                       f"g{j}[xg::st3_plane(p0 + pl + {o0}, sh) * P0 + gr * P1 + min(c0 + cc, sh[2] - 1)];")
            out.append("      }")
            out.append("      }")
            out.append("    }")
            off += NP * R2 * C * self._size(j)
        if lds_dma:
            out.append("    xg::wait_vm();   /* the halo rows loaded straight into LDS */")
        if self.yg:
            out.append("    __syncthreads();")
        if prio:
            out.append("    __builtin_amdgcn_s_setprio(0);")
        out.append("  }")
        # Every work-item of the workgroup runs this for every n, and the lane shifts come
        # before the edge lanes' selects: a shift written inside `lane == 63 ? e : shift`
        # is evaluated only by the other lanes, and a lane reading a disabled lane gets 0.
        out.append("  __device__ static __forceinline__ void st3_plane_n(St3& st, int n) {")
        out.append("    (void)st; (void)n;")
        for g, (j, o0) in enumerate(self.xg):
            t = self.types[j]
            q = f"st.q{j}[n + {o0 - self.pmin(j)}]"
            if 1 in self.xside[(j, o0)]:
                # right of the lane's last column: the next lane's first
                out.append(f"    {{ const {t} cf = {q}[0];")
                out.append(f"      const {t} a = xg::lane_shift<1>(cf);")
                out.append(f"      st.xp{g} = st.lane == 63 ? st.e{g}[n] : a; }}")
            if -1 in self.xside[(j, o0)]:
                # left of the lane's first column: the previous lane's last
                out.append(f"    {{ const {t} cl = {q}[{K - 1}];")
                out.append(f"      const {t} b = xg::lane_shift<-1>(cl);")
                out.append(f"      st.xn{g} = st.lane == 0 ? st.e{g}[n] : b; }}")
        for g, (j, o0) in enumerate(self.yg):
            t = self.types[j]
            out.append(f"    {{ const {t}* L = (const {t}*)(st.lds + {self._offs[g]});")
            out.append("#pragma unroll")
            out.append(f"      for (int k = 0; k < {K}; ++k) {{")
            out.append(f"        st.yn{g}[k] = L[(n * {R2} + st.wr) * {C} + st.lane * {K} + k];")
            out.append(f"        st.yp{g}[k] = L[(n * {R2} + st.wr + 2) * {C} + st.lane * {K} + k];")
            out.append("      }")
            out.append("    }")
        out.append("  }")
        out.append("  __device__ static __forceinline__ T value_st3(const St3& st, const Refs& R, const Params& P, "
                   "const int32_t* sh, int n, int i0, int i1, int i2) {")
        out.append("    constexpr bool SAFE = true;")
        out.append("    const int64_t xc = xg::lin<3>(sh, i0, i1, i2);")
        out.append("    (void)st; (void)R; (void)P; (void)sh; (void)n; (void)xc; (void)i0; (void)i1; (void)i2;")
        out.append(f"    return {expr};")
        out.append("  }")
        return out


class Emitter:
    def __init__(self, info: KernelInfo, overstep: str = "none", comment: bool = False,
                 entry: str = "xg_entry", macro: list | None = None, block_shape: tuple | None = None):
        self.info = info
        self.shape = tuple(block_shape) if block_shape is not None else None   # None: by element size
        self.used_shape = None         # the shape the blockable loops were built with
        self.macro = list(macro or [])
        self.func = info.func
        self.overstep = overstep
        self.comment = comment
        self.entry = entry
        self.functions = []            # user operators reachable from the kernel (dependency order)
        self.device_functions = []     # those reachable from stencil expressions
        self._prefetched = set()       # blocked-loop statements whose other-grid loads are prefetched
        self._tb_reads = {}            # blocked-loop statement -> its loop-grid read offsets, in slot order
        self._dv = {}                  # statement -> has a reciprocal form (uniform divisors; see udiv)
        self._collect_functions()

    # ------------------------------------------------------------- helpers
    def _collect_functions(self):
        def visit_fn(op, into, seen):
            if op in seen:
                return
            seen.add(op)
            if op.mode == "external":
                if op.intrinsic == "tick":
                    raise EmitError("xgrid.tick is declared but never defined by xgrid; it cannot be called")
                into.append(op)
                return
            for n in ir.walk_all(op.ir.body):
                if isinstance(n, ir.Call) and n.callee is not None:
                    visit_fn(n.callee, into, seen)
                if isinstance(n, ir.Cell):
                    raise EmitError(f"{n.loc}: grid access inside @function '{op.name}' is not supported")
            into.append(op)

        seen = set()
        for n in ir.walk_all(self.func.body):
            if isinstance(n, ir.Call) and n.callee is not None:
                visit_fn(n.callee, self.functions, seen)
        dseen = set()
        for st in self.info.stmts:
            for n in ir.walk(st.node.value):
                if isinstance(n, ir.Call) and n.callee is not None:
                    visit_fn(n.callee, self.device_functions, dseen)
        # externals in stencil statements: the <math.h> functions the device has
        for op in self.device_functions:
            if op.mode == "external" and not is_device_math(op):
                raise EmitError(f"external function '{op.name}' cannot be evaluated inside a stencil statement: "
                                f"only <math.h> functions the device provides can ({', '.join(sorted(DEVICE_MATH))})")
        self.device_functions = [op for op in self.device_functions if not is_device_math(op)]
        for op in self.device_functions:
            for n in ir.walk_all(op.ir.body):
                if isinstance(n, ir.Inline):
                    raise EmitError(f"@function '{op.name}' with inline C cannot be evaluated inside a stencil statement")

    def _structs(self) -> list:
        out = []
        funcs = [self.func] + [op.ir for op in self.functions if op.mode != "external"]
        for f in funcs:
            for s in structures_in(f):
                if s not in out:
                    out.append(s)
        return out

    # ------------------------------------------------------------- device
    def _stmt_refs(self, st: StencilStmt) -> list:
        refs = list(st.refs)
        for g in st.info_grids:
            if not any(r[0] == g for r in refs):
                refs.append((g, -1))
        if len(refs) > 16:
            raise EmitError(f"{st.node.loc}: a statement may load at most 16 distinct (grid, time level) arrays")
        return refs

    def _plan_st3(self, st: StencilStmt, refs: list, grids: dict):
        """The staged plan of a 3-D statement, or None (plain loads).
        XGRID_STAGE3=0 turns staging off (for comparisons)."""
        if os.environ.get("XGRID_STAGE3", "1") == "0":
            return None
        types = [device_type_name(grids[g].element) for g, _ in refs]
        sizes = [element_nbytes(grids[g].element) for g, _ in refs]
        esize = element_nbytes(st.gtype.element)
        if esize not in (4, 8):
            return None
        plan = _St3Plan(refs, types, sizes, esize)
        for n in ir.walk(st.node.value):
            if isinstance(n, ir.Cell) and grids[n.grid.name].dimension == 3:
                plan.add(refs.index((n.grid.name, n.level)), tuple(n.offsets))
        return plan if plan.usable() else None

    def _device_stmt(self, st: StencilStmt, refs: list, params: list, out: list, tb: LoopMeta | None = None,
                     fwd_from: str | None = None, meta: StmtMeta | None = None):
        sid = st.sid
        D = st.dimension
        # launch-uniform subexpressions the host computed: read from the parameter block
        hoisted = {id(n): f"P.{h.name}" for h in (meta.hoisted if meta else []) for n in h.nodes}
        T = device_type_name(st.gtype.element)
        grids = {n: t for _, n, t in self.func.grid_args}
        in_lds = [False]     # while lowering value_tb: the loop grid's level 0 lives in LDS
        in_reg = [False]     # while lowering value_rg: the loop grid's level 0 lives in registers
        pre_mode = [False]   # ... and the other loads come from a per-cell prefetch
        pre_slots: list = []  # (grid, level, offsets) of each prefetched load
        rd_slots: list = []   # distinct (row, column) offsets of the loop grid's level-0 reads
        st3 = [None]          # while lowering value_st3: the staged 3-D plan (see _plan_st3)
        fwd = [False]         # while lowering value_fwd: time-0 loads of `fwd_from` at the cell are `fwd`

        def cell(e: ir.Cell) -> str:
            if fwd[0] and e.grid.name == fwd_from and e.level == 0 and not any(e.offsets):
                return "fwd"
            if st3[0] is not None:
                acc = st3[0].access(refs.index((e.grid.name, e.level)), tuple(e.offsets))
                if acc is not None:
                    return acc
            if in_lds[0] and e.grid.name == st.grid and e.level == 0:
                return f"xg::tb_ld<T, CHECK, RHH, RWW>(lds, vld, ok, lr + {e.offsets[0]}, lc + {e.offsets[1]})"
            if in_reg[0] and e.grid.name == st.grid and e.level == 0:
                key = (e.offsets[0], e.offsets[1])
                if key not in rd_slots:
                    rd_slots.append(key)
                return f"s[{rd_slots.index(key)}]"
            if pre_mode[0]:
                key = (e.grid.name, e.level, tuple(e.offsets))
                if key not in pre_slots:
                    pre_slots.append(key)
                return f"pre.v{pre_slots.index(key)}"
            r = refs.index((e.grid.name, e.level))
            gt = grids[e.grid.name]
            if gt.dimension > D:
                raise EmitError(f"{e.loc}: grid '{e.grid.name}' has more dimensions than the store grid")
            offs = [str(e.offsets[d]) if d < gt.dimension else "0" for d in range(3)]
            et = device_type_name(gt.element)
            return (f"xg::ld<{et}, {gt.dimension}, XG_OVERSTEP, SAFE>(R.v[{r}], sh, xc, {', '.join(offs)}, "
                    f"i0, i1, i2)")

        def name(e: ir.Name) -> str:
            return f"P.{e.var.name}"

        def grid_info(e: ir.GridInfo) -> str:
            r = next(i for i, ref in enumerate(refs) if ref[0] == e.grid.name)
            return f"R.v[{r}].shape[{_ExprC(True, cell, name, grid_info)(e.dim)}]"

        udivs: list = []      # divisor texts of the per-cell divisions by a launch-uniform value
        expr = _ExprC(True, cell, name, grid_info, udivs, hoisted)(st.node.value)
        expr_st3 = None
        plan3 = self._plan_st3(st, refs, grids) if D == 3 and tb is None else None
        if plan3 is not None:
            st3[0] = plan3
            expr_st3 = _ExprC(True, cell, name, grid_info, udivs, hoisted)(st.node.value)
            st3[0] = None
        expr_tb = expr_pre = expr_lds = expr_lds_pre = None
        if tb is not None:
            in_reg[0] = True
            expr_tb = _ExprC(True, cell, name, grid_info, udivs, hoisted)(st.node.value)
            in_reg[0] = False
            in_lds[0] = True
            expr_lds = _ExprC(True, cell, name, grid_info, udivs, hoisted)(st.node.value)
            in_lds[0] = False
            if st.klass == 0:
                pre_mode[0] = True
                in_reg[0] = True
                expr_pre = _ExprC(True, cell, name, grid_info, udivs, hoisted)(st.node.value)
                in_reg[0] = False
                in_lds[0] = True
                expr_lds_pre = _ExprC(True, cell, name, grid_info, udivs, hoisted)(st.node.value)
                in_lds[0] = False
                pre_mode[0] = False
                if len(pre_slots) > self.PREFETCH_MAX:
                    expr_pre = expr_lds_pre = None
            self._tb_reads[sid] = list(rd_slots)
        so = list(st.store_offsets) + [0] * (3 - len(st.store_offsets))
        rl, rh = [0, 0, 0], [0, 0, 0]
        for n in ir.walk(st.node.value):
            if isinstance(n, ir.Cell):
                for d, o in enumerate(n.offsets):
                    rl[d] = max(rl[d], -o)
                    rh[d] = max(rh[d], o)
        nr = max(1, len(refs))
        out.append(f"// statement {sid}: {st.node.loc}")
        out.append(f"struct xg_p{sid} {{ {param_fields(meta, device_type_name)} }};")
        out.append(f"struct xg_r{sid} {{ xg_gview v[{nr}]; }};")
        # DV: the statement's uniform divisors were found to be exact powers of two for
        # this launch (dv_ok), so xg::udiv multiplies by their reciprocals
        out.append(f"template <bool DV> struct xg_s{sid}_t {{")
        out.append(f"  typedef {T} T; typedef xg_r{sid} Refs; typedef xg_p{sid} Params;")
        out.append(f"  static constexpr int NDV = {len(udivs)};")
        out.append("  __device__ static __forceinline__ bool dv_ok(const Refs& R, const Params& P) {")
        out.append("    (void)R; (void)P;")
        out.append(f"    return {' && '.join(f'xg::rcp_exact({d})' for d in udivs) if udivs else 'false'};")
        out.append("  }")
        out.append(f"  static constexpr int D = {D}, SO0 = {so[0]}, SO1 = {so[1]}, SO2 = {so[2]};")
        out.append(f"  static constexpr int RL0 = {rl[0]}, RL1 = {rl[1]}, RL2 = {rl[2]}, "
                   f"RH0 = {rh[0]}, RH1 = {rh[1]}, RH2 = {rh[2]};")
        out.append("  template <bool SAFE>")
        out.append(f"  __device__ static __forceinline__ T value(const Refs& R, const Params& P, const int32_t* sh, "
                   f"int i0, int i1, int i2) {{")
        out.append(f"    const int64_t xc = SAFE ? xg::lin<{D}>(sh, i0, i1, i2) : 0;")
        out.append("    (void)R; (void)P; (void)sh; (void)xc; (void)i0; (void)i1; (void)i2;")
        out.append(f"    return {expr};")
        out.append("  }")
        if expr_st3 is not None:
            out.extend(plan3.members(expr_st3))
        else:
            out.append("  static constexpr bool ST3 = false;")
            out.append("  static constexpr int ST3_LDS = 0;")
            out.append("  static constexpr int CPL = 1;")
        if meta is not None:
            # the host sizes this statement's 3-D launches by its tile width
            meta.tile3_cols = plan3.COLS if expr_st3 is not None else 64
        if fwd_from is not None:
            # the same expression with its time-0 loads of `fwd_from` at the cell taken from the
            # value the paired statement just computed for that cell
            fwd[0] = True
            expr_fwd = _ExprC(True, cell, name, grid_info, udivs, hoisted)(st.node.value)
            fwd[0] = False
            ft = device_type_name(grids[fwd_from].element)
            out.append("  template <bool SAFE>")
            out.append(f"  __device__ static __forceinline__ T value_fwd(const Refs& R, const Params& P, const int32_t* sh, "
                       f"int i0, int i1, int i2, const {ft} fwd) {{")
            out.append(f"    const int64_t xc = SAFE ? xg::lin<{D}>(sh, i0, i1, i2) : 0;")
            out.append("    (void)R; (void)P; (void)sh; (void)xc; (void)i0; (void)i1; (void)i2; (void)fwd;")
            out.append(f"    return {expr_fwd};")
            out.append("  }")
        if expr_tb is not None:
            # the same expression with the loop grid's level-0 reads taken from s[], the values
            # at the statement's distinct offsets (gathered from registers by the blocked kernel)
            out.append(f"  static constexpr int NRD = {max(1, len(rd_slots))};")
            out.append("  template <bool SAFE>")
            out.append("  __device__ static __forceinline__ T value_rg(const T* s, const Refs& R, const Params& P, "
                       "const int32_t* sh, int i0, int i1, int i2) {")
            out.append(f"    const int64_t xc = SAFE ? xg::lin<{D}>(sh, i0, i1, i2) : 0;")
            out.append("    (void)s; (void)R; (void)P; (void)sh; (void)xc; (void)i0; (void)i1; (void)i2;")
            out.append(f"    return {expr_tb};")
            out.append("  }")
            # ... and read from the workgroup's LDS region
            out.append("  template <bool SAFE, bool CHECK, int RHH, int RWW>")
            out.append("  __device__ static __forceinline__ T value_tb(const T* lds, const uint8_t* vld, bool& ok, "
                       "int lr, int lc, const Refs& R, const Params& P, const int32_t* sh, int i0, int i1, int i2) {")
            out.append(f"    const int64_t xc = SAFE ? xg::lin<{D}>(sh, i0, i1, i2) : 0;")
            out.append("    (void)lds; (void)vld; (void)ok; (void)lr; (void)lc; (void)R; (void)P; (void)sh; (void)xc;"
                       " (void)i0; (void)i1; (void)i2;")
            out.append(f"    return {expr_lds};")
            out.append("  }")
        if expr_pre is not None:
            # loads of grids the loop does not write, fetched once per launch per cell
            gts = {n: t for _, n, t in self.func.grid_args}
            out.append(f"  static constexpr int NPRE = {len(pre_slots)};")
            fields_ = " ".join(f"{device_type_name(gts[g].element)} v{i};" for i, (g, _, _) in enumerate(pre_slots))
            out.append(f"  struct Pre {{ {fields_ or 'int32_t unused;'} }};")
            out.append("  template <bool SAFE>")
            out.append("  __device__ static __forceinline__ void prefetch(Pre& pre, const Refs& R, const int32_t* sh, "
                       "int i0, int i1, int i2) {")
            out.append(f"    const int64_t xc = SAFE ? xg::lin<{D}>(sh, i0, i1, i2) : 0;")
            out.append("    (void)R; (void)sh; (void)xc; (void)pre;")
            for i, (g, lvl, offs) in enumerate(pre_slots):
                gt = gts[g]
                r = refs.index((g, lvl))
                o3 = [str(offs[d]) if d < gt.dimension else "0" for d in range(3)]
                out.append(f"    pre.v{i} = xg::ld<{device_type_name(gt.element)}, {gt.dimension}, XG_OVERSTEP, SAFE>"
                           f"(R.v[{r}], sh, xc, {', '.join(o3)}, i0, i1, i2);")
            out.append("  }")
            out.append("  __device__ static __forceinline__ T value_rgp(const T* s, const Refs& R, const Params& P, "
                       "const Pre& pre) {")
            out.append("    (void)s; (void)R; (void)P; (void)pre;")
            out.append(f"    return {expr_pre};")
            out.append("  }")
            out.append("  template <bool CHECK, int RHH, int RWW>")
            out.append("  __device__ static __forceinline__ T value_pre(const T* lds, const uint8_t* vld, bool& ok, "
                       "int lr, int lc, const Refs& R, const Params& P, const Pre& pre) {")
            out.append("    (void)lds; (void)vld; (void)ok; (void)lr; (void)lc; (void)R; (void)P; (void)pre;")
            out.append(f"    return {expr_lds_pre};")
            out.append("  }")
            self._prefetched.add(sid)
        out.append("};")
        out.append(f"typedef xg_s{sid}_t<false> xg_s{sid}; typedef xg_s{sid}_t<true> xg_s{sid}_r;")
        # a staged 3-D statement keeps its LDS copy in run_main: a second instantiation
        # would allocate a second copy, so it always divides
        dv = bool(udivs) and expr_st3 is None
        self._dv[sid] = dv

        def disp(call: str, rv: str, pv: str) -> str:
            """`call` (with {S} for the statement type), taking the reciprocal form when
            this launch's divisors allow it (decided once per workgroup)."""
            if not dv:
                return call.format(S=f"xg_s{sid}") + ";"
            return (f"if (xg_s{sid}::dv_ok({rv}, {pv})) {call.format(S=f'xg_s{sid}_r')}; "
                    f"else {call.format(S=f'xg_s{sid}')};")

        args = f"const xg_launch_hdr H, const xg_r{sid} R, const xg_p{sid} P"
        remap = "xg::xcd_remap(blockIdx.x, gridDim.x)"
        # A staged 3-D statement's kernel holds plane registers and an LDS copy: ask for the
        # occupancy its LDS allows (at most 8 waves per SIMD), which the register allocator
        # then meets (the FP64 7-point update: 62 VGPRs instead of 100, no spills).
        # The batched variant keeps more state per work-item (its entry copied into locals):
        # two waves fewer keep it free of spills.
        wv = wvb = ""
        if expr_st3 is not None:
            w = max(1, min(8, 163840 // max(1, plan3.lds_bytes())))
            wv = f" __attribute__((amdgpu_waves_per_eu({w})))"
            wvb = f" __attribute__((amdgpu_waves_per_eu({max(1, w - 2)})))"
        out.append(f'extern "C" __global__ void __launch_bounds__(256){wv} xg_k{sid}({args}) '
                   f"{{ {disp('xg::run_main<{S}, XG_OVERSTEP>(H, R, P, ' + remap + ', gridDim.x)', 'R', 'P')} }}")
        # batched: one launch over the tiles of many instances (entries in a device table)
        # The entry's header, views and parameters are copied into locals first: read
        # through the table's generic pointers they could alias the kernel's own
        # stores, and would be re-read from memory for every cell.
        bargs = "const xg_bent* __restrict__ E, const int32_t n, const int64_t total"
        local = (f"const xg_launch_hdr h = e.hdr; const xg_r{sid} r = *xg::cp<xg_r{sid}>(e.refs); "
                 f"const xg_p{sid} p = *xg::cp<xg_p{sid}>(e.params);")
        out.append(f'extern "C" __global__ void __launch_bounds__(256){wvb} xg_k{sid}_b({bargs}) '
                   f"{{ xg::for_batch_tiles(E, n, total, [&](const xg_bent& e, uint32_t t) "
                   f"{{ {local} {disp('xg::run_main<{S}, XG_OVERSTEP>(h, r, p, t, e.ntiles)', 'r', 'p')} }}); }}")
        if st.mode == "pingpong":
            out.append(f'extern "C" __global__ void __launch_bounds__(256) xg_k{sid}_pp({args}) '
                       f"{{ {disp('xg::run_pingpong<{S}, XG_OVERSTEP>(H, R, P, ' + remap + ')', 'R', 'P')} }}")
            out.append(f'extern "C" __global__ void __launch_bounds__(256) xg_k{sid}_ppb({bargs}) '
                       f"{{ xg::for_batch_tiles(E, n, total, [&](const xg_bent& e, uint32_t t) "
                       f"{{ {local} {disp('xg::run_pingpong<{S}, XG_OVERSTEP>(h, r, p, t)', 'r', 'p')} }}); }}")

    def _device_function(self, op, out: list):
        f = op.ir
        sym = function_symbol(op)
        args = ", ".join(f"{device_type_name(t)} {n}" for n, t in f.args)
        out.append(f"__device__ {device_type_name(f.ret)} {sym}({args}) {{")
        for v in f.locals:
            out.append(f"  {device_type_name(v.type)} {v.name};")
        expr = _ExprC(True, None, lambda e: e.var.name, None)
        self._stmts_text(f.body, out, expr, 1, device=True)
        out.append("}")

    def _tb_kernel_lds(self, L: LoopMeta, out: list):
        """Blocked kernel of one loop: each workgroup keeps core + halo of the
        loop grid in LDS and runs up to `tmax` iterations of the body on it.
        Every statement is computed for all its cells in the region (reading
        the region as it was before the statement), then written back, so a
        statement sees exactly the values the statements before it produced."""
        T = device_type_name(L.stmts[0].stmt.gtype.element)
        c0, c1 = L.core
        m = len(L.stmts)
        params = ["const xg_tb_hdr& TH"] + [f"const xg_r{s.stmt.sid}& R{j}, const xg_p{s.stmt.sid}& P{j}"
                                            for j, s in enumerate(L.stmts)]
        # reach of the loads that come from global memory (other grids / levels)
        rl, rh = [0, 0], [0, 0]
        for s in L.stmts:
            for n in ir.walk(s.stmt.node.value):
                if isinstance(n, ir.Cell) and not (n.grid.name == L.grid and n.level == 0):
                    for d in range(2):
                        rl[d] = max(rl[d], -n.offsets[d])
                        rh[d] = max(rh[d], n.offsets[d])
        o = out.append
        o(f"template <bool CHECK, bool DVOK>")
        o(f"__device__ __forceinline__ void {L.kernel}_body({', '.join(params)}, const int tile) {{")
        o(f"  typedef {T} T;")
        o(f"  constexpr int TBT = {L.threads};")
        o(f"  constexpr int C0 = {c0}, C1 = {c1}, HL = {L.halo}, RHH = C0 + 2 * HL, RWW = C1 + 2 * HL;")
        o("  constexpr int NR = RHH * RWW, NC = (NR + TBT - 1) / TBT;")
        # two copies of the region (each statement reads one and writes the other), the
        # statement-membership bits of every cell and, when validating, exactness bytes
        o("  __shared__ T buf[2 * NR];")
        o("  __shared__ uint8_t mem[NR];")
        o("  __shared__ uint8_t vl[CHECK ? 2 * NR : 1];")
        # cells of the boundary (class != 0) statements, listed once per launch
        sparse_mask = sum(1 << j for j, s in enumerate(L.stmts) if s.stmt.klass != 0)
        o(f"  constexpr uint32_t SPARSE = {sparse_mask}u;")
        o("  constexpr int SL = NR / 4 > 256 ? NR / 4 : (NR < 256 ? NR : 256);   /* overflow: plan declined */")
        o("  __shared__ uint16_t slist[SL];")
        o("  __shared__ int scount;")
        o("  if (threadIdx.x == 0) scount = 0;")
        o("  __syncthreads();")
        o("  const int32_t* sh = TH.src.shape;")
        o("  const int N0 = sh[0], N1 = sh[1];")
        o("  const int nt1 = (N1 + C1 - 1) / C1;")
        o("  const int g0 = (tile / nt1) * C0, g1 = (tile % nt1) * C1;")
        o("  const int r0 = g0 - HL, r1 = g1 - HL;")
        o("  const xg::gptr<const T> __restrict__ src = xg::gp<const T>(TH.src.ptr);")
        o("  uint32_t any = 0;")
        o("#pragma unroll")
        o("  for (int n = 0; n < NC; ++n) {")
        o("    const int idx = threadIdx.x + TBT * n;")
        o("    if (idx < NR) {")
        o("      const int i0 = r0 + idx / RWW, i1 = r1 + idx % RWW;")
        o("      const bool in = i0 >= 0 && i0 < N0 && i1 >= 0 && i1 < N1;")
        o("      const int64_t c = (int64_t)i0 * N1 + i1;")
        o("      buf[idx] = in ? src[c] : T(0);")
        o("      uint32_t mb = 0u;")
        o("      if (in) {")
        o("        const int32_t k = TH.mask ? xg::gp<const int32_t>(TH.mask)[c] : 0;")
        o("        mb = " + " | ".join(f"(k == TH.klass[{j}] ? {1 << j}u : 0u)" for j in range(m)) + ";")
        o("      }")
        o("      mem[idx] = (uint8_t)mb;")
        o("      any |= mb;")
        o("      if (mb & SPARSE) {")
        o("        const int pos = atomicAdd(&scount, 1);")
        o("        if (pos < SL) slist[pos] = (uint16_t)idx;")
        o("      }")
        o("      if (CHECK) vl[idx] = in ? 1 : 0;")
        o("    }")
        o("  }")
        o("  any = xg::block_or(any);")
        o("  const int ns = scount < SL ? scount : SL;")
        o("  if (CHECK && threadIdx.x == 0 && scount > SL) *xg::gp<int32_t>(TH.fail) = 1;   /* list overflow: decline */")
        o(f"  const bool safe = TH.uniform && r0 - {rl[0]} >= 0 && r0 + RHH + {rh[0]} <= N0 && "
          f"r1 - {rl[1]} >= 0 && r1 + RWW + {rh[1]} <= N1;")
        # Cells at least `reach` inside the region are computed; their reads stay in the
        # region, so LDS indices need no clamp.  Cells nearer the edge only feed them
        # (and, being stale, are marked not exact by the validation variant).
        o(f"  constexpr int RE = {L.reach};")
        o("  int cb = 0;                 /* offset of the current copy in buf / vl */")
        # one instantiation per addressing mode of the other grids, chosen once per workgroup
        o("  auto iterate = [&](auto safe_c, auto dv_c) {")
        o("  constexpr bool SAFE = decltype(safe_c)::value;")
        o("  constexpr bool DV = decltype(dv_c)::value;   /* reciprocal form of uniform divisions */")
        o("  (void)DV;")
        for j, s in enumerate(L.stmts):
            if s.stmt.sid in self._prefetched:
                S = f"xg_s{s.stmt.sid}_t<DV>"
                o(f"  typename {S}::Pre pre{j}[NC];")
                o("#pragma unroll")
                o("  for (int k = 0; k < NC; ++k) {")
                o("    const int idx = threadIdx.x + TBT * k;")
                o(f"    if (idx < NR && (mem[idx] & {1 << j}u))")
                o(f"      {S}::template prefetch<SAFE>(pre{j}[k], R{j}, sh, r0 + idx / RWW, r1 + idx % RWW, 0);")
                o("  }")
        o("  for (int it = 0; it < TH.iters; ++it) {")
        for j, s in enumerate(L.stmts):
            S = f"xg_s{s.stmt.sid}_t<DV>"
            bit = f"{1 << j}u"
            if s.stmt.klass != 0:
                # boundary statement: its listed cells into the spare copy, then back
                o(f"    if (any & {bit}) {{")
                o("      T* cur = buf + cb;")
                o("      T* spare = buf + (NR - cb);")
                o("      uint8_t* vcur = vl + (CHECK ? cb : 0);")
                o("      uint8_t* vsp = vl + (CHECK ? NR - cb : 0);")
                o("      for (int e = threadIdx.x; e < ns; e += TBT) {")
                o("        const int idx = slist[e];")
                o(f"        if (mem[idx] & {bit}) {{")
                o("          const int lr = idx / RWW, lc = idx % RWW;")
                o("          if (lr >= RE && lr < RHH - RE && lc >= RE && lc < RWW - RE) {")
                o("            bool ok = true;")
                o(f"            spare[idx] = {S}::template value_tb<SAFE, CHECK, RHH, RWW>(cur, vcur, ok, lr, lc, "
                  f"R{j}, P{j}, sh, r0 + lr, r1 + lc, 0);")
                o("            if (CHECK) vsp[idx] = ok ? 1 : 0;")
                o("          } else if (CHECK) {")
                o("            vsp[idx] = 0;          /* too near the region edge: stays stale */")
                o("            spare[idx] = cur[idx];")
                o("          } else {")
                o("            spare[idx] = cur[idx];")
                o("          }")
                o("        }")
                o("      }")
                o("      __syncthreads();")
                o("      for (int e = threadIdx.x; e < ns; e += TBT) {")
                o("        const int idx = slist[e];")
                o(f"        if (mem[idx] & {bit}) {{")
                o("          cur[idx] = spare[idx];")
                o("          if (CHECK) vcur[idx] = vsp[idx];")
                o("        }")
                o("      }")
                o("      __syncthreads();")
                o("    }")
                continue
            o(f"    if (any & {bit}) {{")
            o("      const T* cur = buf + cb;")
            o("      T* nxt = buf + (NR - cb);")
            o("      const uint8_t* vcur = vl + (CHECK ? cb : 0);")
            o("      uint8_t* vnxt = vl + (CHECK ? NR - cb : 0);")
            # after `it` iterations only cells (it + 1) * reach inside the region can still be
            # exact: nearer ones are left stale (the validation variant marks them)
            o("      const int lo = RE * (it + 1) < RHH / 2 ? RE * (it + 1) : RHH / 2;")
            pre = s.stmt.sid in self._prefetched
            if pre:
                o("#pragma unroll")
                o("      for (int k = 0; k < NC; ++k) {")
                o("        const int idx = threadIdx.x + TBT * k;")
                o("        if (idx >= NR) continue;")
            else:
                o("#pragma unroll 2")
                o("      for (int idx = threadIdx.x; idx < NR; idx += TBT) {")
            o("        const int lr = idx / RWW, lc = idx % RWW;")
            o(f"        const bool mine = (mem[idx] & {bit}) != 0;")
            o("        const bool act = lr >= lo && lr < RHH - lo && lc >= lo && lc < RWW - lo;")
            o("        if (mine && act) {")
            o("          bool ok = true;")
            if pre:
                o(f"          nxt[idx] = {S}::template value_pre<CHECK, RHH, RWW>(cur, vcur, ok, lr, lc, R{j}, P{j}, "
                  f"pre{j}[k]);")
            else:
                o(f"          nxt[idx] = {S}::template value_tb<SAFE, CHECK, RHH, RWW>(cur, vcur, ok, lr, lc, R{j}, "
                  f"P{j}, sh, r0 + lr, r1 + lc, 0);")
            o("          if (CHECK) vnxt[idx] = ok ? 1 : 0;")
            o("        } else {")
            o("          nxt[idx] = cur[idx];")
            o("          if (CHECK) vnxt[idx] = mine ? 0 : vcur[idx];   /* a stale member is not exact */")
            o("        }")
            o("      }")
            o("      __syncthreads();")
            o("      cb = NR - cb;")
            o("    }")
        o("  }")
        o("  };")
        self._tb_iterate_call(L, o)
        o("  const T* lds = buf + cb;")
        o("  const uint8_t* vld = vl + (CHECK ? cb : 0);")
        o("  const xg::gptr<T> __restrict__ dst = xg::gp<T>(TH.dst.ptr);")
        o("  constexpr int NCORE = (C0 * C1 + TBT - 1) / TBT;")
        o("#pragma unroll")
        o("  for (int n = 0; n < NCORE; ++n) {")
        o("    const int q = threadIdx.x + TBT * n;")
        o("    const int i0 = g0 + q / C1, i1 = g1 + q % C1;")
        o("    if (q < C0 * C1 && i0 < N0 && i1 < N1) {")
        o("      const int li = (HL + q / C1) * RWW + HL + q % C1;")
        o("      dst[(int64_t)i0 * N1 + i1] = lds[li];")
        o("      if (CHECK && !vld[li]) *xg::gp<int32_t>(TH.fail) = 1;")
        o("    }")
        o("  }")
        o("}")
        kargs = ["const xg_tb_hdr TH"] + [f"const xg_r{s.stmt.sid} R{j}, const xg_p{s.stmt.sid} P{j}"
                                          for j, s in enumerate(L.stmts)]
        call = ", ".join(["TH"] + [f"R{j}, P{j}" for j in range(m)] + ["(int)xg::xcd_remap(blockIdx.x, gridDim.x)"])
        o(f'extern "C" __global__ void {self._tb_bounds(L)} {L.kernel}({", ".join(kargs)}) '
          f"{{ {L.kernel}_body<false, true>({call}); }}")
        o(f'extern "C" __global__ void {self._tb_bounds(L, True)} {L.kernel}_chk({", ".join(kargs)}) '
          f"{{ {L.kernel}_body<true, true>({call}); }}")
        if L.batched:
            self._tb_batched(L, out)

    def _tb_iterate_call(self, L: LoopMeta, o) -> None:
        """Run a blocked body's iterations in the instantiation for this workgroup:
        safe (no clamped loads) or not, and with the reciprocal form of uniform
        divisions when every statement's divisors allow it."""
        conds = [f"xg_s{s.stmt.sid}::dv_ok(R{j}, P{j})" for j, s in enumerate(L.stmts) if self._dv.get(s.stmt.sid)]
        t, f = "xg::bconst<true>()", "xg::bconst<false>()"
        if not conds:
            o(f"  if (safe) iterate({t}, {f}); else iterate({f}, {f});")
            return
        # DVOK: the batched variants leave the reciprocal form out (their instances
        # seldom have power-of-two divisors, and its second instantiation of the
        # iterations made the batched LDS kernel spill 11 registers instead of 3)
        # Edge tiles (clamped loads) take the reciprocal form too: the 2049^2 cavity step
        # measured 0.648 ms against 0.676 ms when they divided (FP32 fields: 0.506
        # against 0.550), although FP32's register budget then spills
        o("  if constexpr (DVOK) {")
        o(f"    const bool dv = {' && '.join(conds)};")
        o(f"    if (safe) {{ if (dv) iterate({t}, {t}); else iterate({t}, {f}); }}")
        o(f"    else {{ if (dv) iterate({f}, {t}); else iterate({f}, {f}); }}")
        o("  } else {")
        o(f"    if (safe) iterate({t}, {f}); else iterate({f}, {f});")
        o("  }")

    @staticmethod
    def _tb_bounds(L: LoopMeta, check: bool = False) -> str:
        """Launch bounds of a blocked kernel, with the register budget set for the
        occupancy its LDS allows (see _tb_waves_per_simd).  The checking variant runs
        once per plan and needs more LDS: no target."""
        if check:
            return f"__launch_bounds__({L.threads})"
        return (f"__launch_bounds__({L.threads}) "
                f"__attribute__((amdgpu_waves_per_eu({Emitter._tb_waves_per_simd(L)})))")

    @staticmethod
    def _tb_waves_per_simd(L: LoopMeta) -> int:
        """Waves per SIMD the register budget is set for: two workgroups per CU, or
        one when the LDS region (160 KiB per CU) leaves room for only one.  Asking
        for more than two made the compiler spill on the 512-work-item shapes."""
        waves = L.threads // 64
        wg = 2
        if L.variant == "lds":
            elem = 8 if device_type_name(L.stmts[0].stmt.gtype.element) == "double" else 4
            nr = (L.core[0] + 2 * L.halo) * (L.core[1] + 2 * L.halo)
            sl = nr // 4 if nr // 4 > 256 else (nr if nr < 256 else 256)
            lds = 2 * nr * elem + nr + 2 * sl + 64       # both region copies, membership bytes, sparse list
            wg = max(1, min(wg, 163840 // lds))
        return max(1, min(8, -(-wg * waves // 4)))

    def _tb_batched(self, L: LoopMeta, out: list):
        """Batched variants of a blocked loop's kernel: the tiles of many instances.
        Each entry's header, views and parameters are copied into locals before the
        body runs (read through the table's generic pointers they could alias the
        body's stores and would be re-read for every cell); the barrier lets the
        workgroup reuse its LDS for its next tile."""
        loc = ["const xg_tb_hdr th = e.th;"]
        for j, s in enumerate(L.stmts):
            sid = s.stmt.sid
            loc.append(f"const xg_r{sid} r{j} = *xg::cp<xg_r{sid}>(e.refs[{j}]);")
            loc.append(f"const xg_p{sid} p{j} = *xg::cp<xg_p{sid}>(e.params[{j}]);")
        bcall = ", ".join(["th"] + [f"r{j}, p{j}" for j in range(len(L.stmts))] + ["(int)t"])
        for suffix, chk in (("_b", "false"), ("_chkb", "true")):
            out.append(f'extern "C" __global__ void {self._tb_bounds(L, chk == "true")} {L.kernel}{suffix}('
                       f"const xg_btent* __restrict__ E, const int32_t n, const int64_t total) "
                       f"{{ xg::for_batch_tiles(E, n, total, [&](const xg_btent& e, uint32_t t) "
                       f"{{ {' '.join(loc)} {L.kernel}_body<{chk}, false>({bcall}); __syncthreads(); }}); }}")

    def _tb_kernel_reg(self, L: LoopMeta, out: list):
        """Blocked kernel of one loop.  A workgroup owns a core tile of the loop grid
        and keeps core + halo (a region RH rows x 64 columns) in registers for up to
        `tmax` iterations of the loop body: lane c of wave w holds column c of the
        wave's band of BR rows, plus `reach` rows above and below.  A read at
        column offset o comes from lane c + o (a cross-lane DPP move); a read at row
        offset o from the lane's own registers, after the bands' edge rows have been
        exchanged through LDS (once per statement that reads other rows).  Each
        statement is evaluated for every region cell from the values before it, then
        stored where it applies (its class, inside the shrinking exact window), so a
        statement sees exactly the values the statements before it produced.  The
        checking variant carries an exactness bit per cell the same way."""
        T = device_type_name(L.stmts[0].stmt.gtype.element)
        c0, c1 = L.core
        m = len(L.stmts)
        RE = L.reach
        params = ["const xg_tb_hdr& TH"] + [f"const xg_r{s.stmt.sid}& R{j}, const xg_p{s.stmt.sid}& P{j}"
                                            for j, s in enumerate(L.stmts)]
        # reach of the loads that come from global memory (other grids / levels)
        rl, rh = [0, 0], [0, 0]
        for s_ in L.stmts:
            for n in ir.walk(s_.stmt.node.value):
                if isinstance(n, ir.Cell) and not (n.grid.name == L.grid and n.level == 0):
                    for d in range(2):
                        rl[d] = max(rl[d], -n.offsets[d])
                        rh[d] = max(rh[d], n.offsets[d])
        o = out.append
        o("template <bool CHECK, bool DVOK>")
        o(f"__device__ __forceinline__ void {L.kernel}_body({', '.join(params)}, const int tile) {{")
        o(f"  typedef {T} T;")
        o(f"  constexpr int TBT = {L.threads}, WV = TBT / 64;")
        o(f"  constexpr int C0 = {c0}, C1 = {c1}, HL = {L.halo}, RH = C0 + 2 * HL, RW = 64, RE = {RE};")
        o("  constexpr int BR = RH / WV, NQ = BR + 2 * RE;")
        o('  static_assert(C1 + 2 * HL == RW && RH % WV == 0, "region: 64 columns, whole rows per wave");')
        # band-edge rows of every wave, double buffered (alternate exchanges use alternate halves)
        o("  __shared__ T ex[2][WV][2][RE][64];")
        o("  __shared__ uint8_t exv[CHECK ? 2 : 1][WV][2][RE][64];")
        o("  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;")
        o("  const int32_t* sh = TH.src.shape;")
        o("  const int N0 = sh[0], N1 = sh[1];")
        o("  const int nt1 = (N1 + C1 - 1) / C1;")
        o("  const int g0 = (tile / nt1) * C0, g1 = (tile % nt1) * C1;")
        o("  const int r0 = g0 - HL, cg = g1 - HL + lane;   /* region origin row; this lane's grid column */")
        o("  const int br = w * BR;                          /* the band's first region row */")
        o("  const xg::gptr<const T> __restrict__ src = xg::gp<const T>(TH.src.ptr);")
        o("  T V[NQ];")
        o("  uint8_t OK[CHECK ? NQ : 1];")
        o("  uint32_t MB[(BR + 3) / 4];          /* statement-membership bits: 8 per row */")
        o("#pragma unroll")
        o("  for (int q = 0; q < (BR + 3) / 4; ++q) MB[q] = 0u;")
        o("  uint32_t mine_any = 0u;")
        o("#pragma unroll")
        o("  for (int q = 0; q < NQ; ++q) {")
        o("    const int gr = r0 + br + q - RE;")
        o("    const bool in = gr >= 0 && gr < N0 && cg >= 0 && cg < N1;")
        o("    const int64_t c = (int64_t)gr * N1 + cg;")
        o("    V[q] = in ? src[c] : T(0);")
        o("    if (CHECK) OK[q] = in ? 1 : 0;")
        o("    if (q >= RE && q < RE + BR) {")
        o("      uint32_t mb = 0u;")
        o("      if (in) {")
        o("        const int32_t k = TH.mask ? xg::gp<const int32_t>(TH.mask)[c] : 0;")
        o("        mb = " + " | ".join(f"(k == TH.klass[{j}] ? {1 << j}u : 0u)" for j in range(m)) + ";")
        o("      }")
        o("      MB[(q - RE) >> 2] |= mb << (8 * ((q - RE) & 3));")
        o("      mine_any |= mb;")
        o("    }")
        o("  }")
        o("  const uint32_t any = xg::block_or(mine_any);")
        o(f"  const bool safe = TH.uniform && r0 - {rl[0]} >= 0 && r0 + RH + {rh[0]} <= N0 && "
          f"g1 - HL - {rl[1]} >= 0 && g1 - HL + RW + {rh[1]} <= N1;")
        o("  int xb = 0;                 /* exchange buffer half */")
        # refresh the reach rows above and below the band from the neighbouring bands
        o("  auto exchange = [&]() {")
        o("#pragma unroll")
        o("    for (int k = 0; k < RE; ++k) {")
        o("      ex[xb][w][0][k][lane] = V[RE + k];")
        o("      ex[xb][w][1][k][lane] = V[BR + k];")
        o("      if (CHECK) { exv[CHECK ? xb : 0][w][0][k][lane] = OK[RE + k]; "
          "exv[CHECK ? xb : 0][w][1][k][lane] = OK[BR + k]; }")
        o("    }")
        o("    __syncthreads();")
        o("#pragma unroll")
        o("    for (int k = 0; k < RE; ++k) {")
        o("      if (w > 0) { V[k] = ex[xb][w - 1][1][k][lane]; if (CHECK) OK[k] = exv[CHECK ? xb : 0][w - 1][1][k][lane]; }")
        o("      if (w < WV - 1) { V[RE + BR + k] = ex[xb][w + 1][0][k][lane]; "
          "if (CHECK) OK[RE + BR + k] = exv[CHECK ? xb : 0][w + 1][0][k][lane]; }")
        o("    }")
        o("    xb ^= 1;")
        o("  };")
        o("  auto iterate = [&](auto safe_c, auto dv_c) {")
        o("  constexpr bool SAFE = decltype(safe_c)::value;")
        o("  constexpr bool DV = decltype(dv_c)::value;   /* reciprocal form of uniform divisions */")
        o("  (void)DV;")
        for j, s_ in enumerate(L.stmts):
            if s_.stmt.sid in self._prefetched:
                S = f"xg_s{s_.stmt.sid}_t<DV>"
                o(f"  typename {S}::Pre pre{j}[BR];")
                o("#pragma unroll")
                o("  for (int q = 0; q < BR; ++q)")
                o(f"    if ((MB[q >> 2] >> (8 * (q & 3))) & {1 << j}u) "
                  f"{S}::template prefetch<SAFE>(pre{j}[q], R{j}, sh, r0 + br + q, cg, 0);")
        o("  for (int it = 0; it < TH.iters; ++it) {")
        for j, s_ in enumerate(L.stmts):
            S = f"xg_s{s_.stmt.sid}_t<DV>"
            bit = f"{1 << j}u"
            reads = self._tb_reads[s_.stmt.sid]
            vertical = any(r[0] != 0 for r in reads)
            o(f"    if (any & {bit}) {{")
            if vertical:
                o("      exchange();")
            if s_.stmt.klass == 0:
                # after `it` iterations only cells (it + 1) * reach inside the region can still be exact
                o("      const int lo = RE * (it + 1) < RH / 2 ? RE * (it + 1) : RH / 2;")
            else:
                o("      const int lo = RE;")
            # new rows wait `reach` rows before they replace the old ones (later rows still read them)
            o("      T hold[RE];")
            o("      uint8_t hok[CHECK ? RE : 1];")
            o("#pragma unroll")
            o("      for (int q = 0; q < BR; ++q) {")
            o("        const int lr = br + q;")
            o(f"        const bool mine = ((MB[q >> 2] >> (8 * (q & 3))) & {bit}) != 0;")
            o("        const bool act = lr >= lo && lr < RH - lo && lane >= lo && lane < RW - lo;")
            # every lane gathers its reads unconditionally (cross-lane moves need the whole wave)
            o(f"        T s[{S}::NRD];")
            o("        uint32_t okr = 1u;")
            for k, (d0, d1) in enumerate(reads):
                o(f"        s[{k}] = xg::lane_shift<{d1}>(V[q + RE + ({d0})]);")
                o(f"        if (CHECK) okr &= xg::lane_shift_u32<{d1}>((uint32_t)OK[q + RE + ({d0})]);")
            if not reads:
                o("        s[0] = T(0);")
            if s_.stmt.sid in self._prefetched:
                o(f"        const T val = {S}::value_rgp(s, R{j}, P{j}, pre{j}[q]);")
            else:
                o(f"        const T val = {S}::template value_rg<SAFE>(s, R{j}, P{j}, sh, r0 + lr, cg, 0);")
            o("        const T nv = (mine && act) ? val : V[q + RE];")
            o("        const uint8_t nok = CHECK ? ((mine && act) ? (uint8_t)(okr != 0u) : (mine ? (uint8_t)0 : OK[q + RE]))"
              " : (uint8_t)0;")
            o("        if (q >= RE) {")
            o("          V[q] = hold[q % RE];")
            o("          if (CHECK) OK[q] = hok[CHECK ? q % RE : 0];")
            o("        }")
            o("        hold[q % RE] = nv;")
            o("        if (CHECK) hok[CHECK ? q % RE : 0] = nok;")
            o("      }")
            o("#pragma unroll")
            o("      for (int q = BR - RE; q < BR; ++q) {")
            o("        V[q + RE] = hold[q % RE];")
            o("        if (CHECK) OK[q + RE] = hok[CHECK ? q % RE : 0];")
            o("      }")
            o("    }")
        o("  }")
        o("  };")
        self._tb_iterate_call(L, o)
        o("  const xg::gptr<T> __restrict__ dst = xg::gp<T>(TH.dst.ptr);")
        o("#pragma unroll")
        o("  for (int q = 0; q < BR; ++q) {")
        o("    const int lr = br + q, gr = r0 + lr;")
        o("    if (lr >= HL && lr < HL + C0 && lane >= HL && lane < HL + C1 && gr < N0 && cg < N1) {")
        o("      dst[(int64_t)gr * N1 + cg] = V[q + RE];")
        o("      if (CHECK && !OK[q + RE]) *xg::gp<int32_t>(TH.fail) = 1;")
        o("    }")
        o("  }")
        o("}")
        kargs = ["const xg_tb_hdr TH"] + [f"const xg_r{s_.stmt.sid} R{j}, const xg_p{s_.stmt.sid} P{j}"
                                          for j, s_ in enumerate(L.stmts)]
        call = ", ".join(["TH"] + [f"R{j}, P{j}" for j in range(m)] + ["(int)xg::xcd_remap(blockIdx.x, gridDim.x)"])
        o(f'extern "C" __global__ void {self._tb_bounds(L)} {L.kernel}({", ".join(kargs)}) '
          f"{{ {L.kernel}_body<false, true>({call}); }}")
        o(f'extern "C" __global__ void {self._tb_bounds(L, True)} {L.kernel}_chk({", ".join(kargs)}) '
          f"{{ {L.kernel}_body<true, true>({call}); }}")
        if L.batched:
            self._tb_batched(L, out)

    def device_source(self, stmt_meta: list, groups: list, loops: list = ()) -> str:
        out = ["// generated by xgrid_amd for the MI355X backend", '#include "xgrid_device.h"',
               f"#define XG_OVERSTEP {OVERSTEP_CODE[self.overstep]}"]
        loop_of = {m.stmt.sid: L for L in loops for m in L.stmts}
        for s in self._structs():
            fields_ = " ".join(f"{device_type_name(t)} {n};" for n, t in s.elements)
            out.append(f"struct {s.name} {{ {fields_} }};")
        for op in self.device_functions:
            f = op.ir
            args = ", ".join(f"{device_type_name(t)} {n}" for n, t in f.args)
            out.append(f"__device__ {device_type_name(f.ret)} {function_symbol(op)}({args});")
        for op in self.device_functions:
            self._device_function(op, out)
        for m in stmt_meta:
            self._device_stmt(m.stmt, m.refs, m.params, out, loop_of.get(m.stmt.sid), m.fwd_from, m)
        for m, nxt in zip(stmt_meta, stmt_meta[1:]):
            if m.pair:
                a, b = m.stmt.sid, nxt.stmt.sid
                fw = 'true' if m.pair_fwd else 'false'
                run = (f"xg::run_pair<xg_s{a}{{s}}, xg_s{b}{{s}}, XG_OVERSTEP, {fw}>(HA, RA, PA, HB, RB, PB, "
                       f"xg::xcd_remap(blockIdx.x, gridDim.x));")
                conds = [f"xg_s{x}::dv_ok(R{y}, P{y})" for x, y in ((a, "A"), (b, "B")) if self._dv.get(x)]
                body = (f"if ({' && '.join(conds)}) {run.format(s='_r')} else {run.format(s='')}" if conds
                        else run.format(s=""))
                out.append(f'extern "C" __global__ void __launch_bounds__(256) {m.pair}('
                           f"const xg_launch_hdr HA, const xg_r{a} RA, const xg_p{a} PA, "
                           f"const xg_launch_hdr HB, const xg_r{b} RB, const xg_p{b} PB) "
                           f"{{ {body} }}")
        for L in loops:
            for V in (L, L.alt):
                if V is not None:
                    (self._tb_kernel_reg if V.variant == "reg" else self._tb_kernel_lds)(V, out)
        # phase-2 kernels per (element type, dimension)
        done = set()
        for m in stmt_meta:
            T = device_type_name(m.stmt.gtype.element)
            key = (T, m.stmt.dimension)
            if key in done:
                continue
            done.add(key)
            ph2 = f"xg::run_phase2<{T}, {m.stmt.dimension}, XG_OVERSTEP>"
            out.append(f'extern "C" __global__ void __launch_bounds__(256) {m.phase2}(const xg_launch_hdr H) '
                       f"{{ {ph2}(H, xg::xcd_remap(blockIdx.x, gridDim.x), gridDim.x); }}")
            out.append(f'extern "C" __global__ void __launch_bounds__(256) {m.phase2}_b('
                       f"const xg_bent* __restrict__ E, const int32_t n, const int64_t total) "
                       f"{{ xg::for_batch_tiles(E, n, total, [&](const xg_bent& e, uint32_t t) "
                       f"{{ const xg_launch_hdr h = e.hdr; {ph2}(h, t, e.ntiles); }}); }}")
        for g in groups:
            params = ["const uint32_t en"]
            body = []
            for pos in range(g.count):
                m = stmt_meta[g.first + pos]
                sid = m.stmt.sid
                params.append(f"const xg_launch_hdr H{pos}, const xg_r{sid} R{pos}, const xg_p{sid} P{pos}")
                if m.stmt.mode == "inplace":
                    body.append(f"  if (en & {1 << pos}u) {{ xg::run_block<xg_s{sid}, XG_OVERSTEP>(H{pos}, R{pos}, P{pos}); __syncthreads(); }}")
            out.append(f'extern "C" __global__ void __launch_bounds__(1024) {g.kernel}({", ".join(params)}) {{')
            out.extend(body)
            out.append("}")
            # batched: workgroup i chains the statements of instance i (entries E[i * count + pos];
            # an empty domain skips a statement)
            out.append(f'extern "C" __global__ void __launch_bounds__(1024) {g.kernel}_b('
                       f"const xg_bent* __restrict__ E, const int32_t count) {{")
            out.append("  const xg_bent* I = E + (int64_t)blockIdx.x * count;")
            for pos in range(g.count):
                m = stmt_meta[g.first + pos]
                sid = m.stmt.sid
                if m.stmt.mode == "inplace":
                    out.append(f"  if (I[{pos}].hdr.dom.kind != XG_DOM_EMPTY) {{ const xg_launch_hdr h = I[{pos}].hdr; "
                               f"const xg_r{sid} r = *xg::cp<xg_r{sid}>(I[{pos}].refs); "
                               f"const xg_p{sid} p = *xg::cp<xg_p{sid}>(I[{pos}].params); "
                               f"xg::run_block<xg_s{sid}, XG_OVERSTEP>(h, r, p); __syncthreads(); }}")
            out.append("}")
        return "\n".join(out) + "\n"

    # ------------------------------------------------------------- statements (C text)
    def _stmts_text(self, body, out, expr, depth, device=False, stencil=None):
        ind = "  " * depth
        for s in body:
            if self.comment and not device:
                out.append(f'#line {s.loc.line} "{s.loc.file}"')
            if isinstance(s, ir.Assign):
                if isinstance(s.target, ir.Cell):
                    if stencil is None:
                        raise EmitError(f"{s.loc}: grid store outside a kernel")
                    out.append(ind + stencil(s))
                else:
                    out.append(f"{ind}{expr(s.target)} = {expr(s.value)};")
            elif isinstance(s, ir.Eval):
                out.append(f"{ind}{expr(s.value)};")
            elif isinstance(s, ir.If):
                out.append(f"{ind}if ({expr(s.cond)}) {{")
                self._stmts_text(s.body, out, expr, depth + 1, device, stencil)
                if s.orelse:
                    out.append(f"{ind}}} else {{")
                    self._stmts_text(s.orelse, out, expr, depth + 1, device, stencil)
                out.append(f"{ind}}}")
            elif isinstance(s, ir.While):
                out.append(f"{ind}while ({expr(s.cond)}) {{")
                self._stmts_text(s.body, out, expr, depth + 1, device, stencil)
                out.append(f"{ind}}}")
            elif isinstance(s, ir.For) and not device and s in getattr(self, "_stub_loops", {}):
                # offer the whole loop to the runtime (blocked); run it here if declined
                L = self._stub_loops[s]
                v = s.var.name
                ct = ctype_name(s.var.type)
                out.append(f"{ind}{{")
                out.append(f"{ind}  {ct} xg_s = {expr(s.start)}, xg_e = {expr(s.stop)}, xg_st = {expr(s.step)};")
                out.append(f"{ind}  int64_t xg_n = (xg_st > 0 && xg_e > xg_s) ? (((int64_t)xg_e - (int64_t)xg_s + "
                           f"(int64_t)xg_st - 1) / (int64_t)xg_st) : 0;")
                for j, m in enumerate(L.stmts):
                    out.append(f"{ind}  struct xg_p{m.stmt.sid} xg_L{j}; memset(&xg_L{j}, 0, sizeof xg_L{j}); "
                               + self._param_sets(m, f"xg_L{j}"))
                out.append(f"{ind}  const void* xg_LP[{len(L.stmts)}] = {{{', '.join(f'&xg_L{j}' for j in range(len(L.stmts)))}}};")
                out.append(f"{ind}  if (xg_st > 0 && xg_A->loop(xg_K, {L.lid}, xg_n, xg_LP, {len(L.stmts)})) {{")
                out.append(f"{ind}    {v} = xg_n > 0 ? ({ct})(xg_s + ({ct})xg_n * xg_st) : xg_s;")
                out.append(f"{ind}  }} else {{")
                out.append(f"{ind}    for ({v} = xg_s; {v} < {expr(s.stop)}; {v} += {expr(s.step)}) {{")
                self._stmts_text(s.body, out, expr, depth + 3, device, stencil)
                out.append(f"{ind}    }}")
                out.append(f"{ind}  }}")
                out.append(f"{ind}}}")
            elif isinstance(s, ir.For):
                v = s.var.name
                out.append(f"{ind}for ({v} = {expr(s.start)}; {v} < {expr(s.stop)}; {v} += {expr(s.step)}) {{")
                self._stmts_text(s.body, out, expr, depth + 1, device, stencil)
                out.append(f"{ind}}}")
            elif isinstance(s, ir.Return):
                out.append(f"{ind}return;" if s.value is None else f"{ind}return {expr(s.value)};")
            elif isinstance(s, ir.Break):
                out.append(f"{ind}break;")
            elif isinstance(s, ir.Continue):
                out.append(f"{ind}continue;")
            elif isinstance(s, ir.Inline):
                if device:
                    raise EmitError(f"{s.loc}: inline C cannot run on the device")
                out.append(s.source)
            else:
                raise EmitError(f"cannot lower statement {type(s).__name__}")

    # ------------------------------------------------------------- host stub
    @staticmethod
    def _param_sets(m, var: str) -> str:
        """C assignments filling a statement's parameter block `var`."""
        sets = []
        for v in m.params:
            src = f"(*{v.name})" if isinstance(v.type, Pointer) else v.name
            sets.append(f"{var}.{v.name} = {src};")
        for h in m.hoisted:
            sets.append(f"{var}.{h.name} = {h.host};")
        return " ".join(sets)

    def stub_source(self, stmt_meta: list, loops: list = ()) -> str:
        f = self.func
        self._stub_loops = {L.node: L for L in loops}
        out = ["/* generated by xgrid_amd: host control flow of one kernel */",
               "#include <stdio.h>", "#include <stdlib.h>", "#include <stdint.h>", "#include <stdbool.h>",
               "#include <string.h>", "#include <math.h>"]
        for inc in f.includes:
            out.append(f'#include "{inc}"')
        out.append("typedef struct xg_api_t { int (*begin)(void*, void* const*, int32_t, const void*, int64_t);"
                   " int (*stmt)(void*, int32_t, const void*, int64_t); int (*end)(void*);"
                   " const int32_t* (*shape)(void*);"
                   " int (*loop)(void*, int32_t, int64_t, const void* const*, int32_t); } xg_api_t;")
        for s in self._structs():
            fields_ = " ".join(f"{ctype_name(t)} {n};" for n, t in s.elements)
            out.append(f"struct {s.name} {{ {fields_} }};")
        host_expr = self._host_expr()
        # user functions (host versions)
        for op in self.functions:
            fi = op.ir
            args = ", ".join(f"{ctype_name(t)} {n}" for n, t in fi.args)
            decl = f"{ctype_name(fi.ret)} {function_symbol(op)}({args})"
            if op.mode != "external":      # externals are declared by their includes
                out.append("static " + decl + ";")
        for op in self.functions:
            if op.mode == "external":
                continue
            fi = op.ir
            args = ", ".join(f"{ctype_name(t)} {n}" for n, t in fi.args)
            out.append(f"static {ctype_name(fi.ret)} {function_symbol(op)}({args}) {{")
            for v in fi.locals:
                out.append(f"  {ctype_name(v.type)} {v.name};")
            for m in op.macro:
                out.append(m)
            self._stmts_text(fi.body, out, self._host_expr(), 1)
            out.append("}")
        for m in stmt_meta:
            out.append(f"struct xg_p{m.stmt.sid} {{ {param_fields(m, ctype_name)} }};")

        def stencil(s: ir.Assign) -> str:
            m = next(mm for mm in stmt_meta if mm.stmt.node is s)
            return ("{ struct xg_p%d xg_P; memset(&xg_P, 0, sizeof xg_P); %s xg_A->stmt(xg_K, %d, &xg_P, sizeof xg_P); }"
                    % (m.stmt.sid, self._param_sets(m, "xg_P"), m.stmt.sid))

        def param_decl(n, t):
            if isinstance(t, GridType):
                return f"void* {n}"
            return f"{ctype_name(t)} {n}"

        params = ", ".join(["void* xg_K", "const xg_api_t* xg_A"] + [param_decl(n, t) for n, t in f.args])
        ret = ctype_name(f.ret)
        out.append(f"static {ret} xg_body({params}) {{")
        out.append("  (void)xg_K; (void)xg_A;")
        for v in f.locals:
            out.append(f"  {ctype_name(v.type)} {v.name};")
        for m in self.macro:
            out.append(m)
        self._stmts_text(f.body, out, host_expr, 1, stencil=stencil)
        out.append("}")
        # entry: key of the scalar arguments, grid handles, begin/body/end
        scal = [(i, n, t) for i, (n, t) in enumerate(f.args) if not isinstance(t, GridType)]
        grids = [n for _, n, _ in f.grid_args]
        out.append(f"{ret} {self.entry}({params}) {{")
        if scal:
            fields_ = " ".join(f"{ctype_name(t.element if isinstance(t, Pointer) else t)} a{i};" for i, n, t in scal)
            out.append(f"  struct {{ {fields_} }} xg_key;")
            out.append("  memset(&xg_key, 0, sizeof xg_key);")
            for i, n, t in scal:
                out.append(f"  xg_key.a{i} = {'(*' + n + ')' if isinstance(t, Pointer) else n};")
            keyp, keyn = "&xg_key", "sizeof xg_key"
        else:
            keyp, keyn = "NULL", "0"
        if grids:
            out.append(f"  void* xg_G[{len(grids)}] = {{{', '.join(grids)}}};")
            gptr = "xg_G"
        else:
            gptr = "NULL"
        out.append(f"  xg_A->begin(xg_K, {gptr}, {len(grids)}, {keyp}, {keyn});")
        call = f"xg_body({', '.join(['xg_K', 'xg_A'] + [n for n, _ in f.args])})"
        if ret == "void":
            out.append(f"  {call};")
            out.append("  xg_A->end(xg_K);")
        else:
            out.append(f"  {ret} xg_r = {call};")
            out.append("  xg_A->end(xg_K);")
            out.append("  return xg_r;")
        out.append("}")
        # the same entry with every argument in one struct (laid out like a ctypes
        # Structure of the same fields): batches call it from C without Python
        fields_ = " ".join(f"{param_decl(f'a{i}', t)};" for i, (n, t) in enumerate(f.args)) or "int32_t xg_unused;"
        out.append(f"struct xg_args {{ {fields_} }};")
        unpack = ", ".join(["xg_K", "xg_A"] + [f"xg_a->a{i}" for i in range(len(f.args))])
        out.append("void xg_entry_packed(void* xg_K, const xg_api_t* xg_A, const void* xg_p) {")
        out.append("  const struct xg_args* xg_a = (const struct xg_args*)xg_p;")
        out.append("  (void)xg_a;")
        out.append(f"  {self.entry}({unpack});")
        out.append("}")
        out.append("size_t xg_args_size(void) { return sizeof(struct xg_args); }")
        return "\n".join(out) + "\n"

    def _host_expr(self, fold_pow: bool = False):
        f = self.func
        grid_names = {n for _, n, _ in f.grid_args}

        def name(e: ir.Name) -> str:
            if e.var.name in grid_names:
                raise EmitError(f"{e.loc}: grid '{e.var.name}' used as a value")
            return f"(*{e.var.name})" if isinstance(e.var.type, Pointer) else e.var.name

        holder = {}

        def grid_info(e: ir.GridInfo) -> str:
            return f"(xg_A->shape({e.grid.name})[{holder['x'](e.dim)}])"

        x = _ExprC(False, None, name, grid_info, fold_pow=fold_pow)
        holder["x"] = x
        return x

    # ------------------------------------------------------------- driver
    def lower(self) -> Lowered:
        stmt_meta = []
        phase2_names = {}
        for st in self.info.stmts:
            refs = self._stmt_refs(st)
            T = device_type_name(st.gtype.element)
            key = (T, st.dimension)
            if key not in phase2_names:
                phase2_names[key] = f"xg_ph2_{len(phase2_names)}"
            stmt_meta.append(StmtMeta(st, refs, list(st.scalars), f"xg_k{st.sid}",
                                      f"xg_k{st.sid}_pp" if st.mode == "pingpong" else None,
                                      phase2_names[key]))
        for m in stmt_meta:
            m.hoisted = self._hoists(m.stmt)
        groups = self._groups(stmt_meta)
        self._pairs(stmt_meta)
        loops = self._blockable_loops(stmt_meta)
        graphs_ok = not any(isinstance(n, ir.Inline) for n in ir.walk_all(self.func.body)) and \
            not any(op.mode == "external" and not is_device_math(op) for op in self.functions)
        plain_src, batch_src = split_batched(self.device_source(stmt_meta, groups, loops))
        return Lowered(plain_src, self.stub_source(stmt_meta, loops), stmt_meta,
                       groups, list(self.func.grid_args),
                       [(i, n, t) for i, (n, t) in enumerate(self.func.args) if not isinstance(t, GridType)],
                       graphs_ok, loops, self.used_shape, batch_src)

    def _hoists(self, st: StencilStmt) -> list:
        """The statement's maximal launch-uniform floating-point subexpressions that
        do some arithmetic (a bare parameter or literal is already a kernel argument),
        grouped by text.  XGRID_HOIST=0 keeps them on the device (for comparisons)."""
        if os.environ.get("XGRID_HOIST", "1") == "0":
            return []
        host = self._host_expr(fold_pow=True)
        found: dict = {}

        def visit(e):
            if not any(isinstance(n, ir.Cell) for n in ir.walk(e)) and not _is_leaf(e) \
                    and _c_type(e) in _FP_RANK and _host_exact(e):
                text = host(e)
                if text not in found:
                    found[text] = Hoist([], _c_type(e), f"xg_h{len(found)}", text)
                found[text].nodes.append(e)
                return
            for c in ir.children(e):
                visit(c)

        visit(st.node.value)
        return list(found.values())

    PREFETCH_MAX = 4   # most other-grid loads of a blocked statement kept in registers per cell

    # 1024 work-items; the region (core + halo 12 on each side) takes as many rows as
    # leave LDS for two workgroups per CU (32 waves): 72 x 64 with 8-byte cells (80.6
    # KB each), 96 x 64 with 4-byte cells
    DEFAULT_BLOCK_SHAPE = (48, 64, 1024, 10, "lds")
    DEFAULT_BLOCK_SHAPE_4B = (72, 64, 1024, 10, "lds")
    # Plain calls on smaller grids take a second tile: 1024 work-items on a 64 x 64
    # region (two workgroups per CU in either precision), whose launches are shorter.
    # Measured per cavity step: it wins up to 724^2 (FP32) and 1024^2 (FP64) cells and
    # loses from 1024^2 (FP32) and 1448^2 (FP64) up.
    SMALL_GRID_SHAPE = (40, 64, 1024, 10, "lds")
    SMALL_GRID_MAX_CELLS = {4: 800 * 800, 8: 1300 * 1300}

    @classmethod
    def block_shape(cls, spec: str | None = None, elem: int = 8) -> tuple:
        """(core rows, region columns, work-items, most iterations per launch, variant)
        of blocked loops.  The region (core + halo) is a power of two wide so a region
        index splits into row and column with shifts; the core is the region width less
        the halo.  Variant "lds" keeps the region in LDS; "reg" keeps one region column
        per lane in registers (64 columns).  XGRID_TB_SHAPE="rows,cols,threads,tmax[,variant]"
        overrides the default, which depends on the loop grid's element size `elem`."""
        spec = spec if spec is not None else os.environ.get("XGRID_TB_SHAPE")
        if not spec:
            return cls.DEFAULT_BLOCK_SHAPE_4B if elem == 4 else cls.DEFAULT_BLOCK_SHAPE
        parts = spec.split(",")
        r, c, t, m = (int(x) for x in parts[:4])
        v = parts[4].strip() if len(parts) > 4 else "lds"
        if v not in ("lds", "reg"):
            raise EmitError(f"unknown blocked-loop variant '{v}'")
        return r, c, t, m, v

    def _blockable_loops(self, stmt_meta: list) -> list:
        """`for` loops that may run temporally blocked: the body is 1..8 stencil
        statements, all storing unshifted to one 2-D grid, none two-phase, none
        capturing the loop variable; integer loop variable; overstep 'none'."""
        if self.overstep != "none":
            return []
        by_node = {m.stmt.node: m for m in stmt_meta}
        loops = []
        for n in ir.walk_all(self.func.body):
            if not isinstance(n, ir.For) or not isinstance(n.var.type, Integer):
                continue
            if not n.body or len(n.body) > 8 or not all(isinstance(s, ir.Assign) and s in by_node for s in n.body):
                continue
            ms = [by_node[s] for s in n.body]
            g = ms[0].stmt.grid
            if any(m.stmt.grid != g or m.stmt.dimension != 2 or m.stmt.mode == "twophase" or any(m.stmt.store_offsets)
                   for m in ms):
                continue
            if any(any(v is n.var for v in m.params) for m in ms):
                continue
            reach = 0
            for m in ms:
                for c in ir.walk(m.stmt.node.value):
                    if isinstance(c, ir.Cell) and c.grid.name == g and c.level == 0:
                        reach = max(reach, max(abs(o) for o in c.offsets))
            elem = 8 if device_type_name(ms[0].stmt.gtype.element) == "double" else 4
            shape = self.shape if self.shape is not None else self.block_shape(elem=elem)
            L = self._loop_meta(len(loops), n, ms, g, reach, shape, f"xg_tb{len(loops)}")
            if L is None:
                continue
            if self.shape is None and not os.environ.get("XGRID_TB_SHAPE"):
                L.alt = self._loop_meta(len(loops), n, ms, g, reach, self.SMALL_GRID_SHAPE, f"xg_tb{len(loops)}_s")
                if L.alt is not None:
                    L.alt.batched = False
                    L.alt_max_cells = self.SMALL_GRID_MAX_CELLS[elem]
            loops.append(L)
            self.used_shape = tuple(shape)
        return loops

    @staticmethod
    def _loop_meta(lid: int, node, ms: list, g: str, reach: int, shape: tuple, name: str):
        """The blocked kernel of one loop for one tile shape, or None if the shape cannot hold it."""
        rows, region_cols, threads, tm, variant = shape
        tmax = max(1, tm // max(1, reach)) if reach else tm
        # information moves `reach` cells per iteration through the bulk; a boundary
        # statement reading a neighbour computed in the same iteration adds a hop,
        # so leave room for two such hops (a validation launch proves the choice)
        halo = tmax * reach + 2 * reach
        cols = region_cols - 2 * halo
        if halo > 16 or cols < 16 or threads % 64 or threads > 1024:
            return None
        if variant == "reg":
            waves = threads // 64
            if region_cols != 64 or reach < 1:
                return None
            rows += (-(rows + 2 * halo)) % waves      # whole region rows per wave
            if (rows + 2 * halo) // waves > 16:
                return None
        return LoopMeta(lid, node, ms, g, (rows, cols), halo, tmax, name, threads, reach, variant)

    def _pairs(self, stmt_meta: list) -> None:
        """Mark statement pairs that may run as one launch: consecutive in-place 1-D or
        2-D statements of one group, storing unshifted to different grids, the first
        not loading the second's store grid at time 0 (the cavity's u and v updates).
        The second may load the first's store at time 0 at its own cell only: that
        value is forwarded in a register (the cavity's b and first pressure sweep).
        The runtime runs a pair together only when both classes are the same box and
        no two arguments alias (`pairable` in xgrid_runtime.cpp)."""
        i = 0
        while i + 1 < len(stmt_meta):
            a, b = stmt_meta[i], stmt_meta[i + 1]
            sa, sb = a.stmt, b.stmt
            ok = (a.group >= 0 and a.group == b.group and b.group_pos == a.group_pos + 1 and
                  sa.mode == "inplace" and sb.mode == "inplace" and
                  sa.dimension == sb.dimension and sa.dimension in (1, 2) and
                  not any(sa.store_offsets) and not any(sb.store_offsets) and sa.grid != sb.grid and
                  (sb.grid, 0) not in a.refs)
            # b may read a's store at time 0 only at its own cell: that value is forwarded
            fwd = ok and (sa.grid, 0) in b.refs
            if fwd:
                ok = all(not any(n.offsets) for n in ir.walk(sb.node.value)
                         if isinstance(n, ir.Cell) and n.grid.name == sa.grid and n.level == 0)
            if ok:
                a.pair = f"xg_f{sa.sid}"
                a.pair_fwd = fwd
                b.fwd_from = sa.grid if fwd else None
                i += 2
            else:
                i += 1

    def _groups(self, stmt_meta: list) -> list:
        """Runs of consecutive stencil statements (same block, nothing between)."""
        by_node = {m.stmt.node: m for m in stmt_meta}
        groups = []

        def scan(body):
            run = []

            def close():
                inplace = [m for m in run if m.stmt.mode == "inplace"]
                if len(run) >= 2 and len(inplace) >= 2 and len(run) <= 32:
                    g = GroupMeta(run[0].stmt.sid, len(run), f"xg_g{len(groups)}")
                    for pos, m in enumerate(run):
                        m.group, m.group_pos = len(groups), pos
                    groups.append(g)
                run.clear()

            for s in body:
                if isinstance(s, ir.Assign) and s in by_node:
                    run.append(by_node[s])
                    continue
                close()
                for child in (getattr(s, "body", None), getattr(s, "orelse", None)):
                    if isinstance(child, list):
                        scan(child)
            close()

        scan(self.func.body)
        return groups
