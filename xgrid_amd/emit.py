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

from dataclasses import dataclass, field

from . import ir
from .analysis import KernelInfo, StencilStmt, captured_value_type, structures_in
from .types import (Boolean, Floating, GridType, Ignore, Integer, Pointer, Structure, Value, Void)


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


@dataclass
class GroupMeta:
    first: int
    count: int
    kernel: str


@dataclass
class Lowered:
    device_src: str
    stub_src: str
    stmts: list
    groups: list
    grid_args: list               # [(arg index, name, GridType)]
    scalar_args: list             # [(arg index, name, type)]
    graphs_ok: bool


class _ExprC:
    """Expression text with C semantics, shared by host and device lowering."""

    def __init__(self, device: bool, cell=None, name=None, grid_info=None):
        self.device = device
        self.cell = cell              # callback for ir.Cell (device statements only)
        self.name = name              # callback for ir.Name
        self.grid_info = grid_info    # callback for ir.GridInfo

    def __call__(self, e) -> str:
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
            if not self.device:
                return f"{'pow' if double else 'powf'}({self(e.left)}, {self(e.right)})"
            fold = _pow_fold(e.right)
            x = f"(({fp})({self(e.left)}))"
            if fold == "sq":
                return f"xg::sq{x}"
            if fold == "id":
                return x
            if fold == "inv":
                return f"(({'1.0' if double else '1.0f'}) / {x})"
            return f"{'pow' if double else 'powf'}({x}, (({fp})({self(e.right)})))"
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
        return f"{function_symbol(e.callee)}({', '.join(args)})"

    def e_GridInfo(self, e):
        if e.info == "dimension":
            return str(e.grid.type.dimension)
        return self.grid_info(e)


def function_symbol(op) -> str:
    return "xgf_" + op.c_name


class Emitter:
    def __init__(self, info: KernelInfo, overstep: str = "none", comment: bool = False,
                 entry: str = "xg_entry", macro: list | None = None):
        self.info = info
        self.macro = list(macro or [])
        self.func = info.func
        self.overstep = overstep
        self.comment = comment
        self.entry = entry
        self.functions = []            # user operators reachable from the kernel (dependency order)
        self.device_functions = []     # those reachable from stencil expressions
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
        for op in self.device_functions:
            if op.mode == "external":
                raise EmitError(f"external function '{op.name}' cannot be evaluated inside a stencil statement")
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

    def _device_stmt(self, st: StencilStmt, refs: list, params: list, out: list):
        sid = st.sid
        D = st.dimension
        T = device_type_name(st.gtype.element)
        grids = {n: t for _, n, t in self.func.grid_args}

        def cell(e: ir.Cell) -> str:
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

        expr = _ExprC(True, cell, name, grid_info)(st.node.value)
        so = list(st.store_offsets) + [0] * (3 - len(st.store_offsets))
        rl, rh = [0, 0, 0], [0, 0, 0]
        for n in ir.walk(st.node.value):
            if isinstance(n, ir.Cell):
                for d, o in enumerate(n.offsets):
                    rl[d] = max(rl[d], -o)
                    rh[d] = max(rh[d], o)
        nr = max(1, len(refs))
        out.append(f"// statement {sid}: {st.node.loc}")
        if params:
            fields_ = " ".join(f"{device_type_name(captured_value_type(v))} {v.name};" for v in params)
            out.append(f"struct xg_p{sid} {{ {fields_} }};")
        else:
            out.append(f"struct xg_p{sid} {{ int32_t xg_unused; }};")
        out.append(f"struct xg_r{sid} {{ xg_gview v[{nr}]; }};")
        out.append(f"struct xg_s{sid} {{")
        out.append(f"  typedef {T} T; typedef xg_r{sid} Refs; typedef xg_p{sid} Params;")
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
        out.append("};")
        args = f"const xg_launch_hdr H, const xg_r{sid} R, const xg_p{sid} P"
        out.append(f'extern "C" __global__ void __launch_bounds__(256) xg_k{sid}({args}) '
                   f"{{ xg::run_main<xg_s{sid}, XG_OVERSTEP>(H, R, P); }}")
        if st.mode == "pingpong":
            out.append(f'extern "C" __global__ void __launch_bounds__(256) xg_k{sid}_pp({args}) '
                       f"{{ xg::run_pingpong<xg_s{sid}, XG_OVERSTEP>(H, R, P); }}")

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

    def device_source(self, stmt_meta: list, groups: list) -> str:
        out = ["// generated by xgrid_amd for the MI355X backend", '#include "xgrid_device.h"',
               f"#define XG_OVERSTEP {OVERSTEP_CODE[self.overstep]}"]
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
            self._device_stmt(m.stmt, m.refs, m.params, out)
        # phase-2 kernels per (element type, dimension)
        done = set()
        for m in stmt_meta:
            T = device_type_name(m.stmt.gtype.element)
            key = (T, m.stmt.dimension)
            if key in done:
                continue
            done.add(key)
            out.append(f'extern "C" __global__ void __launch_bounds__(256) {m.phase2}(const xg_launch_hdr H) '
                       f"{{ xg::run_phase2<{T}, {m.stmt.dimension}, XG_OVERSTEP>(H); }}")
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
    def stub_source(self, stmt_meta: list) -> str:
        f = self.func
        out = ["/* generated by xgrid_amd: host control flow of one kernel */",
               "#include <stdio.h>", "#include <stdlib.h>", "#include <stdint.h>", "#include <stdbool.h>",
               "#include <string.h>", "#include <math.h>"]
        for inc in f.includes:
            out.append(f'#include "{inc}"')
        out.append("typedef struct xg_api_t { int (*begin)(void*, void* const*, int32_t, const void*, int64_t);"
                   " int (*stmt)(void*, int32_t, const void*, int64_t); int (*end)(void*);"
                   " const int32_t* (*shape)(void*); } xg_api_t;")
        for s in self._structs():
            fields_ = " ".join(f"{ctype_name(t)} {n};" for n, t in s.elements)
            out.append(f"struct {s.name} {{ {fields_} }};")
        host_expr = self._host_expr()
        # user functions (host versions)
        for op in self.functions:
            fi = op.ir
            args = ", ".join(f"{ctype_name(t)} {n}" for n, t in fi.args)
            decl = f"{ctype_name(fi.ret)} {function_symbol(op)}({args})"
            out.append(("extern " if op.mode == "external" else "static ") + decl + ";")
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
            sid = m.stmt.sid
            if m.params:
                fields_ = " ".join(f"{ctype_name(captured_value_type(v))} {v.name};" for v in m.params)
                out.append(f"struct xg_p{sid} {{ {fields_} }};")
            else:
                out.append(f"struct xg_p{sid} {{ int32_t xg_unused; }};")

        def stencil(s: ir.Assign) -> str:
            m = next(mm for mm in stmt_meta if mm.stmt.node is s)
            sets = []
            for v in m.params:
                src = f"(*{v.name})" if isinstance(v.type, Pointer) else v.name
                sets.append(f"xg_P.{v.name} = {src};")
            return ("{ struct xg_p%d xg_P; memset(&xg_P, 0, sizeof xg_P); %s xg_A->stmt(xg_K, %d, &xg_P, sizeof xg_P); }"
                    % (m.stmt.sid, " ".join(sets), m.stmt.sid))

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
        return "\n".join(out) + "\n"

    def _host_expr(self):
        f = self.func
        grid_names = {n for _, n, _ in f.grid_args}

        def name(e: ir.Name) -> str:
            if e.var.name in grid_names:
                raise EmitError(f"{e.loc}: grid '{e.var.name}' used as a value")
            return f"(*{e.var.name})" if isinstance(e.var.type, Pointer) else e.var.name

        holder = {}

        def grid_info(e: ir.GridInfo) -> str:
            return f"(xg_A->shape({e.grid.name})[{holder['x'](e.dim)}])"

        x = _ExprC(False, None, name, grid_info)
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
        groups = self._groups(stmt_meta)
        graphs_ok = not any(isinstance(n, ir.Inline) for n in ir.walk_all(self.func.body)) and \
            not any(op.mode == "external" for op in self.functions)
        return Lowered(self.device_source(stmt_meta, groups), self.stub_source(stmt_meta), stmt_meta, groups,
                       list(self.func.grid_args),
                       [(i, n, t) for i, (n, t) in enumerate(self.func.args) if not isinstance(t, GridType)],
                       graphs_ok)

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
