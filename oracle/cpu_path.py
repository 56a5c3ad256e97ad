"""Restatement of the reference's CPU path (ORACLE - test infrastructure only).

Given a kernel parsed by ``xgrid_amd`` (the IR, not the device lowering), this
module writes the C99/OpenMP translation unit that the reference's
``xgrid/lang/generator.py`` writes, compiles it with gcc and calls it through
ctypes with the reference's by-value grid structs (``xgrid/xgrid/__init__.py:60-68``):

* grid struct ``{int32 time; int32 shape[D]; T** data; int32* boundary_mask}``
  and its ``_at`` accessor with the ``none``/``limit``/``wrap`` overstep modes
  (``generator.py:144-182``);
* one ``#pragma omp parallel for collapse(D)`` nest per stencil statement over
  the store grid's shape, gated by ``mask == k`` (``generator.py:289-304,353-361``);
* implicit statements as the two-phase nest with a malloc'd temporary, an
  OpenMP barrier and a copy back (``generator.py:312-352``);
* scalar statements and control flow on the calling thread
  (``generator.py:251-283,362-374``), ``pow``/``powf`` for ``**``
  (``generator.py:377-383``), constants as ``repr(v).lower()`` (``:393-394``).

``index="reference"`` uses the reference's linear index
``sum_i s_{D-1-i} * (i == 0 ? 1 : shape[i-1])`` (``generator.py:171-179,295-296``);
``index="rowmajor"`` (default) uses true row-major strides, which is what the
device backend implements.  The two agree for 1-D and square 2-D grids.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
import tempfile

import numpy as np

from xgrid_amd import ir
from xgrid_amd.types import (Boolean, Floating, GridType, Integer, Pointer, Structure, Value, Void,
                             numpy_dtype, parse_annotation)

CACHE = os.path.join(tempfile.gettempdir(), f"xgrid_oracle_{os.getuid()}")


class OracleError(RuntimeError):
    pass


# ---------------------------------------------------------------- host grid
class HostGrid:
    """numpy time levels + int32 mask, rotated like xgrid/xgrid/__init__.py:43-54."""

    def __init__(self, shape, dtype):
        el = parse_annotation(dtype)
        if not isinstance(el, Value):
            raise OracleError(f"Grid element should be value instead of '{el}'")
        self.element = el
        self.numpy_dtype = numpy_dtype(el)
        self.shape = tuple(int(s) for s in shape)
        self._data = [np.zeros(self.shape, dtype=self.numpy_dtype)]
        self.boundary = np.zeros(self.shape, dtype=np.int32)

    @property
    def dimension(self):
        return len(self.shape)

    def _extend_time(self, depth):
        while len(self._data) < depth:
            self._data.append(np.zeros(self.shape, dtype=self.numpy_dtype))
        self._data = self._data[:depth]

    def _op_invoke(self, depth, tick):
        self._extend_time(depth)
        if tick:
            last = self._data.pop()
            self._data.insert(0, last)

    @property
    def now(self):
        return self._data[0]

    def fill(self, data, time=0):
        if data.shape != self.now.shape or data.dtype != self.now.dtype:
            raise OracleError("Unable to fill grid with incompatible shape or data type")
        self._extend_time(abs(time))
        self._data[time] = data.copy()

    def __getitem__(self, s):
        return self.now[s]

    def __setitem__(self, s, v):
        self.now[s] = v


# ---------------------------------------------------------------- C text
def _abbr(t) -> str:
    if isinstance(t, Boolean):
        return "b"
    if isinstance(t, Integer):
        return f"i{t.width_bits}"
    if isinstance(t, Floating):
        return "f" if t.width_bytes == 4 else "d"
    if isinstance(t, Structure):
        return f"st{t.name}"
    raise OracleError(f"no abbreviation for {t!r}")


def _ctype(t) -> str:
    if isinstance(t, Void):
        return "void"
    if isinstance(t, Boolean):
        return "bool"
    if isinstance(t, Integer):
        return f"int{t.width_bits}_t"
    if isinstance(t, Floating):
        return "float" if t.width_bytes == 4 else "double"
    if isinstance(t, Structure):
        return f"struct {t.name}"
    if isinstance(t, Pointer):
        return _ctype(t.element) + "*"
    if isinstance(t, GridType):
        return f"struct __Grid{t.dimension}d_{_abbr(t.element)}"
    raise OracleError(f"no C type for {t!r}")


class CSource:
    def __init__(self, op, overstep="none", index="rowmajor", parallel=True):
        self.op = op
        self.overstep = overstep
        self.index = index
        self.parallel = parallel
        self.decls: list[str] = []
        self.types: dict[str, str] = {}
        self.funcs: dict[str, str] = {}
        self.depth = 0
        self.kernel_name = "xg_oracle_kernel"
        body = self.define(op, export=True)
        self.funcs[self.kernel_name] = body

    # -- types
    def type_name(self, t) -> str:
        if isinstance(t, Structure):
            self.define_struct(t)
        if isinstance(t, GridType):
            self.define_grid(t)
        if isinstance(t, Pointer):
            self.type_name(t.element)
        return _ctype(t)

    def define_struct(self, t: Structure):
        name = t.name
        if name in self.types:
            return
        # field types first: the emitted type definitions keep this insertion order
        fields_ = "".join(f"  {self.type_name(ft)} {fn};\n" for fn, ft in t.elements)
        self.types[name] = ""
        self.decls.append(f"struct {name};")
        self.types[name] = f"struct {name} {{\n{fields_}}};\n"

    def _lin(self, g: str, D: int, idx: list) -> str:
        if self.index == "reference":
            terms = [f"{idx[D - 1 - i]} * {'1' if i == 0 else f'{g}.shape[{i - 1}]'}" for i in range(D)]
            return " + ".join(terms)
        expr = idx[0]
        for d in range(1, D):
            expr = f"({expr}) * {g}.shape[{d}] + {idx[d]}"
        return expr

    def define_grid(self, t: GridType):
        name = f"__Grid{t.dimension}d_{_abbr(t.element)}"
        if name in self.types:
            return
        el = self.type_name(t.element)      # a structure element is defined before the grid
        self.types[name] = ""
        D = t.dimension
        self.decls.append(f"struct {name};")
        s = [f"struct {name} {{", "  int32_t time;", f"  int32_t shape[{D}];", f"  {el}** data;",
             "  int32_t* boundary_mask;", "};"]
        params = ", ".join(f"int32_t s{i}" for i in range(D))
        s.append(f"static inline {el}* {name}_at(struct {name} grid, {params}, int32_t t) {{")
        def adj(coord: str, extent: str) -> str:
            if self.overstep == "wrap":
                return f"(({coord} + {extent}) % {extent})"
            if self.overstep == "limit":
                return f"({coord} < 0 ? 0 : {coord} >= {extent} ? {extent} - 1 : {coord})"
            return coord

        if self.index == "reference":
            # term i: coordinate D-1-i, wrapped/clamped with extent i, times shape[i-1]
            # (generator.py:171-179)
            terms = [f"{adj(f's{D - 1 - i}', f'grid.shape[{i}]')} * {'1' if i == 0 else f'grid.shape[{i - 1}]'}"
                     for i in range(D)]
            s.append(f"  return &grid.data[t][{' + '.join(terms)}];")
        else:
            idx = [adj(f"s{d}", f"grid.shape[{d}]") for d in range(D)]
            s.append(f"  return &grid.data[t][{self._lin('grid', D, idx)}];")
        s.append("}")
        self.types[name] = "\n".join(s) + "\n"

    # -- expressions
    def expr(self, e) -> str:
        if isinstance(e, ir.Const):
            return repr(e.value).lower()
        if isinstance(e, ir.Name):
            return f"(*{e.var.name})" if isinstance(e.var.type, Pointer) else e.var.name
        if isinstance(e, ir.Field):
            return f"({self.expr(e.value)}).{e.attr}"
        if isinstance(e, ir.Cell):
            self.depth = max(self.depth, e.level)
            g = e.grid
            idx = ", ".join(f"_d{i} + {e.offsets[i]}" for i in range(g.type.dimension))
            return f"(*__Grid{g.type.dimension}d_{_abbr(g.type.element)}_at({g.name}, {idx}, {e.level}))"
        if isinstance(e, ir.Bin):
            if e.op == "**":
                fn = "pow" if isinstance(e.type, Floating) and e.type.width_bytes == 8 else "powf"
                return f"{fn}({self.expr(e.left)}, {self.expr(e.right)})"
            return f"({self.expr(e.left)} {e.op} {self.expr(e.right)})"
        if isinstance(e, ir.Un):
            return f"({e.op} {self.expr(e.operand)})"
        if isinstance(e, ir.Select):
            return f"({self.expr(e.cond)} ? {self.expr(e.then)} : {self.expr(e.orelse)})"
        if isinstance(e, ir.Cast):
            return f"(({self.type_name(e.type)})({self.expr(e.value)}))"
        if isinstance(e, ir.Call):
            args = [self.expr(a) for a in e.args]
            if e.ctor is not None:
                return f"(({self.type_name(e.ctor)}) {{{', '.join(args)}}})"
            callee = e.callee
            sym = self.define(callee)
            for i, pt in enumerate(e.param_types):
                if isinstance(pt, Pointer) and e.args[i].type == pt.element:
                    args[i] = f"&{args[i]}"
            return f"{sym}({', '.join(args)})"
        if isinstance(e, ir.GridInfo):
            if e.info == "dimension":
                return str(e.grid.type.dimension)
            return f"{e.grid.name}.shape[{self.expr(e.dim)}]"
        raise OracleError(f"cannot translate {type(e).__name__}")

    # -- statements
    def stencil_flag(self, s: ir.Assign):
        tgt = s.target
        implicit = any(isinstance(n, ir.Cell) and n.time == 0 and n.grid.name == tgt.grid.name and n.klass == 0
                       for n in ir.walk(s.value))
        return tgt.grid, tgt.klass, implicit

    def stmts(self, body, out, ind):
        for s in body:
            p = "  " * ind
            if isinstance(s, ir.Assign):
                if isinstance(s.target, ir.Cell):
                    self.nest(s, out, ind)
                else:
                    out.append(f"{p}{self.expr(s.target)} = {self.expr(s.value)};")
            elif isinstance(s, ir.Eval):
                out.append(f"{p}{self.expr(s.value)};")
            elif isinstance(s, ir.If):
                out.append(f"{p}if ({self.expr(s.cond)}) {{")
                self.stmts(s.body, out, ind + 1)
                if s.orelse:
                    out.append(f"{p}}} else {{")
                    self.stmts(s.orelse, out, ind + 1)
                out.append(f"{p}}}")
            elif isinstance(s, ir.While):
                out.append(f"{p}while ({self.expr(s.cond)}) {{")
                self.stmts(s.body, out, ind + 1)
                out.append(f"{p}}}")
            elif isinstance(s, ir.For):
                v = s.var.name
                out.append(f"{p}for ({v} = {self.expr(s.start)}; {v} < {self.expr(s.stop)}; {v} += {self.expr(s.step)}) {{")
                self.stmts(s.body, out, ind + 1)
                out.append(f"{p}}}")
            elif isinstance(s, ir.Return):
                out.append(f"{p}return;" if s.value is None else f"{p}return {self.expr(s.value)};")
            elif isinstance(s, ir.Break):
                out.append(f"{p}break;")
            elif isinstance(s, ir.Continue):
                out.append(f"{p}continue;")
            elif isinstance(s, ir.Inline):
                out.append(s.source)
            else:
                raise OracleError(f"cannot translate {type(s).__name__}")

    def nest(self, s: ir.Assign, out, ind):
        grid, klass, implicit = self.stencil_flag(s)
        D = grid.type.dimension
        g = grid.name
        dims = [f"_d{i}" for i in range(D)]
        ident = self._lin(g, D, dims)
        self.depth = max(self.depth, s.target.level)
        if implicit and not self.parallel:
            raise OracleError("Unable to perform implicit operation while parallel is off")

        def head():
            for i in range(D):
                out.append(f"for (int32_t _d{i} = 0; _d{i} < {g}.shape[{i}]; _d{i}++) {{")
            out.append(f"if ({g}.boundary_mask[{ident}] == {klass}) {{")

        def tail():
            out.append("}" * (D + 1))

        if implicit and klass == 0:
            vt = self.type_name(s.value.type)
            size = " * ".join(f"{g}.shape[{i}]" for i in range(D))
            out.append("{")
            out.append(f"{vt}* _value = malloc(sizeof({vt}) * ({size}));")
            out.append("#pragma omp parallel")
            out.append("{")
            out.append(f"#pragma omp for collapse({D})")
            head()
            out.append(f"_value[{ident}] = {self.expr(s.value)};")
            tail()
            out.append("#pragma omp barrier")
            out.append(f"#pragma omp for collapse({D})")
            head()
            out.append(f"{self.expr(s.target)} = _value[{ident}];")
            tail()
            out.append("}")
            out.append("free(_value);")
            out.append("}")
        else:
            if self.parallel:
                out.append(f"#pragma omp parallel for collapse({D})")
            head()
            out.append(f"{self.expr(s.target)} = {self.expr(s.value)};")
            tail()

    def define(self, op, export=False) -> str:
        sym = self.kernel_name if export else f"xg_fn_{op.c_name}"
        if not export and sym in self.funcs:
            return sym
        f = op.ir
        args = ", ".join(f"{self.type_name(t)} {n}" for n, t in f.args)
        sig = f"{self.type_name(f.ret)} {sym}({args})"
        if op.mode == "external":
            if op.intrinsic == "tick":
                raise OracleError("tick is declared but never defined")
            self.decls.append(f"extern {self.type_name(f.ret)} {op.name}({args});")
            return op.name
        if not export:
            self.funcs[sym] = ""
            self.decls.append(sig + ";")
        out = [sig + " {"]
        for v in f.locals:
            out.append(f"  {self.type_name(v.type)} {v.name};")
        for m in getattr(op, "macro", []):
            out.append(m)
        self.stmts(f.body, out, 1)
        out.append("}")
        text = "\n".join(out) + "\n"
        if not export:
            self.funcs[sym] = text
            return sym
        return text

    def source(self) -> str:
        heads = ["stdio.h", "stdlib.h", "stdint.h", "stdbool.h", "math.h"] + (["omp.h"] if self.parallel else [])
        out = [f"#include <{h}>" for h in heads]
        out += [f'#include "{i}"' for i in self.op.includes]
        out += self.decls
        out += list(self.types.values())
        out += [v for k, v in self.funcs.items() if k != self.kernel_name]
        out.append(self.funcs[self.kernel_name])
        if self.parallel:
            out.append("void xg_oracle_threads(int n) { omp_set_num_threads(n); }")
        return "\n".join(out) + "\n"


def compile_c(src: str, opt_level: int = 3, parallel: bool = True) -> str:
    os.makedirs(CACHE, exist_ok=True)
    flags = ["-shared", "-fpic", f"-O{opt_level}", "-ffp-contract=off"] + (["-fopenmp"] if parallel else [])
    key = hashlib.sha256((src + " ".join(flags)).encode()).hexdigest()[:40]
    lib = os.path.join(CACHE, key + ".so")
    if not os.path.exists(lib):
        c = os.path.join(CACHE, key + ".c")
        with open(c, "w") as f:
            f.write(src)
        tmp = lib + f".tmp{os.getpid()}"
        r = subprocess.run(["gcc", *flags, c, "-o", tmp, "-lm"], capture_output=True, text=True)
        if r.returncode != 0:
            raise OracleError(f"failed to compile oracle source:\n{r.stderr}")
        os.replace(tmp, lib)
    return lib


# ---------------------------------------------------------------- callable
def _grid_ctype(t: GridType):
    el = t.element.ctype
    return type(f"__Grid{t.dimension}d_{_abbr(t.element)}", (ctypes.Structure,), {
        "_fields_": [("time", ctypes.c_int32), ("shape", ctypes.c_int32 * t.dimension),
                     ("data", ctypes.POINTER(ctypes.POINTER(el))), ("boundary_mask", ctypes.POINTER(ctypes.c_int32))]})


class CpuKernel:
    """Run a kernel Operator on the restated CPU path (ticks HostGrid args first)."""

    def __init__(self, op, overstep="none", index="rowmajor", parallel=True, opt_level=3, tick=True):
        self.op = op
        cs = CSource(op, overstep, index, parallel)
        self.source = cs.source()
        self.depth = cs.depth + 1
        self.tick = tick
        self.lib = ctypes.CDLL(compile_c(self.source, opt_level, parallel))
        self.fn = getattr(self.lib, cs.kernel_name)
        self.args = list(op.ir.args)
        self.argtypes = []
        for _, t in self.args:
            if isinstance(t, GridType):
                self.argtypes.append(_grid_ctype(t))
            elif isinstance(t, Pointer):
                self.argtypes.append(ctypes.POINTER(t.element.ctype))
            else:
                self.argtypes.append(t.ctype)
        self.fn.argtypes = self.argtypes
        self.ret = op.ir.ret
        self.fn.restype = None if isinstance(self.ret, Void) else self.ret.ctype

    def threads(self, n: int):
        self.lib.xg_oracle_threads(int(n))

    def __call__(self, *args):
        if len(args) != len(self.args):
            raise TypeError(f"this function takes {len(self.args)} argument ({len(args)} given)")
        for a in args:
            if isinstance(a, HostGrid):
                a._op_invoke(self.depth, self.tick)
        cargs, keep = [], []
        for (n, t), ct, a in zip(self.args, self.argtypes, args):
            if isinstance(t, GridType):
                el = ctypes.POINTER(t.element.ctype)
                for lvl in a._data:
                    assert lvl.flags.c_contiguous
                mask = np.ascontiguousarray(a.boundary, dtype=np.int32)
                keep.append(mask)
                ptrs = (el * len(a._data))(*[lvl.ctypes.data_as(el) for lvl in a._data])
                keep.append(ptrs)
                shape_t = dict(ct._fields_)["shape"]          # the struct's own array type
                cargs.append(ct(len(a._data), shape_t(*a.shape), ptrs,
                                mask.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
            elif isinstance(t, Structure):
                cargs.append(t.to_c(a))
            elif isinstance(t, Pointer):
                cargs.append(a if isinstance(a, ctypes._Pointer) else ctypes.pointer(a))
            else:
                cargs.append(a)
        r = self.fn(*cargs)
        if isinstance(self.ret, Structure):
            return self.ret.from_c(r)
        return r
