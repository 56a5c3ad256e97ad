"""Python AST -> typed IR.

The accepted language and its typing follow the reference parser
(``xgrid/lang/parser.py:54-641``):

* statements: assignment (new locals take the value's type), augmented
  assignment on scalars, ``if``/``while`` (no ``else``)/``for .. in range(a, b[, s])``,
  ``return``, ``break``, ``continue``, ``pass``, ``import a.b`` (an include),
  ``with xgrid.boundary(k)`` and ``with xgrid.c()`` (inline host C text);
* expressions: ``+ - * / ** %``, comparisons (a chain ``a < b < c`` means
  ``a < b && a < c``, ``parser.py:396-414``), ``and``/``or``/``not``,
  conditional expressions, ``cast(T, x)``, structure constructors and methods,
  ``@function`` calls, ``shape(g, i)``/``dimension(g)``, global constants
  folded at first call (``parser.py:548-554``);
* grid subscripts ``g[o0, .., oD-1]`` with integer-literal offsets and an
  optional time subscript ``g[...][t]``; loads default to time -1 and stores
  to time 0 (``parser.py:512-525``);
* exact typing with no implicit promotion (``parser.py:374-376``); ``**``
  yields a double when its left operand is a double, else the configured
  float (``parser.py:378-382``).
"""
from __future__ import annotations

import ast
import inspect
import textwrap
import typing
from functools import reduce

from . import ir
from .config import get_config
from .types import (BaseType, Boolean, Floating, GridType, Ignore, Integer, Number, Pointer,
                    Reference, Structure, Value, Void, parse_annotation)


class DSLSyntaxError(SyntaxError):
    pass


_BINOPS = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/", ast.Pow: "**", ast.Mod: "%"}
_CMPOPS = {ast.Eq: "==", ast.NotEq: "!=", ast.Gt: ">", ast.GtE: ">=", ast.Lt: "<", ast.LtE: "<="}
_BOOLOPS = {ast.And: "&&", ast.Or: "||"}
_UNOPS = {ast.UAdd: "+", ast.USub: "-", ast.Not: "!"}


def _ctx(node) -> str:
    return "load" if isinstance(getattr(node, "ctx", ast.Load()), ast.Load) else "store"


def _int_literal(node):
    """Return the int of ``k`` or ``-k`` (int literal), else None."""
    if isinstance(node, ast.Constant) and type(node.value) is int:
        return node.value
    if (isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub)
            and isinstance(node.operand, ast.Constant) and type(node.operand.value) is int):
        return -node.operand.value
    return None


class Frontend:
    def __init__(self, func, name: str, mode: str, self_type: BaseType | None = None):
        self.func = func
        self.name = name
        self.mode = mode
        self.self_type = self_type
        self.file = inspect.getsourcefile(func) or "<unknown>"
        lines, first = inspect.getsourcelines(func)
        self.first_line = max(first, 1)
        tree = ast.parse(textwrap.dedent("".join(lines)))
        fdef = tree.body[0]
        if not isinstance(fdef, (ast.FunctionDef,)):
            raise DSLSyntaxError(f"{name}: expected a function definition")
        self.globals = dict(func.__globals__)
        try:
            cv = inspect.getclosurevars(func)
            self.globals.update(cv.nonlocals)
        except (TypeError, ValueError):
            pass
        self.scope: dict[str, ir.Var] = {}
        self.args: list = []
        self.klass = 0
        self.context = [mode]
        self.includes: list[str] = []
        self.result = self._function(fdef)

    # ------------------------------------------------------------ utilities
    def loc(self, node) -> ir.Loc:
        return ir.Loc(self.file, self.name, getattr(node, "lineno", 1) + self.first_line - 1)

    def error(self, node, message: str):
        loc = self.loc(node)
        raise DSLSyntaxError(f"File {loc.file}, line {loc.line}, in {self.name}\n  Syntax error: {message}")

    def visit(self, node):
        method = getattr(self, "v_" + type(node).__name__, None)
        if method is None:
            self.error(node, f"Python syntax '{type(node).__name__}' is currently unsupported")
        return method(node)

    def block(self, stmts) -> list:
        out = []
        for s in stmts:
            r = self.visit(s)
            if isinstance(r, list):
                out.extend(r)
            elif r is not None:
                out.append(r)
        return out

    # ------------------------------------------------------------- function
    def _function(self, node: ast.FunctionDef) -> ir.Function:
        sig = inspect.signature(self.func)
        for pname, param in sig.parameters.items():
            if param.kind != inspect.Parameter.POSITIONAL_OR_KEYWORD:
                self.error(node, f"Argument '{pname}' of kind '{param.kind}' is not supported")
            if param.annotation is inspect.Parameter.empty:
                if self.self_type is None:
                    self.error(node, f"Argument '{pname}' requires type annotation")
                t = self.self_type
            else:
                t = parse_annotation(param.annotation, self.globals)
            if t is None or isinstance(t, Void):
                self.error(node, f"Argument '{pname}' requires non-void type annotation ({param.annotation})")
            self.scope[pname] = ir.Var(pname, t)
            self.args.append((pname, t))
        ret = parse_annotation(sig.return_annotation, self.globals) \
            if sig.return_annotation is not inspect.Signature.empty else None
        if ret is None or isinstance(ret, Reference):
            self.error(node, f"Invalid return type '{ret}'")
        self.ret = ret
        body_nodes = list(node.body)
        # a leading docstring is documentation, not a statement
        if (body_nodes and isinstance(body_nodes[0], ast.Expr) and isinstance(body_nodes[0].value, ast.Constant)
                and isinstance(body_nodes[0].value.value, str) and self.mode != "external"):
            body_nodes = body_nodes[1:]
        body = [] if self.mode == "external" else self.block(body_nodes)
        return ir.Function(self.loc(node), self.name, self.mode, list(self.args), ret,
                           self.scope, body, self.includes)

    # ----------------------------------------------------------- statements
    def v_Return(self, node):
        value = None if node.value is None else self.visit(node.value)
        if value is not None and value.type != self.ret:
            self.error(node, f"Incompatible return type '{value.type}' with '{self.ret}'")
        return ir.Return(self.loc(node), value)

    def v_Pass(self, node):
        return []

    def v_Break(self, node):
        return ir.Break(self.loc(node))

    def v_Continue(self, node):
        return ir.Continue(self.loc(node))

    def v_If(self, node):
        cond = self.visit(node.test)
        self.context.append("if")
        body, orelse = self.block(node.body), self.block(node.orelse)
        self.context.pop()
        return ir.If(self.loc(node), cond, body, orelse)

    def v_While(self, node):
        cond = self.visit(node.test)
        if node.orelse:
            self.error(node, "While statement does not support else clause")
        self.context.append("while")
        body = self.block(node.body)
        self.context.pop()
        return ir.While(self.loc(node), cond, body)

    def v_Expr(self, node):
        if self.context[-1] == "c":
            if isinstance(node.value, ast.Constant) and isinstance(node.value.value, str):
                return ir.Inline(self.loc(node), node.value.value)
            self.error(node, "only string literals are allowed inside 'with c()'")
        return ir.Eval(self.loc(node), self.visit(node.value))

    def v_With(self, node):
        from . import lang
        if len(node.items) != 1:
            self.error(node, "Only one pragma switch at once")
        item = node.items[0].context_expr
        if not isinstance(item, ast.Call):
            self.error(node, f"Invalid pragma switch '{ast.dump(item)}'")
        obj = self.resolve_global(item.func)
        if obj is lang.c:
            self.context.append("c")
            body = self.block(node.body)
            self.context.pop()
            return body
        if obj is lang.boundary:
            args = item.args
            if len(args) != 1 or not isinstance(args[0], ast.Constant) or type(args[0].value) is not int:
                self.error(node, "Invalid pragram switch 'boundary'")
            self.klass = args[0].value
            body = self.block(node.body)
            self.klass = 0
            return body
        self.error(node, f"Unknown pragma switch '{obj}'")

    def v_Assign(self, node):
        loc = self.loc(node)
        value = self.visit(node.value)
        if isinstance(value.type, GridType):
            self.error(node, "Incompatible assignment to grid type")
        if len(node.targets) != 1:
            self.error(node, "Multiple assignment is not supported")
        target = node.targets[0]
        term = self.resolve_local(target)
        if term is None:
            if isinstance(target, ast.Name):
                var = ir.Var(target.id, value.type)
                self.scope[target.id] = var
                term = ir.Name(loc, value.type, var, "store")
            else:
                self.error(node, f"Undefined identifier {ast.dump(target)}")
        if term.type != value.type:
            self.error(node, f"Incompatible assignment from type {value.type} to {term.type}")
        return ir.Assign(loc, term, value)

    def v_AugAssign(self, node):
        value = self.visit(node.value)
        target = self.resolve_local(node.target)
        if target is None:
            self.error(node, f"Undefined identifier {ast.dump(node.target)}")
        op = _BINOPS.get(type(node.op))
        if op is None:
            self.error(node, f"Unsupported binary operator '{type(node.op).__name__}'")
        if isinstance(target, ir.Cell):
            self.error(node, "augmented assignment to a grid cell is not supported")
        if not isinstance(target.type, Number) or not isinstance(value.type, Number) or target.type != value.type:
            self.error(node, f"Incompatible binary operator '{op}' with type '{target.type}' and '{value.type}'")
        if op == "**":
            self.error(node, "augmented power assignment is not supported")
        loc = self.loc(node)
        return ir.Assign(loc, target, ir.Bin(loc, target.type, op, target, value))

    def v_For(self, node):
        if not isinstance(node.target, ast.Name):
            self.error(node, "For loop variable should be a name")
        it = node.iter
        if not (isinstance(it, ast.Call) and isinstance(it.func, ast.Name) and it.func.id == "range"):
            self.error(node, "For loop only supports range")
        if len(it.args) not in (2, 3):
            self.error(node, "For loop requires start:end:step or start:end")
        if node.orelse:
            self.error(node, "For statement does not support else clause")
        rng = [self.visit(a) for a in it.args]
        if len(rng) == 2:
            rng.append(ir.Const(self.loc(node), Integer(4), 1))
        rtype = rng[2].type
        for r in rng:
            if r.type != rtype or not isinstance(r.type, Number):
                self.error(node, f"Incompatible loop range type '{r.type}'")
        vid = node.target.id
        if vid not in self.scope:
            self.scope[vid] = ir.Var(vid, rtype)
        var = self.scope[vid]
        if var.type != rtype:
            self.error(node, f"Incompatible loop variable type '{var.type}' with range type '{rtype}'")
        body = self.block(node.body)
        return ir.For(self.loc(node), var, rng[0], rng[1], rng[2], body)

    def v_Import(self, node):
        for alias in node.names:
            if alias.asname is not None:
                self.error(node, f"Using import as include requires no alias for '{alias.name}'")
            self.includes.append(alias.name.replace(".", "/") + ".h")
        return []

    # ---------------------------------------------------------- expressions
    def constant(self, node, value):
        t = {int: Integer(4), float: Floating(get_config().fsize), bool: Boolean()}.get(type(value))
        if t is None:
            self.error(node, f"Incompatible constant '{value}' of type '{type(value)}'")
        return ir.Const(self.loc(node), t, value)

    def v_Constant(self, node):
        return self.constant(node, node.value)

    def v_UnaryOp(self, node):
        op = _UNOPS.get(type(node.op))
        if op is None:
            self.error(node, f"Unsupported unary operator '{type(node.op).__name__}'")
        x = self.visit(node.operand)
        if (op == "!" and not isinstance(x.type, Boolean)) or (op in "+-" and not isinstance(x.type, Number)):
            self.error(node, f"Incompatible unary operator '{op}' with type '{x.type}'")
        return ir.Un(self.loc(node), x.type, op, x)

    def v_BinOp(self, node):
        op = _BINOPS.get(type(node.op))
        if op is None:
            self.error(node, f"Unsupported binary operator '{type(node.op).__name__}'")
        left, right = self.visit(node.left), self.visit(node.right)
        if not isinstance(left.type, Number) or not isinstance(right.type, Number) or left.type != right.type:
            self.error(node, f"Incompatible binary operator '{op}' with type '{left.type}' and '{right.type}'")
        if op == "**":
            rtype = Floating(8) if isinstance(left.type, Floating) and left.type.width_bytes == 8 \
                else Floating(get_config().fsize)
        else:
            rtype = left.type
        return ir.Bin(self.loc(node), rtype, op, left, right)

    def v_BoolOp(self, node):
        loc = self.loc(node)
        op = _BOOLOPS[type(node.op)]
        values = [self.visit(v) for v in node.values]
        for v in values:
            if not isinstance(v.type, Boolean):
                self.error(node, f"Incompatible boolean operator '{op}' with '{v.type}'")
        return reduce(lambda a, b: ir.Bin(loc, Boolean(), op, a, b), values)

    def v_Compare(self, node):
        loc = self.loc(node)
        left = self.visit(node.left)
        if not isinstance(left.type, Number):
            self.error(node, f"Incompatible compare expression with type '{left.type}'")
        comps = [self.visit(c) for c in node.comparators]
        ops = []
        for o in node.ops:
            if type(o) not in _CMPOPS:
                self.error(node, f"Unsupported compare operator '{type(o).__name__}'")
            ops.append(_CMPOPS[type(o)])
        for o, c in zip(ops, comps):
            if c.type != left.type:
                self.error(node, f"Incompatible compare operator '{o}' with type '{left.type}' and '{c.type}'")
        # every comparator is compared with the LEFT operand (parser.py:413-414)
        terms = [ir.Bin(loc, Boolean(), o, left, c) for o, c in zip(ops, comps)]
        return reduce(lambda a, b: ir.Bin(loc, Boolean(), "&&", a, b), terms)

    def v_IfExp(self, node):
        cond = self.visit(node.test)
        if not isinstance(cond.type, Boolean):
            self.error(node, f"Incompatible condition type '{cond.type}' of if expression")
        a, b = self.visit(node.body), self.visit(node.orelse)
        if a.type != b.type or not isinstance(a.type, Value):
            self.error(node, f"Incompatible type '{a.type}' and '{b.type}' of if expression")
        return ir.Select(self.loc(node), a.type, cond, a, b)

    def resolve_global(self, node):
        names = []
        while True:
            if isinstance(node, ast.Name):
                names.append(node.id)
                break
            if isinstance(node, ast.Attribute):
                names.append(node.attr)
                node = node.value
                continue
            self.error(node, f"Python syntax '{type(node).__name__}' is currently unsupported")
        names.reverse()
        scope = self.globals
        resolved = ["globals"]
        builtins = {"int": int, "float": float, "bool": bool, "range": range}
        for i, attr in enumerate(names):
            try:
                if isinstance(scope, dict):
                    if attr in scope:
                        scope = scope[attr]
                    elif i == 0 and attr in builtins:
                        scope = builtins[attr]
                    else:
                        raise KeyError(attr)
                else:
                    scope = getattr(scope, attr)
            except (KeyError, AttributeError):
                self.error(node, f"Undefined attribute '{attr}' of '{'.'.join(resolved)}'")
            resolved.append(attr)
        return scope

    def _cell(self, sub: ast.Subscript, time: int) -> ir.Cell:
        if not isinstance(sub.value, ast.Name):
            self.error(sub, f"Incompatible subscript to '{type(sub.value).__name__}' under context '{self.context[-1]}'")
        gname = sub.value.id
        if gname not in self.scope:
            self.error(sub, f"Undefined identifier '{gname}'")
        var = self.scope[gname]
        if not isinstance(var.type, GridType):
            self.error(sub, f"Incompatible subscript to type '{var.type}'")
        elts = sub.slice.elts if isinstance(sub.slice, ast.Tuple) else [sub.slice]
        offs = []
        for e in elts:
            v = _int_literal(e)
            if v is None:
                self.error(sub, f"Incompatible subscript '{ast.dump(e)}'")
            offs.append(v)
        if len(offs) != var.type.dimension:
            self.error(sub, f"Incompatible subscript length '{len(offs)}' with dimension {var.type.dimension}")
        return ir.Cell(self.loc(sub), var.type.element, var, time, tuple(offs), self.klass, _ctx(sub))

    def resolve_local(self, node):
        if isinstance(node, ast.Subscript):
            if isinstance(node.value, ast.Subscript):
                t = _int_literal(node.slice)
                if t is None:
                    self.error(node.value, f"Invalid time dimension subscript to '{type(node.value).__name__}'")
                return self._cell(node.value, t)
            return self._cell(node, 0 if _ctx(node) == "store" else -1)
        if isinstance(node, ast.Attribute):
            value = self.resolve_local(node.value)
            if value is None or not isinstance(value.type, Structure) or node.attr not in value.type.elements_map:
                return None
            return ir.Field(self.loc(node), value.type.elements_map[node.attr], value, node.attr, _ctx(node))
        if isinstance(node, ast.Name):
            var = self.scope.get(node.id)
            if var is None:
                return None
            t = var.type.element if isinstance(var.type, Pointer) else var.type
            return ir.Name(self.loc(node), t, var, _ctx(node))
        self.error(node, f"Python syntax '{type(node).__name__}' is currently unsupported")

    def v_Name(self, node):
        local = self.resolve_local(node)
        return local if local is not None else self.constant(node, self.resolve_global(node))

    def v_Attribute(self, node):
        local = self.resolve_local(node)
        return local if local is not None else self.constant(node, self.resolve_global(node))

    def v_Subscript(self, node):
        return self.resolve_local(node)

    def v_Call(self, node):
        from .operator import Operator, method_operator
        loc = self.loc(node)
        args: list = []
        func = None
        fname = None
        if isinstance(node.func, ast.Attribute) and (obj := self.resolve_local(node.func.value)) is not None:
            if not isinstance(obj.type, Structure):
                self.error(node, f"Invalid method call on type {obj.type}")
            raw = getattr(obj.type.dataclass, node.func.attr, None)
            fname = f"{obj.type.dataclass.__qualname__}.{node.func.attr}"
            flag = getattr(raw, "__xgrid_method", None)
            if flag is None:
                self.error(node, f"Invalid call to object '{raw}', it is not an operator")
            func = method_operator(raw, obj.type)
            args.append(obj)
        else:
            func = self.resolve_global(node.func)

        if func is typing.cast:
            if len(node.args) != 2:
                self.error(node, f"Cast requires 2 arguments, got {len(node.args)}")
            target = parse_annotation(self.resolve_global(node.args[0]), self.globals)
            if target is None:
                self.error(node, "Invalid cast type")
            return ir.Cast(loc, target, self.visit(node.args[1]))

        if node.keywords:
            self.error(node, "keyword arguments are not supported")
        args.extend(self.visit(a) for a in node.args)

        if isinstance(func, type):
            st = parse_annotation(func, self.globals)
            if not isinstance(st, Structure):
                self.error(node, f"Invalid type constructor '{func.__name__}'")
            params = [t for _, t in st.elements]
            if len(params) != len(args):
                self.error(node, f"Operator '{st.name}.constructor' requires {len(params)} arguments, but got {len(args)}")
            for (pn, pt), a in zip(st.elements, args):
                if pt != a.type:
                    self.error(node, f"Incompatible type '{a.type}' with '{pt}' of argument '{pn}, operator '{st.name}.constructor'")
            return ir.Call(loc, st, None, st, args, params)

        if not isinstance(func, Operator):
            self.error(node, f"Invalid call to object '{func}', it is not an operator")
        fname = fname or func.name
        sig_args = func.signature_args()
        if len(sig_args) != len(args):
            self.error(node, f"Operator '{fname}' requires {len(sig_args)} arguments, but got {len(args)}")
        if func.mode == "external" and func.typecheck_override is not None:
            try:
                rtype = func.typecheck_override([a.type for a in args])
            except Exception as e:  # the override reports the type error
                self.error(node, e.args[0] if e.args else str(e))
            if func.intrinsic in ("shape", "dimension"):
                if not isinstance(args[0], ir.Name):
                    self.error(node, f"Incompatible '{func.intrinsic}' to argument '{args[0]}'")
                return ir.GridInfo(loc, Integer(4), func.intrinsic, args[0].var,
                                   args[1] if func.intrinsic == "shape" else None)
        else:
            for (pn, pt), a in zip(sig_args, args):
                if pt != a.type:
                    self.error(node, f"Incompatible type '{a.type}' with '{pt}' of argument '{pn}, operator '{fname}'")
            rtype = func.signature_ret()
        return ir.Call(loc, rtype, func, None, args, [t for _, t in sig_args])


def parse(func, name: str, mode: str, self_type: BaseType | None = None) -> ir.Function:
    return Frontend(func, name, mode, self_type).result
