"""Python AST -> stencil IR.

Behaviour follows the upstream parser (``cutile/frontend/parser.py:62-664``):

* only the ``return`` expression is lowered (``:465-475``); accesses are also
  collected from plain assignments in the body (``:494-535``);
* array parameters are the subscripted ones, index variables the names used
  inside subscripts, with the positional fallbacks of ``:515-523``;
* an index is ``i``, ``i + c``, ``i - c`` or ``c + i`` with an integer
  constant (``:80-132``);
* ``ndim`` and ``order = 2 * max|offset|`` are inferred when not given
  (``:157-177``);
* supported expressions: ``+ - * /``, unary ``-``, numeric constants, closure
  or module-level numeric names, ``max/maximum``, ``min/minimum``,
  ``abs/fabs``, ``where(cond, a, b)`` and single comparisons
  (``:236-361``);
* constants are de-duplicated by value (``:226-234``), accesses by
  (field, offsets) in first-use order (``:585-608``).
"""

from __future__ import annotations

import ast
import inspect
import textwrap

from cutile.frontend.ir import Access, BoundaryAttr, Const, FuncOp, ModuleOp, Op

_DTYPES = ("float64", "float32")


class ParseError(Exception):
    """The stencil function cannot be lowered."""


# kept under the upstream private name as well
_ParseError = ParseError


def _const_int(node: ast.expr) -> int:
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) \
            and not isinstance(node.value, bool):
        return int(node.value)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_const_int(node.operand)
    raise ParseError(f"Expected compile-time integer constant, got: {ast.dump(node)}")


def _offset_1(node: ast.expr, index_vars: list) -> int:
    if isinstance(node, ast.Name):
        if node.id in index_vars:
            return 0
        raise ParseError(f"Unknown index variable '{node.id}', expected one of {index_vars}")
    if isinstance(node, ast.BinOp):
        if isinstance(node.left, ast.Name) and node.left.id in index_vars:
            c = _const_int(node.right)
            if isinstance(node.op, ast.Add):
                return c
            if isinstance(node.op, ast.Sub):
                return -c
            raise ParseError(f"Unsupported offset operation: {type(node.op).__name__}")
        if isinstance(node.right, ast.Name) and node.right.id in index_vars:
            c = _const_int(node.left)
            if isinstance(node.op, ast.Add):
                return c
            raise ParseError(f"Unsupported offset operation: {type(node.op).__name__}")
    raise ParseError(f"Cannot resolve subscript expression to an offset: {ast.dump(node)}")


def resolve_offsets(node: ast.expr, index_vars: list) -> list:
    elts = node.elts if isinstance(node, ast.Tuple) else [node]
    return [_offset_1(e, index_vars) for e in elts]


def _collect(node: ast.expr, arrays: list, index_vars: list, out: list) -> None:
    if isinstance(node, ast.Subscript) and isinstance(node.value, ast.Name) \
            and node.value.id in arrays:
        out.append((node.value.id, resolve_offsets(node.slice, index_vars)))
        return
    for child in ast.iter_child_nodes(node):
        if isinstance(child, ast.expr):
            _collect(child, arrays, index_vars, out)


class _Builder:
    _BIN = {ast.Add: "add", ast.Sub: "sub", ast.Mult: "mul", ast.Div: "div"}
    _CALL2 = {"max": "max", "maximum": "max", "min": "min", "minimum": "min"}
    _CMP = {ast.Gt: "ogt", ast.GtE: "oge", ast.Lt: "olt", ast.LtE: "ole",
            ast.Eq: "oeq", ast.NotEq: "une"}

    def __init__(self, accesses: dict, constants: dict, index_vars: list, ops: list):
        self.accesses = accesses
        self.constants = constants
        self.index_vars = index_vars
        self.ops = ops
        self._consts: dict = {}

    def const(self, v: float) -> Const:
        v = float(v)
        if v not in self._consts:
            c = Const(v)
            self._consts[v] = c
            self.ops.append(c)
        return self._consts[v]

    def op(self, kind, operands, predicate=None) -> Op:
        o = Op(kind, operands, predicate)
        self.ops.append(o)
        return o

    def build(self, node: ast.expr):
        if isinstance(node, ast.BinOp):
            kind = self._BIN.get(type(node.op))
            if kind is None:
                raise ParseError(f"Unsupported binary operator: {type(node.op).__name__}")
            left = self.build(node.left)
            right = self.build(node.right)
            return self.op(kind, (left, right))
        if isinstance(node, ast.UnaryOp):
            if isinstance(node.op, ast.USub):
                return self.op("neg", (self.build(node.operand),))
            if isinstance(node.op, ast.UAdd):
                return self.build(node.operand)
            raise ParseError(f"Unsupported unary operator: {type(node.op).__name__}")
        if isinstance(node, ast.Constant):
            if isinstance(node.value, (int, float)) and not isinstance(node.value, bool):
                return self.const(node.value)
            raise ParseError(f"Unsupported constant type: {type(node.value)}")
        if isinstance(node, ast.Subscript):
            if not isinstance(node.value, ast.Name):
                raise ParseError("Subscript target must be a simple name")
            key = (node.value.id, tuple(resolve_offsets(node.slice, self.index_vars)))
            if key not in self.accesses:
                raise ParseError(f"Access {key[0]}{list(key[1])} not found in access map")
            return self.accesses[key]
        if isinstance(node, ast.Name):
            if node.id in self.constants:
                return self.const(self.constants[node.id])
            raise ParseError(
                f"Unknown variable '{node.id}'. Not an array parameter or known constant.")
        if isinstance(node, ast.Call):
            f = node.func
            name = f.id if isinstance(f, ast.Name) else (f.attr if isinstance(f, ast.Attribute) else None)
            if node.keywords or name is None:
                raise ParseError(f"Unsupported call: {ast.dump(node)}")
            if name in self._CALL2 and len(node.args) == 2:
                a = self.build(node.args[0])
                b = self.build(node.args[1])
                return self.op(self._CALL2[name], (a, b))
            if name in ("abs", "fabs") and len(node.args) == 1:
                return self.op("abs", (self.build(node.args[0]),))
            if name == "where" and len(node.args) == 3:
                c = self.build(node.args[0])
                a = self.build(node.args[1])
                b = self.build(node.args[2])
                return self.op("select", (c, a, b))
            raise ParseError(
                f"Unsupported call '{name}' with {len(node.args)} argument(s); "
                f"supported: max(a, b), min(a, b), abs(x), where(cond, a, b)")
        if isinstance(node, ast.Compare):
            if len(node.ops) != 1 or len(node.comparators) != 1:
                raise ParseError("Unsupported chained comparison")
            pred = self._CMP.get(type(node.ops[0]))
            if pred is None:
                raise ParseError(f"Unsupported comparison: {type(node.ops[0]).__name__}")
            a = self.build(node.left)
            b = self.build(node.comparators[0])
            return self.op("cmp", (a, b), pred)
        raise ParseError(f"Unsupported expression node: {type(node).__name__} ({ast.dump(node)})")


def _closure_constants(fn) -> dict:
    out = {}
    if fn.__code__.co_freevars and fn.__closure__:
        for name, cell in zip(fn.__code__.co_freevars, fn.__closure__):
            try:
                val = cell.cell_contents
            except ValueError:
                continue
            if isinstance(val, (int, float)) and not isinstance(val, bool):
                out[name] = float(val)
    return out


def parse_stencil(fn, ndim=None, order=None, dtype="float64", boundary=None, constants=None) -> ModuleOp:
    """Parse ``fn`` into a module holding one :class:`FuncOp`."""
    source = textwrap.dedent(inspect.getsource(fn))
    tree = ast.parse(source)
    source_file, line_offset = None, 0
    try:
        source_file = inspect.getfile(fn)
        line_offset = inspect.getsourcelines(fn)[1] - 1
    except (OSError, TypeError):
        pass

    func_def = next((n for n in ast.walk(tree)
                     if isinstance(n, (ast.FunctionDef, ast.AsyncFunctionDef))), None)
    if func_def is None:
        raise ParseError("No function definition found in source")
    name = func_def.name
    params = [a.arg for a in func_def.args.args]
    if len(params) < 2:
        raise ParseError(
            f"Stencil function '{name}' must have at least 2 parameters "
            f"(array + index variables), got {len(params)}: {params}")

    ret = next((n for n in ast.walk(func_def) if isinstance(n, ast.Return) and n.value is not None), None)
    if ret is None:
        raise ParseError(f"Stencil function '{name}' must have a return statement")
    expr = ret.value

    sub_arrays, sub_index = set(), set()
    scan = [expr] + [s.value for s in func_def.body if isinstance(s, ast.Assign)]
    for root in scan:
        for node in ast.walk(root):
            if isinstance(node, ast.Subscript) and isinstance(node.value, ast.Name) \
                    and node.value.id in params:
                sub_arrays.add(node.value.id)
                for s in ast.walk(node.slice):
                    if isinstance(s, ast.Name) and s.id in params:
                        sub_index.add(s.id)
    arrays = [p for p in params if p in sub_arrays]
    index_vars = [p for p in params if p in sub_index]
    if not arrays:
        arrays, index_vars = [params[0]], params[1:]
    if not index_vars:
        index_vars = [p for p in params if p not in set(arrays)]

    found: list = []
    _collect(expr, arrays, index_vars, found)
    for s in func_def.body:
        if isinstance(s, ast.Assign):
            _collect(s.value, arrays, index_vars, found)

    if ndim is None or order is None:
        if not found:
            raise ParseError("No array accesses found; cannot infer ndim/order")
        nd = len(found[0][1])
        mx = 0
        for _, offs in found:
            if len(offs) != nd:
                raise ParseError(f"Inconsistent dimensionality: expected {nd}, got {len(offs)}")
            mx = max([mx] + [abs(o) for o in offs])
        ndim = nd if ndim is None else ndim
        order = 2 * mx if order is None else order

    if dtype not in _DTYPES:
        raise ParseError(f"Unsupported dtype '{dtype}'. Use 'float32' or 'float64'.")

    consts = _closure_constants(fn)
    if constants:
        consts.update({k: float(v) for k, v in constants.items()})
    referenced = {n.id for n in ast.walk(expr) if isinstance(n, ast.Name)}
    for nm in referenced:
        if nm not in consts and nm not in arrays and nm in getattr(fn, "__globals__", {}):
            val = fn.__globals__[nm]
            if isinstance(val, (int, float)) and not isinstance(val, bool):
                consts[nm] = float(val)

    ops: list = []
    unique: dict = {}
    for arr, offs in found:
        key = (arr, tuple(offs))
        if key in unique:
            continue
        a = Access(arr, arrays.index(arr), offs, index_vars)
        unique[key] = a
        ops.append(a)
    accesses = list(unique.values())

    builder = _Builder(unique, consts, index_vars, ops)
    result = builder.build(expr)

    bnd = None
    if boundary is not None:
        if isinstance(boundary, str):
            bnd = BoundaryAttr(boundary)
        elif isinstance(boundary, dict):
            bnd = BoundaryAttr(boundary.get("type", boundary.get("bc_type", "dirichlet")),
                               boundary.get("value"))
        elif isinstance(boundary, BoundaryAttr):
            bnd = boundary
        else:
            from cutile.frontend.types import BoundarySpec

            if isinstance(boundary, BoundarySpec):
                c = boundary.conditions[0][0]
                bnd = BoundaryAttr(c.bc_type.value, c.value if c.value else None)

    func = FuncOp(
        func_name=name, ndim=ndim, order=order, dtype=dtype, arg_names=arrays,
        index_names=index_vars, accesses=accesses, ops=ops, result=result,
        boundary=bnd, constants=consts, source_file=source_file,
        source_line=func_def.lineno + line_offset,
    )
    return ModuleOp([func])
