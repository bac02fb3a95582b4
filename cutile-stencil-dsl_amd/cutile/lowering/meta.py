"""Stencil metadata and expression reconstruction for the HIP emitter.

Follows ``cutile/lowering/stencil_to_target.py:163-465`` in what decides the
arithmetic:

* one load per distinct (field, offset), named ``<field>_<off>_...`` with
  ``0 / p<k> / m<k>`` offset tokens (``:135-142, 211-237``);
* the expression keeps the source operator tree exactly (Python left-to-right
  association, explicit parentheses) and folds constant sub-expressions in
  float64 the way the reference does (``_fold``, ``:260-273, 377-441``):
  binary ops with two constant operands, negation, ``max``/``min``/``abs``
  of constants;
* fused kernels merge inputs by parameter name and share accesses
  (``extract_fused_meta``, ``:968-1031``), named by ``fused_name``
  (``:964-965``).

Instead of Python text for cuTile, the expression is rendered as C++ for
the device: every constant becomes an exact hexadecimal literal in the
compute type (float64, or the float32 rounding NumPy applies to weak Python
scalars), every operation a single IEEE operation, ``max``/``min`` the
NaN-propagating selects NumPy uses.  Compiled with contraction off, the
device result is bit-identical to NumPy on the same inputs.
"""

from __future__ import annotations

import math
import re
import struct
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from cutile.frontend.ir import Access, Const, FuncOp, Op, find_func

VERY_WIDE_LOADS = 64


def format_offset(off: int) -> str:
    if off == 0:
        return "0"
    return f"p{off}" if off > 0 else f"m{-off}"


def default_input_names(n: int) -> list:
    d = list("uvwxyz")
    return [d[i] if i < len(d) else f"arr{i}" for i in range(n)]


def map3(ndim: int, offs: Sequence[int], mapping: str = "rows") -> tuple:
    """Offsets of an ``ndim``-D access on the 3-D (stream, row, column) map.

    2-D arrays stream along their rows by default (``"rows"``: axis 0 is the
    streaming axis); ``"plane"`` maps them to one plane tiled in 2-D."""
    if ndim == 1:
        return (0, 0, offs[0])
    if ndim == 2:
        return (0, offs[0], offs[1]) if mapping == "plane" else (offs[0], 0, offs[1])
    return tuple(offs)


@dataclass
class AccessInfo:
    array_index: int
    offsets: tuple
    view_name: str


@dataclass
class StencilMeta:
    name: str = ""
    ndim: int = 2
    order: int = 2
    dtype: str = "float64"
    input_names: list = field(default_factory=list)
    accesses: list = field(default_factory=list)
    expressions: list = field(default_factory=list)      # C++ per output
    py_expressions: list = field(default_factory=list)   # Python text per output
    output_names: list = field(default_factory=list)
    stencil_inputs: list = field(default_factory=list)
    constants: dict = field(default_factory=dict)
    boundary: Optional[dict] = None
    flops: int = 0
    mapping: str = "rows"
    update: Optional[dict] = None     # time-update epilogue, see apply_update

    @property
    def num_inputs(self) -> int:
        return len(self.input_names)

    @property
    def num_outputs(self) -> int:
        """Arrays the kernel writes (rates' targets, plus compensations)."""
        n = max(1, len(self.expressions))
        return 2 * n if (self.update and self.update["mode"] == "kahan") else n

    @property
    def num_exprs(self) -> int:
        return max(1, len(self.expressions))

    @property
    def expression(self) -> str:
        return self.expressions[0]

    def output_params(self) -> list:
        if self.output_names:
            return [f"{n}_out" for n in self.output_names]
        return ["u_out"]


# ---------------------------------------------------------------------------
# constant rendering
# ---------------------------------------------------------------------------

def f32_exact(v: float) -> bool:
    return math.isfinite(v) and float(np.float32(v)) == v


def cpp_literal(v: float, ctype: str) -> str:
    """Exact C++ literal of ``v`` in the compute type."""
    if ctype == "float":
        v = float(np.float32(v))
        if math.isinf(v):
            return "(-__builtin_huge_valf())" if v < 0 else "__builtin_huge_valf()"
        if math.isnan(v):
            return "__builtin_nanf(\"\")"
        return f"({float.hex(v)}f)" if v < 0 else f"{float.hex(v)}f"
    if math.isinf(v):
        return "(-__builtin_huge_val())" if v < 0 else "__builtin_huge_val()"
    if math.isnan(v):
        return "__builtin_nan(\"\")"
    return f"({float.hex(v)})" if v < 0 else float.hex(v)


def _fold(kind: str, a: float, b: float):
    try:
        if kind == "add":
            r = a + b
        elif kind == "sub":
            r = a - b
        elif kind == "mul":
            r = a * b
        elif kind == "div":
            r = a / b
        else:
            return None
    except (ZeroDivisionError, OverflowError):
        return None
    return r if math.isfinite(r) else None


_CMP_CPP = {"ogt": ">", "oge": ">=", "olt": "<", "ole": "<=", "oeq": "==", "une": "!="}
_BIN_CPP = {"add": "+", "sub": "-", "mul": "*", "div": "/"}


class _Render:
    """Render values of one stencil body as C++ and as Python text."""

    def __init__(self, ndim: int, ctype: str, field_index: dict, view_names: dict, mapping: str = "rows"):
        self.ndim = ndim
        self.mapping = mapping
        self.ctype = ctype
        self.field_index = field_index      # parameter name -> global input index
        self.view_names = view_names        # Access -> view name
        self.const: dict = {}               # Value -> folded float
        self.memo_c: dict = {}
        self.memo_p: dict = {}
        self.flops = 0

    def fold(self, v) -> Optional[float]:
        if v in self.const:
            return self.const[v]
        r = None
        if isinstance(v, Const):
            r = v.value
        elif isinstance(v, Op):
            if v.kind in _BIN_CPP:
                a, b = self.fold(v.lhs), self.fold(v.rhs)
                if a is not None and b is not None:
                    r = _fold(v.kind, a, b)
            elif v.kind == "neg":
                a = self.fold(v.operands[0])
                r = -a if a is not None else None
            elif v.kind in ("max", "min"):
                a, b = self.fold(v.lhs), self.fold(v.rhs)
                if a is not None and b is not None:
                    r = max(a, b) if v.kind == "max" else min(a, b)
            elif v.kind == "abs":
                a = self.fold(v.operands[0])
                r = abs(a) if a is not None else None
        if r is not None:
            self.const[v] = r
        return r

    def cpp(self, v) -> str:
        if v in self.memo_c:
            return self.memo_c[v]
        c = self.fold(v)
        if c is not None:
            s = cpp_literal(c, self.ctype)
        elif isinstance(v, Access):
            dx, dy, dz = map3(self.ndim, v.offsets, self.mapping)
            s = f"a.template g<{self.field_index[v.field]}, {dx}, {dy}, {dz}>()"
        else:
            k = v.kind
            if k in _BIN_CPP:
                s = f"({self.cpp(v.lhs)} {_BIN_CPP[k]} {self.cpp(v.rhs)})"
            elif k == "neg":
                s = f"(-{self.cpp(v.operands[0])})"
            elif k == "max":
                s = f"st::maxp<C>({self.cpp(v.lhs)}, {self.cpp(v.rhs)})"
            elif k == "min":
                s = f"st::minp<C>({self.cpp(v.lhs)}, {self.cpp(v.rhs)})"
            elif k == "abs":
                s = f"st::absp<C>({self.cpp(v.operands[0])})"
            elif k == "cmp":
                s = f"({self.cpp(v.lhs)} {_CMP_CPP[v.predicate]} {self.cpp(v.rhs)})"
            elif k == "select":
                c0, a, b = v.operands
                s = f"(({self.cpp(c0)}) ? C({self.cpp(a)}) : C({self.cpp(b)}))"
            else:
                raise ValueError(f"cannot render op {k}")
        self.memo_c[v] = s
        return s

    def py(self, v) -> str:
        """Python-syntax rendering (fully parenthesised; for display)."""
        if v in self.memo_p:
            return self.memo_p[v]
        c = self.fold(v)
        if c is not None:
            s = repr(c) if c >= 0 else f"({c!r})"
        elif isinstance(v, Access):
            s = f"t_{self.view_names[v]}"
        else:
            k = v.kind
            if k in _BIN_CPP:
                s = f"({self.py(v.lhs)} {_BIN_CPP[k]} {self.py(v.rhs)})"
            elif k == "neg":
                s = f"(-{self.py(v.operands[0])})"
            elif k in ("max", "min"):
                s = f"{'maximum' if k == 'max' else 'minimum'}({self.py(v.lhs)}, {self.py(v.rhs)})"
            elif k == "abs":
                s = f"abs({self.py(v.operands[0])})"
            elif k == "cmp":
                s = f"({self.py(v.lhs)} {_CMP_CPP[v.predicate]} {self.py(v.rhs)})"
            else:
                c0, a, b = v.operands
                s = f"where({self.py(c0)}, {self.py(a)}, {self.py(b)})"
        self.memo_p[v] = s
        return s


def count_flops(func: FuncOp) -> int:
    """Roofline flop count (``passes/analysis/roofline.py:61-104``): add, sub,
    mul, div, max, min, select and abs count one; neg and constants are free."""
    return sum(1 for v in func.ops if isinstance(v, Op) and v.kind not in ("neg", "cmp"))


def _view_name(field_name: str, offsets: tuple) -> str:
    return f"{field_name}_" + "_".join(format_offset(o) for o in offsets)


def extract_meta(func: FuncOp, input_names: Optional[Sequence[str]] = None,
                 compute_dtype: Optional[str] = None, mapping: str = "rows") -> StencilMeta:
    """Metadata of a single stencil (``_extract_meta``, ``stencil_to_target.py:163-252``)."""
    meta = StencilMeta(name=func.func_name, ndim=func.ndim, order=func.order, dtype=func.dtype)
    nin = len(func.arg_names)
    if input_names is not None:
        if len(input_names) != nin:
            raise ValueError(f"{meta.name}: {len(input_names)} input names for {nin} arrays")
        meta.input_names = list(input_names)
    else:
        meta.input_names = list(func.arg_names)
    meta.constants = dict(func.constants)
    if func.boundary is not None:
        meta.boundary = {"type": func.boundary.bc_type}
        if func.boundary.has_value:
            meta.boundary["value"] = func.boundary.value
    names = {}
    for acc in func.accesses:
        fname = meta.input_names[acc.index]
        names[acc] = _view_name(fname, acc.offsets)
        meta.accesses.append(AccessInfo(acc.index, acc.offsets, names[acc]))
    ctype = "float" if (compute_dtype or func.dtype) == "float32" else "double"
    r = _Render(func.ndim, ctype, {p: k for k, p in enumerate(func.arg_names)}, names, mapping)
    meta.mapping = mapping
    meta.expressions = [r.cpp(func.result)]
    meta.py_expressions = [r.py(func.result)]
    meta.flops = count_flops(func)
    meta.stencil_inputs = [list(func.arg_names)]
    return meta


def fused_name(names: Sequence[str]) -> str:
    return "_".join(names) if len(names) <= 3 else f"{names[0]}_and_{len(names) - 1}_others"


def extract_fused_meta(modules: Sequence, compute_dtype: Optional[str] = None,
                       mapping: str = "rows") -> StencilMeta:
    """Merge several stencils into one multi-output kernel description
    (``extract_fused_meta``, ``stencil_to_target.py:968-1031``)."""
    if not modules:
        raise ValueError("At least one module is required for fusion")
    funcs = [find_func(m) for m in modules]
    inputs: list = []
    for f in funcs:
        for n in f.arg_names:
            if n not in inputs:
                inputs.append(n)
    names = [f.func_name for f in funcs]
    if len(set(names)) != len(names):
        raise ValueError(f"fused stencils must have distinct names, got {names}")
    for f in funcs[1:]:
        if f.ndim != funcs[0].ndim:
            raise ValueError(f"cannot fuse {f.ndim}D {f.func_name} with {funcs[0].ndim}D {funcs[0].func_name}")
        if f.dtype != funcs[0].dtype:
            raise ValueError(f"cannot fuse {f.dtype} {f.func_name} with {funcs[0].dtype} {funcs[0].func_name}")
    meta = StencilMeta(name=fused_name(names), ndim=funcs[0].ndim,
                       order=max(f.order for f in funcs), dtype=funcs[0].dtype)
    meta.input_names = inputs
    ctype = "float" if (compute_dtype or meta.dtype) == "float32" else "double"
    seen: dict = {}
    for f in funcs:
        vnames = {}
        for acc in f.accesses:
            vn = _view_name(acc.field, acc.offsets)
            vnames[acc] = vn
            if vn not in seen:
                seen[vn] = AccessInfo(inputs.index(acc.field), acc.offsets, vn)
        r = _Render(f.ndim, ctype, {p: inputs.index(p) for p in f.arg_names}, vnames, mapping)
        meta.expressions.append(r.cpp(f.result))
        meta.py_expressions.append(r.py(f.result))
        meta.flops += count_flops(f)
        meta.constants.update(f.constants)
    meta.accesses = list(seen.values())
    meta.mapping = mapping
    meta.output_names = names
    meta.stencil_inputs = [list(f.arg_names) for f in funcs]
    return meta


UPDATE_MODES = ("euler", "kahan")


def apply_update(meta: StencilMeta, mode: str, dt: float, targets: Sequence[str]) -> StencilMeta:
    """Turn rate stencils into a time update fused into the sweep epilogue.

    Each expression k is read as the rate ``R_k`` of field ``targets[k]``;
    the kernel stores ``u_k + dt * R_k`` (``"euler"``, in the compute type --
    the same operations as a stencil written ``u[i, j] + dt * (...)``), or
    the Kahan-compensated update (``"kahan"``): with the compensation array
    ``c_<target>`` as an extra input and output,
    ``y = dt*R - c;  t = round_storage(u + y);  c' = (t - u) - y;  u' = t``,
    so the low-order part that the storage type drops is carried to the
    next step, and ``t - c'`` equals ``(u - c) + dt*R`` up to the rounding
    of ``c'``.  The rates are evaluated on that compensated state ``u - c``
    (every access of a target field reads both arrays), so the pair behaves
    as one value with about twice the storage mantissa.  Upstream's time
    steppers update in the array's precision (``rk_integrator.py:183-232``)."""
    if mode not in UPDATE_MODES:
        raise ValueError(f"update must be one of {UPDATE_MODES}, got {mode!r}")
    targets = list(targets)
    if len(targets) != meta.num_exprs:
        raise ValueError(f"{meta.name}: {len(targets)} update targets for {meta.num_exprs} rate expression(s)")
    zero = (0,) * meta.ndim
    have = {(a.array_index, tuple(a.offsets)) for a in meta.accesses}

    def centre(idx):
        if (idx, zero) not in have:
            meta.accesses.append(AccessInfo(idx, zero, _view_name(meta.input_names[idx], zero)))
            have.add((idx, zero))

    tf = []
    for t in targets:
        if t not in meta.input_names:
            raise ValueError(f"{meta.name}: update target {t!r} is not an input of the rate stencil(s)")
        tf.append(meta.input_names.index(t))
        centre(tf[-1])
    cf = []
    if mode == "kahan":
        for t in targets:
            meta.input_names.append(f"c_{t}")
            cf.append(len(meta.input_names) - 1)
            centre(cf[-1])
        # the rates see the compensated state u - c at every access of a
        # target field, so the stored pair (u, c) acts as one number with
        # about twice the storage mantissa
        comp_of = {f: c for f, c in zip(tf, cf)}
        for a in list(meta.accesses):
            if a.array_index in comp_of and (comp_of[a.array_index], tuple(a.offsets)) not in have:
                c = comp_of[a.array_index]
                meta.accesses.append(AccessInfo(c, tuple(a.offsets), _view_name(meta.input_names[c], a.offsets)))
                have.add((c, tuple(a.offsets)))
        pat = re.compile(r"a\.template g<(\d+), (-?\d+), (-?\d+), (-?\d+)>\(\)")

        def comp_read(m):
            f = int(m.group(1))
            if f not in comp_of:
                return m.group(0)
            rest = f"{m.group(2)}, {m.group(3)}, {m.group(4)}"
            return f"(a.template g<{f}, {rest}>() - a.template g<{comp_of[f]}, {rest}>())"

        meta.expressions = [pat.sub(comp_read, e) for e in meta.expressions]
        pyn = {meta.input_names[f]: meta.input_names[c] for f, c in comp_of.items()}
        meta.py_expressions = [re.sub(r"\bt_(" + "|".join(pyn) + r")_([0-9mp_]+)",
                                      lambda m: f"(t_{m.group(1)}_{m.group(2)} - t_{pyn[m.group(1)]}_{m.group(2)})", e)
                               for e in meta.py_expressions]
    meta.update = {"mode": mode, "dt": float(dt), "tf": tf, "cf": cf, "targets": targets}
    meta.output_names = targets + ([f"c_{t}" for t in targets] if mode == "kahan" else [])
    return meta


def access_table(meta: StencilMeta) -> list:
    """(field index, dx, dy, dz) of every distinct access on the 3-D map."""
    return [(a.array_index,) + map3(meta.ndim, a.offsets, meta.mapping) for a in meta.accesses]
