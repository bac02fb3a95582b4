"""Stencil IR without xDSL.

The upstream frontend builds an xDSL module with a ``cutile_stencil.func``
op holding ``access`` ops, ``arith``/``math`` ops and a ``yield``
(``cutile/dialects/cutile_stencil/dialect.py:169-369``), normalised to the
``stencil.apply`` form the passes annotate (``cutile/lowering/normalize.py:
57-156``).  xDSL is not available to this build, so the same information is
carried by the small record types below; the attribute names the passes
attach (``halo_widths``, ``tile_sizes``, ``temporal_steps``, ``flops``,
``unique_loads``, ``arithmetic_intensity``, ``bound``, ``padding_mode``,
``boundary_fn``, ...) are kept, with plain Python values.
"""

from __future__ import annotations

import copy
from dataclasses import dataclass, field
from typing import Optional


# ---------------------------------------------------------------------------
# SSA values
# ---------------------------------------------------------------------------

class Value:
    """An SSA value of the stencil body (identity-compared)."""

    __slots__ = ()

    def __hash__(self):
        return id(self)

    def __eq__(self, other):
        return self is other


class Const(Value):
    __slots__ = ("value",)

    def __init__(self, value: float):
        self.value = float(value)

    def __repr__(self):
        return f"Const({self.value!r})"


class Access(Value):
    """Read of field ``field`` (parameter name) at ``offsets``."""

    __slots__ = ("field", "index", "offsets", "index_names")

    def __init__(self, field: str, index: int, offsets: tuple, index_names: tuple = ()):
        self.field = field
        self.index = index
        self.offsets = tuple(int(o) for o in offsets)
        self.index_names = tuple(index_names)

    def __repr__(self):
        return f"Access({self.field}{list(self.offsets)})"


# op kinds -> printed names (the upstream arith/math op names, used by the
# IR printer and by the roofline flop count)
OP_NAMES = {
    "add": "arith.addf", "sub": "arith.subf", "mul": "arith.mulf", "div": "arith.divf",
    "neg": "arith.negf", "max": "arith.maximumf", "min": "arith.minimumf",
    "abs": "math.absf", "cmp": "arith.cmpf", "select": "arith.select",
}
BINARY = ("add", "sub", "mul", "div")
CMP_PREDICATES = ("ogt", "oge", "olt", "ole", "oeq", "une")


class Op(Value):
    __slots__ = ("kind", "operands", "predicate")

    def __init__(self, kind: str, operands: tuple, predicate: Optional[str] = None):
        if kind not in OP_NAMES:
            raise ValueError(f"unknown op kind {kind!r}")
        self.kind = kind
        self.operands = tuple(operands)
        self.predicate = predicate

    @property
    def lhs(self):
        return self.operands[0]

    @property
    def rhs(self):
        return self.operands[1]

    def __repr__(self):
        return f"Op({self.kind})"


# ---------------------------------------------------------------------------
# D1: the stencil function  (dialect.py:169-244)
# ---------------------------------------------------------------------------

@dataclass
class BoundaryAttr:
    """``cutile_stencil.boundary`` (``dialect.py:59-111``)."""

    bc_type: str
    value: Optional[float] = None

    _VALID = ("dirichlet", "neumann", "periodic", "reflecting")

    def __post_init__(self):
        if self.bc_type not in self._VALID:
            raise ValueError(
                f"Invalid boundary type '{self.bc_type}'. Expected one of {sorted(self._VALID)}."
            )

    @property
    def has_value(self) -> bool:
        return self.value is not None

    @property
    def bc_type_str(self) -> str:
        return self.bc_type


@dataclass
class FuncOp:
    """The parsed stencil: unique accesses, SSA ops in program order, result."""

    func_name: str
    ndim: int
    order: int
    dtype: str
    arg_names: list
    index_names: list
    accesses: list                    # unique Access values, first-use order
    ops: list                         # every Const/Access/Op in creation order
    result: Value
    boundary: Optional[BoundaryAttr] = None
    constants: dict = field(default_factory=dict)
    source_file: Optional[str] = None
    source_line: int = 0

    def clone(self) -> "FuncOp":
        return copy.deepcopy(self)

    # -- printing in the upstream textual style (tests grep for op names) --
    def __str__(self) -> str:
        names: dict = {}
        lines = []
        for k, a in enumerate(self.arg_names):
            names[("arg", a)] = f"%arg{k}"
        counter = [0]

        def nm(v):
            if v not in names:
                names[v] = f"%{counter[0]}"
                counter[0] += 1
            return names[v]

        ty = "f64" if self.dtype == "float64" else "f32"
        for v in self.ops:
            if isinstance(v, Const):
                lines.append(f"  {nm(v)} = arith.constant {v.value!r} : {ty}")
            elif isinstance(v, Access):
                idx = ", ".join(f'"{n}"' for n in v.index_names)
                lines.append(
                    f"  {nm(v)} = cutile_stencil.access %arg{v.index} {list(v.offsets)} {{{idx}}} : {ty}"
                )
            else:
                ops = ", ".join(nm(o) for o in v.operands)
                pred = f" {v.predicate}" if v.predicate else ""
                lines.append(f"  {nm(v)} = {OP_NAMES[v.kind]}{pred} {ops} : {ty}")
        lines.append(f"  cutile_stencil.yield {nm(self.result)} : {ty}")
        bnd = f", boundary={self.boundary.bc_type}" if self.boundary else ""
        head = (f'cutile_stencil.func @{self.func_name}(ndim={self.ndim}, order={self.order}, '
                f'dtype="{self.dtype}"{bnd}) {{')
        return "\n".join([head] + lines + ["}"])


@dataclass
class ModuleOp:
    """Container of ops (``builtin.module``)."""

    ops: list

    @property
    def body(self):
        return self

    def clone(self) -> "ModuleOp":
        return copy.deepcopy(self)

    def walk(self):
        for op in self.ops:
            yield op
            if isinstance(op, ApplyOp):
                yield from op.accesses

    def __str__(self) -> str:
        return "builtin.module {\n" + "\n".join(str(o) for o in self.ops) + "\n}"


# ---------------------------------------------------------------------------
# D2: the normalised apply the passes annotate  (normalize.py:57-156)
# ---------------------------------------------------------------------------

@dataclass
class ApplyOp:
    """``stencil.apply``: the stencil body plus the attributes passes attach."""

    func: FuncOp
    attributes: dict = field(default_factory=dict)

    @property
    def accesses(self) -> list:
        return list(self.func.accesses)

    def __str__(self) -> str:
        attrs = ", ".join(f"{k}={v!r}" for k, v in self.attributes.items())
        body = "\n".join(
            "    " + line.replace("cutile_stencil.access", "stencil.access")
            .replace("cutile_stencil.yield", "stencil.return")
            for line in str(self.func).splitlines()[1:-1]
        )
        return (f"func.func @{self.func.func_name}() {{\n  stencil.apply() {{\n{body}\n  }}"
                f" attributes {{{attrs}}}\n}}")


def find_func(module) -> FuncOp:
    for op in module.ops:
        if isinstance(op, FuncOp):
            return op
        if isinstance(op, ApplyOp):
            return op.func
    raise ValueError("No cutile_stencil.FuncOp found in module")


def find_apply(module) -> Optional[ApplyOp]:
    for op in module.ops:
        if isinstance(op, ApplyOp):
            return op
    return None
