"""Boundary pass (upstream ``cutile/passes/boundary.py:45-72``): Dirichlet
-> ``padding_mode = "zero"``, periodic/Neumann/reflecting ->
``boundary_fn = "host_side"``; the raw spec is attached as
``boundary_spec``.  On this backend every kind is applied by the on-device
halo-fill kernel (``st_fill_halo``); the attribute names are kept."""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp


@dataclass(frozen=True)
class BoundaryPass:
    name: ClassVar[str] = "stencil-boundary"

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if not isinstance(op, ApplyOp):
                continue
            b = op.attributes.get("boundary", op.func.boundary)
            if b is None:
                continue
            op.attributes["boundary_spec"] = b
            kind = b.bc_type if hasattr(b, "bc_type") else str(b)
            if kind == "dirichlet":
                op.attributes["padding_mode"] = "zero"
            elif kind in ("periodic", "neumann", "reflecting"):
                op.attributes["boundary_fn"] = "host_side"
