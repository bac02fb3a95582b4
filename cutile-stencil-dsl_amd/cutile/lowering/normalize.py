"""Normalisation: the parsed ``FuncOp`` becomes an ``ApplyOp`` the passes
annotate (upstream ``cutile/lowering/normalize.py:57-156``, which rewrites
the D1 function into ``func.func { stencil.apply }`` and copies the
ndim/order/dtype/boundary/constants metadata)."""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp, FuncOp


@dataclass(frozen=True)
class NormalizePass:
    name: ClassVar[str] = "normalize-stencil"

    def apply(self, ctx, module) -> None:
        new_ops = []
        for op in module.ops:
            if isinstance(op, FuncOp):
                attrs = {"ndim": op.ndim, "order": op.order, "cutile.dtype": op.dtype}
                if op.boundary is not None:
                    attrs["boundary"] = op.boundary
                if op.constants:
                    attrs["constants"] = dict(op.constants)
                new_ops.append(ApplyOp(func=op, attributes=attrs))
            else:
                new_ops.append(op)
        module.ops[:] = new_ops
