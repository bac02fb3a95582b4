"""Roofline pass (upstream ``cutile/passes/analysis/roofline.py:34-104``):
``flops`` (add/sub/mul/div/max/min/select/abs; neg and constants free),
``unique_loads`` (distinct offset patterns), ``arithmetic_intensity =
flops / ((unique_loads + 1) * dtype_bytes)`` and ``bound``."""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp, Op

COMPUTE_BOUND_THRESHOLD = 10.0
DTYPE_BYTES = {"float64": 8, "float32": 4, "f64": 8, "f32": 4}


@dataclass(frozen=True)
class RooflinePass:
    name: ClassVar[str] = "roofline-analysis"

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if isinstance(op, ApplyOp):
                self._analyze(op)

    @staticmethod
    def _analyze(op: ApplyOp) -> None:
        f = op.func
        flops = sum(1 for v in f.ops if isinstance(v, Op) and v.kind not in ("neg", "cmp"))
        unique = len({a.offsets for a in f.accesses})
        nbytes = DTYPE_BYTES.get(f.dtype, 8)
        ai = flops / ((unique + 1) * nbytes)
        op.attributes["flops"] = flops
        op.attributes["unique_loads"] = unique
        op.attributes["arithmetic_intensity"] = ai
        op.attributes["bound"] = "compute" if ai >= COMPUTE_BOUND_THRESHOLD else "memory"
