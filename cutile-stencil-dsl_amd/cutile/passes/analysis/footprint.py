"""Footprint pass: ``halo_widths`` = max |offset| per dimension
(upstream ``cutile/passes/analysis/footprint.py:20-64``)."""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp


@dataclass(frozen=True)
class AnalysisPass:
    name: ClassVar[str] = "stencil-analysis"

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if isinstance(op, ApplyOp):
                self._analyze(op)

    @staticmethod
    def _analyze(op: ApplyOp) -> None:
        accs = op.func.accesses
        if not accs:
            return
        nd = len(accs[0].offsets)
        mx = [0] * nd
        for a in accs:
            for d, o in enumerate(a.offsets[:nd]):
                mx[d] = max(mx[d], abs(o))
        op.attributes["halo_widths"] = mx
