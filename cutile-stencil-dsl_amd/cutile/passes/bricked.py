"""Bricked layout pass (upstream ``cutile/passes/bricked.py:23-46``): tags
``layout = "bricked"`` and ``brick_size``."""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp


@dataclass(frozen=True)
class BrickedLayoutPass:
    name: ClassVar[str] = "bricked-layout"

    brick_size: int = 64

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if isinstance(op, ApplyOp):
                op.attributes["layout"] = "bricked"
                op.attributes["brick_size"] = self.brick_size
