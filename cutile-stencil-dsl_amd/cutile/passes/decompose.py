"""Domain decomposition pass (upstream ``cutile/passes/decompose.py``):
a no-op at ``num_gpus <= 1`` (``:50-51``), which is the only configuration
this single-GPU backend builds; for more GPUs it records the slab split
attributes so the IR stays inspectable."""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp


@dataclass(frozen=True)
class DecompositionPass:
    name: ClassVar[str] = "domain-decomposition"

    num_gpus: int = 1
    split_axis: int = 0
    topology: str = "linear"

    def apply(self, ctx, module) -> None:
        if self.num_gpus <= 1:
            return
        for op in module.ops:
            if isinstance(op, ApplyOp):
                op.attributes["num_gpus"] = self.num_gpus
                op.attributes["split_axis"] = self.split_axis
                op.attributes["topology"] = self.topology
