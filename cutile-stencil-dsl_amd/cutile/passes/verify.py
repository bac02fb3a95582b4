"""IR verification (upstream ``cutile/passes/verify.py:54-103``)."""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp


def _diag(op: ApplyOp, msg: str) -> str:
    f = op.func
    if f.source_file:
        return f"{f.source_file}:{f.source_line} -- {msg}"
    return msg


@dataclass(frozen=True)
class VerifyPass:
    name: ClassVar[str] = "verify-ir"

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if isinstance(op, ApplyOp):
                self._verify(op)

    @staticmethod
    def _pow2(n: int) -> bool:
        return n > 0 and (n & (n - 1)) == 0

    def _verify(self, op: ApplyOp) -> None:
        a = op.attributes
        tile = a.get("tile_sizes")
        if tile is not None:
            for k, t in enumerate(tile):
                if not self._pow2(t):
                    raise ValueError(_diag(op, f"verify-ir: tile_sizes[{k}] = {t} is not a power of 2"))
            if "halo_widths" not in a:
                raise ValueError(_diag(op, "verify-ir: tile_sizes is set but halo_widths is missing "
                                           "(AnalysisPass must run before TilingPass)"))
        if "temporal_steps" in a and a["temporal_steps"] < 1:
            raise ValueError(_diag(op, f"verify-ir: temporal_steps = {a['temporal_steps']} must be >= 1"))
        if "num_gpus" in a and a["num_gpus"] < 1:
            raise ValueError(_diag(op, f"verify-ir: num_gpus = {a['num_gpus']} must be >= 1"))
