"""Temporal blocking pass (upstream ``cutile/passes/temporal.py:54-81``):
the largest ``T <= max_steps`` whose expanded tile, double-buffered, fits
the LDS budget; attaches ``temporal_steps``, ``expanded_halo`` and
``temporal_buffer_count``.  The budget defaults to the MI355X LDS (160 KB
per CU).  ``compile(..., temporal_steps=T)`` overrides the choice."""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp


@dataclass(frozen=True)
class TemporalBlockingPass:
    name: ClassVar[str] = "temporal-blocking"

    max_steps: int = 8
    shared_mem_bytes: int = 160 * 1024
    dtype_bytes: int = 8

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if isinstance(op, ApplyOp):
                self._process(op)

    def _process(self, op: ApplyOp) -> None:
        tile = op.attributes.get("tile_sizes")
        halo = op.attributes.get("halo_widths")
        if tile is None or halo is None:
            return
        best = 1
        for t in range(self.max_steps, 1, -1):
            expanded = [tile[d] + 2 * t * halo[d] for d in range(len(tile))]
            if math.prod(expanded) * self.dtype_bytes * 2 <= self.shared_mem_bytes:
                best = t
                break
        op.attributes["temporal_steps"] = best
        op.attributes["expanded_halo"] = [best * h for h in halo]
        op.attributes["temporal_buffer_count"] = best + 1
