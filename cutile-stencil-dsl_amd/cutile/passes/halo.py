"""Halo-exchange pass (upstream ``cutile/passes/halo.py``): a no-op at one
GPU (``:50-52``); halos are filled on-device by ``st_fill_halo``."""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp


@dataclass(frozen=True)
class HaloExchangePass:
    name: ClassVar[str] = "halo-exchange"

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if isinstance(op, ApplyOp) and op.attributes.get("num_gpus", 1) > 1:
                op.attributes["halo_exchange_enabled"] = 1
                op.attributes["halo_exchange_width"] = list(op.attributes.get("halo_widths", []))
