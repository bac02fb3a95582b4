"""Composable pass pipeline (upstream ``cutile/runtime/pipeline.py:17-65``)."""

from __future__ import annotations

from dataclasses import dataclass, field


@dataclass
class Pipeline:
    passes: list = field(default_factory=list)

    def add(self, pass_) -> "Pipeline":
        self.passes.append(pass_)
        return self

    def run(self, module):
        for p in self.passes:
            p.apply(None, module)
        return module

    @classmethod
    def single_gpu(cls, shared_mem_bytes: int = 160 * 1024, dtype_bytes: int = 8,
                   max_temporal_steps: int = 8) -> "Pipeline":
        """normalize -> analysis -> roofline -> tiling -> boundary -> temporal."""
        from cutile.lowering.normalize import NormalizePass
        from cutile.passes import (AnalysisPass, BoundaryPass, RooflinePass, TemporalBlockingPass,
                                   TilingPass)

        p = cls()
        p.add(NormalizePass())
        p.add(AnalysisPass())
        p.add(RooflinePass())
        p.add(TilingPass(shared_mem_bytes=shared_mem_bytes, dtype_bytes=dtype_bytes))
        p.add(BoundaryPass())
        p.add(TemporalBlockingPass(max_steps=max_temporal_steps, shared_mem_bytes=shared_mem_bytes,
                                   dtype_bytes=dtype_bytes))
        return p
