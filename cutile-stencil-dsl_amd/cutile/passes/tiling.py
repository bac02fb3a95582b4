"""Tiling pass: workgroup tile shapes for the MI355X sweep kernel.

Same interface as the upstream pass (``cutile/passes/tiling.py:68-207``):
``unique_access_count``, ``TilingPass.tier``, ``TilingPass.candidates`` (a
ranked list of ``(tile_sizes, kernel_hints)``) and the ``tile_sizes`` /
``occupancy`` attributes it attaches.  The shapes are re-derived for gfx950:

* a workgroup is ``threads_per_block`` (256) threads = four wave64s, each
  thread owning ``16 // dtype_bytes`` contiguous elements (one 16-byte
  vector), so the contiguous tile extent is 256 columns for float64 rows
  of 2 KB per wave group, and rows are added along axis 1 (3-D);
* ``tile_sizes`` lists, per array dimension, the cells one workgroup
  covers: ``(chunk, rows, cols)`` in 3-D, ``(chunk, cols)`` in 2-D and
  ``(cols,)`` in 1-D; ``chunk`` is how many planes (3-D) or rows (2-D) the
  workgroup streams through with its register queue;
* narrow stencils (<= 16 distinct accesses) take 2 rows per thread, wide
  ones 1 row (the register queue holds 2r+1 planes per owned cell), very
  wide ones add an ``occupancy`` hint, which the emitter turns into
  ``amdgpu_waves_per_eu`` (resident waves per SIMD) and therefore a VGPR
  budget of 512 / occupancy;
* the ladder after the first candidate trades rows for occupancy the same
  way the upstream ladder does, and the runtime probe
  (``cutile.runtime.regcheck``) keeps the first that neither spills nor
  exceeds its register budget.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import ClassVar

from cutile.frontend.ir import ApplyOp


def unique_access_count(apply_op) -> int:
    """Distinct (field, offset) reads of the stencil (one load each)."""
    func = apply_op.func if isinstance(apply_op, ApplyOp) else apply_op
    return len({(a.field, a.offsets) for a in func.accesses})


@dataclass(frozen=True)
class TilingPass:
    name: ClassVar[str] = "stencil-tiling"

    shared_mem_bytes: int = 160 * 1024
    dtype_bytes: int = 8
    threads_per_block: int = 256
    wide_stencil_loads: int = 16
    very_wide_stencil_loads: int = 64
    narrow_rows_per_thread: int = 2
    wide_rows_per_thread: int = 1
    very_wide_occupancy: int = 4
    min_occupancy: int = 2
    chunk_3d: int = 64
    chunk_2d: int = 64

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if isinstance(op, ApplyOp) and "halo_widths" in op.attributes:
                nd = len(op.attributes["halo_widths"])
                tile, hints = self.candidates(nd, unique_access_count(op))[0]
                op.attributes["tile_sizes"] = list(tile)
                if "occupancy" in hints:
                    op.attributes["occupancy"] = hints["occupancy"]

    def tier(self, num_loads: int) -> str:
        if num_loads <= self.wide_stencil_loads:
            return "narrow"
        if num_loads > self.very_wide_stencil_loads:
            return "very_wide"
        return "wide"

    def is_wide(self, num_loads: int) -> bool:
        return self.tier(num_loads) == "wide"

    @property
    def vector(self) -> int:
        return max(1, 16 // self.dtype_bytes)

    def row_tile(self, ndim: int, rows_per_thread: int, chunk: int | None = None) -> tuple:
        """3-D: one wave64 spans a 1 KB row segment (64 vectors) and the
        workgroup's waves stack along axis 1; 2-D/1-D: the whole workgroup
        spans one row segment of ``threads_per_block`` vectors."""
        v = self.vector
        if ndim == 1:
            return (self.threads_per_block * v,)
        if ndim == 2:
            return (chunk or self.chunk_2d, self.threads_per_block * v)
        waves = max(1, self.threads_per_block // 64)
        return (chunk or self.chunk_3d, waves * rows_per_thread, 64 * v)

    def candidates(self, ndim: int, num_loads: int) -> list:
        tier = self.tier(num_loads)
        narrow = self.row_tile(ndim, self.narrow_rows_per_thread)
        one = self.row_tile(ndim, self.wide_rows_per_thread)
        if tier == "narrow":
            return [(narrow, {})]
        ladder = [(one, {"occupancy": occ})
                  for occ in range(self.very_wide_occupancy, self.min_occupancy - 1, -1)]
        unhinted = (one, {})
        ordered = [unhinted] + ladder if tier == "wide" else ladder + [unhinted]
        seen: list = []
        for c in ordered:
            if c not in seen:
                seen.append(c)
        return seen

    def tile_for(self, ndim: int, num_loads: int) -> list:
        return list(self.candidates(ndim, num_loads)[0][0])
