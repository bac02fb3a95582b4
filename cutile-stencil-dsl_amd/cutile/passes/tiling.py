"""Tiling pass: workgroup tile shapes for the MI355X sweep kernel.

Same interface as the upstream pass (``cutile/passes/tiling.py:68-207``):
``unique_access_count``, ``TilingPass.tier``, ``TilingPass.candidates`` (a
ranked list of ``(tile_sizes, kernel_hints)``) and the ``tile_sizes`` /
``occupancy`` attributes it attaches.  The shapes are the gfx950 ones:

* ``tile_sizes`` lists, per array dimension, the cells one workgroup
  covers: ``(chunk, rows, cols)`` in 3-D, ``(chunk, cols)`` in 2-D and
  ``(cols,)`` in 1-D; ``chunk`` is how many planes (3-D) or rows (2-D) the
  workgroup streams through with its register queue;
* a workgroup is 256 threads (four wave64s); each thread owns 16-byte
  vectors along the contiguous axis and ``rows / (256 / (cols / vector))``
  rows, so the per-thread register queue holds ``(2r+1) x rows-per-thread``
  vectors;
* the first candidate per tier and element size is the shape measured
  fastest on MI355X at 512^3 among variants that neither spill nor exceed
  256 VGPRs (7-point float64: 16 x 256 rows x cols, 128 planes per chunk,
  70 % of the 8 TB/s HBM peak; 25-point float64: 8 x 128; 25-point float32:
  4 x 512); the rest of the ladder trades rows for registers, then adds an
  ``occupancy`` hint (``amdgpu_waves_per_eu``, a VGPR budget of 512 /
  occupancy);
* the runtime probe (``cutile.runtime.regcheck``) keeps the first candidate
  whose code object neither spills nor exceeds its register budget.
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
    chunk_3d: int = 128
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

    # (rows, cols) per (element bytes, tier) for 3-D; measured on MI355X
    _TILES_3D = {
        (8, "narrow"): (16, 256), (8, "wide"): (8, 128), (8, "very_wide"): (4, 128),
        (4, "narrow"): (16, 512), (4, "wide"): (4, 512), (4, "very_wide"): (4, 256),
        (2, "narrow"): (16, 1024), (2, "wide"): (4, 1024), (2, "very_wide"): (4, 512),
    }

    def row_tile(self, ndim: int, tier: str, chunk: int | None = None) -> tuple:
        v = self.vector
        if ndim == 1:
            return (self.threads_per_block * v,)
        if ndim == 2:
            return (chunk or self.chunk_2d, self.threads_per_block * v)
        rows, cols = self._TILES_3D.get((self.dtype_bytes, tier), (8, 64 * v))
        return (chunk or self.chunk_3d, rows, cols)

    def candidates(self, ndim: int, num_loads: int) -> list:
        tier = self.tier(num_loads)
        first = self.row_tile(ndim, tier)
        if tier == "narrow":
            return [(first, {})]
        smaller = first if ndim < 3 else (first[0], max(1, first[1] // 2), first[2])
        ladder = [(smaller, {"occupancy": occ})
                  for occ in range(self.very_wide_occupancy, self.min_occupancy - 1, -1)]
        ordered = [(first, {}), (smaller, {})] + ladder if tier == "wide" else \
            [(first, {})] + ladder + [(smaller, {})]
        seen: list = []
        for c in ordered:
            if c not in seen:
                seen.append(c)
        return seen

    def tile_for(self, ndim: int, num_loads: int) -> list:
        return list(self.candidates(ndim, num_loads)[0][0])
