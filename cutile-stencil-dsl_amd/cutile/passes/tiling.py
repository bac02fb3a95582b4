"""Tiling pass: workgroup tile shapes for the MI355X sweep kernel.

Same interface as the upstream pass (``cutile/passes/tiling.py:68-207``):
``unique_access_count``, ``TilingPass.tier``, ``TilingPass.candidates`` (a
ranked list of ``(tile_sizes, kernel_hints)``) and the ``tile_sizes`` /
``occupancy`` (and ``threads``) attributes it attaches.  The shapes are the gfx950 ones:

* ``tile_sizes`` lists, per array dimension, the cells one workgroup
  covers: ``(chunk, rows, cols)`` in 3-D, ``(chunk, cols)`` in 2-D and
  ``(cols,)`` in 1-D; ``chunk`` is the longest run of planes (3-D) or rows
  (2-D) one workgroup streams through with its register queue (the launcher
  shortens it until the grid holds one workgroup per CU);
* a workgroup is 256 threads (four wave64s) unless the candidate carries a
  ``threads`` hint; each thread owns 16-byte vectors along the contiguous
  axis and ``rows / (threads / (cols / vector))`` rows, so the per-thread
  register queue holds ``(2r+1) x rows-per-thread`` vectors;
* the first candidate per tier and element size is the shape measured
  fastest on MI355X at 512^3 among variants that neither spill nor exceed
  256 VGPRs (7-point float64: 16 x 256 rows x cols on 512 threads;
  25-point float64: 32 x 64 on 512 threads; 25-point float32: 32 x 128 on
  512 threads); taller tiles cut the halo rows re-read through L2, which is
  what bounds the wide stencils.  The rest of the ladder halves the rows,
  then adds an ``occupancy`` hint (``amdgpu_waves_per_eu``, a VGPR budget of
  512 / occupancy);
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
    chunk_3d: int = 256
    chunk_2d: int = 64

    def apply(self, ctx, module) -> None:
        for op in module.ops:
            if isinstance(op, ApplyOp) and "halo_widths" in op.attributes:
                nd = len(op.attributes["halo_widths"])
                tile, hints = self.candidates(nd, unique_access_count(op))[0]
                op.attributes["tile_sizes"] = list(tile)
                if "occupancy" in hints:
                    op.attributes["occupancy"] = hints["occupancy"]
                if "threads" in hints:
                    op.attributes["threads"] = hints["threads"]

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

    # (rows, cols, threads) per (element bytes, tier) for 3-D; the float64
    # and float32 entries are the fastest non-spilling shapes measured on
    # MI355X at 512^3, the bfloat16 ones keep the float32 register footprint
    _TILES_3D = {
        (8, "narrow"): (16, 256, 512), (8, "wide"): (32, 64, 512), (8, "very_wide"): (4, 128, 256),
        (4, "narrow"): (16, 256, 512), (4, "wide"): (32, 128, 512), (4, "very_wide"): (4, 256, 256),
        (2, "narrow"): (16, 512, 512), (2, "wide"): (8, 512, 512), (2, "very_wide"): (4, 512, 256),
    }

    def row_tile(self, ndim: int, tier: str, chunk: int | None = None) -> tuple:
        return self.shape(ndim, tier, chunk)[0]

    def shape(self, ndim: int, tier: str, chunk: int | None = None) -> tuple:
        """(tile_sizes, threads per workgroup) of the first candidate."""
        v = self.vector
        if ndim == 1:
            return (self.threads_per_block * v,), self.threads_per_block
        if ndim == 2:
            return (chunk or self.chunk_2d, self.threads_per_block * v), self.threads_per_block
        rows, cols, nt = self._TILES_3D.get((self.dtype_bytes, tier), (8, 64 * v, 256))
        return (chunk or self.chunk_3d, rows, cols), nt

    def candidates(self, ndim: int, num_loads: int) -> list:
        tier = self.tier(num_loads)
        first, nt = self.shape(ndim, tier)
        base = {"threads": nt} if nt != self.threads_per_block else {}
        if tier == "narrow":
            return [(first, dict(base))]
        small = dict(base)
        if ndim == 3:
            smaller = (first[0], max(1, first[1] // 2), first[2])
        elif self.vector > 1:
            # 1-D / 2-D rows: half the cells per thread (half-width vectors)
            smaller = first[:-1] + (first[-1] // 2,)
            small["vector"] = self.vector // 2
        else:
            smaller = first
        ladder = [(smaller, dict(small, occupancy=occ))
                  for occ in range(self.very_wide_occupancy, self.min_occupancy - 1, -1)]
        ordered = [(first, dict(base)), (smaller, small)] + ladder if tier == "wide" else \
            [(first, dict(base))] + ladder + [(smaller, small)]
        seen: list = []
        for c in ordered:
            if c not in seen:
                seen.append(c)
        return seen

    def tile_for(self, ndim: int, num_loads: int) -> list:
        return list(self.candidates(ndim, num_loads)[0][0])
