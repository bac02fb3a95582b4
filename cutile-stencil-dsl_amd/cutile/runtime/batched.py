"""Batched execution: many instances of one stencil, of uneven sizes, in one
persistent launch driven by a device tile work-list.

There is no upstream counterpart (upstream launches one kernel per array);
this is the batching runtime the serving loop and the batched benchmark use.

* Instances are bucketed by size; each bucket has its own tile shape (and,
  for 2-D, its own mapping: small grids are tiled as one plane so every
  thread of a workgroup has work, large grids stream along their rows).
  The kernel contains one sweep instantiation per bucket and dispatches on
  the work item's bucket.
* The host builds, once per set of instances, a device table of per-instance
  launch arguments (``st::SweepArgs``), the first work item of every
  instance and its tile bucket; item i belongs to the instance found by a
  binary search over the first items and is decoded into (tile z, tile y,
  chunk) from that instance's tile counts.  A fixed grid of
  ``occupancy x CUs`` workgroups strides through the items, so the launch
  shape does not depend on the batch and a captured hipGraph stays valid.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from cutile.runtime import rt


@dataclass(frozen=True)
class Bucket:
    """Tile configuration for instances whose contiguous extent is at most
    ``max_cols`` (``None``: unbounded)."""

    mapping: str          # "plane" or "rows" (2-D); "rows" for 3-D
    ty: int
    tz: int
    xch: int = 1
    max_cols: Optional[int] = None


DEFAULT_BUCKETS_2D = (
    Bucket("plane", 16, 64, 1, 160),
    Bucket("plane", 8, 128, 1, 384),
    Bucket("rows", 1, 1024, 16, None),
)


def upload(a: np.ndarray):
    """Device copy of a small host table, queued on the current stream
    without blocking the host: staged through pinned memory, which the
    caller keeps alive (returned second) until the stream has used it."""
    import torch

    host = torch.from_numpy(np.array(a, copy=True)).pin_memory()
    return host.to("cuda", non_blocking=True), host


class _BatchArgs(ctypes.Structure):
    """Kernel arguments of a batched launch (the kernel's parameter list)."""
    _fields_ = [("insts", ctypes.c_void_p), ("first", ctypes.c_void_p), ("bucket", ctypes.c_void_p),
                ("ninst", ctypes.c_int), ("nitems", ctypes.c_int)]


@dataclass
class BatchPlan:
    table: object                 # device uint8 tensor holding SweepArgs[]
    first: object                 # device int32: first work item of every instance (+ total)
    bucket: object                # device int32: tile bucket of every instance
    nitems: int
    grid: int
    cells: int
    instances: int
    keep: list = field(default_factory=list)   # tensors referenced by the table
    # (device, function, packed kernel arguments) of the last launch: a
    # plan replayed every step packs its arguments once
    packed: Optional[tuple] = field(default=None, repr=False)


class BatchedStencil:
    """One persistent kernel for many arrays of the same stencil."""

    def __init__(self, stencil_fn, storage: Optional[str] = None, buckets: Sequence[Bucket] = None,
                 occupancy: int = 8, threads: int = 256):
        from cutile.frontend.ir import find_func
        from cutile.lowering.hip_emitter import KernelConfig, emit_traits
        from cutile.lowering.meta import extract_fused_meta, extract_meta
        from cutile.lowering.stencil_to_hip import storage_compute

        fns = list(stencil_fn) if isinstance(stencil_fn, (list, tuple)) else [stencil_fn]
        funcs = [find_func(f._ir) for f in fns]
        func = funcs[0]
        if len(fns) > 1:
            probe = extract_fused_meta([f._ir for f in fns])
            self.name = probe.name
            self.inputs = tuple(probe.input_names)
            self.outputs = tuple(probe.output_names)
        else:
            self.name = func.func_name
            self.inputs = tuple(func.arg_names)
            self.outputs = (func.func_name,)
        self.ndim = func.ndim
        self.storage, self.compute = storage_compute(func.dtype, storage)
        esz = {"float64": 8, "float32": 4, "bfloat16": 2}[self.storage]
        self.vz = 16 // esz
        if buckets is None:
            if self.ndim == 2:
                buckets = [Bucket(b.mapping, b.ty, b.tz * self.vz // 4, b.xch, b.max_cols)
                           for b in DEFAULT_BUCKETS_2D]
            else:
                buckets = [Bucket("rows", 4, 64 * self.vz, 16, None)]
        self.buckets = list(buckets)
        self.occupancy = occupancy
        self.threads = int(threads)
        self.num_inputs = len(self.inputs)
        self.num_outputs = len(self.outputs)
        parts = ['#include "stencil_core.h"', ""]
        cases = []
        self.cfgs = []
        for k, b in enumerate(self.buckets):
            mp = b.mapping if self.ndim == 2 else "rows"
            meta = (extract_fused_meta([f._ir for f in fns], mapping=mp) if len(fns) > 1
                    else extract_meta(func, mapping=mp))
            ntz = b.tz // self.vz
            if threads % ntz:
                raise ValueError(f"bucket {b}: {ntz} threads per row must divide {threads}")
            cfg = KernelConfig(ty=b.ty, tz=b.tz, vz=self.vz, nt=threads)
            self.cfgs.append(cfg)
            parts.append(emit_traits(meta, cfg, self.storage, self.compute, f"b{k}", partial=True))
            cases.append(f"      case {k}: st::sweep_tile<P_b{k}>(a, tz, ty, ch, reinterpret_cast<P_b{k}::C*>(pool)); break;")
        self.symbol = f"{self.name}_{self.storage}_batched"
        # The work-list is implicit: instance k owns items first[k] ..
        # first[k+1]-1, decoded into (tile z, tile y, chunk) from its own
        # tile counts; a workgroup finds the instance of item i by binary
        # search over first[] (uniform, so scalar loads).  The host uploads
        # only per-instance records, so the list is cheap to rebuild.
        parts.append(
            f'extern "C" __global__ void __launch_bounds__({threads}) {self.symbol}('
            "const st::SweepArgs* __restrict__ insts, const int* __restrict__ first, "
            "const int* __restrict__ bucket, int ninst, int nitems) {\n"
            # one LDS pool for every bucket, sized to the largest: a workgroup
            # runs one item at a time, and each item's first LDS write
            # follows a workgroup barrier
            f"  __shared__ __attribute__((aligned(16))) unsigned char pool[st::sweep_lds_bytes<"
            f"{', '.join(f'P_b{k}' for k in range(len(self.buckets)))}>()];\n"
            "  for (int i = blockIdx.x; i < nitems; i += gridDim.x) {\n"
            "    int lo = 0, hi = ninst - 1;\n"
            "    while (lo < hi) {\n"
            "      const int mid = (lo + hi + 1) >> 1;\n"
            "      if (first[mid] <= i) lo = mid; else hi = mid - 1;\n"
            "    }\n"
            "    const int inst = __builtin_amdgcn_readfirstlane(lo);\n"
            "    const st::SweepArgs& a = insts[inst];\n"
            "    const int local = i - first[inst];\n"
            "    const int tz = local % a.ntz, rest = local / a.ntz;\n"
            "    const int ty = rest % a.nty, ch = rest / a.nty;\n"
            "    switch (bucket[inst]) {\n" + "\n".join(cases) + "\n      default: break;\n    }\n"
            "  }\n}"
        )
        self.source = "\n".join(parts)
        self._fn = {}

    def _function(self):
        import torch

        from cutile.runtime.kernel import load_checked

        dev = torch.cuda.current_device()
        if dev not in self._fn:
            # refused before loading when the code object uses scratch,
            # more than 256 VGPR+AGPR or more than 160 KiB of LDS
            _mod, fns = load_checked(self.source, f"{self.symbol}.hip", [self.symbol])
            self._fn[dev] = fns[self.symbol]
        return self._fn[dev]

    def resources(self):
        return rt.function_resources(self._function())

    # ------------------------------------------------------------------
    def bucket_of(self, n_cols: int) -> int:
        for k, b in enumerate(self.buckets):
            if b.max_cols is None or n_cols <= b.max_cols:
                return k
        return len(self.buckets) - 1

    def record(self, ins: Sequence, outs: Sequence, halo: Sequence[int], inner: Sequence[int]):
        """Launch record of one instance: ``(args bytes, work items, bucket,
        interior cells)`` or ``None`` for an empty interior.  Records can be
        cached and assembled into plans cheaply (the serving loop rebuilds
        its work-lists from cached records)."""
        from cutile.runtime.kernel import SweepArgs

        nd = self.ndim
        if len(ins) != self.num_inputs or len(outs) != self.num_outputs:
            raise ValueError(f"expected {self.num_inputs} inputs and {self.num_outputs} outputs")
        shape = tuple(ins[0].shape)
        inner = tuple(inner)
        if any(n + 2 * h > s for n, h, s in zip(inner, halo, shape)):
            raise ValueError(f"interior {inner} + halo exceeds the array {shape}")
        if min(inner) <= 0:
            return None
        k = self.bucket_of(inner[-1])
        b = self.buckets[k]
        if nd == 2:
            if b.mapping == "plane":
                s3, h3, n3 = (1, shape[0], shape[1]), (0, halo[0], halo[1]), (1, inner[0], inner[1])
            else:
                s3, h3, n3 = (shape[0], 1, shape[1]), (halo[0], 0, halo[1]), (inner[0], 1, inner[1])
        elif nd == 1:
            s3, h3, n3 = (1, 1, shape[0]), (0, 0, halo[0]), (1, 1, inner[0])
        else:
            s3, h3, n3 = shape, tuple(halo), inner
        if h3[2] % self.vz or s3[2] % self.vz:
            raise ValueError(f"contiguous extent/halo not a multiple of {self.vz}")
        if s3[1] * s3[2] * ins[0].element_size() >= 1 << 31:
            raise ValueError("a plane exceeds the 2 GiB range of one buffer resource")
        a = SweepArgs()
        for f, t in enumerate(ins):
            a.inp[f] = t.data_ptr()
        for m, t in enumerate(outs):
            a.out[m] = t.data_ptr()
        for t in list(ins) + list(outs):
            if t.data_ptr() % 16:
                raise ValueError("arrays must be 16-byte aligned")
        a.s0, a.s1, a.s2 = s3
        a.h0, a.h1, a.h2 = h3
        a.n0, a.n1, a.n2 = n3
        a.ntz = -(-n3[2] // b.tz)
        a.nty = -(-n3[1] // b.ty)
        a.xch = max(1, min(b.xch, n3[0]))
        a.nch = -(-n3[0] // a.xch)
        a.nxcd = 1
        return bytes(a), a.nch * a.nty * a.ntz, k, math.prod(inner)

    def assemble(self, records: Sequence, keep: Sequence = ()) -> BatchPlan:
        """One plan (device tables of launch records, first item of every
        instance, bucket of every instance) from records of :meth:`record`;
        ``keep`` holds tensors the plan uses."""
        from cutile.runtime.kernel import cu_count

        recs = [r for r in records if r is not None]
        if recs:
            blob = b"".join(r[0] for r in recs)
            first = np.zeros(len(recs) + 1, dtype=np.int32)
            np.cumsum([r[1] for r in recs], out=first[1:])
            bucket = np.array([r[2] for r in recs], dtype=np.int32)
        else:
            blob, first, bucket = bytes(1), np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32)
        dtab, htab = upload(np.frombuffer(blob, dtype=np.uint8))
        dfirst, hfirst = upload(first)
        dbucket, hbucket = upload(bucket)
        nitems = int(first[-1])
        grid = max(1, min(nitems, self.occupancy * cu_count()))
        return BatchPlan(dtab, dfirst, dbucket, nitems, grid, int(sum(r[3] for r in recs)), len(recs),
                         list(keep) + [htab, hfirst, hbucket])

    def prepare(self, instances: Sequence, halo: Sequence[int], interiors: Optional[Sequence] = None) -> BatchPlan:
        """``instances``: sequence of ``(inputs, outputs)`` tensor lists, all
        with the same leading halo; returns the device tables for one launch.

        ``interiors`` gives each instance's interior extents when its arrays
        are padded beyond the symmetric halo on the high side (rows padded to
        a multiple of the 16-byte vector keep every row aligned)."""
        recs, keep = [], []
        for idx, (ins, outs) in enumerate(instances):
            inner = tuple(interiors[idx]) if interiors is not None else \
                tuple(s - 2 * h for s, h in zip(ins[0].shape, halo))
            try:
                recs.append(self.record(ins, outs, halo, inner))
            except ValueError as e:
                raise ValueError(f"instance {idx}: {e}") from None
            keep += list(ins) + list(outs)
        return self.assemble(recs, keep)

    def launch(self, plan: BatchPlan, stream=None) -> None:
        from cutile.runtime.kernel import default_stream

        if plan.nitems == 0:
            return
        fn = self._function()
        if plan.packed is None or plan.packed[0] != fn:
            args = _BatchArgs(plan.table.data_ptr(), plan.first.data_ptr(), plan.bucket.data_ptr(), plan.instances,
                              plan.nitems)
            plan.packed = (fn, args, ctypes.addressof(args), ctypes.sizeof(args))
        _fn, _args, addr, size = plan.packed
        rt.check(rt.lib().st_launch(fn, plan.grid, 1, 1, self.threads, 1, 1, 0,
                                    stream if stream is not None else default_stream(), addr, size), "st_launch")


class LeapfrogBatch:
    """Two-level (leapfrog) stepping of many instances with one batched
    stencil ``f(u, uold) -> u_new``: three buffers per instance rotate, so
    three prepared plans cover every step and a graph of three steps can be
    replayed indefinitely.  Halos are never written by the sweep, so zero
    (Dirichlet) halos stay zero."""

    def __init__(self, bs: BatchedStencil, pairs: Sequence, halo: Sequence[int],
                 interiors: Optional[Sequence] = None):
        import torch

        self.bs = bs
        self.halo = tuple(halo)
        self.bufs = [(u, uold, torch.zeros_like(u)) for u, uold in pairs]
        rots = [(0, 1, 2), (2, 0, 1), (1, 2, 0)]
        self.plans = [bs.prepare([([b[i], b[j]], [b[k]]) for b in self.bufs], self.halo, interiors)
                      for i, j, k in rots]
        self.steps = 0

    def step(self, stream=None) -> None:
        self.bs.launch(self.plans[self.steps % 3], stream)
        self.steps += 1

    def current(self, idx: int):
        """(u, uold) of instance ``idx`` after the steps taken so far."""
        b = self.bufs[idx]
        r = self.steps % 3
        return [(b[0], b[1]), (b[2], b[0]), (b[1], b[2])][r]

    @property
    def cells(self) -> int:
        return self.plans[0].cells
