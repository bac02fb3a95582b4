"""Device time marching: the step loop around a compiled stencil.

Upstream leaves the loop to user code -- ``apply_boundary_<name>`` then
``launch_<name>`` then swap (``examples/wave_2d.py:62-68``,
``examples/gray_scott.py:78-94``), with the semantics of
``cutile/reference/stencil_ref.py:208-245`` (stencil, then boundary, per
step).  :class:`TimeStepper` runs that loop on the device:

* the halo is refilled on the device after every step (``st_fill_halo``);
* with ``temporal_steps=T`` one launch advances T steps in LDS/registers:
  zero-Dirichlet problems use the zero-halo level mode, periodic ones the
  recompute mode on a halo T*r deep, anything else steps one at a time;
  a remainder of ``steps % T`` steps runs through the single-step kernel;
* the loop can be captured once as a hipGraph and replayed, which removes
  the per-launch host cost that dominates small grids;
* fused multi-field stencils (``compile_fused``) map member k's output back
  to field ``targets[k]`` after every step (all members read the old state).
"""

from __future__ import annotations

from typing import Optional, Sequence

from cutile.runtime import boundary as _boundary
from cutile.runtime import rt


def _bc_kind(spec, ndim) -> str:
    spec = _boundary.normalize_spec(spec, ndim)
    kinds = set()
    for lo, hi in spec.values():
        for s in (lo, hi):
            kinds.add(s if isinstance(s, str) else (s[0] if s[1] == 0.0 else "dirichlet_value"))
    if kinds == {"periodic"} and len(spec) == ndim:
        return "periodic"
    if kinds == {"dirichlet"} and len(spec) == ndim:
        return "dirichlet0"
    if kinds <= {"dirichlet", "dirichlet_value"} and len(spec) == ndim:
        return "dirichlet"
    return "other"


class HaloPolicy:
    """How the halo is kept valid between steps:

    * ``"wrap"``: all axes periodic -- the sweep reads periodically past
      the interior (``periodic_axes``), so no fill between steps; one fill
      at the end leaves the result's halo as a per-step fill would;
    * ``"static"``: every side a Dirichlet value -- the sweep never writes
      halo cells, so both buffers are filled once and stay valid;
    * ``"fill"``: anything else -- a halo fill after every step."""

    def __init__(self, spec, ndim: int, halo):
        kind = _bc_kind(spec, ndim)
        self.mode = {"periodic": "wrap", "dirichlet0": "static", "dirichlet": "static"}.get(kind, "fill")
        self.ndim = ndim
        self.halo = tuple(halo)

    def axes_for(self, u) -> tuple:
        """Periodic axes the sweep may wrap for array ``u`` (interior >= halo)."""
        if self.mode != "wrap":
            return ()
        ok = all(s - 2 * h >= h for s, h in zip(u.shape, self.halo))
        return tuple(range(self.ndim)) if ok else ()


class TimeStepper:
    """Advance one stencil (single field) on the device.

    ``stencil_fn`` is a ``@stencil`` function; ``boundary`` any spec accepted
    by ``apply_boundary``."""

    def __init__(self, stencil_fn, boundary, temporal_steps: int = 1, storage: Optional[str] = None,
                 tile_sizes=None, kernel_hints: Optional[dict] = None, block_hints: Optional[dict] = None,
                 split: str = "auto"):
        """``split``: how a block of ``temporal_steps`` steps is launched.
        ``"fit"``: one launch of depth T, or, where T levels do not fit the
        registers and LDS, T/d launches of the largest depth d dividing T
        that fits.  ``"auto"`` (default): on the first array of a given
        shape, every depth d dividing T (d = 1: single sweeps) is timed on
        scratch arrays of that shape and the fastest is used for that shape
        (the states are the same for every d; cached per stencil, storage,
        size bucket, T and device architecture, see
        ``cutile.runtime.autotune.autotune_split``).  Without a device
        ``"auto"`` behaves as ``"fit"``."""
        from cutile.runtime.launcher import compile as _compile

        if split not in ("auto", "fit"):
            raise ValueError(f"split must be 'auto' or 'fit', got {split!r}")
        self.fn = stencil_fn
        self.spec = boundary
        self.split = split
        self._storage_arg, self._block_hints = storage, block_hints
        self._plans: dict = {}
        self.T = int(temporal_steps)
        from cutile.frontend.ir import find_func

        kind = _bc_kind(boundary, find_func(stencil_fn._ir).ndim)
        self.mode = "recompute" if kind == "periodic" else "zero"
        if self.T > 1 and kind in ("other", "dirichlet"):
            # no exact multi-level semantics: intermediate levels of a
            # temporal block hold 0 (or recomputed values) outside the
            # interior, not a refilled boundary value
            self.T = 1
        self.single = _compile(stencil_fn, temporal_blocking=False, storage=storage, tile_sizes=tile_sizes,
                               kernel_hints=kernel_hints)
        self.block_reps = 1
        if self.T > 1:
            self.block = _compile(stencil_fn, temporal_steps=self.T, temporal_mode=self.mode, storage=storage,
                                  kernel_hints=block_hints)
            if "tblock" not in self.block.load_module()._VARIANTS:
                # the ghost ring of T levels fits neither registers nor LDS:
                # a block of T steps runs as T/d launches of the largest depth
                # d dividing T that fits (same states for periodic and zero
                # Dirichlet problems: the launches chain through zero or
                # periodic halos exactly as T single steps do)
                for d in range(self.T - 1, 1, -1):
                    if self.T % d:
                        continue
                    alt = _compile(stencil_fn, temporal_steps=d, temporal_mode=self.mode, storage=storage,
                                   kernel_hints=block_hints)
                    if "tblock" in alt.load_module()._VARIANTS:
                        self.block, self.block_reps = alt, self.T // d
                        break
            self.halo = self.block.halo_widths
        else:
            self.block = None
            self.halo = self.single.halo_widths
        self._single = self.single.load_module()._KERNEL
        self._block = self.block.load_module()._KERNEL if self.block else None
        self.ndim = self.single.ndim
        self.storage = self.single.storage
        self.policy = HaloPolicy(boundary, self.ndim, self.halo)

    # -- arrays ---------------------------------------------------------
    def empty_like_interior(self, interior_shape, device="cuda"):
        import torch

        shape = tuple(n + 2 * h for n, h in zip(interior_shape, self.halo))
        return torch.zeros(shape, dtype=getattr(torch, self.storage), device=device)

    def fill(self, u, stream=None):
        _boundary.fill(u, self.spec, self.halo, stream)

    # -- stepping -------------------------------------------------------
    def _one(self, a, b, stream, axes=()):
        self._single.launch([a], [b], self.halo, stream, validate=False, periodic_axes=axes)
        if not axes and self.policy.mode != "static":
            self.fill(b, stream)
        return b, a

    # -- depth split by measured speed ------------------------------------
    def depth_kernel(self, d: int):
        """Compiled block kernel of depth ``d`` (a divisor of T, >= 2) in this
        stepper's mode, or None when d levels fit neither registers nor LDS."""
        from cutile.runtime.launcher import compile as _compile

        if d == self.T // self.block_reps and self._block is not None:
            return self._block
        r = _compile(self.fn, temporal_steps=d, temporal_mode=self.mode, storage=self._storage_arg,
                     kernel_hints=self._block_hints)
        k = r.load_module()._KERNEL
        return k if "tblock" in k.variants else None

    def plan_for(self, u) -> tuple:
        """(kernel, launches per block) used for arrays shaped like ``u``:
        the fit-based choice, or under ``split="auto"`` the measured fastest
        (kernel None: T single sweeps)."""
        key = (tuple(u.shape), u.dtype)
        plan = self._plans.get(key)
        if plan is None:
            plan = (self._block, self.block_reps)
            if self.split == "auto" and self._block is not None:
                import torch

                if torch.cuda.is_available() and u.is_cuda:
                    from cutile.runtime.autotune import autotune_split

                    d = autotune_split(self, u)
                    plan = (None, self.T) if d == 1 else (self.depth_kernel(d), self.T // d)
            self._plans[key] = plan
        return plan

    def _blk(self, a, b, stream, wrap=()):
        # the halo of `a` is current (every step or block ends with a fill),
        # or `wrap` names the periodic axes the block reads around
        kern, reps = self.plan_for(a)
        if kern is None:                      # measured fastest: T single sweeps
            axes = self.policy.axes_for(a)
            if wrap and not axes:
                self.fill(a, stream)          # earlier wrapped launches left it stale
            for _ in range(self.T):
                a, b = self._one(a, b, stream, axes)
            return a, b
        for _ in range(reps):
            kern.launch_temporal([a], [b], self.halo, stream,
                                        mode=1 if (self.mode == "recompute" and not wrap) else 0,
                                        validate=False, periodic_axes=wrap)
            if not wrap and self.policy.mode != "static":   # a Dirichlet halo is never written: it stays valid
                self.fill(b, stream)
            a, b = b, a
        return a, b

    def _block_wrap(self, u=None) -> tuple:
        """Periodic axes a temporal block reads around (no halo fills between
        blocks): all of them for a periodic problem on the register/LDS
        kernel, which wraps any distance."""
        kern = self._block if u is None else self.plan_for(u)[0]
        if self.policy.mode == "wrap" and (kern is None or "tblock" in kern.variants):
            return tuple(range(self.ndim))
        return ()

    def enqueue(self, a, b, steps: int, stream=None):
        """Enqueue ``steps`` steps from ``a`` (halo already valid) using ``b``
        as scratch; returns ``(result, other)``."""
        if stream is None:
            from cutile.runtime.kernel import default_stream

            stream = default_stream()
        if self.policy.mode == "static":
            self.fill(b, stream)          # constant halo: valid in both buffers from here on
        stale = False                     # the halo of `a` is out of date (wrapped launches)
        if self._block is not None:
            bw = self._block_wrap(a)
            nblk = steps // self.T
            for _ in range(nblk):
                a, b = self._blk(a, b, stream, bw)
            stale = bool(bw) and nblk > 0
            steps -= nblk * self.T
        axes = self.policy.axes_for(a)
        if stale and steps and not axes:
            self.fill(a, stream)
            stale = False
        for _ in range(steps):
            a, b = self._one(a, b, stream, axes)
        if stale or (axes and steps):
            self.fill(a, stream)          # wrapped launches left the halo stale
        return a, b

    def march(self, u, steps: int, scratch=None, stream=None, graph: bool = False, graph_steps: int = 0,
              residual: bool = False):
        """Run ``steps`` steps starting from device array ``u`` (its halo is
        filled first).  With ``graph=True`` the loop is captured once (or
        ``graph_steps`` steps of it, replayed) as a hipGraph.  With
        ``residual=True`` the last step runs on its own (not inside a
        temporal block, the same state) and ``self.last_residual`` holds the
        FP64 norms of its change (``cutile.runtime.norms.ResidualNorms``);
        the call then waits for the stream."""
        import torch

        from cutile.runtime.kernel import default_stream

        if stream is None:
            stream = default_stream()
        b = scratch if scratch is not None else torch.empty_like(u)
        self.fill(u, stream)
        if residual and steps > 0:
            from cutile.runtime.norms import residual_norms

            head = u
            if steps > 1:
                head = self.march(u, steps - 1, b, stream, graph, graph_steps)
                other = b if head is u else u
            else:
                other = b
            new, old = self.enqueue(head, other, 1, stream)
            self.last_residual = residual_norms(new, old, self.halo, stream)
            return new
        if not graph:
            return self.enqueue(u, b, steps, stream)[0]
        # graphs are captured on a private stream ordered after the caller's
        cur = torch.cuda.current_stream()
        side = torch.cuda.Stream()
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            out = self._march_graph(u, b, steps, side.cuda_stream, graph_steps)
        cur.wait_stream(side)
        return out

    def _march_graph(self, u, b, steps, stream, graph_steps):
        import torch

        # warm the JIT (and the depth choice) outside the capture
        self._single._function(self._single.pick([u], [b], self.halo))
        if self._block is not None:
            kern = self.plan_for(u)[0]
            if kern is not None:
                kern._function(kern.temporal_key([u], [b], self.halo))
        chunk = graph_steps or steps
        per = chunk if self._block is None else (chunk // self.T) * self.T or chunk
        reps, rest = divmod(steps, per)
        a = u
        if reps:
            g = rt.Graph(stream)
            with g:
                a2, b2 = self.enqueue(a, b, per, stream)
            # after one replay the result sits in a2; replays must start from
            # the same buffer, so an odd-length chunk ends with a copy back
            for _ in range(reps):
                g.launch()
                if a2 is not a:
                    a.copy_(a2)
            self._graph = g
            # the graph reads and writes these buffers on every replay: they
            # must outlive it (a freed scratch would be handed to the next
            # allocation while replays still write into it)
            self._graph_buffers = (a, b, a2, b2)
        if rest:
            a, _ = self.enqueue(a, b if b is not a else torch.empty_like(a), rest, stream)
        return a


class FusedStepper:
    """Advance several coupled fields with a fused kernel (``compile_fused``):
    every member reads the old state; member k's output becomes field
    ``targets[k]``; each field in ``fill`` (default: all targets) gets its
    halo refilled.  A time-update result (``update="euler"|"kahan"``) writes
    its fields and compensations in ``result.outputs`` order, so
    ``targets=result.outputs`` and ``fill`` = the fields whose halo the
    stencil reads (compensations are read at the centre only)."""

    def __init__(self, result, boundary, targets: Sequence[str], fill: Optional[Sequence[str]] = None,
                 block=None):
        self.r = result
        self.spec = boundary
        self.targets = list(targets)
        self.fill_names = list(fill) if fill is not None else list(targets)
        self.halo = result.halo_widths
        self._k = result.load_module()._KERNEL
        self.policy = HaloPolicy(boundary, result.ndim, self.halo)
        # ``block``: ``compile_fused_block`` of the same members, T steps per
        # launch.  Exact for a closed system (targets = inputs, in order)
        # under periodic or zero-Dirichlet boundaries only
        self._blk, self.T = None, 1
        if block is not None:
            kind = _bc_kind(boundary, result.ndim)
            if self.targets != list(result.inputs) or list(block.inputs) != list(result.inputs):
                raise ValueError("a fused temporal block needs targets equal to the inputs, in order")
            if tuple(block.halo_widths) != tuple(self.halo):
                raise ValueError("the fused temporal block must share the one-step kernel's halo")
            if kind in ("periodic", "dirichlet0"):
                self._blk, self.T = block.load_module()._KERNEL, int(block.temporal_steps)

    def enqueue(self, fields: dict, scratch: dict, steps: int, stream=None, residual: bool = False):
        """Advance ``steps`` fused steps (fields <-> scratch swap per step).
        With ``residual=True`` the last step runs as a single fused step and
        ``self.last_residual`` maps each target to the FP64 norms of its
        change (``ResidualNorms``); the call then waits for the stream."""
        from cutile.runtime.kernel import default_stream

        if stream is None:
            stream = default_stream()
        if residual and steps > 0:
            from cutile.runtime.norms import residual_norms

            if steps > 1:
                fields, scratch = self.enqueue(fields, scratch, steps - 1, stream)
            blk, self._blk = self._blk, None
            try:
                fields, scratch = self.enqueue(fields, scratch, 1, stream)
            finally:
                self._blk = blk
            self.last_residual = {t: residual_norms(fields[t], scratch[t], self.halo, stream)
                                  for t in self.targets}
            return fields, scratch
        if self.policy.mode == "static":
            for t in self.fill_names:
                _boundary.fill(scratch[t], self.spec, self.halo, stream)
        axes = self.policy.axes_for(fields[self.targets[0]])
        nblk = steps // self.T if self._blk is not None else 0
        wrap = tuple(range(self.r.ndim)) if self.policy.mode == "wrap" else ()
        for _ in range(nblk):
            ins = [fields[n] for n in self.r.inputs]
            outs = [scratch[t] for t in self.targets]
            self._blk.launch_temporal(ins, outs, self.halo, stream, mode=0, validate=False, periodic_axes=wrap)
            for t in self.targets:
                fields[t], scratch[t] = scratch[t], fields[t]
        blocked = nblk > 0
        steps -= nblk * self.T
        if blocked and wrap and steps and not axes:
            for t in self.fill_names:     # single steps that read the halo follow wrapped blocks
                _boundary.fill(fields[t], self.spec, self.halo, stream)
        for _ in range(steps):
            ins = [fields[n] for n in self.r.inputs]
            outs = [scratch[t] for t in self.targets]
            self._k.launch(ins, outs, self.halo, stream, validate=False, periodic_axes=axes)
            if not axes and self.policy.mode != "static":
                for t in self.fill_names:
                    _boundary.fill(scratch[t], self.spec, self.halo, stream)
            for t in self.targets:
                fields[t], scratch[t] = scratch[t], fields[t]
        if (axes and steps) or (blocked and wrap):
            for t in self.fill_names:     # the wrapped sweeps left the halos stale
                _boundary.fill(fields[t], self.spec, self.halo, stream)
        return fields, scratch
