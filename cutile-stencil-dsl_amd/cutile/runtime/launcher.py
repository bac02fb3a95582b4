"""Compilation and execution API: ``compile``, ``compile_fused``, ``CompileResult``.

Same signatures, fields and policy order as upstream
``cutile/runtime/launcher.py:27-728``: run the pass pipeline on a clone,
read halo/tile/temporal/roofline attributes, pad the innermost halo to 128
bytes (``pad_inner_halo``), optionally autotune, rank the tiling pass's
candidates and keep the first that fits its register budget (wide) or the
fastest measured (very wide), then lower to the emitted launcher module.
Arrays are torch tensors on the ROCm device instead of CuPy arrays.
"""

from __future__ import annotations

import importlib.util
import math
import os
import tempfile
from dataclasses import dataclass, field
from typing import Any, Optional, Sequence

from cutile.runtime.pipeline import Pipeline

_DTYPE_BYTES = {"float64": 8, "float32": 4, "bfloat16": 2}


@dataclass
class CompileResult:
    name: str
    code: str
    ndim: int
    halo_widths: tuple
    tile_sizes: tuple
    temporal_steps: int
    analysis: dict = field(default_factory=dict)
    stencil_halo: tuple = field(default=())
    kernel_hints: dict = field(default_factory=dict)
    resources: Any = field(default=None)
    inputs: tuple = ()
    outputs: tuple = ()
    dtype: str = "float64"
    storage: str = "float64"
    hip_source: str = ""
    temporal_mode: str = "zero"
    _ref_fn: Any = field(default=None, repr=False)
    _ref_fns: Any = field(default=None, repr=False)
    _ref_inputs: Any = field(default=None, repr=False)
    _module: Any = field(default=None, repr=False)

    def emit_to_file(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.code)

    def load_module(self):
        """Import the emitted launcher module (cached per result)."""
        if self._module is None:
            with tempfile.NamedTemporaryFile(suffix=".py", delete=False, mode="w") as f:
                f.write(self.code)
                path = f.name
            spec = importlib.util.spec_from_file_location(f"_gen_{self.name}", path)
            mod = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(mod)
            try:
                os.unlink(path)
            except OSError:
                pass
            self._module = mod
        return self._module

    @property
    def launcher(self):
        return getattr(self.load_module(), f"launch_{self.name}")

    # ------------------------------------------------------------------
    def _torch_dtype(self):
        import torch

        return getattr(torch, self.storage)

    def validate(self, u_input=None, atol: float = 1e-10, reference=None, seed: int = 0) -> bool:
        """Run the kernel once on random input and compare the interior at
        distance ``T*h`` from the edge with a CPU reference
        (``launcher.py:76-169``).  ``reference`` is the checker module that
        evaluates stencils on NumPy arrays (it must provide ``apply_stencil``
        and ``eval_multi``); this package ships no CPU evaluator of its own."""
        import numpy as np
        import torch

        if reference is None:
            raise RuntimeError("validate() needs the CPU checker: pass reference=<oracle module>")
        hw, T = self.halo_widths, self.temporal_steps
        rng = np.random.default_rng(seed)
        npdt = np.float32 if self.dtype == "float32" else np.float64
        base = 128 if self.ndim == 1 else 64
        fns = self._ref_fns or [self._ref_fn]
        if len(fns) > 1 or len(self.inputs) > 1:
            shape = tuple(base + 2 * h for h in hw)
            fields = {n: (rng.random(shape) + 0.5).astype(npdt) for n in self.inputs}
            dev = [torch.from_numpy(fields[n]).to("cuda", self._torch_dtype()) for n in self.inputs]
            outs = [torch.zeros_like(dev[0]) for _ in fns]
            self.launcher(*dev, *outs)
            torch.cuda.synchronize()
            inner = tuple(slice(h, s - h) for h, s in zip(hw, shape))
            ok = True
            for fn, out in zip(fns, outs):
                ref = reference.eval_multi(fields, fn, self.ndim, hw)
                got = out.to(torch.float64).cpu().numpy()[inner]
                diff = float(np.max(np.abs(ref - got))) if ref.size else 0.0
                good = bool(np.allclose(ref, got, atol=atol))
                print(f"Validation[{fn.__name__}]: max_diff={diff:.2e}, pass={good}")
                ok = ok and good
            return ok
        if u_input is None:
            shape = tuple(base + 2 * T * h for h in hw)
            u = rng.random(shape).astype(npdt)
        else:
            u = np.asarray(u_input)
        ref = u.copy()
        for _ in range(T):
            ref = reference.apply_stencil(ref, self._ref_fn, self.ndim, hw)
        du = torch.from_numpy(u).to("cuda", self._torch_dtype())
        out = torch.zeros_like(du)
        self.launcher(du, out)
        torch.cuda.synchronize()
        got = out.to(torch.float64).cpu().numpy()
        inner = tuple(slice(T * h, s - T * h) for h, s in zip(hw, u.shape))
        diff = float(np.max(np.abs(ref[inner] - got[inner])))
        ok = bool(np.allclose(ref[inner], got[inner], atol=atol))
        print(f"Validation: max_diff={diff:.2e}, pass={ok}")
        return ok

    def benchmark(self, u_input=None, warmup: int = 5, iters: int = 20, domain=None) -> dict:
        """Time the launcher with HIP events (``launcher.py:171-231``);
        ``gbytes_per_s`` counts one read per input and one write per output
        in the storage type."""
        import torch

        from cutile.runtime import rt
        from cutile.runtime.kernel import default_stream

        hw = self.halo_widths
        if u_input is not None:
            shape = tuple(u_input.shape)
        else:
            dom = domain or ((128,) if self.ndim == 1 else (64,) * self.ndim)
            shape = tuple(d + 2 * h for d, h in zip(dom, hw))
        dt = self._torch_dtype()
        ins = [torch.rand(shape, device="cuda", dtype=torch.float64).to(dt) for _ in (self.inputs or ("u",))]
        outs = [torch.zeros_like(ins[0]) for _ in (self.outputs or ("out",))]
        fn = self.launcher
        for _ in range(warmup):
            fn(*ins, *outs)
        torch.cuda.synchronize()
        s = default_stream()
        e0, e1 = rt.Event(), rt.Event()
        e0.record(s)
        for _ in range(iters):
            fn(*ins, *outs)
        e1.record(s)
        ms = e0.elapsed_ms(e1) / iters
        cells = math.prod(n - 2 * h for n, h in zip(shape, hw))
        nbytes = cells * (len(ins) + len(outs)) * _DTYPE_BYTES[self.storage] / max(1, self.temporal_steps)
        return {
            "time_ms": ms,
            "gpoints_per_s": cells * self.temporal_steps / ms * 1e-6,
            "gbytes_per_s": nbytes * self.temporal_steps / ms * 1e-6,
        }


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------

def pad_inner_halo(halo_widths: tuple, dtype_bytes: int, align_bytes: int = 128) -> tuple:
    """Round the innermost halo up to a multiple of ``align_bytes`` so every
    interior row starts on a 128-byte line (``launcher.py:257-273``)."""
    if not halo_widths:
        return halo_widths
    align = max(1, align_bytes // dtype_bytes)
    inner = -(-halo_widths[-1] // align) * align
    return tuple(halo_widths[:-1]) + (inner,)


def _budgeted_configs(tiling, candidates):
    from cutile.runtime import regcheck

    return [regcheck.Config(tuple(t), dict(h), max_registers=regcheck.register_budget(h.get("occupancy")))
            for t, h in candidates]


_SELECT_MODES = ("probe", "measure")


def _selection_mode(select, tiling, num_loads) -> str:
    if select is not None:
        if select not in _SELECT_MODES:
            raise ValueError(f"select must be one of {_SELECT_MODES}, got {select!r}")
        return select
    return "measure" if tiling.tier(num_loads) == "very_wide" else "probe"


def _choose_config(candidates, lower, *, kernel_name, ndim, num_inputs, halo_widths, dtype,
                   storage, max_registers=None, num_outputs=1, select="probe"):
    from cutile.runtime import regcheck

    probed: dict = {}
    time_shape = regcheck.timing_shape(ndim) if select == "measure" else None

    def _probe(cfg):
        code = lower(cfg)
        res = regcheck.probe_resources(code=code, kernel_name=kernel_name, ndim=ndim,
                                       num_inputs=num_inputs, num_outputs=num_outputs,
                                       tile_sizes=cfg.tile_sizes, halo_widths=halo_widths,
                                       dtype=storage, time_shape=time_shape)
        probed[id(cfg)] = res
        return res

    if select == "measure":
        chosen = regcheck.select_fastest(candidates, _probe, max_registers=max_registers)
    else:
        chosen = regcheck.select_config(candidates, _probe, max_registers=max_registers)
    return chosen.tile_sizes, dict(chosen.kernel_hints), probed.get(id(chosen))


def _input_names(ir) -> tuple:
    from cutile.frontend.ir import find_func

    return tuple(find_func(ir).arg_names)


# ---------------------------------------------------------------------------
# compile()
# ---------------------------------------------------------------------------

def compile(stencil_fn, domain=None, hw=None, pipeline: Optional[Pipeline] = None,
            temporal_blocking: bool = True, autotune: bool = True, num_gpus: int = 1,
            split_axis: int = 0, overlap: bool = True, layout: Optional[str] = None,
            brick_size: int = 32, align_halo: bool = True, occupancy: Optional[int] = None,
            select: Optional[str] = None, temporal_steps: Optional[int] = None,
            storage: Optional[str] = None, tile_sizes: Optional[tuple] = None,
            temporal_mode: str = "zero", kernel_hints: Optional[dict] = None,
            update: Optional[str] = None, dt: float = 1.0, targets: Optional[Sequence[str]] = None) -> CompileResult:
    """Compile a ``@stencil`` function (``launcher.py:342-602``).

    Extra keyword arguments of this backend: ``temporal_steps`` forces the
    temporal depth, ``storage="bfloat16"`` stores a float32 stencil's arrays
    in bfloat16 (arithmetic stays float32), ``tile_sizes`` forces the tile,
    ``kernel_hints`` adds emitter options (``occupancy``, ``nontemporal``),
    ``temporal_mode="recompute"`` (with ``temporal_steps > 1``) widens the
    halo to ``T * stencil halo`` and recomputes intermediate levels there
    (T periodic steps per launch after a halo fill).
    ``update="euler"|"kahan"`` (with ``dt`` and ``targets``) reads the
    stencil as the rate of field ``targets[0]`` and fuses the time update
    into the sweep (see ``cutile.lowering.meta.apply_update``).
    ``num_gpus > 1`` is rejected: this backend drives one GPU."""
    from cutile.frontend.ir import ApplyOp
    from cutile.lowering.stencil_to_hip import lower_stencil_to_python
    from cutile.passes.tiling import TilingPass, unique_access_count

    if num_gpus != 1:
        raise NotImplementedError("this backend runs on one GPU (num_gpus=1)")
    ir = stencil_fn._ir
    name = stencil_fn._fn.__name__
    dtype = getattr(stencil_fn, "_dtype", "float64") or "float64"
    store = storage or dtype
    dtype_bytes = _DTYPE_BYTES[store]
    if hw is not None:
        dtype_bytes = hw.dtype_bytes

    clone = ir.clone()
    if pipeline is None:
        lds = 160 * 1024 if hw is None else hw.shared_mem_bytes
        pipeline = Pipeline.single_gpu(shared_mem_bytes=lds, dtype_bytes=dtype_bytes,
                                       max_temporal_steps=8 if temporal_blocking else 1)
    pipeline.run(clone)

    halo = (1,)
    tiles = None
    T = 1
    analysis: dict = {}
    hints: dict = {}
    num_loads = 0
    for op in clone.ops:
        if isinstance(op, ApplyOp):
            a = op.attributes
            halo = tuple(a.get("halo_widths", halo))
            tiles = tuple(a["tile_sizes"]) if "tile_sizes" in a else None
            T = a.get("temporal_steps", 1)
            if "occupancy" in a:
                hints["occupancy"] = a["occupancy"]
            if "threads" in a and tile_sizes is None:
                hints["threads"] = a["threads"]
            num_loads = unique_access_count(op)
            for k in ("flops", "unique_loads", "arithmetic_intensity", "bound"):
                if k in a:
                    analysis[k] = a[k]
            break
    if not temporal_blocking:
        T = 1
    if temporal_steps is not None:
        T = int(temporal_steps)
    if T > 1 and len(_input_names(ir)) != 1:
        T = 1  # temporal chains need a single input (upstream breaks there too)
    if update is not None:
        T = 1
        targets = list(targets) if targets else [_input_names(ir)[0]]

    stencil_halo = halo
    if T > 1 and temporal_mode == "recompute":
        halo = tuple(T * h for h in halo)
    if align_halo and layout != "bricked":
        halo = pad_inner_halo(halo, dtype_bytes)

    from cutile.frontend.ir import find_func
    from cutile.lowering.meta import extract_meta
    from cutile.lowering.stencil_to_hip import fit_lds

    meta = extract_meta(find_func(ir))

    def _fit(t, h):
        return fit_lds(meta, t, h, storage)

    tiling = next((p for p in pipeline.passes if isinstance(p, TilingPass)), TilingPass(dtype_bytes=dtype_bytes))
    candidates = [_fit(t, h) for t, h in tiling.candidates(len(halo), num_loads)]
    if tiles is None:
        tiles = tuple(candidates[0][0])

    resources = None
    if tile_sizes is not None:
        tiles = tuple(tile_sizes)
    elif autotune and domain is not None:
        from cutile.runtime.autotune import autotune as _autotune

        res = _autotune(stencil_fn, domain, hw=hw, storage=store, halo=halo)
        tiles = tuple(res.tile_sizes)
        hints = dict(res.kernel_hints or {})
    if occupancy is not None:
        hints["occupancy"] = occupancy
    if kernel_hints:
        hints.update(kernel_hints)

    boundary_spec = getattr(stencil_fn, "_boundary", None)
    tiles, hints = _fit(tiles, hints)

    probe_wanted = (len(candidates) > 1 and occupancy is None and tile_sizes is None
                    and not (autotune and domain is not None) and layout != "bricked")
    if probe_wanted:
        def _lower(cfg):
            return lower_stencil_to_python(ir.clone(), tile_sizes=cfg.tile_sizes, halo_widths=halo,
                                           kernel_hints=cfg.kernel_hints, storage=storage, update=update,
                                           dt=dt, targets=targets)

        if kernel_hints:   # emitter options the caller asked for apply to every candidate
            candidates = [(t, dict(h, **kernel_hints)) for t, h in candidates]
        tiles, hints, resources = _choose_config(
            _budgeted_configs(tiling, candidates), _lower, kernel_name=f"{name}_kernel",
            ndim=len(halo), num_inputs=len(_input_names(ir)), halo_widths=halo, dtype=dtype,
            storage=store, select=_selection_mode(select, tiling, num_loads))

    code = lower_stencil_to_python(ir.clone(), domain=domain, tile_sizes=tiles, halo_widths=halo,
                                   temporal_steps=T, boundary_spec=boundary_spec,
                                   kernel_hints=hints, storage=storage, temporal_mode=temporal_mode,
                                   update=update, dt=dt, targets=targets,
                                   brick_size=brick_size if layout == "bricked" else None)
    if layout == "bricked":
        from cutile.lowering.bricked import bricked_suffix

        code += bricked_suffix(name, brick_size)
    src = code.split("_HIP_SOURCE = r'''\n", 1)[1].split("'''", 1)[0] if "_HIP_SOURCE" in code else ""
    inputs, outputs = _input_names(ir), ()
    if update is not None:
        from cutile.lowering.meta import apply_update

        um = apply_update(extract_meta(find_func(ir)), update, dt, targets)
        inputs, outputs = tuple(um.input_names), tuple(um.output_names)
    return CompileResult(
        name=name, code=code, ndim=stencil_fn._ndim or len(halo), halo_widths=tuple(halo),
        tile_sizes=tuple(tiles), temporal_steps=T, analysis=analysis, stencil_halo=tuple(stencil_halo),
        kernel_hints=hints, resources=resources, inputs=inputs, outputs=outputs, dtype=dtype, storage=store,
        hip_source=src, temporal_mode=temporal_mode, _ref_fn=stencil_fn._fn, _ref_fns=[stencil_fn._fn],
    )


# ---------------------------------------------------------------------------
# compile_fused()
# ---------------------------------------------------------------------------

def compile_fused(stencil_fns: list, domain=None, hw=None, tile_sizes=None, halo_widths=None,
                  temporal_steps: int = 1, occupancy: Optional[int] = None, align_halo: bool = True,
                  select: Optional[str] = None, storage: Optional[str] = None, update: Optional[str] = None,
                  dt: float = 1.0, targets: Optional[Sequence[str]] = None,
                  kernel_hints: Optional[dict] = None) -> CompileResult:
    """Several stencils over one domain in one kernel (``launcher.py:610-728``):
    inputs merged by parameter name, one load per distinct access, one
    output per member, launcher ``launch_<name>(*inputs, *outputs)``.
    With ``update`` the members are rates of fields ``targets`` and the
    kernel stores the advanced fields (and, for ``"kahan"``, their
    compensations ``c_<target>``, which are inputs and outputs as well).
    ``kernel_hints`` (as for ``compile``) are merged into every candidate."""
    from cutile.frontend.ir import ApplyOp
    from cutile.lowering.meta import extract_fused_meta
    from cutile.lowering.stencil_to_hip import fit_lds, lower_fused_stencils_to_python
    from cutile.passes.tiling import TilingPass
    from cutile.runtime import regcheck

    if not stencil_fns:
        raise ValueError("At least one stencil function is required")
    if temporal_steps != 1:
        raise NotImplementedError(
            "temporal blocking of fused kernels is not supported (temporal_steps must be 1)")
    modules = [fn._ir for fn in stencil_fns]
    meta = extract_fused_meta(modules)
    if update is not None:
        from cutile.lowering.meta import apply_update

        apply_update(meta, update, dt, targets)
    ndim = meta.ndim
    dtype = getattr(stencil_fns[0], "_dtype", "float64") or "float64"
    store = storage or dtype
    dtype_bytes = _DTYPE_BYTES[store]
    lds = 160 * 1024
    if hw is not None:
        lds, dtype_bytes = hw.shared_mem_bytes, hw.dtype_bytes

    if halo_widths is None:
        per = []
        for m in modules:
            c = m.clone()
            Pipeline.single_gpu(shared_mem_bytes=lds, dtype_bytes=dtype_bytes, max_temporal_steps=1).run(c)
            for op in c.ops:
                if isinstance(op, ApplyOp) and "halo_widths" in op.attributes:
                    per.append(tuple(op.attributes["halo_widths"]))
                    break
        halo_widths = tuple(max(hs) for hs in zip(*per)) if per else (1,) * ndim
    stencil_halo = tuple(halo_widths)
    if align_halo:
        halo_widths = pad_inner_halo(tuple(halo_widths), dtype_bytes)

    num_loads = len(meta.accesses)
    tiling = TilingPass(dtype_bytes=dtype_bytes)
    candidates = [(tuple(tile_sizes), {})] if tile_sizes is not None else tiling.candidates(ndim, num_loads)
    if occupancy is not None:
        candidates = [(candidates[0][0], dict(candidates[0][1], occupancy=occupancy))]
    if kernel_hints:
        candidates = [(t, dict(h, **kernel_hints)) for t, h in candidates]
    candidates = [fit_lds(meta, t, h, storage, lds) for t, h in candidates]

    def _lower(cfg):
        return lower_fused_stencils_to_python(modules, tile_sizes=cfg.tile_sizes, halo_widths=halo_widths,
                                              kernel_hints=cfg.kernel_hints, storage=storage, update=update,
                                              dt=dt, targets=targets)

    resources = None
    if len(candidates) > 1:
        tiles, hints, resources = _choose_config(
            _budgeted_configs(tiling, candidates), _lower, kernel_name=f"{meta.name}_kernel", ndim=ndim,
            num_inputs=len(meta.input_names), num_outputs=len(meta.output_names),
            halo_widths=halo_widths, dtype=dtype, storage=store,
            select=_selection_mode(select, tiling, num_loads))
    else:
        tiles, hints = tuple(candidates[0][0]), dict(candidates[0][1])
    code = _lower(regcheck.Config(tuple(tiles), dict(hints)))
    src = code.split("_HIP_SOURCE = r'''\n", 1)[1].split("'''", 1)[0]
    return CompileResult(
        name=meta.name, code=code, ndim=ndim, halo_widths=tuple(halo_widths), tile_sizes=tuple(tiles),
        temporal_steps=1, analysis={"unique_loads": num_loads}, stencil_halo=stencil_halo,
        kernel_hints=hints, resources=resources, inputs=tuple(meta.input_names),
        outputs=tuple(meta.output_names), dtype=dtype, storage=store, hip_source=src,
        _ref_fns=[fn._fn for fn in stencil_fns], _ref_inputs=meta.stencil_inputs,
    )


def compile_fused_block(stencil_fns: list, temporal_steps: int, like: CompileResult,
                        targets: Sequence[str]) -> CompileResult:
    """Temporal-blocking kernel of a fused closed system: member m's result
    is the next state of field ``targets[m]`` (fused Gray-Scott: ``gs_u``,
    ``gs_v`` advancing ``u``, ``v``), ``temporal_steps`` steps per launch
    with the halo, tile and storage of the one-step result ``like``.  The
    kernel feeds result m back as input field m at every intermediate level,
    so ``targets[m]`` must be the m-th merged input name; any other order
    raises.  Used by ``FusedStepper``; ``compile_fused`` keeps the upstream
    restriction to one step (``launcher.py:652-655``), this goes beyond it.
    The launch semantics are the single-field kernel's (``launch_temporal``:
    zero-halo levels, or periodic wrap), so a block equals
    ``temporal_steps`` fused steps bit for bit under zero-Dirichlet or
    periodic boundaries."""
    from cutile.lowering.meta import extract_fused_meta
    from cutile.lowering.stencil_to_hip import lower_fused_stencils_to_python

    modules = [fn._ir for fn in stencil_fns]
    meta = extract_fused_meta(modules)
    if list(meta.input_names) != list(like.inputs) or len(meta.output_names) != len(meta.input_names):
        raise ValueError("a fused temporal block needs one member per input, member m advancing input m")
    targets = list(targets)
    if len(targets) != len(stencil_fns):
        raise ValueError(f"a fused temporal block needs one target field per member, got {len(targets)} "
                         f"for {len(stencil_fns)} members")
    for m, (t, f) in enumerate(zip(targets, meta.input_names)):
        if t != f:
            raise ValueError(f"member {m} ({meta.output_names[m]}) advances field {t!r}, but the block feeds "
                             f"result {m} back as input field {f!r}: order the members as the inputs "
                             f"{tuple(meta.input_names)}")
    T = int(temporal_steps)
    if T < 2:
        raise ValueError("temporal_steps must be at least 2")
    store = like.storage if like.storage != like.dtype else None
    code = lower_fused_stencils_to_python(modules, tile_sizes=like.tile_sizes, halo_widths=like.halo_widths,
                                          temporal_steps=T, storage=store, temporal_mode="zero")
    src = code.split("_HIP_SOURCE = r'''\n", 1)[1].split("'''", 1)[0]
    return CompileResult(
        name=meta.name, code=code, ndim=meta.ndim, halo_widths=tuple(like.halo_widths),
        tile_sizes=tuple(like.tile_sizes), temporal_steps=T, analysis=dict(like.analysis),
        stencil_halo=like.stencil_halo, kernel_hints={}, resources=None, inputs=tuple(meta.input_names),
        outputs=tuple(meta.output_names), dtype=like.dtype, storage=like.storage, hip_source=src,
        temporal_mode="zero", _ref_fns=[fn._fn for fn in stencil_fns], _ref_inputs=meta.stencil_inputs,
    )
