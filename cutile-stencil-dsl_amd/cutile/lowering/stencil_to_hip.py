"""Lowering entry points: stencil IR -> emitted launcher module.

Same entry points and arguments as upstream
``lower_stencil_to_python`` (``cutile/lowering/stencil_to_cutile.py:250-331``)
and ``lower_fused_stencils_to_python`` (``cutile/lowering/fusion_emitter.py:
27-56``); the returned text is a Python module that embeds the emitted HIP
source and exposes the upstream launcher names:

* ``launch_<name>(u_in, u_out, stream=None)`` (single input),
  ``launch_<name>(u, v, ..., u_out, stream=None)`` (several inputs),
  ``launch_<fused>(*inputs, *<member>_out, stream=None)`` (fused);
* ``apply_boundary_<name>(u, stream=None)`` when the stencil declares a
  boundary condition;
* with ``temporal_steps > 1`` one ``launch_<name>`` call advances T steps
  in one kernel (in-kernel temporal blocking).  ``temporal_mode="zero"``
  keeps upstream's semantics of T sweeps through zero-halo scratch arrays
  (``stencil_to_target.py:740-843``); ``"recompute"`` evaluates the
  intermediate levels inside a halo T*r deep, which equals T steps with a
  periodic refill when that halo was filled before the call.
"""

from __future__ import annotations

import json
from typing import Optional, Sequence

from cutile.frontend.ir import find_func
from cutile.lowering.emitter import CodeEmitter
from cutile.lowering.hip_emitter import (ELEM_BYTES, KernelConfig, brick_symbol, emit_hip_source, kernel_symbol,
                                        lds_bytes, plan_accesses, radius3, sweep_tag, tblock_queue_range,
                                        select_tblock, tblock_symbol, temporal_symbol, temporal_tile)
from cutile.lowering.meta import StencilMeta, apply_update, extract_fused_meta, extract_meta


def storage_compute(dtype: str, storage: Optional[str]) -> tuple:
    if storage in (None, dtype):
        return dtype, dtype
    if storage == "bfloat16":
        if dtype != "float32":
            raise ValueError("bfloat16 storage computes in float32: declare the stencil dtype='float32'")
        return "bfloat16", "float32"
    raise ValueError(f"unsupported storage {storage!r} for dtype {dtype!r}")


def make_configs(ndim: int, tile: Sequence[int], storage: str, threads: Optional[int] = None,
                 vector: Optional[int] = None) -> tuple:
    """(vector config, scalar config, chunk) for a ``tile_sizes`` tuple;
    ``threads`` overrides the workgroup size (a multiple of the threads one
    tile row needs); ``vector`` the cells per thread along the contiguous
    axis (default: one 16-byte vector; fewer cells hold fewer registers)."""
    tile = tuple(int(t) for t in tile)
    if ndim == 3:
        xch, ty, tz = tile
    elif ndim == 2:
        (xch, tz), ty = tile, 1
    else:
        (tz,), ty, xch = tile, 1, 1
    sb = ELEM_BYTES[storage]
    vz = max(1, min(int(vector) if vector else 16 // sb, 16 // sb, tz))

    def cfg(tz_, vz_):
        ntz = max(1, tz_ // vz_)
        if ntz > 1024:
            raise ValueError(f"tile row of {tz_} elements needs {ntz} threads (> 1024)")
        if threads:
            nt = max(int(threads), ntz)
            if nt % ntz or nt % 64:
                raise ValueError(f"{threads} threads do not tile rows of {ntz} threads in wave64s")
        elif ntz >= 64:
            nt = ntz * max(1, min(ty, 256 // ntz))
        else:
            nt = 64
        return KernelConfig(ty=ty, tz=tz_, vz=vz_, nt=nt)

    vec = cfg(tz, vz)
    scal = cfg(max(1, tz // vz), 1)
    return vec, scal, xch


def fit_lds(meta: StencilMeta, tile: Sequence[int], hints: Optional[dict], storage: Optional[str],
            budget: int = 160 * 1024) -> tuple:
    """``(tile, hints)`` shrunk until the sweep's LDS plane buffers fit in
    ``budget`` bytes (one CU's LDS): halve the rows, then the columns down to
    64 vectors; a ``threads`` hint is dropped once the tile rows no longer
    give every thread a row.  Stencils with many distinct planes and fields
    need one buffer per (field, plane offset)."""
    storage_t, compute = storage_compute(meta.dtype, storage)
    tile = tuple(int(t) for t in tile)
    hints = dict(hints or {})
    nd = len(tile)
    while True:
        vec, _, _ = make_configs(nd, tile, storage_t, threads=hints.get("threads"), vector=hints.get("vector"))
        if lds_bytes(meta, vec, compute) <= budget:
            return tile, hints
        if nd == 3 and tile[1] > 1:
            tile = (tile[0], tile[1] // 2, tile[2])
        elif tile[-1] > 64 * vec.vz:
            tile = tile[:-1] + (tile[-1] // 2,)
        else:
            raise ValueError(f"{meta.name}: LDS plane buffers exceed {budget} bytes even at tile {tile}")
        if "threads" in hints:
            ntz = max(1, tile[-1] // vec.vz)
            rows = tile[1] if nd == 3 else 1
            if hints["threads"] // ntz > rows or hints["threads"] % ntz:
                hints.pop("threads")


def _attrs(kernel_hints: Optional[dict]) -> str:
    occ = (kernel_hints or {}).get("occupancy")  # other hints ("nontemporal") are emitter options
    return f"__attribute__((amdgpu_waves_per_eu({int(occ)}))) " if occ else ""


def brick_dims(ndim: int, brick_size: int) -> tuple:
    """Brick extents in the 3-D (stream, row, column) map of an ``ndim``-D array."""
    b = int(brick_size)
    return {1: (1, 1, b), 2: (b, 1, b), 3: (b, b, b)}[ndim]


def emit_module(meta: StencilMeta, tile_sizes: Sequence[int], halo_widths: Sequence[int],
                temporal_steps: int = 1, boundary_spec=None, kernel_hints: Optional[dict] = None,
                storage: Optional[str] = None, fused: bool = False, temporal_mode: str = "zero",
                brick_size: Optional[int] = None) -> str:
    storage_t, compute_t = storage_compute(meta.dtype, storage)
    vec, scal, xch = make_configs(meta.ndim, tile_sizes, storage_t, (kernel_hints or {}).get("threads"),
                                  (kernel_hints or {}).get("vector"))
    temporal = None
    tblocks = []
    tblock_depth = 0
    if temporal_steps > 1:
        if meta.num_outputs != meta.num_inputs or meta.num_exprs != meta.num_inputs or meta.update:
            raise NotImplementedError("in-kernel temporal blocking needs one result per field "
                                      "(result m is the next state of field m)")
        # register/LDS kernel at the vector width and at scalar width (arrays
        # whose rows are not vector-aligned); the LDS-ring kernel when the
        # ghost ring of T levels does not fit the registers or the LDS
        # (a forced geometry applies to the vector variant only)
        hints = dict(kernel_hints or {})
        for vz in dict.fromkeys((min(vec.vz, 4), 1)):
            h = hints if vz > 1 or vec.vz == 1 else {k: v for k, v in hints.items()
                                                      if k not in ("tblock_tile", "tbws_tile")}
            tb = select_tblock(meta, int(temporal_steps), storage_t, compute_t, vz, h)
            if tb is None:
                tblocks = []
                break
            tblocks.append(tb)
        if tblocks:
            qlo, qhi = tblock_queue_range(meta)
            tblock_depth = int(temporal_steps) * (qhi - qlo + 1)
        elif meta.num_inputs != 1:
            raise NotImplementedError("coupled fields at this temporal depth fit neither the registers nor the LDS")
        else:
            ty, tz = temporal_tile(meta.ndim, temporal_steps, radius3(meta), compute_t)
            temporal = (KernelConfig(ty=ty, tz=tz, vz=1, nt=256), int(temporal_steps))
    # z-neighbours across lanes (DPP) instead of the LDS plane: an emitter
    # option, off unless measured faster for the stencil
    zdpp = bool((kernel_hints or {}).get("zdpp", 0))
    # 3-D stencils reaching 2+ cells in the plane: two sets of LDS plane
    # buffers (one barrier per plane) and halo rings fetched QHI + PD planes
    # ahead, with the neighbouring tiles' queue heads of the same plane
    # (together on MI355X: 25-point star at 512^3 FP64 401.8 -> 384.7 us per
    # sweep, FP32 204.7 -> 192.4; either alone within 1 %; the 7-point star
    # gains nothing from them)
    plan = plan_accesses(meta)
    wide = meta.ndim == 3 and max(radius3(meta)[1:]) >= 2
    pf = int((kernel_hints or {}).get("prefetch", 1))
    sdb = int((kernel_hints or {}).get("sdb", 1 if wide else 0))
    if sdb and 2 * max(lds_bytes(meta, c, compute_t) for c in (vec, scal)) > 160 * 1024:
        sdb = 0           # the second set of plane buffers does not fit the LDS
    ringpd = int((kernel_hints or {}).get("ringpd", max(1, plan["qhi"] + pf) if wide else 1))
    bricked = []
    if brick_size:
        bd = brick_dims(meta.ndim, brick_size)
        bricked = [(c, bd) for c in dict.fromkeys((vec, scal)) if bd[2] % c.vz == 0]
    src = emit_hip_source(meta, [vec, scal], storage_t, compute_t, temporal,
                          nontemporal=bool((kernel_hints or {}).get("nontemporal", 1)),
                          prefetch=pf,
                          ntload=bool((kernel_hints or {}).get("ntload", 0)), tblocks=tblocks,
                          bricked=bricked, zdpp=zdpp, tbskip=int((kernel_hints or {}).get("tbskip", 0)),
                          tbdb=(kernel_hints or {}).get("tbdb"), tbpd=int((kernel_hints or {}).get("tbpd", 2)),
                          ringpd=ringpd, sdb=sdb, tbrp=(kernel_hints or {}).get("tbrp"))
    attr = _attrs(kernel_hints)
    if attr:
        src = src.replace('extern "C" __global__ void ', f'extern "C" __global__ void {attr}')
    variants = {
        "vec": (kernel_symbol(meta, sweep_tag(vec, zdpp)), vec.ty, vec.tz, vec.vz, vec.nt),
        "scalar": (kernel_symbol(meta, sweep_tag(scal, zdpp)), scal.ty, scal.tz, scal.vz, scal.nt),
    }
    if temporal is not None:
        tcfg, T = temporal
        variants["temporal"] = (temporal_symbol(meta, tcfg.tag, T), tcfg.ty, tcfg.tz, 1, tcfg.nt)
    for tb, key in zip(tblocks, ("tblock", "tblock_s")):
        # (symbol, output tile rows, output tile columns, vector, threads)
        variants[key] = (tblock_symbol(meta, tb), tb.oty, tb.otz, tb.vz, tb.nt)
    for c, bd in bricked:
        # brick-major sweeps: "brick" at the vector width, "brick_s" scalar
        variants["brick" if c.vz > 1 or vec.vz == 1 else "brick_s"] = (brick_symbol(meta, c.tag, bd), c.ty, c.tz,
                                                                      c.vz, c.nt)
    if bricked and "brick_s" not in variants:
        variants["brick_s"] = variants["brick"]
    plan = plan_accesses(meta)
    qdepth = max(1, plan["qhi"] - plan["qlo"] + 1)
    nd = meta.ndim
    tvars = ["TX", "TY", "TZ"][:nd]
    hvars = ["HX", "HY", "HZ"][:nd]
    tr = "," if nd == 1 else ""
    multi = meta.num_inputs > 1 or fused
    inputs = list(meta.input_names) if multi else ["u_in"]
    outs = meta.output_params() if (fused or meta.num_outputs > 1 or meta.update) else ["u_out"]

    e = CodeEmitter()
    e.line(f'"""HIP kernel module for the {meta.name} stencil (generated for gfx950)."""')
    e.blank()
    e.line("from cutile.runtime.kernel import StencilKernel")
    e.line("from cutile.runtime import boundary as _boundary")
    if meta.constants:
        e.blank()
        e.line("# Captured constants from stencil definition scope")
        for k, v in sorted(meta.constants.items()):
            e.line(f"{k} = {v!r}")
    e.blank()
    e.line(f"{', '.join(tvars)}{tr} = {', '.join(str(int(t)) for t in tile_sizes)}{tr}")
    e.line(f"{', '.join(hvars)}{tr} = {', '.join(str(int(h)) for h in halo_widths)}{tr}")
    e.line(f"TEMPORAL_STEPS = {int(temporal_steps)}")
    if temporal_mode not in ("zero", "recompute"):
        raise ValueError(f"temporal_mode must be 'zero' or 'recompute', got {temporal_mode!r}")
    e.line(f"TEMPORAL_MODE = {0 if temporal_mode == 'zero' else 1}  # {temporal_mode}")
    e.line(f"STORAGE = {storage_t!r}")
    e.line(f"INPUTS = {tuple(meta.input_names)!r}")
    e.line(f"OUTPUTS = {tuple(meta.output_names)!r}")
    e.line(f"EXPRESSIONS = {meta.py_expressions!r}")
    e.blank()
    e.line("_HIP_SOURCE = r'''")
    for ln in src.splitlines():
        e.raw(ln)
    e.raw("'''")
    e.blank()
    e.line(f"_VARIANTS = {json.dumps(variants)}")
    e.line(f"_KERNEL = StencilKernel({meta.name!r}, {nd}, {meta.num_inputs}, {meta.num_outputs}, "
           f"_HIP_SOURCE, {{k: tuple(v) for k, v in _VARIANTS.items()}}, STORAGE, {int(xch)}, "
           f"queue_depth={qdepth}, temporal_steps={int(temporal_steps) if temporal_steps > 1 else 1}, "
           f"tblock_depth={tblock_depth}, brick={tuple(brick_dims(nd, brick_size)) if brick_size else None})")
    e.line(f"_HALO = ({', '.join(hvars)}{tr})")

    if boundary_spec is not None and meta.num_outputs == 1 and not fused:
        e.blank()
        e.blank()
        e.line(f"_BOUNDARY = {_spec_repr(boundary_spec)}")
        e.blank()
        e.blank()
        e.line(f"def apply_boundary_{meta.name}(u, stream=None):")
        with e.indent():
            e.line('"""Apply boundary conditions to halo cells (on device)."""')
            e.line("_boundary.fill(u, _BOUNDARY, _HALO, stream)")

    e.blank()
    e.blank()
    e.line(f"def launch_{meta.name}({', '.join(inputs + outs)}, stream=None):")
    with e.indent():
        ins = f"[{', '.join(inputs)}]"
        ots = f"[{', '.join(outs)}]"
        if temporal_steps > 1:
            e.line(f"# Temporal blocking: {temporal_steps} steps in one launch (levels in LDS rings)")
            e.line(f"_KERNEL.launch_temporal({ins}, {ots}, _HALO, stream, mode=TEMPORAL_MODE)")
        else:
            e.line(f"_KERNEL.launch({ins}, {ots}, _HALO, stream)")
    return e.render()


def _spec_repr(spec) -> str:
    from cutile.frontend.ir import BoundaryAttr

    if isinstance(spec, BoundaryAttr):
        spec = {"type": spec.bc_type, "value": spec.value}
    if hasattr(spec, "to_dict"):
        return repr(spec.to_dict(3))
    return repr(spec)


def lower_stencil_to_python(module, domain=None, tile_sizes=None, halo_widths=None,
                            temporal_steps: int = 1, boundary_spec=None, kernel_hints=None,
                            spatial_term_order=None, storage=None, temporal_mode: str = "zero",
                            update=None, dt: float = 1.0, targets=None, brick_size=None) -> str:
    func = find_func(module)
    if tile_sizes is None:
        from cutile.passes.tiling import TilingPass, unique_access_count

        tp = TilingPass(dtype_bytes=4 if func.dtype == "float32" else 8)
        tile_sizes = tuple(tp.candidates(func.ndim, unique_access_count(func))[0][0])
    if halo_widths is None:
        h = func.order // 2 if func.order else 1
        halo_widths = (h,) * func.ndim
    meta = extract_meta(func)
    if update is not None:
        if temporal_steps != 1:
            raise NotImplementedError("a time-update epilogue runs one step per launch (temporal_steps=1)")
        apply_update(meta, update, dt, targets or [meta.input_names[0]])
    if boundary_spec is None and meta.boundary is not None:
        boundary_spec = meta.boundary
    return emit_module(meta, tile_sizes, halo_widths, temporal_steps, boundary_spec, kernel_hints, storage,
                       temporal_mode=temporal_mode, brick_size=brick_size)


def lower_fused_stencils_to_python(modules: Sequence, tile_sizes=None, halo_widths=None,
                                   temporal_steps: int = 1, kernel_hints=None,
                                   spatial_term_order=None, storage=None, update=None, dt: float = 1.0,
                                   targets=None, temporal_mode: str = "zero") -> str:
    """Module of a fused kernel.  ``temporal_steps > 1`` (reached through
    ``FusedStepper``; ``compile_fused`` keeps the upstream restriction to
    one step) needs a closed system: member m's result is the next state of
    input m."""
    if temporal_steps != 1 and update is not None:
        raise NotImplementedError("a time-update epilogue runs one step per launch (temporal_steps=1)")
    meta = extract_fused_meta(modules)
    if update is not None:
        apply_update(meta, update, dt, targets)
    if halo_widths is None:
        halo_widths = (max(1, meta.order // 2),) * meta.ndim
    if tile_sizes is None:
        from cutile.passes.tiling import TilingPass

        tp = TilingPass(dtype_bytes=4 if meta.dtype == "float32" else 8)
        tile_sizes = tuple(tp.tile_for(meta.ndim, len(meta.accesses)))
    return emit_module(meta, tile_sizes, halo_widths, temporal_steps, None, kernel_hints, storage, fused=True,
                       temporal_mode=temporal_mode)
