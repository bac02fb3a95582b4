"""The temporal-blocking launches the device tests and the benchmark make,
replayed on the host model (CPU only).

Each case compiles the stencil as the device path does, builds the launch
arguments through ``StencilKernel.args`` (256 CUs, the MI355X count), and runs
the model with the variant and chunk length the launch would use: values
against the oracle for the test sizes, bounds of the edge workgroups for the
benchmark sizes."""

import numpy as np
import pytest
import torch

import oracle
import tblock_model as tm
from cutile import compile
from cutile import workloads as w
from cutile.frontend.ir import find_func
from cutile.lowering.hip_emitter import (plan_accesses, select_tblock, tblock_queue_range, tblock_symbol)
from cutile.lowering.meta import extract_meta
from cutile.runtime import kernel as K


@pytest.fixture(autouse=True)
def _mi355x(monkeypatch):
    from cutile.runtime import regcheck, rt

    monkeypatch.setattr(K, "cu_count", lambda: 256)
    monkeypatch.setattr(K, "current_arch", lambda: "gfx950")

    def resources(self, key="vec"):
        # what the device would report, from the code object's notes
        res = regcheck.code_object_resources(K.code_object(self.source, "gfx950"))
        v = res[self.variants[key][0]]
        return rt.Resources(num_regs=v.registers, local_bytes=v.stack_bytes, shared_bytes=v.lds_bytes)

    monkeypatch.setattr(K.StencilKernel, "resources", resources)


def _launch_geometry(r, shape, storage):
    """(variant config, chunk length) of a temporal launch on arrays of ``shape``."""
    mod = r.load_module()
    kern = mod._KERNEL
    dt = getattr(torch, K.STORAGE_TORCH[kern.storage])
    a, b = torch.zeros(shape, dtype=dt), torch.zeros(shape, dtype=dt)
    key = kern.temporal_key([a], [b], r.halo_widths)
    assert key.startswith("tblock"), kern.variants.keys()
    args, _ = kern.args([a], [b], r.halo_widths, key)
    return kern, key, args


def _cfg_for(fn, r, key, storage, hints=None):
    meta = extract_meta(find_func(fn._ir))
    vz = r.load_module()._KERNEL.variants[key][3]
    if hints and vz == 1 and r.load_module()._KERNEL.variants["vec"][3] > 1:
        hints = {k: v for k, v in hints.items() if k not in ("tblock_tile", "tbws_tile")}
    cfg = select_tblock(meta, r.temporal_steps, storage, storage if storage != "bfloat16" else "float32", vz, hints)
    assert tblock_symbol(meta, cfg) == r.load_module()._KERNEL.variants[key][0]
    return meta, cfg


DEVICE_CASES = [
    # the shapes of tests/test_gpu_temporal.py and tests/test_gpu_parity.py
    ("heat2d", 4, (70, 150), 0),
    ("diffusion7", 2, (23, 19, 45), 0),
    ("diffusion7", 3, (23, 19, 45), 0),
    ("star25", 2, (23, 19, 45), 0),
    ("star25", 2, (24, 20, 36), 1),
    ("diffusion7", 4, (30, 18, 50), 1),
    ("heat2d", 4, (100, 100), 0),
    ("heat2d", 4, (96, 96), 0),                    # the prepared-record case of test_gpu_parity.py
    ("diffusion7", 2, (16, 20, 24), 1),
    ("heat2d", 3, (33, 47), 0),
    ("heat2d", 3, "validate", 0),       # CompileResult.validate: 64 + 2*T*h per axis, halo included
    # TimeStepper on a periodic problem: blocks read around every axis
    ("star25", 2, (24, 20, 36), "wrap"),
    ("star25", 2, (20, 24, 40), "wrap"),          # a depth-4 block split into two depth-2 launches
    ("diffusion7", 4, (30, 18, 50), "wrap"),
    ("diffusion7", 2, (16, 20, 24), "wrap"),
    # the two-level kernel: hints of tests/test_gpu_temporal.py
    ("star25", 2, (23, 19, 45), 0, {"tbws": 0}),
    ("star25", 2, (23, 19, 45), 0, {"tbws_tile": [24, 32, 640, 384]}),
    ("star25", 2, (23, 19, 45), 0, {"tbws_tile": [24, 32, 320, 192]}),
    ("diffusion7", 2, (23, 19, 45), 0, {"tbws": 1}),
    ("star25-float32", 2, (23, 19, 45), 0, None),
    ("star25-float32", 2, (23, 19, 45), 0, {"tbws_tile": [24, 64, 576, 384]}),
    ("star25", 2, (37, 45, 58), "wrap", {"tbws_tile": [24, 32, 320, 192]}),
    ("star25", 2, (37, 45, 58), "wrap", {"tbws_tile": [16, 32, 512, 256]}),
    ("star25", 2, (23, 19, 45), 0, {"tbws_tile": [24, 32, 320, 192], "tbya": 1}),
    ("star25-float32", 2, (23, 19, 45), 0, {"tbws_tile": [24, 64, 320, 192], "tbya": 1}),
    ("star25", 2, (37, 45, 58), "wrap", {"tbws_tile": [24, 32, 320, 192], "tbya": 1}),
    ("diffusion7", 2, (21, 45, 58), "wrap", {"tbws": 1}),
]
DEVICE_CASES = [c if len(c) == 5 else c + (None,) for c in DEVICE_CASES]


@pytest.mark.parametrize("which,T,n,tmode,hints", DEVICE_CASES,
                         ids=[f"{c[0]}-T{c[1]}-{c[2] if isinstance(c[2], str) else 'x'.join(map(str, c[2]))}-m{c[3]}"
                              f"{'-' + str(sorted(c[4].items())) if c[4] else ''}" for c in DEVICE_CASES])
def test_device_case_matches_oracle(which, T, n, tmode, hints):
    name, _, dt = which.partition("-")
    fn = getattr(w, name)(dt or "float64")
    r = compile(fn, temporal_steps=T, temporal_mode="zero" if tmode == 0 else "recompute", kernel_hints=hints)
    wrap = 0
    if tmode == "wrap":
        tmode, wrap = 0, {1: 4, 2: 5, 3: 7}[r.ndim]
    halo = r.halo_widths
    if n == "validate":
        n = tuple(64 + 2 * T * h - 2 * h for h in halo)
    shape = tuple(a + 2 * h for a, h in zip(n, halo))
    kern, key, args = _launch_geometry(r, shape, r.storage)
    meta, cfg = _cfg_for(fn, r, key, r.storage, hints)
    u = np.random.default_rng(T).random(shape)     # the model computes in float64 (any storage's geometry)
    ndim = len(n)
    if tmode == 0 and not wrap:
        from test_tblock_model import _ref_zero
        ref = _ref_zero(u, fn._fn, ndim, halo, T)
    else:
        spec = w.periodic_spec(ndim)
        oracle.apply_boundary(u, spec, halo)
        ref = oracle.time_march(u, fn._fn, ndim, halo, T, spec)
    inner = oracle.interior(shape, halo)
    if wrap:
        stale = np.random.default_rng(T + 1).random(shape)   # the halo is stale: it must not be read
        stale[inner] = u[inner]
        u = stale
    got, errs = tm.run(fn._fn, ndim, cfg, plan_accesses(meta), tblock_queue_range(meta), u, halo, args.xch,
                       tmode=tmode, wrap=wrap)
    assert not errs, errs[:5]
    np.testing.assert_array_equal(got[inner], ref[inner])


BENCH_CASES = [
    ("star25", "float64", 2, (512, 512, 512), 7),
    ("star25", "float32", 2, (512, 512, 512), 7),
    ("diffusion7", "float64", 2, (512, 512, 512), 7),
    ("diffusion7", "float64", 4, (512, 512, 512), 7),
    ("heat2d", "float64", 4, (1024, 1024), 0),
    ("heat2d", "float64", 10, (1024, 1024), 0),
    # two-level geometries with two cells per thread on adjacent rows (tbya)
    # (the other geometries timed with this option -- 16x32 and 24x64/16x64
    # in FP32, 32x32 at 704 threads with shallower prefetch -- were checked
    # the same way before they ran)
    ("star25", "float64", 2, (512, 512, 512), 7, {"tbws_tile": [24, 32, 320, 192], "tbya": 1}),
    ("star25", "float64", 2, (512, 512, 512), 7, {"tbws_tile": [32, 32, 448, 256], "tbya": 1, "tbrp": 3}),
]
BENCH_CASES = [c if len(c) == 6 else c + (None,) for c in BENCH_CASES]


@pytest.mark.parametrize("which,dtype,T,n,wrap,hints", BENCH_CASES,
                         ids=[f"{c[0]}-{c[1]}-T{c[2]}-w{c[4]}{'-' + str(sorted(c[5].items())) if c[5] else ''}"
                              for c in BENCH_CASES])
def test_bench_launch_in_bounds(which, dtype, T, n, wrap, hints):
    fn = getattr(w, which)(dtype)
    r = compile(fn, temporal_steps=T, temporal_mode="zero", kernel_hints=hints)
    halo = r.halo_widths
    shape = tuple(a + 2 * h for a, h in zip(n, halo))
    kern, key, args = _launch_geometry(r, shape, r.storage)
    meta, cfg = _cfg_for(fn, r, key, r.storage, hints)
    u = np.zeros(shape)
    got, errs = tm.run(fn._fn, len(n), cfg, plan_accesses(meta), tblock_queue_range(meta), u, halo, args.xch,
                       tmode=0, wrap=wrap, edges=True)
    assert not errs, errs[:5]


def _solo_cases():
    from cutile.serving import SOLO

    out = []
    for (kind, dtype), (lo, T) in SOLO.items():
        if kind == "grayscott":
            # coupled-field blocks: their solo sides are checked by
            # test_fused_block_launch_in_bounds through the fused kernel
            continue
        top = 384 if kind == "diffusion3d" else 4096
        for side in (lo, lo + 1, top - 1, top):
            out.append(pytest.param(kind, dtype, side, T, id=f"{kind}-{dtype}-{side}"))
    return out


@pytest.mark.parametrize("kind,dtype,side,T", _solo_cases())
def test_serving_solo_launch_in_bounds(kind, dtype, side, T):
    """The temporal launches the serving loop makes for large jobs (zero
    Dirichlet, rows aligned or ragged: vector or scalar variant), from the
    smallest solo side to the largest side of the job stream."""
    fn = w.diffusion7(dtype, coef=0.1) if kind == "diffusion3d" else w.heat2d(dtype)
    r = compile(fn, temporal_steps=T, temporal_mode="zero")
    halo = r.halo_widths
    shape = tuple(side + 2 * h for h in halo)
    kern, key, args = _launch_geometry(r, shape, r.storage)
    meta, cfg = _cfg_for(fn, r, key, r.storage)
    got, errs = tm.run(fn._fn, len(shape), cfg, plan_accesses(meta), tblock_queue_range(meta), np.zeros(shape),
                       halo, args.xch, tmode=0, wrap=0, edges=True)
    assert not errs, errs[:5]


@pytest.mark.parametrize("dtype,T,n,wrap", [("float64", 2, 4096, 5), ("float64", 4, 4096, 5), ("float32", 2, 4096, 5),
                                            ("float32", 4, 4096, 5), ("float64", 2, 1000, 0), ("float64", 4, 999, 0),
                                            # the serving loop's solo Gray-Scott jobs
                                            ("float64", 4, 1536, 5), ("float64", 4, 1537, 5), ("float64", 4, 4095, 5),
                                            ("float32", 4, 2800, 5), ("float32", 4, 2801, 5), ("float32", 4, 4095, 5)])
def test_fused_block_launch_in_bounds(dtype, T, n, wrap):
    """Launches of the coupled-field temporal kernel (fused Gray-Scott) that
    ``FusedStepper`` makes: the benchmark size wrapped, and zero-Dirichlet
    ones, on the model's edge workgroups."""
    from cutile import compile_fused
    from cutile.lowering.meta import extract_fused_meta
    from cutile.runtime.launcher import compile_fused_block

    fns = list(w.gray_scott(dtype))
    one = compile_fused(fns)
    blk = compile_fused_block(fns, T, one, ["u", "v"])
    halo = blk.halo_widths
    shape = tuple(n + 2 * h for h in halo)
    kern = blk.load_module()._KERNEL
    dt = getattr(torch, K.STORAGE_TORCH[kern.storage])
    ins = [torch.zeros(shape, dtype=dt) for _ in range(2)]
    key = kern.temporal_key(ins, ins, halo)
    assert key.startswith("tblock"), kern.variants.keys()
    args, _ = kern.args(ins, ins, halo, key)
    meta = extract_fused_meta([f._ir for f in fns])
    vz = kern.variants[key][3]
    cfg = select_tblock(meta, T, blk.storage, blk.storage, vz)
    assert tblock_symbol(meta, cfg) == kern.variants[key][0]
    names = list(meta.input_names)
    argmap = [tm._arg_fields(f._fn, 2, names) for f in fns]
    _, errs = tm.run_fields([f._fn for f in fns], argmap, 2, cfg, plan_accesses(meta), tblock_queue_range(meta),
                            [np.zeros(shape), np.zeros(shape)], halo, args.xch, tmode=0, wrap=wrap, edges=True)
    assert not errs, errs[:5]
