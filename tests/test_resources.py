"""Register and scratch footprint of the default kernels (CPU: hiprtc + the
code-object notes, no device).  A default variant that spills or needs more
than the 256 architectural VGPRs is rejected before it can reach a GPU."""

import pytest

from cutile import compile, compile_fused
from cutile import workloads as w
from cutile.runtime import regcheck, rt
from cutile.runtime.kernel import code_object

pytestmark = pytest.mark.skipif(not rt.available() or regcheck.readelf_path() is None,
                                reason="needs libstencil_rt.so (hiprtc) and llvm-readelf")


def _fused_block(dtype, T):
    from cutile.runtime.launcher import compile_fused_block

    fns = list(w.gray_scott(dtype))
    return compile_fused_block(fns, T, compile_fused(fns), ["u", "v"])


def _cases():
    return [
        ("heat2d", lambda: compile(w.heat2d())),
        ("diffusion7_f64", lambda: compile(w.diffusion7())),
        ("diffusion7_f32", lambda: compile(w.diffusion7("float32"))),
        ("diffusion7_bf16", lambda: compile(w.diffusion7("float32"), storage="bfloat16")),
        ("star25_f64", lambda: compile(w.star25())),
        ("star25_f32", lambda: compile(w.star25("float32"))),
        ("wave2d", lambda: compile(w.wave2d())),
        ("gray_scott_f64", lambda: compile_fused(list(w.gray_scott()))),
        ("gray_scott_f32", lambda: compile_fused(list(w.gray_scott("float32")))),
        ("star25_f64_T2", lambda: compile(w.star25(), temporal_steps=2, temporal_mode="recompute")),
        ("heat2d_T4", lambda: compile(w.heat2d(), temporal_steps=4, temporal_mode="recompute")),
        ("heat2d_T10", lambda: compile(w.heat2d(), temporal_steps=10)),
        ("diffusion7_f64_T2", lambda: compile(w.diffusion7(), temporal_steps=2)),
        ("diffusion7_f64_T4", lambda: compile(w.diffusion7(), temporal_steps=4)),
        ("diffusion7_f32_T4", lambda: compile(w.diffusion7("float32"), temporal_steps=4)),
        ("star25_f32_T2", lambda: compile(w.star25("float32"), temporal_steps=2)),
        ("star25_f32_T4", lambda: compile(w.star25("float32"), temporal_steps=4)),
        ("star25_f64_T4", lambda: compile(w.star25(), temporal_steps=4)),
        ("diffusion7_f32_T2", lambda: compile(w.diffusion7("float32"), temporal_steps=2)),
        ("heat2d_f32_T4", lambda: compile(w.heat2d("float32"), temporal_steps=4)),
        ("heat2d_f32_T10", lambda: compile(w.heat2d("float32"), temporal_steps=10)),
        ("gray_scott_f64_T2", lambda: _fused_block("float64", 2)),
        ("gray_scott_f64_T4", lambda: _fused_block("float64", 4)),
        ("gray_scott_f32_T4", lambda: _fused_block("float32", 4)),
        ("gray_scott_kahan_bf16", lambda: compile_fused(list(w.gray_scott_rates("float32")), update="kahan",
                                                        targets=["u", "v"], storage="bfloat16")),
        ("gray_scott_kahan_f32", lambda: compile_fused(list(w.gray_scott_rates("float32")), update="kahan",
                                                       targets=["u", "v"])),
    ]


@pytest.mark.parametrize("name,make", _cases(), ids=[c[0] for c in _cases()])
def test_default_variants_fit_registers(name, make):
    r = make()
    res = regcheck.code_object_resources(code_object(r.hip_source, "gfx950"))
    kernels = {k: v for k, v in res.items() if "_sweep_" in k or "_temporal_" in k or "_tblock_" in k}
    assert kernels, res
    for k, v in kernels.items():
        assert v.stack_bytes == 0, f"{k} spills ({v.stack_bytes} B scratch)"
        assert v.registers <= 256, f"{k} uses {v.registers} VGPRs"
        assert v.lds_bytes <= 160 * 1024, f"{k} uses {v.lds_bytes} B LDS"


_RANDOM_SHAPES = [(1, 1, 1), (1, 2, 2), (2, 1, 3), (2, 2, 2), (2, 4, 4), (3, 1, 2), (3, 2, 1), (3, 3, 3),
                  (3, 4, 2), (3, 1, 4)]


@pytest.mark.parametrize("seed", [2, 3, 10, 30, 41, 57, 73, 99])
def test_random_stencils_build_within_lds(seed):
    """Many fields x many plane offsets need one LDS buffer each: the tile
    must shrink until they fit the 160 KiB of a CU, and the kernel build."""
    from helpers import random_stencil

    nd, nf, rad = _RANDOM_SHAPES[seed % len(_RANDOM_SHAPES)]
    fn, _, _ = random_stencil(seed, nd, nf, rad, 12)
    r = compile(fn, temporal_blocking=False)
    res = regcheck.code_object_resources(code_object(r.hip_source, "gfx950"))
    for k, v in res.items():
        assert v.lds_bytes <= 160 * 1024, (k, v)


@pytest.mark.parametrize("fn,ty,tz,vz,nt", [
    (lambda: w.star25(), 8, 40, 2, 160),          # 20 lanes a row: padded pitch
    (lambda: w.star25(), 32, 64, 2, 512),         # 32 lanes a row: unpadded
    (lambda: w.diffusion7("float32"), 16, 64, 4, 256),   # 16 lanes a row, 4-byte cells
    (lambda: w.heat2d(), 1, 512, 2, 256),
])
def test_sweep_lds_matches_the_host_estimate(fn, ty, tz, vz, nt):
    """The LDS the compiled sweep reserves equals the emitter's estimate,
    row pitch padding included (the estimate decides tiles and occupancy)."""
    from cutile.frontend.ir import find_func
    from cutile.lowering.hip_emitter import KernelConfig, emit_hip_source, lds_bytes
    from cutile.lowering.meta import extract_meta
    from cutile.lowering.stencil_to_hip import storage_compute

    meta = extract_meta(find_func(fn()._ir))
    storage, compute = storage_compute(meta.dtype, None)
    cfg = KernelConfig(ty=ty, tz=tz, vz=vz, nt=nt)
    src = emit_hip_source(meta, [cfg], storage, compute)
    res = regcheck.code_object_resources(code_object(src, "gfx950"))
    (v,) = res.values()
    assert v.lds_bytes == lds_bytes(meta, cfg, compute)


@pytest.mark.parametrize("tiles", [(32, 128, 128), (16, 128, 128), (32, 64, 64)])
def test_oversized_tiles_are_refused_before_launch(tiles):
    """The 7-point tiles whose tuning run ended in a GPU memory fault spill
    or exceed 256 VGPR+AGPR in at least one variant; the product launch path
    refuses such a variant from its code object before it looks at the
    arrays or a device, and the variants that fit stay launchable."""
    import torch

    from cutile.runtime.kernel import code_resources

    r = compile(w.diffusion7(), tile_sizes=tiles)
    kern = r.load_module()._KERNEL
    res = code_resources(r.hip_source, "gfx950")
    bad = {k for k, (sym, *_rest) in kern.variants.items()
           if res[sym].stack_bytes > 0 or res[sym].registers > 256}
    assert bad, {k: res[v[0]] for k, v in kern.variants.items()}
    for key in kern.variants:
        if key in bad:
            with pytest.raises(RuntimeError, match="refused"):
                kern.assert_launchable(key)
        else:
            kern.assert_launchable(key)
    if "vec" in bad:
        # CPU tensors: the refusal comes first, before the device check
        shape = tuple(16 + 2 * h for h in r.halo_widths)
        a, b = torch.zeros(shape, dtype=torch.float64), torch.zeros(shape, dtype=torch.float64)
        with pytest.raises(RuntimeError, match="refused"):
            r.launcher(a, b)


def test_default_tiles_pass_the_launch_guard():
    from cutile.runtime.kernel import assert_launchable

    for r in (compile(w.diffusion7()), compile(w.star25()), compile(w.star25("float32"))):
        for sym, *_ in r.load_module()._KERNEL.variants.values():
            assert_launchable(r.hip_source, sym, "gfx950")
