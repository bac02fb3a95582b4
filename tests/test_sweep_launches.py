"""The sweep launches of the application drivers and the benchmark suite,
replayed on the host bounds model (CPU only).

Each case compiles the stencil as the application script does, builds the launch
arguments through ``StencilKernel.args`` (256 CUs, the MI355X count; the
variant the launch picks) and checks every global and LDS access of the
edge workgroups with ``kernel_model.check_launch`` before any of them runs
on a device."""

import pytest
import torch

import kernel_model as km
from cutile import apps, compile, compile_fused, suite
from cutile.frontend.ir import find_func
from cutile.lowering.hip_emitter import KernelConfig, plan_accesses
from cutile.lowering.meta import access_table, extract_fused_meta, extract_meta
from cutile.runtime import kernel as K


@pytest.fixture(autouse=True)
def _mi355x(monkeypatch):
    monkeypatch.setattr(K, "cu_count", lambda: 256)
    monkeypatch.setattr(K, "current_arch", lambda: "gfx950")


def _check(r, meta, shape):
    kern = r.load_module()._KERNEL
    dt = getattr(torch, K.STORAGE_TORCH[kern.storage])
    ins = [torch.zeros(shape, dtype=dt) for _ in range(kern.num_inputs)]
    outs = [torch.zeros(shape, dtype=dt) for _ in range(kern.num_outputs)]
    key = kern.pick(ins, outs, r.halo_widths)
    a, nblocks = kern.args(ins, outs, r.halo_widths, key)
    _, ty, tz, vz, nt = kern.variants[key]
    cfg = KernelConfig(ty, tz, vz, nt)
    errs = km.check_launch(cfg, plan_accesses(meta), (a.s0, a.s1, a.s2), (a.h0, a.h1, a.h2), (a.n0, a.n1, a.n2),
                           a.xch, a.yslab, accesses=access_table(meta), wrap=a.wrap,
                           elem_bytes=8 if kern.storage == "float64" else 4)
    assert not errs, errs[:5]
    assert nblocks == a.ntz * a.nty * a.nch


def _single(fn):
    return extract_meta(find_func(fn._ir))


APP_CASES = [
    ("heat_1d", lambda: apps.heat_1d_stencil(), (128,)),
    ("advection", lambda: apps.advection_stencil(), (200,)),
    ("fdtd_E", lambda: apps.fdtd_stencils()[0], (500,)),
    ("fdtd_H", lambda: apps.fdtd_stencils()[1], (500,)),
    ("wave_2d", lambda: apps.wave_2d_stencil(), (64, 64)),
    ("laplacian_3d", lambda: apps.laplacian_3d_stencil(), (16, 16, 16)),
    ("heat_2d", lambda: apps.heat_2d_stencil(), (128, 128)),
]


@pytest.mark.parametrize("name,make,n", APP_CASES, ids=[c[0] for c in APP_CASES])
def test_app_sweeps_in_bounds(name, make, n):
    fn = make()
    r = compile(fn, temporal_blocking=False)
    _check(r, _single(fn), tuple(a + 2 * h for a, h in zip(n, r.halo_widths)))


@pytest.mark.parametrize("which,n", [("shallow_water", 32), ("gray_scott", 64)])
def test_app_fused_sweeps_in_bounds(which, n):
    fns = list(apps.shallow_water_stencils() if which == "shallow_water" else apps.gray_scott_stencils())
    r = compile_fused(fns)
    meta = extract_fused_meta([f._ir for f in fns])
    _check(r, meta, tuple(n + 2 * h for h in r.halo_widths))


@pytest.mark.parametrize("name", list(suite.DOMAINS))
def test_suite_sweeps_in_bounds(name):
    fn, nd, order = suite.stencils()[name]
    r = compile(fn, temporal_blocking=False)
    for domain in suite.DOMAINS[name]:
        _check(r, _single(fn), tuple(d + 2 * h for d, h in zip(domain, r.halo_widths)))


@pytest.mark.parametrize("which,b,n", [("diffusion7", 8, (18, 19, 21)), ("star25", 8, (20, 24, 36)),
                                       ("heat2d", 32, (100, 130)), ("diffusion7", 32, (64, 64, 64))])
def test_brick_sweeps_in_bounds(which, b, n):
    """Brick-major launches use the flat variants' tiles on the logical
    (flat) shape; the brick address map is checked separately
    (``test_brick_layout.py``)."""
    from cutile import workloads as w

    fn = getattr(w, which)()
    r = compile(fn, temporal_blocking=False, layout="bricked", brick_size=b)
    kern = r.load_module()._KERNEL
    shape = tuple(a + 2 * h for a, h in zip(n, r.halo_widths))

    class V:
        def __init__(self):
            self.shape = shape

        def data_ptr(self):
            return 256

        def element_size(self):
            return 8

    key = "brick" if kern.vector_ok([V()], [V()], r.halo_widths, kern.variants["brick"][3]) else "brick_s"
    a, nblocks = kern.args([V()], [V()], r.halo_widths, key)
    _, ty, tz, vz, nt = kern.variants[key]
    assert kern.brick[2] % vz == 0
    meta = _single(fn)
    errs = km.check_launch(KernelConfig(ty, tz, vz, nt), plan_accesses(meta), (a.s0, a.s1, a.s2),
                           (a.h0, a.h1, a.h2), (a.n0, a.n1, a.n2), a.xch, a.yslab, accesses=access_table(meta),
                           elem_bytes=8 if kern.storage == "float64" else 4)
    assert not errs, errs[:5]


@pytest.mark.parametrize("n0,tiles,qd,wgs", [(512, 128, 9, 256), (512, 64, 3, 256), (384, 48, 3, 240),
                                            (320, 40, 3, 240), (448, 56, 3, 504), (256, 16, 3, 256)])
def test_3d_sweep_chunks_fill_whole_rounds(n0, tiles, qd, wgs):
    """3-D sweeps launch whole rounds of one workgroup per CU: the 512^3
    benchmark grids stay at 256 workgroups, and 384^3 takes 240 workgroups
    of 77 planes instead of 288 of 64 (a second, nearly empty round)."""
    x = K.sweep_chunk_length(n0, tiles, 256, qd)
    assert tiles * -(-n0 // x) == wgs


@pytest.mark.parametrize("n0,tiles,depth,slots,x", [(1024, 4, 9, 1280, 8), (1024, 4, 12, 1280, 16),
                                                    (4096, 16, 9, 1280, 52), (512, 496, 9, 256, 512),
                                                    (512, 176, 9, 256, 128)])
def test_temporal_chunks_weigh_workgroups_sharing_a_cu(n0, tiles, depth, slots, x):
    """Temporal chunks: whole rounds first, then warm-up steps against
    workgroups sharing a CU.  2-D heat at 1024^2 (4 column tiles) takes
    chunks of 8 or 16 rows (2 or 1 workgroups a CU) instead of 4 (4 a CU,
    measured 10 % slower); the 512^3 launches, which fill whole rounds,
    keep their chunks."""
    assert K.tblock_chunk_length(n0, tiles, n0, depth, slots, 256) == x


@pytest.mark.parametrize("dtype,n", [("float64", 1536), ("float64", 1537), ("float64", 4095), ("float32", 2800),
                                     ("float32", 2801), ("float32", 4095)])
def test_serving_solo_gray_scott_steps_in_bounds(dtype, n):
    """The single fused steps (periodic, wrapped reads) that end the serving
    loop's solo Gray-Scott jobs."""
    from cutile import workloads as w

    fns = list(w.gray_scott(dtype))
    r = compile_fused(fns)
    kern = r.load_module()._KERNEL
    shape = tuple(n + 2 * h for h in r.halo_widths)
    dt = getattr(torch, K.STORAGE_TORCH[kern.storage])
    ins = [torch.zeros(shape, dtype=dt) for _ in range(2)]
    key = kern.pick(ins, ins, r.halo_widths)
    a, nblocks = kern.args(ins, ins, r.halo_widths, key)
    a.wrap = kern.wrap_mask((0, 1), a)
    _, ty, tz, vz, nt = kern.variants[key]
    meta = extract_fused_meta([f._ir for f in fns])
    errs = km.check_launch(KernelConfig(ty, tz, vz, nt), plan_accesses(meta), (a.s0, a.s1, a.s2), (a.h0, a.h1, a.h2),
                           (a.n0, a.n1, a.n2), a.xch, a.yslab, accesses=access_table(meta), wrap=a.wrap,
                           elem_bytes=8 if dtype == "float64" else 4)
    assert not errs, errs[:5]


def test_prepared_wrapped_heat_sweep_in_bounds():
    """The wrapped single sweep that ``test_gpu_parity.py`` replays from a
    prepared record (heat2d FP64, 96^2, periodic reads on both axes); the
    arguments are built as ``StencilKernel.prepare`` builds them."""
    from cutile import workloads as w

    fn = w.heat2d()
    r = compile(fn, temporal_blocking=False)
    kern = r.load_module()._KERNEL
    shape = tuple(96 + 2 * h for h in r.halo_widths)
    u = torch.zeros(shape, dtype=torch.float64)
    key = kern.pick([u], [u], r.halo_widths)
    a, nblocks = kern.args([u], [u], r.halo_widths, key)
    a.wrap = kern.wrap_mask((0, 1), a)
    _, ty, tz, vz, nt = kern.variants[key]
    assert nblocks == a.ntz * a.nty * a.nch
    meta = _single(fn)
    errs = km.check_launch(KernelConfig(ty, tz, vz, nt), plan_accesses(meta), (a.s0, a.s1, a.s2), (a.h0, a.h1, a.h2),
                           (a.n0, a.n1, a.n2), a.xch, a.yslab, accesses=access_table(meta), wrap=a.wrap, elem_bytes=8)
    assert not errs, errs[:5]
