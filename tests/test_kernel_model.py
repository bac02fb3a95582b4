"""Host bounds model of the sweep kernel (CPU only).

Every tile shape the tiling pass hands out, at every queue prefetch depth the
emitter accepts, is checked against the array bounds before any of them runs
on a device: register-queue loads, halo-ring fetches, LDS stores/reads and
output stores for the edge tiles of ragged and aligned shapes."""

import pytest

import kernel_model as km
from cutile import workloads as w
from cutile.frontend.ir import find_func
from cutile.lowering.hip_emitter import plan_accesses
from cutile.lowering.meta import access_table, extract_meta
from cutile.lowering.stencil_to_hip import make_configs
from cutile.passes.tiling import TilingPass


@pytest.mark.parametrize("qlo,qhi", [(-1, 1), (-4, 4), (0, 0), (-2, 0)])
@pytest.mark.parametrize("pd", [1, 2, 3])
@pytest.mark.parametrize("xb,xe", [(0, 1), (0, 2), (0, 7), (64, 128), (10, 13)])
def test_queue_schedule_loads_each_plane_once(qlo, qhi, pd, xb, xe):
    planes = km.queue_planes(xb, xe, qlo, qhi, pd)
    assert sorted(planes) == list(range(xb + qlo, xe + qhi))
    assert len(set(planes)) == len(planes)


def _cases():
    out = []
    for fn in (w.diffusion7(), w.star25(), w.star25("float32")):
        func = find_func(fn._ir)
        meta = extract_meta(func)
        eb = 8 if func.dtype == "float64" else 4
        tiles = {(t, h.get("threads")) for t, h in TilingPass(dtype_bytes=eb).candidates(3, len(access_table(meta)))}
        for t, nt in sorted(tiles, key=str):
            out.append(pytest.param(meta, t, nt, id=f"{meta.name}-{func.dtype}-{'x'.join(map(str, t))}-n{nt}"))
    return out


@pytest.mark.parametrize("meta,tile,threads", _cases())
@pytest.mark.parametrize("pd", [1, 2])
@pytest.mark.parametrize("n", [(512, 512, 512), (37, 45, 70), (130, 9, 260)])
def test_tuned_tiles_in_bounds(meta, tile, threads, pd, n):
    plan = plan_accesses(meta)
    h = max(abs(o) for a in meta.accesses for o in a.offsets)
    halo3 = (h, h, h)
    shape3 = tuple(k + 2 * h for k in n)
    vec, scal, xch = make_configs(3, tile, meta.dtype, threads=threads)
    acc = access_table(meta)
    for cfg, rows_ok in ((vec, shape3[2] % vec.vz == 0), (scal, True)):
        if not rows_ok:
            continue
        for x in (xch, 64, 16):
            errs = km.check_launch(cfg, plan, shape3, halo3, n, x, 0, accesses=acc, pd=pd,
                                   elem_bytes=8 if meta.dtype == "float64" else 4)
            assert not errs, errs[:5]


def _autotune_cases():
    from cutile.runtime.autotune import generate_candidates

    out = []
    for fn in (w.diffusion7(), w.star25(), w.star25("float32")):
        func = find_func(fn._ir)
        meta = extract_meta(func)
        eb = 8 if func.dtype == "float64" else 4
        for tile, T, hints in generate_candidates(3, eb, len(access_table(meta))):
            if T == 1:
                out.append(pytest.param(meta, tile, hints.get("threads"),
                                        id=f"{meta.name}-{func.dtype}-{'x'.join(map(str, tile))}-n{hints.get('threads')}"))
    return out


@pytest.mark.parametrize("meta,tile,threads", _autotune_cases())
def test_autotune_candidates_in_bounds(meta, tile, threads):
    plan = plan_accesses(meta)
    h = max(abs(o) for a in meta.accesses for o in a.offsets)
    acc = access_table(meta)
    vec, scal, xch = make_configs(3, tile, meta.dtype, threads=threads)
    for n in ((64, 64, 64), (37, 45, 70)):
        halo3 = (h, h, -(-h // vec.vz) * vec.vz)
        shape3 = tuple(k + 2 * hh for k, hh in zip(n, halo3))
        for cfg in (vec, scal):
            if shape3[2] % cfg.vz:
                continue
            for x in (8, 64):
                errs = km.check_launch(cfg, plan, shape3, halo3, n, x, 0, accesses=acc,
                                       elem_bytes=8 if meta.dtype == "float64" else 4)
                assert not errs, errs[:5]


@pytest.mark.parametrize("seed", [2, 3, 10, 30, 41, 57, 73, 99])
def test_random_stencil_launches_in_bounds(seed):
    """The compiled tile of the random device-parity stencils (several fields,
    several plane offsets, radius up to 4) at the parity test's shapes."""
    from helpers import random_stencil

    from cutile import compile

    shapes = [(1, 1, 1), (1, 2, 2), (2, 1, 3), (2, 2, 2), (2, 4, 4), (3, 1, 2), (3, 2, 1), (3, 3, 3),
              (3, 4, 2), (3, 1, 4)]
    nd, nf, rad = shapes[seed % len(shapes)]
    fn, _, _ = random_stencil(seed, nd, nf, rad, 12)
    r = compile(fn, temporal_blocking=False)
    meta = extract_meta(find_func(fn._ir))
    plan = plan_accesses(meta)
    acc = access_table(meta)
    vec, scal, xch = make_configs(nd, r.tile_sizes, meta.dtype, threads=r.kernel_hints.get("threads"),
                                  vector=r.kernel_hints.get("vector"))
    base = {1: (300,), 2: (37, 70), 3: (11, 13, 40)}[nd]
    shape = tuple(b + 2 * h for b, h in zip(base, r.halo_widths))
    if nd == 1:
        s3, h3, n3 = (1, 1, shape[0]), (0, 0, r.halo_widths[0]), (1, 1, base[0])
    elif nd == 2:
        s3, h3, n3 = (shape[0], 1, shape[1]), (r.halo_widths[0], 0, r.halo_widths[1]), (base[0], 1, base[1])
    else:
        s3, h3, n3 = shape, tuple(r.halo_widths), base
    for cfg in (vec, scal):
        if s3[2] % cfg.vz or h3[2] % cfg.vz:
            continue
        for x in (xch, 4, 1):
            errs = km.check_launch(cfg, plan, s3, h3, n3, x, 0, accesses=acc,
                                   elem_bytes=8 if meta.dtype == "float64" else 4)
            assert not errs, errs[:5]


@pytest.mark.parametrize("n", [(512, 512, 512), (4, 4, 16), (37, 45, 70), (9, 10, 34)])
def test_wrapped_periodic_reads_in_bounds(n):
    """Periodic (wrapped) reads stay inside the array for interiors down to
    the halo width."""
    for fn in (w.diffusion7(), w.star25()):
        meta = extract_meta(find_func(fn._ir))
        plan = plan_accesses(meta)
        h = max(abs(o) for a in meta.accesses for o in a.offsets)
        t, hints = TilingPass(dtype_bytes=8).candidates(3, len(access_table(meta)))[0]
        vec, scal, xch = make_configs(3, t, meta.dtype, threads=hints.get("threads"))
        halo3 = (h, h, 16)
        if any(nn < hh for nn, hh in zip(n, halo3)):
            continue
        shape3 = tuple(k + 2 * hh for k, hh in zip(n, halo3))
        for cfg in (vec, scal):
            if shape3[2] % cfg.vz:
                continue
            for x in (xch, 3):
                errs = km.check_launch(cfg, plan, shape3, halo3, n, x, 0, accesses=access_table(meta), wrap=7,
                                       elem_bytes=8 if meta.dtype == "float64" else 4)
                assert not errs, errs[:5]


@pytest.mark.parametrize("tile", [(32, 128, 128), (16, 128, 128), (32, 64, 64)])
@pytest.mark.parametrize("n", [(512, 512, 512), (128, 256, 256), (64, 64, 64), (37, 45, 70)])
def test_fault_record_tiles_in_bounds(tile, n):
    """The 7-point tiles of the tuning run that ended in a GPU memory fault
    (vector and scalar variants, with and without wrapped reads): every
    global and LDS access of the model stays inside its buffer.  The launch
    path refuses these variants from their code objects (scratch or more
    than 256 VGPR+AGPR); this pins that the fault does not follow from the
    kernel's indexing at these shapes."""
    meta = extract_meta(find_func(w.diffusion7()._ir))
    plan = plan_accesses(meta)
    acc = access_table(meta)
    vec, scal, xch = make_configs(3, tile, meta.dtype)
    halo3 = (1, 1, 16)
    shape3 = tuple(k + 2 * hh for k, hh in zip(n, halo3))
    for cfg in (vec, scal):
        if shape3[2] % cfg.vz:
            continue
        for x in (xch, 64, 7):
            for wrap in (0, 7):
                errs = km.check_launch(cfg, plan, shape3, halo3, n, x, 0, accesses=acc, wrap=wrap)
                assert not errs, errs[:5]


def test_launch_block_matches_launch_bounds():
    """Every variant is launched with exactly the workgroup size its kernel
    was compiled for (``__launch_bounds__``), the fault-record tiles and the
    temporal kernels included."""
    import re

    from cutile import compile

    results = [compile(w.diffusion7(), tile_sizes=t) for t in ((32, 128, 128), (16, 128, 128), (32, 64, 64))]
    results += [compile(w.star25()), compile(w.star25(), temporal_steps=2), compile(w.heat2d(), temporal_steps=4),
                compile(w.diffusion7(), layout="bricked")]
    for r in results:
        bounds = dict((s, int(n)) for n, s in
                      re.findall(r"__launch_bounds__\((\d+)\)\s+(\w+)\(", r.hip_source))
        kern = r.load_module()._KERNEL
        for key, (sym, _ty, _tz, _vz, nt) in kern.variants.items():
            assert bounds[sym] == nt, (key, sym, bounds[sym], nt)
