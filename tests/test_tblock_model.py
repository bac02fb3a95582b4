"""Host model of the register/LDS temporal-blocking kernel (CPU only).

For the geometries the emitter picks, the model (``tblock_model.py``, the
kernel's schedule step by step) must produce exactly the oracle's T-step
results in all three semantics (zero halo between levels, recompute in a
deep filled halo, periodic wrap) with no access outside its buffer, for
several chunk lengths and ragged tile counts.  Only geometries that pass here
run on a device."""

import dataclasses

import numpy as np
import pytest

import oracle
import tblock_model as tm
from cutile import workloads as w
from cutile.frontend.decorator import stencil_from_source
from cutile.frontend.ir import find_func
from cutile.lowering.hip_emitter import (plan_accesses, radius3, select_tblock, tblock_config, tblock_queue_range,
                                        tbws_adjacent)
from cutile.lowering.meta import extract_meta


def box27():
    return stencil_from_source("""
def box27(u, i, j, k):
    return (0.3 * u[i, j, k] + 0.05 * (u[i - 1, j - 1, k] + u[i + 1, j + 1, k] + u[i - 1, j, k + 1]
            + u[i + 1, j, k - 1] + u[i, j - 1, k - 1] + u[i, j + 1, k + 1]) + 0.01 * u[i + 1, j, k]
            + 0.02 * u[i, j, k - 1] - 0.015 * u[i - 1, j + 1, k - 1])
""")


def line1d():
    return stencil_from_source("""
def line1d(u, i):
    return 0.5 * u[i] + 0.2 * (u[i - 2] + u[i + 1]) + 0.1 * u[i + 2]
""")


def _setup(fn, T, storage, vz, hints=None):
    """The variant the emitter builds (``hints``: its kernel hints)."""
    meta = extract_meta(find_func(fn._ir))
    cfg = select_tblock(meta, T, storage, storage, vz, hints)
    assert cfg is not None
    return meta, cfg, plan_accesses(meta), tblock_queue_range(meta)


def _ref_zero(u, fn, ndim, halo, T):
    """T chained sweeps; levels >= 1 pass through zero-halo arrays."""
    cur = u
    inner = oracle.interior(u.shape, halo)
    for t in range(T):
        nxt = oracle.apply_stencil(cur, fn, ndim, halo)
        if t + 1 < T:
            cur = np.zeros_like(nxt)
            cur[inner] = nxt[inner]
        else:
            cur = nxt
    return cur


CASES = [
    # name, stencil factory, T, storage for the geometry, vz, interior, halo
    ("heat2d-T4-v2", w.heat2d, 4, "float64", 2, (37, 300), (4, 16)),
    ("heat2d-T4-v1", w.heat2d, 4, "float64", 1, (37, 301), (4, 5)),
    ("heat2d-T10-v2", w.heat2d, 10, "float64", 2, (23, 260), (10, 16)),
    ("diff7-T2-v2", w.diffusion7, 2, "float64", 2, (9, 50, 70), (2, 2, 16)),
    ("diff7-T4-v1", w.diffusion7, 4, "float64", 1, (9, 41, 37), (4, 4, 4)),
    ("star25-T2-v2", w.star25, 2, "float64", 2, (10, 40, 36), (8, 8, 16)),
    ("star25-T2-v1", w.star25, 2, "float64", 1, (10, 37, 35), (8, 8, 8)),
    ("star25f-T2-v4", w.star25, 2, "float32", 4, (10, 56, 72), (8, 8, 32)),
    # the two-level kernel with wave roles: the default for the 25-point
    # star at depth 2 (above), forced for the 7-point star and forced
    # geometries with two cells per thread
    ("star25-T2-v2-ws24x32", w.star25, 2, "float64", 2, (9, 50, 70), (8, 8, 16), {"tbws_tile": (24, 32, 640, 384)}),
    ("star25-T2-v2-ws2c", w.star25, 2, "float64", 2, (9, 50, 70), (8, 8, 16), {"tbws_tile": (24, 32, 320, 192)}),
    ("star25-T2-v2-ws2y", w.star25, 2, "float64", 2, (9, 50, 70), (8, 8, 16),
     {"tbws_tile": (24, 32, 320, 192), "tbya": 1}),
    ("star25f-T2-v4-ws2y", w.star25, 2, "float32", 4, (9, 50, 136), (8, 8, 32),
     {"tbws_tile": (24, 64, 320, 192), "tbya": 1}),
    ("diff7-T2-v2-ws", w.diffusion7, 2, "float64", 2, (9, 30, 200), (2, 2, 16), {"tbws": 1}),
    ("diff7-T2-v1-ws", w.diffusion7, 2, "float64", 1, (9, 29, 101), (2, 2, 2), {"tbws": 1}),
    ("diff7-T2-v2-tb", w.diffusion7, 2, "float64", 2, (9, 50, 70), (2, 2, 16), {"tbws": 0}),
    ("star25-T2-v2-tb", w.star25, 2, "float64", 2, (10, 40, 36), (8, 8, 16), {"tbws": 0}),
    ("box27-T2-v2", box27, 2, "float64", 2, (8, 20, 30), (2, 2, 16)),
    ("box27-T3-v1", box27, 3, "float64", 1, (8, 21, 31), (3, 3, 3)),
    ("line1d-T3-v1", line1d, 3, "float64", 1, (50,), (6,)),
]


CASES = [c if len(c) == 8 else c + (None,) for c in CASES]


@pytest.mark.parametrize("mode", ["zero", "recompute", "wrap"])
@pytest.mark.parametrize("name,make,T,storage,vz,n,halo,hints", CASES, ids=[c[0] for c in CASES])
def test_model_matches_oracle(name, make, T, storage, vz, n, halo, hints, mode):
    fn = make()
    ndim = len(n)
    meta, cfg, plan, qr = _setup(fn, T, storage, vz, hints)
    rng = np.random.default_rng(len(name) + T)
    shape = tuple(a + 2 * h for a, h in zip(n, halo))
    u = rng.random(shape)
    spec = w.periodic_spec(ndim)
    if mode == "zero":
        ref = _ref_zero(u, fn._fn, ndim, halo, T)
        tmode, wrap = 0, 0
    else:
        up = u.copy()
        oracle.apply_boundary(up, spec, halo)
        ref = oracle.time_march(up, fn._fn, ndim, halo, T, spec)
        if mode == "recompute":
            u, tmode, wrap = up, 1, 0
        else:
            tmode, wrap = 0, 7
    inner = oracle.interior(shape, halo)
    for xch in sorted({3, 7, n[0]} if ndim > 1 else {1}):
        got, errs = tm.run(fn._fn, ndim, cfg, plan, qr, u, halo, xch, tmode=tmode, wrap=wrap)
        assert not errs, (xch, errs[:5])
        np.testing.assert_array_equal(got[inner], ref[inner], err_msg=f"xch={xch}")


@pytest.mark.parametrize("name,make,T,storage,vz,n,halo,hints", CASES, ids=[c[0] for c in CASES])
def test_geometry_invariants(name, make, T, storage, vz, n, halo, hints):
    fn = make()
    meta, cfg, plan, (qlo, qhi) = _setup(fn, T, storage, vz, hints)
    rx, ry, rz = radius3(meta)
    assert qlo <= 0 <= qhi and qhi - qlo >= 2 * rx if rx else True
    assert cfg.gy >= (T - 1) * ry and cfg.gz >= (T - 1) * rz
    assert cfg.gz % cfg.vz == 0 and cfg.tz % cfg.vz == 0 and cfg.otz % cfg.vz == 0
    assert cfg.oty > 0 and cfg.otz > 0
    if cfg.na:
        # roles of whole waves; every region / output vector has a thread
        nb = cfg.nt - cfg.na
        assert cfg.na % 64 == 0 and nb % 64 == 0 and cfg.nt <= 1024
        assert cfg.kc * cfg.na >= cfg.ty * cfg.tz // cfg.vz > (cfg.kc - 1) * cfg.na
        assert cfg.kb * nb >= cfg.oty * cfg.otz // cfg.vz > (cfg.kb - 1) * nb
        assert T == 2 and cfg.gy >= ry and cfg.gz >= rz
    else:
        assert cfg.kc * cfg.nt >= cfg.ty * cfg.tz // cfg.vz > (cfg.kc - 1) * cfg.nt


def _ref_zero_multi(fields, fns, names, ndim, halo, T):
    """T chained fused sweeps of coupled fields; levels >= 1 pass through
    zero-halo arrays (result m is the next state of field ``names[m]``)."""
    cur = {k: v.copy() for k, v in fields.items()}
    shape = next(iter(fields.values())).shape
    inner = oracle.interior(shape, halo)
    for t in range(T):
        outs = oracle.apply_multi(cur, fns, ndim, halo, base=[cur[n] for n in names])
        nxt = {}
        for n, o in zip(names, outs):
            if t + 1 < T:
                z = np.zeros_like(o)
                z[inner] = o[inner]
                nxt[n] = z
            else:
                nxt[n] = o
        cur = nxt
    return cur


@pytest.mark.parametrize("mode", ["zero", "wrap"])
@pytest.mark.parametrize("T,vz,n,halo", [(2, 2, (37, 300), (1, 16)), (4, 2, (23, 262), (1, 16)),
                                         (2, 1, (30, 97), (1, 3)), (3, 1, (19, 140), (1, 1))])
def test_coupled_fields_model_matches_oracle(T, vz, n, halo, mode):
    """Fused Gray-Scott (two coupled fields) at temporal depth T: the model
    of the generalised kernel (a queue and LDS planes per field and level)
    equals T fused steps of the oracle, zero-halo and periodic-wrap."""
    from cutile.lowering.meta import extract_fused_meta

    fns = list(w.gray_scott("float64"))
    meta = extract_fused_meta([f._ir for f in fns])
    names = list(meta.input_names)
    assert names == ["u", "v"] and meta.num_exprs == 2
    cfg = tblock_config(meta, T, "float64", "float64", vz)
    assert cfg is not None
    plan, qr = plan_accesses(meta), tblock_queue_range(meta)
    rng = np.random.default_rng(T * 10 + vz)
    shape = tuple(a + 2 * h for a, h in zip(n, halo))
    fields = {k: rng.random(shape) for k in names}
    if mode == "zero":
        ref = _ref_zero_multi(fields, [f._fn for f in fns], names, 2, halo, T)
        tmode, wrap = 0, 0
    else:
        fp = {k: v.copy() for k, v in fields.items()}
        for v in fp.values():
            oracle.apply_boundary(v, w.periodic_spec(2), halo)
        ref = oracle.time_march_multi(fp, [f._fn for f in fns], names, 2, halo, T, w.periodic_spec(2))
        tmode, wrap = 0, 5
    argmap = [tm._arg_fields(f._fn, 2, names) for f in fns]
    inner = oracle.interior(shape, halo)
    for xch in (3, 7, n[0]):
        got, errs = tm.run_fields([f._fn for f in fns], argmap, 2, cfg, plan, qr, [fields[k] for k in names], halo,
                                  xch, tmode=tmode, wrap=wrap)
        assert not errs, (xch, errs[:5])
        for k, g in zip(names, got):
            np.testing.assert_array_equal(g[inner], ref[k][inner], err_msg=f"{k} xch={xch}")


def test_zdpp_hint_selects_the_single_role_kernel():
    """Cross-lane z-neighbours (hint ``zdpp``) are an option of the
    single-role kernel: the two-level kernel does not take them (its
    level-1 view has no neighbour-lane vectors), so the hint selects the
    other kernel even where the two-level one is the default."""
    meta = extract_meta(find_func(w.star25()._ir))
    assert select_tblock(meta, 2, "float64", "float64", 2).na > 0
    assert select_tblock(meta, 2, "float64", "float64", 2, {"zdpp": 1}).na == 0
    assert select_tblock(meta, 2, "float64", "float64", 2, {"zdpp": 1, "tbws": 1}).na == 0


def _ws_cells(cfg):
    """(level-1, level-2) cells each thread owns in ``st::tbws_tile``:
    lists of (thread, k, row, vector column) in region coordinates, active
    cells only, with the kernel's mapping (flat, or adjacent rows when
    ``cfg.ya``)."""
    ntz, ontz = cfg.tz // cfg.vz, cfg.otz // cfg.vz
    ka, kb, na, nbt = cfg.kc, cfg.kb, cfg.na, cfg.nt - cfg.na
    one, two = [], []
    for t in range(na):
        for k in range(ka):
            if cfg.ya:
                grp = t // ntz
                if grp < cfg.ty // ka:
                    one.append((t, k, grp * ka + k, t % ntz))
            elif t + k * na < cfg.ty * ntz:
                c = t + k * na
                one.append((t, k, c // ntz, c % ntz))
    for t in range(nbt):
        for k in range(kb):
            if cfg.ya:
                grp = t // ontz
                if grp < cfg.oty // kb:
                    two.append((t, k, cfg.gy + grp * kb + k, cfg.gz // cfg.vz + t % ontz))
            elif t + k * nbt < cfg.oty * ontz:
                c = t + k * nbt
                two.append((t, k, cfg.gy + c // ontz, cfg.gz // cfg.vz + c % ontz))
    return one, two


@pytest.mark.parametrize("storage,vz,force", [("float64", 2, (24, 32, 320, 192)), ("float32", 4, (24, 64, 320, 192)),
                                              ("float64", 2, (16, 32, 256, 128))])
def test_adjacent_rows_cover_each_cell_once(storage, vz, force):
    """Option ``tbya`` (each thread's cells on adjacent rows of one column):
    every region cell has exactly one level-1 owner and every output cell
    exactly one level-2 owner, as in the flat order; a thread's cells k and
    k + 1 are one row apart in the same column (what lets the kernel read a
    neighbour that is its own cell from its queue); and a geometry whose
    row groups do not divide the tile keeps the flat order."""
    meta = extract_meta(find_func(w.star25(storage)._ir))
    flat = select_tblock(meta, 2, storage, storage, vz, {"tbws_tile": force})
    adj = select_tblock(meta, 2, storage, storage, vz, {"tbws_tile": force, "tbya": 1})
    assert flat.ya == 0 and adj.ya == 1 and adj.tag == flat.tag + "y"
    for cfg in (flat, adj):
        one, two = _ws_cells(cfg)
        region = sorted((r, c) for _, _, r, c in one)
        assert region == sorted((r, c) for r in range(cfg.ty) for c in range(cfg.tz // cfg.vz))
        out = sorted((r, c) for _, _, r, c in two)
        assert out == sorted((r, c) for r in range(cfg.gy, cfg.gy + cfg.oty)
                             for c in range(cfg.gz // cfg.vz, (cfg.gz + cfg.otz) // cfg.vz))
    for cells in _ws_cells(adj):
        by = {(t, k): (r, c) for t, k, r, c in cells}
        for (t, k), (r, c) in by.items():
            if (t, k + 1) in by:
                assert by[t, k + 1] == (r + 1, c)
    odd = dataclasses.replace(flat, ty=flat.ty - 1)       # an odd region height: no whole row pairs
    assert odd.kc == 2 and tbws_adjacent(odd) is None
