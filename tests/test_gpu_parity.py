"""Device parity: every HIP kernel against the oracle on the same seeded
inputs (``-m gpu``).  Integer-exact comparisons (``max|diff| == 0``) where
the arithmetic is the same IEEE sequence (float64 and float32 stencils);
stated tolerances for the bfloat16-storage variant."""

import math

import numpy as np
import pytest

from helpers import random_stencil, to_dev, to_host

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _bc(u, spec, halo):
    from cutile.runtime import boundary

    boundary.fill(u, spec, halo)


def _march_compare(oracle, result, u0, steps, spec, exact=True, rtol=0.0):
    """Step on the device (launch + halo fill) and compare with the oracle
    after every step."""
    import torch

    fn = result._ref_fn
    halo = result.halo_widths
    dt = getattr(torch, result.storage)
    a = to_dev(u0, dt)
    b = torch.zeros_like(a)
    ref = u0.copy()
    launch = result.launcher
    worst = 0.0
    for s in range(steps):
        launch(a, b)
        _bc(b, spec, halo)
        a, b = b, a
        ref = oracle.apply_stencil(ref, fn, result.ndim, halo)
        oracle.apply_boundary(ref, spec, halo)
        got = to_host(a)
        if exact:
            diff = float(np.max(np.abs(got - ref)))
            assert diff == 0.0, f"step {s + 1}: max|diff| = {diff:.3e}"
        else:
            err = oracle.rel_max_err(got, ref)
            worst = max(worst, err)
            assert err <= rtol, f"step {s + 1}: rel err {err:.3e} > {rtol:.1e}"
    return worst


def test_heat2d_small_dirichlet_every_step(oracle):
    from cutile import compile
    from cutile.workloads import dirichlet_spec, heat2d, heat2d_init

    _flush("heat2d 64x64 fp64 x20")
    r = compile(heat2d())
    assert r.halo_widths == (1, 16)
    u0 = heat2d_init(64, r.halo_widths, seed=0)
    _march_compare(oracle, r, u0, 20, dirichlet_spec(2), exact=True)


def test_heat2d_c1_full_100_steps(oracle):
    """Configuration C1: 1024^2 + halo 1 (inner padded to 16), zero Dirichlet,
    uniform[0,1) from seed 0, 100 steps, bit-exact after every step."""
    from cutile import compile
    from cutile.workloads import dirichlet_spec, heat2d, heat2d_init

    _flush("heat2d 1024x1024 fp64 x100")
    r = compile(heat2d())
    u0 = heat2d_init(1024, r.halo_widths, seed=0)
    _march_compare(oracle, r, u0, 100, dirichlet_spec(2), exact=True)


@pytest.mark.parametrize("shape", [(40, 36, 48), (17, 9, 130), (64, 64, 64)])
def test_diffusion7_periodic(oracle, shape):
    from cutile import compile
    from cutile.workloads import diffusion7, periodic_spec, sinusoid_init

    _flush(f"diffusion7 {shape} fp64 x4")
    r = compile(diffusion7())
    assert r.halo_widths == (1, 1, 16)
    u0 = sinusoid_init(shape, r.halo_widths, seed=1)
    oracle.apply_boundary(u0, periodic_spec(3), r.halo_widths)
    _march_compare(oracle, r, u0, 4, periodic_spec(3), exact=True)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_star25_periodic_exact(oracle, dtype):
    from cutile import compile
    from cutile.workloads import periodic_spec, sinusoid_init, star25

    _flush(f"star25 (24,40,72) {dtype} x3")
    r = compile(star25(dtype), temporal_blocking=False)
    assert r.stencil_halo == (4, 4, 4)
    u0 = sinusoid_init((24, 40, 72), r.halo_widths, seed=2, dtype=np.dtype(dtype).type)
    oracle.apply_boundary(u0, periodic_spec(3), r.halo_widths)
    _march_compare(oracle, r, u0, 3, periodic_spec(3), exact=True)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_gray_scott_fused_exact(oracle, dtype):
    import torch

    from cutile import compile_fused
    from cutile.workloads import gray_scott, gray_scott_init, periodic_spec

    _flush(f"gray-scott fused 96x80 {dtype} x10")
    gu, gv = gray_scott(dtype)
    r = compile_fused([gu, gv])
    assert r.inputs == ("u", "v") and r.outputs == ("gs_u", "gs_v")
    halo = r.halo_widths
    u, v = gray_scott_init(96, halo, seed=3, squares=4, dtype=np.dtype(dtype).type)
    u, v = u[:, : 80 + 2 * halo[1]].copy(), v[:, : 80 + 2 * halo[1]].copy()
    spec = periodic_spec(2)
    oracle.apply_boundary(u, spec, halo)
    oracle.apply_boundary(v, spec, halo)
    du, dv = to_dev(u), to_dev(v)
    nu, nv = torch.zeros_like(du), torch.zeros_like(dv)
    f = {"u": u, "v": v}
    for s in range(10):
        r.launcher(du, dv, nu, nv)
        _bc(nu, spec, halo)
        _bc(nv, spec, halo)
        du, nu = nu, du
        dv, nv = nv, dv
        outs = oracle.apply_multi(f, [gu._fn, gv._fn], 2, halo, base=[f["u"], f["v"]])
        for o in outs:
            oracle.apply_boundary(o, spec, halo)
        f = {"u": outs[0], "v": outs[1]}
        assert float(np.max(np.abs(to_host(du) - f["u"]))) == 0.0, f"u step {s + 1}"
        assert float(np.max(np.abs(to_host(dv) - f["v"]))) == 0.0, f"v step {s + 1}"


def test_wave2d_leapfrog_exact(oracle):
    import torch

    from cutile import compile
    from cutile.workloads import dirichlet_spec, wave2d, wave_init

    _flush("wave2d 100x130 fp32 x12")
    w = wave2d("float32")
    r = compile(w)
    halo = r.halo_widths
    u, uold = wave_init(100, halo, seed=4)
    u = np.pad(u, ((0, 0), (0, 30)))[:, : 130 + 2 * halo[1]].copy()
    uold = u.copy()
    spec = dirichlet_spec(2)
    oracle.apply_boundary(u, spec, halo)
    oracle.apply_boundary(uold, spec, halo)
    a, b = to_dev(u), to_dev(uold)
    c = torch.zeros_like(a)
    ra, rb = oracle.leapfrog(u, uold, w._fn, 2, halo, 0, spec)
    for s in range(12):
        r.launcher(a, b, c)
        _bc(c, spec, halo)
        a, b, c = c, a, b
        ra, rb = oracle.leapfrog(ra, rb, w._fn, 2, halo, 1, spec)
        assert float(np.max(np.abs(to_host(a) - ra))) == 0.0, f"step {s + 1}"


@pytest.mark.parametrize("seed", [2, 3, 10, 30, 41, 57, 73, 99])
def test_random_expressions_match_oracle(oracle, seed):
    """Random expressions (products, divisions, max/min/abs, where, nested
    differences) over 1-4 fields in 1-3 D, radius up to 4: bit-exact."""
    import torch

    from cutile import compile

    shapes = [(1, 1, 1), (1, 2, 2), (2, 1, 3), (2, 2, 2), (2, 4, 4), (3, 1, 2), (3, 2, 1), (3, 3, 3),
              (3, 4, 2), (3, 1, 4)]
    nd, nf, rad = shapes[seed % len(shapes)]
    fn, fields, body = random_stencil(seed, nd, nf, rad, 12)
    r = compile(fn, temporal_blocking=False)
    _flush(f"random seed {seed}: {nd}D {len(r.inputs)} fields halo {r.halo_widths}")
    rng = np.random.default_rng(seed)
    base = {1: (300,), 2: (37, 70), 3: (11, 13, 40)}[nd]
    shape = tuple(b + 2 * h for b, h in zip(base, r.halo_widths))
    arrs = {f: rng.random(shape) + 0.5 for f in r.inputs}
    dev = [to_dev(arrs[f]) for f in r.inputs]
    out = torch.zeros_like(dev[0])
    r.launcher(*dev, out)
    ref = oracle.eval_multi(arrs, fn._fn, nd, r.halo_widths)
    got = to_host(out)[oracle.interior(shape, r.halo_widths)]
    assert np.array_equal(got, ref), body


def test_unaligned_arrays_use_the_scalar_variant(oracle):
    import torch

    from cutile import compile
    from cutile.workloads import diffusion7

    _flush("diffusion7 unpadded halo (scalar variant)")
    r = compile(diffusion7(), align_halo=False)
    assert r.halo_widths == (1, 1, 1)
    rng = np.random.default_rng(5)
    u = rng.random((20, 21, 35))
    du = to_dev(u)
    out = torch.zeros_like(du)
    r.launcher(du, out)
    ref = oracle.apply_stencil(u, r._ref_fn, 3, (1, 1, 1))
    inner = oracle.interior(u.shape, (1, 1, 1))
    assert np.array_equal(to_host(out)[inner], ref[inner])


@pytest.mark.parametrize("kinds", [
    ("periodic",) * 3,
    ("neumann",) * 3,
    ("reflecting",) * 3,
    (("dirichlet", 0.5),) * 3,
    ("periodic", "neumann", ("dirichlet", -1.0)),
])
def test_halo_fill_matches_oracle(oracle, kinds):
    from cutile.runtime import boundary

    rng = np.random.default_rng(7)
    halo = (2, 3, 4)
    u = rng.random((9, 11, 16))
    spec = {d: (k, k) for d, k in enumerate(kinds)}
    d = to_dev(u)
    boundary.fill(d, spec, halo)
    ref = u.copy()
    oracle.apply_boundary(ref, spec, halo)
    assert np.array_equal(to_host(d), ref)


def test_mixed_sides_and_narrow_interior(oracle):
    from cutile.runtime import boundary

    rng = np.random.default_rng(8)
    halo = (3, 2)
    u = rng.random((8, 10))  # interior 2 < halo 3 along dim 0
    spec = {0: ("reflecting", "periodic"), 1: (("dirichlet", 2.0), "neumann")}
    d = to_dev(u)
    boundary.fill(d, spec, halo)
    ref = u.copy()
    oracle.apply_boundary(ref, spec, halo)
    assert np.array_equal(to_host(d), ref)


def test_temporal_chain_matches_repeated_sweeps(oracle):
    import torch

    from cutile import compile
    from cutile.workloads import heat2d

    _flush("heat2d temporal_steps=3")
    r = compile(heat2d(), temporal_steps=3)
    assert r.temporal_steps == 3
    assert r.validate(reference=oracle, atol=0.0)


def test_bfloat16_storage_error_is_bounded(oracle):
    import torch

    from cutile import compile
    from cutile.workloads import diffusion7, periodic_spec, sinusoid_init

    _flush("diffusion7 bf16 storage x5")
    r = compile(diffusion7("float32"), storage="bfloat16")
    assert r.storage == "bfloat16"
    u0 = sinusoid_init((32, 32, 64), r.halo_widths, seed=9)
    oracle.apply_boundary(u0, periodic_spec(3), r.halo_widths)
    ref64 = compile(diffusion7())
    # reference in float64 with the bf16-rounded initial field
    u0b = to_host(to_dev(u0, __import__("torch").bfloat16))
    worst = 0.0
    a = to_dev(u0, __import__("torch").bfloat16)
    b = __import__("torch").zeros_like(a)
    ref = u0b.copy()
    for s in range(5):
        r.launcher(a, b)
        _bc(b, periodic_spec(3), r.halo_widths)
        a, b = b, a
        ref = oracle.apply_stencil(ref, ref64._ref_fn, 3, r.halo_widths)
        oracle.apply_boundary(ref, periodic_spec(3), r.halo_widths)
        worst = max(worst, oracle.rel_max_err(to_host(a), ref))
    _flush(f"bf16 worst rel err over 5 steps: {worst:.3e}")
    assert worst < 2e-2


def test_diff_norms_and_copy():
    import torch

    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream

    rng = np.random.default_rng(10)
    a = rng.random((7, 9, 33))
    b = a.copy()
    b[3, 4, 20] += 0.25
    da, db = to_dev(a), to_dev(b)
    md, mb = rt.diff_norms(da.data_ptr(), rt.ST_F64, db.data_ptr(), rt.ST_F64, a.shape, (1, 1, 1),
                           default_stream())
    inner = (slice(1, -1),) * 3
    assert md == pytest.approx(0.25)
    assert mb == float(np.max(np.abs(b[inner])))
    dc = torch.empty_like(da)
    rt.copy(dc.data_ptr(), da.data_ptr(), da.numel() * 8, default_stream())
    torch.cuda.synchronize()
    assert np.array_equal(to_host(dc), a)


def test_bricks_round_trip():
    from cutile.runtime.bricks import from_bricks, to_bricks

    rng = np.random.default_rng(11)
    a = rng.random((37, 20, 45))
    d = to_dev(a)
    b = to_bricks(d, 16)
    assert tuple(b.shape) == (3, 2, 3, 16, 16, 16)
    bh = to_host(b)
    assert np.array_equal(bh[1, 0, 2, 3, 4, 5], a[16 + 3, 4, 32 + 5])
    assert bh[2, 1, 2, 15, 15, 15] == 0.0  # padding
    back = from_bricks(b, a.shape, 16)
    assert np.array_equal(to_host(back), a)


def test_bricked_launcher_matches_flat(oracle):
    import torch

    from cutile import compile
    from cutile.workloads import diffusion7

    r = compile(diffusion7(), layout="bricked", brick_size=8)
    assert r.halo_widths == (1, 1, 1)
    mod = r.load_module()
    rng = np.random.default_rng(12)
    u = rng.random((18, 19, 21))
    out = torch.zeros(u.shape, dtype=torch.float64, device="cuda")
    mod.launch_diffusion7_bricked(to_dev(u), out, brick_size=8)
    ref = oracle.apply_stencil(u, r._ref_fn, 3, (1, 1, 1))
    inner = oracle.interior(u.shape, (1, 1, 1))
    assert np.array_equal(to_host(out)[inner], ref[inner])


@pytest.mark.parametrize("which,b,n,steps", [("diffusion7", 8, (18, 19, 21), 3), ("star25", 8, (20, 24, 36), 2),
                                             ("heat2d", 32, (100, 130), 4)])
def test_brick_major_sweeps(oracle, which, b, n, steps):
    """Steps run on the bricks themselves (ping-pong between two brick-major
    arrays) equal the oracle's steps on the flat array, halos untouched."""
    import torch

    from cutile import compile
    from cutile import workloads as w
    from cutile.runtime.bricks import from_bricks, to_bricks

    _flush(f"brick-major {which} bricks of {b} {n} x {steps}")
    fn = getattr(w, which)()
    r = compile(fn, temporal_blocking=False, layout="bricked", brick_size=b)
    mod = r.load_module()
    halo = r.halo_widths
    shape = tuple(a + 2 * h for a, h in zip(n, halo))
    u = np.random.default_rng(21).random(shape)
    ua = to_bricks(to_dev(u), b)
    ub = to_bricks(to_dev(u), b)
    launch = getattr(mod, f"launch_{r.name}_on_bricks")
    for _ in range(steps):
        launch(ua, ub, shape)
        ua, ub = ub, ua
    got = to_host(from_bricks(ua, shape, b))
    ref = u
    for _ in range(steps):
        ref = oracle.apply_stencil(ref, fn._fn, len(n), halo)
    assert np.array_equal(got, ref)


def test_graph_replay_matches_eager(oracle):
    """A captured step loop (sweep + halo fill, 3 steps) replayed several
    times with eager work in between gives the eager result every time."""
    import torch

    from cutile import compile
    from cutile.runtime import boundary, rt
    from cutile.runtime.kernel import default_stream
    from cutile.workloads import diffusion7, periodic_spec, sinusoid_init

    r = compile(diffusion7())
    spec = periodic_spec(3)
    u0 = sinusoid_init((24, 20, 40), r.halo_widths, seed=13)
    oracle.apply_boundary(u0, spec, r.halo_widths)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        a = to_dev(u0)
        b = torch.zeros_like(a)
        torch.cuda.synchronize()
        launch = r.launcher
        # warm up compile outside the capture
        launch(a, b)
        a.copy_(to_dev(u0))
        torch.cuda.synchronize()
        g = rt.Graph(default_stream())
        with g:
            for _ in range(3):
                launch(a, b)
                boundary.fill(b, spec, r.halo_widths)
                launch(b, a)
                boundary.fill(a, spec, r.halo_widths)
        ref = u0.copy()
        for rep in range(3):
            g.launch()
            torch.cuda.synchronize()
            for _ in range(6):
                ref = oracle.apply_stencil(ref, r._ref_fn, 3, r.halo_widths)
                oracle.apply_boundary(ref, spec, r.halo_widths)
            assert np.array_equal(to_host(a), ref), f"replay {rep}"
            b.zero_()  # eager work between replays
            torch.cuda.synchronize()


def test_compile_reports_resources_and_benchmark():
    from cutile import compile
    from cutile.workloads import diffusion7

    r = compile(diffusion7())
    mod = r.load_module()
    res = mod._KERNEL.resources("vec")
    assert res.local_bytes == 0, "sweep kernel spills to scratch"
    m = r.benchmark(domain=(64, 64, 64))
    assert m["gpoints_per_s"] > 0 and m["gbytes_per_s"] > 0


def test_prepared_launch_records_match_the_oracle(oracle):
    """``StencilKernel.prepare`` + ``run_prepared`` (the packed records the
    serving loop keeps per job): a sweep with periodic wrapped reads and a
    zero-Dirichlet temporal block of 4 steps, each replayed from its record
    several times, equal the oracle bit for bit."""
    import torch

    from cutile import compile
    from cutile.runtime.kernel import run_prepared
    from cutile.workloads import heat2d, heat2d_init, periodic_spec

    _flush("prepared records: heat2d 96x96 fp64, wrapped sweep x6 and depth-4 block x3")
    r = compile(heat2d(), temporal_blocking=False)
    kern = r.load_module()._KERNEL
    halo = r.halo_widths
    u0 = heat2d_init(96, halo, seed=3)
    a, b = to_dev(u0, torch.float64), torch.zeros(u0.shape, dtype=torch.float64, device="cuda")
    recs = [kern.prepare([a], [b], halo, periodic_axes=(0, 1)), kern.prepare([b], [a], halo, periodic_axes=(0, 1))]
    for s in range(6):
        run_prepared(recs[s % 2])
    torch.cuda.synchronize()
    ref = u0.copy()
    for _ in range(6):
        oracle.apply_boundary(ref, periodic_spec(2), halo)
        ref = oracle.apply_stencil(ref, r._ref_fn, 2, halo)
    inner = tuple(slice(h, h + 96) for h in halo)
    assert np.array_equal(to_host(a)[inner], ref[inner])

    blk = compile(heat2d(), temporal_steps=4, temporal_mode="zero")
    bk = blk.load_module()._KERNEL
    bh = blk.halo_widths
    v0 = heat2d_init(96, bh, seed=4)
    c, d = to_dev(v0, torch.float64), torch.zeros(v0.shape, dtype=torch.float64, device="cuda")
    brecs = [bk.prepare([c], [d], bh, temporal=True, mode=0), bk.prepare([d], [c], bh, temporal=True, mode=0)]
    for s in range(3):
        run_prepared(brecs[s % 2])
    torch.cuda.synchronize()
    ref = v0.copy()
    for _ in range(12):
        ref = oracle.apply_stencil(ref, blk._ref_fn, 2, bh)
    assert np.array_equal(to_host(d), ref)
