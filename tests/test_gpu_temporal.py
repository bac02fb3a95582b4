"""In-kernel temporal blocking and the device time-marching loop (``-m gpu``).

Zero-halo level mode: one launch of depth T equals T chained sweeps through
zero-halo scratch arrays (the upstream temporal launcher's semantics).
Recompute mode on a T*r-deep periodic halo, and zero mode for zero
Dirichlet, equal T steps of (sweep + halo refill).  All comparisons are
bit-exact against the oracle."""

import numpy as np
import pytest

from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _chain_zero_halo(oracle, u, fn, ndim, halo, T):
    """T sweeps: the first reads u's halo, later ones read zero halos."""
    cur = u.copy()
    for t in range(T):
        nxt = oracle.apply_stencil(cur, fn, ndim, halo)
        if t < T - 1:
            z = np.zeros_like(nxt)
            inner = oracle.interior(nxt.shape, halo)
            z[inner] = nxt[inner]
            nxt = z
        cur = nxt
    return cur


ZERO_CASES = [
    ("heat2d", 4, None), ("diffusion7", 2, None), ("diffusion7", 3, None), ("star25", 2, None),
    # the per-wave level skip of the single-role kernel (option tbskip 1 and 2)
    ("heat2d", 4, {"tbskip": 1}), ("diffusion7", 3, {"tbskip": 1}), ("star25", 2, {"tbskip": 1, "tbws": 0}),
    ("heat2d", 4, {"tbskip": 2}), ("diffusion7", 3, {"tbskip": 2}),
    # the two-level kernel with wave roles (the 25-point star's default at
    # depth 2) against the single-role kernel, forced geometries, and the
    # 7-point star forced onto it
    ("star25", 2, {"tbws": 0}), ("star25", 2, {"tbws_tile": [24, 32, 640, 384]}),
    ("star25", 2, {"tbws_tile": [24, 32, 320, 192]}), ("diffusion7", 2, {"tbws": 1}),
    ("star25-float32", 2, None), ("star25-float32", 2, {"tbws_tile": [24, 64, 576, 384]}),
    # two cells per thread on adjacent rows (option tbya)
    ("star25", 2, {"tbws_tile": [24, 32, 320, 192], "tbya": 1}),
    ("star25-float32", 2, {"tbws_tile": [24, 64, 320, 192], "tbya": 1}),
]


@pytest.mark.parametrize("which,T,hints", ZERO_CASES,
                         ids=[f"{c[0]}-T{c[1]}-{'-'.join(f'{k}{v}' for k, v in (c[2] or {}).items()) or 'default'}"
                              .replace(" ", "").replace("[", "").replace("]", "").replace(",", "x")
                              for c in ZERO_CASES])
def test_zero_mode_equals_zero_halo_chain(oracle, which, T, hints):
    import torch

    from cutile import compile
    from cutile import workloads as w

    name, _, dt = which.partition("-")
    fn = getattr(w, name)(dt or "float64")
    r = compile(fn, temporal_steps=T, kernel_hints=hints)
    v = r.load_module()._VARIANTS
    key = "tblock" if "tblock" in v else "temporal"
    assert r.temporal_steps == T and key in v
    if hints and ("tbws_tile" in hints or hints.get("tbws")):
        assert "_tblock_tw" in v[key][0], v[key]
    if hints and hints.get("tbya"):
        assert v[key][0].endswith("y"), v[key]
    if hints and hints.get("tbws") == 0:
        assert "_tblock_tb" in v[key][0], v[key]
    print(f"{which} T={T} hints {hints} variant {v[key]}", flush=True)
    rng = np.random.default_rng(T)
    base = {2: (70, 150), 3: (23, 19, 45)}[r.ndim]
    shape = tuple(b + 2 * h for b, h in zip(base, r.halo_widths))
    u = rng.random(shape).astype(dt or "float64")      # the oracle computes in the array's precision
    du = to_dev(u)
    out = torch.zeros_like(du)
    r.launcher(du, out)
    ref = _chain_zero_halo(oracle, u, fn._fn, r.ndim, r.halo_widths, T)
    inner = oracle.interior(shape, r.halo_widths)
    got = to_host(out)
    assert np.array_equal(got[inner], ref[inner])


@pytest.mark.parametrize("which,T,steps,shape,hints", [
    ("star25", 2, 5, (24, 20, 36), None),
    ("star25", 4, 6, (20, 24, 40), None),
    ("diffusion7", 4, 9, (30, 18, 50), None),
    # blocks of the two-level kernel at a ragged shape (partial output tiles,
    # rows read around every periodic axis), forced geometries
    ("star25", 2, 7, (37, 45, 58), {"tbws_tile": [24, 32, 320, 192]}),
    ("star25", 2, 7, (37, 45, 58), {"tbws_tile": [16, 32, 512, 256]}),
    ("star25", 2, 7, (37, 45, 58), {"tbws_tile": [24, 32, 320, 192], "tbya": 1}),
    ("diffusion7", 2, 7, (21, 45, 58), {"tbws": 1}),
])
def test_periodic_recompute_march(oracle, which, T, steps, shape, hints):
    from cutile import workloads as w
    from cutile.runtime.timestep import TimeStepper

    fn = getattr(w, which)()
    spec = w.periodic_spec(3)
    ts = TimeStepper(fn, spec, temporal_steps=T, split="fit", block_hints=hints)
    assert ts.T == T and ts.mode == "recompute"
    halo = ts.halo
    print(f"{which} T={T} halo {halo} steps {steps}", flush=True)
    u0 = w.sinusoid_init(shape, halo, seed=3)
    oracle.apply_boundary(u0, spec, halo)
    got = to_host(ts.march(to_dev(u0), steps))
    ref = oracle.time_march(u0, fn._fn, 3, halo, steps, spec)
    assert np.array_equal(got, ref)


def test_dirichlet_zero_blocks_match_march(oracle):
    from cutile import workloads as w
    from cutile.runtime.timestep import TimeStepper

    fn = w.heat2d()
    spec = w.dirichlet_spec(2)
    ts = TimeStepper(fn, spec, temporal_steps=4, split="fit")
    assert ts.mode == "zero" and ts.T == 4
    u0 = w.heat2d_init(100, ts.halo, seed=0)
    got = to_host(ts.march(to_dev(u0), 11))
    ref = oracle.time_march(u0, fn._fn, 2, ts.halo, 11, spec)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("T", [1, 2])
def test_graph_march_matches_eager(oracle, T):
    from cutile import workloads as w
    from cutile.runtime.timestep import TimeStepper

    fn = w.diffusion7()
    spec = w.periodic_spec(3)
    ts = TimeStepper(fn, spec, temporal_steps=T, split="fit")
    u0 = w.sinusoid_init((16, 20, 24), ts.halo, seed=5)
    oracle.apply_boundary(u0, spec, ts.halo)
    eager = to_host(ts.march(to_dev(u0), 10))
    graphed = to_host(ts.march(to_dev(u0), 10, graph=True, graph_steps=4))
    ref = oracle.time_march(u0, fn._fn, 3, ts.halo, 10, spec)
    assert np.array_equal(eager, ref)
    assert np.array_equal(graphed, ref)


def test_graph_keeps_its_buffers_alive():
    """The captured step loop writes its scratch array on every replay: the
    stepper must hold it, or the caching allocator hands its memory to the
    next tensor of that size and replays overwrite that tensor."""
    import torch

    from cutile import workloads as w
    from cutile.runtime.timestep import TimeStepper

    ts = TimeStepper(w.diffusion7(), w.periodic_spec(3), temporal_steps=1)
    u0 = w.sinusoid_init((16, 20, 24), ts.halo, seed=5)
    u = to_dev(u0)
    ts.march(u, 4, graph=True, graph_steps=4)
    torch.cuda.synchronize()
    held = [torch.full_like(u, 7.0) for _ in range(4)]
    ts._graph.launch()
    torch.cuda.synchronize()
    assert all(bool((t == 7.0).all()) for t in held)


def test_nonzero_dirichlet_with_a_temporal_depth_matches_march(oracle):
    """A boundary value other than 0 cannot be carried through the zeroed
    ghost levels of a temporal block: the stepper steps one at a time and
    matches the per-step refill definition."""
    from cutile import workloads as w
    from cutile.runtime.timestep import TimeStepper

    fn = w.heat2d()
    spec = w.dirichlet_spec(2, value=1.5)
    ts = TimeStepper(fn, spec, temporal_steps=4, split="fit")
    u0 = w.heat2d_init(64, ts.halo, seed=3)
    oracle.apply_boundary(u0, spec, ts.halo)
    got = to_host(ts.march(to_dev(u0), 7))
    ref = oracle.time_march(u0, fn._fn, 2, ts.halo, 7, spec)
    assert np.array_equal(got, ref)


def test_split_deep_block_matches_march(oracle):
    """Depth 4 of the radius-4 star runs as two depth-2 launches per block
    (periodic, wrapped): equal to 8 single periodic steps."""
    from cutile import workloads as w
    from cutile.runtime.timestep import TimeStepper

    fn = w.star25()
    spec = w.periodic_spec(3)
    ts = TimeStepper(fn, spec, temporal_steps=4, split="fit")
    assert ts.block_reps == 2
    print(f"star25 depth 4 as 2 x 2, halo {ts.halo}", flush=True)
    u0 = w.sinusoid_init((20, 24, 40), ts.halo, seed=4)
    oracle.apply_boundary(u0, spec, ts.halo)
    got = to_host(ts.march(to_dev(u0), 8))
    ref = oracle.time_march(u0, fn._fn, 3, ts.halo, 8, spec)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("T", [2, 4])
@pytest.mark.parametrize("bc", ["periodic", "dirichlet0"])
@pytest.mark.parametrize("n", [(160, 256), (97, 262)])
def test_fused_blocks_match_single_steps(oracle, dtype, T, bc, n):
    """Fused Gray-Scott through the coupled-field temporal kernel (blocks of
    T steps, then single steps for the remainder) equals single fused steps
    bit for bit; float64 also equals the oracle."""
    import torch

    from cutile import compile_fused
    from cutile import workloads as w
    from cutile.runtime.launcher import compile_fused_block
    from cutile.runtime.timestep import FusedStepper

    print(f"fused blocks: gray-scott {dtype} {n} T={T} {bc}", flush=True)
    fns = list(w.gray_scott(dtype))
    one = compile_fused(fns)
    blk = compile_fused_block(fns, T, one, ["u", "v"])
    h = one.halo_widths
    spec = w.periodic_spec(2) if bc == "periodic" else w.dirichlet_spec(2)
    shape = tuple(a + 2 * b for a, b in zip(n, h))
    rng = np.random.default_rng(T)
    init = {k: rng.random(shape) for k in ("u", "v")}
    if bc == "dirichlet0":
        for v in init.values():
            oracle.apply_boundary(v, spec, h)
    steps = 2 * T + 1
    res = []
    for block in (None, blk):
        dt = getattr(torch, dtype)
        f = {k: to_dev(v).to(dt) for k, v in init.items()}
        sc = {k: torch.empty_like(f[k]) for k in f}
        from cutile.runtime import boundary
        for k in f:
            boundary.fill(f[k], spec, h)
        st = FusedStepper(one, spec, ["u", "v"], block=block)
        assert (st._blk is not None) == (block is not None)
        st.enqueue(f, sc, steps)
        torch.cuda.synchronize()
        res.append({k: to_host(f[k].to(torch.float64)) for k in f})
    inner = oracle.interior(shape, h)
    for k in ("u", "v"):
        assert np.array_equal(res[0][k][inner], res[1][k][inner]), k
    if dtype == "float64":
        fp = {k: v.copy() for k, v in init.items()}
        for v in fp.values():
            oracle.apply_boundary(v, spec, h)
        ref = oracle.time_march_multi(fp, [g._fn for g in fns], ["u", "v"], 2, h, steps, spec)
        for k in ("u", "v"):
            assert np.array_equal(res[1][k][inner], ref[k][inner]), k
