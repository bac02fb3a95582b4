"""Sweeps and temporal levels whose z-neighbours are read across lanes (DPP
wave shifts, ``kernel_hints={"zdpp": 1}``) instead of from the LDS plane:
bit-exact against the oracle on ragged shapes (tile-row and wave edges fall
back to the LDS halo), on random multi-field stencils, on wrapped periodic
steps, in float64 and float32, and for depth-T temporal launches."""

import numpy as np
import pytest

from helpers import random_stencil, to_dev, to_host

pytestmark = pytest.mark.gpu


def _one(oracle, fn, n, seed):
    import torch

    from cutile import compile

    r = compile(fn, temporal_blocking=False, kernel_hints={"zdpp": 1})
    assert r.load_module()._KERNEL.variants["vec"][0].endswith("z")
    print(f"zdpp {r.name} {r.storage} {n} halo {r.halo_widths}", flush=True)
    shape = tuple(a + 2 * h for a, h in zip(n, r.halo_widths))
    rng = np.random.default_rng(seed)
    npdt = np.float32 if r.storage == "float32" else np.float64
    arrs = {f: (rng.random(shape) + 0.5).astype(npdt) for f in r.inputs}
    dev = [to_dev(arrs[f]) for f in r.inputs]
    out = torch.zeros_like(dev[0])
    r.launcher(*dev, out)
    ref = oracle.eval_multi(arrs, fn._fn, len(n), r.halo_widths)
    got = to_host(out)[oracle.interior(shape, r.halo_widths)]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("which,dtype,n", [
    ("star25", "float64", (9, 45, 70)), ("star25", "float64", (5, 33, 200)),
    ("star25", "float32", (7, 19, 200)), ("diffusion7", "float64", (6, 33, 300)),
    ("diffusion7", "float32", (5, 17, 270)), ("heat2d", "float64", (70, 700)),
])
def test_zdpp_workloads_match_oracle(oracle, which, dtype, n):
    from cutile import workloads as w

    _one(oracle, getattr(w, which)(dtype), n, len(n) * 11 + n[-1])


@pytest.mark.parametrize("seed", [2, 3, 10, 30, 41, 57])
def test_zdpp_random_stencils_match_oracle(oracle, seed):
    shapes = [(1, 1, 1), (1, 2, 2), (2, 1, 3), (2, 2, 2), (2, 4, 4), (3, 1, 2), (3, 2, 1), (3, 3, 3),
              (3, 4, 2), (3, 1, 4)]
    nd, nf, rad = shapes[seed % len(shapes)]
    fn, _, _ = random_stencil(seed, nd, nf, rad, 12)
    _one(oracle, fn, {1: (300,), 2: (37, 70), 3: (11, 13, 40)}[nd], seed)


def test_zdpp_wrapped_steps_match_oracle(oracle):
    import torch

    from cutile import compile
    from cutile import workloads as w

    fn = w.star25()
    r = compile(fn, temporal_blocking=False, kernel_hints={"zdpp": 1})
    K = r.load_module()._KERNEL
    n = (12, 40, 136)
    print(f"zdpp wrapped star25 {n} x3", flush=True)
    h = r.halo_widths
    shape = tuple(a + 2 * b for a, b in zip(n, h))
    u = np.random.default_rng(5).random(shape)
    spec = w.periodic_spec(3)
    oracle.apply_boundary(u, spec, h)
    a, b = to_dev(u), torch.zeros(shape, dtype=torch.float64, device="cuda")
    ref = u
    for _ in range(3):
        K.launch([a], [b], h, None, periodic_axes=(0, 1, 2))
        a, b = b, a
        ref = oracle.apply_stencil(ref, fn._fn, 3, h)
        oracle.apply_boundary(ref, spec, h)
    I = oracle.interior(shape, h)
    assert np.array_equal(to_host(a)[I], ref[I])


@pytest.mark.parametrize("which,T,n", [("star25", 2, (23, 19, 45)), ("diffusion7", 3, (23, 19, 45)),
                                       ("heat2d", 4, (70, 150)), ("heat2d", 3, (33, 47))])
def test_zdpp_temporal_levels_match_oracle(oracle, which, T, n):
    """The temporal kernel's levels with z-neighbours read across lanes
    (level L-1's queue vectors of the neighbour lanes; region-row and wave
    edges fall back to the level's LDS plane): one depth-T launch equals T
    chained zero-halo sweeps, bit for bit."""
    import torch

    from cutile import compile
    from cutile import workloads as w
    from test_gpu_temporal import _chain_zero_halo

    fn = getattr(w, which)()
    r = compile(fn, temporal_steps=T, kernel_hints={"zdpp": 1})
    v = r.load_module()._VARIANTS
    assert "tblock" in v and "ZDPP = 1" in r.hip_source
    print(f"zdpp temporal {which} T={T} {n} variant {v['tblock']}", flush=True)
    shape = tuple(b + 2 * h for b, h in zip(n, r.halo_widths))
    u = np.random.default_rng(T + len(n)).random(shape)
    du = to_dev(u)
    out = torch.zeros_like(du)
    r.launcher(du, out)
    ref = _chain_zero_halo(oracle, u, fn._fn, r.ndim, r.halo_widths, T)
    inner = oracle.interior(shape, r.halo_widths)
    assert np.array_equal(to_host(out)[inner], ref[inner])
