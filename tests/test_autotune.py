"""Autotuner: candidate generation, ranking and cache keys on the CPU; the
device search and its cache under ``-m gpu``."""

import numpy as np
import pytest

from helpers import to_dev, to_host
from cutile import workloads as w
from cutile.runtime import autotune as at
from cutile.lowering.stencil_to_hip import make_configs


@pytest.mark.parametrize("ndim,esz,loads", [(3, 8, 7), (3, 8, 25), (3, 4, 25), (2, 8, 5), (1, 4, 3)])
def test_candidates_are_launchable_and_default_first(ndim, esz, loads):
    from cutile.passes.tiling import TilingPass

    c = at.generate_candidates(ndim, esz, loads, temporal=(1, 2))
    first, hints = TilingPass(dtype_bytes=esz).candidates(ndim, loads)[0]
    assert c[0] == (tuple(first), 1, hints)
    storage = {8: "float64", 4: "float32"}[esz]
    for tile, T, h in c:
        vec, scal, xch = make_configs(ndim, tile, storage, threads=h.get("threads"))
        assert vec.nt % 64 == 0 and vec.nt <= 1024
        if ndim == 3:
            assert vec.nt // (vec.tz // vec.vz) <= vec.ty
    assert any(T == 2 for _, T, _ in c)
    p = at.prioritize(c, 6)
    assert p[0] == c[0] and len(p) == min(6, len(c))
    if len(c) > 6:
        assert all(T == 1 for _, T, _ in p)    # temporal candidates rank last


def test_cache_key_separates_what_changes_the_answer():
    f64, f32 = w.diffusion7(), w.diffusion7("float32")
    k = at.cache_key(f64, "float64", (512, 512, 512), (1,), "gfx950")
    assert k == at.cache_key(f64, "float64", (500, 520, 512), (1,), "gfx950")   # same size bucket
    assert k != at.cache_key(f64, "float64", (256, 512, 512), (1,), "gfx950")
    assert k != at.cache_key(f32, "float32", (512, 512, 512), (1,), "gfx950")
    assert k != at.cache_key(f64, "float64", (512, 512, 512), (1, 2), "gfx950")
    assert k != at.cache_key(f64, "float64", (512, 512, 512), (1,), "gfx942")
    assert at.size_bucket((500, 100, 3)) == (512, 128, 4)


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setenv("CUTILE_AMD_AUTOTUNE_CACHE", str(tmp_path))
    r = at.AutotuneResult((256, 16, 256), 1, 300.0, 5000.0, 0.4, {"threads": 512})
    at._save("abc", r)
    got = at._load("abc")
    assert got == r


def test_without_a_device_the_default_that_fits_is_returned(monkeypatch):
    import torch

    monkeypatch.setattr(torch.cuda, "is_available", lambda: False)
    r = at.autotune(w.diffusion7(), domain=(64, 64, 64))
    assert r.tile_sizes == (256, 16, 256) and r.temporal_steps == 1 and r.kernel_hints == {"threads": 512}


@pytest.mark.gpu
def test_device_search_caches_and_result_is_exact(oracle, tmp_path, monkeypatch):
    from cutile import compile
    from cutile.workloads import periodic_spec, sinusoid_init
    from cutile.runtime import boundary

    monkeypatch.setenv("CUTILE_AMD_AUTOTUNE_CACHE", str(tmp_path))
    print("autotune diffusion7 64^3 fp64, 8 candidates", flush=True)
    fn = w.diffusion7()
    r1 = at.autotune(fn, domain=(64, 64, 64), max_candidates=8, iters=5)
    assert r1.time_ms > 0 and r1.throughput_gpoints > 0
    assert list(tmp_path.iterdir()), "result not cached"
    r2 = at.autotune(fn, domain=(64, 64, 64), max_candidates=8, iters=5)
    assert r2 == r1
    res = compile(fn, domain=(64, 64, 64))
    assert res.tile_sizes == r1.tile_sizes
    u0 = sinusoid_init(64, res.halo_widths, seed=3)
    spec = periodic_spec(3)
    oracle.apply_boundary(u0, spec, res.halo_widths)
    a = to_dev(u0)
    b = a.clone()
    ref = u0.copy()
    for _ in range(3):
        res.launcher(a, b)
        boundary.fill(b, spec, res.halo_widths)
        a, b = b, a
        ref = oracle.apply_stencil(ref, res._ref_fn, 3, res.halo_widths)
        oracle.apply_boundary(ref, spec, res.halo_widths)
    assert float(np.max(np.abs(to_host(a) - ref))) == 0.0


@pytest.mark.gpu
def test_bucket_autotune_returns_one_bucket_per_class():
    print("autotune_buckets wave2d fp32", flush=True)
    sides = [70, 100, 150, 200, 300, 500, 900]
    bks = at.autotune_buckets(w.wave2d("float32"), sides, iters=3)
    assert len(bks) == 3 and bks[0].max_cols == 160 and bks[-1].max_cols is None
