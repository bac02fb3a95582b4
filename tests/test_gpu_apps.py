"""The upstream example applications on the HIP path (``-m gpu``).

Each driver of ``cutile.apps`` runs its example's time loop on the device;
the oracle runs the same loop as the example writes it (copy, then the
interior from the stencil, H then E, the source term, the whole-array
leapfrog and the zeroed outer ring), from the same initial arrays.  All are
float64 and compared bit for bit."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_heat_1d(oracle):
    from cutile import apps

    _flush("app heat_1d 128 x 100")
    out = apps.heat_1d(128, 100)
    h = 16
    u0 = apps.heat_1d_init(128, h)
    ref = oracle.time_march(u0, apps.heat_1d_stencil()._fn, 1, (h,), 100)
    assert np.array_equal(out["fields"]["u"], ref)
    rep = out["report"]
    assert rep["energy_final"] < rep["energy_initial"]


def test_advection_upwind(oracle):
    from cutile import apps

    _flush("app advection_upwind 200 x 80")
    out = apps.advection_upwind(200, 80)
    h = 16
    ref = oracle.time_march(apps.advection_init(200, h), apps.advection_stencil()._fn, 1, (h,), 80)
    assert np.array_equal(out["fields"]["u"], ref)
    rep = out["report"]
    assert rep["centre_final"] > rep["centre_initial"]       # moves right (c > 0)


def test_fdtd_maxwell_1d(oracle):
    from cutile import apps

    _flush("app fdtd_maxwell_1d 500 x 200")
    out = apps.fdtd_maxwell_1d(500, 200)
    fe, fh = apps.fdtd_stencils()
    h = 16
    E = np.zeros(500 + 2 * h)
    H = np.zeros(500 + 2 * h)
    inner = oracle.interior(E.shape, (h,))
    for s in range(200):
        Hn = H.copy()
        Hn[inner] = oracle.eval_multi({"E": E, "H": H}, fh._fn, 1, (h,))
        H = Hn
        En = E.copy()
        En[inner] = oracle.eval_multi({"E": E, "H": H}, fe._fn, 1, (h,))
        E = En
        E[500 // 4 + h] += np.exp(-0.5 * ((s - 40) / 12) ** 2)
    assert np.array_equal(out["fields"]["E"], E)
    assert np.array_equal(out["fields"]["H"], H)
    assert out["report"]["max_abs_E"] > 0.1


def test_shallow_water(oracle):
    from cutile import apps

    _flush("app shallow_water 32x32 x 50")
    out = apps.shallow_water(32, 50)
    fns = [f._fn for f in apps.shallow_water_stencils()]
    halo = (1, 16)
    shape = (32 + 2, 32 + 32)
    f = {"h": apps.bump_2d(shape, 3.0, 0.1), "hu": np.zeros(shape), "hv": np.zeros(shape)}
    for _ in range(50):
        new = oracle.apply_multi(f, fns, 2, halo, base=[f["h"], f["hu"], f["hv"]])
        f = dict(zip(("h", "hu", "hv"), new))
    for k in ("h", "hu", "hv"):
        assert np.array_equal(out["fields"][k], f[k]), k
    assert np.max(np.abs(out["fields"]["h"])) < 10.0


def test_wave_2d(oracle):
    from cutile import apps

    _flush("app wave_2d 64x64 x 50")
    out = apps.wave_2d(64, 50)
    fn = apps.wave_2d_stencil()._fn
    halo = (2, 16)
    u = apps.bump_2d((64 + 4, 64 + 32), 5.0)
    u_old = u.copy()
    for _ in range(50):
        lap = oracle.apply_stencil(u, fn, 2, halo)
        u_new = 2 * u - u_old + lap
        u_new[0, :] = 0
        u_new[-1, :] = 0
        u_new[:, 0] = 0
        u_new[:, -1] = 0
        u_old, u = u, u_new
    assert np.array_equal(out["fields"]["u"], u)


def test_gray_scott_example(oracle):
    from cutile import apps

    _flush("app gray_scott 64x64 x 100")
    out = apps.gray_scott(64, 100)
    fns = [f._fn for f in apps.gray_scott_stencils()]
    halo = (1, 16)
    u, v = apps.gray_scott_init((64 + 2, 64 + 32))
    for _ in range(100):
        u, v = oracle.apply_multi({"u": u, "v": v}, fns, 2, halo, base=[u, v])
    assert np.array_equal(out["fields"]["u"], u)
    assert np.array_equal(out["fields"]["v"], v)


def test_laplacian_3d_impulse(oracle):
    from cutile import apps

    _flush("app laplacian_3d 16^3")
    out = apps.laplacian_3d(16)
    assert out["report"]["centre"] == -6.0
    halo = (1, 1, 16)
    u0 = np.zeros(tuple(16 + 2 * h for h in halo))
    u0[tuple(8 + h for h in halo)] = 1.0
    ref = oracle.apply_stencil(u0, apps.laplacian_3d_stencil()._fn, 3, halo)
    inner = oracle.interior(u0.shape, halo)
    assert np.array_equal(out["fields"]["lap"][inner], ref[inner])


def test_heat_2d_bricked(oracle):
    from cutile import apps

    _flush("app heat_2d_bricked 128x128")
    out = apps.heat_2d_bricked(128)
    assert out["report"]["bricked_equals_flat"] and out["report"]["round_trip_exact"]
    halo = (1, 16)
    shape = (130, 160)
    cx, cy = shape[0] / 2.0, shape[1] / 2.0
    ii = np.arange(shape[0], dtype=np.float64)[:, None]
    jj = np.arange(shape[1], dtype=np.float64)[None, :]
    u0 = np.exp(-(((ii - cx) / 10.0) ** 2 + ((jj - cy) / 10.0) ** 2))
    ref = oracle.apply_stencil(u0, apps.heat_2d_stencil()._fn, 2, halo)
    inner = oracle.interior(shape, halo)
    assert np.array_equal(out["fields"]["flat"][inner], ref[inner])


def test_suite_rows_and_one_step_parity(oracle):
    """The suite's report rows, and one launch of each of its stencils equals
    the oracle's sweep on the smallest domain."""
    import torch

    from cutile import compile, suite

    _flush("suite: smallest domain of each stencil")
    res = suite.run(max_domains=1, warmup=2, iters=5, log=lambda *a: None)
    assert len(res["results"]) == 7 and res["peak_bw_gbs"] > 1000
    for row in res["results"]:
        assert set(row) == {"stencil", "domain", "tiles", "cutile_gps", "cutile_ms", "eff_bw_pct"}
        assert row["cutile_ms"] > 0
    rng = np.random.default_rng(0)
    for name, (fn, nd, order) in suite.stencils().items():
        r = compile(fn, temporal_blocking=False)
        shape = tuple(d + 2 * h for d, h in zip(suite.DOMAINS[name][0], r.halo_widths))
        u = rng.standard_normal(shape)
        a = torch.from_numpy(u).cuda()
        b = torch.zeros_like(a)
        getattr(r.load_module(), f"launch_{r.name}")(a, b)
        ref = oracle.apply_stencil(u, fn._fn, nd, r.halo_widths)
        inner = oracle.interior(shape, r.halo_widths)
        assert np.array_equal(b.cpu().numpy()[inner], ref[inner]), name
