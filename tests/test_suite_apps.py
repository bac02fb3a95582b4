"""Benchmark-suite and application-driver tables (CPU: no device).

The suite carries the upstream ``run_benchmarks.py`` stencils and domains
(``run_benchmarks.py:87-143``) and every one of them compiles for gfx950;
the application drivers' stencils compile with the launcher signatures the
drivers call."""

import pytest

from cutile import apps, compile, compile_fused, suite
from cutile.runtime import rt
from cutile.runtime.kernel import compile_options


def test_suite_table_matches_upstream():
    assert list(suite.DOMAINS) == ["heat_1d", "heat_2d", "lap_2d_5pt", "lap_1d_4th", "wave_2d_4th", "lap_3d_7pt",
                                   "advect_1d"]
    assert suite.DOMAINS["lap_3d_7pt"] == [(32, 32, 32), (64, 64, 64), (128, 128, 128)]
    assert all(len(v) == 4 for k, v in suite.DOMAINS.items() if k != "lap_3d_7pt")
    st = suite.stencils()
    assert {k: (v[1], v[2]) for k, v in st.items()} == {
        "heat_1d": (1, 2), "heat_2d": (2, 2), "lap_2d_5pt": (2, 2), "lap_1d_4th": (1, 4), "wave_2d_4th": (2, 4),
        "lap_3d_7pt": (3, 2), "advect_1d": (1, 2)}


@pytest.mark.skipif(not rt.available(), reason="needs libstencil_rt.so (hiprtc)")
def test_suite_stencils_build_for_gfx950():
    for name, (fn, nd, order) in suite.stencils().items():
        r = compile(fn, temporal_blocking=False)
        assert r.stencil_halo == (order // 2,) * nd, name
        rt.compile_only(r.hip_source, compile_options("gfx950"))


def test_app_launchers():
    fe, fh = apps.fdtd_stencils()
    assert "def launch_update_H(E, H, u_out, stream=None):" in compile(fh, temporal_blocking=False).code
    r = compile_fused(list(apps.shallow_water_stencils()))
    assert r.inputs == ("h", "hu", "hv") and len(r.outputs) == 3
    r = compile_fused(list(apps.gray_scott_stencils()))
    assert r.inputs == ("u", "v")
    assert set(apps.APPS) == {"heat_1d", "advection_upwind", "fdtd_maxwell_1d", "shallow_water", "wave_2d",
                              "gray_scott", "laplacian_3d", "heat_2d_bricked"}


def test_initial_conditions_follow_the_examples():
    import numpy as np

    u = apps.heat_1d_init(128, 16)
    assert u.shape == (160,) and u[0] == 0.0 and u[-1] == 0.0 and abs(u[80] - 1.0) < 1e-3
    a = apps.advection_init(200, 16)
    assert a.sum() == 50.0 and a[16 + 50] == 1.0 and a[16 + 100] == 0.0
    h = apps.bump_2d((34, 64), 3.0, 0.1)
    assert h[17, 32] == 0.1 and h.shape == (34, 64)
    gu, gv = apps.gray_scott_init((66, 96))
    assert gu[33, 48] == 0.5 and gv[33, 48] == 0.25 and gu[0, 0] == 1.0 and gv[0, 0] == 0.0
    assert np.isclose(apps.fdtd_source(40), 1.0)
