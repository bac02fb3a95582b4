"""Pins the CPU oracle (CPU only).

Known answers from the upstream tests (``tests/test_cutile_new.py:780-847``,
``examples/laplacian_3d.py:54-57``), an independent scalar brute-force
evaluator, closed-form boundary fills, and the committed golden vectors.
A plausible slip -- a dropped term, a wrong sign or offset, a transposed
axis, a wrong halo side -- fails at least one of these."""

import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _fn(src, **env):
    ns = dict(env)
    exec(src, ns)
    return next(v for k, v in ns.items() if callable(v) and not k.startswith("__") and k not in env)


class TestKnownAnswers:
    def test_1d_heat_smoothing(self, oracle):
        f = _fn("def heat(u, i):\n    return 0.25 * u[i - 1] + 0.5 * u[i] + 0.25 * u[i + 1]\n")
        u = np.zeros(10)
        u[5] = 1.0
        r = oracle.apply_stencil(u, f, 1, (1,))
        assert r[5] == 0.5 and r[4] == 0.25 and r[6] == 0.25 and r[3] == 0.0

    def test_2d_laplacian_of_constant_is_zero(self, oracle):
        f = _fn("def lap(u, i, j):\n    return u[i-1,j] + u[i+1,j] + u[i,j-1] + u[i,j+1] - 4 * u[i,j]\n")
        r = oracle.apply_stencil(np.ones((10, 10)), f, 2, (1, 1))
        assert np.allclose(r[1:-1, 1:-1], 0.0, atol=1e-14)

    def test_3d_impulse_response(self, oracle):
        f = _fn("def lap(u, i, j, k):\n    return (u[i-1,j,k] + u[i+1,j,k] + u[i,j-1,k] + u[i,j+1,k]"
                " + u[i,j,k-1] + u[i,j,k+1] - 6 * u[i,j,k])\n")
        u = np.zeros((7, 7, 7))
        u[3, 3, 3] = 1.0
        r = oracle.apply_stencil(u, f, 3, (1, 1, 1))
        assert r[3, 3, 3] == -6.0
        for d in [(2, 3, 3), (4, 3, 3), (3, 2, 3), (3, 4, 3), (3, 3, 2), (3, 3, 4)]:
            assert r[d] == 1.0
        assert r[2, 2, 3] == 0.0

    def test_identity_and_halo_untouched(self, oracle):
        f = _fn("def ident(u, i):\n    return u[i]\n")
        u = np.arange(10, dtype=float)
        r = oracle.apply_stencil(u, f, 1, (1,))
        np.testing.assert_array_equal(r, u)

    def test_averaging_of_constant(self, oracle):
        f = _fn("def avg(u, i, j):\n    return 0.25 * (u[i-1,j] + u[i+1,j] + u[i,j-1] + u[i,j+1])\n")
        r = oracle.apply_stencil(5.0 * np.ones((8, 8)), f, 2, (1, 1))
        assert np.allclose(r[1:-1, 1:-1], 5.0, atol=1e-14)

    def test_time_march_is_monotone(self, oracle):
        f = _fn("def heat(u, i):\n    return 0.25 * u[i - 1] + 0.5 * u[i] + 0.25 * u[i + 1]\n")
        u0 = np.zeros(20)
        u0[10] = 1.0
        prev = [1.0]
        oracle.time_march(u0, f, 1, (1,), 5, callback=lambda s, u: (prev.append(u.max())))
        assert all(b <= a + 1e-14 for a, b in zip(prev, prev[1:]))
        assert len(prev) == 6

    def test_offsets_are_per_axis(self, oracle):
        """u[i+1, j] reads the next row, u[i, j+1] the next column."""
        f = _fn("def sh(u, i, j):\n    return 10.0 * u[i + 1, j] + u[i, j + 1]\n")
        u = np.arange(30, dtype=float).reshape(5, 6)
        r = oracle.apply_stencil(u, f, 2, (1, 1))
        assert r[1, 1] == 10.0 * u[2, 1] + u[1, 2]

    def test_offset_beyond_halo_is_rejected(self, oracle):
        f = _fn("def far(u, i):\n    return u[i + 2]\n")
        with pytest.raises(IndexError):
            oracle.apply_stencil(np.zeros(10), f, 1, (1,))


CASES = [
    ("def s(u, i, j):\n    return 0.25 * (u[i-1,j] + u[i+1,j] + u[i,j-1] + u[i,j+1])\n", 2, (1, 1)),
    ("def s(u, i, j, k):\n    return u[i,j,k] + 0.1 * (u[i-1,j,k] + u[i+1,j,k] + u[i,j-1,k] + u[i,j+1,k]"
     " + u[i,j,k-1] + u[i,j,k+1] - 6 * u[i,j,k])\n", 3, (1, 1, 1)),
    ("def s(u, i):\n    return max(u[i-2], u[i+1]) - min(u[i], 0.3) * abs(u[i-1] - u[i+2])\n", 1, (2,)),
    ("def s(u, i, j):\n    return where(u[i,j] > u[i+1,j-1], u[i-1,j+1], 0.1 * u[i,j]) / (u[i,j+1] + 1.5)\n",
     2, (1, 1)),
    ("def s(u, i, j, k):\n    return -u[i-2,j+1,k-3] + 7 * u[i+2,j,k+3] * u[i,j-1,k]\n", 3, (2, 1, 3)),
]


@pytest.mark.parametrize("src,nd,halo", CASES)
def test_apply_stencil_matches_scalar_brute_force(oracle, src, nd, halo):
    f = _fn(src, where=oracle.where)
    rng = np.random.default_rng(0)
    u = rng.random(tuple(s + 2 * h for s, h in zip([6, 7, 9][:nd], halo))) + 0.5
    a = oracle.apply_stencil(u, f, nd, halo)
    b = oracle.apply_stencil_scalar(u, f, nd, halo)
    assert np.array_equal(a, b)


class TestBoundary:
    def test_periodic_1d_closed_form(self, oracle):
        u = np.arange(8, dtype=float)  # halo 2: interior 2..5
        oracle.apply_boundary(u, {0: ("periodic", "periodic")}, (2,))
        np.testing.assert_array_equal(u, [4, 5, 2, 3, 4, 5, 2, 3])

    def test_neumann_1d_closed_form(self, oracle):
        u = np.arange(8, dtype=float)
        oracle.apply_boundary(u, {0: ("neumann", "neumann")}, (2,))
        np.testing.assert_array_equal(u, [2, 2, 2, 3, 4, 5, 5, 5])

    def test_reflecting_1d_closed_form(self, oracle):
        u = np.arange(8, dtype=float)
        oracle.apply_boundary(u, {0: ("reflecting", "reflecting")}, (2,))
        np.testing.assert_array_equal(u, [3, 2, 2, 3, 4, 5, 5, 4])

    def test_dirichlet_value_and_corner_order(self, oracle):
        u = np.ones((4, 4))
        oracle.apply_boundary(u, {0: (("dirichlet", 2.0), ("dirichlet", 3.0)),
                                  1: (("dirichlet", 5.0), "neumann")}, (1, 1))
        # dim 1 is applied after dim 0: it owns the corners
        assert u[0, 0] == 5.0 and u[3, 0] == 5.0
        assert u[0, 3] == u[0, 2] == 2.0 and u[3, 3] == u[3, 2] == 3.0
        assert u[1, 3] == 1.0 and u[0, 1] == 2.0

    def test_periodic_2d_wraps_corners(self, oracle):
        u = np.arange(36, dtype=float).reshape(6, 6)
        ref = u.copy()
        oracle.apply_boundary(u, {0: ("periodic", "periodic"), 1: ("periodic", "periodic")}, (1, 1))
        inner = ref[1:5, 1:5]
        tiled = np.tile(inner, (3, 3))[3:9, 3:9]
        np.testing.assert_array_equal(u, tiled)


def test_golden_vectors(oracle):
    """Fixtures written by tests/golden/make_golden.py: inputs and outputs
    after several steps of the benchmark stencils at small sizes."""
    from cutile import workloads as w

    path = os.path.join(GOLDEN, "golden.npz")
    g = np.load(path)
    spec2d = w.dirichlet_spec(2)
    f = w.heat2d()
    out = oracle.time_march(g["heat_in"], f._fn, 2, (1, 1), 10, spec2d)
    assert np.array_equal(out, g["heat_out"])
    f = w.star25()
    out = oracle.time_march(g["star_in"], f._fn, 3, (4, 4, 4), 2, w.periodic_spec(3))
    assert np.array_equal(out, g["star_out"])
    gu, gv = w.gray_scott()
    res = oracle.time_march_multi({"u": g["gs_u_in"], "v": g["gs_v_in"]}, [gu._fn, gv._fn], ["u", "v"], 2,
                                  (1, 1), 5, w.periodic_spec(2))
    assert np.array_equal(res["u"], g["gs_u_out"]) and np.array_equal(res["v"], g["gs_v_out"])
    # independent spot check of the fixture itself: one heat step by hand
    u = g["heat_in"]
    c = 0.25 * (u[0, 1] + u[2, 1] + u[1, 0] + u[1, 2])
    one = oracle.time_march(u, w.heat2d()._fn, 2, (1, 1), 1, spec2d)
    assert one[1, 1] == c


class TestMultiFieldPins:
    """Pins of the multi-input / fused / two-level paths: closed forms on
    uniform states and a scalar loop over every interior cell."""

    def test_unread_parameter_is_zero(self, oracle):
        f = _fn("def s(u, w, i):\n    return u[i] + 2.0 * w[i]\n")
        u = np.arange(6, dtype=float)
        got = oracle.eval_multi({"u": u}, f, 1, (1,))
        np.testing.assert_array_equal(got, u[1:5])

    def test_gray_scott_uniform_closed_form(self, oracle):
        from cutile import workloads as w

        gu, gv = w.gray_scott()
        GS = w.GS
        u0, v0 = 0.7, 0.2
        u = np.full((6, 6), u0)
        v = np.full((6, 6), v0)
        res = oracle.time_march_multi({"u": u, "v": v}, [gu._fn, gv._fn], ["u", "v"], 2, (1, 1), 1,
                                      w.periodic_spec(2))
        # the Laplacian of a uniform field is 0: only the reaction terms act,
        # both evaluated from the old state
        eu = u0 + GS["dt"] * (GS["Du"] * (u0 + u0 + u0 + u0 - 4 * u0) - u0 * v0 * v0 + GS["F"] * (1.0 - u0))
        ev = v0 + GS["dt"] * (GS["Dv"] * (v0 + v0 + v0 + v0 - 4 * v0) + u0 * v0 * v0 - (GS["F"] + GS["k"]) * v0)
        assert np.all(res["u"] == eu) and np.all(res["v"] == ev)
        # u = 1, v = 0 is a fixed point
        res = oracle.time_march_multi({"u": np.ones((6, 6)), "v": np.zeros((6, 6))}, [gu._fn, gv._fn],
                                      ["u", "v"], 2, (1, 1), 3, w.periodic_spec(2))
        assert np.all(res["u"] == 1.0) and np.all(res["v"] == 0.0)

    def test_gray_scott_step_scalar_loop(self, oracle):
        from cutile import workloads as w

        gu, gv = w.gray_scott()
        GS = w.GS
        rng = np.random.default_rng(5)
        u = rng.random((7, 8))
        v = rng.random((7, 8))
        res = oracle.time_march_multi({"u": u, "v": v}, [gu._fn, gv._fn], ["u", "v"], 2, (1, 1), 1, None)
        for i in range(1, 6):
            for j in range(1, 7):
                lu = u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1] - 4 * u[i, j]
                lv = v[i - 1, j] + v[i + 1, j] + v[i, j - 1] + v[i, j + 1] - 4 * v[i, j]
                eu = u[i, j] + GS["dt"] * (GS["Du"] * lu - u[i, j] * v[i, j] * v[i, j] + GS["F"] * (1.0 - u[i, j]))
                ev = v[i, j] + GS["dt"] * (GS["Dv"] * lv + u[i, j] * v[i, j] * v[i, j] - (GS["F"] + GS["k"]) * v[i, j])
                assert res["u"][i, j] == eu and res["v"][i, j] == ev
        # halos are the old values (no BC given): the device writes only interiors
        assert res["u"][0, 3] == u[0, 3]

    def test_leapfrog_closed_forms_and_scalar_loop(self, oracle):
        from cutile import workloads as w

        f = w.wave2d("float64")
        # uniform levels under a periodic refill: L = 0, so u_new = 2u - uold
        # (a straight line in time)
        u, uold = np.full((5, 5), 3.0), np.full((5, 5), 1.0)
        a, b = oracle.leapfrog(u, uold, f._fn, 2, (1, 1), 2, w.periodic_spec(2))
        assert np.all(a[1:4, 1:4] == 7.0) and np.all(b[1:4, 1:4] == 5.0)
        rng = np.random.default_rng(9)
        u, uold = rng.random((6, 7)), rng.random((6, 7))
        a, b = oracle.leapfrog(u, uold, f._fn, 2, (1, 1), 1, w.dirichlet_spec(2))
        for i in range(1, 5):
            for j in range(1, 6):
                e = 2.0 * u[i, j] - uold[i, j] + 0.25 * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1]
                                                          - 4.0 * u[i, j])
                assert a[i, j] == e
        assert np.all(a[0] == 0.0) and np.array_equal(b, u)

    def test_rel_max_err(self, oracle):
        assert oracle.rel_max_err(np.array([1.0, 2.0]), np.array([1.0, 4.0])) == 0.5
        assert oracle.rel_max_err(np.zeros(3), np.zeros(3)) == 0.0


class TestTimeUpdatePins:
    def test_bf16_round_matches_torch(self, oracle):
        import torch

        rng = np.random.default_rng(1)
        x = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 5000),
                            np.float32([0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, np.inf, -np.inf, 3e38])])
        x = x.astype(np.float32)
        ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
        np.testing.assert_array_equal(oracle.bf16_round(x), ref)
        assert np.isnan(oracle.bf16_round(np.float32([np.nan]))).all()

    def test_euler_update_equals_the_full_form_stencil(self, oracle):
        from cutile import workloads as w

        rng = np.random.default_rng(3)
        u, v = rng.random((10, 12)), rng.random((10, 12))
        full = oracle.time_march_multi({"u": u, "v": v}, [f._fn for f in w.gray_scott()], ["u", "v"], 2,
                                       (1, 1), 3, w.periodic_spec(2))
        f = {"u": u.copy(), "v": v.copy()}
        for _ in range(3):
            oracle.time_update(f, [r._fn for r in w.gray_scott_rates()], ["u", "v"], 2, (1, 1), w.GS["dt"])
            for n in ("u", "v"):
                oracle.apply_boundary(f[n], w.periodic_spec(2), (1, 1))
        assert np.array_equal(f["u"], full["u"]) and np.array_equal(f["v"], full["v"])

    def test_kahan_carries_what_bf16_drops(self, oracle):
        # a constant rate of 2^-10 on u = 1: bf16 (8-bit mantissa) drops every
        # increment, the compensated sum keeps u - c equal to the exact sum
        f = _fn("def r(u, i):\n    return u[i] * 0.0 + 0.0009765625\n")
        u = np.ones(6, dtype=np.float32)
        c = np.zeros(6, dtype=np.float32)
        plain = np.ones(6, dtype=np.float32)
        for n in range(1, 40):
            oracle.time_update({"u": u}, [f], ["u"], 1, (1,), 1.0, "kahan", {"u": c}, "bfloat16")
            oracle.time_update({"u": plain}, [f], ["u"], 1, (1,), 1.0, "euler", None, "bfloat16")
            assert np.all(u[1:5] - c[1:5] == np.float32(1.0 + n * 2 ** -10))
            assert np.all(plain[1:5] == 1.0)
