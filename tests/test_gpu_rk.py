"""Runge-Kutta stage algebra on the device (``-m gpu``): float64 RK steps
equal the upstream NumPy statements evaluated by the oracle, bit for bit."""

import numpy as np
import pytest

from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _rk_oracle(oracle, u, fn, ndim, halo, dt, method, steps):
    inner = oracle.interior(u.shape, halo)

    def L(x):
        k = np.zeros_like(x)
        k[inner] = oracle.apply_stencil(x, fn, ndim, halo)[inner]
        return k

    u = u.copy()
    for _ in range(steps):
        k1 = L(u)
        if method == "euler":
            u = u + dt * k1
        elif method == "rk2":
            u1 = u + 0.5 * dt * k1
            u = u + dt * L(u1)
        elif method == "rk4":
            u1 = u + 0.5 * dt * k1
            k2 = L(u1)
            u2 = u + 0.5 * dt * k2
            k3 = L(u2)
            u3 = u + dt * k3
            k4 = L(u3)
            u = u + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        else:
            u1 = u + dt * k1
            k2 = L(u1)
            u2 = 0.75 * u + 0.25 * (u1 + dt * k2)
            k3 = L(u2)
            u = (1.0 / 3.0) * u + (2.0 / 3.0) * (u2 + dt * k3)
    return u


@pytest.mark.parametrize("method", ["euler", "rk2", "rk4", "ssprk3"])
def test_rk_methods_match_numpy_statements(oracle, method):
    from cutile import RKIntegrator
    from cutile.frontend.decorator import stencil_from_source

    lap = stencil_from_source(
        "def lap(u, i, j):\n    return u[i-1, j] + u[i+1, j] + u[i, j-1] + u[i, j+1] - 4.0 * u[i, j]\n")
    res = RKIntegrator(lap, method).compile()
    mod = res.load_module()
    halo = res.stencil_result.halo_widths
    rng = np.random.default_rng(1)
    u0 = rng.random((40 + 2 * halo[0], 70 + 2 * halo[1]))
    dt = 0.1
    got = getattr(mod, f"rk_{method}_solve")(to_dev(u0), dt, 3)
    ref = _rk_oracle(oracle, u0, lap._fn, 2, halo, dt, method, 3)
    assert np.array_equal(to_host(got), ref)
