"""The benchmark workloads: stencil definitions and seeded initial fields.

Each workload is the stencil of one benchmark configuration plus a
deterministic generator of its input.  Generators return NumPy arrays with
the full (padded) halo so the same data feeds the device and any checker.

* heat2d        -- 2-D 5-point Jacobi ``0.25*(N+S+E+W)``, zero Dirichlet,
                   interior uniform[0,1) from seed 0 (1024^2, 100 steps).
* diffusion7    -- 3-D 7-point ``u + 0.1*(sum of 6 neighbours - 6u)``.
* star25        -- 3-D 25-point 8th-order Laplacian ``u + 0.05*L8(u)``,
                   per-axis weights -205/72, 8/5, -1/5, 8/315, -1/560,
                   periodic, 3 seeded sinusoids + 1e-3 uniform noise.
* gray_scott    -- fused ``gs_u``/``gs_v`` (Du=.16, Dv=.08, F=.035,
                   k=.065, dt=1), periodic, 64 seeded 10x10 squares.
* wave2d        -- leapfrog ``2u - uold + 0.25*(N+S+E+W-4u)``, zero
                   Dirichlet, Gaussian pulse (sigma = 0.05*side).
"""

from __future__ import annotations

import math
from typing import Optional

import numpy as np

from cutile.frontend.decorator import stencil_from_source

# 8th-order central second-derivative weights, offsets 0..4
W8 = (-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0)


def heat2d(dtype: str = "float64"):
    return stencil_from_source("""
def heat2d(u, i, j):
    return 0.25 * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1])
""", dtype=dtype)


def diffusion7(dtype: str = "float64", coef: float = 0.1, boundary=None):
    return stencil_from_source(f"""
def diffusion7(u, i, j, k):
    return u[i, j, k] + {coef!r} * (u[i - 1, j, k] + u[i + 1, j, k] + u[i, j - 1, k] + u[i, j + 1, k]
                                   + u[i, j, k - 1] + u[i, j, k + 1] - 6 * u[i, j, k])
""", dtype=dtype, boundary=boundary)


def _star_terms(r: int) -> str:
    parts = [f"C0 * u[i, j, k]"]
    for o in range(1, r + 1):
        nb = (f"u[i - {o}, j, k] + u[i + {o}, j, k] + u[i, j - {o}, k] + u[i, j + {o}, k]"
              f" + u[i, j, k - {o}] + u[i, j, k + {o}]")
        parts.append(f"C{o} * ({nb})")
    return " + ".join(parts)


def star25(dtype: str = "float64", coef: float = 0.05, boundary="periodic"):
    env = {"C0": 3.0 * W8[0], "C1": W8[1], "C2": W8[2], "C3": W8[3], "C4": W8[4]}
    return stencil_from_source(f"""
def star25(u, i, j, k):
    return u[i, j, k] + {coef!r} * ({_star_terms(4)})
""", env=env, dtype=dtype, boundary=boundary)


GS = {"Du": 0.16, "Dv": 0.08, "F": 0.035, "k": 0.065, "dt": 1.0}


def gray_scott(dtype: str = "float64"):
    env = dict(GS)
    u = stencil_from_source("""
def gs_u(u, v, i, j):
    return u[i, j] + dt * (Du * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1] - 4 * u[i, j])
                          - u[i, j] * v[i, j] * v[i, j] + F * (1.0 - u[i, j]))
""", env=env, dtype=dtype)
    v = stencil_from_source("""
def gs_v(u, v, i, j):
    return v[i, j] + dt * (Dv * (v[i - 1, j] + v[i + 1, j] + v[i, j - 1] + v[i, j + 1] - 4 * v[i, j])
                          + u[i, j] * v[i, j] * v[i, j] - (F + k) * v[i, j])
""", env=env, dtype=dtype)
    return u, v


def gray_scott_rates(dtype: str = "float64"):
    """The Gray-Scott right-hand sides R_u, R_v (the same operator trees as
    the parenthesised parts of ``gray_scott``): compiled with
    ``compile_fused(..., update="euler"|"kahan", dt=GS["dt"], targets=["u", "v"])``
    the kernel stores ``u + dt*R_u`` and ``v + dt*R_v``."""
    env = dict(GS)
    ru = stencil_from_source("""
def gs_ru(u, v, i, j):
    return (Du * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1] - 4 * u[i, j])
            - u[i, j] * v[i, j] * v[i, j] + F * (1.0 - u[i, j]))
""", env=env, dtype=dtype)
    rv = stencil_from_source("""
def gs_rv(u, v, i, j):
    return (Dv * (v[i - 1, j] + v[i + 1, j] + v[i, j - 1] + v[i, j + 1] - 4 * v[i, j])
            + u[i, j] * v[i, j] * v[i, j] - (F + k) * v[i, j])
""", env=env, dtype=dtype)
    return ru, rv


def wave2d(dtype: str = "float32", c2: float = 0.25):
    return stencil_from_source(f"""
def wave2d(u, uold, i, j):
    return 2.0 * u[i, j] - uold[i, j] + {c2!r} * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1]
                                                 - 4.0 * u[i, j])
""", dtype=dtype)


# ---------------------------------------------------------------------------
# initial fields
# ---------------------------------------------------------------------------

def padded(interior: tuple, halo: tuple) -> tuple:
    return tuple(n + 2 * h for n, h in zip(interior, halo))


def inner(halo: tuple, shape: tuple) -> tuple:
    return tuple(slice(h, s - h) for h, s in zip(halo, shape))


def heat2d_init(n: int = 1024, halo=(1, 16), seed: int = 0, dtype=np.float64) -> np.ndarray:
    u = np.zeros(padded((n, n), halo), dtype=dtype)
    u[inner(halo, u.shape)] = np.random.default_rng(seed).random((n, n))
    return u


def sinusoid_init(n, halo, seed: int = 0, dtype=np.float64, kmax: int = 8, noise: float = 1e-3) -> np.ndarray:
    """Sum of three sinusoids with integer wavevectors |k| <= kmax (periodic on
    the ``n`` interior) and random phases, plus ``noise`` * uniform[0,1)."""
    if isinstance(n, int):
        n = (n, n, n)
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < 3:
        k = rng.integers(-kmax, kmax + 1, size=3)
        if 0 < float(np.dot(k, k)) <= kmax * kmax:
            ks.append(k)
    phases = rng.random(3) * 2.0 * math.pi
    x = [np.arange(m, dtype=np.float64) / m for m in n]
    u = np.zeros(padded(tuple(n), halo), dtype=dtype)
    acc = np.zeros(tuple(n), dtype=np.float64)
    for k, ph in zip(ks, phases):
        arg = (2.0 * math.pi * k[0] * x[0])[:, None, None] + (2.0 * math.pi * k[1] * x[1])[None, :, None] \
            + (2.0 * math.pi * k[2] * x[2])[None, None, :] + ph
        acc += np.sin(arg)
    acc += noise * rng.random(tuple(n))
    u[inner(halo, u.shape)] = acc.astype(dtype)
    return u


def gray_scott_init(n: int = 4096, halo=(1, 16), seed: int = 0, squares: int = 64, side: int = 10,
                    noise: float = 0.01, dtype=np.float64) -> tuple:
    rng = np.random.default_rng(seed)
    shape = padded((n, n), halo)
    u = np.zeros(shape, dtype=dtype)
    v = np.zeros(shape, dtype=dtype)
    ui = np.ones((n, n))
    vi = np.zeros((n, n))
    for _ in range(squares):
        cx, cy = rng.integers(0, n - side + 1, size=2)
        ui[cx:cx + side, cy:cy + side] = 0.5
        vi[cx:cx + side, cy:cy + side] = 0.25
    ui += noise * rng.random((n, n))
    vi += noise * rng.random((n, n))
    u[inner(halo, shape)] = ui
    v[inner(halo, shape)] = vi
    return u, v


def wave_sides(count: int = 512, lo: int = 64, hi: int = 2048, seed: int = 0) -> list:
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(x))) for x in rng.uniform(math.log(lo), math.log(hi), size=count)]


def wave_init(side: int, halo=(1, 32), seed: int = 0, dtype=np.float32) -> tuple:
    rng = np.random.default_rng(seed)
    cx, cy = rng.uniform(0, side, size=2)
    sig = 0.05 * side
    x = np.arange(side, dtype=np.float64)
    g = np.exp(-(((x - cx) ** 2)[:, None] + ((x - cy) ** 2)[None, :]) / (2.0 * sig * sig))
    u = np.zeros(padded((side, side), halo), dtype=dtype)
    u[inner(halo, u.shape)] = g
    return u, u.copy()


def periodic_spec(ndim: int) -> dict:
    return {d: ("periodic", "periodic") for d in range(ndim)}


def dirichlet_spec(ndim: int, value: float = 0.0) -> dict:
    return {d: (("dirichlet", value), ("dirichlet", value)) for d in range(ndim)}
