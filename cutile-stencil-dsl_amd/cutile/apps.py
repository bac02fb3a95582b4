"""Application drivers of the upstream examples, run on the HIP path.

Each driver builds its stencils with the public API (``compile`` /
``compile_fused``), runs the example's time loop on ``cuda:0`` with the
HIP kernels of this package (sweeps, halo fills, point-wise updates) and
returns the final fields (host NumPy arrays) with a JSON-ready report:
grid, steps, device time of the loop, Gcell-updates/s and the example's own
diagnostics.  Initial conditions, parameters and update order follow the
upstream examples:

* ``heat_1d``           -- ``examples/heat_1d.py`` (Gaussian, halo kept)
* ``advection_upwind``  -- ``examples/advection_upwind.py`` (mass, centre)
* ``fdtd_maxwell_1d``   -- ``examples/fdtd_maxwell_1d.py`` (H then E,
                           Gaussian soft source at N/4)
* ``shallow_water``     -- ``examples/shallow_water.py`` (h, hu, hv from
                           the old state: one fused kernel)
* ``wave_2d``           -- ``examples/wave_2d.py`` (4th-order Laplacian,
                           leapfrog over the whole array, outer ring zeroed)
* ``gray_scott``        -- ``examples/gray_scott.py`` (u, v from the old
                           state: one fused kernel)
* ``laplacian_3d``      -- ``examples/laplacian_3d.py`` (unit impulse)
* ``heat_2d_bricked``   -- ``examples/heat_2d_bricked.py`` (bricked launch
                           against the flat one, brick round trip)

``python -m cutile.apps <name> [--size N] [--steps K]`` prints the report.
"""

from __future__ import annotations

import argparse
import json
import sys
from typing import Callable, Dict

import numpy as np

from cutile.frontend.decorator import stencil_from_source


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the application drivers run on a ROCm device")
    return torch


def _timer():
    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream

    class T:
        def __enter__(self):
            self.s = default_stream()
            self.a, self.b = rt.Event(), rt.Event()
            self.a.record(self.s)
            return self

        def __exit__(self, *exc):
            self.b.record(self.s)
            self.ms = self.a.elapsed_ms(self.b)

    return T()


def _report(app, grid, steps, ms, **diag) -> dict:
    cells = int(np.prod(grid))
    return {"app": app, "grid": list(grid), "steps": steps, "device_ms": ms,
            "gcell_updates_per_s": cells * steps / (ms * 1e-3) / 1e9 if ms > 0 else None, **diag}


def _launcher(result):
    mod = result.load_module()
    return getattr(mod, f"launch_{result.name}")


# ---------------------------------------------------------------------------
# stencils of the examples
# ---------------------------------------------------------------------------

def heat_1d_stencil():
    return stencil_from_source("""
def heat_1d(u, i):
    return 0.25 * u[i - 1] + 0.5 * u[i] + 0.25 * u[i + 1]
""")


def advection_stencil(cfl: float = 0.5):
    return stencil_from_source("""
def advection_upwind(u, i):
    return u[i] - CFL * (u[i] - u[i - 1])
""", env={"CFL": cfl})


FDTD = {"c0": 1.0, "dx": 0.01, "dt": 0.005, "eps": 1.0, "mu": 1.0}


def fdtd_stencils():
    e = stencil_from_source("""
def update_E(E, H, i):
    return E[i] + (dt / (eps * dx)) * (H[i] - H[i - 1])
""", env=dict(FDTD))
    h = stencil_from_source("""
def update_H(E, H, i):
    return H[i] + (dt / (mu * dx)) * (E[i + 1] - E[i])
""", env=dict(FDTD))
    return e, h


SWE = {"g": 9.81, "H0": 1.0, "dt": 0.01, "dx": 0.1}


def shallow_water_stencils():
    h = stencil_from_source("""
def shallow_water_h(h, hu, hv, i, j):
    return h[i, j] - dt * H0 * ((hu[i + 1, j] - hu[i - 1, j]) / (2 * dx) + (hv[i, j + 1] - hv[i, j - 1]) / (2 * dx))
""", env=dict(SWE))
    hu = stencil_from_source("""
def shallow_water_hu(h, hu, i, j):
    return hu[i, j] - dt * g * (h[i + 1, j] - h[i - 1, j]) / (2 * dx)
""", env=dict(SWE))
    hv = stencil_from_source("""
def shallow_water_hv(h, hv, i, j):
    return hv[i, j] - dt * g * (h[i, j + 1] - h[i, j - 1]) / (2 * dx)
""", env=dict(SWE))
    return h, hu, hv


def wave_2d_stencil():
    return stencil_from_source("""
def wave_2d(u, i, j):
    return 0.1 * ((-1/12) * u[i - 2, j] + (4/3) * u[i - 1, j] + (-5/2) * u[i, j] + (4/3) * u[i + 1, j] + (-1/12) * u[i + 2, j] + (-1/12) * u[i, j - 2] + (4/3) * u[i, j - 1] + (-5/2) * u[i, j] + (4/3) * u[i, j + 1] + (-1/12) * u[i, j + 2])
""")


GS_EXAMPLE = {"Du": 0.16, "Dv": 0.08, "F": 0.035, "k": 0.065, "dt": 1.0}


def gray_scott_stencils():
    u = stencil_from_source("""
def gray_scott_u(u, v, i, j):
    return u[i, j] + dt * (Du * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1] - 4 * u[i, j]) - u[i, j] * v[i, j] * v[i, j] + F * (1.0 - u[i, j]))
""", env=dict(GS_EXAMPLE))
    v = stencil_from_source("""
def gray_scott_v(u, v, i, j):
    return v[i, j] + dt * (Dv * (v[i - 1, j] + v[i + 1, j] + v[i, j - 1] + v[i, j + 1] - 4 * v[i, j]) + u[i, j] * v[i, j] * v[i, j] - (F + k) * v[i, j])
""", env=dict(GS_EXAMPLE))
    return u, v


def laplacian_3d_stencil():
    return stencil_from_source("""
def laplacian_3d(u, i, j, k):
    return (u[i - 1, j, k] + u[i + 1, j, k]
            + u[i, j - 1, k] + u[i, j + 1, k]
            + u[i, j, k - 1] + u[i, j, k + 1]
            - 6 * u[i, j, k])
""")


def heat_2d_stencil():
    return stencil_from_source("""
def heat_2d(u, i, j):
    return 0.25 * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1])
""")


# ---------------------------------------------------------------------------
# initial conditions (host, float64, halo included as in the examples)
# ---------------------------------------------------------------------------

def heat_1d_init(n: int, h: int) -> np.ndarray:
    x = np.linspace(-1, 1, n + 2 * h)
    u = np.exp(-20 * x ** 2)
    u[0] = 0.0
    u[-1] = 0.0
    return u


def advection_init(n: int, h: int) -> np.ndarray:
    u = np.zeros(n + 2 * h)
    u[h + n // 4:h + n // 2] = 1.0
    return u


def fdtd_source(s: int, t0: float = 40, spread: float = 12) -> float:
    return float(np.exp(-0.5 * ((s - t0) / spread) ** 2))


def bump_2d(shape, width: float, amp: float = 1.0) -> np.ndarray:
    """``amp * exp(-r^2)`` about the array centre (integer centre for odd
    halves as the examples compute it), over the whole padded array."""
    cx, cy = shape[0] // 2, shape[1] // 2
    ii = np.arange(shape[0], dtype=np.float64)[:, None]
    jj = np.arange(shape[1], dtype=np.float64)[None, :]
    r2 = ((ii - cx) / width) ** 2 + ((jj - cy) / width) ** 2
    return amp * np.exp(-r2)


def gray_scott_init(shape, r: int = 5):
    u = np.ones(shape)
    v = np.zeros(shape)
    cx, cy = shape[0] // 2, shape[1] // 2
    u[cx - r:cx + r, cy - r:cy + r] = 0.50
    v[cx - r:cx + r, cy - r:cy + r] = 0.25
    return u, v


# ---------------------------------------------------------------------------
# drivers
# ---------------------------------------------------------------------------

def _pair(torch, a: np.ndarray):
    """Two device buffers holding ``a``: a sweep writes only the interior,
    so both halos stay what the example's arrays hold."""
    d = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return d, d.clone()


def heat_1d(size: int = 128, steps: int = 100) -> dict:
    from cutile import compile

    torch = _torch()
    fn = heat_1d_stencil()
    r = compile(fn, temporal_blocking=False)     # one step per launch, as the example loops
    launch = _launcher(r)
    h = r.halo_widths[0]
    u0 = heat_1d_init(size, h)
    a, b = _pair(torch, u0)
    with _timer() as t:
        for _ in range(steps):
            launch(a, b)
            a, b = b, a
    u = a.cpu().numpy()
    return {"fields": {"u": u},
            "report": _report("heat_1d", (size,), steps, t.ms, energy_initial=float(np.sum(u0 ** 2)),
                              energy_final=float(np.sum(u ** 2)))}


def advection_upwind(size: int = 200, steps: int = 80, cfl: float = 0.5) -> dict:
    from cutile import compile

    torch = _torch()
    r = compile(advection_stencil(cfl), temporal_blocking=False)
    launch = _launcher(r)
    h = r.halo_widths[0]
    u0 = advection_init(size, h)
    a, b = _pair(torch, u0)
    with _timer() as t:
        for _ in range(steps):
            launch(a, b)
            a, b = b, a
    u = a.cpu().numpy()
    inner0, inner = u0[h:-h], u[h:-h]
    m0, m1 = float(np.sum(inner0)), float(np.sum(inner))
    x = np.arange(size)
    return {"fields": {"u": u},
            "report": _report("advection_upwind", (size,), steps, t.ms, mass_initial=m0, mass_final=m1,
                              centre_initial=float(np.sum(x * inner0) / max(m0, 1e-10)),
                              centre_final=float(np.sum(x * inner) / max(m1, 1e-10)))}


def fdtd_maxwell_1d(size: int = 500, steps: int = 200) -> dict:
    from cutile import compile

    torch = _torch()
    fe, fh = fdtd_stencils()
    re_, rh = compile(fe, temporal_blocking=False), compile(fh, temporal_blocking=False)
    le, lh = _launcher(re_), _launcher(rh)
    h = max(re_.halo_widths[0], rh.halo_widths[0])
    if re_.halo_widths != rh.halo_widths:
        raise ValueError("E and H updates need the same halo")
    E, E2 = _pair(torch, np.zeros(size + 2 * h))
    H, H2 = _pair(torch, np.zeros(size + 2 * h))
    src = size // 4 + h
    with _timer() as t:
        for s in range(steps):
            lh(E, H, H2)           # H first (leapfrog), from the old E
            H, H2 = H2, H
            le(E, H, E2)           # then E from the new H
            E, E2 = E2, E
            E[src] += fdtd_source(s)
    e, hh = E.cpu().numpy(), H.cpu().numpy()
    return {"fields": {"E": e, "H": hh},
            "report": _report("fdtd_maxwell_1d", (size,), steps, t.ms, max_abs_E=float(np.max(np.abs(e))),
                              max_abs_H=float(np.max(np.abs(hh))))}


def shallow_water(size: int = 32, steps: int = 50) -> dict:
    from cutile import compile_fused

    torch = _torch()
    r = compile_fused(list(shallow_water_stencils()))
    launch = _launcher(r)
    hx, hy = r.halo_widths
    shape = (size + 2 * hx, size + 2 * hy)
    h0 = bump_2d(shape, 3.0, 0.1)
    f = {"h": _pair(torch, h0), "hu": _pair(torch, np.zeros(shape)), "hv": _pair(torch, np.zeros(shape))}
    cur = {k: v[0] for k, v in f.items()}
    nxt = {k: v[1] for k, v in f.items()}
    with _timer() as t:
        for _ in range(steps):
            launch(cur["h"], cur["hu"], cur["hv"], nxt["h"], nxt["hu"], nxt["hv"])
            cur, nxt = nxt, cur
    out = {k: v.cpu().numpy() for k, v in cur.items()}
    inner = (slice(hx, -hx), slice(hy, -hy))
    return {"fields": out,
            "report": _report("shallow_water", (size, size), steps, t.ms,
                              energy_initial=float(np.sum(h0[inner] ** 2)),
                              energy_final=float(np.sum(out["h"][inner] ** 2)))}


def wave_2d(size: int = 64, steps: int = 50) -> dict:
    from cutile import compile
    from cutile.runtime import boundary
    from cutile.runtime.pointwise import PointwiseKernel

    torch = _torch()
    r = compile(wave_2d_stencil(), temporal_blocking=False)
    launch = _launcher(r)
    hx, hy = r.halo_widths
    shape = (size + 2 * hx, size + 2 * hy)
    u0 = bump_2d(shape, 5.0)
    u = torch.from_numpy(u0).to("cuda")
    uold = u.clone()
    lap = u.clone()
    unew = torch.empty_like(u)
    # u_new = 2*u - u_old + lap over the whole array, as the example writes it
    leap = PointwiseKernel("2.0 * x0 - x1 + x2", nin=3, name="wave_leap")
    ring = {0: (("dirichlet", 0.0), ("dirichlet", 0.0)), 1: (("dirichlet", 0.0), ("dirichlet", 0.0))}
    with _timer() as t:
        for _ in range(steps):
            lap.copy_(u)                       # apply_stencil: a copy of u, interior replaced
            launch(u, lap)
            leap([unew], [u, uold, lap])
            boundary.fill(unew, ring, (1, 1))  # the outermost ring only
            uold, u, unew = u, unew, uold
    out = u.cpu().numpy()
    return {"fields": {"u": out},
            "report": _report("wave_2d", (size, size), steps, t.ms, max_abs_u=float(np.max(np.abs(out))))}


def gray_scott(size: int = 64, steps: int = 100) -> dict:
    from cutile import compile_fused

    torch = _torch()
    r = compile_fused(list(gray_scott_stencils()))
    launch = _launcher(r)
    hx, hy = r.halo_widths
    shape = (size + 2 * hx, size + 2 * hy)
    u0, v0 = gray_scott_init(shape)
    (u, u2), (v, v2) = _pair(torch, u0), _pair(torch, v0)
    with _timer() as t:
        for _ in range(steps):
            launch(u, v, u2, v2)
            u, u2, v, v2 = u2, u, v2, v
    uu, vv = u.cpu().numpy(), v.cpu().numpy()
    return {"fields": {"u": uu, "v": vv},
            "report": _report("gray_scott", (size, size), steps, t.ms, u_range=[float(uu.min()), float(uu.max())],
                              v_range=[float(vv.min()), float(vv.max())])}


def laplacian_3d(size: int = 16, steps: int = 1) -> dict:
    from cutile import compile

    torch = _torch()
    r = compile(laplacian_3d_stencil(), temporal_blocking=False)
    launch = _launcher(r)
    halo = r.halo_widths
    u0 = np.zeros(tuple(size + 2 * h for h in halo))
    c = tuple(size // 2 + h for h in halo)
    u0[c] = 1.0
    a, b = _pair(torch, u0)
    with _timer() as t:
        launch(a, b)
    out = b.cpu().numpy()
    return {"fields": {"lap": out},
            "report": _report("laplacian_3d", (size,) * 3, 1, t.ms, centre=float(out[c]))}


def heat_2d_bricked(size: int = 128, steps: int = 1, brick: int = 32) -> dict:
    from cutile import compile
    from cutile.runtime.bricks import from_bricks, to_bricks

    torch = _torch()
    fn = heat_2d_stencil()
    flat_r = compile(fn, temporal_blocking=False)
    br_r = compile(fn, temporal_blocking=False, layout="bricked", brick_size=brick)
    hx, hy = flat_r.halo_widths
    shape = (size + 2 * hx, size + 2 * hy)
    cx, cy = shape[0] / 2.0, shape[1] / 2.0
    ii = np.arange(shape[0], dtype=np.float64)[:, None]
    jj = np.arange(shape[1], dtype=np.float64)[None, :]
    u0 = np.exp(-(((ii - cx) / 10.0) ** 2 + ((jj - cy) / 10.0) ** 2))
    u = torch.from_numpy(u0).to("cuda")
    flat_out = u.clone()
    _launcher(flat_r)(u, flat_out)
    br_mod = br_r.load_module()
    br_out = torch.zeros_like(u)
    with _timer() as t:
        getattr(br_mod, f"launch_{br_r.name}_bricked")(u, br_out, brick_size=brick)
    rt_u = from_bricks(to_bricks(u, brick), u.shape, brick)
    inner = (slice(hx, -hx), slice(hy, -hy))
    fo, bo = flat_out.cpu().numpy(), br_out.cpu().numpy()
    return {"fields": {"flat": fo, "bricked": bo},
            "report": _report("heat_2d_bricked", (size, size), 1, t.ms,
                              bricked_equals_flat=bool(np.array_equal(fo[inner], bo[inner])),
                              round_trip_exact=bool(torch.equal(rt_u, u)))}


APPS: Dict[str, Callable[..., dict]] = {
    "heat_1d": heat_1d, "advection_upwind": advection_upwind, "fdtd_maxwell_1d": fdtd_maxwell_1d,
    "shallow_water": shallow_water, "wave_2d": wave_2d, "gray_scott": gray_scott,
    "laplacian_3d": laplacian_3d, "heat_2d_bricked": heat_2d_bricked,
}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="upstream example applications on the HIP path")
    ap.add_argument("app", choices=sorted(APPS) + ["all"])
    ap.add_argument("--size", type=int, default=None, help="interior extent per axis")
    ap.add_argument("--steps", type=int, default=None)
    args = ap.parse_args(argv)
    names = sorted(APPS) if args.app == "all" else [args.app]
    for name in names:
        kw = {}
        if args.size is not None:
            kw["size"] = args.size
        if args.steps is not None:
            kw["steps"] = args.steps
        print(json.dumps(APPS[name](**kw)["report"]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
