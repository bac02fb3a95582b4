"""The stencil benchmark suite of the upstream ``run_benchmarks.py``, on the
HIP path.

Same stencils (``run_benchmarks.py:87-133``: heat 1-D/2-D, 5-point
Laplacian, 4th-order 1-D Laplacian, 4th-order 2-D wave operator, 7-point 3-D
Laplacian, upwind advection), same domains (``:135-143``), same measurement
(one launch ``launch(u, out)`` on random float64 input, 50 warm-up and 200
timed launches between two events, ``:61-74``) and the same JSON report:
``{"gpu", "n_gpus", "peak_bw_gbs", "autotuned", "results": [{"stencil",
"domain", "tiles", "cutile_gps", "cutile_ms", "eff_bw_pct"}, ...]}``
(``:230-300``), where ``eff_bw_pct`` is one read and one write of every
interior point (8 bytes each) against the measured device-to-device copy
bandwidth (``:38-59``; here ``st_copy``).  Kernels are compiled with
``temporal_blocking=False`` as upstream, so one launch is one step.

``python -m cutile.suite [--output FILE] [--max-domains K] [--autotune]``.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

from cutile.frontend.decorator import stencil_from_source

SOURCES = {
    "heat_1d": ("def heat_1d(u, i):\n    return 0.25 * u[i - 1] + 0.5 * u[i] + 0.25 * u[i + 1]\n", 2),
    "heat_2d": ("def heat_2d(u, i, j):\n    return 0.25 * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1])\n", 2),
    "lap_2d_5pt": ("def lap_2d(u, i, j):\n    return u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1] - 4.0 * u[i, j]\n",
                   2),
    "lap_1d_4th": ("def lap_1d_4th(u, i):\n    return (-u[i-2] + 16*u[i-1] - 30*u[i] + 16*u[i+1] - u[i+2]) / 12.0\n", 4),
    "wave_2d_4th": ("def wave_2d(u, i, j):\n    return (-u[i-2,j] + 16*u[i-1,j] - 30*u[i,j] + 16*u[i+1,j] - u[i+2,j]\n"
                    "            - u[i,j-2] + 16*u[i,j-1] - 30*u[i,j] + 16*u[i,j+1] - u[i,j+2]) / 12.0\n", 4),
    "lap_3d_7pt": ("def lap_3d(u, i, j, k):\n    return (u[i-1,j,k] + u[i+1,j,k] + u[i,j-1,k] + u[i,j+1,k]\n"
                   "            + u[i,j,k-1] + u[i,j,k+1]) / 6.0\n", 2),
    "advect_1d": ("def advect_1d(u, i):\n    return u[i] - 0.5 * (u[i] - u[i - 1])\n", 2),
}

DOMAINS = {
    "heat_1d": [(2 ** 16,), (2 ** 18,), (2 ** 20,), (2 ** 22,)],
    "heat_2d": [(512, 512), (1024, 1024), (2048, 2048), (4096, 4096)],
    "lap_2d_5pt": [(512, 512), (1024, 1024), (2048, 2048), (4096, 4096)],
    "lap_1d_4th": [(2 ** 16,), (2 ** 18,), (2 ** 20,), (2 ** 22,)],
    "wave_2d_4th": [(512, 512), (1024, 1024), (2048, 2048), (4096, 4096)],
    "lap_3d_7pt": [(32, 32, 32), (64, 64, 64), (128, 128, 128)],
    "advect_1d": [(2 ** 16,), (2 ** 18,), (2 ** 20,), (2 ** 22,)],
}


def stencils() -> dict:
    """name -> (stencil, ndim, order)."""
    out = {}
    for name, (src, order) in SOURCES.items():
        ndim = len(DOMAINS[name][0])
        out[name] = (stencil_from_source(src, ndim=ndim, order=order), ndim, order)
    return out


def copy_bandwidth(nbytes: int = 256 * 1024 * 1024, iters: int = 100) -> float:
    """Device-to-device copy GB/s (read + write), as upstream's peak figure."""
    import torch

    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream

    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").normal_()
    b = torch.empty_like(a)
    s = default_stream()
    for _ in range(10):
        rt.copy(b.data_ptr(), a.data_ptr(), nbytes, s)
    e0, e1 = rt.Event(), rt.Event()
    e0.record(s)
    for _ in range(iters):
        rt.copy(b.data_ptr(), a.data_ptr(), nbytes, s)
    e1.record(s)
    return 2 * nbytes / (e0.elapsed_ms(e1) / iters * 1e-3) / 1e9


def time_launch(launch, shape, warmup: int = 50, iters: int = 200) -> float:
    """ms per launch of ``launch(u, out)`` on random float64 input."""
    import torch

    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream

    u = torch.randn(shape, dtype=torch.float64, device="cuda")
    out = torch.zeros_like(u)
    for _ in range(warmup):
        launch(u, out)
    torch.cuda.synchronize()
    s = default_stream()
    e0, e1 = rt.Event(), rt.Event()
    e0.record(s)
    for _ in range(iters):
        launch(u, out)
    e1.record(s)
    return e0.elapsed_ms(e1) / iters


def run(max_domains: int = 4, warmup: int = 50, iters: int = 200, names=None, log=print,
        autotune: bool = False) -> dict:
    import torch

    from cutile import compile
    from cutile.runtime import rt

    if not torch.cuda.is_available():
        raise RuntimeError("the benchmark suite runs on a ROCm device")
    info = rt.device_info(0)
    peak = copy_bandwidth()
    rows = []
    for name, (fn, ndim, order) in stencils().items():
        if names and name not in names:
            continue
        r = compile(fn, temporal_blocking=False)
        launch = getattr(r.load_module(), f"launch_{r.name}")
        for domain in DOMAINS[name][:max_domains]:
            if autotune:      # tiles tuned for this domain (cached), as upstream's --autotune
                r = compile(fn, temporal_blocking=False, autotune=True, domain=domain)
                launch = getattr(r.load_module(), f"launch_{r.name}")
            shape = tuple(d + 2 * h for d, h in zip(domain, r.halo_widths))
            ms = time_launch(launch, shape, warmup, iters)
            npts = math.prod(domain)
            gps = npts / (ms * 1e-3) / 1e9
            bw = npts * 8 * 2 / (ms * 1e-3) / 1e9
            row = {"stencil": name, "domain": list(domain), "tiles": list(r.tile_sizes),
                   "cutile_gps": round(gps, 2), "cutile_ms": round(ms, 4), "eff_bw_pct": round(bw / peak * 100, 1)}
            log(f"{name:>12s} | {'x'.join(map(str, domain)):>15s} | {'x'.join(map(str, r.tile_sizes)):>12s} | "
                f"{gps:9.1f} GP/s | {ms:9.4f} ms | {row['eff_bw_pct']:6.1f}%")
            rows.append(row)
    return {"gpu": info.name.decode(), "arch": info.arch.decode(), "n_gpus": rt.device_count(),
            "peak_bw_gbs": round(peak, 1), "autotuned": autotune, "timestamp": time.strftime("%Y-%m-%dT%H:%M:%S"),
            "results": rows}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="stencil benchmark suite on the HIP path")
    ap.add_argument("--output", default="benchmark_results.json")
    ap.add_argument("--max-domains", type=int, default=4)
    ap.add_argument("--stencil", action="append", help="only these stencils (repeatable)")
    ap.add_argument("--autotune", action="store_true", help="tune tiles per domain on the device (cached)")
    args = ap.parse_args(argv)
    res = run(args.max_domains, names=args.stencil, autotune=args.autotune)
    with open(args.output, "w") as f:
        json.dump(res, f, indent=2)
    print(f"wrote {args.output}: {len(res['results'])} rows, copy peak {res['peak_bw_gbs']} GB/s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
