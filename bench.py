#!/usr/bin/env python3
"""Benchmark of the stencil hot path on one MI355X.

Headline workload: 3-D 25-point star (8th-order Laplacian, radius 4)
diffusion step ``u += 0.05*L8(u)`` in float64 on a 512^3 interior with a
periodic halo, one step = one sweep kernel + one on-device halo fill,
temporal depth 1.  ``value`` is interior cell-updates per second over the
timed steps (inputs resident in HBM).  The same JSON line carries:

* ``roofline``: the sweep kernel's algorithmic HBM bytes per launch
  (16 B per interior cell: one float64 read, one write) divided by its
  average duration (HIP events around every sweep inside the timed loop),
  against the 8 TB/s MI355X peak;
* ``cpu_baseline``: the NumPy oracle (single core) on a 256^3 sample of
  the same workload;
* ``extra``: the 3-D 7-point float64 kernel at 512^3 (target: >= 80 % of
  the HBM roofline), the float32 25-point kernel, a device copy ceiling.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cutile-stencil-dsl_amd"))

PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of ``kernel`` measured by rocprofv3 PMC
    (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the counter
    reads for the device copy kernel of known size), as recorded in
    profiles/traffic.txt (``<kernel symbol> <bytes per launch>`` lines)."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def run_steps(launch, fill, a, b, steps, stream, rt, record_kernel=True):
    """Run ``steps`` steps (sweep a->b, halo fill b, swap); returns
    (total_ms, mean sweep ms, a, b)."""
    starts = [rt.Event() for _ in range(steps)] if record_kernel else []
    ends = [rt.Event() for _ in range(steps)] if record_kernel else []
    e0, e1 = rt.Event(), rt.Event()
    e0.record(stream)
    for k in range(steps):
        if record_kernel:
            starts[k].record(stream)
        launch(a, b)
        if record_kernel:
            ends[k].record(stream)
        fill(b)
        a, b = b, a
    e1.record(stream)
    total = e0.elapsed_ms(e1)
    kern = (sum(s.elapsed_ms(e) for s, e in zip(starts, ends)) / steps) if record_kernel else None
    return total, kern, a, b


def measure(name, stencil_fn, interior_dev, n, steps, warmup, dtype_name, spec, **compile_kw):
    import torch

    from cutile import compile
    from cutile.runtime import boundary, rt
    from cutile.runtime.kernel import default_stream

    r = compile(stencil_fn, temporal_blocking=False, **compile_kw)
    halo = r.halo_widths
    shape = tuple(n + 2 * h for h in halo)
    tdt = getattr(torch, r.storage)
    a = torch.zeros(shape, dtype=tdt, device="cuda")
    a[tuple(slice(h, h + n) for h in halo)] = interior_dev.to(tdt)
    b = torch.zeros_like(a)
    launch = r.launcher
    fill = (lambda t: boundary.fill(t, spec, halo))
    fill(a)
    s = default_stream()
    log(f"{name}: compile + warmup ({warmup} steps), halo {halo}, tile {r.tile_sizes}")
    _, _, a, b = run_steps(launch, fill, a, b, max(1, warmup), s, rt, record_kernel=False)
    torch.cuda.synchronize()
    total_ms, kern_ms, a, b = run_steps(launch, fill, a, b, steps, s, rt)
    torch.cuda.synchronize()
    cells = n ** 3
    esz = a.element_size()
    K = r.load_module()._KERNEL
    key = K.pick([a], [b], halo)
    res = K.resources(key)
    out = {
        "kernel": K.variants[key][0],
        "workload": name, "dtype": dtype_name, "interior": [n, n, n], "halo": list(halo),
        "tile": list(r.tile_sizes), "steps": steps, "ms_per_step": total_ms / steps,
        "kernel_ms": kern_ms, "gcell_updates_per_s": cells * steps / (total_ms * 1e-3) / 1e9,
        "kernel_gcells_per_s": cells / (kern_ms * 1e-3) / 1e9,
        "alg_bytes_per_launch": cells * 2 * esz,
        "kernel_gbs": cells * 2 * esz / (kern_ms * 1e-3) / 1e9,
        "vgprs": res.num_regs, "scratch_bytes": res.local_bytes, "lds_bytes": res.shared_bytes,
    }
    out["hbm_frac"] = out["kernel_gbs"] / PEAK_HBM_GBS
    log(f"{name}: {out['ms_per_step']:.4f} ms/step, kernel {kern_ms:.4f} ms "
        f"= {out['kernel_gbs']:.0f} GB/s ({100 * out['hbm_frac']:.1f}% of 8 TB/s), "
        f"{out['gcell_updates_per_s']:.1f} Gcell-updates/s")
    del a, b
    torch.cuda.empty_cache()
    return out


def copy_ceiling(nbytes, iters=20):
    import torch

    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream

    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    s = default_stream()
    for _ in range(3):
        rt.copy(dst.data_ptr(), src.data_ptr(), nbytes, s)
    e0, e1 = rt.Event(), rt.Event()
    e0.record(s)
    for _ in range(iters):
        rt.copy(dst.data_ptr(), src.data_ptr(), nbytes, s)
    e1.record(s)
    ms = e0.elapsed_ms(e1) / iters
    del src, dst
    return {"bytes_moved": 2 * nbytes, "ms": ms, "gbs": 2 * nbytes / (ms * 1e-3) / 1e9}


def cpu_baseline(n=256, steps=8):
    """The NumPy oracle (one core: NumPy ufuncs are single-threaded) on a
    256^3 sample of the headline workload."""
    sys.path.insert(0, ROOT)
    import numpy as np

    import oracle
    from cutile.workloads import periodic_spec, sinusoid_init, star25

    f = star25("float64")
    halo = (4, 4, 16)
    u = sinusoid_init(n, halo, seed=0)
    oracle.apply_boundary(u, periodic_spec(3), halo)
    u = oracle.apply_stencil(u, f._fn, 3, halo)  # warm-up step
    t0 = time.perf_counter()
    for _ in range(steps):
        u = oracle.apply_stencil(u, f._fn, 3, halo)
        oracle.apply_boundary(u, periodic_spec(3), halo)
    dt = time.perf_counter() - t0
    try:
        cpu_model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        cpu_model = platform.processor()
    return {
        "value": n ** 3 * steps / dt / 1e9, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
        "sample": f"star25 float64 {n}^3 interior, {steps} steps (stencil + periodic fill), NumPy "
                  f"{np.__version__}, {cpu_model}, os.cpu_count()={os.cpu_count()}, "
                  f"affinity={len(os.sched_getaffinity(0))}",
        "seconds": dt,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py drives one GPU: this build runs every workload on a single MI355X "
              "(no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cutile.runtime import rt
    from cutile.workloads import diffusion7, periodic_spec, sinusoid_init, star25

    assert torch.cuda.is_available(), "no ROCm device"
    n = args.n
    info = rt.device_info(0)
    log(f"device {info.name.decode()} ({info.arch.decode()}), {info.compute_units} CUs")
    t0 = time.time()
    host = sinusoid_init(n, (0, 0, 0), seed=0)
    interior = torch.from_numpy(host).to("cuda")
    del host
    log(f"initial field {n}^3 generated in {time.time() - t0:.1f}s")
    spec = periodic_spec(3)

    head = measure("star25_fp64", star25("float64"), interior, n, args.steps, args.warmup, "f64", spec)
    extra = {}
    if not args.no_extra:
        extra["diffusion7_fp64"] = measure("diffusion7_fp64", diffusion7("float64"), interior, n, args.steps,
                                           args.warmup, "f64", spec)
        extra["star25_fp32"] = measure("star25_fp32", star25("float32"), interior, n, args.steps,
                                       args.warmup, "f32", spec)
        extra["copy_ceiling"] = copy_ceiling(head["alg_bytes_per_launch"] // 2)
        log(f"copy ceiling: {extra['copy_ceiling']['gbs']:.0f} GB/s")
    del interior
    torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (NumPy oracle, 256^3 sample) ...")
        cpu = cpu_baseline()
        log(f"cpu baseline: {cpu['value']:.4f} Gcell-updates/s in {cpu['seconds']:.1f}s")

    traffic = traffic_from_profiles(head["kernel"])
    line = {
        "metric": "Gcell-updates/s + %HBM roofline, batched FP32 vs plain FP64; serving p99 ms",
        "value": head["gcell_updates_per_s"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sum of 3 sinusoids + 1e-3 uniform noise)",
        "config": {"workload": "3D 25-point star (8th-order Laplacian, r=4) u += 0.05*L(u), FP64, "
                               f"{n}^3 interior + halo 4 (inner padded to 16), periodic, temporal depth 1",
                   "interior": [n, n, n], "halo": head["halo"], "tile": head["tile"]},
        "roofline": {"bound": "hbm", "achieved": head["kernel_gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": head["hbm_frac"], "traffic": traffic,
                     "kernel_ms": head["kernel_ms"], "alg_bytes_per_launch": head["alg_bytes_per_launch"]},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
