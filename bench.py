#!/usr/bin/env python3
"""Benchmark of the stencil hot path on one MI355X.

Headline workload: 3-D 25-point star (8th-order Laplacian, radius 4)
diffusion step ``u += 0.05*L8(u)`` in float64 on a 512^3 interior with
periodic boundaries, temporal depth 1: one step = one sweep kernel whose
reads wrap around the periodic axes (no halo fill between steps).
``value`` is interior cell-updates per second over the timed steps
(inputs resident in HBM): the K steps are captured as one hipGraph and
one replay is timed between one pair of HIP events.  The same JSON line
carries:

* ``roofline``: the sweep kernel's algorithmic HBM bytes per launch
  (16 B per interior cell: one float64 read, one write) divided by its
  average duration (the timed replay / K, one kernel per step), against
  the 8 TB/s MI355X peak; ``traffic`` is the kernel's measured HBM bytes
  per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated, from
  profiles/traffic.txt);
* ``cpu_baseline``: the NumPy oracle (single core) on a 256^3 sample of
  the same workload;
* ``extra``: the 3-D 7-point float64 kernel at 512^3, the float32
  25-point kernel, copy ceilings (three copy forms), temporal depths 1, 2,
  4 of the 25-point and 7-point stars (50 steps), 2-D heat, Gray-Scott,
  the batched wave and the 60-s serving stream.

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


def measure(name, stencil_fn, interior_dev, n, steps, warmup, dtype_name, spec, warmup_s=0.5, **compile_kw):
    """``steps`` periodic steps of one stencil, each one wrapped sweep (reads
    past the interior wrap around, so no halo fill runs between steps),
    captured once as a hipGraph on a side stream and replayed.  Warm-up:
    ``warmup`` eager steps, then graph replays until ``warmup_s`` seconds
    have passed.  The timed region is one replay of the ``steps``-step
    graph between one pair of HIP events on the launch stream; every step
    is one kernel, so ``kernel_ms`` = that time / ``steps``."""
    import torch

    from cutile import compile
    from cutile.runtime import boundary, rt

    r = compile(stencil_fn, temporal_blocking=False, **compile_kw)
    halo = r.halo_widths
    shape = tuple(n + 2 * h for h in halo)
    tdt = getattr(torch, r.storage)
    a = torch.zeros(shape, dtype=tdt, device="cuda")
    a[tuple(slice(h, h + n) for h in halo)] = interior_dev.to(tdt)
    b = torch.zeros_like(a)
    boundary.fill(a, spec, halo)
    K = r.load_module()._KERNEL
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    s = side.cuda_stream

    def enqueue(k):
        x, y = (a, b) if k % 2 == 0 else (b, a)
        K.launch([x], [y], halo, s, validate=False, periodic_axes=(0, 1, 2))

    log(f"{name}: compile + warmup ({warmup} steps, then {warmup_s:.1f} s of graph replays), halo {halo}, "
        f"tile {r.tile_sizes}")
    with torch.cuda.stream(side):
        for k in range(max(2, warmup + warmup % 2)):       # an even count: the state is back in a
            enqueue(k)
        g = rt.Graph(s)
        with g:
            for k in range(steps):
                enqueue(k)
        t0 = time.perf_counter()
        reps = 0
        while reps < 1 or time.perf_counter() - t0 < warmup_s:
            g.launch(s)
            rt.stream_sync(s)
            reps += 1
        torch.cuda.synchronize()
        total_ms = _timed(s, lambda: g.launch(s), rt)
        torch.cuda.synchronize()
    kern_ms = total_ms / steps
    cells = n ** 3
    esz = a.element_size()
    key = K.pick([a], [b], halo)
    res = K.resources(key)
    out = {
        "kernel": K.variants[key][0],
        "workload": name, "dtype": dtype_name, "interior": [n, n, n], "halo": list(halo),
        "tile": list(r.tile_sizes), "steps": steps, "ms_per_step": total_ms / steps,
        "kernel_ms": kern_ms, "timing": f"one hipGraph of {steps} sweeps, one event pair, "
                                        f"after {reps} warm-up replays",
        "gcell_updates_per_s": cells * steps / (total_ms * 1e-3) / 1e9,
        "alg_bytes_per_launch": cells * 2 * esz,
        "kernel_gbs": cells * 2 * esz / (kern_ms * 1e-3) / 1e9,
        "vgprs": res.num_regs, "scratch_bytes": res.local_bytes, "lds_bytes": res.shared_bytes,
    }
    out["hbm_frac"] = out["kernel_gbs"] / PEAK_HBM_GBS
    log(f"{name}: {out['ms_per_step']:.4f} ms/step = {out['kernel_gbs']:.0f} GB/s algorithmic "
        f"({100 * out['hbm_frac']:.1f}% of 8 TB/s), {out['gcell_updates_per_s']:.1f} Gcell-updates/s")
    del a, b, g
    torch.cuda.empty_cache()
    return out


def copy_ceiling(nbytes, iters=20):
    """Device-to-device copy bandwidth of ``nbytes`` three ways: the
    runtime's one-vector-per-thread kernel (``st_copy``), its multi-vector
    form (``st_copy_wide``, 4 and 8 vectors per thread) and the HIP
    runtime's own ``hipMemcpyAsync`` (called directly in libamdhip64)."""
    import ctypes

    import torch

    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream

    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    s = default_stream()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    hip.hipMemcpyAsync.restype = ctypes.c_int
    d, p = dst.data_ptr(), src.data_ptr()

    def memcpy():
        rc = hip.hipMemcpyAsync(d, p, nbytes, 3, s)      # 3 = hipMemcpyDeviceToDevice
        if rc != 0:
            raise RuntimeError(f"hipMemcpyAsync failed ({rc})")

    kinds = {"k_copy16": lambda: rt.copy(d, p, nbytes, s),
             "k_copy16_wide_x4": lambda: rt.copy_wide(d, p, nbytes, 4, s),
             "k_copy16_wide_x8": lambda: rt.copy_wide(d, p, nbytes, 8, s),
             "hipMemcpyAsync": memcpy}
    out = {"bytes_moved": 2 * nbytes}
    for name, fn in kinds.items():
        for _ in range(3):
            fn()
        ms = _timed(s, lambda: [fn() for _ in range(iters)], rt) / iters
        out[name] = {"ms": ms, "gbs": 2 * nbytes / (ms * 1e-3) / 1e9}
    out["gbs"] = max(v["gbs"] for k, v in out.items() if isinstance(v, dict))
    del src, dst
    return out


def _timed(stream, fn, rt):
    e0, e1 = rt.Event(), rt.Event()
    e0.record(stream)
    fn()
    e1.record(stream)
    return e0.elapsed_ms(e1)


def _full_graph(st, u, steps, rt):
    """One hipGraph of all ``steps`` steps of stepper ``st`` from ``u``:
    whole blocks and the single steps after the last whole block (50 steps
    at depth 4: 12 blocks and 2 sweeps), for timing (``march`` captures
    whole blocks only and runs the remainder once, outside its graph)."""
    import torch

    b = torch.empty_like(u)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        g = rt.Graph(side.cuda_stream)
        with g:
            st.enqueue(u, b, steps, side.cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    g._buffers = (u, b)               # the replays read and write these
    return g


def _replay_ms(g, rt, warm_s=0.25, reps=5):
    """Median time of one replay of graph ``g``: replays for ``warm_s``
    seconds first (clocks and caches settle, code objects are resident),
    then ``reps`` replays, each timed by its own event pair on the graph's
    stream.  The first replay after a capture ran up to 16 % slower than the
    ones after it, so a single timed replay is not used."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < warm_s:
        g.launch()
        rt.stream_sync(g.stream)
    ms = sorted(_timed(g.stream, g.launch, rt) for _ in range(reps))
    return ms[len(ms) // 2]


def temporal_depths(interior_dev, n, steps=50):
    """25-point star (the C2 workload) and 7-point diffusion, float64 and
    float32, temporal depth 1, 2 and 4, periodic: all ``steps`` steps as one
    captured hipGraph (at depth 4, 50 steps are 12 blocks and 2 sweeps),
    timed as the median replay after warm-up.  At depth 1 every step is one
    wrapped sweep; deeper blocks advance T steps per launch in registers and
    LDS, reading around the periodic axes (no halo fills); where a depth's
    ghost ring does not fit (the 25-point star at depth 4) a block runs as
    launches of the largest depth that fits (2).  Every depth's state after
    ``steps`` steps is compared bit for bit with depth 1's."""
    import torch

    from cutile.runtime import rt
    from cutile.runtime.timestep import TimeStepper
    from cutile.workloads import diffusion7, periodic_spec, star25

    out = {}
    for sname, mk in (("star25", star25), ("diffusion7", diffusion7)):
        ref64 = None                  # the float64 state after `steps` steps (bit-equal to the oracle's)
        for dt_name in ("float64", "float32"):
            first = None
            for T in (1, 2, 4):
                key = f"{sname}_{dt_name}_T{T}"
                st = TimeStepper(mk(dt_name), periodic_spec(3), temporal_steps=T, split="fit")
                u = st.empty_like_interior((n, n, n))
                inner = tuple(slice(h, h + n) for h in st.halo)
                u[inner] = interior_dev.to(u.dtype)
                # capture the step loop once (JIT build, one replay), then time replays
                st.march(u, steps, graph=True, graph_steps=steps)
                torch.cuda.synchronize()
                state = u[inner].clone()
                if first is None:
                    first = state
                    if dt_name == "float64":
                        ref64 = state
                same = bool(torch.equal(state, first))
                err = None
                if dt_name != "float64" and ref64 is not None:
                    err = float((state.to(torch.float64) - ref64).abs().max() / ref64.abs().max())
                g = _full_graph(st, u, steps, rt)
                ms = _replay_ms(g, rt)
                torch.cuda.synchronize()
                kind = "sweep" if st._block is None else (
                    "registers+LDS" if "tblock" in st._block.variants else "LDS rings")
                if st.block_reps > 1:
                    kind += f", {st.block_reps} launches of depth {st.T // st.block_reps} per block"
                out[key] = {"T": T, "dtype": dt_name, "steps": steps, "halo": list(st.halo), "kernel": kind,
                            **({"rel_err_vs_float64": err} if err is not None else {}),
                            "ms_per_step": ms / steps, "gcell_updates_per_s": n ** 3 * steps / (ms * 1e-3) / 1e9,
                            "alg_bytes_per_update": 2 * u.element_size() / T, "state_equals_T1": same}
                log(f"{key}: {ms / steps:.4f} ms/step, {out[key]['gcell_updates_per_s']:.1f} Gcell-updates/s "
                    f"({kind}; state after {steps} steps equal to T=1: {same}"
                    + (f"; rel max err vs float64 {err:.2e})" if err is not None else ")"))
                del u, st, g, state
                torch.cuda.empty_cache()
            # the depth split the autotuner picks for a block of 4 steps
            st = TimeStepper(mk(dt_name), periodic_spec(3), temporal_steps=4, split="auto")
            u = st.empty_like_interior((n, n, n))
            kern, reps = st.plan_for(u)
            depth = 1 if kern is None else 4 // reps
            out[f"{sname}_{dt_name}_T4_tuned_split"] = f"{4 // depth} x depth {depth}"
            log(f"{sname} {dt_name}: autotuned split of a 4-step block: {4 // depth} x depth {depth}")
            del u, st
            torch.cuda.empty_cache()
            del first
    return out


def heat2d_c1(steps=100, n=1024):
    """2-D heat 5-point, float64, 1024^2, zero Dirichlet, 100 steps as one
    captured hipGraph, at temporal depth 1 (one sweep per step, halo filled
    once), 4 and 10 (zero-halo levels in LDS: the same as a zero-Dirichlet
    refill every step); each result checked bit for bit against the NumPy
    oracle, which is also timed on the same run."""
    import numpy as np
    import torch

    sys.path.insert(0, ROOT)
    import oracle
    from cutile.runtime import rt
    from cutile.runtime.timestep import TimeStepper
    from cutile.workloads import dirichlet_spec, heat2d, heat2d_init

    spec = dirichlet_spec(2)
    ref, out, cpu_s = None, {}, None
    for T, split in ((1, "fit"), (4, "fit"), (10, "fit"), (4, "auto"), (10, "auto")):
        st = TimeStepper(heat2d(), spec, temporal_steps=T, split=split)
        u0 = heat2d_init(n, st.halo, seed=0)
        if ref is None or ref[0] != st.halo:
            t0 = time.perf_counter()
            r = oracle.time_march(u0, heat2d()._fn, 2, st.halo, steps, spec)
            cpu_s = time.perf_counter() - t0
            ref = (st.halo, r)
        u = torch.from_numpy(u0).cuda()
        res = st.march(u, steps, graph=True, graph_steps=steps)     # capture + one replay
        torch.cuda.synchronize()
        got = res.cpu().numpy()
        g = _full_graph(st, u, steps, rt)
        ms = _replay_ms(g, rt)
        torch.cuda.synchronize()
        I = oracle.interior(u0.shape, st.halo)
        exact = bool(np.array_equal(got[I], ref[1][I]))
        name = f"T{T}" if split == "fit" else f"T{T}_tuned"
        kern, reps = st.plan_for(u) if T > 1 else (None, 1)
        depth = 1 if kern is None else T // reps
        out[name] = {"ms_per_step": ms / steps, "gcell_updates_per_s": n * n * steps / (ms * 1e-3) / 1e9,
                     "bit_exact_vs_numpy": exact, "halo": list(st.halo),
                     "launches": f"{T // depth} x depth {depth} per block of {T}"}
        log(f"heat2d 1024^2 x{steps} T={T} ({split}: {out[name]['launches']}): "
            f"{out[name]['gcell_updates_per_s']:.1f} Gcell-updates/s, bit-exact={exact}")
        del u, res, st, g
    out["cpu_numpy_gcell_updates_per_s"] = n * n * steps / cpu_s / 1e9
    log(f"heat2d NumPy: {out['cpu_numpy_gcell_updates_per_s']:.3f} Gcell-updates/s")
    return out


def gray_scott_c3(n=4096, steps=1000):
    """Fused Gray-Scott 4096^2, periodic, 1000 steps: float64, float32,
    bfloat16 storage with a Kahan-compensated update (float32 arithmetic),
    float32 with a Kahan update.  Errors per field after all steps, against
    float64 runs from the exact (float64) initial state and from the same
    initial state rounded to the variant's storage.  With these parameters
    the 64 seeded squares die out: v decays to ~1e-41 (subnormal) by step
    1000 and u tends to 1, so the final errors are small for u and measure
    subnormal resolution for v; the errors at step 200, while the pattern is
    alive, are the informative precision comparison.  float64 and float32
    also run through the coupled-field temporal kernel (``_T2``, ``_T4``:
    blocks of 2 or 4 fused steps per launch), whose final state must equal
    the one-step run of the same precision bit for bit."""
    import torch

    from cutile import compile_fused
    from cutile.runtime import boundary, rt
    from cutile.runtime.launcher import compile_fused_block
    from cutile.runtime.timestep import FusedStepper
    from cutile.workloads import GS, gray_scott, gray_scott_init, gray_scott_rates, periodic_spec

    spec = periodic_spec(2)
    r64 = compile_fused(list(gray_scott("float64")))
    r32 = compile_fused(list(gray_scott("float32")))
    blocks = {(dt, T): compile_fused_block(list(gray_scott(dt)), T, r, ["u", "v"]) for dt, r in (("float64", r64), ("float32", r32))
              for T in (2, 4)}
    variants = [
        ("float64", r64, None, torch.float64),
        ("float32", r32, None, torch.float32),
        ("bfloat16_kahan", compile_fused(list(gray_scott_rates("float32")), update="kahan", dt=GS["dt"],
                                         targets=["u", "v"], storage="bfloat16"), "kahan", torch.bfloat16),
        ("float32_kahan", compile_fused(list(gray_scott_rates("float32")), update="kahan", dt=GS["dt"],
                                        targets=["u", "v"]), "kahan", torch.float32),
        # float64 runs from the initial state rounded to float32 / bfloat16
        ("float64_from_float32_init", r64, None, torch.float32),
        ("float64_from_bfloat16_init", r64, None, torch.bfloat16),
    ] + [(f"{dt}_T{T}", r64 if dt == "float64" else r32, None, getattr(torch, dt)) for dt in ("float64", "float32")
         for T in (2, 4)]
    out, finals, early = {}, {}, {}
    for name, r, upd, init_dt in variants:
        blk = None
        if "_T" in name:
            bdt, bT = name.split("_T")
            blk = blocks[bdt, int(bT)]
        u0, v0 = gray_scott_init(n, r.halo_widths, seed=0)
        tdt = getattr(torch, r.storage)
        f = {"u": torch.from_numpy(u0).cuda().to(init_dt).to(tdt), "v": torch.from_numpy(v0).cuda().to(init_dt).to(tdt)}
        if upd:
            f["c_u"], f["c_v"] = torch.zeros_like(f["u"]), torch.zeros_like(f["v"])
        targets = list(r.outputs) if upd else ["u", "v"]
        scratch = {t: torch.empty_like(f[t]) for t in targets}
        stp = FusedStepper(r, spec, targets, block=blk)
        for t in targets:
            boundary.fill(f[t], spec, r.halo_widths)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            stp.enqueue(f, scratch, 2, side.cuda_stream)    # JIT of the one-step kernel
            if blk is not None:                              # and of the block kernel, before any capture
                bk = blk.load_module()._KERNEL
                bk._function(bk.temporal_key([f[t] for t in r.inputs], [scratch[t] for t in targets],
                                             r.halo_widths))
            # two graphs: 198 steps (to a snapshot at step 200, while the
            # pattern is alive) and 200 steps (replayed to the end).  The
            # wrapped sweeps leave halos stale; each graph refills them once,
            # at its end
            g198, g200 = rt.Graph(side.cuda_stream), rt.Graph(side.cuda_stream)
            with g198:
                stp.enqueue(f, scratch, 198, side.cuda_stream)
            with g200:
                stp.enqueue(f, scratch, 200, side.cuda_stream)
            ms = _timed(side.cuda_stream, lambda: g198.launch(), rt)
            side.synchronize()
            h = r.halo_widths
            # copy=True: a float64 slice would otherwise alias the live buffer
            early[name] = {t: f[t][h[0]:h[0] + n, h[1]:h[1] + n].to(torch.float64, copy=True) for t in ("u", "v")}
            ms += _timed(side.cuda_stream, lambda: [g200.launch() for _ in range((steps - 200) // 200)], rt)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        finals[name] = {t: f[t][h[0]:h[0] + n, h[1]:h[1] + n].to(torch.float64) for t in ("u", "v")}
        timed = steps - 2
        if not name.startswith("float64_from"):
            esz = f["u"].element_size()
            depth = int(name.split("_T")[1]) if "_T" in name else 1
            # two fields read and written (plus two compensations for the
            # Kahan form) once per launch, i.e. once per ``depth`` steps
            nbytes = (2 + 2 + (4 if upd else 0)) * esz / depth
            out[name] = {"ms_per_step": ms / timed, "gcell_updates_per_s": n * n * timed / (ms * 1e-3) / 1e9,
                         "alg_bytes_per_update": nbytes, "steps": steps, "steps_per_launch": depth,
                         "alg_gbs": n * n * nbytes * timed / (ms * 1e-3) / 1e9}
            log(f"gray-scott {name}: {ms / timed * 1e3:.1f} us/step, {out[name]['gcell_updates_per_s']:.1f} "
                f"Gcell-updates/s")
        del f, scratch, g198, g200, stp
        torch.cuda.empty_cache()

    def err(a, b, t):
        return float((a[t] - b[t]).abs().max() / b[t].abs().max())

    for dt in ("float64", "float32"):
        for T in (2, 4):
            name = f"{dt}_T{T}"
            same = all(torch.equal(finals[name][t], finals[dt][t]) for t in ("u", "v"))
            out[name]["state_equals_T1"] = bool(same)
            out[name]["kernel"] = f"coupled-field temporal kernel, {T} fused steps per launch"
            log(f"gray-scott {name}: final state equal to the one-step run: {same}")
    ref = finals["float64"]
    out["float64"]["max_abs_u"] = float(ref["u"].abs().max())
    out["float64"]["max_abs_v"] = float(ref["v"].abs().max())
    for name, same in (("float32", "float64_from_float32_init"), ("float32_kahan", "float64_from_float32_init"),
                       ("bfloat16_kahan", "float64_from_bfloat16_init")):
        for t in ("u", "v"):
            out[name][f"rel_err_{t}_vs_float64"] = err(finals[name], ref, t)
            out[name][f"rel_err_{t}_vs_float64_same_init"] = err(finals[name], finals[same], t)
            out[name][f"abs_err_{t}_vs_float64"] = float((finals[name][t] - ref[t]).abs().max())
            out[name][f"rel_err_{t}_vs_float64_step200"] = err(early[name], early["float64"], t)
        log(f"gray-scott {name}: after {steps} steps rel max err u {out[name]['rel_err_u_vs_float64']:.3e} "
            f"(same init {out[name]['rel_err_u_vs_float64_same_init']:.3e}), v "
            f"{out[name]['rel_err_v_vs_float64']:.3e} (abs {out[name]['abs_err_v_vs_float64']:.2e}; "
            f"max|v| of the float64 run {out['float64']['max_abs_v']:.2e}); at step 200: u "
            f"{out[name]['rel_err_u_vs_float64_step200']:.3e}, v {out[name]['rel_err_v_vs_float64_step200']:.3e}")
    return out


def batched_wave_c4(count=512, steps=200, seed=0, graph=True, plain=True):
    """512 leapfrog wave instances (sides log-uniform in [64, 2048], Gaussian
    pulse, zero Dirichlet): the batched float32 persistent kernel (graph of a
    full buffer rotation) against plain float64 one-launch-per-instance
    (also graph-captured), same step count; error of float32 vs float64.
    ``graph=False`` launches every step eagerly (for profilers whose
    counter collection does not follow graph launches); ``plain=False``
    skips the float64 one-launch-per-instance run (its ~10^5 launches are
    not needed to count the batched kernel's traffic)."""
    import math

    import numpy as np
    import torch

    from cutile import compile
    from cutile.runtime import rt
    from cutile.runtime.batched import BatchedStencil, LeapfrogBatch
    from cutile.runtime.kernel import default_stream
    from cutile.workloads import wave2d, wave_sides

    sides = wave_sides(count, seed=seed)
    rng = np.random.default_rng(seed + 1)
    centres = rng.uniform(0, 1, size=(count, 2))
    cells = sum(s * s for s in sides)

    def pulse(s, c, dtype, halo, cols):
        x = torch.arange(s, dtype=torch.float64, device="cuda")
        sig = 0.05 * s
        g = torch.exp(-(((x - c[0] * s) ** 2)[:, None] + ((x - c[1] * s) ** 2)[None, :]) / (2 * sig * sig))
        a = torch.zeros((s + 2 * halo[0], cols + 2 * halo[1]), dtype=dtype, device="cuda")
        a[halo[0]:halo[0] + s, halo[1]:halo[1] + s] = g.to(dtype)
        return a

    # --- batched float32 --------------------------------------------------
    # tile buckets per size class: the autotuner's measured choice for these
    # sides (the package's tuned table for this arch, or measured here)
    from cutile.runtime.autotune import tuned_buckets

    buckets = tuned_buckets(wave2d("float32"), sides)
    bs = BatchedStencil(wave2d("float32"), buckets=buckets)
    h32 = (1, bs.vz)
    pairs, inter = [], []
    for s, c in zip(sides, centres):
        cols = -(-s // bs.vz) * bs.vz
        u = pulse(s, c, torch.float32, h32, cols)
        pairs.append((u, u.clone()))
        inter.append((s, s))
    lf = LeapfrogBatch(bs, pairs, h32, inter)
    lf.step()
    lf.step()
    lf.step()                                      # JIT + one rotation (back to buffer order 0)
    torch.cuda.synchronize()
    s = default_stream()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        g = rt.Graph(side.cuda_stream) if graph else None
        if graph:
            with g:
                for _ in range(3):
                    lf.step(side.cuda_stream)
        reps, rest = divmod(steps, 3) if graph else (0, steps)
        def run32():
            for _ in range(reps):
                g.launch()
            for _ in range(rest):
                lf.step(side.cuda_stream)
        ms32 = _timed(side.cuda_stream, run32, rt)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    if not plain:
        g32 = cells * steps / (ms32 * 1e-3) / 1e9
        log(f"batched wave: fp32 batched {g32:.1f} Gcell-updates/s (plain float64 run skipped)")
        return {"instances": count, "cells": cells, "steps": steps,
                "batched_fp32": {"ms_per_step": ms32 / steps, "gcell_updates_per_s": g32, "launches_per_step": 1}}
    # --- plain float64, one launch per instance ---------------------------------
    r64 = compile(wave2d("float64"))
    k64 = r64.load_module()._KERNEL
    h64 = r64.halo_widths
    bufs = []
    for sd, c in zip(sides, centres):
        u = pulse(sd, c, torch.float64, h64, sd)
        bufs.append([u, u.clone(), torch.zeros_like(u)])
    rots = [(0, 1, 2), (2, 0, 1), (1, 2, 0)]

    def step64(k, stream):
        i, j, o = rots[k % 3]
        for b in bufs:
            k64.launch([b[i], b[j]], [b[o]], h64, stream, validate=False)

    for k in range(3):
        step64(k, s)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        g64 = rt.Graph(side.cuda_stream) if graph else None
        if graph:
            with g64:
                for k in range(3):
                    step64(k, side.cuda_stream)
        def run64():
            for _ in range(reps):
                g64.launch()
            for k in range(rest):
                step64(k, side.cuda_stream)
        ms64 = _timed(side.cuda_stream, run64, rt)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    # both did 3 + steps steps; compare u of every instance
    err = 0.0
    for idx, (sd, b) in enumerate(zip(sides, bufs)):
        u32 = lf.current(idx)[0][1:1 + sd, h32[1]:h32[1] + sd].to(torch.float64)
        r = (3 + steps) % 3
        u64 = b[rots[r][0]][1:1 + sd, h64[1]:h64[1] + sd]
        den = float(u64.abs().max())
        err = max(err, float((u32 - u64).abs().max()) / den if den > 0 else 0.0)
    g32 = cells * steps / (ms32 * 1e-3) / 1e9
    g64v = cells * steps / (ms64 * 1e-3) / 1e9
    out = {"instances": count, "cells": cells, "steps": steps,
           "batched_fp32": {"ms_per_step": ms32 / steps, "gcell_updates_per_s": g32, "launches_per_step": 1},
           "plain_fp64": {"ms_per_step": ms64 / steps, "gcell_updates_per_s": g64v, "launches_per_step": count},
           "speedup": g32 / g64v, "rel_err_fp32_vs_fp64": err,
           "buckets": [[b.mapping, b.ty, b.tz, b.xch, b.max_cols] for b in bs.buckets],
           "buckets_from": "autotuner" if buckets is not None else "defaults"}
    log(f"batched wave: fp32 batched {g32:.1f} vs fp64 plain {g64v:.1f} Gcell-updates/s "
        f"= {g32 / g64v:.2f}x, rel err {err:.2e}")
    return out


def serving_c5(seconds=60.0, rate=200.0, seed=0):
    """The serving stream (Poisson 200 jobs/s, mixed heat / 3-D diffusion /
    Gray-Scott, 80 % float32) for ``seconds`` of arrivals: p50/p99 latency
    and aggregate Gcell-updates/s; every class's kernels are built before
    the clock starts."""
    from cutile.serving import StencilServer, generate_jobs, warmup_jobs

    srv = StencilServer()
    srv.run(warmup_jobs(srv), realtime=False)
    jobs = generate_jobs(rate, seconds, seed)
    stats = srv.run(jobs, realtime=True, log=log)
    stats.update(rate=rate, seconds=seconds, offered_jobs=len(jobs),
                 offered_gcell_updates_per_s=sum(j.cells * j.steps for j in jobs) / seconds / 1e9)
    log(f"serving {seconds:.0f}s @ {rate:.0f} jobs/s: {stats['jobs']} jobs, p50 {stats['p50_ms']:.1f} ms, "
        f"p99 {stats['p99_ms']:.1f} ms, {stats['gcell_updates_per_s']:.1f} Gcell-updates/s "
        f"(offered {stats['offered_gcell_updates_per_s']:.1f})")
    return stats


def cpu_baseline(n=256, steps=32):
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


def _cpu_c4_instance(args):
    """One wave instance of the batched configuration, float64, on the NumPy
    oracle: (cell-updates, seconds of the step loop)."""
    side, cx, cy, steps = args
    import numpy as np

    import oracle
    from cutile.workloads import dirichlet_spec, wave2d

    x = np.arange(side, dtype=np.float64)
    sig = 0.05 * side
    u = np.zeros((side + 2, side + 2))
    u[1:-1, 1:-1] = np.exp(-(((x - cx * side) ** 2)[:, None] + ((x - cy * side) ** 2)[None, :]) / (2 * sig * sig))
    t0 = time.perf_counter()
    oracle.leapfrog(u, u.copy(), wave2d("float64")._fn, 2, (1, 1), steps, dirichlet_spec(2))
    return side * side * steps, time.perf_counter() - t0


def _cpu_c5_job(args):
    """One job of the serving stream, float64, on the NumPy oracle (its own
    seeded initial field): (cell-updates, seconds of the step loop)."""
    kind, side, steps, seed = args
    from cutile import workloads as w

    import oracle

    if kind == "heat2d":
        u = w.heat2d_init(side, (1, 1), seed=seed)
        t0 = time.perf_counter()
        oracle.time_march(u, w.heat2d()._fn, 2, (1, 1), steps, w.dirichlet_spec(2))
        cells = side * side
    elif kind == "diffusion3d":
        u = w.sinusoid_init(side, (1, 1, 1), seed=seed)
        t0 = time.perf_counter()
        oracle.time_march(u, w.diffusion7(coef=0.1)._fn, 3, (1, 1, 1), steps, w.dirichlet_spec(3))
        cells = side ** 3
    else:
        u, v = w.gray_scott_init(side, (1, 1), seed=seed)
        t0 = time.perf_counter()
        oracle.time_march_multi({"u": u, "v": v}, [f._fn for f in w.gray_scott()], ["u", "v"], 2, (1, 1), steps,
                                w.periodic_spec(2))
        cells = side * side
    return cells * steps, time.perf_counter() - t0


def cpu_pools(c4_steps=8, c5_jobs=96, c5_steps=3):
    """The NumPy oracle over independent work on a process pool (all cores of
    the affinity set, at most 16): every instance of the batched wave
    configuration for ``c4_steps`` steps, and the first ``c5_jobs`` jobs of
    the serving stream for ``c5_steps`` steps each.  Run before the process
    touches the GPU (the workers are forked).  Per-core rates come from the
    summed task times."""
    import multiprocessing as mp

    import numpy as np

    sys.path.insert(0, ROOT)
    from cutile.serving import generate_jobs
    from cutile.workloads import wave_sides

    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    sides = wave_sides(512, seed=0)
    centres = np.random.default_rng(1).uniform(0, 1, size=(512, 2))
    c4 = [(int(sd), float(c[0]), float(c[1]), c4_steps) for sd, c in zip(sides, centres)]
    c4.sort(key=lambda t: -t[0])                      # largest first: an even finish
    c5 = [(j.kind, j.side, c5_steps, j.id) for j in generate_jobs(200.0, 60.0, 0)[:c5_jobs]]
    c5.sort(key=lambda t: -(t[1] ** (3 if t[0] == "diffusion3d" else 2)))
    out = {}
    with mp.get_context("fork").Pool(cores) as pool:
        for name, fn, tasks, what in (("c4", _cpu_c4_instance, c4, f"512 wave instances x {c4_steps} steps, float64"),
                                      ("c5", _cpu_c5_job, c5, f"first {c5_jobs} jobs of the 60-s stream x "
                                                              f"{c5_steps} steps, float64")):
            t0 = time.perf_counter()
            res = pool.map(fn, tasks, chunksize=1)
            wall = time.perf_counter() - t0
            upd = sum(r[0] for r in res)
            busy = sum(r[1] for r in res)
            out[name] = {"value": upd / wall / 1e9, "unit": "Gcell-updates/s", "cores": cores, "kind": "port",
                         "per_core": upd / busy / 1e9, "seconds": wall, "sample": what}
            log(f"cpu pool {name}: {out[name]['value']:.3f} Gcell-updates/s on {cores} cores "
                f"({out[name]['per_core']:.3f} per core), {wall:.1f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--core", action="store_true",
                    help="extras limited to the other 3-D sweeps and the copy reference (for profiling)")
    ap.add_argument("--only", choices=["temporal", "heat2d", "gray_scott", "batched", "serving"],
                    help="the headline plus this one extra section (for profiling)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched section: eager launches instead of graph replays (for counter collection)")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched section: skip the float64 one-launch-per-instance run (for counter collection)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py drives one GPU: this build runs every workload on a single MI355X "
              "(no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # the CPU pools fork their workers: before anything touches the GPU
    pools = None if args.no_cpu or args.no_extra or args.core or args.only else cpu_pools()

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
    if args.only:
        args.no_extra = True
        sect = {"temporal": ("temporal_depths_c2", lambda: temporal_depths(interior, n)),
                "heat2d": ("heat2d_c1", heat2d_c1), "gray_scott": ("gray_scott_c3", gray_scott_c3),
                "batched": ("batched_wave_c4", lambda: batched_wave_c4(graph=not args.no_graph,
                                                                       plain=not args.no_plain)),
                "serving": ("serving_c5", serving_c5)}
        key, fn = sect[args.only]
        extra[key] = fn()
    if not args.no_extra:
        extra["diffusion7_fp64"] = measure("diffusion7_fp64", diffusion7("float64"), interior, n, args.steps,
                                           args.warmup, "f64", spec)
        extra["star25_fp32"] = measure("star25_fp32", star25("float32"), interior, n, args.steps,
                                       args.warmup, "f32", spec)
        extra["copy_ceiling"] = copy_ceiling(head["alg_bytes_per_launch"] // 2)
        log("copy ceiling: " + ", ".join(f"{k} {v['gbs']:.0f} GB/s" for k, v in extra["copy_ceiling"].items()
                                         if isinstance(v, dict)))
        for key, fn in (() if args.core else (("temporal_depths_c2", lambda: temporal_depths(interior, n)),)):
            try:
                extra[key] = fn()
            except Exception as e:  # an extra must not hide the headline line
                extra[key] = {"error": f"{type(e).__name__}: {e}"}
                log(f"{key} failed: {e}")
    del interior
    torch.cuda.empty_cache()
    if not args.no_extra and not args.core:
        for key, fn in (("heat2d_c1", heat2d_c1), ("gray_scott_c3", gray_scott_c3),
                        ("batched_wave_c4", batched_wave_c4), ("serving_c5", serving_c5)):
            try:
                extra[key] = fn()
            except Exception as e:
                extra[key] = {"error": f"{type(e).__name__}: {e}"}
                log(f"{key} failed: {e}")
            torch.cuda.empty_cache()

    if pools is not None:
        extra["cpu_pools"] = pools
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
