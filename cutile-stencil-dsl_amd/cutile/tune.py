"""Measure and store the autotuner's choices for the configured workloads.

``python -m cutile.tune [--export PATH] [--fresh]`` runs, on the current device, every
measurement the runtime looks up in its tuned tables:

* the batched kernel's tile bucket per size class for the 512 wave
  instances (sides log-uniform in [64, 2048]) and for every serving class;
* with ``--serving-solo``, the serving classes' solo thresholds and
  temporal depths (otherwise the hand-set ``cutile.serving.SOLO`` table);
* the depth split of temporal blocks (``TimeStepper(split="auto")``) for
  2-D heat at 1024^2 (T = 4, 10) and the 25-point and 7-point stars at
  512^3 (T = 2, 4, float64 and float32).

Results go to the user's autotune cache; ``--export`` also merges them into
a table file (the package ships ``cutile/tuned/<arch>.json``, written this
way on an MI355X).
"""

from __future__ import annotations

import argparse
import time


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--export", help="write the merged tuned table for this device's arch here")
    ap.add_argument("--serving-solo", action="store_true",
                    help="also measure the serving classes' solo thresholds (used only with this flag)")
    ap.add_argument("--fresh", action="store_true",
                    help="measure every choice again, ignoring the shipped table (after the kernels changed)")
    args = ap.parse_args()

    import torch

    from cutile import workloads as w
    from cutile.runtime.autotune import _arch, export_tuned, tuned_buckets
    from cutile.runtime.timestep import TimeStepper
    from cutile import serving
    from cutile.serving import KINDS, Job, StencilServer

    assert torch.cuda.is_available(), "tuning needs a device"
    from cutile.runtime import autotune

    autotune.VERBOSE = True
    autotune.IGNORE_SHIPPED = bool(args.fresh)
    t0 = time.time()

    def log(msg):
        print(f"[tune {time.time() - t0:6.1f}s] {msg}", flush=True)

    b = tuned_buckets(w.wave2d("float32"), w.wave_sides(512, seed=0))
    log(f"wave2d float32, 512 instances: buckets {b}")
    serving.SOLO_TUNING = bool(args.serving_solo)
    srv = StencilServer()
    for kind in KINDS:
        for dtype in ("float32", "float64"):
            cls = srv._class(Job(-1, kind, dtype, 1, 1))
            log(f"serving {kind} {dtype}: solo from side {cls.solo_min} at depth {cls.depth}; "
                f"buckets {cls.bs.buckets}")
    for fn, spec, n, Ts in ((w.heat2d(), w.dirichlet_spec(2), (1024, 1024), (4, 10)),
                            (w.star25(), w.periodic_spec(3), (512,) * 3, (2, 4)),
                            (w.star25("float32"), w.periodic_spec(3), (512,) * 3, (2, 4)),
                            (w.diffusion7(), w.periodic_spec(3), (512,) * 3, (2, 4)),
                            (w.diffusion7("float32"), w.periodic_spec(3), (512,) * 3, (2, 4))):
        for T in Ts:
            st = TimeStepper(fn, spec, temporal_steps=T, split="auto")
            u = st.empty_like_interior(n)
            # timed on the workload's own field (the temporal kernels' clock
            # depends on the values; see autotune_split)
            inner = tuple(slice(h, h + k) for h, k in zip(st.halo, n))
            init = w.heat2d_init(n[0], (0, 0)) if len(n) == 2 else w.sinusoid_init(n[0], (0, 0, 0), seed=0)
            u[inner] = torch.from_numpy(init).to(u.device).to(u.dtype)
            kern, reps = st.plan_for(u)
            d = 1 if kern is None else T // reps
            log(f"{fn._fn.__name__} {st.storage} {n} T={T}: {T // d} x depth {d}")
            del u, st
            torch.cuda.empty_cache()
    if args.export:
        log(f"exported {export_tuned(_arch(), args.export)}")


if __name__ == "__main__":
    main()
