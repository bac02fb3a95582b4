"""Write ``fullsize_digests.json``: interior digests (``oracle.digest``) of
the oracle's runs at the benchmark configurations' full sizes.

* 25-point star (u += 0.05*L8(u)), 512^3, periodic, float64 and float32:
  the digest of the interior after every one of 50 steps (step 0 = the
  initial field).
* Fused Gray-Scott, 4096^2, periodic, 1000 steps: float64 and float32 (u
  and v after every 50 steps), and bfloat16 storage with the
  Kahan-compensated float32 update (u, v and both compensations).

Runs on the CPU with the oracle only (about half an hour); the ``-m gpu``
tests in ``tests/test_gpu_fullsize_digests.py`` replay the same runs on the
device and compare digests.  Usage: python tests/golden/make_fullsize_digests.py
"""

from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "cutile-stencil-dsl_amd"), ROOT, os.path.dirname(HERE)):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import fullsize_cases as fc  # noqa: E402
import oracle  # noqa: E402
from oracle.digest import field_digest  # noqa: E402

OUT = os.path.join(HERE, "fullsize_digests.json")


def log(msg):
    print(f"[golden] {time.strftime('%H:%M:%S')} {msg}", flush=True)


def c2_run(spec: dict, dtype: str) -> list:
    from cutile import workloads as w

    fn = w.star25(dtype)
    halo = (4, 4, 4)
    npdt = np.float64 if dtype == "float64" else np.float32
    u = fc.place(fc.c2_interior(spec), halo, npdt)
    bc = w.periodic_spec(3)
    oracle.apply_boundary(u, bc, halo)
    inner = oracle.interior(u.shape, halo)
    out = [field_digest(u[inner])]
    for s in range(fc.C2_STEPS):
        u = oracle.apply_stencil(u, fn._fn, 3, halo)
        oracle.apply_boundary(u, bc, halo)
        out.append(field_digest(u[inner]))
        log(f"c2 {dtype} step {s + 1}: {out[-1]:016x}")
    return [f"{d:016x}" for d in out]


def c3_run(dtype: str) -> dict:
    from cutile import workloads as w

    halo = (1, 1)
    bc = w.periodic_spec(2)
    u0, v0 = w.gray_scott_init(fc.C3_N, halo, seed=0)
    inner = oracle.interior(u0.shape, halo)
    names = ["u", "v"] if dtype != "bfloat16_kahan" else ["u", "v", "c_u", "c_v"]
    rec = {k: [] for k in names}

    def digest(fields, comps=None):
        for k in ("u", "v"):
            x = fields[k][inner]
            rec[k].append(f"{field_digest(fc.bf16_bits(x) if comps is not None else x):016x}")
        if comps is not None:
            for k in ("u", "v"):
                rec["c_" + k].append(f"{field_digest(fc.bf16_bits(comps[k][inner])):016x}")

    if dtype in ("float64", "float32"):
        npdt = np.float64 if dtype == "float64" else np.float32
        f = {"u": u0.astype(npdt), "v": v0.astype(npdt)}
        for a in f.values():
            oracle.apply_boundary(a, bc, halo)
        digest(f)
        fns = [g._fn for g in w.gray_scott(dtype)]

        def cb(step, fields):
            if step % fc.C3_EVERY == 0:
                digest(fields)
                log(f"c3 {dtype} step {step}: u {rec['u'][-1]} v {rec['v'][-1]}")

        oracle.time_march_multi(f, fns, ["u", "v"], 2, halo, fc.C3_STEPS, bc, callback=cb)
    else:
        f = {k: oracle.bf16_round(a.astype(np.float32)) for k, a in (("u", u0), ("v", v0))}
        comps = {k: np.zeros_like(f[k]) for k in ("u", "v")}
        for a in f.values():
            oracle.apply_boundary(a, bc, halo)
        digest(f, comps)
        rates = [g._fn for g in w.gray_scott_rates("float32")]
        for s in range(fc.C3_STEPS):
            oracle.time_update(f, rates, ["u", "v"], 2, halo, w.GS["dt"], "kahan", comps, "bfloat16")
            for k in ("u", "v"):
                oracle.apply_boundary(f[k], bc, halo)
                oracle.apply_boundary(comps[k], bc, halo)
            if (s + 1) % fc.C3_EVERY == 0:
                digest(f, comps)
                log(f"c3 {dtype} step {s + 1}: u {rec['u'][-1]} v {rec['v'][-1]}")
    return rec


def main():
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    spec = fc.c2_tables()
    data["c2_init"] = spec
    data.setdefault("c2", {})
    data.setdefault("c3", {})
    for dt in ("float64", "float32"):
        if dt not in data["c2"]:
            data["c2"][dt] = c2_run(spec, dt)
            json.dump(data, open(OUT, "w"))
    for dt in ("float64", "float32", "bfloat16_kahan"):
        if dt not in data["c3"]:
            data["c3"][dt] = {"every": fc.C3_EVERY, **c3_run(dt)}
            json.dump(data, open(OUT, "w"))
    log(f"wrote {OUT}")


if __name__ == "__main__":
    main()
