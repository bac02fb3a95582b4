"""Full-size parity against the oracle, step by step (``-m gpu``).

The golden digests in ``tests/golden/fullsize_digests.json`` come from the
oracle's runs on the CPU (``tests/golden/make_fullsize_digests.py``):

* 25-point star (u += 0.05*L8(u)), 512^3, periodic, 50 steps, float64 and
  float32: the interior digest after every step.  The device runs the same
  50 steps at temporal depth 1 (one wrapped sweep per step), 2 and 4
  (register/LDS blocks; the last two steps of the depth-4 run are single
  steps) and must match at every step it exposes (every step, every 2nd,
  every 4th then 49 and 50).
* Fused Gray-Scott, 4096^2, periodic, 1000 steps: float64 and float32 at
  depths 1, 2 and 4 (coupled-field blocks), and bfloat16 storage with the
  Kahan float32 update: u and v (and the compensations) after every 50
  steps.

A digest is a position-dependent 64-bit checksum of the stored bits
(``oracle.digest``), so equality here is equality of the whole interior,
bit for bit, at that step."""

import json
import os

import numpy as np
import pytest

import fullsize_cases as fc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fullsize_digests.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


def _digest(t, halo):
    from cutile.runtime import rt
    from cutile.runtime.kernel import dtype_code

    return f"{rt.field_digest(t.data_ptr(), dtype_code(t), tuple(t.shape), tuple(halo)):016x}"


_C2_INIT: dict = {}


def _c2_interior(spec):
    if "u" not in _C2_INIT:
        _C2_INIT["u"] = fc.c2_interior(spec)
    return _C2_INIT["u"]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("T", [1, 2, 4])
def test_star25_512_every_step(dtype, T):
    import torch

    from cutile import workloads as w
    from cutile.runtime.timestep import TimeStepper

    g = _golden()
    want = g["c2"][dtype]
    print(f"star25 512^3 {dtype} T={T}: 50 steps vs golden digests", flush=True)
    st = TimeStepper(w.star25(dtype), w.periodic_spec(3), temporal_steps=T, split="fit")
    n = fc.C2_N
    a = st.empty_like_interior((n, n, n))
    inner = tuple(slice(h, h + n) for h in st.halo)
    a[inner] = torch.from_numpy(_c2_interior(g["c2_init"])).to("cuda").to(a.dtype)
    st.fill(a)
    b = torch.empty_like(a)
    assert _digest(a, st.halo) == want[0], "initial field differs from the golden run's"
    plan = {1: [1] * 50, 2: [2] * 25, 4: [4] * 12 + [1, 1]}[T]
    step = 0
    for k in plan:
        a, b = st.enqueue(a, b, k)
        step += k
        got = _digest(a, st.halo)
        assert got == want[step], f"{dtype} T={T}: step {step} differs from the oracle ({got} != {want[step]})"
    assert step == fc.C2_STEPS
    del a, b
    torch.cuda.empty_cache()


def _c3_setup(dtype, T):
    import torch

    from cutile import compile_fused
    from cutile import workloads as w
    from cutile.runtime import boundary
    from cutile.runtime.launcher import compile_fused_block
    from cutile.runtime.timestep import FusedStepper

    spec = w.periodic_spec(2)
    if dtype == "bfloat16_kahan":
        r = compile_fused(list(w.gray_scott_rates("float32")), update="kahan", dt=w.GS["dt"], targets=["u", "v"],
                          storage="bfloat16")
        targets, blk = list(r.outputs), None
    else:
        fns = list(w.gray_scott(dtype))
        r = compile_fused(fns)
        targets = ["u", "v"]
        blk = compile_fused_block(fns, T, r, targets) if T > 1 else None
    u0, v0 = w.gray_scott_init(fc.C3_N, r.halo_widths, seed=0)
    tdt = getattr(torch, r.storage)
    f = {k: torch.from_numpy(a).to("cuda").to(torch.float32 if dtype != "float64" else torch.float64).to(tdt)
         for k, a in (("u", u0), ("v", v0))}
    if dtype == "bfloat16_kahan":
        f["c_u"], f["c_v"] = torch.zeros_like(f["u"]), torch.zeros_like(f["v"])
    for t in targets:
        boundary.fill(f[t], spec, r.halo_widths)
    scratch = {t: torch.empty_like(f[t]) for t in targets}
    return FusedStepper(r, spec, targets, block=blk), f, scratch, r.halo_widths, targets


@pytest.mark.parametrize("dtype,T", [("float64", 1), ("float64", 2), ("float64", 4), ("float32", 1),
                                     ("float32", 2), ("float32", 4), ("bfloat16_kahan", 1)])
def test_gray_scott_4096_every_50_steps(dtype, T):
    import torch

    want = _golden()["c3"][dtype]
    print(f"gray-scott 4096^2 {dtype} T={T}: 1000 steps vs golden digests", flush=True)
    st, f, scratch, halo, targets = _c3_setup(dtype, T)
    names = [t for t in ("u", "v", "c_u", "c_v") if t in want]
    for k in names:
        assert _digest(f[k], halo) == want[k][0], f"initial {k} differs from the golden run's"
    for i in range(1, fc.C3_STEPS // fc.C3_EVERY + 1):
        f, scratch = st.enqueue(f, scratch, fc.C3_EVERY)
        for k in names:
            got = _digest(f[k], halo)
            assert got == want[k][i], f"{dtype} T={T}: {k} at step {i * fc.C3_EVERY} differs from the oracle"
    del f, scratch
    torch.cuda.empty_cache()
