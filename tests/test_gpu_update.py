"""Time-update epilogue fused into the sweep (``update="euler"|"kahan"``):
bit-exact against the oracle's restatement (float32 arithmetic, bf16 stores
rounded to nearest even), and the Kahan state buys back accuracy."""

import numpy as np
import pytest

from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _run(r, fields, steps, spec, storage):
    import torch

    from cutile.runtime.timestep import FusedStepper

    dt = getattr(torch, storage)
    f = {n: to_dev(a.astype(np.float32) if storage != "float64" else a, dt) for n, a in fields.items()}
    scratch = {n: torch.zeros_like(f[n]) for n in r.outputs}
    st = FusedStepper(r, spec, list(r.outputs))
    f, _ = st.enqueue(f, scratch, steps)
    torch.cuda.synchronize()
    return {n: to_host(f[n]).astype(np.float64 if storage == "float64" else np.float32) for n in f}


def _oracle(oracle, fields, rate_fns, steps, spec, halo, mode, storage, dtype):
    f = {n: fields[n].astype(dtype).copy() for n in ("u", "v")}
    if storage == "bfloat16":
        f = {n: oracle.bf16_round(a) for n, a in f.items()}
    comps = {n: np.zeros_like(f[n]) for n in ("u", "v")}
    for _ in range(steps):
        oracle.time_update(f, rate_fns, ["u", "v"], 2, halo, 1.0, mode, comps, storage)
        for n in ("u", "v"):
            oracle.apply_boundary(f[n], spec, halo)
            if mode == "kahan":
                oracle.apply_boundary(comps[n], spec, halo)
    return f, comps


def _init(halo, n=64, seed=0):
    from cutile import workloads as w

    u, v = w.gray_scott_init(n, halo, seed=seed, squares=6)
    oracle_spec = w.periodic_spec(2)
    return u, v, oracle_spec


def test_euler_update_fp64_equals_full_form(oracle):
    from cutile import compile_fused
    from cutile import workloads as w

    print("gray-scott fp64: rates + euler epilogue vs full-form fused stencil, 64^2 x 30", flush=True)
    ru = compile_fused(list(w.gray_scott_rates()), update="euler", dt=w.GS["dt"], targets=["u", "v"])
    full = compile_fused(list(w.gray_scott()))
    assert ru.halo_widths == full.halo_widths
    u, v, spec = _init(ru.halo_widths)
    for a in (u, v):
        oracle.apply_boundary(a, spec, ru.halo_widths)
    got = _run(ru, {"u": u, "v": v}, 30, spec, "float64")
    ref = oracle.time_march_multi({"u": u, "v": v}, [f._fn for f in w.gray_scott()], ["u", "v"], 2,
                                  ru.halo_widths, 30, spec)
    I = oracle.interior(u.shape, ru.halo_widths)
    assert np.array_equal(got["u"][I], ref["u"][I]) and np.array_equal(got["v"][I], ref["v"][I])


@pytest.mark.parametrize("storage,mode", [("float32", "kahan"), ("bfloat16", "kahan"), ("bfloat16", "euler"),
                                          ("float32", "euler")])
def test_update_modes_match_oracle_bit_for_bit(oracle, storage, mode):
    from cutile import compile_fused
    from cutile import workloads as w

    print(f"gray-scott {storage} {mode}: 64^2 x 25 vs oracle", flush=True)
    rates = w.gray_scott_rates("float32")
    r = compile_fused(list(rates), update=mode, dt=1.0, targets=["u", "v"],
                      storage=None if storage == "float32" else storage)
    u, v, spec = _init(r.halo_widths, seed=2)
    for a in (u, v):
        oracle.apply_boundary(a, spec, r.halo_widths)
    fields = {"u": u, "v": v}
    if mode == "kahan":
        fields.update({"c_u": np.zeros_like(u), "c_v": np.zeros_like(v)})
    got = _run(r, fields, 25, spec, storage)
    ref, comps = _oracle(oracle, {"u": u, "v": v}, [f._fn for f in rates], 25, spec, r.halo_widths, mode,
                         storage, np.float32)
    I = oracle.interior(u.shape, r.halo_widths)
    for n in ("u", "v"):
        assert np.array_equal(got[n][I], ref[n][I]), n
        if mode == "kahan":
            assert np.array_equal(got["c_" + n][I], comps[n][I]), "c_" + n


def test_kahan_bf16_is_closer_to_fp64_than_plain_bf16(oracle):
    from cutile import compile_fused
    from cutile import workloads as w

    print("gray-scott bf16 kahan vs plain: error against fp64 after 300 steps, 96^2", flush=True)
    rates = w.gray_scott_rates("float32")
    rk = compile_fused(list(rates), update="kahan", dt=1.0, targets=["u", "v"], storage="bfloat16")
    rp = compile_fused(list(rates), update="euler", dt=1.0, targets=["u", "v"], storage="bfloat16")
    r64 = compile_fused(list(w.gray_scott()))
    u, v, spec = _init(rk.halo_widths, n=96, seed=4)
    for a in (u, v):
        oracle.apply_boundary(a, spec, rk.halo_widths)
    steps = 300
    gk = _run(rk, {"u": u, "v": v, "c_u": np.zeros_like(u), "c_v": np.zeros_like(v)}, steps, spec, "bfloat16")
    gp = _run(rp, {"u": u, "v": v}, steps, spec, "bfloat16")
    u64, v64, _ = _init(r64.halo_widths, n=96, seed=4)
    for a in (u64, v64):
        oracle.apply_boundary(a, spec, r64.halo_widths)
    ref = oracle.time_march_multi({"u": u64, "v": v64}, [f._fn for f in w.gray_scott()], ["u", "v"], 2,
                                  r64.halo_widths, steps, spec)
    I = oracle.interior(u.shape, rk.halo_widths)
    I64 = oracle.interior(u64.shape, r64.halo_widths)
    ek = max(oracle.rel_max_err(gk[n][I], ref[n][I64]) for n in ("u", "v"))
    ep = max(oracle.rel_max_err(gp[n][I], ref[n][I64]) for n in ("u", "v"))
    print(f"bf16 kahan err {ek:.3e}  bf16 plain err {ep:.3e}", flush=True)
    # measured on MI355X: kahan 2.0e-2, plain 2.3e-1 (the bf16 inputs alone
    # are 2^-9 off); the compensated state must stay several times closer
    assert ek < ep / 5 and ek < 5e-2
