"""Parity at the benchmark configurations' full sizes (``-m gpu``).

The 25-point star (C2) and the 7-point star at 512^3, periodic, as the
benchmark runs them (wrapped sweeps, and the register/LDS temporal kernel
at depth 2), and fused Gray-Scott (C3) at 4096^2 in float64: a few steps
each, compared bit for bit with the oracle at the same size.  The oracle
needs a few seconds per 512^3 step, so the step counts are small; the
benchmark line itself checks 48 steps of every temporal depth against depth
1 on the device."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _march_vs_oracle(oracle, fn, n3, steps, T):
    import torch

    from cutile import workloads as w
    from cutile.runtime.timestep import TimeStepper

    spec = w.periodic_spec(len(n3))
    st = TimeStepper(fn, spec, temporal_steps=T, split="fit")
    h = st.halo
    inter = np.random.default_rng(11).random(n3)
    u = st.empty_like_interior(n3)
    inner = tuple(slice(b, b + a) for a, b in zip(n3, h))
    u[inner] = torch.from_numpy(inter).cuda()
    got = st.march(u, steps)[inner].cpu().numpy()
    del u
    uo = np.zeros(tuple(a + 2 * b for a, b in zip(n3, h)))
    uo[inner] = inter
    oracle.apply_boundary(uo, spec, h)
    ref = oracle.time_march(uo, fn._fn, len(n3), h, steps, spec)[inner]
    return got, ref


@pytest.mark.parametrize("which,T,steps", [("star25", 1, 2), ("star25", 2, 2), ("diffusion7", 2, 4)])
def test_512_cube_periodic(oracle, which, T, steps):
    from cutile import workloads as w

    _flush(f"full size: {which} 512^3 periodic, depth {T}, {steps} steps")
    got, ref = _march_vs_oracle(oracle, getattr(w, which)(), (512, 512, 512), steps, T)
    assert np.array_equal(got, ref)


def test_gray_scott_4096_float64(oracle):
    import torch

    from cutile import compile_fused
    from cutile import workloads as w
    from cutile.runtime.timestep import FusedStepper

    _flush("full size: Gray-Scott 4096^2 float64, 4 steps")
    r = compile_fused(list(w.gray_scott("float64")))
    h = r.halo_widths
    spec = w.periodic_spec(2)
    u0, v0 = w.gray_scott_init(4096, h, seed=0)
    f = {"u": torch.from_numpy(u0).cuda(), "v": torch.from_numpy(v0).cuda()}
    scratch = {t: torch.empty_like(f[t]) for t in ("u", "v")}
    stp = FusedStepper(r, spec, ["u", "v"])
    for t in ("u", "v"):
        oracle.apply_boundary(u0 if t == "u" else v0, spec, h)
    from cutile.runtime import boundary

    for t in ("u", "v"):
        boundary.fill(f[t], spec, h)
    stp.enqueue(f, scratch, 4)
    torch.cuda.synchronize()
    fields = {"u": u0, "v": v0}
    fns = [g._fn for g in w.gray_scott("float64")]
    for _ in range(4):
        new = oracle.apply_multi(fields, fns, 2, h, base=[fields["u"], fields["v"]])
        fields = {"u": new[0], "v": new[1]}
        for t in ("u", "v"):
            oracle.apply_boundary(fields[t], spec, h)
    inner = oracle.interior(u0.shape, h)
    for t in ("u", "v"):
        assert np.array_equal(f[t].cpu().numpy()[inner], fields[t][inner]), t
