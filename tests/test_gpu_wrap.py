"""Halo policies of the steppers against the oracle, whole arrays (halo
included), bit for bit: periodic reads wrapped inside the sweep with one
fill at the end ("wrap"), Dirichlet halos filled once ("static"), and the
per-step fill."""

import numpy as np
import pytest

from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _stepper_vs_oracle(oracle, fn, shape_interior, spec, steps, dtype="float64", seed=0, expect=None):
    import torch

    from cutile.runtime.timestep import TimeStepper

    st = TimeStepper(fn, spec)
    if expect is not None:
        assert st.policy.mode == expect
    rng = np.random.default_rng(seed)
    npdt = np.float64 if dtype == "float64" else np.float32
    u0 = np.zeros(tuple(n + 2 * h for n, h in zip(shape_interior, st.halo)), dtype=npdt)
    u0[oracle.interior(u0.shape, st.halo)] = rng.random(shape_interior) + 0.5
    oracle.apply_boundary(u0, spec, st.halo)
    u = to_dev(u0)
    out = st.march(u, steps)
    torch.cuda.synchronize()
    ref = oracle.time_march(u0, fn._fn, len(shape_interior), st.halo, steps, spec)
    got = to_host(out)
    assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
    return st


@pytest.mark.parametrize("shape", [(40, 36, 48), (2, 2, 32), (9, 5, 33)])
def test_wrap_diffusion7_fp64(oracle, shape):
    from cutile import workloads as w

    print(f"wrap diffusion7 fp64 {shape}", flush=True)
    st = _stepper_vs_oracle(oracle, w.diffusion7(), shape, w.periodic_spec(3), 5, expect="wrap")
    assert st.policy.axes_for(st.empty_like_interior(shape)) == (0, 1, 2)


@pytest.mark.parametrize("shape", [(20, 24, 64), (17, 9, 37)])
def test_wrap_star25_fp32(oracle, shape):
    from cutile import workloads as w

    print(f"wrap star25 fp32 {shape}", flush=True)
    _stepper_vs_oracle(oracle, w.star25("float32"), shape, w.periodic_spec(3), 4, dtype="float32", expect="wrap")


def test_periodic_interior_as_narrow_as_the_halo(oracle):
    from cutile import workloads as w

    print("periodic star25 fp64 interior == halo on axis 0 (wrap path)", flush=True)
    st = _stepper_vs_oracle(oracle, w.star25(), (4, 12, 40), w.periodic_spec(3), 3, expect="wrap")
    assert st.policy.axes_for(st.empty_like_interior((4, 12, 40))) == (0, 1, 2)


def test_neumann_uses_a_fill_per_step(oracle):
    from cutile import workloads as w

    print("neumann star25 fp64 (fill path)", flush=True)
    spec = {d: ("neumann", "neumann") for d in range(3)}
    st = _stepper_vs_oracle(oracle, w.star25(), (9, 12, 40), spec, 3, expect="fill")
    assert st.policy.axes_for(st.empty_like_interior((9, 12, 40))) == ()


def test_wrap_heat1d_and_2d(oracle):
    from cutile import workloads as w
    from cutile.frontend.decorator import stencil_from_source

    print("wrap heat 1d / 2d fp64", flush=True)
    h1 = stencil_from_source("def heat1(u, i):\n    return 0.5 * u[i] + 0.25 * (u[i - 1] + u[i + 1])\n", ndim=1)
    _stepper_vs_oracle(oracle, h1, (300,), w.periodic_spec(1), 6, expect="wrap")
    _stepper_vs_oracle(oracle, w.heat2d(), (70, 90), w.periodic_spec(2), 6, expect="wrap")


def test_static_dirichlet_value(oracle):
    from cutile import workloads as w

    print("static dirichlet 0.5 heat2d fp64", flush=True)
    _stepper_vs_oracle(oracle, w.heat2d(), (64, 80), w.dirichlet_spec(2, 0.5), 7, expect="static")


def test_fused_wrap_gray_scott(oracle):
    import torch

    from cutile import compile_fused
    from cutile import workloads as w
    from cutile.runtime.timestep import FusedStepper

    print("wrap fused gray-scott fp64 48x40", flush=True)
    r = compile_fused(list(w.gray_scott()))
    spec = w.periodic_spec(2)
    u, v = w.gray_scott_init(48, r.halo_widths, seed=1, squares=4)
    u = u[:, : 40 + 2 * r.halo_widths[1]].copy()
    v = v[:, : 40 + 2 * r.halo_widths[1]].copy()
    for a in (u, v):
        oracle.apply_boundary(a, spec, r.halo_widths)
    st = FusedStepper(r, spec, ["u", "v"])
    assert st.policy.mode == "wrap"
    f = {"u": to_dev(u), "v": to_dev(v)}
    scratch = {k: torch.zeros_like(t) for k, t in f.items()}
    f, _ = st.enqueue(f, scratch, 9)
    torch.cuda.synchronize()
    ref = oracle.time_march_multi({"u": u, "v": v}, [g._fn for g in w.gray_scott()], ["u", "v"], 2,
                                  r.halo_widths, 9, spec)
    for n in ("u", "v"):
        assert np.array_equal(to_host(f[n]), ref[n]), n
