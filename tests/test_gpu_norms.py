"""FP64 residual norms (``st_residual_norms``) and the max-norm comparison
(``st_diff_norms``) against NumPy on the same arrays: the maxima exactly,
the 2-norms to 1e-12 relative (NumPy sums in another order); NaN propagates;
an empty interior gives zeros; two calls give the same bits (the reduction
order depends on the shape only)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    import torch

    return torch.from_numpy(a).to("cuda").to(getattr(torch, dtype))


def _host64(t):
    import torch

    return t.to(torch.float64).cpu().numpy()


@pytest.mark.parametrize("shape,halo,dtype", [
    ((37, 45, 70), (4, 4, 16), "float64"),
    ((9, 70), (1, 32), "float32"),
    ((5, 6, 130), (2, 1, 32), "bfloat16"),
    ((3, 2, 70000), (1, 1, 16), "float64"),
    ((300,), (16,), "float64"),
])
def test_residual_norms_match_numpy(shape, halo, dtype):
    from cutile.runtime import rt
    from cutile.runtime.kernel import dtype_code

    print(f"norms {shape} {halo} {dtype}", flush=True)
    full = tuple(n + 2 * h for n, h in zip(shape, halo))
    rng = np.random.default_rng(full[-1])
    a, b = _dev(rng.standard_normal(full), dtype), _dev(rng.standard_normal(full), dtype)
    got = rt.residual_norms(a.data_ptr(), dtype_code(a), b.data_ptr(), dtype_code(b), a.shape, halo)
    inner = tuple(slice(h, h + n) for n, h in zip(shape, halo))
    x, y = _host64(a)[inner], _host64(b)[inner]
    d = x - y
    assert got[0] == np.max(np.abs(d)) and got[1] == np.max(np.abs(y))
    assert abs(got[2] - np.sqrt(np.sum(d * d))) <= 1e-12 * np.sqrt(np.sum(d * d))
    assert abs(got[3] - np.sqrt(np.sum(y * y))) <= 1e-12 * np.sqrt(np.sum(y * y))
    again = rt.residual_norms(a.data_ptr(), dtype_code(a), b.data_ptr(), dtype_code(b), a.shape, halo)
    assert again == got
    mx = rt.diff_norms(a.data_ptr(), dtype_code(a), b.data_ptr(), dtype_code(b), a.shape, halo)
    assert mx == (got[0], got[1])


def test_norms_nan_and_empty():
    import torch

    from cutile.runtime import rt

    a = torch.zeros((6, 40), dtype=torch.float64, device="cuda")
    b = torch.ones_like(a)
    a[3, 20] = float("nan")
    got = rt.residual_norms(a.data_ptr(), rt.ST_F64, b.data_ptr(), rt.ST_F64, a.shape, (1, 16))
    assert all(np.isnan(v) for v in got)
    assert rt.residual_norms(a.data_ptr(), rt.ST_F64, b.data_ptr(), rt.ST_F64, a.shape, (3, 16)) == (0.0,) * 4


def _np_norms(new, old):
    d = new - old
    return (float(np.max(np.abs(d))), float(np.max(np.abs(new))), float(np.sqrt(np.sum(d * d))),
            float(np.sqrt(np.sum(new * new))))


def _close(got, want):
    r = (got.max_abs, got.max_ref, got.l2, got.l2_ref)
    assert r[0] == want[0] and r[1] == want[1], (r, want)
    for a, b in zip(r[2:], want[2:]):
        assert abs(a - b) <= 1e-12 * max(b, 1e-300), (r, want)


def test_stepper_residuals_match_the_oracle(oracle):
    """TimeStepper (temporal blocks, last step single), FusedStepper
    (Gray-Scott blocks) and RK euler: the reported residual is the last
    step's change of the oracle's states."""
    import torch

    from cutile import compile_fused
    from cutile import workloads as w
    from cutile.runtime.launcher import compile_fused_block
    from cutile.runtime.rk_integrator import RKIntegrator
    from cutile.runtime.timestep import FusedStepper, TimeStepper

    print("residuals: heat2d T=4 x 9, gray-scott T=2 x 7, rk euler", flush=True)
    spec = w.dirichlet_spec(2)
    st = TimeStepper(w.heat2d(), spec, temporal_steps=4)
    u0 = w.heat2d_init(64, st.halo, seed=1)
    got = st.march(torch.from_numpy(u0).cuda(), 9, residual=True)
    prev = oracle.time_march(u0, w.heat2d()._fn, 2, st.halo, 8, spec)
    last = oracle.time_march(prev, w.heat2d()._fn, 2, st.halo, 1, spec)
    I = oracle.interior(u0.shape, st.halo)
    assert np.array_equal(got.cpu().numpy()[I], last[I])
    _close(st.last_residual, _np_norms(last[I], prev[I]))

    fns = list(w.gray_scott())
    one = compile_fused(fns)
    blk = compile_fused_block(fns, 2, one, ["u", "v"])
    pspec = w.periodic_spec(2)
    fs = FusedStepper(one, pspec, ["u", "v"], block=blk)
    u, v = w.gray_scott_init(96, one.halo_widths, seed=2, squares=4)
    for a in (u, v):
        oracle.apply_boundary(a, pspec, one.halo_widths)
    f = {"u": torch.from_numpy(u).cuda(), "v": torch.from_numpy(v).cuda()}
    sc = {k: torch.empty_like(t) for k, t in f.items()}
    f, sc = fs.enqueue(f, sc, 7, residual=True)
    ref6 = oracle.time_march_multi({"u": u, "v": v}, [g._fn for g in fns], ["u", "v"], 2,
                                   one.halo_widths, 6, pspec)
    ref7 = oracle.time_march_multi(ref6, [g._fn for g in fns], ["u", "v"], 2, one.halo_widths, 1, pspec)
    I = oracle.interior(u.shape, one.halo_widths)
    for k in ("u", "v"):
        assert np.array_equal(f[k].cpu().numpy()[I], ref7[k][I])
        _close(fs.last_residual[k], _np_norms(ref7[k][I], ref6[k][I]))

    rk = RKIntegrator(w.heat2d(), "euler").compile()
    u = torch.from_numpy(w.heat2d_init(48, rk.stencil_result.halo_widths, seed=3)).cuda()
    before = u.cpu().numpy().copy()
    rk.runner.step(u, 0.1, residual=True)
    after = u.cpu().numpy()
    I = oracle.interior(after.shape, rk.stencil_result.halo_widths)
    _close(rk.runner.last_residual, _np_norms(after[I], before[I]))
