"""The allocation-free residual norms (``st_residual_norms_static``): the
same outputs, bit for bit, as ``st_residual_norms`` on two calls in a row
(the second on a smaller interior, so stale partials of the first call would
show), through the C ABI; interiors beyond the static scratch are refused."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(name, a, b, shape, halo):
    from cutile.runtime import rt

    out = (ctypes.c_double * 4)()
    rc = getattr(rt.lib(), name)(a.data_ptr(), rt.ST_F64, b.data_ptr(), rt.ST_F64, len(shape), rt._i64(shape),
                                 rt._i64(halo), None, out)
    return rc, tuple(out)


def test_static_norms_equal_the_allocating_entry_point():
    import torch

    print("no-alloc residual norms: 2 calls, second smaller", flush=True)
    rng = np.random.default_rng(7)
    for shape, halo in (((130, 200, 96), (4, 4, 16)), ((17, 33, 40), (1, 1, 16))):
        full = tuple(n + 2 * h for n, h in zip(shape, halo))
        a = torch.from_numpy(rng.standard_normal(full)).to("cuda")
        b = torch.from_numpy(rng.standard_normal(full)).to("cuda")
        rc1, want = _call("st_residual_norms", a, b, full, halo)
        rc2, got = _call("st_residual_norms_static", a, b, full, halo)
        assert rc1 == 0 and rc2 == 0
        assert got == want, (shape, got, want)


def test_static_norms_refuse_oversized_interiors():
    import torch

    from cutile.runtime import rt

    # 16385 rows of 65536 cells need 16385 partials: one more than the scratch
    # holds.  The refusal happens on the host before anything is launched;
    # the 8.6 GB array is real, so nothing could be read out of bounds anyway
    a = torch.empty((16385, 65536), dtype=torch.float64, device="cuda")
    rc, _ = _call("st_residual_norms_static", a, a, tuple(a.shape), (0, 0))
    assert rc != 0 and "static scratch" in rt.lib().st_last_error().decode()
    del a
    torch.cuda.empty_cache()
