"""Code-object cache keys and the ahead-of-time folder (``-m gpu``).

The runtime's disk cache (C++) and the Python ``cache_key`` must name the
same file, so code objects built ahead of time without a device are found
by launches on one."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_runtime_and_python_cache_keys_agree(tmp_path):
    from cutile import compile
    from cutile import workloads as w
    from cutile.runtime import rt
    from cutile.runtime.kernel import compile_options, current_arch

    r = compile(w.heat2d())
    opts = compile_options(current_arch())
    print(f"cache key check in {tmp_path}", flush=True)
    rt.compile_module(r.hip_source + "\n// key check\n", opts, str(tmp_path), name="k.hip")
    assert os.path.exists(tmp_path / (rt.cache_key(r.hip_source + "\n// key check\n", opts) + ".hsaco"))


def test_launch_uses_a_precompiled_code_object(tmp_path, monkeypatch, oracle):
    import torch

    from cutile import compile
    from cutile import workloads as w
    from cutile.runtime import kernel as K

    monkeypatch.setenv("CUTILE_AMD_AOT", str(tmp_path))
    r = compile(w.diffusion7())
    src = r.hip_source
    path = K.precompile(src, K.current_arch())
    assert os.path.dirname(path) == str(tmp_path)
    print(f"launch from {path}", flush=True)
    u = np.random.default_rng(0).random(tuple(20 + 2 * h for h in r.halo_widths))
    a = torch.from_numpy(u).cuda()
    b = torch.zeros_like(a)
    r.launcher(a, b)
    torch.cuda.synchronize()
    inner = oracle.interior(u.shape, r.halo_widths)
    ref = oracle.apply_stencil(u, w.diffusion7()._fn, 3, r.halo_widths)
    assert np.array_equal(b.cpu().numpy()[inner], ref[inner])
