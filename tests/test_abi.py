"""The C-ABI library loads and exports every entry point its header
declares; the emitted HIP sources build for gfx950 with hiprtc (no GPU)."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "stencil_rt.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(st_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    from cutile.runtime import rt

    assert set(_declared()) == set(rt.EXPORTS)


def test_library_exports_every_declared_symbol():
    import ctypes

    from cutile.runtime import rt

    assert rt.available(), "build the runtime first (__graft_entry__.build())"
    lib = ctypes.CDLL(rt.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_cpu_calls_without_a_device():
    from cutile.runtime import rt

    assert rt.device_count() >= 0
    with pytest.raises(rt.RuntimeLibraryError, match="hiprtc"):
        rt.compile_only("this is not C++", ["--offload-arch=gfx950"])


def test_workload_kernels_build_without_scratch():
    from cutile import compile, compile_fused
    from cutile import workloads as w
    from cutile.runtime import regcheck
    from cutile.runtime.kernel import code_object

    results = [compile(w.heat2d()), compile(w.diffusion7()), compile(w.star25()), compile(w.star25("float32")),
               compile(w.wave2d()), compile_fused(list(w.gray_scott()))]
    for r in results:
        res = regcheck.code_object_resources(code_object(r.hip_source, "gfx950"))
        assert res, r.name
        for name, v in res.items():
            assert v.stack_bytes == 0, f"{name} spills"
            assert v.registers <= 256, f"{name} uses {v.registers} VGPRs"
