"""Time-update epilogue: what the emitter writes (CPU only)."""

import pytest

from cutile import compile, compile_fused
from cutile import workloads as w


def test_kahan_bf16_signature_and_traits():
    r = compile_fused(list(w.gray_scott_rates("float32")), update="kahan", dt=1.0, targets=["u", "v"],
                      storage="bfloat16")
    assert r.inputs == ("u", "v", "c_u", "c_v") and r.outputs == ("u", "v", "c_u", "c_v")
    assert "def launch_gs_ru_gs_rv(u, v, c_u, c_v, u_out, v_out, c_u_out, c_v_out, stream=None):" in r.code
    assert "UPD = 2, NOUT = 4" in r.hip_source and "TF[2] = {0, 1}, CF[2] = {2, 3}" in r.hip_source
    assert "using S = st::bf16_t;" in r.hip_source and "using C = float;" in r.hip_source


def test_euler_single_stencil_defaults_to_first_input():
    lap = w.diffusion7("float64")
    r = compile(lap, update="euler", dt=0.5, temporal_blocking=False)
    assert r.outputs == ("u",) and "UPD = 1, NOUT = 1" in r.hip_source
    assert "DT = 0x1.0000000000000p-1;" in r.hip_source
    assert "def launch_diffusion7(u_in, u_out, stream=None):" in r.code


def test_plain_kernels_declare_no_update():
    r = compile(w.heat2d())
    assert "UPD = 0, NOUT = 1" in r.hip_source


def test_bad_update_arguments():
    with pytest.raises(ValueError, match="update target"):
        compile_fused(list(w.gray_scott_rates()), update="euler", targets=["u", "q"])
    with pytest.raises(ValueError, match="update must be"):
        compile_fused(list(w.gray_scott_rates()), update="rk9", targets=["u", "v"])
    with pytest.raises(ValueError, match="targets"):
        compile_fused(list(w.gray_scott_rates()), update="euler", targets=["u"])
