"""Frontend, passes and compile API (CPU only; no device calls).

Mirrors the behaviour pinned by the upstream unit tests
(``tests/test_cutile_new.py:42-934``, ``tests/test_fused.py:40-100``,
``tests/test_perf_layout.py``, ``tests/test_expr_precedence.py``) for the
parts this backend keeps: parser legality rules, recorded names, inferred
ndim/order, footprint and roofline numbers, tiling/temporal/boundary
attributes, 128-byte inner-halo padding, fused naming and signatures."""

import ast

import numpy as np
import pytest

from cutile import BoundarySpec, Pipeline, compile, compile_fused, stencil
from cutile.frontend.decorator import stencil_from_source
from cutile.frontend.ir import ApplyOp, find_func
from cutile.frontend.parser import ParseError
from cutile.lowering.normalize import NormalizePass
from cutile.passes import (AnalysisPass, BoundaryPass, RooflinePass, TemporalBlockingPass, TilingPass,
                           VerifyPass, unique_access_count)
from cutile.runtime.launcher import pad_inner_halo


def S(src, **kw):
    return stencil_from_source(src, **kw)


HEAT2D = "def heat2d(u, i, j):\n    return 0.25 * (u[i - 1, j] + u[i + 1, j] + u[i, j - 1] + u[i, j + 1])\n"
LAP3D = ("def lap3d(u, i, j, k):\n    return (u[i-1,j,k] + u[i+1,j,k] + u[i,j-1,k] + u[i,j+1,k]"
         " + u[i,j,k-1] + u[i,j,k+1] - 6.0 * u[i,j,k])\n")


def _attrs(fn, *passes):
    m = fn._ir.clone()
    NormalizePass().apply(None, m)
    for p in passes:
        p.apply(None, m)
    return next(op for op in m.ops if isinstance(op, ApplyOp)).attributes


class TestParser:
    def test_infers_ndim_and_order(self):
        f = S("def lap4(u, i):\n    return -u[i - 2] + 16 * u[i - 1] - 30 * u[i] + 16 * u[i + 1] - u[i + 2]\n")
        fo = find_func(f._ir)
        assert fo.ndim == 1 and fo.order == 4

    def test_decorator_attributes(self):
        f = S(HEAT2D, ndim=2, order=2, dtype="float32")
        assert f._ndim == 2 and f._order == 2 and f._dtype == "float32" and callable(f._fn)

    def test_ir_prints_upstream_op_names(self):
        text = str(S(HEAT2D)._ir)
        assert "cutile_stencil.access" in text and "arith.mulf" in text and "arith.addf" in text

    def test_only_subscripted_parameters_are_arrays(self):
        f = S("def f(u, unused, i):\n    return u[i - 1] + u[i + 1]\n")
        assert find_func(f._ir).arg_names == ["u"]

    def test_arg_names_keep_parameter_order(self):
        f = S("def a(u, v, i, j):\n    return u[i + 1, j] + 0.3 * v[i, j]\n")
        assert find_func(f._ir).arg_names == ["u", "v"]

    def test_accesses_are_unique_in_first_use_order(self):
        f = S("def sq(u, i):\n    return u[i] * u[i] + u[i + 1] - u[i]\n")
        acc = find_func(f._ir).accesses
        assert [a.offsets for a in acc] == [(0,), (1,)]
        assert unique_access_count(find_func(f._ir)) == 2

    def test_closure_and_global_constants(self):
        f = S("def g(u, i):\n    return c * u[i - 1] + d\n", env={"c": 0.5, "d": 2})
        assert find_func(f._ir).constants == {"c": 0.5, "d": 2.0}

    @pytest.mark.parametrize("body,msg", [
        ("u[i * 2]", "offset"),
        ("u[i] ** 2", "Unsupported binary operator"),
        ("sin(u[i])", "Unsupported call"),
        ("u[i] if u[i] > 0 else 0.0", "Unsupported expression"),
        ("u[i] + zz", "Unknown variable"),
        ("0 < u[i] < 1", "chained"),
    ])
    def test_rejects_unsupported_forms(self, body, msg):
        with pytest.raises(ParseError, match=msg):
            S(f"def bad(u, i):\n    return {body}\n")

    def test_inconsistent_dimensionality(self):
        with pytest.raises(ParseError, match="Inconsistent"):
            S("def bad(u, i, j):\n    return u[i, j] + u[i]\n")

    def test_bad_dtype(self):
        with pytest.raises(ParseError, match="dtype"):
            S(HEAT2D, dtype="float16")

    def test_boundary_attr(self):
        f = S(HEAT2D, boundary={"type": "dirichlet", "value": 1.5})
        b = find_func(f._ir).boundary
        assert b.bc_type == "dirichlet" and b.value == 1.5
        with pytest.raises(ValueError):
            S(HEAT2D, boundary="bogus")

    def test_decorator_on_a_real_function(self):
        @stencil
        def heat(u, i):
            return 0.5 * u[i - 1] + 0.5 * u[i + 1]

        assert find_func(heat._ir).func_name == "heat"


class TestPasses:
    def test_halo_widths(self):
        assert _attrs(S(HEAT2D), AnalysisPass())["halo_widths"] == [1, 1]
        f = S("def s(u, i, j, k):\n    return u[i - 3, j, k + 2] + u[i, j + 1, k]\n")
        assert _attrs(f, AnalysisPass())["halo_widths"] == [3, 1, 2]

    def test_roofline_numbers_of_2d_heat(self):
        a = _attrs(S(HEAT2D), AnalysisPass(), RooflinePass())
        assert a["flops"] == 4 and a["unique_loads"] == 4
        assert abs(a["arithmetic_intensity"] - 0.1) < 1e-12
        assert a["bound"] == "memory"

    def test_tiling_shapes_for_gfx950(self):
        t = TilingPass()
        assert t.candidates(3, 7) == [((256, 16, 256), {"threads": 512})]
        assert t.candidates(2, 4) == [((64, 512), {})]
        assert t.candidates(1, 3) == [((512,),  {})]
        assert TilingPass(dtype_bytes=4).candidates(3, 7) == [((256, 16, 256), {"threads": 512})]
        assert TilingPass(dtype_bytes=4).candidates(3, 25)[0] == ((256, 32, 128), {"threads": 512})
        wide = t.candidates(3, 25)
        assert wide[:2] == [((256, 32, 64), {"threads": 512}), ((256, 16, 64), {"threads": 512})]
        assert wide[2][1] == {"threads": 512, "occupancy": 4}
        very = t.candidates(3, 125)
        assert very[0] == ((256, 4, 128), {}) and very[1][1] == {"occupancy": 4}
        assert very[-1] == ((256, 2, 128), {})
        assert t.tier(16) == "narrow" and t.tier(17) == "wide" and t.tier(65) == "very_wide"

    def test_tiling_attribute_and_verify(self):
        a = _attrs(S(LAP3D), AnalysisPass(), TilingPass())
        assert a["tile_sizes"] == [256, 16, 256] and a["threads"] == 512
        m = S(LAP3D)._ir.clone()
        NormalizePass().apply(None, m)
        AnalysisPass().apply(None, m)
        m.ops[0].attributes["tile_sizes"] = [64, 6, 128]
        with pytest.raises(ValueError, match="power of 2"):
            VerifyPass().apply(None, m)

    def test_temporal_pass_grows_with_budget(self):
        f = S("def h(u, i):\n    return 0.25 * u[i - 1] + 0.5 * u[i] + 0.25 * u[i + 1]\n")
        small = _attrs(f, AnalysisPass(), TilingPass(), TemporalBlockingPass(shared_mem_bytes=8192))
        big = _attrs(f, AnalysisPass(), TilingPass(), TemporalBlockingPass(shared_mem_bytes=1 << 20))
        assert big["temporal_steps"] >= small["temporal_steps"] >= 1
        assert big["expanded_halo"] == [big["temporal_steps"]]
        assert big["temporal_buffer_count"] == big["temporal_steps"] + 1

    def test_boundary_pass(self):
        assert _attrs(S(HEAT2D, boundary="dirichlet"), BoundaryPass())["padding_mode"] == "zero"
        assert _attrs(S(HEAT2D, boundary="periodic"), BoundaryPass())["boundary_fn"] == "host_side"

    def test_pipeline(self):
        p = Pipeline.single_gpu()
        assert len(p.passes) == 6
        m = S(HEAT2D)._ir.clone()
        assert p.run(m) is m
        a = m.ops[0].attributes
        assert {"halo_widths", "flops", "tile_sizes", "temporal_steps"} <= set(a)
        assert Pipeline().add(AnalysisPass()).add(RooflinePass()).passes[1].name == "roofline-analysis"


class TestCompileAPI:
    def test_pad_inner_halo(self):
        assert pad_inner_halo((1, 1, 1), 8) == (1, 1, 16)
        assert pad_inner_halo((4, 4, 4), 4) == (4, 4, 32)
        assert pad_inner_halo((2,), 8) == (16,)

    def test_compile_result_fields(self):
        r = compile(S(LAP3D), temporal_blocking=False)
        assert r.name == "lap3d" and r.ndim == 3
        assert r.halo_widths == (1, 1, 16) and r.stencil_halo == (1, 1, 1)
        assert r.tile_sizes == (256, 16, 256) and r.temporal_steps == 1
        assert r.inputs == ("u",)
        assert "flops" in r.analysis and "bound" in r.analysis
        ast.parse(r.code)
        assert "def launch_lap3d(u_in, u_out, stream=None):" in r.code
        assert "HX, HY, HZ = 1, 1, 16" in r.code
        assert 'extern "C" __global__' in r.hip_source and "lap3d_sweep_" in r.hip_source

    def test_align_halo_off(self):
        r = compile(S(LAP3D), align_halo=False)
        assert r.halo_widths == (1, 1, 1)

    def test_boundary_function_is_emitted(self):
        r = compile(S(HEAT2D, boundary="periodic"))
        assert "def apply_boundary_heat2d(u, stream=None):" in r.code

    def test_temporal_launcher(self):
        r = compile(S(HEAT2D), temporal_steps=3)
        ast.parse(r.code)
        assert r.temporal_steps == 3 and "TEMPORAL_STEPS = 3" in r.code
        assert "_KERNEL.launch_temporal([u_in], [u_out]" in r.code
        # register/LDS kernel at the vector width and at scalar width
        assert r.hip_source.count("st::tblock_entry<P_tb3r") == 2
        assert "st::temporal_entry" not in r.hip_source
        r2 = compile(S(HEAT2D), temporal_steps=2, temporal_mode="recompute")
        assert r2.halo_widths == (2, 16) and r2.stencil_halo == (1, 1)

    def test_deep_wide_temporal_falls_back_to_lds_rings(self):
        # 4 levels of a radius-4 star need a 12-cell ghost ring per side: no
        # register/LDS geometry fits, the LDS-ring kernel is emitted instead
        from cutile import workloads as w

        r = compile(w.star25(), temporal_steps=4)
        assert "st::temporal_entry<P_tb4_" in r.hip_source and "tblock" not in r.hip_source

    def test_multi_input_launcher_signature(self):
        r = compile(S("def wave(u, uold, i, j):\n    return 2.0 * u[i, j] - uold[i, j]\n"))
        assert "def launch_wave(u, uold, u_out, stream=None):" in r.code

    @pytest.mark.parametrize("src", [
        "def h1(u, i):\n    return 0.5 * u[i - 1] + 0.5 * u[i + 1]\n", HEAT2D, LAP3D,
        "def w(u, uold, i, j):\n    return 2.0 * u[i, j] - uold[i, j]\n"])
    def test_emitted_module_imports(self, src):
        r = compile(S(src))
        mod = r.load_module()
        assert callable(getattr(mod, f"launch_{r.name}"))
        assert mod._HALO == r.halo_widths
        r3 = compile(S(src), temporal_steps=2) if "uold" not in src else r
        assert callable(getattr(r3.load_module(), f"launch_{r3.name}"))

    def test_emit_to_file(self, tmp_path):
        r = compile(S(HEAT2D))
        p = tmp_path / "k.py"
        r.emit_to_file(str(p))
        assert p.read_text() == r.code

    def test_num_gpus_above_one_is_rejected(self):
        with pytest.raises(NotImplementedError):
            compile(S(HEAT2D), num_gpus=2)

    def test_bfloat16_storage_requires_float32(self):
        with pytest.raises(ValueError):
            compile(S(HEAT2D), storage="bfloat16")
        r = compile(S(HEAT2D, dtype="float32"), storage="bfloat16")
        assert r.storage == "bfloat16" and "st::bf16_t" in r.hip_source

    def test_boundary_spec_expansion(self):
        spec = BoundarySpec.dirichlet(2.0).to_dict(2)
        assert spec == {0: (("dirichlet", 2.0), ("dirichlet", 2.0)), 1: (("dirichlet", 2.0), ("dirichlet", 2.0))}
        assert BoundarySpec.periodic().to_dict(1) == {0: ("periodic", "periodic")}


class TestFused:
    def _pair(self):
        a = S("def a(u, v, i, j):\n    return u[i + 1, j] + 0.3 * v[i, j]\n")
        b = S("def b(v, w, i, j):\n    return 0.3 * v[i, j] - w[i, j - 1]\n")
        return a, b

    def test_result_describes_the_fused_kernel(self):
        a, b = self._pair()
        r = compile_fused([a, b])
        assert r.name == "a_b" and r.inputs == ("u", "v", "w") and r.outputs == ("a", "b")
        assert r.halo_widths == (1, 16) and r.kernel_hints == {}
        assert "def launch_a_b(u, v, w, a_out, b_out, stream=None):" in r.code
        ast.parse(r.code)

    def test_shared_access_loaded_once(self):
        a, b = self._pair()
        r = compile_fused([a, b])
        from cutile.lowering.meta import extract_fused_meta

        meta = extract_fused_meta([a._ir, b._ir])
        assert [x.view_name for x in meta.accesses] == ["u_p1_0", "v_0_0", "w_0_m1"]
        assert "NO = 2" in r.hip_source

    def test_naming_of_many_members(self):
        from cutile.lowering.meta import fused_name

        assert fused_name(["a", "b", "c"]) == "a_b_c"
        assert fused_name(["a", "b", "c", "d"]) == "a_and_3_others"

    def test_temporal_rejected_and_distinct_names(self):
        a, b = self._pair()
        with pytest.raises(NotImplementedError):
            compile_fused([a, b], temporal_steps=2)
        with pytest.raises(ValueError):
            compile_fused([a, a])

    def test_explicit_occupancy(self):
        a, b = self._pair()
        r = compile_fused([a, b], occupancy=8)
        assert r.kernel_hints == {"occupancy": 8}
        assert "amdgpu_waves_per_eu(8)" in r.hip_source


class TestPrecedence:
    """The emitted C++ keeps the source tree (fully parenthesised)."""

    def test_product_of_difference(self):
        from cutile.lowering.meta import extract_meta

        f = S("def f(u, i):\n    return 0.961953 * (-0.125 * abs(u[i + 1] - u[i]) - u[i - 1])\n")
        py = extract_meta(find_func(f._ir)).py_expressions[0]
        assert py.startswith("(0.961953 * (") and py.endswith("))")
        assert "((-0.125) * abs((t_u_p1 - t_u_0)))" in py
