"""The emitted per-cell expression is bit-identical to the oracle (CPU).

The C++ expression the HIP emitter generates is compiled with g++ (contraction
off, SSE2 arithmetic) inside a host harness whose accessor reads the same
shifted cells the device kernel reads, run over every interior cell, and
compared with the oracle's NumPy evaluation of the user's function with
``np.array_equal``.  This pins operator structure, association, constant
folding and constant rounding (float64 exact hex literals; float32 weak-scalar
rounding) of the emitter without a GPU."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from helpers import random_stencil

HARNESS = r"""
#define ST_HD
#include "st_math.h"
#include <cstdint>
namespace st { typedef unsigned short bf16_t; }
"""

RUNNER = r"""
struct Acc_{tag} {{
  using C = P_{tag}::C;
  const C* const* f; long s1, s2, x, y, z;
  template <int F, int DX, int DY, int DZ> C g() const {{
    return f[F][((x + DX) * s1 + (y + DY)) * s2 + (z + DZ)];
  }}
}};
extern "C" void run_{tag}(const void* const* fv, void* const* ov, long s0, long s1, long s2,
                          int h0, int h1, int h2) {{
  using C = P_{tag}::C;
  const C* const* f = reinterpret_cast<const C* const*>(fv);
  C* const* o = reinterpret_cast<C* const*>(ov);
  for (long x = h0; x < s0 - h0; ++x)
    for (long y = h1; y < s1 - h1; ++y)
      for (long z = h2; z < s2 - h2; ++z) {{
        Acc_{tag} a{{f, s1, s2, x, y, z}};
        C out[P_{tag}::NO];
        P_{tag}::eval(a, out);
        for (int m = 0; m < P_{tag}::NO; ++m) o[m][(x * s1 + y) * s2 + z] = out[m];
      }}
}}
"""


def _build(cases, tmp_path):
    from cutile.lowering.hip_emitter import KernelConfig, emit_traits
    from cutile.runtime.rt import CSRC_DIR

    src = [HARNESS]
    for tag, meta, dtype in cases:
        src.append(emit_traits(meta, KernelConfig(1, 64, 1, 64), dtype, dtype, tag))
        src.append(RUNNER.format(tag=tag))
    cpp = tmp_path / "harness.cpp"
    so = tmp_path / "harness.so"
    cpp.write_text("\n".join(src))
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                    f"-I{CSRC_DIR}", str(cpp), "-o", str(so)], check=True)
    return ctypes.CDLL(str(so))


def _run(lib, tag, ndim, arrays, halo, nout, dtype):
    shape = arrays[0].shape
    if ndim == 1:
        s3, h3 = (1, 1, shape[0]), (0, 0, halo[0])
    elif ndim == 2:
        s3, h3 = (shape[0], 1, shape[1]), (halo[0], 0, halo[1])
    else:
        s3, h3 = tuple(shape), tuple(halo)
    outs = [np.zeros(shape, dtype=dtype) for _ in range(nout)]
    fv = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    ov = (ctypes.c_void_p * nout)(*[o.ctypes.data for o in outs])
    getattr(lib, f"run_{tag}")(fv, ov, ctypes.c_long(s3[0]), ctypes.c_long(s3[1]), ctypes.c_long(s3[2]),
                               h3[0], h3[1], h3[2])
    return outs


SHAPES = [(1, 1, 1), (1, 2, 2), (2, 1, 3), (2, 2, 2), (2, 4, 4), (3, 1, 2), (3, 2, 1), (3, 3, 3), (3, 4, 2),
          (3, 1, 4)]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_random_expressions_bit_exact(tmp_path, oracle, dtype):
    from cutile.frontend.ir import find_func
    from cutile.lowering.meta import extract_meta

    seeds = range(40) if dtype == "float64" else range(0, 40, 2)
    cases, stencils = [], {}
    for seed in seeds:
        nd, nf, rad = SHAPES[seed % len(SHAPES)]
        fn, fields, body = random_stencil(seed, nd, nf, rad, [3, 6, 12][seed % 3], dtype=dtype)
        meta = extract_meta(find_func(fn._ir))
        cases.append((f"s{seed}", meta, dtype))
        stencils[seed] = (fn, nd, rad, body, meta)
    lib = _build(cases, tmp_path)
    npdt = np.dtype(dtype).type
    for seed, (fn, nd, rad, body, meta) in stencils.items():
        rng = np.random.default_rng(seed)
        halo = tuple([rad] * nd)
        shape = tuple([9] * (nd - 1) + [17]) if nd > 1 else (40,)
        shape = tuple(s + 2 * h for s, h in zip(shape, halo))
        fields = {f: (rng.random(shape) + 0.5).astype(npdt) for f in meta.input_names}
        got = _run(lib, f"s{seed}", nd, [fields[f] for f in meta.input_names], halo, 1, npdt)[0]
        ref = oracle.eval_multi(fields, fn._fn, nd, halo)
        inner = oracle.interior(shape, halo)
        assert got[inner].dtype == ref.dtype or np.asarray(ref).dtype == npdt
        assert np.array_equal(got[inner], ref), f"seed {seed}: {body}"


def test_workload_expressions_bit_exact(tmp_path, oracle):
    from cutile import workloads as w
    from cutile.frontend.ir import find_func
    from cutile.lowering.meta import extract_fused_meta, extract_meta

    gu, gv = w.gray_scott()
    gu32, gv32 = w.gray_scott("float32")
    items = [
        ("heat", extract_meta(find_func(w.heat2d()._ir)), "float64", [w.heat2d()], 2, (1, 1)),
        ("d7", extract_meta(find_func(w.diffusion7()._ir)), "float64", [w.diffusion7()], 3, (1, 1, 1)),
        ("s25", extract_meta(find_func(w.star25()._ir)), "float64", [w.star25()], 3, (4, 4, 4)),
        ("s25f", extract_meta(find_func(w.star25("float32")._ir)), "float32", [w.star25("float32")], 3, (4, 4, 4)),
        ("gs", extract_fused_meta([gu._ir, gv._ir]), "float64", [gu, gv], 2, (1, 1)),
        ("gsf", extract_fused_meta([gu32._ir, gv32._ir]), "float32", [gu32, gv32], 2, (1, 1)),
        ("wave", extract_meta(find_func(w.wave2d()._ir)), "float32", [w.wave2d()], 2, (1, 1)),
    ]
    lib = _build([(t, m, d) for t, m, d, *_ in items], tmp_path)
    rng = np.random.default_rng(0)
    for tag, meta, dtype, fns, nd, halo in items:
        npdt = np.dtype(dtype).type
        shape = tuple(([12] * (nd - 1) + [21])[k] + 2 * halo[k] for k in range(nd))
        fields = {f: (rng.random(shape) + 0.25).astype(npdt) for f in meta.input_names}
        outs = _run(lib, tag, nd, [fields[f] for f in meta.input_names], halo, len(fns), npdt)
        inner = oracle.interior(shape, halo)
        for fn, got in zip(fns, outs):
            ref = oracle.eval_multi(fields, fn._fn, nd, halo)
            assert np.array_equal(got[inner], ref), tag


def test_constant_folding_follows_python(tmp_path, oracle):
    """``c * (-1/12) * u`` folds ``c*(-1/12)`` in float64 like Python; an
    inexact constant such as 0.1 reaches the kernel as its exact double."""
    from cutile.frontend.decorator import stencil_from_source
    from cutile.frontend.ir import find_func
    from cutile.lowering.meta import extract_meta

    fn = stencil_from_source("""
def fold(u, i):
    return 3.0 * (-1 / 12) * u[i - 1] + 0.1 * u[i] - max(2.0, 3.0) * abs(-0.5) * u[i + 1] + (1 / 3)
""")
    meta = extract_meta(find_func(fn._ir))
    expr = meta.expressions[0]
    assert float.hex(3.0 * (-1 / 12)) in expr
    assert float.hex(0.1) in expr
    assert float.hex(1 / 3) in expr
    assert "st::maxp" not in expr and "st::absp" not in expr  # folded
    lib = _build([("f", meta, "float64")], tmp_path)
    u = np.random.default_rng(1).random(50)
    got = _run(lib, "f", 1, [u], (1,), 1, np.float64)[0]
    ref = oracle.apply_stencil(u, fn._fn, 1, (1,))
    assert np.array_equal(got[1:-1], ref[1:-1])


def test_nan_propagation_in_max_min(tmp_path, oracle):
    from cutile.frontend.decorator import stencil_from_source
    from cutile.frontend.ir import find_func
    from cutile.lowering.meta import extract_meta

    fn = stencil_from_source("""
def mm(u, v, i):
    return max(u[i], v[i]) + min(v[i - 1], u[i + 1])
""")
    meta = extract_meta(find_func(fn._ir))
    lib = _build([("mm", meta, "float64")], tmp_path)
    rng = np.random.default_rng(2)
    u, v = rng.random(30), rng.random(30)
    u[5] = np.nan
    v[9] = np.nan
    got = _run(lib, "mm", 1, [u, v], (1,), 1, np.float64)[0][1:-1]
    ref = oracle.eval_multi({"u": u, "v": v}, fn._fn, 1, (1,))
    assert np.array_equal(got, ref, equal_nan=True)
    assert np.isnan(got[4]) and np.isnan(got[8])
