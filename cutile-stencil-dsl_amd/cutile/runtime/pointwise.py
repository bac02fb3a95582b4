"""JIT element-wise kernels for the time-update algebra.

``PointwiseKernel("x0 + s0 * x1", nin=2)`` builds ``out = E(x0, x1, ...)`` with
runtime scalars ``s0..s7``, evaluated in ``accumulate`` precision (float64 by
default) over every element of the arrays, written back in the storage type.
Operations are applied exactly as written (contraction off), so a float64
update is bit-identical to the NumPy statement it restates."""

from __future__ import annotations

import ctypes
import re
from typing import Sequence

from cutile.runtime import rt

_CT = {"float64": "double", "float32": "float", "bfloat16": "st::bf16_t"}


class PointArgs(ctypes.Structure):
    _fields_ = [("inp", ctypes.c_void_p * 12), ("out", ctypes.c_void_p * 4), ("sc", ctypes.c_double * 8),
                ("n", ctypes.c_int64)]


class PointwiseKernel:
    def __init__(self, exprs: Sequence[str], nin: int, storage: str = "float64",
                 accumulate: str = "float64", name: str = "pw"):
        if isinstance(exprs, str):
            exprs = [exprs]
        self.exprs = list(exprs)
        self.nin = nin
        self.storage = storage
        compute = "float32" if storage == "bfloat16" else storage
        body = []
        for m, e in enumerate(self.exprs):
            e = re.sub(r"\bx(\d+)\b", r"x[\1]", e)
            e = re.sub(r"\bs(\d+)\b", r"A(sc[\1])", e)
            body.append(f"    o[{m}] = {e};")
        import hashlib

        digest = hashlib.sha1(repr((self.exprs, nin, storage, accumulate)).encode()).hexdigest()[:12]
        self.symbol = f"{name}_{digest}"
        self.source = "\n".join([
            '#include "pointwise.h"',
            "struct PW {",
            f"  using S = {_CT[storage]}; using C = {_CT[compute]}; using A = {_CT[accumulate]};",
            f"  static constexpr int NI = {nin}, NO = {len(self.exprs)};",
            "  static __device__ __forceinline__ void eval(const A* x, const double* sc, A* o) {",
            *body,
            "  }",
            "};",
            f'extern "C" __global__ void __launch_bounds__(256) {self.symbol}(st::PointArgs a) '
            "{ st::pointwise_entry<PW>(a); }",
        ])
        self._fn = {}

    def _function(self):
        import torch

        from cutile.runtime.kernel import load_checked

        dev = torch.cuda.current_device()
        if dev not in self._fn:
            # refused before loading when the code object uses scratch,
            # more than 256 VGPR+AGPR or more than 160 KiB of LDS
            _mod, fns = load_checked(self.source, "pointwise.hip", [self.symbol])
            self._fn[dev] = fns[self.symbol]
        return self._fn[dev]

    def __call__(self, outs: Sequence, ins: Sequence, scalars: Sequence[float] = (), stream=None):
        from cutile.runtime.kernel import cu_count, default_stream

        a = PointArgs()
        n = ins[0].numel()
        for k, t in enumerate(ins):
            if t.numel() != n or not t.is_contiguous():
                raise ValueError("pointwise: inputs must be contiguous and of one size")
            a.inp[k] = t.data_ptr()
        for m, t in enumerate(outs):
            a.out[m] = t.data_ptr()
        for k, v in enumerate(scalars):
            a.sc[k] = float(v)
        a.n = n
        grid = max(1, min(-(-n // 256), 16 * cu_count()))
        rt.launch(self._function(), (grid, 1, 1), (256, 1, 1), a,
                  stream if stream is not None else default_stream())
