"""Runge-Kutta time integration around a compiled stencil operator
(upstream ``cutile/runtime/rk_integrator.py:54-232``).

The stage algebra of each method is the upstream one, statement for
statement (euler, midpoint rk2, classic rk4, ssprk3), but each statement is
one fused element-wise HIP kernel (:class:`PointwiseKernel`) accumulating in
float64 -- for float64 fields bit-identical to the NumPy statement, for
float32/bfloat16 fields the time update is accumulated in float64 and
rounded once.  Stage buffers are allocated once per shape and reused.
"""

from __future__ import annotations

import importlib.util
import os
import tempfile
from dataclasses import dataclass, field
from typing import Any

from cutile.runtime.launcher import CompileResult
from cutile.runtime.launcher import compile as stencil_compile

# statement -> (pointwise expression over x0.., scalar builder from dt)
# each statement of rk_integrator.py:183-232 as (expression over x0.., inputs,
# scalars from dt); Python folds the scalar factors before they meet an array
_STMT = {
    "axpy_half": ("x0 + s0 * x1", 2, lambda dt: [0.5 * dt]),              # u + 0.5 * dt * k
    "axpy": ("x0 + s0 * x1", 2, lambda dt: [dt]),                          # u + dt * k
    "rk4_final": ("x0 + s0 * (((x1 + 2.0 * x2) + 2.0 * x3) + x4)", 5,
                  lambda dt: [dt / 6.0]),                                  # u + (dt/6)*(k1+2k2+2k3+k4)
    "ssp2": ("0.75 * x0 + 0.25 * (x1 + s0 * x2)", 3, lambda dt: [dt]),     # 0.75u + 0.25(u1 + dt k2)
    "ssp3": ("s1 * x0 + s2 * (x1 + s0 * x2)", 3,
             lambda dt: [dt, 1.0 / 3.0, 2.0 / 3.0]),                       # u/3 + 2/3 (u2 + dt k3)
}


class RKRunner:
    """Device RK stepping with a compiled stencil ``launch_fn(u, out)``."""

    def __init__(self, stencil_result: CompileResult, method: str):
        from cutile.runtime.pointwise import PointwiseKernel

        self.r = stencil_result
        self.method = method
        st = stencil_result.storage
        self.k = {name: PointwiseKernel(expr, nin, storage=st, name=f"rk_{name}")
                  for name, (expr, nin, _) in _STMT.items()}
        self._bufs: dict = {}

    def _buf(self, u, name):
        import torch

        key = (tuple(u.shape), u.dtype, u.device, name)
        b = self._bufs.get(key)
        if b is None:
            b = torch.zeros_like(u)
            self._bufs[key] = b
        return b

    def _pw(self, name, out, ins, dt):
        self.k[name]([out], ins, _STMT[name][2](dt))

    def step(self, u, dt, launch_fn=None, residual: bool = False):
        """One step in place on ``u``.  With ``residual=True``, u is kept
        before the step and ``self.last_residual`` holds the FP64 norms of
        the step's change (``ResidualNorms``); the call then waits for the
        current stream."""
        if residual:
            from cutile.runtime.norms import residual_norms

            prev = self._buf(u, "prev")
            prev.copy_(u)
            self.step(u, dt, launch_fn)
            self.last_residual = residual_norms(u, prev, self.r.halo_widths)
            return u
        L = launch_fn or self.r.launcher
        m = self.method
        k1 = self._buf(u, "k1")
        L(u, k1)
        if m == "euler":
            self._pw("axpy", u, [u, k1], dt)
            return u
        if m == "rk2":
            u1, k2 = self._buf(u, "u1"), self._buf(u, "k2")
            self._pw("axpy_half", u1, [u, k1], dt)
            L(u1, k2)
            self._pw("axpy", u, [u, k2], dt)
            return u
        if m == "rk4":
            u1, u2, u3 = self._buf(u, "u1"), self._buf(u, "u2"), self._buf(u, "u3")
            k2, k3, k4 = self._buf(u, "k2"), self._buf(u, "k3"), self._buf(u, "k4")
            self._pw("axpy_half", u1, [u, k1], dt)
            L(u1, k2)
            self._pw("axpy_half", u2, [u, k2], dt)
            L(u2, k3)
            self._pw("axpy", u3, [u, k3], dt)
            L(u3, k4)
            self._pw("rk4_final", u, [u, k1, k2, k3, k4], dt)
            return u
        if m == "ssprk3":
            u1, u2 = self._buf(u, "u1"), self._buf(u, "u2")
            k2, k3 = self._buf(u, "k2"), self._buf(u, "k3")
            self._pw("axpy", u1, [u, k1], dt)
            L(u1, k2)
            self._pw("ssp2", u2, [u, u1, k2], dt)
            L(u2, k3)
            self._pw("ssp3", u, [u, u2, k3], dt)
            return u
        raise ValueError(m)

    def solve(self, u0, dt, num_steps, launch_fn=None):
        u = u0.clone()
        for _ in range(num_steps):
            self.step(u, dt, launch_fn)
        return u


@dataclass
class RKCompileResult:
    stencil_result: CompileResult
    method: str
    num_stages: int
    code: str = ""
    _runner: Any = field(default=None, repr=False)

    def emit_to_file(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.code)

    def load_module(self):
        with tempfile.NamedTemporaryFile(suffix=".py", delete=False, mode="w") as f:
            f.write(self.code)
            path = f.name
        spec = importlib.util.spec_from_file_location("_rk", path)
        mod = importlib.util.module_from_spec(spec)
        mod.__dict__["_RESULT"] = self
        spec.loader.exec_module(mod)
        os.unlink(path)
        return mod

    @property
    def runner(self) -> RKRunner:
        if self._runner is None:
            self._runner = RKRunner(self.stencil_result, self.method)
        return self._runner


class RKIntegrator:
    """``RKIntegrator(stencil, method="rk4").compile(domain)`` -> module with
    ``rk_<method>_step(u, dt, launch_fn=launch_<name>)`` and
    ``rk_<method>_solve(u0, dt, num_steps, launch_fn=launch_<name>)``."""

    _STAGES = {"euler": 1, "rk2": 2, "rk4": 4, "ssprk3": 3}

    def __init__(self, stencil_fn, method: str = "rk4"):
        self.stencil_fn = stencil_fn
        self.method = method.lower()
        if self.method not in self._STAGES:
            raise ValueError(f"Unknown RK method {method!r}. Supported: {', '.join(self._STAGES)}")

    def compile(self, domain=None, hw=None) -> RKCompileResult:
        res = stencil_compile(self.stencil_fn, domain=domain, hw=hw, temporal_blocking=False, autotune=False)
        m = self.method
        code = "\n".join([
            f'"""RK ({m}) stepping around the {res.name} stencil (generated)."""',
            "",
            res.code,
            "",
            f"def rk_{m}_step(u, dt, launch_fn=launch_{res.name}):",
            '    """One RK step, in place on device array u."""',
            "    return _RESULT.runner.step(u, dt, launch_fn)",
            "",
            "",
            f"def rk_{m}_solve(u0, dt, num_steps, launch_fn=launch_{res.name}):",
            f'    """Run {m.upper()} for num_steps steps from a copy of u0."""',
            "    return _RESULT.runner.solve(u0, dt, num_steps, launch_fn)",
            "",
        ])
        return RKCompileResult(stencil_result=res, method=m, num_stages=self._STAGES[m], code=code)
