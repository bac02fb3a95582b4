"""Runge-Kutta time integration around a compiled stencil operator
(upstream ``cutile/runtime/rk_integrator.py:54-232``)."""

from __future__ import annotations

from dataclasses import dataclass


@dataclass
class RKCompileResult:
    stencil_result: object
    method: str
    num_stages: int
    code: str = ""


class RKIntegrator:
    _STAGES = {"euler": 1, "rk2": 2, "rk4": 4, "ssprk3": 3}

    def __init__(self, stencil_fn, method: str = "rk4"):
        self.stencil_fn = stencil_fn
        self.method = method.lower()
        if self.method not in self._STAGES:
            raise ValueError(f"Unknown RK method {method!r}. Supported: {', '.join(self._STAGES)}")
