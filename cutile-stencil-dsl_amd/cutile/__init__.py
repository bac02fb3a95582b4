"""MI355X-native stencil DSL: the upstream ``cutile`` API on a HIP backend.

``@stencil`` functions are parsed into the stencil IR, annotated by the
pass pipeline and lowered by the HIP emitter to gfx950 kernels that are
JIT-built with hiprtc and launched through ``libstencil_rt.so``.
"""

from cutile.frontend.decorator import stencil
from cutile.elementwise import maximum, minimum, where
from cutile.runtime.launcher import compile, compile_fused, CompileResult
from cutile.runtime.pipeline import Pipeline
from cutile.runtime.rk_integrator import RKIntegrator
from cutile.frontend.types import BoundarySpec

__all__ = ["stencil", "compile", "compile_fused", "Pipeline", "RKIntegrator", "BoundarySpec",
           "maximum", "minimum", "where", "CompileResult"]
