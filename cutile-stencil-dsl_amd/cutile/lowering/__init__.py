"""Lowering: metadata extraction, HIP emission, launcher modules."""
from cutile.lowering.emitter import CodeEmitter  # noqa: F401
from cutile.lowering.stencil_to_hip import lower_stencil_to_python, lower_fused_stencils_to_python  # noqa: F401
from cutile.lowering.meta import extract_meta, extract_fused_meta, fused_name  # noqa: F401
from cutile.lowering.hip_emitter import emit_hip_source, KernelConfig  # noqa: F401
