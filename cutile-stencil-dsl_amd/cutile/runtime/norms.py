"""FP64 residual norms of a time step on the device (``st_residual_norms``).

No upstream counterpart: the reference's steppers report nothing; the
stepping API here (``TimeStepper.march``, ``FusedStepper.enqueue``,
``RKRunner.step`` with ``residual=True``) reports the change of the last
step, ``u^{n+1} - u^n``, over the interior, accumulated in float64 in a
fixed order (the same bits on every run)."""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class ResidualNorms:
    """max|u_new - u_old|, max|u_new|, ||u_new - u_old||_2, ||u_new||_2
    over the interior (float64)."""

    max_abs: float
    max_ref: float
    l2: float
    l2_ref: float

    @property
    def rel_max(self) -> float:
        return self.max_abs / self.max_ref if self.max_ref > 0 else self.max_abs

    @property
    def rel_l2(self) -> float:
        return self.l2 / self.l2_ref if self.l2_ref > 0 else self.l2


def residual_norms(new, old, halo, stream=None, no_alloc: bool = False) -> ResidualNorms:
    """Norms of ``new - old`` relative to ``new`` over the interior of two
    device arrays of the same shape (halo ``halo``); synchronizes
    ``stream`` (default: torch's current stream).  ``no_alloc`` uses the
    runtime's fixed scratch (``st_residual_norms_static``) instead of its
    workspace: no device allocation, interiors up to about 1.07e9 cells."""
    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream, dtype_code

    if tuple(new.shape) != tuple(old.shape):
        raise ValueError(f"residual_norms: shapes differ ({tuple(new.shape)} vs {tuple(old.shape)})")
    if len(halo) != new.dim():
        raise ValueError(f"residual_norms: halo {tuple(halo)} does not match a {new.dim()}-D array")
    for t in (new, old):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("residual_norms: arrays must be C-contiguous device tensors")
    s = default_stream() if stream is None else stream
    out = rt.residual_norms(old.data_ptr(), dtype_code(old), new.data_ptr(), dtype_code(new), tuple(new.shape),
                            tuple(halo), s, no_alloc=no_alloc)
    return ResidualNorms(*out)
