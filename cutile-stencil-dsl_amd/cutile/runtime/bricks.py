"""Flat <-> brick-major layout on the device (``st_bricks``), the
conversion upstream performs with CuPy pad/reshape/transpose
(``multigpu_emitter.py:370-434``): a flat array of shape ``s`` becomes
``(nb_0, .., nb_{d-1}, bs, .., bs)`` with ``nb_i = ceil(s_i / bs)``, padding
cells zero."""

from __future__ import annotations

from cutile.runtime import rt


def to_bricks(flat, brick_size: int, stream=None):
    import torch

    from cutile.runtime.kernel import default_stream, dtype_code

    flat = flat.contiguous()
    nb = [-(-s // brick_size) for s in flat.shape]
    out = torch.empty(nb + [brick_size] * flat.dim(), dtype=flat.dtype, device=flat.device)
    rt.bricks(flat.data_ptr(), out.data_ptr(), dtype_code(flat), list(flat.shape), brick_size, True,
              stream if stream is not None else default_stream())
    return out


def from_bricks(bricked, original_shape, brick_size: int, stream=None):
    import torch

    from cutile.runtime.kernel import default_stream, dtype_code

    out = torch.empty(tuple(original_shape), dtype=bricked.dtype, device=bricked.device)
    rt.bricks(out.data_ptr(), bricked.contiguous().data_ptr(), dtype_code(bricked), list(original_shape),
              brick_size, False, stream if stream is not None else default_stream())
    return out
