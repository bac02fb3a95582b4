"""Bricked-layout glue appended to an emitted module when
``compile(..., layout="bricked")`` (upstream
``cutile/lowering/multigpu_emitter.py:370-538``): ``to_bricks``,
``from_bricks`` and ``launch_<name>_bricked``, with the layout conversion
done by the ``st_bricks`` device kernel instead of CuPy reshapes."""

from __future__ import annotations


def bricked_suffix(name: str, brick_size: int) -> str:
    return f'''

from cutile.runtime.bricks import to_bricks, from_bricks  # noqa: E402,F401


def launch_{name}_bricked(u_in, u_out, brick_size={int(brick_size)}, stream=None):
    """Run {name} on a flat array through the brick-major layout: the input
    goes to bricks and back (the round trip the upstream wrapper performs),
    the flat kernel runs, and u_out receives the interior (halo zero)."""
    import torch
    original_shape = tuple(u_in.shape)
    u_flat = from_bricks(to_bricks(u_in, brick_size, stream), original_shape, brick_size, stream)
    out_flat = torch.zeros_like(u_flat)
    launch_{name}(u_flat, out_flat, stream)
    res = from_bricks(to_bricks(out_flat, brick_size, stream), original_shape, brick_size, stream)
    u_out.copy_(res)
'''
