"""Bricked-layout glue appended to an emitted module when
``compile(..., layout="bricked")`` (upstream
``cutile/lowering/multigpu_emitter.py:370-538``): ``to_bricks``,
``from_bricks``, ``launch_<name>_bricked`` (flat in, flat out, as upstream)
and ``launch_<name>_on_bricks`` (brick-major in and out).  The sweep runs on
the bricks themselves (a brick-major variant of the kernel addresses every
cell through its brick), where upstream converts back to flat to launch;
the conversions are the ``st_bricks`` device kernel."""

from __future__ import annotations


def bricked_suffix(name: str, brick_size: int) -> str:
    return f'''

from cutile.runtime.bricks import to_bricks, from_bricks  # noqa: E402,F401


def launch_{name}_on_bricks(ub_in, ub_out, shape, stream=None):
    """One sweep of {name} on brick-major arrays (``to_bricks`` of flat
    arrays of ``shape``, halo included), read and written in place; only
    interior cells of ub_out are written."""
    _KERNEL.launch_bricks([ub_in], [ub_out], shape, _HALO, stream)


def launch_{name}_bricked(u_in, u_out, brick_size={int(brick_size)}, stream=None):
    """Run {name} on a flat array through the brick-major layout (the
    upstream wrapper's interface): the input goes to bricks, the brick-major
    kernel runs on them, and u_out receives the result (interior; halo and
    padding zero)."""
    import torch
    if brick_size != {int(brick_size)}:
        raise ValueError("this module's brick kernels were built for bricks of {int(brick_size)}")
    original_shape = tuple(u_in.shape)
    ub = to_bricks(u_in, brick_size, stream)
    ob = torch.zeros_like(ub)
    launch_{name}_on_bricks(ub, ob, original_shape, stream)
    u_out.copy_(from_bricks(ob, original_shape, brick_size, stream))
'''
