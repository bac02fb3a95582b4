"""How the time stepper plans its launches (CPU: compilation only)."""

from cutile import workloads as w
from cutile.runtime.timestep import TimeStepper


def test_deep_block_of_a_wide_star_runs_as_shallower_register_launches():
    # 4 levels of the radius-4 star fit neither registers nor LDS; a block of
    # 4 steps runs as 2 launches of the register/LDS kernel at depth 2
    st = TimeStepper(w.star25(), w.periodic_spec(3), temporal_steps=4)
    assert st.T == 4 and st.block_reps == 2
    assert "tblock" in st._block.variants and st._block.temporal_steps == 2
    st = TimeStepper(w.star25(), w.dirichlet_spec(3), temporal_steps=4)
    assert st.block_reps == 2 and st.policy.mode == "static"


def test_register_kernel_used_directly_when_it_fits():
    st = TimeStepper(w.diffusion7(), w.periodic_spec(3), temporal_steps=4)
    assert st.block_reps == 1 and "tblock" in st._block.variants and st._block.temporal_steps == 4


def test_nonzero_dirichlet_steps_one_at_a_time():
    # intermediate levels of a block would hold 0 outside the interior, not the
    # boundary value
    st = TimeStepper(w.heat2d(), w.dirichlet_spec(2, value=1.5), temporal_steps=4)
    assert st.T == 1 and st.block is None
    st = TimeStepper(w.heat2d(), w.dirichlet_spec(2), temporal_steps=4)
    assert st.T == 4


def test_fused_block_needs_a_closed_system():
    """``compile_fused_block``: one member per input, member m advancing
    input m, depth at least 2; ``compile_fused`` itself keeps upstream's
    one-step restriction."""
    import pytest

    from cutile import compile_fused
    from cutile.runtime.launcher import compile_fused_block

    fns = list(w.gray_scott())
    one = compile_fused(fns)
    blk = compile_fused_block(fns, 2, one, ["u", "v"])
    assert blk.temporal_steps == 2 and blk.inputs == ("u", "v") and blk.halo_widths == one.halo_widths
    assert "tblock" in blk.load_module()._KERNEL.variants
    with pytest.raises(ValueError):
        compile_fused_block(fns[:1], 2, compile_fused(fns[:1]), ["u"])    # gs_u alone: one result, two inputs
    with pytest.raises(ValueError):
        compile_fused_block(fns, 1, one, ["u", "v"])
    with pytest.raises(NotImplementedError):
        compile_fused(fns, temporal_steps=2)


def test_fused_block_rejects_members_out_of_input_order():
    """Members listed as (gs_v, gs_u) merge to inputs (u, v) but advance
    (v, u): the block would feed gs_v's result back as field u, so it is
    refused, and so is a target list whose length does not match."""
    import pytest

    from cutile import compile_fused
    from cutile.runtime.launcher import compile_fused_block

    gs_u, gs_v = w.gray_scott()
    swapped = compile_fused([gs_v, gs_u])
    assert swapped.inputs == ("u", "v") and swapped.outputs == ("gs_v", "gs_u")
    with pytest.raises(ValueError, match="advances field"):
        compile_fused_block([gs_v, gs_u], 2, swapped, ["v", "u"])
    with pytest.raises(ValueError, match="one target field per member"):
        compile_fused_block([gs_u, gs_v], 2, compile_fused([gs_u, gs_v]), ["u"])


def test_fused_stepper_uses_blocks_only_where_they_are_exact():
    import pytest

    from cutile import compile_fused
    from cutile.runtime.launcher import compile_fused_block
    from cutile.runtime.timestep import FusedStepper

    fns = list(w.gray_scott())
    one = compile_fused(fns)
    blk = compile_fused_block(fns, 4, one, ["u", "v"])
    assert FusedStepper(one, w.periodic_spec(2), ["u", "v"], block=blk).T == 4
    assert FusedStepper(one, w.dirichlet_spec(2), ["u", "v"], block=blk).T == 4
    # a refilled (Neumann) halo between steps is not what zero-halo levels compute
    neumann = {0: ("neumann", "neumann"), 1: ("neumann", "neumann")}
    st = FusedStepper(one, neumann, ["u", "v"], block=blk)
    assert st.T == 1 and st._blk is None
    with pytest.raises(ValueError):
        FusedStepper(one, w.periodic_spec(2), ["v", "u"], block=blk)   # results must advance inputs in order
