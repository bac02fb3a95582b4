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
