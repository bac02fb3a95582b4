"""CPU oracle (test infrastructure only; see stencil_oracle.py)."""
from .stencil_oracle import *  # noqa: F401,F403
from .stencil_oracle import (ArrayProxy, apply_stencil, apply_multi, eval_multi, apply_boundary,
                             uniform_spec, time_march, time_march_multi, leapfrog, rel_max_err,
                             apply_stencil_scalar, array_aware, stencil_params, interior)
