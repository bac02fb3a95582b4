"""Passes over the normalised stencil IR."""
from cutile.passes.analysis.footprint import AnalysisPass  # noqa: F401
from cutile.passes.analysis.roofline import RooflinePass  # noqa: F401
from cutile.passes.tiling import TilingPass, unique_access_count  # noqa: F401
from cutile.passes.temporal import TemporalBlockingPass  # noqa: F401
from cutile.passes.boundary import BoundaryPass  # noqa: F401
from cutile.passes.verify import VerifyPass  # noqa: F401
from cutile.passes.decompose import DecompositionPass  # noqa: F401
from cutile.passes.halo import HaloExchangePass  # noqa: F401
from cutile.passes.bricked import BrickedLayoutPass  # noqa: F401
