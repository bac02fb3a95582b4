"""Frontend: ``@stencil`` decorator and AST parser."""
from cutile.frontend.decorator import stencil  # noqa: F401
from cutile.frontend.parser import parse_stencil, ParseError  # noqa: F401
from cutile.frontend.types import BoundarySpec, BoundaryType, BoundaryCondition  # noqa: F401
