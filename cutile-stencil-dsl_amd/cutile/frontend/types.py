"""Boundary condition types of the user API (upstream ``cutile/frontend/types.py:10-65``)."""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Union


class BoundaryType(Enum):
    DIRICHLET = "dirichlet"
    NEUMANN = "neumann"
    PERIODIC = "periodic"
    REFLECTING = "reflecting"


@dataclass(frozen=True)
class BoundaryCondition:
    bc_type: BoundaryType
    value: float = 0.0


@dataclass
class BoundarySpec:
    """Per-dimension ``(low, high)`` conditions; a single entry means
    "the same on every side of every dimension"."""

    conditions: list

    @classmethod
    def uniform(cls, bc_type: Union[str, BoundaryType], value: float = 0.0) -> "BoundarySpec":
        if isinstance(bc_type, str):
            bc_type = BoundaryType(bc_type)
        bc = BoundaryCondition(bc_type=bc_type, value=value)
        return cls(conditions=[(bc, bc)])

    @classmethod
    def periodic(cls) -> "BoundarySpec":
        return cls.uniform(BoundaryType.PERIODIC)

    @classmethod
    def dirichlet(cls, value: float = 0.0) -> "BoundarySpec":
        return cls.uniform(BoundaryType.DIRICHLET, value)

    @classmethod
    def neumann(cls) -> "BoundarySpec":
        return cls.uniform(BoundaryType.NEUMANN)

    def to_dict(self, ndim: int) -> dict:
        """Expand to the ``{dim: (low, high)}`` form used by halo fills."""
        conds = self.conditions
        if len(conds) == 1:
            conds = conds * ndim

        def side(c):
            return (c.bc_type.value, float(c.value)) if c.bc_type is BoundaryType.DIRICHLET \
                else c.bc_type.value

        return {d: (side(lo), side(hi)) for d, (lo, hi) in enumerate(conds)}
