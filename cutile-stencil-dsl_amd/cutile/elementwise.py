"""Element-wise helpers usable inside ``@stencil`` bodies.

Inside a stencil body the frontend lowers ``max``/``min``/``abs``/``where``
to device operations; these functions are what the names resolve to when
the same Python function runs on host values (upstream
``cutile/reference/elementwise.py:17-36``): scalars use the builtins,
NumPy/torch arrays the NaN-propagating library functions.
"""

from __future__ import annotations

import builtins


def _lib(*args):
    for a in args:
        mod = type(a).__module__.split(".")[0]
        if mod == "numpy":
            import numpy

            return numpy
        if mod == "torch":
            import torch

            return torch
    return None


def maximum(a, b):
    xp = _lib(a, b)
    return builtins.max(a, b) if xp is None else xp.maximum(xp.asarray(a) if xp.__name__ == "numpy" else a, b)


def minimum(a, b):
    xp = _lib(a, b)
    return builtins.min(a, b) if xp is None else xp.minimum(xp.asarray(a) if xp.__name__ == "numpy" else a, b)


def where(cond, a, b):
    xp = _lib(cond, a, b)
    return (a if cond else b) if xp is None else xp.where(cond, a, b)
