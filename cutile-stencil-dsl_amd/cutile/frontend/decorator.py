"""The ``@stencil`` decorator (upstream ``cutile/frontend/decorator.py:16-61``).

The wrapped callable keeps the user's function and gains ``_ir`` (the parsed
module), ``_fn``, ``_ndim``, ``_order``, ``_dtype`` and ``_boundary``.
Both ``@stencil`` and ``@stencil(ndim=..., ...)`` forms are accepted.
"""

from __future__ import annotations

import functools

from cutile.frontend.parser import parse_stencil


def stencil(fn=None, *, ndim=None, order=None, dtype="float64", boundary=None):
    def _wrap(func):
        ir = parse_stencil(func, ndim=ndim, order=order, dtype=dtype, boundary=boundary)

        @functools.wraps(func)
        def wrapper(*args, **kwargs):
            return func(*args, **kwargs)

        wrapper._ir = ir
        wrapper._fn = func
        wrapper._ndim = ndim
        wrapper._order = order
        wrapper._dtype = dtype
        wrapper._boundary = boundary
        return wrapper

    if fn is not None:
        return _wrap(fn)
    return _wrap


_SOURCE_COUNTER = [0]


def stencil_from_source(source: str, name: str = None, env: dict = None, **kwargs):
    """Define a stencil from source text (``def f(u, i): return ...``).

    The text is registered with ``linecache`` under a synthetic file name so
    the parser can read it back like any other function's source."""
    import linecache
    import textwrap

    source = textwrap.dedent(source)
    _SOURCE_COUNTER[0] += 1
    fname = f"<cutile-stencil-{_SOURCE_COUNTER[0]}>"
    linecache.cache[fname] = (len(source), None, source.splitlines(True), fname)
    ns = dict(env or {})
    exec(compile(source, fname, "exec"), ns)
    if name is None:
        import ast

        name = next(n.name for n in ast.parse(source).body if isinstance(n, ast.FunctionDef))
    return stencil(ns[name], **kwargs)
