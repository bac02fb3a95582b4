"""CPU oracle for the stencil hot path -- TEST INFRASTRUCTURE ONLY.

This module is the parity checker.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it; the product
package (``cutile``) never does, and nothing it computes is ever returned by a
product entry point.

It restates, in our own code, the NumPy reference executor of the upstream
project:

* ``ArrayProxy``      follows ``cutile/reference/stencil_ref.py:21-45``: a
  subscript ``u[o0, o1, ...]`` is the interior slice shifted by the offsets,
  ``u[h+o : n-h+o]`` per dimension.
* ``apply_stencil``   follows ``stencil_ref.py:52-87``: copy the array, write
  the interior with the user's function evaluated on proxies (index args 0).
* ``apply_multi``     follows ``launcher.py:134-169`` and
  ``tests/test_cutile_gpu_apps.py:49-77``: several inputs by parameter name,
  unread parameters get zero arrays.
* ``apply_boundary``  follows ``stencil_ref.py:102-201`` (dims in order 0..n-1,
  low side then high side; Dirichlet fills the whole slab, Neumann copies the
  first/last interior layer, periodic wraps, reflecting mirrors).
* ``time_march``      follows ``stencil_ref.py:208-245`` (stencil then BC per
  step, no BC before step 1) but streams instead of keeping history.
* ``array_aware``     follows ``cutile/reference/elementwise.py:17-52``
  (``max``/``min`` become NaN-propagating ``np.maximum``/``np.minimum`` on
  arrays; ``where`` is added to the globals).

How each function is pinned (tests/test_oracle.py): reference known answers
(``tests/test_cutile_new.py:780-847`` -- delta smoothing, constant Laplacian
is 0, identity, averaging of a constant, monotone time march), a scalar
brute-force evaluator over every interior cell (independent of NumPy
slicing), the 3D 7-point unit-impulse closed form (``examples/laplacian_3d.py:
54-57``), closed-form boundary fills on small arrays, and the committed golden
vectors under ``tests/golden/``.
"""

from __future__ import annotations

import builtins
import types
from typing import Callable, Sequence

import numpy as np


# --------------------------------------------------------------------------
# array-aware builtins  (cutile/reference/elementwise.py:17-52)
# --------------------------------------------------------------------------

def _is_array(x) -> bool:
    return isinstance(x, np.ndarray)


def maximum(a, b):
    if _is_array(a) or _is_array(b):
        return np.maximum(a, b)
    return builtins.max(a, b)


def minimum(a, b):
    if _is_array(a) or _is_array(b):
        return np.minimum(a, b)
    return builtins.min(a, b)


def where(cond, a, b):
    if _is_array(cond) or _is_array(a) or _is_array(b):
        return np.where(cond, a, b)
    return a if cond else b


def array_aware(fn: Callable) -> Callable:
    """Return ``fn`` re-bound so that ``max``/``min`` work on arrays and
    ``where`` resolves (``elementwise.py:42-52``)."""
    g = dict(fn.__globals__)
    b = dict(vars(builtins))
    b["max"] = maximum
    b["min"] = minimum
    g["__builtins__"] = b
    g.setdefault("where", where)
    # the product package exports its own where/maximum/minimum; on NumPy
    # arrays they must behave like the oracle's
    for key, impl in (("where", where), ("maximum", maximum), ("minimum", minimum)):
        if getattr(g.get(key), "__module__", "").startswith("cutile"):
            g[key] = impl
    return types.FunctionType(fn.__code__, g, fn.__name__, fn.__defaults__, fn.__closure__)


# --------------------------------------------------------------------------
# evaluation  (stencil_ref.py:21-87)
# --------------------------------------------------------------------------

class ArrayProxy:
    """``proxy[offs]`` is the interior of ``data`` shifted by ``offs``."""

    def __init__(self, data: np.ndarray, halo: Sequence[int]):
        self._data = data
        self._halo = tuple(halo)

    def __getitem__(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        sl = []
        for dim, off in enumerate(key):
            h = self._halo[dim]
            n = self._data.shape[dim]
            if abs(off) > h:
                raise IndexError(f"offset {off} exceeds halo {h} in dim {dim}")
            sl.append(slice(h + off, n - h + off))
        return self._data[tuple(sl)]


def interior(shape: Sequence[int], halo: Sequence[int]) -> tuple:
    return tuple(slice(h, s - h) for h, s in zip(halo, shape))


def apply_stencil(u: np.ndarray, fn: Callable, ndim: int, halo: Sequence[int]) -> np.ndarray:
    """One sweep of a single-input stencil (``stencil_ref.py:52-87``)."""
    res = array_aware(fn)(ArrayProxy(u, halo), *([0] * ndim))
    out = u.copy()
    out[interior(u.shape, halo)] = res
    return out


def stencil_params(fn: Callable, ndim: int) -> list[str]:
    """All array parameter names of ``fn`` (its parameters minus the trailing
    ``ndim`` index variables), as the reference's multi-input validation
    enumerates them (``launcher.py:161``)."""
    import inspect

    return list(inspect.signature(fn).parameters)[: len(inspect.signature(fn).parameters) - ndim]


def eval_multi(fields: dict, fn: Callable, ndim: int, halo: Sequence[int]) -> np.ndarray:
    """Interior result of ``fn`` on named fields; unread parameters are zero
    arrays (``launcher.py:161-163``)."""
    shape = next(iter(fields.values())).shape
    dtype = next(iter(fields.values())).dtype
    params = stencil_params(fn, ndim)
    args = [ArrayProxy(fields[p] if p in fields else np.zeros(shape, dtype), halo) for p in params]
    res = array_aware(fn)(*args, *([0] * ndim))
    if np.isscalar(res) or np.ndim(res) == 0:
        res = np.full(tuple(s - 2 * h for s, h in zip(shape, halo)), res, dtype=dtype)
    return res


def apply_multi(fields: dict, fns: Sequence[Callable], ndim: int, halo: Sequence[int],
                base: dict | None = None) -> list[np.ndarray]:
    """Evaluate every member of a fused group on the same (old) fields and
    return one full array per member; each output's halo is copied from
    ``base[name]`` if given, else from zeros (the device writes only the
    interior)."""
    shape = next(iter(fields.values())).shape
    dtype = next(iter(fields.values())).dtype
    outs = []
    for k, fn in enumerate(fns):
        res = eval_multi(fields, fn, ndim, halo)
        out = (base[k].copy() if base is not None else np.zeros(shape, dtype))
        out[interior(shape, halo)] = res
        outs.append(out)
    return outs


# --------------------------------------------------------------------------
# boundary conditions  (stencil_ref.py:102-201)
# --------------------------------------------------------------------------

def _parse_bc(spec):
    if isinstance(spec, str):
        return spec, 0.0
    if isinstance(spec, tuple):
        return spec[0], float(spec[1])
    raise TypeError(f"unsupported boundary spec {spec!r}")


def _side(u: np.ndarray, dim: int, h: int, kind: str, value: float, low: bool) -> None:
    n = u.shape[dim]
    nd = u.ndim

    def at(i):
        s = [slice(None)] * nd
        s[dim] = i
        return tuple(s)

    if h == 0:
        return
    if kind == "dirichlet":
        u[at(slice(0, h) if low else slice(n - h, n))] = value
    elif kind == "neumann":
        for k in range(h):
            u[at(h - 1 - k if low else n - h + k)] = u[at(h if low else n - h - 1)]
    elif kind == "periodic":
        for k in range(h):
            u[at(k if low else n - h + k)] = u[at(n - 2 * h + k if low else h + k)]
    elif kind == "reflecting":
        for k in range(h):
            u[at(k if low else n - h + k)] = u[at(2 * h - 1 - k if low else n - h - 1 - k)]
    else:
        raise ValueError(f"unknown boundary kind {kind!r}")


def apply_boundary(u: np.ndarray, spec: dict, halo: Sequence[int]) -> None:
    """In-place halo fill; ``spec`` maps dim -> (low, high), each a kind
    string or ``(kind, value)``."""
    for dim in range(u.ndim):
        if dim not in spec:
            continue
        lo, hi = spec[dim]
        lk, lv = _parse_bc(lo)
        hk, hv = _parse_bc(hi)
        _side(u, dim, halo[dim], lk, lv, True)
        _side(u, dim, halo[dim], hk, hv, False)


def uniform_spec(kind: str, ndim: int, value: float = 0.0) -> dict:
    side = (kind, value) if kind == "dirichlet" else kind
    return {d: (side, side) for d in range(ndim)}


# --------------------------------------------------------------------------
# time marching  (stencil_ref.py:208-245, streamed)
# --------------------------------------------------------------------------

def time_march(u0: np.ndarray, fn: Callable, ndim: int, halo: Sequence[int], steps: int,
               boundary_spec: dict | None = None, callback: Callable | None = None) -> np.ndarray:
    u = u0.copy()
    for s in range(steps):
        u = apply_stencil(u, fn, ndim, halo)
        if boundary_spec is not None:
            apply_boundary(u, boundary_spec, halo)
        if callback is not None:
            callback(s + 1, u)
    return u


def time_march_multi(fields: dict, fns: Sequence[Callable], targets: Sequence[str], ndim: int,
                     halo: Sequence[int], steps: int, boundary_spec: dict | None = None,
                     callback: Callable | None = None) -> dict:
    """Fused multi-field stepping (``examples/gray_scott.py:78-94``): every
    member reads the old state, member k's result becomes field
    ``targets[k]``; then the BC is applied to each updated field."""
    f = {k: v.copy() for k, v in fields.items()}
    for s in range(steps):
        outs = apply_multi(f, fns, ndim, halo, base=[f[t] for t in targets])
        for t, o in zip(targets, outs):
            f[t] = o
            if boundary_spec is not None:
                apply_boundary(f[t], boundary_spec, halo)
        if callback is not None:
            callback(s + 1, f)
    return f


def leapfrog(u: np.ndarray, uold: np.ndarray, fn: Callable, ndim: int, halo: Sequence[int],
             steps: int, boundary_spec: dict | None = None) -> tuple:
    """Two-level stepping where ``fn(u, uold, ...)`` returns the new level
    (``examples/wave_2d.py:62-68`` with the update folded into the stencil)."""
    a, b = u.copy(), uold.copy()
    for _ in range(steps):
        new = apply_multi({"u": a, "uold": b}, [fn], ndim, halo, base=[a])[0]
        if boundary_spec is not None:
            apply_boundary(new, boundary_spec, halo)
        a, b = new, a
    return a, b


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 values rounded to the nearest bfloat16 (ties to even, NaN kept
    quiet), returned as float32 -- the rounding of the device's bf16 stores."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def to_storage(x: np.ndarray, storage: str) -> np.ndarray:
    return bf16_round(x) if storage == "bfloat16" else x


def time_update(fields: dict, rate_fns: Sequence[Callable], targets: Sequence[str], ndim: int,
                halo: Sequence[int], dt: float, mode: str = "euler", comps: dict | None = None,
                storage: str | None = None) -> None:
    """One step of ``u_k <- u_k + dt * R_k`` in place on the interiors.

    ``R_k`` are the rate functions evaluated on the old fields (the compute
    type is the arrays' dtype; bfloat16 storage is held as float32 arrays of
    bf16 values).  ``"euler"`` stores ``round(u + dt*R)``; ``"kahan"`` keeps a
    compensation per field, evaluates the rates on the compensated fields
    ``u - c`` (whole arrays, halos included), and updates
    ``y = dt*R - c; t = round(u + y); c = round((t - u) - y); u = t``."""
    shape = next(iter(fields.values())).shape
    dtype = next(iter(fields.values())).dtype
    I = interior(shape, halo)
    dtc = dtype.type(dt)
    seen = dict(fields)
    if mode == "kahan":
        for t in targets:
            seen[t] = fields[t] - comps[t]
    rates = [eval_multi(seen, fn, ndim, halo) for fn in rate_fns]
    for t, R in zip(targets, rates):
        base = fields[t][I].copy()
        if mode == "euler":
            fields[t][I] = to_storage(base + dtc * R, storage)
        elif mode == "kahan":
            y = dtc * R - comps[t][I]
            tn = to_storage(base + y, storage)
            comps[t][I] = to_storage((tn - base) - y, storage)
            fields[t][I] = tn
        else:
            raise ValueError(mode)


def rel_max_err(x: np.ndarray, ref: np.ndarray) -> float:
    """||x - ref||_inf / ||ref||_inf."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = float(np.max(np.abs(ref))) if ref.size else 0.0
    num = float(np.max(np.abs(x - ref))) if ref.size else 0.0
    return num / den if den > 0 else num


# --------------------------------------------------------------------------
# brute-force scalar evaluator (pins apply_stencil independently of NumPy
# slicing: every interior cell evaluated with Python floats)
# --------------------------------------------------------------------------

class _ScalarProxy:
    def __init__(self, data, halo, idx):
        self._d, self._h, self._i = data, halo, idx

    def __getitem__(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        pos = tuple(i + h + o for i, h, o in zip(self._i, self._h, key))
        return float(self._d[pos])


def apply_stencil_scalar(u: np.ndarray, fn: Callable, ndim: int, halo: Sequence[int]) -> np.ndarray:
    out = u.copy()
    g = dict(fn.__globals__)
    g["where"] = lambda c, a, b: a if c else b
    f = types.FunctionType(fn.__code__, g, fn.__name__, fn.__defaults__, fn.__closure__)
    for idx in np.ndindex(*[s - 2 * h for s, h in zip(u.shape, halo)]):
        val = f(_ScalarProxy(u, halo, idx), *([0] * ndim))
        out[tuple(i + h for i, h in zip(idx, halo))] = val
    return out
