"""On-device halo fills (the emitted ``apply_boundary_<name>``).

Upstream emits CuPy slice copies per dimension and side
(``cutile/lowering/target_to_python.py:199-297``) with the semantics of
``cutile/reference/stencil_ref.py:148-201``.  Here one fixed HIP kernel
(``st_fill_halo``) fills both sides of a dimension per launch, dimensions in
order 0..n-1 so corners see the earlier dimensions' values.
"""

from __future__ import annotations

from typing import Sequence

from cutile.runtime import rt


def normalize_spec(spec, ndim: int) -> dict:
    """Accept ``"periodic"``, ``{"type": ..., "value": ...}``, a
    ``BoundarySpec`` or the oracle's ``{dim: (low, high)}`` dict."""
    from cutile.frontend.ir import BoundaryAttr
    from cutile.frontend.types import BoundarySpec

    if spec is None:
        return {}
    if isinstance(spec, BoundarySpec):
        return spec.to_dict(ndim)
    if isinstance(spec, BoundaryAttr):
        spec = {"type": spec.bc_type, "value": spec.value}
    if isinstance(spec, str):
        spec = {"type": spec}
    if isinstance(spec, dict) and ("type" in spec or "bc_type" in spec):
        kind = spec.get("type", spec.get("bc_type"))
        val = spec.get("value") or 0.0
        side = (kind, float(val)) if kind == "dirichlet" else kind
        return {d: (side, side) for d in range(ndim)}
    return dict(spec)


def _side(s):
    if isinstance(s, str):
        return rt.BC_CODES[s], 0.0
    return rt.BC_CODES[s[0]], float(s[1])


def encode(spec: dict, ndim: int) -> tuple:
    kinds, values = [], []
    for d in range(ndim):
        if d in spec:
            lo, hi = spec[d]
            (kl, vl), (kh, vh) = _side(lo), _side(hi)
        else:
            kl = kh = 0
            vl = vh = 0.0
        kinds += [kl, kh]
        values += [vl, vh]
    return kinds, values


def fill(u, spec, halo: Sequence[int], stream=None) -> None:
    """Fill the halo of torch device tensor ``u`` in place (asynchronous)."""
    from cutile.runtime.kernel import default_stream, dtype_code

    if not u.is_contiguous():
        raise ValueError("apply_boundary: array must be C-contiguous")
    nd = u.dim()
    kinds, values = encode(normalize_spec(spec, nd), nd)
    if stream is None:
        stream = default_stream()
    rt.fill_halo(u.data_ptr(), dtype_code(u), list(u.shape), list(halo), kinds, values, stream)
