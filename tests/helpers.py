"""Shared helpers for the test suite (host <-> device, random stencils)."""

from __future__ import annotations

import random

import numpy as np

FIELDS = ["a", "b", "c", "d"]
INDEX = ["i", "j", "k"]
KINDS = ["lin", "lin", "lin", "prod", "div", "max", "min", "abs", "where", "nested", "neg"]


def to_dev(a: np.ndarray, dtype=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return t if dtype is None else t.to(dtype)


def to_host(t) -> np.ndarray:
    import torch

    return t.to(torch.float64).cpu().numpy() if t.dtype == torch.bfloat16 else t.cpu().numpy()


def _ix(v, o):
    return v if o == 0 else (f"{v} + {o}" if o > 0 else f"{v} - {-o}")


def random_body(seed: int, ndim: int, nfields: int, radius: int, nterms: int):
    """A random stencil expression: linear terms, products, divisions by a
    positive denominator, max/min/abs, where on a comparison, nested
    differences, negations; constants exact in float32, inexact, or integer."""
    rng = random.Random(seed)
    fields, idx = FIELDS[:nfields], INDEX[:ndim]

    def coef():
        k = rng.random()
        if k < 0.3:
            return repr(rng.choice([0.25, 0.5, 2.0, -0.125, 1.0, 4.0]))
        if k < 0.6:
            return repr(round(rng.uniform(-1, 1), 6))
        if k < 0.75:
            return str(rng.choice([2, 3, -1, 7]))
        return repr(rng.choice([1 / 3, 0.1, -1 / 12, 4 / 105, 0.8]))

    def acc():
        f = rng.choice(fields)
        return f"{f}[{', '.join(_ix(v, rng.randint(-radius, radius)) for v in idx)}]"

    def term(depth=0):
        kind = rng.choice(KINDS)
        if kind == "lin":
            return f"{coef()} * {acc()}"
        if kind == "prod":
            return f"{coef()} * {acc()} * {acc()}"
        if kind == "div":
            return f"{coef()} * {acc()} / ({acc()} + 1.5)"
        if kind == "max":
            return f"{coef()} * max({acc()}, {acc()})"
        if kind == "min":
            return f"{coef()} * min({acc()}, {coef()})"
        if kind == "abs":
            return f"{coef()} * abs({acc()} - {acc()})"
        if kind == "where":
            return f"where({acc()} > {acc()}, {acc()}, {coef()} * {acc()})"
        if kind == "neg":
            return f"-{acc()}"
        if kind == "nested" and depth < 2:
            return f"{coef()} * ({term(depth + 1)} - ({term(depth + 1)}))"
        return f"{coef()} * {acc()}"

    parts = [term() for _ in range(nterms)]
    body = parts[0] + "".join(f" {rng.choice(['+', '-'])} {p}" for p in parts[1:])
    return fields, idx, body


def random_stencil(seed: int, ndim: int, nfields: int, radius: int, nterms: int, dtype="float64"):
    from cutile.elementwise import where
    from cutile.frontend.decorator import stencil_from_source

    fields, idx, body = random_body(seed, ndim, nfields, radius, nterms)
    src = f"def r{seed}({', '.join(fields)}, {', '.join(idx)}):\n    return ({body})\n"
    return stencil_from_source(src, env={"where": where}, ndim=ndim, dtype=dtype), fields, body
