"""Shared helpers for the test suite (host <-> device, random stencils)."""

from __future__ import annotations

import numpy as np

def to_dev(a: np.ndarray, dtype=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return t if dtype is None else t.to(dtype)


def to_host(t) -> np.ndarray:
    import torch

    return t.to(torch.float64).cpu().numpy() if t.dtype == torch.bfloat16 else t.cpu().numpy()


FIELD_NAMES = ("u", "v", "w", "s")
INDEX_NAMES = ("i", "j", "k")


def _index(var: str, off: int) -> str:
    return var if off == 0 else f"{var} {'+' if off > 0 else '-'} {abs(off)}"


class _ExprGen:
    """Random expression trees over field accesses.

    The tree is built top down: a node with a budget of n leaves splits it
    at a random point into two operands joined by + or - (sometimes * under
    a parenthesised constant scale), so sums nest at random depths and the
    emitter's parenthesisation and left-to-right folding are exercised.  A
    leaf is one of: a scaled access, a product of two accesses, a quotient
    whose denominator is kept positive, max/min/abs, a where() on a
    comparison, a negation, or a nested difference.  Constants come from
    four families: dyadic (exact in float32), short decimals and reciprocals
    (inexact in float32), and integers."""

    def __init__(self, seed, fields, idx, radius):
        import numpy as np

        self.rng = np.random.default_rng([int(seed), len(fields), len(idx), int(radius), 0x5eed])
        self.fields, self.idx, self.radius = fields, idx, radius

    def pick(self, seq):
        return seq[int(self.rng.integers(len(seq)))]

    def const(self) -> str:
        fam = int(self.rng.integers(4))
        sign = -1 if self.rng.random() < 0.3 else 1
        if fam == 0:
            val = sign * int(self.rng.integers(1, 16)) / float(2 ** int(self.rng.integers(0, 5)))
            return repr(val)
        if fam == 1:
            digits = int(self.rng.integers(1, 5))
            return repr(sign * round(float(self.rng.uniform(0.05, 2.0)), digits))
        if fam == 2:
            return repr(sign / float(self.pick((3, 6, 7, 9, 11, 12, 15))))
        return str(sign * int(self.rng.integers(1, 9)))

    def access(self) -> str:
        f = self.pick(self.fields)
        offs = [int(self.rng.integers(-self.radius, self.radius + 1)) for _ in self.idx]
        return f"{f}[{', '.join(_index(v, o) for v, o in zip(self.idx, offs))}]"

    def leaf(self, depth: int) -> str:
        r = self.rng.random()
        if r < 0.34:
            return f"{self.const()} * {self.access()}"
        if r < 0.44:
            return f"{self.access()} * {self.access()}"
        if r < 0.52:
            return f"{self.access()} / ({self.access()} * {self.access()} + {abs(float(self.const())) + 0.5!r})"
        if r < 0.60:
            return f"{self.pick(('max', 'min'))}({self.access()}, {self.pick((self.access(), self.const()))})"
        if r < 0.67:
            return f"{self.const()} * abs({self.access()} - {self.access()})"
        if r < 0.74:
            op = self.pick(("<", ">", "<=", ">="))
            return f"where({self.access()} {op} {self.access()}, {self.access()}, {self.const()})"
        if r < 0.82:
            return f"-{self.access()}"
        if depth < 2:
            return f"({self.node(2, depth + 1)})"
        return self.access()

    def node(self, n: int, depth: int = 0) -> str:
        if n <= 1:
            return self.leaf(depth)
        k = int(self.rng.integers(1, n))
        left, right = self.node(k, depth), self.node(n - k, depth)
        r = self.rng.random()
        if r < 0.12 and depth < 2:
            return f"{self.const()} * ({left} - {right})"
        op = "+" if r < 0.6 else "-"
        # the right operand of - (and, at random, of +) is parenthesised when
        # it is itself a sum, so the tree is not flattened into one chain
        if k < n - 1 and (op == "-" or self.rng.random() < 0.4):
            right = f"({right})"
        return f"{left} {op} {right}"


def random_body(seed: int, ndim: int, nfields: int, radius: int, nterms: int):
    """Fields, index names and a random expression of about ``nterms``
    leaves over ``nfields`` fields within ``radius`` (``_ExprGen``)."""
    fields, idx = list(FIELD_NAMES[:nfields]), list(INDEX_NAMES[:ndim])
    return fields, idx, _ExprGen(seed, fields, idx, radius).node(nterms)


def random_stencil(seed: int, ndim: int, nfields: int, radius: int, nterms: int, dtype="float64"):
    from cutile.elementwise import where
    from cutile.frontend.decorator import stencil_from_source

    fields, idx, body = random_body(seed, ndim, nfields, radius, nterms)
    src = f"def r{seed}({', '.join(fields)}, {', '.join(idx)}):\n    return ({body})\n"
    return stencil_from_source(src, env={"where": where}, ndim=ndim, dtype=dtype), fields, body
