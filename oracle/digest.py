"""Interior checksum used to pin full-size device states against the oracle.

Test infrastructure only (like the rest of ``oracle/``): nothing in the
product imports it.  ``field_digest`` is the NumPy statement of the
runtime's ``st_field_digest``: the sum modulo 2**64, over interior cells e
in row-major order, of ``mix64(bits(x_e) ^ (e * 0x9e3779b97f4a7c15))``,
where ``bits`` are the stored bits of the element (float64, float32 or
bfloat16, zero-extended to 64 bits) and ``mix64`` is the finaliser of the
SplitMix64 generator (Steele, Lea and Flood, OOPSLA 2014).  The sum does not
depend on the order in which cells are added, so the device may reduce in
any order.  Pinned in ``tests/test_digest.py``: SplitMix64's published
first output for seed 0, a scalar loop, chunk additivity, and sensitivity to
a one-bit change and to swapped cells.
"""

from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= _M1
        z ^= z >> np.uint64(27)
        z *= _M2
        z ^= z >> np.uint64(31)
    return z


def stored_bits(x: np.ndarray) -> np.ndarray:
    """The stored bits of every element, zero-extended to uint64.  bfloat16
    arrays are passed as their uint16 bit patterns."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return x.view(np.uint64).ravel()
    if x.dtype == np.float32:
        return x.view(np.uint32).ravel().astype(np.uint64)
    if x.dtype == np.uint16:
        return x.ravel().astype(np.uint64)
    raise TypeError(f"unsupported dtype {x.dtype}")


def field_digest(interior: np.ndarray, chunk: int = 1 << 24) -> int:
    """Digest of an interior array (halo already removed)."""
    bits = stored_bits(interior)
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for lo in range(0, bits.size, chunk):
            b = bits[lo:lo + chunk]
            e = np.arange(lo, lo + b.size, dtype=np.uint64)
            total += mix64(b ^ (e * GOLDEN)).sum(dtype=np.uint64)
    return int(total)
