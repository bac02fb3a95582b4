"""Initial fields of the full-size parity runs (the 25-point star at 512^3
and fused Gray-Scott at 4096^2), shared by the golden-digest generator
(``tests/golden/make_fullsize_digests.py``, run on the CPU with the oracle)
and the ``-m gpu`` tests that replay the same runs on the device.

The 25-point star's field is the benchmark's: three sinusoids with integer
wavevectors |k| <= 8 and random phases plus 1e-3 uniform noise, drawn from
``numpy.random.default_rng(seed)`` in the same order as
``cutile.workloads.sinusoid_init``.  Each sinusoid is read from a 1-D table
S[m] = sin(2*pi*m/n + phase), m = (k . index) mod n, stored in the golden
fixture, so building the field on another host needs no transcendental
function (NumPy's vectorised sin may differ in the last bit between CPUs).
Gray-Scott uses ``cutile.workloads.gray_scott_init`` (squares and uniform
noise only: the same bits on every host)."""

from __future__ import annotations

import math

import numpy as np

C2_N = 512
C2_STEPS = 50
C3_N = 4096
C3_STEPS = 1000
C3_EVERY = 50


def c2_waves(n: int = C2_N, seed: int = 0, kmax: int = 8):
    """(wavevectors, phases, rng) drawn as ``sinusoid_init`` draws them; the
    returned generator continues with the noise draw."""
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < 3:
        k = rng.integers(-kmax, kmax + 1, size=3)
        if 0 < float(np.dot(k, k)) <= kmax * kmax:
            ks.append([int(x) for x in k])
    phases = [float(p) for p in rng.random(3) * 2.0 * math.pi]
    return ks, phases, rng


def c2_tables(n: int = C2_N, seed: int = 0) -> dict:
    """Wavevectors, phases and the three sine tables (as float.hex strings)."""
    ks, phases, _ = c2_waves(n, seed)
    tabs = [[float(v).hex() for v in np.sin(2.0 * math.pi * np.arange(n) / n + ph)] for ph in phases]
    return {"n": n, "seed": seed, "k": ks, "phase": phases, "tables": tabs}


def c2_interior(spec: dict, noise: float = 1e-3) -> np.ndarray:
    """The float64 interior (n, n, n) built from the stored tables."""
    n = spec["n"]
    _, _, rng = c2_waves(n, spec["seed"])
    tabs = [np.array([float.fromhex(h) for h in t]) for t in spec["tables"]]
    idx = np.arange(n, dtype=np.int64)
    acc = np.zeros((n, n, n))
    for k, tab in zip(spec["k"], tabs):
        plane = (k[1] * idx[:, None] + k[2] * idx[None, :])
        for i in range(n):
            acc[i] += tab[(k[0] * i + plane) % n]
    acc += noise * rng.random((n, n, n))
    return acc


def place(interior: np.ndarray, halo, dtype=np.float64) -> np.ndarray:
    shape = tuple(s + 2 * h for s, h in zip(interior.shape, halo))
    u = np.zeros(shape, dtype=dtype)
    u[tuple(slice(h, h + s) for s, h in zip(interior.shape, halo))] = interior.astype(dtype)
    return u


def bf16_bits(x32: np.ndarray) -> np.ndarray:
    """uint16 bit patterns of float32 values that are exactly bfloat16."""
    b = np.ascontiguousarray(x32, dtype=np.float32).view(np.uint32)
    assert not np.any(b & 0xFFFF), "values are not bfloat16"
    return (b >> 16).astype(np.uint16)
