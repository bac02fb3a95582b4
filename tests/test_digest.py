"""The interior checksum (``oracle.digest.field_digest`` on the CPU,
``st_field_digest`` on the device) that pins full-size device states
against the oracle's golden digests.

CPU: SplitMix64's published outputs for seed 0 (0xe220a8397b1dcdaf,
0x6e789e6aa1b965f4) fix the mixer and the index weight; a scalar
Python-integer loop fixes the sum; splitting the array into chunks leaves
the sum unchanged; one flipped bit, two swapped cells and a transposed
array each change it.  GPU: the device digest of random float64, float32
and bfloat16 arrays with halos, 1-D to 3-D, equals the NumPy digest."""

import numpy as np
import pytest

from oracle.digest import GOLDEN, field_digest, mix64, stored_bits

MASK = (1 << 64) - 1


def _mix_int(z: int) -> int:
    z ^= z >> 30
    z = (z * 0xBF58476D1CE4E5B9) & MASK
    z ^= z >> 27
    z = (z * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def test_mixer_matches_splitmix64_published_outputs():
    # SplitMix64 seeded with 0: state += 0x9e3779b97f4a7c15 before each output
    got = mix64(np.array([int(GOLDEN), (2 * int(GOLDEN)) & MASK], dtype=np.uint64))
    assert [int(v) for v in got] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4]
    # a cell holding +0.0 at interior index e contributes the e-th output
    x = np.zeros(3)
    assert field_digest(x) == (_mix_int(0) + 0xE220A8397B1DCDAF + 0x6E789E6AA1B965F4) & MASK


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.uint16])
def test_digest_equals_scalar_loop(dtype):
    rng = np.random.default_rng(3)
    x = (rng.random((3, 4, 5)) * 1000).astype(dtype)
    want = 0
    for e, b in enumerate(stored_bits(x)):
        want = (want + _mix_int(int(b) ^ ((e * int(GOLDEN)) & MASK))) & MASK
    assert field_digest(x) == want


def test_digest_is_chunk_additive():
    x = np.random.default_rng(4).random(1000)
    assert field_digest(x, chunk=7) == field_digest(x, chunk=1000) == field_digest(x, chunk=64)


def test_digest_sees_small_changes():
    x = np.random.default_rng(5).random((6, 7))
    d = field_digest(x)
    y = x.copy()
    y.view(np.uint64)[3, 4] ^= np.uint64(1)          # one bit of one cell
    assert field_digest(y) != d
    z = x.copy()
    z[0, 0], z[5, 6] = x[5, 6], x[0, 0]               # same multiset of values, other places
    assert field_digest(z) != d
    sq = np.random.default_rng(6).random((5, 5))
    assert field_digest(sq.T) != field_digest(sq)
    assert field_digest(-0.0 * np.ones(4)) != field_digest(np.zeros(4))


@pytest.mark.gpu
@pytest.mark.parametrize("shape,halo,dtype", [
    ((40,), (16,), "float64"),
    ((9, 70), (1, 16), "float32"),
    ((5, 6, 130), (2, 1, 32), "bfloat16"),
    ((37, 45, 70), (4, 4, 16), "float64"),
    ((3, 2, 70000), (1, 1, 16), "float32"),     # rows longer than one partial
    ((1, 1, 1), (0, 0, 0), "float64"),
])
def test_device_digest_equals_numpy(shape, halo, dtype):
    import torch

    from cutile.runtime import rt
    from cutile.runtime.kernel import dtype_code

    print(f"digest {shape} halo {halo} {dtype}", flush=True)
    full = tuple(n + 2 * h for n, h in zip(shape, halo))
    rng = np.random.default_rng(len(shape) * 7 + full[-1])
    host = rng.standard_normal(full)
    t = torch.from_numpy(host).to("cuda").to(getattr(torch, dtype))
    got = rt.field_digest(t.data_ptr(), dtype_code(t), t.shape, halo)
    inner = tuple(slice(h, h + n) for n, h in zip(shape, halo))
    ref = t.cpu()
    ref = ref.view(torch.int16).numpy().view(np.uint16) if dtype == "bfloat16" else ref.numpy()
    assert got == field_digest(np.ascontiguousarray(ref[inner]))
