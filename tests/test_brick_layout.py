"""Brick-major addressing of the sweep (CPU only).

``st::Layout`` (``csrc/stencil_core.h``) addresses cell (X, Y, Z) of a
brick-major array as base(X) + ioff(Y, Z).  Restated here, it must name
exactly the element that the brick conversion (pad to whole bricks, split
every axis into (bricks, brick), bricks first -- upstream
``multigpu_emitter.py:370-406``) puts there, stay inside the plane's buffer
range bytes(X), and keep every aligned vector inside one brick row."""

import numpy as np
import pytest

from cutile.lowering.stencil_to_hip import brick_dims


def layout(shape3, dims):
    B0, B1, B2 = dims
    nby, nbz = -(-shape3[1] // B1), -(-shape3[2] // B2)
    vol = B0 * B1 * B2

    def base(X):
        return (X // B0) * (nby * nbz * vol) + (X % B0) * (B1 * B2)

    def ioff(Y, Z):
        return ((Y // B1) * nbz + Z // B2) * vol + (Y % B1) * B2 + (Z % B2)

    def nbytes(X):
        return nby * nbz * vol - (X % B0) * (B1 * B2)

    return base, ioff, nbytes


def bricked_index(shape, b):
    """Element index of every cell of a flat array in its brick-major copy."""
    nd = len(shape)
    padded = tuple(-(-s // b) * b for s in shape)
    idx = np.full(padded, -1, dtype=np.int64)
    idx[tuple(slice(0, s) for s in shape)] = np.arange(int(np.prod(shape))).reshape(shape)
    split = idx.reshape(sum(([p // b, b] for p in padded), []))
    order = list(range(0, 2 * nd, 2)) + list(range(1, 2 * nd, 2))
    flat = split.transpose(order).reshape(-1)
    pos = np.empty(int(np.prod(shape)), dtype=np.int64)
    valid = flat >= 0
    pos[flat[valid]] = np.nonzero(valid)[0]
    return pos.reshape(shape), int(flat.size)


@pytest.mark.parametrize("shape,b", [((18, 19, 21), 8), ((40, 33, 70), 16), ((9, 9, 9), 4),
                                     ((130, 160), 32), ((37, 45), 8), ((100,), 32), ((64, 64, 64), 32)])
def test_layout_matches_brick_conversion(shape, b):
    nd = len(shape)
    dims = brick_dims(nd, b)
    shape3 = (1, 1, shape[0]) if nd == 1 else ((shape[0], 1, shape[1]) if nd == 2 else tuple(shape))
    base, ioff, nbytes = layout(shape3, dims)
    pos, total = bricked_index(shape, b)
    pos3 = pos.reshape(shape3)
    X, Y, Z = np.meshgrid(*[np.arange(s) for s in shape3], indexing="ij")
    got = base(X) + ioff(Y, Z)
    assert np.array_equal(got, pos3)
    assert got.max() < total
    # every in-plane offset stays within the plane's buffer range
    assert (ioff(Y, Z) < nbytes(X)).all()


@pytest.mark.parametrize("b,vz", [(8, 2), (16, 4), (32, 8), (4, 2)])
def test_aligned_vectors_stay_in_one_brick_row(b, vz):
    dims = brick_dims(3, b)
    base, ioff, _ = layout((b * 2, b * 3, b * 4), dims)
    for z in range(0, b * 4, vz):
        offs = [ioff(5 % (b * 3), z + v) for v in range(vz)]
        assert offs == list(range(offs[0], offs[0] + vz))
