"""Batched persistent kernel over a device tile work-list (``-m gpu``):
many uneven instances in one launch give, per instance, exactly what the
oracle gives for that instance alone."""

import numpy as np
import pytest

from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_batched_leapfrog_matches_each_instance(oracle, dtype):
    import torch

    from cutile import workloads as w
    from cutile.runtime.batched import BatchedStencil, LeapfrogBatch
    from cutile.runtime.launcher import pad_inner_halo

    fn = w.wave2d(dtype)
    npdt = np.dtype(dtype).type
    halo = pad_inner_halo((1, 1), np.dtype(dtype).itemsize)
    sides = [37, 64, 130, 300, 700, 1100]
    bs = BatchedStencil(fn)
    vec = 16 // np.dtype(dtype).itemsize
    pairs, host = [], []
    for k, s in enumerate(sides):
        u, uold = w.wave_init(s, halo, seed=k, dtype=npdt)
        host.append((u, uold))
        cols = -(-u.shape[1] // vec) * vec  # rows padded to whole 16-byte vectors
        pad = ((0, 0), (0, cols - u.shape[1]))
        pairs.append((to_dev(np.pad(u, pad)), to_dev(np.pad(uold, pad))))
    lf = LeapfrogBatch(bs, pairs, halo, interiors=[(s, s) for s in sides])
    print(f"{dtype}: {lf.plans[0].nitems} work items, grid {lf.plans[0].grid}", flush=True)
    steps = 7
    for _ in range(steps):
        lf.step()
    torch.cuda.synchronize()
    spec = w.dirichlet_spec(2)
    for k, (u, uold) in enumerate(host):
        ru, ruold = oracle.leapfrog(u, uold, fn._fn, 2, halo, steps, spec)
        gu, guold = lf.current(k)
        n1 = u.shape[1]
        assert np.array_equal(to_host(gu)[:, :n1], ru), f"instance {k} (side {sides[k]})"
        assert np.array_equal(to_host(guold)[:, :n1], ruold), f"instance {k} (side {sides[k]}) old level"


def test_batched_single_field_3d(oracle):
    import torch

    from cutile import workloads as w
    from cutile.runtime.batched import BatchedStencil

    fn = w.diffusion7()
    halo = (1, 1, 16)
    bs = BatchedStencil(fn)
    rng = np.random.default_rng(0)
    shapes = [(10, 12, 40), (33, 17, 70), (5, 64, 130)]
    ins, outs, host = [], [], []
    for sh in shapes:
        u = rng.random(tuple(n + 2 * h for n, h in zip(sh, halo)))
        host.append(u)
        ins.append(to_dev(u))
        outs.append(torch.zeros_like(ins[-1]))
    plan = bs.prepare([([a], [b]) for a, b in zip(ins, outs)], halo)
    bs.launch(plan)
    torch.cuda.synchronize()
    for u, o in zip(host, outs):
        ref = oracle.apply_stencil(u, fn._fn, 3, halo)
        inner = oracle.interior(u.shape, halo)
        assert np.array_equal(to_host(o)[inner], ref[inner])
