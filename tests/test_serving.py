"""Serving loop: the job generator on the CPU; continuous batching,
admission and per-job results on the device (``-m gpu``)."""

import numpy as np
import pytest

from helpers import to_host


def test_job_stream_statistics():
    from cutile.serving import KINDS, generate_jobs

    jobs = generate_jobs(rate=200, seconds=20, seed=3)
    assert 3600 < len(jobs) < 4400                       # Poisson(4000)
    arr = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(arr) > 0) and arr[-1] < 20
    kinds = {k: sum(j.kind == k for j in jobs) for k in KINDS}
    assert all(abs(c / len(jobs) - 1 / 3) < 0.05 for c in kinds.values())
    fp32 = sum(j.dtype == "float32" for j in jobs) / len(jobs)
    assert abs(fp32 - 0.8) < 0.03
    s2 = [j.side for j in jobs if j.kind != "diffusion3d"]
    s3 = [j.side for j in jobs if j.kind == "diffusion3d"]
    assert min(s2) >= 128 and max(s2) <= 4096 and min(s3) >= 32 and max(s3) <= 384
    assert abs(np.median(np.log(s2)) - np.log(np.sqrt(128 * 4096))) < 0.2    # log-uniform
    assert min(j.steps for j in jobs) >= 10 and max(j.steps for j in jobs) <= 1000
    assert generate_jobs(rate=200, seconds=2, seed=3)[:5] == jobs[:5]


def _mix():
    from cutile.serving import Job

    spec = [("heat2d", "float32", 70, 12), ("heat2d", "float64", 200, 25), ("diffusion3d", "float32", 20, 9),
            ("diffusion3d", "float64", 33, 17), ("grayscott", "float32", 50, 30), ("grayscott", "float64", 130, 6),
            ("heat2d", "float32", 333, 3), ("grayscott", "float32", 129, 21), ("diffusion3d", "float32", 45, 14),
            ("diffusion3d", "float32", 230, 5), ("heat2d", "float32", 2050, 6), ("grayscott", "float32", 2801, 5)]
    return [Job(i, k, d, s, n, arrival=0.0) for i, (k, d, s, n) in enumerate(spec)]


@pytest.mark.gpu
def test_each_job_is_the_same_alone_and_in_the_mix():
    import torch

    from cutile.serving import StencilServer

    print(f"serving: {len(_mix())}-job mix vs each job alone", flush=True)
    mix = StencilServer(keep_results=True)
    stats = mix.run(_mix(), realtime=False)
    assert stats["jobs"] == len(_mix())
    for job in _mix():
        alone = StencilServer(keep_results=True)
        alone.run([job], realtime=False)
        for f, t in alone.results[job.id].items():
            assert torch.equal(t, mix.results[job.id][f]), (job, f)


@pytest.mark.gpu
def test_class_pacing_equals_lockstep():
    """Classes stepping at their own pace (the default) give every job the
    same bits as all classes stepping together, one tick at a time."""
    import torch

    from cutile.serving import StencilServer

    print("serving: class pacing vs lockstep ticks", flush=True)
    paced = StencilServer(keep_results=True, pacing="class")
    lock = StencilServer(keep_results=True, pacing="lockstep")
    a = paced.run(_mix(), realtime=False)
    b = lock.run(_mix(), realtime=False)
    assert a["jobs"] == b["jobs"] == len(_mix())
    for jid, fields in lock.results.items():
        for f, t in fields.items():
            assert torch.equal(t, paced.results[jid][f]), (jid, f)


@pytest.mark.gpu
def test_float64_jobs_match_the_oracle(oracle):
    import torch

    from cutile import workloads as w
    from cutile.serving import KINDS, SOLO, Job, StencilServer

    print("serving: float64 jobs vs oracle", flush=True)
    # jobs 3-5 are stepped alone: depth-2 launches (two, then one sweep for
    # the fifth step) and depth-4 launches (one, then two sweeps; job 5
    # through the coupled-field kernel)
    srv = StencilServer(keep_results=True)
    # the classes' solo thresholds (measured and cached, or the defaults)
    solo = {k: srv._class(Job(-1, k, "float64", 1, 1)).solo_min or SOLO[k, "float64"][0] for k in KINDS}
    jobs = [Job(0, "heat2d", "float64", 90, 11), Job(1, "diffusion3d", "float64", 26, 7),
            Job(2, "grayscott", "float64", 61, 9),
            Job(3, "diffusion3d", "float64", solo["diffusion3d"], 5),
            Job(4, "heat2d", "float64", solo["heat2d"], 6),
            Job(5, "grayscott", "float64", solo["grayscott"], 6)]
    srv.run([Job(j.id, j.kind, j.dtype, j.side, j.steps) for j in jobs], realtime=False)
    for j in jobs:
        cls = srv._class(j)
        cls.allocate(j)                              # the same seeded initial fields
        assert j.solo == (j.id >= 3)
        n, halo = j.side, cls.halo_of(j)
        logical = tuple(slice(0, n + 2 * h) for h in halo)
        init = {f: to_host(j.bufs[f][0])[logical].copy() for f in cls.fields}
        if j.kind == "heat2d":
            ref = {"u": oracle.time_march(init["u"], w.heat2d()._fn, 2, halo, j.steps, w.dirichlet_spec(2))}
        elif j.kind == "diffusion3d":
            ref = {"u": oracle.time_march(init["u"], w.diffusion7(coef=0.1)._fn, 3, halo, j.steps,
                                          w.dirichlet_spec(3))}
        else:
            ref = oracle.time_march_multi(init, [g._fn for g in w.gray_scott()], ["u", "v"], 2, halo, j.steps,
                                          w.periodic_spec(2))
        I = oracle.interior(init["u"].shape, halo)
        for f in cls.fields:
            got = to_host(srv.results[j.id][f])[logical]
            assert np.array_equal(got[I], ref[f][I]), (j.kind, f)
        torch.cuda.synchronize()


@pytest.mark.gpu
def test_admission_waits_for_memory():
    import torch

    from cutile.serving import StencilServer

    print("serving: low HBM budget forces queueing", flush=True)
    jobs = _mix()
    probe = StencilServer()
    biggest = max(probe.job_bytes(j) for j in jobs)
    srv = StencilServer(hbm_budget=int(2.2 * biggest), keep_results=True)
    stats = srv.run(jobs, realtime=False)
    assert stats["jobs"] == len(jobs)
    assert stats["peak_reserved_gib"] * 2**30 <= 2.2 * biggest
    admitted = sorted(j.admitted for j in jobs)
    assert admitted[-1] > admitted[0]                 # some jobs had to wait for others to finish
    ref = StencilServer(keep_results=True)
    ref.run(_mix(), realtime=False)
    for jid, fields in ref.results.items():
        for f, t in fields.items():
            assert torch.equal(t, srv.results[jid][f])
