"""Continuous-batching stencil server on one MI355X.

There is no upstream counterpart (upstream launches one kernel per array per
step); this is the request loop of the serving configuration:

* a job is a time-stepping request -- a stencil kind (2-D heat, 3-D 7-point
  diffusion ``u += 0.1*L(u)``, fused Gray-Scott), a storage type, a side
  length and a step count -- arriving at some time;
* **admission by free HBM**: a job enters when its resident fields (two
  buffers per field, halos included) fit in what the budget has left;
  arrivals wait in FIFO order until completions free memory;
* **resident fields**: an admitted job's buffers are allocated once and
  initialised on the device from ``seed = job id``; the job steps in place
  (ping-pong buffers) until done, then its buffers are released;
* **packing**: a step of a (stencil, storage) class advances every active
  job of the class by one step -- all of them go through ONE launch of the
  batched persistent kernel (``cutile.runtime.batched``) over a device
  work-list of their tiles, with per-size-class tile buckets; periodic
  halos of a class are refilled by one batched fill launch per axis;
  Dirichlet halos are zero once and stay so (the sweep writes interiors
  only);
* **pacing**: each class has its own stream and steps whenever fewer than
  ``inflight`` of its steps are queued on the device (completion is polled
  with events, never waited on while another class can step), so a class
  of cheap steps is not held to the pace of the slowest class
  (``pacing="lockstep"``: all classes step together, one tick at a time);
* **solo temporal stepping**: large jobs (``SOLO``) are stepped alone,
  several steps per launch of the register/LDS temporal kernel (coupled
  fields for Gray-Scott), where that outruns the batched sweep; their last
  steps short of a block run as single sweeps;
* **continuous batching**: jobs join their class at its next step after
  they are admitted and leave after their last step, so work of jobs that
  arrived at different times overlaps on the device; the work-lists are
  rebuilt only when a class's active set changes (two plans per set:
  buffer parities flip every step);
* latency = completion (the class step holding a job's last step has
  finished on the device, as the host sees it) minus arrival; the loop
  reports p50/p99 latency and aggregate Gcell-updates/s.

Each job's result depends only on its own fields: the batched kernel
evaluates every instance with the same arithmetic as the single-instance
kernel, and a temporal block gives the same bits as its T sweeps, so a job
gives the same bits alone and in any mix.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

KINDS = ("heat2d", "diffusion3d", "grayscott")


# marker of a solo job's launch record that has not been packed yet (a packed
# record is None when the job's interior is empty)
_UNPREPARED = object()

@dataclass
class Job:
    id: int
    kind: str
    dtype: str
    side: int
    steps: int
    arrival: float = 0.0          # seconds after the start of the run
    # runtime state
    bufs: dict = field(default_factory=dict, repr=False)
    done: int = 0
    admitted: Optional[float] = None
    finished: Optional[float] = None
    nbytes: int = 0
    keep: list = field(default_factory=list, repr=False)   # staging the device still reads
    recs: list = field(default_factory=list, repr=False)   # launch records, both parities
    fdesc: Optional[list] = field(default=None, repr=False)
    solo: bool = False            # stepped alone, several steps per launch (large jobs)
    cur: int = 0                  # solo jobs: the buffer holding the current state
    lane_ready: bool = False      # solo jobs (class pacing): the solo stream has waited for its set-up

    @property
    def ndim(self) -> int:
        return 3 if self.kind == "diffusion3d" else 2

    @property
    def cells(self) -> int:
        return self.side ** self.ndim

    @property
    def fields(self) -> tuple:
        return ("u", "v") if self.kind == "grayscott" else ("u",)


def generate_jobs(rate: float = 200.0, seconds: float = 60.0, seed: int = 0, fp32_share: float = 0.8,
                  sides_2d=(128, 4096), sides_3d=(32, 384), steps=(10, 1000)) -> List[Job]:
    """Poisson arrivals at ``rate`` jobs/s for ``seconds`` (seeded): kind
    uniform over ``KINDS``, 2-D sides log-uniform in ``sides_2d``, 3-D sides
    log-uniform in ``sides_3d``, steps uniform in ``steps``, float32 with
    probability ``fp32_share`` else float64."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        kind = KINDS[int(rng.integers(0, len(KINDS)))]
        lo, hi = sides_3d if kind == "diffusion3d" else sides_2d
        side = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        dtype = "float32" if rng.random() < fp32_share else "float64"
        jobs.append(Job(len(jobs), kind, dtype, side, nsteps, t))
    return jobs


# Jobs of these classes with at least the given side are stepped alone, T
# steps per launch (the register/LDS temporal kernel, zero Dirichlet),
# instead of through the batched sweep.  The sides are where the single-array
# temporal launch overtakes the batched sweep on MI355X: a depth-2 launch on
# one 3-D array advances 320-830 Gcell/s from side 191 to 383 against 230-540
# for the batched sweep (FP64 gains from ~180, FP32 from ~220); a depth-4
# launch on one 2-D heat array 320-1030 from side 1531 to 4096 against
# 250-600 (FP64 gains from ~1300, FP32 from ~2000).
SOLO = {                      # (kind, storage) -> (smallest side, steps per launch)
    ("diffusion3d", "float32"): (224, 2),
    ("diffusion3d", "float64"): (176, 2),
    ("heat2d", "float32"): (2048, 4),
    ("heat2d", "float64"): (1280, 4),
    # fused Gray-Scott through the coupled-field temporal kernel (periodic:
    # wrapped reads): 410-480 / 160-250 Gcell/s from these sides to 4096
    # against ~300 / ~140 batched
    ("grayscott", "float32"): (2800, 4),
    ("grayscott", "float64"): (1536, 4),
}


# Solo thresholds measured per class (``_Class._measure_solo``: one job alone
# against the per-job time in a loaded batched launch) instead of ``SOLO``.
# Off by default: on the 60-s stream (two runs each) the measured thresholds
# gave p99 1.20 / 1.43 s against 1.32 / 1.00 s with ``SOLO``, which was set
# from class streams under load; ``python -m cutile.tune --serving-solo``
# turns it on.
SOLO_TUNING = False

# Candidates the solo threshold is measured over (``_Class._measure_solo``):
# the depths of the temporal kernel and the sides of a job of the class.
SOLO_DEPTHS = {"diffusion3d": (2, 4), "heat2d": (2, 4), "grayscott": (2, 4)}
SOLO_SIDES = {"diffusion3d": (128, 160, 176, 192, 224, 256, 320, 384),
              "heat2d": (768, 1024, 1280, 1536, 2048, 2800, 4096),
              "grayscott": (1024, 1536, 2048, 2800, 4096)}


def class_sides(lo: int, hi: int, count: int = 24) -> list:
    """Log-spaced sides in [lo, hi]: the batched jobs of a class whose tile
    buckets are tuned."""
    if hi <= lo:
        return [lo]
    return sorted({int(round(lo * (hi / lo) ** (k / (count - 1)))) for k in range(count)})


@dataclass
class _Solo:
    """Kernels of a class's solo jobs: one step, and a block of ``T`` steps
    (the register/LDS temporal kernel), on arrays with halo ``halo``;
    ``wrap``: the axes both read around (periodic jobs)."""

    single: object
    block: object
    halo: tuple
    T: int
    wrap: tuple


def warmup_jobs(server: "Optional[StencilServer]" = None) -> List[Job]:
    """Jobs covering every class's kernels (a two-step small job of every
    (kind, storage) class and a solo-sized job per solo kind and storage,
    one temporal launch and one sweep), run before a timed stream so that no
    kernel is built while the clock runs.  With ``server`` the solo sides
    and depths are that server's classes' (measured) ones."""
    out = []
    for kind in KINDS:
        for dtype in ("float32", "float64"):
            out.append(Job(len(out), kind, dtype, 40 if kind == "diffusion3d" else 128, 2))
    for kind, dtype in SOLO:
        side, depth = SOLO[(kind, dtype)]
        if server is not None:
            cls = server._class(Job(-1, kind, dtype, 1, 1))
            side, depth = cls.solo_min, cls.depth
        if side:
            out.append(Job(len(out), kind, dtype, side, depth + 1))
    return out


# ---------------------------------------------------------------------------
# one (stencil, storage) class
# ---------------------------------------------------------------------------

class _Class:
    def __init__(self, kind: str, dtype: str):
        from cutile import workloads as w
        from cutile.runtime.batched import BatchedStencil, Bucket

        self.kind, self.dtype = kind, dtype
        if kind == "heat2d":
            fn = w.heat2d(dtype)
        elif kind == "diffusion3d":
            fn = w.diffusion7(dtype, coef=0.1)
        else:
            fn = list(w.gray_scott(dtype))
        probe = BatchedStencil(fn)
        vz = probe.vz
        self.fn = fn
        self.solo_min, self.depth = SOLO.get((kind, dtype), (None, 0))
        tuned_solo = self._tuned_solo() if (kind, dtype) in SOLO and SOLO_TUNING else None
        if tuned_solo is not None:
            self.solo_min, self.depth = tuned_solo
        # tile buckets of the batched kernel per size class: measured (the
        # autotuner's cached choice for this stencil, storage, sizes and
        # arch) where available, else these defaults
        from cutile.runtime.autotune import tuned_buckets

        top = self.solo_min - 1 if self.solo_min else (384 if kind == "diffusion3d" else 4096)
        if kind == "diffusion3d":
            buckets = [Bucket("rows", 8, 32 * vz, 16, 96), Bucket("rows", 8, 64 * vz, 32, None)]
            tb = tuned_buckets(fn, class_sides(32, top), bounds=(96, None))
        else:
            buckets = None                       # the 2-D plane/rows buckets by row length
            tb = tuned_buckets(fn, class_sides(128, top))
        self.bs = BatchedStencil(fn, buckets=tb or buckets)
        self.vz = vz
        self.ndim = 3 if kind == "diffusion3d" else 2
        self.halo = (1,) * (self.ndim - 1) + (vz,)   # contiguous-axis halo one vector wide
        self.periodic = kind == "grayscott"
        self.fields = ("u", "v") if kind == "grayscott" else ("u",)
        self.esize = 4 if dtype == "float32" else 8
        self.jobs: List[Job] = []
        self.plans = None
        self.fills = None
        self.tick = 0
        self._nb = 0                  # batched jobs in the plans (rebuild)
        self._end_tick = 0            # class step at which the first of them ends
        self.stream = None            # the class's own stream in the server
        self.solo = None
        if self.solo_min:
            self.solo = self._solo_kernels(self.depth)

    def _solo_kernels(self, depth: int) -> "_Solo":
        """One-step and ``depth``-step kernels of this class's solo jobs."""
        from cutile import workloads as w

        fn, kind = self.fn, self.kind
        if True:
            if kind == "grayscott":
                from cutile import compile_fused
                from cutile.runtime.launcher import compile_fused_block

                # periodic: blocks and single steps read around the interior
                one = compile_fused(fn)
                blk = compile_fused_block(fn, depth, one, ["u", "v"])
                return _Solo(one.load_module()._KERNEL, blk.load_module()._KERNEL, tuple(one.halo_widths),
                             depth, (0, 1))
            from cutile.runtime.timestep import TimeStepper

            # zero-Dirichlet blocks of T steps: the same states as T sweeps
            ndim = 3 if kind == "diffusion3d" else 2
            st = TimeStepper(fn, w.dirichlet_spec(ndim), temporal_steps=depth, split="fit")
            if st.T != depth or st.block_reps != 1:
                raise RuntimeError(f"serving class {kind}: no one-launch temporal kernel of depth {depth}")
            return _Solo(st._single, st._block, tuple(st.halo), depth, ())

    # -- measured solo threshold ---------------------------------------------------
    def _tuned_solo(self):
        """(smallest solo side, depth) from the autotuner's cached
        measurement for this class (``tuned``), or None."""
        from cutile.runtime.autotune import _arch, tuned

        parts = {"what": "serving_solo_loaded", "kind": self.kind, "storage": self.dtype,
                 "depths": list(SOLO_DEPTHS[self.kind]), "sides": list(SOLO_SIDES[self.kind])}
        val = tuned(parts, self._measure_solo, _arch())
        return None if not val else (int(val[0]), int(val[1]))

    def _measure_solo(self):
        """Per candidate side (``SOLO_SIDES``): the time per job-step of a
        batch of such jobs in one launch of the batched sweep (enough jobs
        to fill the device, as the class runs under load), against one job
        stepped alone by each candidate depth's temporal kernel; the solo
        side is the smallest candidate from which the best depth wins at
        every larger candidate too."""
        import torch

        from cutile.runtime import rt
        from cutile.runtime.batched import BatchedStencil
        from cutile.runtime.kernel import default_stream, run_prepared

        tdt = getattr(torch, self.dtype)
        nd = 3 if self.kind == "diffusion3d" else 2
        bs = BatchedStencil(self.fn)
        s = default_stream()
        nf = 2 if self.kind == "grayscott" else 1

        def timed(fn, reps=8):
            fn()
            torch.cuda.synchronize()
            e0, e1 = rt.Event(), rt.Event()
            e0.record(s)
            for _ in range(reps):
                fn()
            e1.record(s)
            return e0.elapsed_ms(e1) / reps

        wins = []
        for side in SOLO_SIDES[self.kind]:
            cols = -(-side // bs.vz) * bs.vz
            halo = (1,) * (nd - 1) + (bs.vz,)
            shp = tuple(n + 2 * h for n, h in zip((side,) * (nd - 1) + (cols,), halo))
            jobs = max(1, min(16, -(-(32 << 20) // side ** nd)))      # ~32M cells per launch
            bufs = [([torch.zeros(shp, dtype=tdt, device="cuda") for _ in range(nf)],
                     [torch.zeros(shp, dtype=tdt, device="cuda") for _ in range(nf)]) for _ in range(jobs)]
            plan = bs.assemble([bs.record(a, b, halo, (side,) * nd) for a, b in bufs])
            batched = timed(lambda: bs.launch(plan, s)) / jobs
            del bufs, plan
            best = None
            for depth in SOLO_DEPTHS[self.kind]:
                sol = self._solo_kernels(depth)
                shp = tuple(side + 2 * h for h in sol.halo)
                a = [torch.zeros(shp, dtype=tdt, device="cuda") for _ in range(nf)]
                b = [torch.zeros(shp, dtype=tdt, device="cuda") for _ in range(nf)]
                rec = sol.block.prepare(a, b, sol.halo, temporal=True, mode=0, periodic_axes=sol.wrap)
                ms = timed(lambda: run_prepared(rec, s)) / depth
                del a, b, rec
                if best is None or ms < best[0]:
                    best = (ms, depth)
            wins.append((side, best[0] < batched, best[1]))
            from cutile.runtime import autotune as _at

            if _at.VERBOSE:
                print(f"  solo {self.kind} {self.dtype} side {side}: batched {batched * 1e3:.1f} us/step, "
                      f"alone {best[0] * 1e3:.1f} us/step at depth {best[1]}", flush=True)
            torch.cuda.empty_cache()
        for k, (side, _win, depth) in enumerate(wins):
            if all(w for _, w, _ in wins[k:]):
                return [side, depth]
        return [0, 0]

    def is_solo(self, side: int) -> bool:
        return self.solo is not None and bool(self.solo_min) and side >= self.solo_min

    def shape(self, side: int) -> tuple:
        if self.is_solo(side):                           # the temporal kernel's own layout
            return tuple(side + 2 * h for h in self.solo.halo)
        cols = -(-side // self.vz) * self.vz             # rows padded to whole vectors
        inner = (side,) * (self.ndim - 1) + (cols,)
        return tuple(n + 2 * h for n, h in zip(inner, self.halo))

    def halo_of(self, job: Job) -> tuple:
        return self.solo.halo if job.solo else self.halo

    def current(self, job: Job) -> int:
        """Index of the buffer that holds the job's state after the steps taken."""
        return job.cur if job.solo else job.done % 2

    def job_bytes(self, side: int) -> int:
        return 2 * len(self.fields) * int(np.prod(self.shape(side))) * self.esize

    # -- buffers ----------------------------------------------------------------
    def allocate(self, job: Job) -> None:
        """Resident buffers, seeded initial fields, and the job's cached
        launch records and fill descriptors for both buffer parities."""
        import torch

        tdt = getattr(torch, self.dtype)
        shp = self.shape(job.side)
        job.solo = self.is_solo(job.side)
        job.cur = 0
        job.bufs = {f: [torch.zeros(shp, dtype=tdt, device="cuda") for _ in range(2)] for f in self.fields}
        job.nbytes = self.job_bytes(job.side)
        init_fields(job, self, job.bufs)
        if job.solo:
            return
        inner = (job.side,) * self.ndim
        job.recs = [self.bs.record([job.bufs[f][p] for f in self.fields], [job.bufs[f][1 - p] for f in self.fields],
                                   self.halo, inner) for p in (0, 1)]
        job.fdesc = [self._fill_descs(job, 1 - p) for p in (0, 1)] if self.periodic else None

    # -- plans -------------------------------------------------------------------
    def rebuild(self) -> None:
        """Work-lists (and periodic fill tables) for the active set at both
        buffer parities, assembled from the jobs' cached records."""
        from cutile.runtime.batched import upload

        self.plans, self.fills = [], []
        batched = [j for j in self.jobs if not j.solo]
        # batched jobs advance one step per class step from here (tick 0):
        # the first of them ends at this tick
        self._end_tick = min((j.steps - j.done for j in batched), default=0)
        self._nb = len(batched)
        for flip in (0, 1):
            if not batched:
                self.plans.append(None)
                self.fills.append(None)
                continue
            par = [(j.done + flip) % 2 for j in batched]
            self.plans.append(self.bs.assemble([j.recs[p] for j, p in zip(batched, par)]))
            if self.periodic:
                blob = b"".join(j.fdesc[p][0] for j, p in zip(batched, par))
                dev, host = upload(np.frombuffer(blob, dtype=np.uint8))
                count = len(batched) * len(self.fields)
                max_slab = max(j.fdesc[p][1] for j, p in zip(batched, par))
                self.fills.append((dev, count, max_slab, host))
            else:
                self.fills.append(None)

    def _fill_descs(self, job: Job, which: int) -> tuple:
        """Periodic fill descriptors (bytes) of the job's buffers ``which``
        (the outputs of a step from the other parity) and their largest slab."""
        from cutile.runtime import rt
        from cutile.runtime.kernel import dtype_code

        descs = (rt.FillDesc * len(self.fields))()
        logical = tuple(job.side + 2 * h for h in self.halo)
        max_slab = 1
        for k, f in enumerate(self.fields):
            t = job.bufs[f][which]
            d = descs[k]
            d.ptr = t.data_ptr()
            for a in range(self.ndim):
                d.shape[a] = logical[a]
                d.halo[a] = self.halo[a]
                d.kinds[2 * a] = d.kinds[2 * a + 1] = rt.BC_CODES["periodic"]
                slab = 2 * self.halo[a] * int(np.prod([logical[b] for b in range(self.ndim) if b != a]))
                max_slab = max(max_slab, slab)
            d.ndim, d.dtype, d.pitch = self.ndim, dtype_code(t), t.shape[-1]
        return bytes(descs), max_slab

    def step(self, stream) -> None:
        """Advance every active job of the class: batched jobs by one step
        (one launch for all), solo jobs by T steps (one temporal launch each;
        single sweeps for a job's last steps short of a block)."""
        self.step_batched(stream)
        for j in self.jobs:
            if j.solo:
                self.step_solo(j, stream)

    def has_batched(self) -> bool:
        """Batched jobs in the active set (as of the last ``rebuild``)."""
        return self._nb > 0

    def batched_ending(self) -> bool:
        """Some batched job has taken its last step (since the last ``rebuild``)."""
        return self._nb > 0 and self.tick >= self._end_tick

    def step_batched(self, stream) -> None:
        """Advance the class's batched jobs by one step (one launch for all)."""
        from cutile.runtime import rt

        flip = self.tick % 2
        plan = self.plans[flip] if self.plans else None
        if plan is not None:
            self.bs.launch(plan, stream)
            if self.fills[flip] is not None:
                blob, count, max_slab, _host = self.fills[flip]
                rt.fill_halo_batched(blob.data_ptr(), count, self.ndim, max_slab, stream)
        self.tick += 1
        for j in self.jobs:
            if not j.solo:
                j.done += 1

    def step_solo(self, j: Job, stream) -> None:
        """Advance solo job ``j`` by T steps (one temporal launch), or by one
        sweep for its last steps short of a block."""
        from cutile.runtime.kernel import run_prepared

        sol = self.solo
        block = j.steps - j.done >= sol.T
        rec = j.recs[j.cur][block] if j.recs else _UNPREPARED
        if rec is _UNPREPARED:                # packed once per buffer parity and kind
            if not j.recs:
                j.recs = [[_UNPREPARED, _UNPREPARED], [_UNPREPARED, _UNPREPARED]]
            ins = [j.bufs[f][j.cur] for f in self.fields]
            outs = [j.bufs[f][1 - j.cur] for f in self.fields]
            rec = (sol.block.prepare(ins, outs, sol.halo, temporal=True, mode=0, periodic_axes=sol.wrap)
                   if block else sol.single.prepare(ins, outs, sol.halo, periodic_axes=sol.wrap))
            j.recs[j.cur][block] = rec       # None: empty interior, nothing to launch
        run_prepared(rec, stream)
        j.done += sol.T if block else 1
        j.cur = 1 - j.cur


def init_fields(job: Job, cls: _Class, bufs: dict) -> None:
    """Initial fields on the device from ``seed = job.id`` (the generators of
    the benchmark configurations): heat -- uniform[0,1); 3-D diffusion --
    three seeded sinusoids + 1e-3 uniform noise; Gray-Scott -- u = 1, v = 0,
    64 seeded 10x10 squares (u = 0.5, v = 0.25) + 1 % uniform noise.  The
    periodic Gray-Scott halo is filled once here."""
    import torch

    g = torch.Generator(device="cuda").manual_seed(int(job.id))
    rng = np.random.default_rng(int(job.id))
    s, h = job.side, cls.halo_of(job)
    sl = tuple(slice(hh, hh + s) for hh in h)
    if job.kind == "heat2d":
        bufs["u"][0][sl] = torch.rand((s, s), generator=g, device="cuda", dtype=torch.float64).to(bufs["u"][0].dtype)
    elif job.kind == "diffusion3d":
        x = torch.arange(s, device="cuda", dtype=torch.float64) * (2.0 * math.pi / s)
        acc = torch.zeros((s, s, s), device="cuda", dtype=torch.float64)
        for _ in range(3):
            k = rng.integers(-8, 9, size=3)
            ph = float(rng.uniform(0, 2 * math.pi))
            acc += torch.sin(k[0] * x[:, None, None] + k[1] * x[None, :, None] + k[2] * x[None, None, :] + ph)
        acc += 1e-3 * torch.rand((s, s, s), generator=g, device="cuda", dtype=torch.float64)
        bufs["u"][0][sl] = acc.to(bufs["u"][0].dtype)
    else:
        u = torch.ones((s, s), device="cuda", dtype=torch.float64)
        v = torch.zeros((s, s), device="cuda", dtype=torch.float64)
        side = min(10, s)
        for _ in range(64):
            cx, cy = (int(c) for c in rng.integers(0, s - side + 1, size=2))
            u[cx:cx + side, cy:cy + side] = 0.5
            v[cx:cx + side, cy:cy + side] = 0.25
        u += 0.01 * torch.rand((s, s), generator=g, device="cuda", dtype=torch.float64)
        v += 0.01 * torch.rand((s, s), generator=g, device="cuda", dtype=torch.float64)
        bufs["u"][0][sl] = u.to(bufs["u"][0].dtype)
        bufs["v"][0][sl] = v.to(bufs["v"][0].dtype)
        job.keep = [_fill_padded(bufs[f][0], s, h) for f in ("u", "v")]


def _fill_padded(t, side: int, halo):
    """Periodic fill of an array whose rows are padded past the logical
    width (one batched-fill descriptor with a row pitch); returns the
    descriptor buffers, which must live until the stream has run the fill."""
    from cutile.runtime import rt
    from cutile.runtime.batched import upload
    from cutile.runtime.kernel import default_stream, dtype_code

    nd = t.dim()
    d = rt.FillDesc()
    d.ptr = t.data_ptr()
    logical = tuple(side + 2 * h for h in halo)
    max_slab = 1
    for a in range(nd):
        d.shape[a], d.halo[a] = logical[a], halo[a]
        d.kinds[2 * a] = d.kinds[2 * a + 1] = rt.BC_CODES["periodic"]
        max_slab = max(max_slab, 2 * halo[a] * int(np.prod([logical[b] for b in range(nd) if b != a])))
    d.ndim, d.dtype, d.pitch = nd, dtype_code(t), t.shape[-1]
    blob, host = upload(np.frombuffer(bytes(d), dtype=np.uint8))
    rt.fill_halo_batched(blob.data_ptr(), 1, nd, max_slab, default_stream())
    return blob, host


# ---------------------------------------------------------------------------
# the server
# ---------------------------------------------------------------------------

class StencilServer:
    """Admit, pack and step time-stepping jobs until all are done.

    ``hbm_budget``: bytes of resident fields allowed at once (default: 90 %
    of the device's free memory at start).  ``keep_results``: keep each
    job's final fields (device tensors) in ``results`` instead of freeing
    them.  ``inflight``: steps a class has queued on the device ahead of the
    host.  ``pacing``: ``"class"`` (default) lets every (stencil, storage)
    class step as soon as its own queue has room, so a class of cheap steps
    is not held to the pace of the slowest class; ``"lockstep"`` advances
    all classes together, one step per tick."""

    def __init__(self, hbm_budget: Optional[int] = None, keep_results: bool = False, inflight: int = 3,
                 pacing: str = "class"):
        import torch

        from cutile.runtime import rt

        if pacing not in ("class", "lockstep"):
            raise ValueError(f"pacing must be 'class' or 'lockstep', got {pacing!r}")
        if hbm_budget is None:
            free, _total = rt.mem_info()
            hbm_budget = int(free * 0.9)
        self.budget = int(hbm_budget)
        self.keep = keep_results
        self.inflight = max(1, inflight)
        self.pacing = pacing
        self.classes: Dict[tuple, _Class] = {}
        self.results: Dict[int, dict] = {}
        self.reserved = 0
        self.peak_reserved = 0
        self.main = torch.cuda.current_stream()
        self.stream = self.main.cuda_stream

    def _class(self, job: Job) -> _Class:
        key = (job.kind, job.dtype)
        if key not in self.classes:
            import torch

            cls = _Class(*key)
            cls.stream = torch.cuda.Stream()
            cls.solo_stream = torch.cuda.Stream()        # the class's solo jobs (class pacing)
            self.classes[key] = cls
        return self.classes[key]

    def job_bytes(self, job: Job) -> int:
        return self._class(job).job_bytes(job.side)

    def run(self, jobs: List[Job], realtime: bool = True, log=None) -> dict:
        """Serve ``jobs`` (sorted by arrival).  ``realtime``: arrivals follow
        the wall clock; otherwise every job is present at time 0."""
        import torch

        from cutile.runtime import rt

        for j in jobs:
            if self.job_bytes(j) > self.budget:
                raise ValueError(f"job {j.id} needs {self.job_bytes(j)} bytes, more than the budget {self.budget}")
        pending = sorted(jobs, key=lambda j: j.arrival)
        waiting: List[Job] = []
        active: List[Job] = []
        ticks: list = []                       # (event, [jobs finishing in that tick])
        finished: List[Job] = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        nticks = 0
        host = {"admit": 0.0, "rebuild": 0.0, "launch": 0.0, "wait": 0.0, "idle": 0.0}

        spare: list = []                       # events of retired class-paced steps, recorded again

        def new_event():
            return spare.pop() if spare else rt.Event()

        def retire(ev, done_jobs):
            # ev: the event of the step that finished these jobs (None: known complete)
            if ev is not None:
                tw = time.perf_counter()
                ev.synchronize()
                host["wait"] += time.perf_counter() - tw
            tnow = now()
            for j in done_jobs:
                j.finished = tnow
                cls = self._class(j)
                if self.keep:
                    p = cls.current(j)
                    self.results[j.id] = {f: j.bufs[f][p] for f in cls.fields}
                j.bufs = {}
                j.keep = []
                j.recs = []
                j.fdesc = None
                self.reserved -= j.nbytes
                finished.append(j)

        # class pacing: per class (its batched jobs) and per solo job, the
        # steps queued on the device, oldest first: (event recorded after the
        # step, jobs whose last step it was)
        queues: Dict[tuple, list] = {}
        finishing: set = set()                 # queues holding a step that ends some job

        def poll(every: bool = False) -> bool:
            """Retire the completed steps of every full queue and of every
            queue that ends a job (``every``: of all queues), without
            waiting; True if any.  (A queue with room and no ending job needs
            no retirement to go on.)"""
            got = False
            for key, q in list(queues.items()):
                if not every and len(q) < self.inflight and key not in finishing:
                    continue
                while q and q[0][0].query():
                    ev, done_jobs = q.pop(0)
                    retire(None, done_jobs)
                    spare.append(ev)
                    got = True
                if key in finishing and not any(d for _, d in q):
                    finishing.discard(key)
                if not q and key[0] == "job":
                    del queues[key]                # a solo job's queue, re-created while it runs
            return got

        def wait_any() -> None:
            """Wait until some queued class step completes and retire it."""
            tw = time.perf_counter()
            while any(queues.values()) and not poll(every=True):
                pass
            host["wait"] += time.perf_counter() - tw

        def outstanding() -> bool:
            return bool(ticks) or any(queues.values())

        last_log = 0.0
        while pending or waiting or active or outstanding():
            t = now() if realtime else float("inf")
            while pending and pending[0].arrival <= t:
                waiting.append(pending.pop(0))
            # admission (FIFO) by free budget
            changed = set()
            ta = time.perf_counter()
            while waiting and self.reserved + self.job_bytes(waiting[0]) <= self.budget:
                j = waiting.pop(0)
                cls = self._class(j)
                with torch.cuda.stream(cls.stream):       # buffers and initial fields on the class's stream
                    cls.allocate(j)
                j.admitted = now()
                self.reserved += j.nbytes
                self.peak_reserved = max(self.peak_reserved, self.reserved)
                cls.jobs.append(j)
                active.append(j)
                changed.add(cls)
            host["admit"] += time.perf_counter() - ta
            if not active:
                if ticks:
                    retire(*ticks.pop(0))
                    continue
                if any(queues.values()):
                    wait_any()
                    continue
                if pending and realtime:
                    ti = time.perf_counter()
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
                    host["idle"] += time.perf_counter() - ti
                continue
            tr = time.perf_counter()
            for cls in changed:
                cls.tick = 0
                with torch.cuda.stream(cls.stream):       # work-list uploads on the class's stream
                    cls.rebuild()
            host["rebuild"] += time.perf_counter() - tr
            # A class's buffers, uploads and launches all live on its own
            # stream, so the classes never wait for each other on the device.
            tl = time.perf_counter()
            if self.pacing == "lockstep":
                # one tick: every class advances its jobs by one step; the
                # main stream joins the classes to mark the tick for
                # retirement (and the host's in-flight limit), so every class
                # runs at the pace of the slowest
                busy = [c for c in self.classes.values() if c.jobs]
                for cls in busy:
                    cls.step(cls.stream.cuda_stream)
                for cls in busy:
                    self.main.wait_stream(cls.stream)
                ev = rt.Event()
                ev.record(self.stream)
                nticks += 1
                done_now = [j for j in active if j.done >= j.steps]
                if done_now:
                    active = [j for j in active if j.done < j.steps]
                    for cls in {self._class(j) for j in done_now}:
                        cls.jobs = [j for j in cls.jobs if j.done < j.steps]
                        cls.tick = 0
                        with torch.cuda.stream(cls.stream):
                            cls.rebuild()
                ticks.append((ev, done_now))
                host["launch"] += time.perf_counter() - tl
                while len(ticks) > self.inflight:
                    retire(*ticks.pop(0))
            else:
                # class pacing: a class's batched jobs step (one launch for
                # all) whenever fewer than `inflight` of their steps are
                # queued, each step marked by an event on the class's stream;
                # each solo job likewise, on the class's solo stream, so a
                # large job's temporal launches do not hold back the small ones
                poll()
                stepped = False
                for key, cls in self.classes.items():
                    if not cls.jobs:
                        continue
                    ended = []
                    q = queues.setdefault(("class",) + key, [])
                    if len(q) < self.inflight and cls.has_batched():
                        st = cls.stream.cuda_stream
                        cls.step_batched(st)
                        ev = new_event()
                        ev.record(st)
                        nticks += 1
                        done_now = ([j for j in cls.jobs if not j.solo and j.done >= j.steps]
                                    if cls.batched_ending() else [])
                        q.append((ev, done_now))
                        if done_now:
                            finishing.add(("class",) + key)
                        ended += done_now
                        stepped = True
                    for j in cls.jobs:
                        if not j.solo or j.done >= j.steps:
                            continue
                        qj = queues.setdefault(("job", j.id), [])
                        if len(qj) >= self.inflight:
                            continue
                        if not j.lane_ready:           # its buffers were set up on the class's stream
                            cls.solo_stream.wait_stream(cls.stream)
                            j.lane_ready = True
                        st = cls.solo_stream.cuda_stream
                        cls.step_solo(j, st)
                        ev = new_event()
                        ev.record(st)
                        nticks += 1
                        last = j.done >= j.steps
                        qj.append((ev, [j] if last else []))
                        if last:
                            ended.append(j)
                            finishing.add(("job", j.id))
                        stepped = True
                    if ended:
                        cls.jobs = [j for j in cls.jobs if j.done < j.steps]
                        active = [j for j in active if j.done < j.steps]
                        if any(not j.solo for j in ended):
                            cls.tick = 0
                            with torch.cuda.stream(cls.stream):
                                cls.rebuild()
                host["launch"] += time.perf_counter() - tl
                if not stepped:
                    wait_any()                    # every busy class has its queue full
            if log and now() - last_log >= 0.5:
                last_log = now()
                log(f"serving: t={now():.1f}s steps={nticks} active={len(active)} waiting={len(waiting)} "
                    f"done={len(finished)} reserved={self.reserved / 2**30:.2f} GiB")
        torch.cuda.synchronize()
        wall = now()
        lat = np.array([(j.finished - j.arrival) * 1e3 for j in finished]) if finished else np.zeros(1)
        updates = sum(j.cells * j.steps for j in finished)
        return {
            "jobs": len(finished), "ticks": nticks, "wall_s": wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "gcell_updates_per_s": updates / wall / 1e9, "cell_updates": updates,
            "peak_reserved_gib": self.peak_reserved / 2**30, "budget_gib": self.budget / 2**30,
            # host seconds: admitting (allocation + on-device init), rebuilding work-lists,
            # enqueueing ticks, waiting for the device, idle before arrivals
            "host_s": {k: round(v, 3) for k, v in host.items()},
        }


def main(argv=None) -> int:
    """``python -m cutile.serving [--rate R] [--seconds S] [--seed N] [--inflight K]
    [--pacing class|lockstep]``: the
    serving configuration (Poisson arrivals, mixed jobs) on the current
    device; prints one JSON line of latency and throughput figures."""
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--seconds", type=float, default=60.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--budget-gib", type=float, default=None)
    ap.add_argument("--inflight", type=int, default=3, help="steps a class queues on the device ahead of the host")
    ap.add_argument("--pacing", choices=("class", "lockstep"), default="class",
                    help="classes step at their own pace (default) or together, one step per tick")
    args = ap.parse_args(argv)
    jobs = generate_jobs(args.rate, args.seconds, args.seed)
    srv = StencilServer(hbm_budget=None if args.budget_gib is None else int(args.budget_gib * 2**30),
                        inflight=args.inflight, pacing=args.pacing)
    srv.run(warmup_jobs(srv), realtime=False)       # build every class's kernels before the clock starts
    stats = srv.run(jobs, realtime=True,
                    log=lambda m: print(f"[serving] {m}", file=sys.stderr, flush=True))
    stats.update(rate=args.rate, seconds=args.seconds, seed=args.seed, offered_jobs=len(jobs))
    print(json.dumps(stats), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
