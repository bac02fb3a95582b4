"""Empirical autotuner for the HIP sweep (upstream ``cutile/runtime/autotune.py``).

Same entry point and result type as upstream (``autotune(stencil_fn,
domain, hw, max_candidates, warmup, iters, verbose) -> AutotuneResult``,
``autotune.py:23-33, 207-309``): candidates are generated, the most
promising ones timed on the device, the fastest kept and cached as JSON.
What is tuned is the gfx950 launch shape instead of cuTile tile shapes:

* 3-D: rows x cols of the in-plane tile, threads per workgroup (256 or
  512), and the temporal depth ``T`` when the caller allows it (a temporal
  candidate is timed per step: one launch advances ``T`` steps);
* 2-D: the row-tile width; 1-D: the segment width;
* every candidate's code object is checked first (no scratch, at most 256
  VGPRs, LDS within the 160 KiB of a CU), so nothing that spills is timed.

The cache key is (stencil name, digest of its IR, storage type, size
bucket = each interior side rounded to a power of two, temporal depths
tried, device architecture).  Without a device the first candidate that
passes the resource check is returned and nothing is cached.

``autotune_buckets`` does the same for the batched persistent kernel: each
size class of a batch gets the fastest tile/mapping among a few candidates,
timed on synthetic instances of that class.
"""

from __future__ import annotations

import hashlib
import json
import math
import os
from dataclasses import asdict, dataclass, field
from typing import Optional, Sequence

_CACHE_VERSION = "hip-v1"

# print every measured candidate (``python -m cutile.tune`` sets it)
VERBOSE = False


@dataclass
class AutotuneResult:
    tile_sizes: tuple
    temporal_steps: int
    throughput_gpoints: float
    bandwidth_gbs: float
    time_ms: float = 0.0
    kernel_hints: dict = field(default_factory=dict)


def cache_root() -> str:
    from cutile.runtime.kernel import cache_dir

    d = os.environ.get("CUTILE_AMD_AUTOTUNE_CACHE") or os.path.join(os.path.dirname(cache_dir() or "."), "autotune")
    try:
        os.makedirs(d, exist_ok=True)
    except OSError:
        return ""
    return d


def size_bucket(domain: Sequence[int]) -> tuple:
    return tuple(1 << max(0, round(math.log2(max(1, int(n))))) for n in domain)


def _ir_digest(stencil_fn) -> str:
    return hashlib.sha1(str(stencil_fn._ir).encode()).hexdigest()[:16]


def cache_key(stencil_fn, storage: str, domain: Sequence[int], temporal: Sequence[int], arch: str) -> str:
    raw = "|".join([_CACHE_VERSION, stencil_fn._fn.__name__, _ir_digest(stencil_fn), storage,
                    str(size_bucket(domain)), str(tuple(temporal)), arch])
    return hashlib.sha1(raw.encode()).hexdigest()


def _load(key: str) -> Optional[AutotuneResult]:
    root = cache_root()
    path = os.path.join(root, key + ".json") if root else ""
    if path and os.path.exists(path):
        try:
            d = json.loads(open(path).read())
            d["tile_sizes"] = tuple(d["tile_sizes"])
            return AutotuneResult(**d)
        except Exception:
            return None
    return None


def _save(key: str, res: AutotuneResult) -> None:
    root = cache_root()
    if not root:
        return
    tmp = os.path.join(root, key + ".json.tmp")
    with open(tmp, "w") as f:
        json.dump(asdict(res), f)
    os.replace(tmp, os.path.join(root, key + ".json"))


# ---------------------------------------------------------------------------
# candidates
# ---------------------------------------------------------------------------

def generate_candidates(ndim: int, dtype_bytes: int, num_loads: int,
                        temporal: Sequence[int] = (1,)) -> list:
    """``(tile_sizes, temporal_steps, kernel_hints)`` candidates, the tiling
    pass's choice first, then the rest by closeness to it."""
    from cutile.passes.tiling import TilingPass

    tp = TilingPass(dtype_bytes=dtype_bytes)
    v = tp.vector
    first, first_hints = tp.candidates(ndim, num_loads)[0]
    out = [(tuple(first), 1, dict(first_hints))]
    if ndim == 3:
        chunk = first[0]
        for rows in (4, 8, 16, 32):
            for cols in (32 * v, 64 * v, 128 * v):
                for nt in (256, 512):
                    ntz = cols // v
                    if nt % ntz or nt // ntz > rows:
                        continue
                    hints = {"threads": nt} if nt != 256 else {}
                    out.append(((chunk, rows, cols), 1, hints))
    elif ndim == 2:
        for chunk in (16, 64):
            for cols in (64 * v, 128 * v, 256 * v):
                out.append(((chunk, cols), 1, {}))
    else:
        for cols in (64 * v, 128 * v, 256 * v, 512 * v):
            out.append(((cols,), 1, {}))
    for T in temporal:
        if T > 1:
            out.append((tuple(first), int(T), {}))
    seen, uniq = set(), []
    for c in out:
        k = (c[0], c[1], tuple(sorted(c[2].items())))
        if k not in seen:
            seen.add(k)
            uniq.append(c)
    return uniq


def prioritize(candidates: list, max_candidates: int) -> list:
    """Keep the first (the tiling pass's choice) and the ``max_candidates -
    1`` others closest to it in tile area, temporal candidates last."""
    if not candidates:
        return []
    head, rest = candidates[0], candidates[1:]
    area0 = math.prod(head[0][1:]) if len(head[0]) > 1 else head[0][0]

    def rank(c):
        area = math.prod(c[0][1:]) if len(c[0]) > 1 else c[0][0]
        return (c[1] > 1, abs(math.log2(area / area0)), sorted(c[2].items()), c[0])

    return [head] + sorted(rest, key=rank)[: max(0, max_candidates - 1)]


def _compiled(stencil_fn, tile, T, hints, storage):
    from cutile.runtime.launcher import compile

    return compile(stencil_fn, tile_sizes=tuple(tile), kernel_hints=dict(hints), storage=storage,
                   temporal_steps=T, temporal_mode="recompute" if T > 1 else "zero",
                   temporal_blocking=T > 1, autotune=False)


def fits(result) -> bool:
    """Resource check of a compiled result's kernels (hiprtc code object)."""
    from cutile.runtime import regcheck, rt
    from cutile.runtime.kernel import compile_options

    try:
        res = regcheck.code_object_resources(rt.compile_only(result.hip_source, compile_options("gfx950")))
    except Exception:
        return False
    ks = [r for k, r in res.items() if "_sweep_" in k or "_temporal_" in k]
    return bool(ks) and all(r.stack_bytes == 0 and r.registers <= 256 and r.lds_bytes <= 160 * 1024 for r in ks)


def _time_candidate(result, domain, warmup: int, iters: int) -> Optional[float]:
    """Milliseconds per step of ``result``'s launcher on random data."""
    import torch

    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream

    halo = result.halo_widths
    shape = tuple(int(n) + 2 * h for n, h in zip(domain, halo))
    dt = getattr(torch, result.storage)
    nin = len(result.inputs)
    g = torch.Generator(device="cuda").manual_seed(0)
    ins = [torch.rand(shape, device="cuda", dtype=torch.float32, generator=g).to(dt) for _ in range(nin)]
    out = torch.zeros(shape, device="cuda", dtype=dt)
    launch = result.launcher
    args = ins + [out]
    for _ in range(warmup):
        launch(*args)
    s = default_stream()
    torch.cuda.synchronize()
    e0, e1 = rt.Event(), rt.Event()
    e0.record(s)
    for _ in range(iters):
        launch(*args)
    e1.record(s)
    ms = e0.elapsed_ms(e1) / iters
    del ins, out
    return ms / max(1, result.temporal_steps)


def autotune(stencil_fn, domain=None, hw=None, max_candidates: int = 30, warmup: int = 5,
             iters: int = 20, verbose: bool = False, storage: Optional[str] = None, halo=None,
             temporal: Sequence[int] = (1,)) -> AutotuneResult:
    """Fastest (tile, temporal depth, hints) for ``stencil_fn`` on ``domain``
    (interior extents); cached per (stencil, storage, size bucket, depths,
    arch).  ``temporal`` lists the depths to try: depths above 1 use the
    recompute mode, whose halo is T times the stencil's (the caller's arrays
    must then have that halo)."""
    import torch

    from cutile.frontend.ir import find_func
    from cutile.passes.tiling import unique_access_count
    from cutile.frontend.ir import ApplyOp

    func = find_func(stencil_fn._ir)
    ndim = func.ndim
    store = storage or func.dtype
    esz = {"float64": 8, "float32": 4, "bfloat16": 2}[store]
    if domain is None:
        domain = {1: (1 << 22,), 2: (2048, 2048), 3: (256, 256, 256)}[ndim]
    domain = tuple(int(n) for n in domain)
    loads = next((unique_access_count(op) for op in stencil_fn._ir.ops if isinstance(op, ApplyOp)), 1)
    have_gpu = torch.cuda.is_available()
    arch = "gfx950"
    if have_gpu:
        from cutile.runtime.kernel import current_arch

        arch = current_arch().split(":")[0]
    key = cache_key(stencil_fn, store, domain, temporal, arch)
    if have_gpu:
        hit = _load(key)
        if hit is not None:
            if verbose:
                print(f"autotune: cache hit for {stencil_fn._fn.__name__} ({key[:10]})")
            return hit

    cands = prioritize(generate_candidates(ndim, esz, loads, temporal), max_candidates)
    best = None
    for tile, T, hints in cands:
        try:
            r = _compiled(stencil_fn, tile, T, hints, store)
        except (ValueError, NotImplementedError):
            continue
        if not fits(r):
            continue
        if not have_gpu:
            return AutotuneResult(tuple(r.tile_sizes), T, 0.0, 0.0, 0.0, dict(r.kernel_hints))
        try:
            ms = _time_candidate(r, domain, warmup, iters)
        except Exception:
            ms = None
        if ms is None:
            continue
        if verbose:
            print(f"  {'x'.join(map(str, tile))} T={T} {hints}: {ms:.4f} ms/step")
        if best is None or ms < best[0]:
            best = (ms, r, T)
    if best is None:
        tile, T, hints = cands[0]
        return AutotuneResult(tuple(tile), T, 0.0, 0.0, 0.0, dict(hints))
    ms, r, T = best
    cells = math.prod(domain)
    res = AutotuneResult(tuple(r.tile_sizes), T, cells / (ms * 1e-3) / 1e9,
                         cells * esz * (len(r.inputs) + 1) / (ms * 1e-3) / 1e9, ms, dict(r.kernel_hints))
    _save(key, res)
    if verbose:
        print(f"autotune: best {res.tile_sizes} T={res.temporal_steps} {res.kernel_hints} "
              f"{res.throughput_gpoints:.1f} Gcell/s")
    return res


# ---------------------------------------------------------------------------
# batched kernel: tiles per size class
# ---------------------------------------------------------------------------

def bucket_candidates(vz: int, max_cols: Optional[int], ndim: int = 2) -> list:
    """Bucket shapes tried for instances whose rows hold at most
    ``max_cols`` cells (``None``: the open-ended largest class): 2-D
    plane-mapped tiles and row tiles, or 3-D row tiles streaming along the
    first axis."""
    from cutile.runtime.batched import Bucket

    out = []
    if ndim == 3:
        for ty, nz, xch in ((8, 32, 16), (8, 64, 32), (4, 64, 16), (16, 32, 32), (4, 32, 8)):
            out.append(Bucket("rows", ty, nz * vz, xch, max_cols))
        return out
    for ty, tz in ((16, 16 * vz), (8, 32 * vz), (4, 64 * vz), (16, 32 * vz)):
        out.append(Bucket("plane", ty, tz, 1, max_cols))
    for tz, xch in ((256 * vz, 16), (256 * vz, 64), (128 * vz, 32)):
        out.append(Bucket("rows", 1, tz, xch, max_cols))
    return out


def autotune_buckets(stencil_fn, sides: Sequence[int], storage: Optional[str] = None,
                     bounds: Sequence[Optional[int]] = (160, 384, None), halo: int = 1,
                     iters: int = 10, verbose: bool = False) -> list:
    """Per size class (rows of at most ``bounds[k]`` cells) the bucket shape
    of the batched 2-D kernel that steps that class's instances fastest,
    timed on instances with the given ``sides`` that fall in the class."""
    import torch

    from cutile.runtime import rt
    from cutile.runtime.batched import BatchedStencil, Bucket
    from cutile.runtime.kernel import default_stream

    probe = BatchedStencil(stencil_fn, storage=storage)
    vz = probe.vz
    nd = probe.ndim
    defaults = probe.buckets
    chosen = []
    lo = 0
    for k, bound in enumerate(bounds):
        members = [s for s in sides if s > lo and (bound is None or s <= bound)]
        lo = bound if bound is not None else lo
        if nd == 3 and len(members) > 8:              # 3-D instances are large: a spread of 8
            members = sorted(members)[:: max(1, len(members) // 8)][:8]
        cands = bucket_candidates(vz, bound, nd)
        if not members or not torch.cuda.is_available():
            d = defaults[min(k, len(defaults) - 1)]
            chosen.append(Bucket(d.mapping, d.ty, d.tz, d.xch, bound))
            continue
        dt = getattr(torch, probe.storage)
        arrays = []
        for s in members[:64]:
            cols = -(-s // vz) * vz
            shp = (s + 2 * halo,) * (nd - 1) + (cols + 2 * max(halo, vz),)
            ins = [torch.rand(shp, device="cuda", dtype=torch.float32).to(dt) for _ in range(probe.num_inputs)]
            outs = [torch.zeros(shp, device="cuda", dtype=dt) for _ in range(probe.num_outputs)]
            arrays.append((ins, outs, (s,) * nd))
        best = None
        for b in cands:
            try:
                bs = BatchedStencil(stencil_fn, storage=storage, buckets=[Bucket(b.mapping, b.ty, b.tz, b.xch, None)])
                plan = bs.prepare([(i, o) for i, o, _ in arrays], (halo,) * (nd - 1) + (max(halo, vz),),
                                  interiors=[n for _, _, n in arrays])
                s = default_stream()
                bs.launch(plan, s)
                torch.cuda.synchronize()
                e0, e1 = rt.Event(), rt.Event()
                e0.record(s)
                for _ in range(iters):
                    bs.launch(plan, s)
                e1.record(s)
                ms = e0.elapsed_ms(e1) / iters
            except (ValueError, RuntimeError):
                continue
            if verbose or VERBOSE:
                print(f"  class <= {bound}: {b} {ms:.4f} ms", flush=True)
            if best is None or ms < best[0]:
                best = (ms, b)
        if best is None:
            d = defaults[min(k, len(defaults) - 1)]
            chosen.append(Bucket(d.mapping, d.ty, d.tz, d.xch, bound))
        else:
            chosen.append(best[1])
    return chosen


# ---------------------------------------------------------------------------
# temporal depth: how a block of T steps is split into launches
# ---------------------------------------------------------------------------

def split_candidates(T: int) -> list:
    """Depths d dividing T (d = T: one launch per block; d = 1: T sweeps)."""
    return [d for d in range(T, 0, -1) if T % d == 0]


def autotune_split(stepper, u, min_ms: float = 20.0, verbose: bool = False) -> int:
    """The depth d (a divisor of ``stepper.T``; 1 = single sweeps) whose
    launches advance arrays shaped like ``u`` fastest per step.  Each
    candidate runs whole blocks of T steps, captured in a hipGraph and
    replayed as step loops run, on two scratch copies of ``u`` (its values
    matter: the arithmetic-bound temporal kernels draw less power on an
    all-zero field and keep a higher clock -- on MI355X the 25-point star's
    depth-2 kernel ran 16 % faster on zeros than on the benchmark's field,
    the memory-bound sweep 1 %) for about ``min_ms``; depths whose levels
    fit neither registers nor LDS are skipped.  The choice is kept in the
    tuned tables per (stencil, storage, size bucket, T, mode, arch).  Every
    candidate gives the same states."""
    from cutile.frontend.ir import find_func

    T = stepper.T
    interior = tuple(int(n) - 2 * h for n, h in zip(u.shape, stepper.halo))
    parts = {"what": "temporal_split", "stencil": stepper.fn._fn.__name__, "ir": _ir_digest(stepper.fn),
             "storage": stepper.storage, "size": list(size_bucket(interior)), "T": T, "mode": stepper.mode,
             "ndim": find_func(stepper.fn._ir).ndim}
    val = tuned(parts, lambda: [_measure_split(stepper, u, min_ms, verbose)], _arch())
    return int(val[0]) if val else T // stepper.block_reps


def _measure_split(stepper, u, min_ms: float, verbose: bool) -> int:
    import torch

    from cutile.runtime import rt

    T = stepper.T
    a = u.clone()
    b = u.clone()
    best = None
    saved = dict(stepper._plans)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    for d in split_candidates(T):
        kern = None if d == 1 else stepper.depth_kernel(d)
        if d > 1 and kern is None:
            continue
        stepper._plans[(tuple(u.shape), u.dtype)] = (kern, T // d)
        try:
            # the blocks are captured in a graph and replayed, as step loops
            # run (eager timing would measure host launch overhead)
            with torch.cuda.stream(side):
                s = side.cuda_stream
                stepper.enqueue(a, b, 2 * T, s)                 # build and warm up
                rt.stream_sync(s)
                e0, e1 = rt.Event(), rt.Event()
                e0.record(s)
                stepper.enqueue(a, b, 2 * T, s)
                e1.record(s)
                once = max(e0.elapsed_ms(e1), 1e-3)
                blocks = 2 * max(1, min(32, int(min_ms / once)))
                g = rt.Graph(s)
                with g:
                    stepper.enqueue(a, b, blocks * T, s)
                g.launch(s)
                rt.stream_sync(s)
                # the median of three replays (the depths of the 25-point
                # star are within a few per cent of each other, about the
                # spread of single replays)
                times = []
                for _ in range(3):
                    e0, e1 = rt.Event(), rt.Event()
                    e0.record(s)
                    g.launch(s)
                    e1.record(s)
                    times.append(e0.elapsed_ms(e1))
                ms = sorted(times)[1]
                del g
        except (ValueError, RuntimeError):
            continue
        per_step = ms / (blocks * T)
        if verbose or VERBOSE:
            print(f"  split T={T} as {T // d} x depth {d}: {per_step * 1e3:.2f} us/step", flush=True)
        if best is None or per_step < best[0]:
            best = (per_step, d)
    stepper._plans.clear()
    stepper._plans.update(saved)
    del a, b
    torch.cuda.synchronize()
    return best[1] if best is not None else T // stepper.block_reps


# ---------------------------------------------------------------------------
# tuned tables: measured choices kept per (what, stencil, storage, sizes, arch)
# ---------------------------------------------------------------------------
#
# A measured choice is looked up first in the user's autotune cache, then in
# the table shipped with the package (``cutile/tuned/<arch>.json``, written
# by ``export_tuned`` from a tuning run on that architecture); a miss with a
# device present is measured and cached.  Without a device and without an
# entry the caller's default applies.

# True: entries of the shipped table are ignored, so a miss in the user cache
# is measured (a re-tune after the kernels changed)
IGNORE_SHIPPED = False


def tuned_path(arch: str) -> str:
    return os.path.join(os.path.dirname(os.path.dirname(__file__)), "tuned", f"{arch}.json")


def _parts_key(parts: dict) -> str:
    return hashlib.sha1(json.dumps(parts, sort_keys=True).encode()).hexdigest()


def _shipped(arch: str) -> dict:
    path = tuned_path(arch)
    if not os.path.exists(path):
        return {}
    try:
        return {_parts_key(e["key"]): e["value"] for e in json.load(open(path))["entries"]}
    except (OSError, ValueError, KeyError):
        return {}


def tuned(parts: dict, measure, arch: str):
    """The cached value of the measured choice described by ``parts`` (a
    JSON-able dict), measuring it with ``measure()`` on a miss when a device
    is present; None when there is neither an entry nor a device."""
    import torch

    parts = dict(parts, arch=arch, version=_CACHE_VERSION)
    key = _parts_key(parts)
    root = cache_root()
    path = os.path.join(root, f"tuned-{key}.json") if root else ""
    if path and os.path.exists(path):
        try:
            return json.load(open(path))["value"]
        except (OSError, ValueError, KeyError):
            pass
    hit = None if IGNORE_SHIPPED else _shipped(arch).get(key)
    if hit is not None:
        return hit
    if not torch.cuda.is_available():
        return None
    value = measure()
    if path and value is not None:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump({"key": parts, "value": value}, f)
        os.replace(tmp, path)
    return value


def export_tuned(arch: str, path: Optional[str] = None) -> str:
    """Merge the shipped table for ``arch`` (the package's
    ``tuned/<arch>.json``), the file at ``path`` and every tuned entry of
    the user cache into ``path`` (default: the shipped table)."""
    import glob

    path = path or tuned_path(arch)
    entries = {}
    for src in dict.fromkeys((tuned_path(arch), path)):     # the shipped table, then the target's
        if os.path.exists(src):
            for e in json.load(open(src))["entries"]:
                entries[_parts_key(e["key"])] = e
    for f in glob.glob(os.path.join(cache_root(), "tuned-*.json")):
        e = json.load(open(f))
        if e["key"].get("arch") == arch:
            entries[_parts_key(e["key"])] = e
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"arch": arch, "entries": sorted(entries.values(), key=lambda e: json.dumps(e["key"],
                                                                                               sort_keys=True))},
                  f, indent=1)
    return path


def _arch() -> str:
    import torch

    if torch.cuda.is_available():
        from cutile.runtime.kernel import current_arch

        return current_arch().split(":")[0]
    return "gfx950"


def tuned_buckets(stencil_fn, sides: Sequence[int], storage: Optional[str] = None,
                  bounds: Sequence[Optional[int]] = (160, 384, None), halo: int = 1) -> Optional[list]:
    """Tile buckets of the batched kernel for instances with these ``sides``
    (one per size class of ``bounds``), from the tuned tables or measured
    with ``autotune_buckets``; None when neither is available (the
    kernel's defaults then apply).  Keyed by the stencil, its storage, the
    class bounds, the power-of-two size buckets of the sides and the arch."""
    from cutile.frontend.ir import find_func
    from cutile.runtime.batched import Bucket

    fns = list(stencil_fn) if isinstance(stencil_fn, (list, tuple)) else [stencil_fn]
    func = find_func(fns[0]._ir)
    parts = {"what": "batched_buckets", "stencil": [f._fn.__name__ for f in fns],
             "ir": [_ir_digest(f) for f in fns], "storage": storage or func.dtype,
             "bounds": [b for b in bounds], "sizes": sorted({size_bucket([int(s)])[0] for s in sides})}

    def measure():
        got = autotune_buckets(stencil_fn, sides, storage=storage, bounds=bounds, halo=halo)
        return [[b.mapping, b.ty, b.tz, b.xch, b.max_cols] for b in got]

    val = tuned(parts, measure, _arch())
    if val is None:
        return None
    return [Bucket(m, int(ty), int(tz), int(xch), mc) for m, ty, tz, xch, mc in val]
