"""Runtime object behind every emitted ``launch_<name>``.

Replaces the cuTile ``ct.launch`` + CuPy glue of the upstream launcher
(``cutile/lowering/target_to_python.py:305-399``): it JIT-builds the emitted
HIP source with hiprtc on first use (cached in process and on disk), picks
the vector or scalar variant from the arrays' alignment, packs the by-value
``SweepArgs`` block and launches through the C ABI on the caller's stream
(``None`` = torch's current stream, as upstream used CuPy's current stream,
``stencil_to_target.py:671-672``).  Launches are asynchronous.
"""

from __future__ import annotations

import ctypes
import functools
import math
import os
import threading
from typing import Optional, Sequence

from cutile.runtime import rt

MAXF = 12

_TORCH_CODES = {}


def _torch():
    import torch

    return torch


def dtype_code(t) -> int:
    torch = _torch()
    if t.dtype == torch.float64:
        return rt.ST_F64
    if t.dtype == torch.float32:
        return rt.ST_F32
    if t.dtype == torch.bfloat16:
        return rt.ST_BF16
    raise TypeError(f"unsupported dtype {t.dtype}")


STORAGE_TORCH = {"float64": "float64", "float32": "float32", "bfloat16": "bfloat16"}


class SweepArgs(ctypes.Structure):
    """Mirror of ``st::SweepArgs`` in ``csrc/stencil_core.h``."""

    _fields_ = [
        ("inp", ctypes.c_void_p * MAXF), ("out", ctypes.c_void_p * MAXF),
        ("s0", ctypes.c_int64), ("s1", ctypes.c_int64), ("s2", ctypes.c_int64),
        ("h0", ctypes.c_int), ("h1", ctypes.c_int), ("h2", ctypes.c_int),
        ("n0", ctypes.c_int), ("n1", ctypes.c_int), ("n2", ctypes.c_int),
        ("ntz", ctypes.c_int), ("nty", ctypes.c_int), ("nch", ctypes.c_int),
        ("xch", ctypes.c_int), ("nxcd", ctypes.c_int), ("tmode", ctypes.c_int),
        ("yslab", ctypes.c_int), ("wrap", ctypes.c_int), ("nby", ctypes.c_int), ("nbz", ctypes.c_int),
    ]


def map3_shape(ndim: int, shape: Sequence[int]) -> tuple:
    if ndim == 1:
        return (1, 1, shape[0])
    if ndim == 2:
        return (shape[0], 1, shape[1])
    return tuple(shape)


def default_stream():
    """Raw handle of torch's current stream on the current device."""
    torch = _torch()
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return raw(torch.cuda.current_device())
    return torch.cuda.current_stream().cuda_stream


def cache_dir() -> str:
    d = os.environ.get("CUTILE_AMD_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "cutile_amd", "code"))
    try:
        os.makedirs(d, exist_ok=True)
    except OSError:
        return ""
    return d


_hdr_digest: list = []


def header_digest() -> str:
    """Hash of the device headers the generated sources include: part of the
    options, so cached code objects are rebuilt when a header changes."""
    if not _hdr_digest:
        names = sorted(n for n in os.listdir(rt.CSRC_DIR) if n.endswith(".h"))
        blob = b"".join(n.encode() + b"\0" + open(os.path.join(rt.CSRC_DIR, n), "rb").read() for n in names)
        _hdr_digest.append(rt.cache_key(blob.decode("latin-1"), []))
    return _hdr_digest[0]


def compile_options(arch: Optional[str] = None) -> list:
    arch = arch or "gfx950"
    return [f"--offload-arch={arch}", "-O3", "-std=c++17", "-ffp-contract=off",
            "-fno-gpu-flush-denormals-to-zero",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            f"-DST_HDR_DIGEST=0x{header_digest()}ull", f"-I{rt.CSRC_DIR}"]


def aot_dir() -> str:
    """Code objects written ahead of time (``precompile``; read-only at run
    time): ``$CUTILE_AMD_AOT`` or the package's ``_aot`` folder."""
    return os.environ.get("CUTILE_AMD_AOT", os.path.join(os.path.dirname(os.path.dirname(__file__)), "_aot"))


def precompile(source: str, arch: str = "gfx950", directory: Optional[str] = None) -> str:
    """Build ``source`` for ``arch`` without a device and store the code
    object where a later launch on such a device looks first.  Returns its
    path."""
    d = directory or aot_dir()
    os.makedirs(d, exist_ok=True)
    opts = compile_options(arch)
    path = os.path.join(d, rt.cache_key(source, opts) + ".hsaco")
    if not os.path.exists(path):
        code = rt.compile_only(source, opts)
        tmp = f"{path}.tmp{os.getpid()}"
        with open(tmp, "wb") as f:
            f.write(code)
        os.replace(tmp, path)
    return path


def code_object(source: str, arch: str, name: str = "stencil.hip") -> bytes:
    """The code object of ``source`` for ``arch``: the one written ahead of
    time when there is one, else the on-disk cache, else a hiprtc build
    (then cached).  Needs no device."""
    opts = compile_options(arch)
    key = rt.cache_key(source, opts)
    path = os.path.join(aot_dir(), key + ".hsaco")
    if os.path.exists(path):
        with open(path, "rb") as f:
            return f.read()
    d = cache_dir()
    path = os.path.join(d, key + ".hsaco") if d else ""
    if path and os.path.exists(path):
        with open(path, "rb") as f:
            code = f.read()
        if code:
            return code
    code = rt.compile_only(source, opts, name=name)
    if path:
        tmp = f"{path}.tmp{os.getpid()}"
        try:
            with open(tmp, "wb") as f:
                f.write(code)
            os.replace(tmp, path)
        except OSError:
            pass
    return code


_notes: dict = {}

# Limits every launched kernel meets: no scratch (register spills), at most
# 256 VGPR+AGPR per lane, the 160 KiB LDS of a gfx950 CU.  A tuning run of
# 7-point tiles whose code objects broke the first two ended in a GPU memory
# fault; such variants are refused before they can reach a device.
MAX_REGISTERS = 256
MAX_LDS_BYTES = 160 * 1024


def code_resources(source: str, arch: Optional[str] = None) -> dict:
    """Kernel symbol -> ``regcheck.Resources`` (VGPR+AGPR, scratch, LDS)
    read from the notes of ``source``'s code object for ``arch`` (default:
    the current device's, or gfx950 without a device)."""
    from cutile.runtime import regcheck

    arch = arch or default_arch()
    k = (rt.cache_key(source, compile_options(arch)), arch)
    res = _notes.get(k)
    if res is None:
        res = regcheck.code_object_resources(code_object(source, arch))
        _notes[k] = res
    return res


def default_arch() -> str:
    torch = _torch()
    return current_arch() if torch.cuda.is_available() else "gfx950"


def assert_launchable(source: str, symbol: str, arch: Optional[str] = None) -> None:
    """Raise ``RuntimeError`` unless kernel ``symbol`` of ``source`` uses no
    scratch, at most ``MAX_REGISTERS`` VGPR+AGPR and at most
    ``MAX_LDS_BYTES`` of LDS (read from its code object, before anything is
    loaded on a device)."""
    res = code_resources(source, arch).get(symbol)
    if res is None:
        if regcheck_available():
            raise RuntimeError(f"{symbol}: kernel not found in its code object")
        return          # no notes reader: the loaded function's attributes are checked instead
    why = []
    if res.stack_bytes > 0:
        why.append(f"{res.stack_bytes} B of scratch per thread (register spills)")
    if res.registers > MAX_REGISTERS:
        why.append(f"{res.registers} VGPR+AGPR > {MAX_REGISTERS}")
    if res.lds_bytes > MAX_LDS_BYTES:
        why.append(f"{res.lds_bytes} B of LDS > {MAX_LDS_BYTES}")
    if why:
        raise RuntimeError(f"{symbol}: {', '.join(why)}; variant refused, not launched "
                           f"(choose a smaller tile)")


def regcheck_available() -> bool:
    from cutile.runtime import regcheck

    return regcheck.readelf_path() is not None


def check_loaded(function, symbol: str) -> None:
    """The same limits on a loaded function's attributes (scratch, VGPRs)."""
    res = rt.function_resources(function)
    if res.local_bytes > 0 or res.num_regs > MAX_REGISTERS:
        raise RuntimeError(f"{symbol}: {res.local_bytes} B of scratch per thread, {res.num_regs} VGPRs; "
                           f"variant refused, not launched")


def load_checked(source: str, name: str, symbols: Sequence[str]):
    """Module of ``source`` on the current device, with every kernel in
    ``symbols`` checked against the launch limits first: from the code
    object's notes before it is loaded, and from the loaded functions'
    attributes.  Returns (module, {symbol: function})."""
    arch = current_arch()
    for s in symbols:
        assert_launchable(source, s, arch)
    mod = rt.load_code(code_object(source, arch, name))
    fns = {}
    for s in symbols:
        fn = rt.get_function(mod, s)
        check_loaded(fn, s)
        fns[s] = fn
    return mod, fns


def load_or_compile(source: str, name: str):
    """Module for ``source`` on the current device: a code object written
    ahead of time when there is one, else a JIT build (cached on disk)."""
    return rt.load_code(code_object(source, current_arch(), name))


_arch_cache: dict = {}


def current_arch() -> str:
    dev = _torch().cuda.current_device()
    if dev not in _arch_cache:
        _arch_cache[dev] = rt.device_arch(dev)
    return _arch_cache[dev]


_cu_cache: dict = {}


def cu_count() -> int:
    dev = _torch().cuda.current_device()
    if dev not in _cu_cache:
        _cu_cache[dev] = max(1, rt.device_info(dev).compute_units)
    return _cu_cache[dev]


def chunk_length(n0: int, tiles: int, default: int, queue_depth: int, per_cu: int = 1) -> int:
    """Planes per workgroup along the streaming axis (1-D and 2-D sweeps;
    3-D sweeps use ``sweep_chunk_length``).

    The grid gets about ``per_cu`` workgroups per CU: ``ceil(per_cu * CUs /
    tiles)`` chunks of ``ceil(n0 / chunks)`` planes, capped at the compiled
    chunk ``default``.  Chunks are kept at 4 queue refills or more (the
    register-queue prologue, ``queue_depth - 1`` planes, then stays a small
    fraction of the chunk) when the grid still fills the CUs that way; a
    small domain gets shorter chunks instead, since there the number of
    workgroups, not the prologue's re-reads, sets the time.  On MI355X one
    long-streaming workgroup per CU beat shorter chunks and several resident
    workgroups per CU for the 3-D sweeps at 512^3 (fewer prologue re-reads,
    neighbouring tiles advancing in step); the one-row tiles of 2-D sweeps
    ran 3-8 % faster at two per CU (fused Gray-Scott at 4096^2, all four
    precisions)."""
    slots = max(1, per_cu) * cu_count()
    chunks = max(1, -(-slots // max(1, tiles)))
    x = min(max(1, default), -(-n0 // chunks))
    floor = 4 * max(1, queue_depth - 1)
    if x < floor and max(1, tiles) * -(-n0 // floor) >= slots:
        x = floor
    return max(1, min(x, n0))


def run_prepared(rec, stream=None) -> None:
    """Launch a packed record of ``StencilKernel.prepare`` on ``stream``; a
    ``None`` record (empty interior) launches nothing, as ``launch`` does."""
    if rec is None:
        return
    fn, nblocks, nt, addr, size, _args = rec
    if stream is None:
        stream = default_stream()
    rt.check(rt.lib().st_launch(fn, nblocks, 1, 1, nt, 1, 1, 0, stream, addr, size), "st_launch")


def sweep_chunk_length(n0: int, tiles: int, default: int, queue_depth: int) -> int:
    """Planes per workgroup of a 3-D sweep: one long-streaming workgroup
    per CU, the chunk count chosen by the round model of
    ``tblock_chunk_length`` (prologue of ``queue_depth - 1`` planes per
    chunk, one slot per CU).  A grid a little larger than the CU count
    leaves a few CUs with two workgroups and the launch takes about twice as
    long: at 384^3 the 7-point sweep's 288 workgroups (48 tiles x 6 chunks)
    ran at 406 Gcell/s against 597 at 256^3, where 256 workgroups fill the
    CUs once.  Of chunk counts within 3 % of the best modelled time the
    smallest is taken (fewer prologue re-reads)."""
    slots, pro = cu_count(), max(0, queue_depth - 1)
    cands = []
    for c in range(max(1, -(-n0 // max(1, default))), n0 + 1):
        x = -(-n0 // c)
        cands.append((-(-(tiles * c) // slots) * (x + pro), c, x))
        if tiles * c > 64 * slots:
            break
    tmin = min(t for t, _, _ in cands)
    return max(1, min((c, x) for t, c, x in cands if t <= 1.03 * tmin)[1])


@functools.lru_cache(maxsize=4096)
def tblock_chunk_length(n0: int, tiles: int, default: int, depth: int, slots: int, cus: int = 256) -> int:
    """Planes per workgroup of a temporal-blocking launch.

    ``slots`` workgroups run at once (CUs x workgroups per CU); a launch of
    ``tiles * c`` workgroups takes ``ceil(tiles * c / slots)`` rounds of
    ``ceil(n0 / c) + depth`` steps each (every chunk warms the skewed levels
    up over ``depth`` steps), and a step takes longer the more workgroups
    share a CU (about as the square root of their number: 2-D heat at
    1024^2, depth 4, ran 4.0 us/step with 2 workgroups a CU and 4.5 with
    4).  The chunk count c minimising that product is taken: a partly
    filled last round and short chunks both cost time (on MI355X a
    2.25-round launch ran 25 % slower than a 2-round one of the same
    kernel).  Chunks stay within the compiled chunk length ``default``."""
    cus = max(1, cus)
    per_cu_max = max(1, slots // cus)
    best = None
    for c in range(max(1, -(-n0 // max(1, default))), n0 + 1):
        x = -(-n0 // c)
        shared = min(-(-(tiles * c) // cus), per_cu_max)
        t = -(-(tiles * c) // max(1, slots)) * (x + depth) * math.sqrt(shared)
        if best is None or t < best[0]:
            best = (t, x)
        if tiles * c > 64 * max(1, slots):
            break
    return max(1, best[1])


def workgroups_per_cu(res, threads: int) -> int:
    """Workgroups of ``threads`` threads one CU holds at once, from a
    kernel's registers and LDS (4 SIMDs of 512 VGPRs per lane, 160 KiB LDS,
    at most 8 waves per SIMD)."""
    waves = -(-threads // 64)
    per_simd = -(-waves // 4)
    regs = max(8, -(-max(1, res.num_regs) // 8) * 8)
    by_regs = 512 // (regs * per_simd) if regs * per_simd <= 512 else 0
    by_waves = 8 // per_simd
    by_lds = (160 * 1024) // res.shared_bytes if res.shared_bytes > 0 else by_waves
    return max(1, min(by_regs, by_waves, by_lds))


def xcd_count(arch: str) -> int:
    # CDNA3/CDNA4 parts are built from 8 XCDs, each with its own L2
    return 8 if arch.startswith(("gfx942", "gfx950")) else 1


class StencilKernel:
    """A compiled-on-demand family of sweep variants for one stencil."""

    def __init__(self, name: str, ndim: int, num_inputs: int, num_outputs: int, source: str,
                 variants: dict, storage: str, xch: int, queue_depth: int = 3, temporal_steps: int = 1,
                 tblock_depth: int = 0, brick: Optional[tuple] = None):
        # variants: {"vec": (symbol, ty, tz, vz, nt), "scalar": (...)}; temporal blocking:
        # "tblock" / "tblock_s" (register/LDS kernel, vector / scalar; ty, tz = output tile)
        # or "temporal" (LDS-ring kernel)
        self.name = name
        self.ndim = ndim
        self.num_inputs = num_inputs
        self.num_outputs = num_outputs
        self.source = source
        self.variants = variants
        self.storage = storage
        self.xch = xch
        self.queue_depth = queue_depth
        self.temporal_steps = temporal_steps
        self.tblock_depth = tblock_depth
        self.brick = tuple(brick) if brick else None     # brick extents in the 3-D map
        self._tb_chunks: dict = {}
        self._launches: dict = {}
        self._modules: dict = {}
        self._fns: dict = {}
        self._lock = threading.Lock()

    # -- compilation ------------------------------------------------------
    def assert_launchable(self, key: str) -> None:
        """Refuse variant ``key`` (every kind: vector, scalar, temporal,
        brick-major) when its code object uses scratch, more than 256
        VGPR+AGPR or more than 160 KiB of LDS.  Needs no device."""
        assert_launchable(self.source, self.variants[key][0])

    def _function(self, key: str):
        dev = _torch().cuda.current_device()
        k = (dev, key)
        fn = self._fns.get(k)
        if fn is None:
            with self._lock:
                sym = self.variants[key][0]
                arch = current_arch()
                assert_launchable(self.source, sym, arch)
                mod = self._modules.get(dev)
                if mod is None:
                    rt.check(rt.lib().st_init(dev), "st_init")
                    mod = load_or_compile(self.source, f"{self.name}.hip")
                    self._modules[dev] = mod
                fn = rt.get_function(mod, sym)
                check_loaded(fn, sym)
                self._fns[k] = fn
        return fn

    def resources(self, key: str = "vec") -> rt.Resources:
        return rt.function_resources(self._function(key))

    # -- launch -----------------------------------------------------------
    def check_arrays(self, inputs: Sequence, outputs: Sequence) -> None:
        torch = _torch()
        if len(inputs) != self.num_inputs or len(outputs) != self.num_outputs:
            raise ValueError(f"launch_{self.name}: expected {self.num_inputs} inputs and "
                             f"{self.num_outputs} outputs, got {len(inputs)} and {len(outputs)}")
        ref = inputs[0]
        want = getattr(torch, STORAGE_TORCH[self.storage])
        for t in list(inputs) + list(outputs):
            if not isinstance(t, torch.Tensor) or not t.is_cuda:
                raise TypeError(f"launch_{self.name}: arrays must be torch tensors on a ROCm device")
            if t.shape != ref.shape:
                raise ValueError(f"launch_{self.name}: shape {tuple(t.shape)} != {tuple(ref.shape)}")
            if t.dtype != want:
                raise TypeError(f"launch_{self.name}: dtype {t.dtype} != {want}")
            if not t.is_contiguous():
                raise ValueError(f"launch_{self.name}: arrays must be C-contiguous")
            if t.dim() != self.ndim:
                raise ValueError(f"launch_{self.name}: expected {self.ndim}-D arrays")
        for o in outputs:
            for i in inputs:
                if o.data_ptr() == i.data_ptr():
                    raise ValueError(f"launch_{self.name}: output aliases an input (out-of-place only)")

    def vector_ok(self, inputs, outputs, halo, vz: int) -> bool:
        """Rows of every array start and end on a ``vz``-element boundary."""
        s = map3_shape(self.ndim, inputs[0].shape)
        h = map3_shape(self.ndim, halo) if self.ndim != 1 else (0, 0, halo[0])
        if self.ndim == 2:
            h = (halo[0], 0, halo[1])
        ok = (h[2] % vz == 0) and (s[2] % vz == 0)
        esz = inputs[0].element_size()
        return ok and all((t.data_ptr() % (vz * esz)) == 0 for t in list(inputs) + list(outputs))

    def pick(self, inputs, outputs, halo) -> str:
        return "vec" if self.vector_ok(inputs, outputs, halo, self.variants["vec"][3]) else "scalar"

    def temporal_key(self, inputs, outputs, halo) -> str:
        if "tblock" in self.variants:
            if self.vector_ok(inputs, outputs, halo, self.variants["tblock"][3]):
                return "tblock"
            return "tblock_s"
        return "temporal"

    yslab = 0

    def args(self, inputs, outputs, halo, key: str, xch: Optional[int] = None) -> tuple:
        _, ty, tz, vz, nt = self.variants[key]
        s = map3_shape(self.ndim, inputs[0].shape)
        if self.ndim == 1:
            h = (0, 0, halo[0])
        elif self.ndim == 2:
            h = (halo[0], 0, halo[1])
        else:
            h = tuple(halo)
        n = tuple(a - 2 * b for a, b in zip(s, h))
        if min(n) <= 0:
            return None, 0
        a = SweepArgs()
        for k, t in enumerate(inputs):
            a.inp[k] = t.data_ptr()
        for k, t in enumerate(outputs):
            a.out[k] = t.data_ptr()
        if s[1] * s[2] * inputs[0].element_size() >= 1 << 31:
            raise ValueError(f"a plane of {s[1]}x{s[2]} cells exceeds the 2 GiB range of one buffer resource")
        a.s0, a.s1, a.s2 = s
        a.h0, a.h1, a.h2 = h
        a.n0, a.n1, a.n2 = n
        a.ntz = -(-n[2] // tz)
        a.nty = -(-n[1] // ty)
        if key.startswith("tblock"):
            # T levels skewed by a queue length each: depth planes of warm-up per chunk
            if not xch:
                ck = (key, n[0], a.ntz * a.nty)
                xch = self._tb_chunks.get(ck)
                if xch is None:
                    slots = cu_count() * workgroups_per_cu(self.resources(key), nt)
                    xch = tblock_chunk_length(n[0], a.ntz * a.nty, self.xch, self.tblock_depth, slots, cu_count())
                    self._tb_chunks[ck] = xch
            x = xch
        else:
            # the LDS-ring temporal kernel warms up (queue_depth - 1) * T planes per chunk
            depth = self.queue_depth if key != "temporal" else (self.queue_depth - 1) * self.temporal_steps + 1
            if xch:
                x = xch
            elif self.ndim == 3:
                x = sweep_chunk_length(n[0], a.ntz * a.nty, self.xch, depth)
            else:
                x = chunk_length(n[0], a.ntz * a.nty, self.xch, depth, 2)
        a.nch = -(-n[0] // x)
        a.xch = x
        a.nxcd = xcd_count(current_arch())
        a.yslab = self.yslab
        return a, a.ntz * a.nty * a.nch

    def launch_temporal(self, inputs: Sequence, outputs: Sequence, halo: Sequence[int], stream=None,
                        mode: int = 0, xch: Optional[int] = None, validate: bool = True,
                        periodic_axes: Sequence[int] = ()) -> None:
        """Advance ``T`` steps in one launch (in-kernel temporal blocking).
        ``mode`` 0: intermediate levels are zero outside the interior (the
        semantics of T chained sweeps through zero-halo scratch arrays);
        1: intermediate levels are recomputed in a halo at least T*r deep
        (T steps with a periodic refill when the halo was filled first).
        ``periodic_axes`` (register/LDS kernel only): axes along which reads
        past the interior wrap around, so T periodic steps need no halo
        at all; the output's halo is not written."""
        # repeated launches on the same arrays reuse the packed arguments (as
        # in ``launch``; the mode is part of the signature)
        sig = (tuple([(t.data_ptr(), t.shape, t.stride(), t.dtype, t.get_device()) for t in (*inputs, *outputs)]),
               tuple(halo), xch, tuple(periodic_axes), int(mode), "temporal")
        hit = self._launches.get(sig)
        if hit is None:
            hit = self.prepare(inputs, outputs, halo, temporal=True, mode=mode, xch=xch,
                               periodic_axes=periodic_axes, validate=validate)
            if hit is None:
                return
            if len(self._launches) >= 256:
                self._launches.clear()
            self._launches[sig] = hit
        run_prepared(hit, stream)

    def prepare(self, inputs: Sequence, outputs: Sequence, halo: Sequence[int], temporal: bool = False,
                mode: int = 0, xch: Optional[int] = None, periodic_axes: Sequence[int] = (),
                validate: bool = True):
        """The packed launch of one sweep (``temporal``: one block of T
        steps, as ``launch_temporal``) on these arrays, for
        ``run_prepared``; None for an empty interior.  A caller that
        launches the same arrays many times (the serving loop's solo jobs)
        keeps it rather than relying on the signature cache."""
        if len(inputs) != self.num_inputs or len(outputs) != self.num_outputs:
            self.check_arrays(inputs, outputs)
        key = self.temporal_key(inputs, outputs, halo) if temporal else self.pick(inputs, outputs, halo)
        # the variant's resources are checked before the arrays and the device
        self.assert_launchable(key)
        if validate:
            self.check_arrays(inputs, outputs)
        args, nblocks = self.args(inputs, outputs, halo, key, xch)
        if args is None:
            return None
        if temporal:
            args.tmode = int(mode)
            if periodic_axes:
                if not key.startswith("tblock"):
                    raise ValueError(f"launch_{self.name}: periodic reads need the register/LDS temporal kernel")
                args.wrap = self.wrap_mask(periodic_axes, args, need_halo=False)
        elif periodic_axes:
            args.wrap = self.wrap_mask(periodic_axes, args)
        # everything st_launch takes but the stream (args stays referenced)
        return (self._function(key), nblocks, self.variants[key][4], ctypes.addressof(args), ctypes.sizeof(args),
                args)

    def launch_bricks(self, inputs: Sequence, outputs: Sequence, shape: Sequence[int], halo: Sequence[int],
                      stream=None, xch: Optional[int] = None) -> None:
        """One sweep on brick-major arrays (``cutile.runtime.bricks.to_bricks``
        of flat arrays of ``shape``, halo included): the kernel reads and
        writes the bricks in place, addressing cell (x, y, z) through its
        brick; only interior cells of the outputs are written."""
        torch = _torch()
        if self.brick is None:
            raise ValueError(f"launch_{self.name}: compiled without a brick layout (layout='bricked')")
        shape = tuple(int(x) for x in shape)
        if len(shape) != self.ndim:
            raise ValueError(f"launch_{self.name}: expected a {self.ndim}-D shape, got {shape}")
        b = self.brick[2]
        want = tuple(-(-x // b) for x in shape) + (b,) * self.ndim
        dt = getattr(torch, STORAGE_TORCH[self.storage])
        if len(inputs) != self.num_inputs or len(outputs) != self.num_outputs:
            raise ValueError(f"launch_{self.name}: expected {self.num_inputs} inputs and {self.num_outputs} outputs")
        for t in list(inputs) + list(outputs):
            if not isinstance(t, torch.Tensor) or not t.is_cuda or tuple(t.shape) != want or t.dtype != dt \
                    or not t.is_contiguous():
                raise ValueError(f"launch_{self.name}: brick arrays must be contiguous {dt} device tensors of "
                                 f"shape {want} (bricks of {b} for a {shape} array)")
        esz = inputs[0].element_size()

        class _View:
            def __init__(self, t):
                self.shape, self._p = shape, t.data_ptr()

            def data_ptr(self):
                return self._p

            def element_size(self):
                return esz

        vi, vo = [_View(t) for t in inputs], [_View(t) for t in outputs]
        key = "brick" if self.vector_ok(vi, vo, halo, self.variants["brick"][3]) else "brick_s"
        args, nblocks = self.args(vi, vo, halo, key, xch)
        if args is None:
            return
        s3 = map3_shape(self.ndim, shape)
        args.nby = -(-s3[1] // self.brick[1])
        args.nbz = -(-s3[2] // self.brick[2])
        fn = self._function(key)
        rt.launch(fn, (nblocks, 1, 1), (self.variants[key][4], 1, 1), args,
                  stream if stream is not None else default_stream())

    def wrap_mask(self, periodic_axes: Sequence[int], args, need_halo: bool = True) -> int:
        """Bits of the 3-D map for the periodic array axes (the sweep wraps
        reads past the interior instead of reading the halo); for the sweep
        every such axis needs an interior at least as wide as its halo (the
        temporal kernel wraps any distance)."""
        if self.ndim == 1:
            bits = {0: 4}
        elif self.ndim == 2:
            bits = {0: 1, 1: 4}
        else:
            bits = {0: 1, 1: 2, 2: 4}
        mask = 0
        for ax in periodic_axes:
            b = bits[int(ax)]
            n, h = {1: (args.n0, args.h0), 2: (args.n1, args.h1), 4: (args.n2, args.h2)}[b]
            if need_halo and n < h:
                raise ValueError(f"periodic reads in the sweep need interior >= halo on axis {ax} "
                                 f"(interior {n}, halo {h})")
            mask |= b
        return mask

    def launch(self, inputs: Sequence, outputs: Sequence, halo: Sequence[int], stream=None,
               xch: Optional[int] = None, validate: bool = True, periodic_axes: Sequence[int] = ()) -> None:
        """One sweep.  ``periodic_axes``: array axes along which the sweep
        reads periodically (past the interior it wraps around instead of
        reading the halo), so no halo fill is needed between steps; the
        output's halo is not written."""
        # repeated launches on the same arrays (a time loop) reuse the
        # validated, packed arguments: the signature holds everything the
        # checks and the arguments depend on
        sig = (tuple([(t.data_ptr(), t.shape, t.stride(), t.dtype, t.get_device()) for t in (*inputs, *outputs)]),
               tuple(halo), xch, tuple(periodic_axes))
        hit = self._launches.get(sig)
        if hit is None:
            hit = self.prepare(inputs, outputs, halo, xch=xch, periodic_axes=periodic_axes, validate=validate)
            if hit is None:
                return
            if len(self._launches) >= 256:
                self._launches.clear()
            self._launches[sig] = hit
        run_prepared(hit, stream)
