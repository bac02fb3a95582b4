"""ctypes binding of ``libstencil_rt.so`` (the C ABI in ``include/stencil_rt.h``).

Every device action of the package goes through this module.  The library
is built in-tree by ``__graft_entry__.build()``; if it is missing, importing
works but any call raises :class:`RuntimeLibraryError` -- there is no
fallback path.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional, Sequence

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libstencil_rt.so")
CSRC_DIR = os.path.join(os.path.dirname(PKG_DIR), "csrc")

ST_F64, ST_F32, ST_BF16 = 0, 1, 2
BC_CODES = {None: 0, "none": 0, "dirichlet": 1, "neumann": 2, "periodic": 3, "reflecting": 4}

# exported symbols (kept in step with include/stencil_rt.h; tests check both)
EXPORTS = (
    "st_last_error", "st_init", "st_device_count", "st_device_info_get", "st_mem_info",
    "st_compile_only", "st_free", "st_compile", "st_load_code", "st_unload", "st_get_function",
    "st_launch", "st_function_resources", "st_fill_halo", "st_fill_halo_batched", "st_copy", "st_copy_wide",
    "st_diff_norms", "st_residual_norms", "st_residual_norms_static", "st_field_digest", "st_bricks",
    "st_event_create", "st_event_record", "st_event_elapsed_ms", "st_event_destroy", "st_event_sync",
    "st_event_query", "st_stream_sync", "st_device_sync", "st_graph_begin", "st_graph_end", "st_graph_launch",
    "st_graph_destroy",
)


class RuntimeLibraryError(RuntimeError):
    pass


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int), ("max_threads_per_block", ctypes.c_int),
        ("lds_per_block", ctypes.c_size_t), ("total_mem", ctypes.c_size_t),
        ("free_mem", ctypes.c_size_t), ("clock_khz", ctypes.c_int),
        ("mem_clock_khz", ctypes.c_int), ("mem_bus_width", ctypes.c_int),
        ("warp_size", ctypes.c_int),
    ]


class FillDesc(ctypes.Structure):
    """Mirror of ``st_fill_desc``."""

    _fields_ = [("ptr", ctypes.c_void_p), ("shape", ctypes.c_int64 * 4), ("halo", ctypes.c_int64 * 4),
                ("kinds", ctypes.c_int * 8), ("values", ctypes.c_double * 8), ("ndim", ctypes.c_int),
                ("dtype", ctypes.c_int), ("pitch", ctypes.c_int64)]


class Resources(ctypes.Structure):
    _fields_ = [("num_regs", ctypes.c_int), ("local_bytes", ctypes.c_int),
                ("shared_bytes", ctypes.c_int), ("max_threads", ctypes.c_int),
                ("const_bytes", ctypes.c_int)]


_lib = None
_lock = threading.Lock()


def _declare(lib):
    vp, cp, i64p = ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)
    sig = {
        "st_last_error": ([], cp),
        "st_init": ([ctypes.c_int], ctypes.c_int),
        "st_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "st_device_info_get": ([ctypes.c_int, ctypes.POINTER(DeviceInfo)], ctypes.c_int),
        "st_mem_info": ([ctypes.POINTER(ctypes.c_size_t)] * 2, ctypes.c_int),
        "st_compile_only": ([cp, cp, ctypes.POINTER(cp), ctypes.c_int, ctypes.POINTER(vp),
                             ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "st_free": ([vp], None),
        "st_compile": ([cp, cp, ctypes.POINTER(cp), ctypes.c_int, cp, ctypes.POINTER(vp)], ctypes.c_int),
        "st_load_code": ([vp, ctypes.c_size_t, ctypes.POINTER(vp)], ctypes.c_int),
        "st_unload": ([vp], ctypes.c_int),
        "st_get_function": ([vp, cp, ctypes.POINTER(vp)], ctypes.c_int),
        "st_launch": ([vp] + [ctypes.c_uint] * 7 + [vp, vp, ctypes.c_size_t], ctypes.c_int),
        "st_function_resources": ([vp, ctypes.POINTER(Resources)], ctypes.c_int),
        "st_fill_halo": ([vp, ctypes.c_int, ctypes.c_int, i64p, i64p, ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_double), vp], ctypes.c_int),
        "st_fill_halo_batched": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, vp], ctypes.c_int),
        "st_copy": ([vp, vp, ctypes.c_size_t, vp], ctypes.c_int),
        "st_copy_wide": ([vp, vp, ctypes.c_size_t, ctypes.c_int, vp], ctypes.c_int),
        "st_diff_norms": ([vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, i64p, i64p, vp,
                           ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "st_residual_norms": ([vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, i64p, i64p, vp,
                               ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "st_residual_norms_static": ([vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, i64p, i64p, vp,
                                      ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "st_field_digest": ([vp, ctypes.c_int, ctypes.c_int, i64p, i64p, vp, ctypes.POINTER(ctypes.c_uint64)],
                            ctypes.c_int),
        "st_bricks": ([vp, vp, ctypes.c_int, ctypes.c_int, i64p, ctypes.c_int, ctypes.c_int, vp],
                      ctypes.c_int),
        "st_event_create": ([ctypes.POINTER(vp)], ctypes.c_int),
        "st_event_record": ([vp, vp], ctypes.c_int),
        "st_event_elapsed_ms": ([vp, vp, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "st_event_sync": ([vp], ctypes.c_int),
        "st_event_query": ([vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "st_event_destroy": ([vp], ctypes.c_int),
        "st_stream_sync": ([vp], ctypes.c_int),
        "st_device_sync": ([], ctypes.c_int),
        "st_graph_begin": ([vp], ctypes.c_int),
        "st_graph_end": ([vp, ctypes.POINTER(vp)], ctypes.c_int),
        "st_graph_launch": ([vp, vp], ctypes.c_int),
        "st_graph_destroy": ([vp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def lib():
    """The loaded runtime library (raises if it was not built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeLibraryError(
                        f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                        "(hipcc --offload-arch=gfx950); there is no CPU fallback")
                l = ctypes.CDLL(LIB_PATH)
                _declare(l)
                _lib = l
    return _lib


def available() -> bool:
    return os.path.exists(LIB_PATH)


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().st_last_error().decode(errors="replace")
        raise RuntimeLibraryError(f"{what}: error {rc}: {msg}")


def _strs(items: Sequence[str]):
    arr = (ctypes.c_char_p * max(1, len(items)))()
    for k, s in enumerate(items):
        arr[k] = s.encode()
    return arr


def _i64(vals: Sequence[int]):
    return (ctypes.c_int64 * max(1, len(vals)))(*[int(v) for v in vals])


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------

def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().st_device_count(ctypes.byref(n)), "st_device_count")
    return n.value


def device_info(dev: int = 0) -> DeviceInfo:
    info = DeviceInfo()
    check(lib().st_device_info_get(dev, ctypes.byref(info)), "st_device_info_get")
    return info


def mem_info() -> tuple:
    f, t = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(lib().st_mem_info(ctypes.byref(f), ctypes.byref(t)), "st_mem_info")
    return f.value, t.value


def device_arch(dev: int = 0) -> str:
    """Offload architecture of device ``dev`` without target features."""
    return device_info(dev).arch.decode().split(":")[0]


# ---------------------------------------------------------------------------
# compilation / launch
# ---------------------------------------------------------------------------

def compile_only(src: str, opts: Sequence[str], name: str = "stencil.hip") -> bytes:
    code, size = ctypes.c_void_p(), ctypes.c_size_t(0)
    rc = lib().st_compile_only(src.encode(), name.encode(), _strs(opts), len(opts),
                               ctypes.byref(code), ctypes.byref(size))
    check(rc, "hiprtc compile")
    try:
        return ctypes.string_at(code, size.value)
    finally:
        lib().st_free(code)


def compile_module(src: str, opts: Sequence[str], cache_dir: Optional[str], name: str = "stencil.hip"):
    mod = ctypes.c_void_p()
    rc = lib().st_compile(src.encode(), name.encode(), _strs(opts), len(opts),
                          cache_dir.encode() if cache_dir else None, ctypes.byref(mod))
    check(rc, "st_compile")
    return mod.value


def load_code(code: bytes):
    """Load a code object (e.g. one written ahead of time) as a module on the
    current device."""
    mod = ctypes.c_void_p()
    buf = ctypes.create_string_buffer(code, len(code))
    check(lib().st_load_code(buf, len(code), ctypes.byref(mod)), "st_load_code")
    return mod.value


def cache_key(src: str, opts: Sequence[str]) -> str:
    """File stem under which ``st_compile`` caches the code object of ``src``
    built with ``opts`` (FNV-1a over the source and the options other than
    include paths, each followed by 0x1f)."""
    h = 0xcbf29ce484222325
    mask = (1 << 64) - 1

    def feed(data: bytes, h: int) -> int:
        for c in data:
            h = ((h ^ c) * 0x100000001b3) & mask
        return h

    h = feed(src.encode(), h)
    for o in opts:
        if o.startswith("-I"):
            continue
        h = feed(o.encode(), h)
        h = feed(b"\x1f", h)
    return f"{h:016x}"


def get_function(module, kernel: str):
    fn = ctypes.c_void_p()
    check(lib().st_get_function(module, kernel.encode(), ctypes.byref(fn)), f"st_get_function({kernel})")
    return fn.value


def launch(function, grid, block, args_blob, stream=None, dyn_lds: int = 0) -> None:
    g = list(grid) + [1] * (3 - len(grid))
    b = list(block) + [1] * (3 - len(block))
    rc = lib().st_launch(function, g[0], g[1], g[2], b[0], b[1], b[2], dyn_lds,
                         stream, ctypes.addressof(args_blob), ctypes.sizeof(args_blob))
    check(rc, "st_launch")


def function_resources(function) -> Resources:
    r = Resources()
    check(lib().st_function_resources(function, ctypes.byref(r)), "st_function_resources")
    return r


# ---------------------------------------------------------------------------
# fixed kernels
# ---------------------------------------------------------------------------

def fill_halo(ptr: int, dtype_code: int, shape, halo, kinds, values, stream=None) -> None:
    nd = len(shape)
    k = (ctypes.c_int * (2 * nd))(*kinds)
    v = (ctypes.c_double * (2 * nd))(*values)
    check(lib().st_fill_halo(ptr, dtype_code, nd, _i64(shape), _i64(halo), k, v, stream), "st_fill_halo")


def fill_halo_batched(descs_dev_ptr: int, count: int, ndim: int, max_slab: int, stream=None) -> None:
    check(lib().st_fill_halo_batched(descs_dev_ptr, count, ndim, max_slab, stream), "st_fill_halo_batched")


def copy(dst: int, src: int, nbytes: int, stream=None) -> None:
    check(lib().st_copy(dst, src, nbytes, stream), "st_copy")


def copy_wide(dst: int, src: int, nbytes: int, vectors: int, stream=None) -> None:
    check(lib().st_copy_wide(dst, src, nbytes, vectors, stream), "st_copy_wide")


def diff_norms(a: int, da: int, b: int, db: int, shape, halo, stream=None) -> tuple:
    out = (ctypes.c_double * 2)()
    check(lib().st_diff_norms(a, da, b, db, len(shape), _i64(shape), _i64(halo), stream, out),
          "st_diff_norms")
    return out[0], out[1]


def residual_norms(a: int, da: int, b: int, db: int, shape, halo, stream=None, no_alloc: bool = False) -> tuple:
    """(max|a-b|, max|b|, ||a-b||_2, ||b||_2) over the interior, in FP64
    (``st_residual_norms``; ``no_alloc``: ``st_residual_norms_static``)."""
    out = (ctypes.c_double * 4)()
    fn = lib().st_residual_norms_static if no_alloc else lib().st_residual_norms
    check(fn(a, da, b, db, len(shape), _i64(shape), _i64(halo), stream, out),
          "st_residual_norms_static" if no_alloc else "st_residual_norms")
    return tuple(out)


def field_digest(a: int, dtype_code: int, shape, halo, stream=None) -> int:
    """Checksum of the interior's stored bits (``st_field_digest``)."""
    out = ctypes.c_uint64(0)
    check(lib().st_field_digest(a, dtype_code, len(shape), _i64(shape), _i64(halo), stream, ctypes.byref(out)),
          "st_field_digest")
    return int(out.value)


def bricks(flat: int, bricked: int, dtype_code: int, shape, bs: int, to_bricks: bool, stream=None) -> None:
    check(lib().st_bricks(flat, bricked, dtype_code, len(shape), _i64(shape), bs, int(to_bricks), stream),
          "st_bricks")


# ---------------------------------------------------------------------------
# events / sync / graphs
# ---------------------------------------------------------------------------

class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().st_event_create(ctypes.byref(h)), "st_event_create")
        self.h = h.value

    def record(self, stream=None):
        check(lib().st_event_record(self.h, stream), "st_event_record")

    def synchronize(self) -> None:
        check(lib().st_event_sync(self.h), "st_event_sync")

    def query(self) -> bool:
        """True once the work recorded before the event has completed
        (does not wait)."""
        done = ctypes.c_int(0)
        check(lib().st_event_query(self.h, ctypes.byref(done)), "st_event_query")
        return bool(done.value)

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        check(lib().st_event_elapsed_ms(self.h, end.h, ctypes.byref(ms)), "st_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            if self.h and _lib is not None:
                _lib.st_event_destroy(self.h)
        except Exception:
            pass


def stream_sync(stream=None) -> None:
    check(lib().st_stream_sync(stream), "st_stream_sync")


def device_sync() -> None:
    check(lib().st_device_sync(), "st_device_sync")


class Graph:
    """hipGraph captured from everything launched on ``stream`` in the block."""

    def __init__(self, stream):
        self.stream = stream
        self.h = None

    def __enter__(self):
        check(lib().st_graph_begin(self.stream), "st_graph_begin")
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_void_p()
        rc = lib().st_graph_end(self.stream, ctypes.byref(h))
        if exc_type is None:
            check(rc, "st_graph_end")
            self.h = h.value
        return False

    def launch(self, stream=None):
        check(lib().st_graph_launch(self.h, self.stream if stream is None else stream), "st_graph_launch")

    def __del__(self):
        try:
            if self.h and _lib is not None:
                _lib.st_graph_destroy(self.h)
        except Exception:
            pass
