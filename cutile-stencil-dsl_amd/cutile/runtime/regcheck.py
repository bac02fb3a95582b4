"""Spill-aware selection between kernel configurations, on AMD code objects.

Upstream (``cutile/runtime/regcheck.py:136-306``) captures the cubin cuTile
builds and reads it with ``cuobjdump -res-usage``.  Here the probe compiles
the emitted HIP source with hiprtc (no GPU needed) and reads the kernel
descriptor metadata of our own code object (``.vgpr_count``,
``.agpr_count``, ``.vgpr_spill_count``, ``.private_segment_fixed_size``,
``.group_segment_fixed_size``) with ``llvm-readelf --notes``; with a GPU and
``time_shape`` it also times the kernel.  The selection rules
(``select_config``, ``select_fastest``) are the upstream ones; the register
budget is the gfx950 one: 512 VGPR+AGPR per lane per SIMD shared by the
resident waves, one wave per SIMD per 256-thread workgroup.
"""

from __future__ import annotations

import os
import re
import shutil
import subprocess
import tempfile
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence


@dataclass(frozen=True)
class Resources:
    registers: int
    stack_bytes: int
    ms: Optional[float] = None
    lds_bytes: int = 0

    @property
    def spills(self) -> bool:
        return self.stack_bytes > 0


@dataclass(frozen=True)
class Config:
    tile_sizes: tuple
    kernel_hints: dict = field(default_factory=dict)
    max_registers: Optional[int] = None


def readelf_path() -> Optional[str]:
    for p in ("/opt/rocm/llvm/bin/llvm-readelf", shutil.which("llvm-readelf") or ""):
        if p and os.path.isfile(p):
            return p
    return None


def parse_notes(text: str) -> dict:
    """Kernel name -> Resources from ``llvm-readelf --notes`` of a code object."""
    out = {}
    for block in re.split(r"\n\s*- \.", text):
        m = re.search(r"\.name:\s+(\S+)", block)
        if not m:
            continue

        def num(key):
            mm = re.search(rf"\.?{key}:\s+(\d+)", block)
            return int(mm.group(1)) if mm else 0

        vgpr = num("vgpr_count") + num("agpr_count")
        # VGPR spills go to scratch memory; SGPR spills go to VGPR lanes
        # (v_writelane/v_readlane) and use no memory, so they do not count
        spill = num("vgpr_spill_count")
        scratch = num("private_segment_fixed_size")
        out[m.group(1)] = Resources(vgpr, max(scratch, 4 * spill), None, num("group_segment_fixed_size"))
    return out


def code_object_resources(code: bytes) -> dict:
    tool = readelf_path()
    if tool is None:
        return {}
    with tempfile.NamedTemporaryFile(suffix=".co", delete=False) as f:
        f.write(code)
        path = f.name
    try:
        r = subprocess.run([tool, "--notes", path], capture_output=True, text=True, timeout=60)
        return parse_notes(r.stdout)
    finally:
        os.unlink(path)


def timing_shape(ndim: int) -> tuple:
    return {1: (1 << 22,), 2: (1024, 2048), 3: (128, 256, 256)}[ndim]


def _source_of(code: str) -> str:
    return code.split("_HIP_SOURCE = r'''\n", 1)[1].split("'''", 1)[0]


def probe_resources(*, code: str, kernel_name: str, ndim: int, num_inputs: int, num_outputs: int = 1,
                    has_consts: bool = False, tile_sizes: Sequence[int], halo_widths: Sequence[int],
                    dtype: str = "float64", time_shape: Optional[Sequence[int]] = None,
                    iterations: int = 5) -> Optional[Resources]:
    """Resources of the vector variant of the kernel in ``code``; ``None``
    when the probe cannot run.  Never raises."""
    try:
        from cutile.runtime import rt
        from cutile.runtime.kernel import compile_options

        if not rt.available():
            return None
        src = _source_of(code)
        blob = rt.compile_only(src, compile_options("gfx950"))
        res = code_object_resources(blob)
        vec = [v for k, v in res.items() if "_sweep_" in k]
        if not vec:
            return None
        base = vec[0]
        ms = None
        if time_shape is not None:
            ms = _time(code, kernel_name, ndim, num_inputs, num_outputs, halo_widths, dtype,
                       time_shape, iterations)
        return Resources(base.registers, base.stack_bytes, ms, base.lds_bytes)
    except Exception:
        return None


def _time(code, kernel_name, ndim, num_inputs, num_outputs, halo_widths, dtype, time_shape, iterations):
    import importlib.util

    import torch

    from cutile.runtime import rt
    from cutile.runtime.kernel import default_stream

    if not torch.cuda.is_available() or rt.device_count() == 0:
        return None
    with tempfile.NamedTemporaryFile("w", suffix=".py", delete=False) as f:
        f.write(code)
        path = f.name
    try:
        spec = importlib.util.spec_from_file_location(f"_probe_{kernel_name}", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
    finally:
        os.unlink(path)
    launcher = getattr(mod, "launch_" + kernel_name.removesuffix("_kernel"))
    shape = tuple(int(n) + 2 * h for n, h in zip(time_shape, halo_widths))
    dt = getattr(torch, dtype)
    arrays = [(torch.rand(shape, device="cuda", dtype=torch.float64) + 0.5).to(dt) for _ in range(num_inputs)]
    outs = [torch.zeros(shape, device="cuda", dtype=dt) for _ in range(num_outputs)]
    for _ in range(2):
        launcher(*arrays, *outs)
    torch.cuda.synchronize()
    s = default_stream()
    e0, e1 = rt.Event(), rt.Event()
    e0.record(s)
    for _ in range(iterations):
        launcher(*arrays, *outs)
    e1.record(s)
    return e0.elapsed_ms(e1) / iterations


def register_budget(occupancy: Optional[int], threads_per_block: int = 256,
                    registers_per_lane: int = 512) -> int:
    """VGPRs per lane that keep ``occupancy`` workgroups of ``threads_per_block``
    threads resident per CU (each wave64 of a 256-thread group lands on its
    own SIMD, so ``occupancy`` waves share one SIMD's 512 registers/lane).
    Without an occupancy target the budget is the 256 architectural VGPRs:
    kernels that need accumulation registers as spill space are rejected."""
    if not occupancy:
        return 256
    waves_per_simd = max(1, occupancy * max(1, threads_per_block // 256))
    return min(256, registers_per_lane // waves_per_simd)


def select_config(configs: Sequence[Config], probe: Callable, max_registers: Optional[int] = None) -> Config:
    probed = []
    for cfg in configs:
        res = probe(cfg)
        if res is None:
            return cfg
        budget = cfg.max_registers if cfg.max_registers is not None else max_registers
        if not res.spills and (budget is None or res.registers <= budget):
            return cfg
        probed.append((cfg, res))
    for cfg, res in probed:
        if not res.spills:
            return cfg
    return configs[-1]


def _fits(cfg, res, max_registers):
    budget = cfg.max_registers if cfg.max_registers is not None else max_registers
    return not res.spills and (budget is None or res.registers <= budget)


def select_fastest(configs: Sequence[Config], probe: Callable, max_spill_bytes: int = 512,
                   margin: float = 0.03, max_registers: Optional[int] = None) -> Config:
    probed = []
    for k, cfg in enumerate(configs):
        res = probe(cfg)
        if res is None:
            if k == 0:
                return cfg
            continue
        probed.append((cfg, res))
    if not probed:
        return configs[0]
    timed = [(c, r) for c, r in probed if r.ms is not None]
    if timed:
        trusted = [(c, r) for c, r in timed if r.stack_bytes <= max_spill_bytes] or timed
        best = min(r.ms for _, r in trusted)
        for c, r in trusted:
            if r.ms <= best * (1.0 + margin):
                return c
    for c, r in probed:
        if _fits(c, r, max_registers):
            return c
    for c, r in probed:
        if not r.spills:
            return c
    return configs[-1]
