"""GPU presets and hardware specs (upstream ``cutile/config.py:17-238``),
with the AMD Instinct parts this backend targets and ROCm name detection."""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple


@dataclass(frozen=True)
class GPUPreset:
    name: str
    peak_bandwidth_gbs: float
    shared_mem_bytes: int      # LDS per CU
    sm_count: int              # compute units
    peak_gflops_fp64: float    # dense vector FP64
    peak_gflops_fp32: float    # dense vector FP32
    l2_bytes: int = 0          # per XCD
    infinity_cache_bytes: int = 0
    xcds: int = 1
    arch: str = ""

    def peak_gflops(self, dtype: str = "float64") -> float:
        return self.peak_gflops_fp32 if dtype in ("float32", "fp32", "bfloat16") else self.peak_gflops_fp64


# AMD Instinct MI355X (CDNA4, gfx950): 8 TB/s HBM3E, 256 CUs in 8 XCDs,
# 160 KB LDS per CU, 4 MB L2 per XCD, 256 MB Infinity Cache.
MI355X = GPUPreset("MI355X", 8000.0, 160 * 1024, 256, 78600.0, 157300.0,
                   4 * 1024 * 1024, 256 * 1024 * 1024, 8, "gfx950")
MI300X = GPUPreset("MI300X", 5300.0, 64 * 1024, 304, 81700.0, 163400.0,
                   4 * 1024 * 1024, 256 * 1024 * 1024, 8, "gfx942")
H100_SXM = GPUPreset("H100_SXM", 3350.0, 232 * 1024, 132, 33500.0, 67000.0)
A100_80GB = GPUPreset("A100_80GB", 2039.0, 164 * 1024, 108, 9746.0, 19492.0)
RTX_PRO_6000 = GPUPreset("RTX_PRO_6000", 1792.0, 228 * 1024, 188, 1953.0, 125000.0)

GPU_PRESETS: Dict[str, GPUPreset] = {p.name: p for p in (MI355X, MI300X, H100_SXM, A100_80GB, RTX_PRO_6000)}


@dataclass
class HardwareSpec:
    shared_mem_bytes: int = 160 * 1024
    dtype_bytes: int = 8
    peak_bandwidth_gbs: float = 8000.0
    peak_gflops: float = 78600.0
    sm_count: int = 256

    @classmethod
    def from_preset(cls, name: str, dtype: str = "float64") -> "HardwareSpec":
        p = GPU_PRESETS[name]
        return cls(p.shared_mem_bytes, _bytes(dtype), p.peak_bandwidth_gbs, p.peak_gflops(dtype), p.sm_count)

    @classmethod
    def auto_detect(cls, dtype: str = "float64") -> "HardwareSpec":
        p = detect_preset() or MI355X
        return cls(p.shared_mem_bytes, _bytes(dtype), p.peak_bandwidth_gbs, p.peak_gflops(dtype), p.sm_count)


def _bytes(dtype: str) -> int:
    return {"float32": 4, "fp32": 4, "bfloat16": 2}.get(dtype, 8)


def detect_preset() -> Optional[GPUPreset]:
    """Match the ROCm device (gcnArchName / marketing name) to a preset."""
    try:
        from cutile.runtime import rt

        if not rt.available() or rt.device_count() == 0:
            return None
        info = rt.device_info(0)
        arch = info.arch.decode().split(":")[0]
        name = info.name.decode().upper()
    except Exception:
        return None
    if "MI355" in name or arch == "gfx950":
        return MI355X
    if "MI300" in name or arch == "gfx942":
        return MI300X
    return None


@dataclass
class TilingConfig:
    candidate_sizes: List[int] = field(default_factory=lambda: [64, 128, 256, 512, 1024])
    max_temporal_steps: int = 8


@dataclass
class BenchmarkConfig:
    grid_sizes: Dict[int, Tuple[int, ...]] = field(default_factory=lambda: {
        1: (1 << 22,), 2: (4096, 4096), 3: (512, 512, 512)})
    warmup_iters: int = 5
    benchmark_iters: int = 20


DEFAULT_HW = HardwareSpec()
DEFAULT_TILING = TilingConfig()
DEFAULT_BENCHMARK = BenchmarkConfig()
