"""HIP emitter: stencil metadata -> HIP C++ for gfx950 (replaces the cuTile
emitter ``cutile/lowering/target_to_python.py:34-526``).

The emitted translation unit includes the hand-written device templates in
``csrc/stencil_core.h`` and defines, per variant, a traits struct ``P_<v>``
(fields, access tables, tile shape, storage/compute types and the per-cell
expression) plus an ``extern "C"`` kernel that instantiates the sweep
template.  Variants differ in the contiguous-axis vector width (16-byte
vectors when the array rows allow it, scalars otherwise) and, for temporal
blocking, in the depth.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

from cutile.lowering.meta import StencilMeta, access_table

STORAGE_CTYPES = {"float64": "double", "float32": "float", "bfloat16": "st::bf16_t"}
COMPUTE_CTYPES = {"float64": "double", "float32": "float"}
ELEM_BYTES = {"float64": 8, "float32": 4, "bfloat16": 2}


@dataclass(frozen=True)
class KernelConfig:
    """Launch geometry of one sweep variant.

    ``ty``/``tz`` are the in-plane tile (rows x contiguous columns) of a
    workgroup, ``vz`` the contiguous-axis vector width, ``nt`` the threads per
    workgroup (a multiple of 64)."""

    ty: int
    tz: int
    vz: int
    nt: int

    @property
    def tag(self) -> str:
        return f"t{self.ty}x{self.tz}v{self.vz}n{self.nt}"


def plan_accesses(meta: StencilMeta) -> dict:
    """Split accesses into register-queue columns and LDS plane buffers."""
    tab = access_table(meta)
    nf = meta.num_inputs
    col = [t for t in tab if t[2] == 0 and t[3] == 0]
    qlo = min((t[1] for t in col), default=1)
    qhi = max((t[1] for t in col), default=0)
    qf = [0] * nf
    for f, dx, dy, dz in col:
        qf[f] = 1
    bufs: dict = {}
    for f, dx, dy, dz in tab:
        if dy == 0 and dz == 0:
            continue
        ry, rz = bufs.get((f, dx), (0, 0))
        bufs[(f, dx)] = (max(ry, abs(dy)), max(rz, abs(dz)))
    blist = sorted(bufs.items())
    return {"qlo": qlo, "qhi": qhi, "qf": qf, "bufs": [(f, dx, ry, rz) for (f, dx), (ry, rz) in blist]}


def lds_pitch(tz: int, vz: int, rzv: int, elem_bytes: int) -> int:
    """LDS row pitch (elements) of a plane buffer, as ``st::Tab::pitch``:
    the row (tile plus halo columns) is padded to a pitch congruent to the
    tile width modulo 256 bytes when a row of lanes (tz / vz of them) does
    not end on a 32-lane boundary."""
    p = tz + 2 * rzv
    if (tz // vz) % 32 == 0:
        return p
    while ((p - tz) * elem_bytes) % 256:
        p += 1
    return p


def lds_bytes(meta: StencilMeta, cfg: KernelConfig, compute: str) -> int:
    plan = plan_accesses(meta)
    cb = ELEM_BYTES[compute]
    tot = 0
    for f, dx, ry, rz in plan["bufs"]:
        rzv = -(-rz // cfg.vz) * cfg.vz
        tot += (cfg.ty + 2 * ry) * lds_pitch(cfg.tz, cfg.vz, rzv, cb) * cb
    return tot


def _arr(vals: Sequence[int]) -> str:
    vals = list(vals) or [0]
    return "{" + ", ".join(str(int(v)) for v in vals) + "}"


def radius3(meta: StencilMeta) -> tuple:
    """Stencil radius along each axis of the 3-D (stream, row, column) map."""
    tab = access_table(meta)
    return tuple(max([abs(t[k]) for t in tab] + [0]) for k in (1, 2, 3))


def temporal_lds_bytes(ty: int, tz: int, T: int, r3: tuple, compute: str) -> int:
    rx, ry, rz = r3
    nr = 2 * rx + 1
    return sum(nr * (ty + 2 * (T - l) * ry) * (tz + 2 * (T - l) * rz) for l in range(1, T)) * ELEM_BYTES[compute]


def temporal_tile(ndim: int, T: int, r3: tuple, compute: str, budget: int = 160 * 1024) -> tuple:
    """Largest in-plane tile whose level rings fit the LDS budget."""
    if ndim == 3:
        cands = [(8, 128), (8, 64), (4, 64), (4, 32), (2, 32), (4, 16), (2, 16), (1, 16), (1, 8)]
    else:
        cands = [(1, 1024), (1, 512), (1, 256), (1, 128), (1, 64), (1, 32), (1, 16)]
    for ty, tz in cands:
        if temporal_lds_bytes(ty, tz, T, r3, compute) <= budget:
            return ty, tz
    raise ValueError(f"temporal depth {T} does not fit the LDS budget for radius {r3}")


@dataclass(frozen=True)
class TBlockConfig:
    """Geometry of one register/LDS temporal-blocking variant.

    The workgroup evaluates ``T`` levels on a ``ty`` x ``tz`` region: its
    output tile widened by ``gy`` rows and ``gz`` columns on each side.
    ``nt`` threads own the region's ``ty * tz / vz`` vector cells flat,
    ``kc`` each (the last ones may be idle)."""

    ty: int
    tz: int
    vz: int
    nt: int
    gy: int
    gz: int
    T: int
    # > 0: the two-level kernel with wave roles (``st::tbws_tile``): ``na``
    # threads evaluate level 1 on the region, the other ``nt - na`` level 2
    # on the output tile
    na: int = 0
    # two-level kernel: a thread's ``kc`` (level 1) or ``kb`` (level 2) cells
    # lie on adjacent rows of one column instead of ``na`` (``nt - na``)
    # cells apart, so the in-plane neighbours they share are read once
    ya: int = 0

    @property
    def oty(self) -> int:
        return self.ty - 2 * self.gy

    @property
    def otz(self) -> int:
        return self.tz - 2 * self.gz

    @property
    def kc(self) -> int:
        """Region vector cells per (level-1) thread."""
        return -(-(self.ty * (self.tz // self.vz)) // (self.na or self.nt))

    @property
    def kb(self) -> int:
        """Output vector cells per level-2 thread of the two-level kernel."""
        return -(-(self.oty * (self.otz // self.vz)) // max(1, self.nt - self.na)) if self.na else 0

    @property
    def tag(self) -> str:
        if self.na:
            return f"tw{self.T}r{self.ty}x{self.tz}v{self.vz}n{self.nt}a{self.na}" + ("y" if self.ya else "")
        return f"tb{self.T}r{self.ty}x{self.tz}v{self.vz}n{self.nt}"

    def kernel_config(self) -> KernelConfig:
        return KernelConfig(self.ty, self.tz, self.vz, self.nt)


def tblock_queue_range(meta: StencilMeta) -> tuple:
    """Plane offsets every level's register queue holds: all accessed dx and
    0 (levels >= 1 fill their LDS planes from the queue)."""
    plan = plan_accesses(meta)
    dxs = [0] + [b[1] for b in plan["bufs"]]
    if plan["qhi"] >= plan["qlo"]:
        dxs += [plan["qlo"], plan["qhi"]]
    return min(dxs), max(dxs)


def _vgpr_limit(nt: int) -> int:
    """Architectural VGPRs a wave may use so that a workgroup of ``nt``
    threads fits one CU (4 SIMDs, 512 VGPRs per lane each)."""
    waves_per_simd = -(-nt // 256)
    return min(256, 512 // waves_per_simd)


def tblock_estimate(meta: StencilMeta, cfg: TBlockConfig, storage: str, compute: str) -> dict:
    """LDS bytes and a VGPR estimate (queues, prefetch, ring, offsets) of a
    temporal-blocking variant."""
    plan = plan_accesses(meta)
    qlo, qhi = tblock_queue_range(meta)
    qn = qhi - qlo + 1
    cw = ELEM_BYTES[compute] // 4
    lds = cfg.T * lds_bytes(meta, cfg.kernel_config(), compute)
    nvb = 1
    for f, dx, ry, rz in plan["bufs"]:
        rzv = -(-rz // cfg.vz) * cfg.vz
        w = (cfg.tz + 2 * rzv) // cfg.vz
        ring = 2 * ry * w + cfg.ty * 2 * (rzv // cfg.vz)
        nvb = max(nvb, -(-ring // cfg.nt))
    nb = max(1, len(plan["bufs"]))
    # queues of T levels + prefetched plane and ring + per-cell offsets and
    # flags + the expression's working set (its loaded operands)
    work = cw * (min(len(access_table(meta)), 40) + 8) + 24
    nf = max(1, meta.num_inputs)
    regs = cw * cfg.vz * cfg.kc * nf * (cfg.T * qn + 1) + cw * cfg.vz * nb * nvb + 16 * cfg.kc + nb * nvb + work
    return {"lds": lds, "vgpr": regs, "qn": qn}


def tblock_config(meta: StencilMeta, T: int, storage: str, compute: str, vz: int,
                  budget: int = 160 * 1024, max_otz: int = 256, force: "tuple | None" = None) -> "TBlockConfig | None":
    """Cheapest register/LDS temporal-blocking geometry for ``T`` levels at
    vector width ``vz``: fewest lane-cells evaluated per output cell (ghost
    ring and idle lanes included) whose queues fit the VGPR budget of its
    workgroup size and whose level planes fit the LDS.  None when nothing
    fits (deep ghosts of wide stencils); the caller falls back to the
    LDS-ring kernel.  ``force`` = (output rows, output columns, threads)
    restricts the search to that shape (a tuning hint; it must satisfy the
    same limits)."""
    # one field, or coupled fields whose result m is the next state of field m
    if meta.num_exprs != meta.num_inputs or meta.update or T < 2:
        return None
    rx, ry, rz = radius3(meta)
    nd = meta.ndim
    gy = (T - 1) * ry
    gz = -(-((T - 1) * rz) // vz) * vz
    best = None
    otzs = [t for t in (16, 32, 64, 128, 256, 512, 1024) if t <= max_otz and t % vz == 0]
    otys = range(1, 65) if nd == 3 else (1,)
    nts = range(64, 769, 64)     # 16-wave workgroups measured slower (barriers across 16 waves)
    if force is not None:
        otys, otzs, nts = (int(force[0]),), (int(force[1]),), (int(force[2]),)
        if otzs[0] % vz or (nd < 3 and otys[0] != 1):
            return None
    for otz in otzs:
        tz = otz + 2 * gz
        for oty in otys:
            ty = oty + 2 * gy
            cells = ty * (tz // vz)
            for nt in nts:
                kc = -(-cells // nt)
                if kc > 4:
                    continue
                if kc > 1 and (kc - 1) * nt >= cells:
                    continue
                cfg = TBlockConfig(ty, tz, vz, nt, gy, gz, T)
                est = tblock_estimate(meta, cfg, storage, compute)
                # a forced shape skips the register estimate (the code
                # object's actual registers are checked before any launch)
                if est["lds"] > budget or (force is None and est["vgpr"] > 0.85 * _vgpr_limit(nt)):
                    continue
                # lane-cells evaluated per output cell (ghost ring and idle
                # lanes), with the penalties measured on MI355X for narrow
                # output columns (halo rows per LDS row, short contiguous
                # runs) and for several cells per thread (register pressure)
                cost = kc * nt * vz / (oty * otz) * (1 + 16 / otz) * (1 + 0.05 * (kc - 1))
                key = (round(cost, 3), -(oty * otz), nt)
                if best is None or key < best[0]:
                    best = (key, cfg)
    # more than ~6 lane-cells per output cell loses to T separate sweeps
    if best is None or best[0][0] > 6.0:
        return None
    return best[1]


def tbws_eligible(meta: StencilMeta, T: int) -> bool:
    """The two-level kernel with wave roles applies: two levels of a 3-D
    closed system (result m is the next state of field m) whose in-plane
    accesses all lie in the evaluated plane and whose column accesses reach
    both neighbouring planes."""
    if T != 2 or meta.ndim != 3 or meta.num_exprs != meta.num_inputs or meta.update:
        return False
    plan = plan_accesses(meta)
    qlo, qhi = tblock_queue_range(meta)
    return qlo <= -1 and qhi >= 1 and all(b[1] == 0 for b in plan["bufs"])


def ring_pitch(cfg: TBlockConfig, elem_bytes: int, nr: int, nf: int) -> int:
    """Elements per field and plane in a row of the two-level kernel's
    level-1 ring (``nr`` planes of ``nf`` fields per row), as
    ``st::ring_pitch``: the row span is congruent to the output width modulo
    256 bytes (the level-2 threads' 16-byte reads free of bank conflicts),
    unless an output row fills whole half-waves."""
    if (cfg.otz // cfg.vz) % 32 == 0:
        return cfg.tz
    p = cfg.tz
    while ((nr * nf * p - cfg.otz) * elem_bytes) % 256:
        p += 1
    return p


def tbws_estimate(meta: StencilMeta, cfg: TBlockConfig, storage: str, compute: str) -> dict:
    """LDS bytes (two level-0 planes and the level-1 ring) and a VGPR
    estimate of the busier role of a two-level variant."""
    plan = plan_accesses(meta)
    qlo, qhi = tblock_queue_range(meta)
    qn = qhi - qlo + 1
    nr = qhi + 2
    cb = ELEM_BYTES[compute]
    cw = cb // 4
    nf = max(1, meta.num_inputs)
    lds = 2 * lds_bytes(meta, cfg.kernel_config(), compute) + nr * nf * cfg.ty * ring_pitch(cfg, cb, nr, nf) * cb
    nbt = cfg.nt - cfg.na
    nvb = 1
    for f, dx, ry, rz in plan["bufs"]:
        rzv = -(-rz // cfg.vz) * cfg.vz
        w = (cfg.tz + 2 * rzv) // cfg.vz
        ring = 2 * ry * w + cfg.ty * 2 * (rzv // cfg.vz)
        nvb = max(nvb, -(-ring // nbt))
    nb = max(1, len(plan["bufs"]))
    # queues (+ two prefetched heads, or two prefetched rings), offsets, and
    # the expression's loaded operands (about six vectors in flight; gfx950
    # code objects of the 25-point star: 102-104 VGPRs at one cell each)
    work = 6 * cw * cfg.vz + 24
    reg_a = cw * cfg.vz * cfg.kc * nf * (qn + 2) + 8 * cfg.kc + work
    reg_b = cw * cfg.vz * cfg.kb * nf * qn + 2 * cw * cfg.vz * nb * nvb + 4 * cfg.kb + 2 * nb * nvb + work
    return {"lds": lds, "vgpr": max(reg_a, reg_b), "qn": qn}


def tbws_config(meta: StencilMeta, T: int, storage: str, compute: str, vz: int,
                budget: int = 160 * 1024, force: "tuple | None" = None) -> "TBlockConfig | None":
    """Geometry of the two-level kernel with wave roles for vector width
    ``vz``: fewest lanes per output cell (both roles, idle lanes included)
    whose roles fit one workgroup (at most 1024 threads, whole waves, one
    vector cell per thread) with the queues in the VGPRs its waves may use
    and the planes in the LDS.  ``force`` = (output rows,
    output columns, level-1 threads, level-2 threads) restricts the search
    to that shape.  None when the kernel does not apply or nothing fits."""
    if not tbws_eligible(meta, T):
        return None
    rx, ry, rz = radius3(meta)
    gy = ry
    gz = -(-rz // vz) * vz
    best = None
    otys, otzs = range(4, 65, 4), [t for t in (16, 24, 32, 40, 48, 64, 96, 128) if t % vz == 0]
    if force is not None:
        otys, otzs = (int(force[0]),), (int(force[1]),)
        if otzs[0] % vz:
            return None
    for otz in otzs:
        tz = otz + 2 * gz
        for oty in otys:
            ty = oty + 2 * gy
            cells, ocells = ty * (tz // vz), oty * (otz // vz)
            # two cells per thread only when forced: on gfx950 their doubled
            # queues spill at the workgroup sizes that give the region one
            # thread per cell
            for ka in ((1, 2) if force is not None else (1,)):
                na = -(-cells // ka)
                na = -(-na // 64) * 64
                for kb in ((1, 2) if force is not None else (1,)):
                    nbt = -(-ocells // kb)
                    nbt = -(-nbt // 64) * 64
                    if force is not None:
                        na, nbt = int(force[2]), int(force[3])
                        if na % 64 or nbt % 64 or na <= 0 or nbt <= 0:
                            return None
                    nt = na + nbt
                    if nt > 1024:
                        continue
                    cfg = TBlockConfig(ty, tz, vz, nt, gy, gz, T, na)
                    if cfg.kc > 2 or cfg.kb > 2:
                        continue
                    est = tbws_estimate(meta, cfg, storage, compute)
                    if est["lds"] > budget or (force is None and est["vgpr"] > 0.9 * _vgpr_limit(nt)):
                        continue
                    # lanes of both roles per output cell (gfx950, 25-point star
                    # at 512^3 per depth-2 launch: FP64 24x32 tiles 771 us,
                    # 16x40 840, 16x32 856; FP32 24x64 391 us, 16x96 424)
                    cost = (na * cfg.kc + nbt * cfg.kb) * vz / (oty * otz)
                    # ties: the larger tile, then the wider one (longer
                    # contiguous rows, fewer halo columns per row)
                    key = (round(cost, 3), -(oty * otz), -otz, nt)
                    if best is None or key < best[0]:
                        best = (key, cfg)
    return None if best is None else best[1]


def tbws_adjacent(cfg: "TBlockConfig") -> "TBlockConfig | None":
    """``cfg`` with each thread's cells on adjacent rows (hint ``tbya``),
    or None where the row groups do not divide the region or the output
    tile, or a role has fewer threads than (row groups x vector columns)."""
    ka, kb = cfg.kc, cfg.kb
    nbt = cfg.nt - cfg.na
    if cfg.ty % ka or cfg.oty % kb:
        return None
    if cfg.na < (cfg.ty // ka) * (cfg.tz // cfg.vz) or nbt < (cfg.oty // kb) * (cfg.otz // cfg.vz):
        return None
    return TBlockConfig(cfg.ty, cfg.tz, cfg.vz, cfg.nt, cfg.gy, cfg.gz, cfg.T, cfg.na, ya=1)


def select_tblock(meta: StencilMeta, T: int, storage: str, compute: str, vz: int,
                  hints: "dict | None" = None) -> "TBlockConfig | None":
    """The register/LDS temporal-blocking variant the emitter builds for
    depth ``T`` at vector width ``vz``: for two levels of a 3-D stencil
    wider than one cell, the kernel with wave roles (hint ``tbws`` 0/1
    overrides the choice, ``tbws_tile`` forces its geometry; the ``zdpp``
    hint, which that kernel does not take, selects the other); otherwise, or
    when that kernel does not apply, the single-role kernel (``tblock_tile``
    forces its geometry).  None: neither fits."""
    hints = hints or {}
    ws = hints.get("tbws")
    ws = max(radius3(meta)[1:]) >= 2 if ws is None else bool(ws)
    if hints.get("zdpp"):
        ws = False        # cross-lane z-neighbours: an option of the single-role kernel only
    tb = None
    if ws and tbws_eligible(meta, T):
        tb = tbws_config(meta, T, storage, compute, vz, force=hints.get("tbws_tile"))
        if tb is not None and int(hints.get("tbya", 0)):
            tb = tbws_adjacent(tb) or tb       # a geometry without whole row groups keeps the flat order
    if tb is None:
        tb = tblock_config(meta, T, storage, compute, vz, force=hints.get("tblock_tile"))
    return tb


def emit_traits(meta: StencilMeta, cfg: KernelConfig, storage: str, compute: str, tag: str,
                nontemporal: bool = True, prefetch: int = 1, partial: bool = False,
                ntload: bool = False, tblock: "TBlockConfig | None" = None, brick: tuple = (0, 0, 0),
                zdpp: bool = False, tbskip: int = 0, tbdb: int = 0, tbpd: int = 2, ringpd: int = 1,
                sdb: int = 0, tbrp: int = 0) -> str:
    """Traits struct of one kernel variant.  ``partial``: rows may end in a
    partial vector (batched instances whose interior is not a multiple of
    the vector width).  ``tblock``: traits of a temporal-blocking variant
    (levels, ghost widths; the queue holds every accessed plane offset).
    ``brick``: (B0, B1, B2) brick of a brick-major variant in the 3-D map
    ((0, 0, 0): flat arrays)."""
    plan = plan_accesses(meta)
    bufs = plan["bufs"]
    r3 = radius3(meta)
    qlo, qhi = (plan["qlo"], plan["qhi"]) if tblock is None else tblock_queue_range(meta)
    qf = plan["qf"] if tblock is None else [1] * meta.num_inputs
    lines = [
        f"struct P_{tag} {{",
        f"  using S = {STORAGE_CTYPES[storage]};",
        f"  using C = {COMPUTE_CTYPES[compute]};",
        f"  static constexpr int NF = {meta.num_inputs}, NO = {meta.num_exprs};",
        f"  static constexpr int NT = {cfg.nt}, TY = {cfg.ty}, TZ = {cfg.tz}, VZ = {cfg.vz};",
        f"  static constexpr int QLO = {qlo}, QHI = {qhi};",
        f"  static constexpr int QF[{max(1, meta.num_inputs)}] = {_arr(qf)};",
        f"  static constexpr int NB = {len(bufs)};",
        f"  static constexpr int BF[{max(1, len(bufs))}] = {_arr([b[0] for b in bufs])};",
        f"  static constexpr int BDX[{max(1, len(bufs))}] = {_arr([b[1] for b in bufs])};",
        f"  static constexpr int BRY[{max(1, len(bufs))}] = {_arr([b[2] for b in bufs])};",
        f"  static constexpr int BRZ[{max(1, len(bufs))}] = {_arr([b[3] for b in bufs])};",
        f"  static constexpr int RX = {r3[0]}, RY = {r3[1]}, RZ = {r3[2]};",
        f"  static constexpr int NTS = {1 if nontemporal else 0}, PD = {int(prefetch)}, "
        f"PARTIAL = {1 if partial else 0}, NTL = {1 if ntload else 0}, RPD = {max(1, int(ringpd))}, "
        f"SDB = {1 if sdb else 0};",
        f"  static constexpr int BR0 = {int(brick[0])}, BR1 = {int(brick[1])}, BR2 = {int(brick[2])};",
        f"  static constexpr int ZDPP = {1 if zdpp and cfg.vz > 0 else 0};",
    ]
    if tblock is not None:
        lines.append(f"  static constexpr int TT = {tblock.T}, GY = {tblock.gy}, GZ = {tblock.gz}, "
                     f"TBSKIP = {int(tbskip)}, TBDB = {int(tbdb)}, TBPD = {int(tbpd)}, NA = {int(tblock.na)}, "
                     f"TBRP = {int(tbrp)}, TBYA = {int(tblock.ya)};")
    upd = meta.update
    ne = meta.num_exprs
    if upd is None:
        lines.append(f"  static constexpr int UPD = 0, NOUT = {meta.num_outputs};")
        lines.append("  static constexpr int TF[1] = {0}, CF[1] = {0};")
        lines.append("  static constexpr C DT = C(0);")
    else:
        mode = {"euler": 1, "kahan": 2}[upd["mode"]]
        from cutile.lowering.meta import cpp_literal

        lines.append(f"  static constexpr int UPD = {mode}, NOUT = {meta.num_outputs};")
        lines.append(f"  static constexpr int TF[{ne}] = {_arr(upd['tf'])}, CF[{ne}] = {_arr(upd['cf'] or [0] * ne)};")
        lines.append(f"  static constexpr C DT = {cpp_literal(upd['dt'], COMPUTE_CTYPES[compute])};")
    lines += [
        "  template <class A> ST_HD static void eval(const A& a, C* out) {",
    ]
    for k, e in enumerate(meta.expressions):
        lines.append(f"    out[{k}] = {e};")
    lines += ["  }", "};"]
    return "\n".join(lines)


def sweep_tag(cfg: KernelConfig, zdpp: bool = False) -> str:
    """Tag of a sweep variant; ``z`` marks z-neighbours read across lanes."""
    return cfg.tag + ("z" if zdpp else "")


def kernel_symbol(meta: StencilMeta, tag: str) -> str:
    return f"{meta.name}_sweep_{tag}"


def temporal_symbol(meta: StencilMeta, tag: str, T: int) -> str:
    return f"{meta.name}_temporal_tb{T}_{tag}"


def tblock_symbol(meta: StencilMeta, cfg: TBlockConfig) -> str:
    return f"{meta.name}_tblock_{cfg.tag}"


def brick_symbol(meta: StencilMeta, tag: str, brick: tuple) -> str:
    return f"{meta.name}_bricks{'x'.join(str(int(b)) for b in brick)}_{tag}"


def emit_hip_source(meta: StencilMeta, variants: Sequence[KernelConfig], storage: str,
                    compute: str, temporal=None, nontemporal: bool = True, prefetch: int = 1,
                    ntload: bool = False, tblocks: Sequence[TBlockConfig] = (),
                    bricked: Sequence[tuple] = (), zdpp: bool = False, tbskip: int = 0,
                    tbdb: "int | None" = None, tbpd: int = 2, ringpd: int = 1, sdb: int = 0,
                    tbrp: "int | None" = None) -> str:
    """Full HIP translation unit with one sweep kernel per variant (plus the
    temporal-blocking kernels: ``temporal`` = (config, T) of the LDS-ring
    kernel, ``tblocks`` the register/LDS ones)."""
    out = [
        f"// {meta.name}: generated by the HIP emitter for gfx950",
        f"// expression(s): " + " | ".join(meta.py_expressions),
        '#include "stencil_core.h"',
        "",
    ]
    for cfg in dict.fromkeys(variants):        # the vector and scalar variants may coincide
        tag = sweep_tag(cfg, zdpp)
        out.append(emit_traits(meta, cfg, storage, compute, tag, nontemporal, prefetch, ntload=ntload, zdpp=zdpp,
                               ringpd=ringpd, sdb=sdb))
        out.append(
            f'extern "C" __global__ void __launch_bounds__({cfg.nt}) '
            f"{kernel_symbol(meta, tag)}(st::SweepArgs a) {{ st::sweep_entry<P_{tag}>(a); }}"
        )
        out.append("")
    if temporal is not None:
        cfg, T = temporal
        tag = f"tb{T}_{cfg.tag}"
        out.append(emit_traits(meta, cfg, storage, compute, tag))
        out.append(
            f'extern "C" __global__ void __launch_bounds__({cfg.nt}) '
            f"{temporal_symbol(meta, cfg.tag, T)}(st::SweepArgs a) {{ st::temporal_entry<P_{tag}, {T}>(a); }}"
        )
        out.append("")
    for cfg, brick in bricked:
        tag = f"br_{cfg.tag}"
        out.append(emit_traits(meta, cfg, storage, compute, tag, nontemporal, prefetch, ntload=ntload, brick=brick))
        out.append(
            f'extern "C" __global__ void __launch_bounds__({cfg.nt}) '
            f"{brick_symbol(meta, cfg.tag, brick)}(st::SweepArgs a) {{ st::sweep_entry<P_{tag}>(a); }}"
        )
        out.append("")
    for tb in tblocks:
        tag = tb.tag
        # two sets of level planes (one barrier per step) where both fit the
        # LDS: 1-6 % faster on MI355X (tbdb 0/1 forces the choice).  Not for
        # the scalar variants (ragged rows): their queues already hold up to
        # 250 registers, and the second set's addresses (or a second
        # prefetch stage) spill them
        fits = tb.na == 0 and 2 * tblock_estimate(meta, tb, storage, compute)["lds"] <= 160 * 1024
        db = int(fits and (bool(tbdb) if tbdb is not None else tb.vz > 1))
        # the deeper prefetch likewise only for the vector variants
        pd = tbpd if tb.vz > 1 else 1
        # the two-level kernel's halo rings: QHI + PD planes ahead (0: PD)
        rp = (tblock_queue_range(meta)[1] + pd if tbrp is None else int(tbrp)) if tb.na else 0
        out.append(emit_traits(meta, tb.kernel_config(), storage, compute, tag, nontemporal, tblock=tb,
                               zdpp=zdpp, tbskip=tbskip, tbdb=db, tbpd=pd, tbrp=rp))
        out.append(
            f'extern "C" __global__ void __launch_bounds__({tb.nt}) '
            f"{tblock_symbol(meta, tb)}(st::SweepArgs a) {{ st::tblock_entry<P_{tag}>(a); }}"
        )
        out.append("")
    return "\n".join(out)
