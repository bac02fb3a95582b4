"""Host-side model of the register/LDS temporal-blocking kernel.

Mirrors ``st::tblock_tile`` in ``csrc/stencil_core.h`` step by step at cell
granularity: the per-level register queues and their slot rotation, the LDS
planes of every level (interior copied from the queue, level-0 halo ring
fetched from memory), the skew of the levels along the stream axis, the
masking of intermediate levels, and the stores of the last level.  Every
global access is checked against the array (plane inside the array, vector
inside the plane) and every LDS read against its buffer, and the values it
produces are returned so the tests can compare them with the oracle before
the kernel runs on a device.
"""

from __future__ import annotations

import itertools

import numpy as np

import oracle


def _pmod(a, n):
    return ((a % n) + n) % n


def _wrapm(i, n):
    return np.mod(i, n)


def map3(ndim, seq):
    seq = tuple(seq)
    if ndim == 1:
        return (1, 1, seq[0])
    if ndim == 2:
        return (seq[0], 1, seq[1])
    return seq


def halo3(ndim, halo):
    halo = tuple(halo)
    if ndim == 1:
        return (0, 0, halo[0])
    if ndim == 2:
        return (halo[0], 0, halo[1])
    return halo


def _offset3(ndim, offs):
    offs = tuple(offs)
    if ndim == 1:
        return (0, 0, offs[0])
    if ndim == 2:
        return (offs[0], 0, offs[1])
    return offs


class _Proxy:
    """The stencil's field argument: u[i+a, j+b] -> the level's values at
    that offset for every cell of the region."""

    def __init__(self, ndim, get):
        self.ndim, self.get = ndim, get

    def __getitem__(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        return self.get(*_offset3(self.ndim, key))


def gvec_rc(b, i):
    """Python copy of ``Tab::gvec_rc`` for a buffer whose interior comes from
    the queue (every temporal buffer does)."""
    w = b["w"]
    band = b["ry"] * w
    if i < band:
        return i // w, i % w
    i -= band
    if i < band:
        return b["ry"] + b["TY"] + i // w, i % w
    i -= band
    sv = b["sv"]
    r = b["ry"] + i // (2 * sv)
    k = i % (2 * sv)
    return r, (k if k < sv else w - 2 * sv + k)


def work_items(ntz, nty, n0, xch, edges=False):
    """(tile_z, tile_y, xb, xe) chunks the workgroups of a launch advance
    (chunks of ``xch`` planes of every tile); ``edges``: only the first two
    and last two tiles / chunks along each axis."""
    def pick(n):
        return sorted({0, 1, n - 2, n - 1} & set(range(n))) if edges else range(n)

    nch = -(-n0 // xch)
    return [(tz, ty, c * xch, min(c * xch + xch, n0))
            for tz, ty, c in itertools.product(pick(ntz), pick(nty), pick(nch))]


def run(fn, ndim, cfg, plan, qrange, u, halo, xch, tmode=0, wrap=0, out=None, edges=False, rad=None):
    """Run one launch of the temporal kernel on ``u`` (the padded input).

    ``cfg``: the variant's TBlockConfig; ``plan``: ``plan_accesses`` of the
    stencil; ``qrange``: ``tblock_queue_range``.  ``edges``: only the
    workgroups of the first two and last two tiles / chunks along each axis
    (bounds of a large launch; the output is then partial).  Returns
    (output array, list of error strings)."""
    outs, errs = run_fields([fn], [[0]], ndim, cfg, plan, qrange, [u], halo, xch, tmode, wrap,
                            None if out is None else [out], edges, rad)
    return outs[0], errs


def _arg_fields(fn, ndim, names):
    """Field index of each of ``fn``'s array parameters (by name)."""
    code = fn.__code__
    params = code.co_varnames[:code.co_argcount - ndim]
    return [names.index(p) for p in params]


def run_fields(fns, argmap, ndim, cfg, plan, qrange, us, halo, xch, tmode=0, wrap=0, outs=None, edges=False,
               rad=None):
    """``run`` for coupled fields: ``fns[m]`` gives the next state of field
    m from the fields listed by ``argmap[m]`` (one field index per array
    parameter; a single-field stencil: ``[0]``), ``us[f]`` is field f.
    Returns (one output array per field, error strings)."""
    if getattr(cfg, "na", 0):
        return run_fields_ws(fns, argmap, ndim, cfg, plan, qrange, us, halo, xch, tmode, wrap, outs, edges)
    NF = len(us)
    TT, TY, TZ, VZ, NT = cfg.T, cfg.ty, cfg.tz, cfg.vz, cfg.nt
    if rad is None:      # stencil radius in the 3-D map, from the LDS buffers and queue range
        rad = (max(-qrange[0], qrange[1]), max([b[2] for b in plan["bufs"]] + [0]),
               max([b[3] for b in plan["bufs"]] + [0]))
    GY, GZ, OTY, OTZ = cfg.gy, cfg.gz, cfg.oty, cfg.otz
    QLO, QHI = qrange
    QN, D = QHI - QLO + 1, QHI + 1
    errs = []
    u3s = [np.asarray(u).reshape(map3(ndim, u.shape)) for u in us]
    u3 = u3s[0]
    s0, s1, s2 = u3.shape
    h0, h1, h2 = halo3(ndim, halo)
    n0, n1, n2 = s0 - 2 * h0, s1 - 2 * h1, s2 - 2 * h2
    w0, w1, w2 = bool(wrap & 1), bool(wrap & 2), bool(wrap & 4)
    res = [np.array(x, copy=True) for x in u3s] if outs is None else [np.asarray(o).reshape(u3.shape) for o in outs]
    written = np.zeros(u3.shape, dtype=np.int32)
    fs = [oracle.array_aware(fn) for fn in fns]
    bufs = []
    for (fi, dx, ry, rz) in plan["bufs"]:
        rzv = -(-rz // VZ) * VZ
        cols = TZ + 2 * rzv
        w = cols // VZ
        sv = rzv // VZ
        bufs.append(dict(f=fi, dx=dx, ry=ry, rz=rz, rzv=rzv, rows=TY + 2 * ry, cols=cols, w=w, sv=sv, TY=TY,
                         gvecs=2 * ry * w + TY * 2 * sv))
    bidx = {(b["f"], b["dx"]): k for k, b in enumerate(bufs)}
    cells = TY * (TZ // VZ)
    if -(-cells // NT) != cfg.kc:
        errs.append("cells per thread")
    ntz, nty, nch = -(-n2 // OTZ), -(-n1 // OTY), -(-n0 // xch)
    ly = np.arange(TY)[:, None] * np.ones((1, TZ), dtype=int)
    lz = np.ones((TY, 1), dtype=int) * np.arange(TZ)[None, :]
    lzv = (lz // VZ) * VZ                       # first cell of each cell's vector

    def plane(x):
        px = int(_wrapm(x, n0)) if w0 else x
        return px, (0 <= h0 + px < s0)

    def pick(n):
        return sorted({0, 1, n - 2, n - 1} & set(range(n))) if edges else range(n)

    for tile_z, tile_y, chunk in itertools.product(pick(ntz), pick(nty), pick(nch)):
        y0, z0 = tile_y * OTY - GY, tile_z * OTZ - GZ
        xb = chunk * xch
        xe = min(xb + xch, n0)
        y = y0 + ly
        z = z0 + lz
        zv = z0 + lzv
        iny, inz = (y >= 0) & (y < n1), (z >= 0) & (z < n2)
        keep = (w1 | iny) & (w2 | inz)
        yq = _wrapm(y, n1) if w1 else y
        zq = _wrapm(zv, n2) if w2 else zv
        okq = (h1 + yq >= 0) & (h1 + yq < s1) & (h2 + zq >= 0) & (h2 + zq + VZ <= s2)
        mine = (ly >= GY) & (ly < GY + OTY) & (lz >= GZ) & (lz < GZ + OTZ) & iny & inz
        vz_off = lz - lzv

        def load_q0(fi, x, live):
            px, ok = plane(x)
            ok = ok and live
            vals = np.zeros((TY, TZ), dtype=np.float64)
            if ok:
                m = okq
                if m.any():
                    if (h2 + zq[m] + VZ > s2).any() or (h1 + yq[m] >= s1).any():
                        errs.append(f"queue load outside plane {x}")
                    vals[m] = u3s[fi][h0 + px, h1 + yq[m], h2 + zq[m] + vz_off[m]]
            return vals

        # level-0 halo ring of every buffer: LDS (row, column) and in-plane
        # array coordinates of each fetched vector (fixed per workgroup)
        rings = []
        for b in bufs:
            rc = np.array([gvec_rc(b, i) for i in range(b["gvecs"])], dtype=int).reshape(-1, 2)
            rr, cc = rc[:, 0], rc[:, 1]
            inside = (rr >= b["ry"]) & (rr < b["ry"] + TY) & (cc * VZ >= b["rzv"]) & (cc * VZ < b["rzv"] + TZ)
            if inside.any():
                errs.append("ring vector inside the tile interior")
            cover = np.zeros((b["rows"], b["w"]), dtype=int)
            np.add.at(cover, (rr, cc), 1)
            cover[b["ry"]:b["ry"] + TY, b["sv"]:b["sv"] + TZ // VZ] += 1
            if not (cover == 1).all():
                errs.append("ring and interior do not tile the LDS plane")
            gy, gz = y0 - b["ry"] + rr, z0 - b["rzv"] + cc * VZ
            if w1:
                gy = _wrapm(gy, n1)
            if w2:
                gz = _wrapm(gz, n2)
            gy, gz = gy + h1, gz + h2
            ok = (gy >= 0) & (gy < s1) & (gz >= 0) & (gz + VZ <= s2)
            rings.append((rr, cc, gy, gz, ok))

        def load_ring(x, live):
            ring = []
            for b, (rr, cc, gy, gz, ok) in zip(bufs, rings):
                px, pok = plane(x + b["dx"])
                v = np.zeros((len(rr), VZ))
                if pok and live and ok.any():
                    cols = gz[ok][:, None] + np.arange(VZ)[None, :]
                    v[ok] = u3s[b["f"]][h0 + px, gy[ok][:, None], cols]
                ring.append(v)
            return ring

        s_first = xb + (TT - 1) * QLO
        s_last = xe - 1 + (TT - 1) * D
        q = [[[np.zeros((TY, TZ)) for _ in range(QN)] for _ in range(NF)] for _ in range(TT)]
        lds = [[np.zeros((b["rows"], b["cols"])) for b in bufs] for _ in range(TT)]
        for fi in range(NF):
            for i in range(QN - 1):
                q[0][fi][i] = load_q0(fi, s_first + QLO + i, True)
        pq = [load_q0(fi, s_first + QHI, True) for fi in range(NF)]
        pb = load_ring(s_first, True)
        for s in range(s_first, s_last + 1):
            R = (s - s_first) % QN
            for fi in range(NF):
                q[0][fi][_pmod(R - 1, QN)] = pq[fi]
            for L in range(TT):
                for k, b in enumerate(bufs):
                    sl = _pmod(R - L * D + b["dx"] - QLO, QN)
                    lds[L][k][b["ry"]:b["ry"] + TY, b["rzv"]:b["rzv"] + TZ] = q[L][b["f"]][sl]
                    if L == 0:
                        rr, cc = rings[k][0], rings[k][1]
                        cols = cc[:, None] * VZ + np.arange(VZ)[None, :]
                        lds[0][k][rr[:, None], cols] = pb[k]
            pq = [load_q0(fi, s + 1 + QHI, s + 1 <= s_last) for fi in range(NF)]
            pb = load_ring(s + 1, s + 1 <= s_last)
            for L in range(TT, 0, -1):
                RL = _pmod(R - (L - 1) * D, QN)
                p = s - (L - 1) * D

                def get(fi, dx, dy, dz, L=L, RL=RL):
                    if dy == 0 and dz == 0:
                        if not QLO <= dx <= QHI:
                            errs.append(f"queue read dx={dx} outside [{QLO}, {QHI}]")
                            return np.zeros((TY, TZ))
                        return q[L - 1][fi][_pmod(dx - QLO + RL, QN)]
                    if (fi, dx) not in bidx:
                        errs.append(f"no LDS buffer for field {fi}, dx={dx}")
                        return np.zeros((TY, TZ))
                    b = bufs[bidx[fi, dx]]
                    r0, c0 = b["ry"] + dy, b["rzv"] + dz
                    if r0 < 0 or r0 + TY > b["rows"] or c0 < 0 or c0 + TZ > b["cols"]:
                        errs.append(f"LDS read ({fi},{dx},{dy},{dz}) outside buffer")
                        return np.zeros((TY, TZ))
                    return lds[L - 1][bidx[fi, dx]][r0:r0 + TY, c0:c0 + TZ]

                vals = []
                for fm, amap in zip(fs, argmap):
                    proxies = [_Proxy(ndim, lambda dx, dy, dz, fi=fi: get(fi, dx, dy, dz)) for fi in amap]
                    vals.append(np.broadcast_to(np.asarray(fm(*proxies, *([0] * ndim)), dtype=np.float64),
                                                (TY, TZ)))
                # the kernel may skip (and zero) cells outside the output tile
                # widened by (TT-L) radii; zero them all here, so a needed
                # cell skipped there fails the comparison with the oracle
                wy, wz = (TT - L) * rad[1], (TT - L) * rad[2]
                need = (ly >= GY - wy) & (ly < GY + OTY + wy) & (lzv + VZ > GZ - wz) & (lzv < GZ + OTZ + wz)
                vals = [np.where(need, v, 0.0) for v in vals]
                if L == TT:
                    if p >= xb:
                        if not (0 <= p < n0):
                            errs.append(f"store plane {p} outside the interior")
                            continue
                        m = mine
                        for fi in range(NF):
                            res[fi][h0 + p, h1 + y[m], h2 + z[m]] = vals[fi][m]
                        written[h0 + p, h1 + y[m], h2 + z[m]] += 1
                else:
                    pin = w0 or (0 <= p < n0)
                    sl = _pmod(R - (L - 1) * D - QLO, QN)
                    for fi in range(NF):
                        if tmode == 0:
                            q[L][fi][sl] = np.where(pin & keep, vals[fi], 0.0)
                        else:
                            q[L][fi][sl] = np.array(vals[fi], copy=True)
    inner = tuple(slice(h, s - h) for h, s in zip((h0, h1, h2), u3.shape))
    if written.max() > 1 or (not edges and not (written[inner] == 1).all()):
        errs.append(f"interior cells written {written[inner].min()}..{written[inner].max()} times")
    wmask = np.ones(u3.shape, dtype=bool)
    wmask[inner] = False
    if written[wmask].any():
        errs.append("halo cells written")
    return [r.reshape(np.shape(us[0])) for r in res], errs


def run_fields_ws(fns, argmap, ndim, cfg, plan, qrange, us, halo, xch, tmode=0, wrap=0, outs=None, edges=False):
    """``run_fields`` for the two-level kernel with wave roles
    (``st::tbws_tile``): the level-1 role's queue over the region, the two
    level-0 LDS planes (interior from that queue, halo ring fetched by the
    level-2 role), the level-1 ring of QHI + 2 planes, the level-2 role's
    queue over the output tile and its stores, step by step.  Within a step
    both roles run between two barriers, so no LDS plane or ring slot may be
    written by one role and read by the other in the same step: that is
    checked too."""
    NF = len(us)
    TT, TY, TZ, VZ, NT, NA = cfg.T, cfg.ty, cfg.tz, cfg.vz, cfg.nt, cfg.na
    NBT = NT - NA
    GY, GZ, OTY, OTZ = cfg.gy, cfg.gz, cfg.oty, cfg.otz
    QLO, QHI = qrange
    QN, D = QHI - QLO + 1, QHI + 1
    NR = D + 1
    PD = 2 if VZ > 1 else 1
    RB = QHI + PD                 # the level-2 threads' halo rings: QHI + PD planes ahead
    errs = []
    if TT != 2 or not (QLO <= -1 and QHI >= 1):
        errs.append("two-level kernel: T = 2 and column accesses on both sides")
    u3s = [np.asarray(u).reshape(map3(ndim, u.shape)) for u in us]
    u3 = u3s[0]
    s0, s1, s2 = u3.shape
    h0, h1, h2 = halo3(ndim, halo)
    n0, n1, n2 = s0 - 2 * h0, s1 - 2 * h1, s2 - 2 * h2
    w0, w1, w2 = bool(wrap & 1), bool(wrap & 2), bool(wrap & 4)
    res = [np.array(x, copy=True) for x in u3s] if outs is None else [np.asarray(o).reshape(u3.shape) for o in outs]
    written = np.zeros(u3.shape, dtype=np.int32)
    fs = [oracle.array_aware(fn) for fn in fns]
    bufs = []
    for (fi, dx, ry, rz) in plan["bufs"]:
        if dx != 0:
            errs.append(f"in-plane access at dx={dx}")
        rzv = -(-rz // VZ) * VZ
        cols = TZ + 2 * rzv
        w = cols // VZ
        sv = rzv // VZ
        bufs.append(dict(f=fi, dx=dx, ry=ry, rz=rz, rzv=rzv, rows=TY + 2 * ry, cols=cols, w=w, sv=sv, TY=TY,
                         gvecs=2 * ry * w + TY * 2 * sv))
    bidx = {fi: k for k, b in enumerate(bufs) for fi in [b["f"]]}
    if -(-(TY * (TZ // VZ)) // NA) != cfg.kc or -(-(OTY * (OTZ // VZ)) // NBT) != cfg.kb:
        errs.append("cells per thread")
    if GY < max([b["ry"] for b in bufs] + [0]) or GZ < max([b["rz"] for b in bufs] + [0]):
        errs.append("ghost ring narrower than the radius")
    ntz, nty = -(-n2 // OTZ), -(-n1 // OTY)
    ly = np.arange(TY)[:, None] * np.ones((1, TZ), dtype=int)
    lz = np.ones((TY, 1), dtype=int) * np.arange(TZ)[None, :]
    lzv = (lz // VZ) * VZ
    vz_off = lz - lzv
    osl = (slice(GY, GY + OTY), slice(GZ, GZ + OTZ))   # the output tile in region coordinates

    def plane(x):
        px = int(_wrapm(x, n0)) if w0 else x
        return px, (0 <= h0 + px < s0)

    for tile_z, tile_y, xb, xe in work_items(ntz, nty, n0, xch, edges):
        y0, z0 = tile_y * OTY - GY, tile_z * OTZ - GZ
        y, z, zv = y0 + ly, z0 + lz, z0 + lzv
        # level-1 role: keep (tmode 0) from the vector's first cell, as the kernel
        keep = (w1 | ((y >= 0) & (y < n1))) & (w2 | ((zv >= 0) & (zv < n2)))
        yq = _wrapm(y, n1) if w1 else y
        zq = _wrapm(zv, n2) if w2 else zv
        okq = (h1 + yq >= 0) & (h1 + yq < s1) & (h2 + zq >= 0) & (h2 + zq + VZ <= s2)
        mine = np.zeros((TY, TZ), dtype=bool)
        mine[osl] = True
        mine &= (y >= 0) & (y < n1) & (z >= 0) & (z < n2)

        def load_q0(fi, x, live):
            px, ok = plane(x)
            vals = np.zeros((TY, TZ), dtype=np.float64)
            if ok and live and okq.any():
                m = okq
                vals[m] = u3s[fi][h0 + px, h1 + yq[m], h2 + zq[m] + vz_off[m]]
            return vals

        rings = []
        for b in bufs:
            rc = np.array([gvec_rc(b, i) for i in range(b["gvecs"])], dtype=int).reshape(-1, 2)
            rr, cc = rc[:, 0], rc[:, 1]
            cover = np.zeros((b["rows"], b["w"]), dtype=int)
            np.add.at(cover, (rr, cc), 1)
            cover[b["ry"]:b["ry"] + TY, b["sv"]:b["sv"] + TZ // VZ] += 1
            if not (cover == 1).all():
                errs.append("ring and interior do not tile the LDS plane")
            if -(-b["gvecs"] // NBT) > 64:
                errs.append("ring vectors per level-2 thread")
            gy, gz = y0 - b["ry"] + rr, z0 - b["rzv"] + cc * VZ
            if w1:
                gy = _wrapm(gy, n1)
            if w2:
                gz = _wrapm(gz, n2)
            gy, gz = gy + h1, gz + h2
            ok = (gy >= 0) & (gy < s1) & (gz >= 0) & (gz + VZ <= s2)
            rings.append((rr, cc, gy, gz, ok))

        def load_ring(x, live):
            out = []
            for b, (rr, cc, gy, gz, ok) in zip(bufs, rings):
                px, pok = plane(x)
                v = np.zeros((len(rr), VZ))
                if pok and live and ok.any():
                    cols = gz[ok][:, None] + np.arange(VZ)[None, :]
                    v[ok] = u3s[b["f"]][h0 + px, gy[ok][:, None], cols]
                out.append(v)
            return out

        def put_ring(dst, vals):
            for k, (b, (rr, cc, _, _, _)) in enumerate(zip(bufs, rings)):
                cols = cc[:, None] * VZ + np.arange(VZ)[None, :]
                dst[k][rr[:, None], cols] = vals[k]

        def put_l0(dst, slot):
            for k, b in enumerate(bufs):
                dst[k][b["ry"]:b["ry"] + TY, b["rzv"]:b["rzv"] + TZ] = qa[b["f"]][slot]

        s_first, s_last = xb + QLO, xe - 1 + D
        l0 = [[np.zeros((b["rows"], b["cols"])) for b in bufs] for _ in range(2)]
        ring = [[np.zeros((TY, TZ)) for _ in range(NF)] for _ in range(NR)]
        qa = [[np.zeros((TY, TZ)) for _ in range(QN)] for _ in range(NF)]
        qb = [[np.zeros((OTY, OTZ)) for _ in range(QN)] for _ in range(NF)]
        for fi in range(NF):
            for i in range(QN - 1):
                qa[fi][i] = load_q0(fi, s_first + QLO + i, True)
        pqa = [[load_q0(fi, s_first + QHI + p, True) for fi in range(NF)] for p in range(PD)]
        pb = [load_ring(s_first + p, s_first + p <= s_last) for p in range(RB)]
        put_l0(l0[0], -QLO)
        put_ring(l0[0], pb[0])
        pb[0] = load_ring(s_first + RB, s_first + RB <= s_last)
        for s in range(s_first, s_last + 1):
            r = s - s_first
            reads, writes = set(), set()
            # ---- level 1, plane s
            for fi in range(NF):
                qa[fi][_pmod(r - 1, QN)] = pqa[r % PD][fi]
            pqa[r % PD] = [load_q0(fi, s + QHI + PD, s + PD <= s_last) for fi in range(NF)]

            def get_a(fi, dx, dy, dz):
                if dy == 0 and dz == 0:
                    if not QLO <= dx <= QHI:
                        errs.append(f"queue read dx={dx} outside [{QLO}, {QHI}]")
                        return np.zeros((TY, TZ))
                    return qa[fi][_pmod(dx - QLO + r, QN)]
                if dx != 0 or fi not in bidx:
                    errs.append(f"no level-0 LDS buffer for field {fi}, dx={dx}")
                    return np.zeros((TY, TZ))
                b = bufs[bidx[fi]]
                r0, c0 = b["ry"] + dy, b["rzv"] + dz
                if r0 < 0 or r0 + TY > b["rows"] or c0 < 0 or c0 + TZ > b["cols"]:
                    errs.append(f"level-0 LDS read ({fi},{dy},{dz}) outside its buffer")
                    return np.zeros((TY, TZ))
                reads.add(("l0", r & 1))
                return l0[r & 1][bidx[fi]][r0:r0 + TY, c0:c0 + TZ]

            vals = []
            for fm, amap in zip(fs, argmap):
                proxies = [_Proxy(ndim, lambda dx, dy, dz, fi=fi: get_a(fi, dx, dy, dz)) for fi in amap]
                vals.append(np.broadcast_to(np.asarray(fm(*proxies, *([0] * ndim)), dtype=np.float64), (TY, TZ)))
            pin = w0 or (0 <= s < n0)
            for fi in range(NF):
                ring[r % NR][fi] = np.where(pin & keep, vals[fi], 0.0) if tmode == 0 else np.array(vals[fi], copy=True)
            writes.add(("ring", r % NR))
            put_l0(l0[(r + 1) & 1], _pmod(r + 1 - QLO, QN))
            writes.add(("l0", (r + 1) & 1))
            # ---- level 2, plane s - D
            for fi in range(NF):
                qb[fi][_pmod(r - 1, QN)] = ring[_pmod(r - 1, NR)][fi][osl].copy()
            reads.add(("ring", _pmod(r - 1, NR)))
            put_ring(l0[(r + 1) & 1], pb[(r + 1) % RB])
            pb[(r + 1) % RB] = load_ring(s + 1 + RB, s + 1 + RB <= s_last)
            p = s - D
            sl = _pmod(r - D, NR)

            def get_b(fi, dx, dy, dz):
                if dy == 0 and dz == 0:
                    if not QLO <= dx <= QHI:
                        errs.append(f"queue read dx={dx} outside [{QLO}, {QHI}]")
                        return np.zeros((OTY, OTZ))
                    return qb[fi][_pmod(dx - QLO + r, QN)]
                if dx != 0:
                    errs.append(f"ring read at dx={dx}")
                    return np.zeros((OTY, OTZ))
                r0, c0 = GY + dy, GZ + dz
                if r0 < 0 or r0 + OTY > TY or c0 < 0 or c0 + OTZ > TZ:
                    errs.append(f"ring read ({fi},{dy},{dz}) outside the region")
                    return np.zeros((OTY, OTZ))
                reads.add(("ring", sl))
                return ring[sl][fi][r0:r0 + OTY, c0:c0 + OTZ]

            ovals = []
            for fm, amap in zip(fs, argmap):
                proxies = [_Proxy(ndim, lambda dx, dy, dz, fi=fi: get_b(fi, dx, dy, dz)) for fi in amap]
                ovals.append(np.broadcast_to(np.asarray(fm(*proxies, *([0] * ndim)), dtype=np.float64),
                                             (OTY, OTZ)))
            if reads & writes:
                errs.append(f"step {s}: read and written in one step: {sorted(reads & writes)}")
            if p >= xb:
                if not (0 <= p < n0):
                    errs.append(f"store plane {p} outside the interior")
                    continue
                m = mine[osl]
                yy, zz = y[osl][m], z[osl][m]
                for fi in range(NF):
                    res[fi][h0 + p, h1 + yy, h2 + zz] = ovals[fi][m]
                written[h0 + p, h1 + yy, h2 + zz] += 1
    inner = tuple(slice(h, s - h) for h, s in zip((h0, h1, h2), u3.shape))
    if written.max() > 1 or (not edges and not (written[inner] == 1).all()):
        errs.append(f"interior cells written {written[inner].min()}..{written[inner].max()} times")
    wmask = np.ones(u3.shape, dtype=bool)
    wmask[inner] = False
    if written[wmask].any():
        errs.append("halo cells written")
    return [r.reshape(np.shape(us[0])) for r in res], errs
