"""Host-side model of the sweep kernel's address arithmetic.

Mirrors, index for index, ``st::sweep_tile`` in ``csrc/stencil_core.h``
(register-queue loads, LDS halo-ring fetches, LDS stores and reads, output
stores) for a given variant, array shape and launch, and reports every
access that would leave its buffer.  Used by the CPU tests to prove a
launch configuration in bounds before it runs on a device."""

from __future__ import annotations

import itertools

import numpy as np


def _tab(cfg, plan):
    NT, TY, TZ, VZ = cfg.nt, cfg.ty, cfg.tz, cfg.vz
    NTZ = TZ // VZ
    NTY = NT // NTZ
    KY = -(-TY // NTY)
    bufs = plan["bufs"]
    t = dict(NT=NT, TY=TY, TZ=TZ, VZ=VZ, NTZ=NTZ, NTY=NTY, KY=KY, bufs=[])
    off = 0
    hasq = plan["qhi"] >= plan["qlo"]
    for (f, dx, ry, rz) in bufs:
        rzv = -(-rz // VZ) * VZ
        rows, cols = TY + 2 * ry, TZ + 2 * rzv
        fromq = hasq and plan["qf"][f] and plan["qlo"] <= dx <= plan["qhi"]
        w = cols // VZ
        sv = rzv // VZ
        gvecs = 2 * ry * w + TY * 2 * sv if fromq else rows * w
        t["bufs"].append(dict(f=f, dx=dx, ry=ry, rz=rz, rzv=rzv, rows=rows, cols=cols, w=w, sv=sv,
                              fromq=fromq, gvecs=gvecs, nvb=-(-gvecs // NT), off=off))
        off += rows * cols
    t["lds"] = max(off, 1)
    t["nvb"] = max([1] + [b["nvb"] for b in t["bufs"]])
    return t


def queue_planes(xb, xe, qlo, qhi, pd):
    """Planes the register queue loads for chunk [xb, xe) with prefetch depth
    ``pd``: the prologue fills xb+qlo .. xb+qhi-1 and pre-loads xb+qhi+k for
    k < pd while xb+k < xe; step x then loads x+pd+qhi while x+pd < xe."""
    planes = [xb + qlo + i for i in range(qhi - qlo)]
    planes += [xb + qhi + k for k in range(pd) if xb + k < xe]
    planes += [x + pd + qhi for x in range(xb, xe) if x + pd < xe]
    return planes


def check_launch(cfg, plan, shape3, halo3, n3, xch, yslab, nxcd=8, accesses=None, tiles="edges", pd=1, wrap=0):
    """Return a list of out-of-bounds descriptions (empty when clean)."""
    s0, s1, s2 = shape3
    h0, h1, h2 = halo3
    n0, n1, n2 = n3
    t = _tab(cfg, plan)
    NT, TY, TZ, VZ, NTZ, NTY, KY = (t[k] for k in ("NT", "TY", "TZ", "VZ", "NTZ", "NTY", "KY"))
    qlo, qhi = plan["qlo"], plan["qhi"]
    hasq = qhi >= qlo
    numel = s0 * s1 * s2
    ntz, nty = -(-n2 // TZ), -(-n1 // TY)
    nch = -(-n0 // xch)
    errs = []
    tid = np.arange(NT)
    tz, ty = tid % NTZ, tid // NTZ

    def gcheck(what, idx, mask):
        idx = np.asarray(idx)[mask]
        if idx.size and (idx.min() < 0 or idx.max() >= numel):
            errs.append(f"{what}: index range [{idx.min()}, {idx.max()}] outside [0, {numel})")

    def lcheck(what, idx, mask):
        idx = np.asarray(idx)[mask]
        if idx.size and (idx.min() < 0 or idx.max() >= t["lds"]):
            errs.append(f"{what}: LDS index range [{idx.min()}, {idx.max()}] outside [0, {t['lds']})")

    def pick(n):
        return sorted({0, 1, n - 2, n - 1} & set(range(n))) if tiles == "edges" else range(n)

    # per-buffer fetch / LDS coordinates do not depend on the plane
    rc = {}
    for bi, b in enumerate(t["bufs"]):
        rr, cc, ok = [], [], []
        for j in range(t["nvb"]):
            i = tid + j * NT
            m = (j < b["nvb"]) & (i < b["gvecs"])
            pairs = [gvec_rc(b, int(ii)) if mm else (0, 0) for ii, mm in zip(i, m)]
            rr.append([p_[0] for p_ in pairs])
            cc.append([p_[1] for p_ in pairs])
            ok.append(m)
        rc[bi] = (np.array(rr), np.array(cc), np.array(ok))
        r_, c_, m_ = rc[bi]
        idx = b["off"] + r_ * b["cols"] + c_ * VZ
        lcheck("LDS halo store", idx, m_)
        lcheck("LDS halo store (end)", idx + VZ - 1, m_)
        if b["fromq"]:
            for k in range(KY):
                m = ty + k * NTY < TY
                d = b["off"] + (b["ry"] + ty + k * NTY) * b["cols"] + b["rzv"] + tz * VZ
                lcheck("LDS interior store", d, m)
                lcheck("LDS interior store (end)", d + VZ - 1, m)

    for tile_z, tile_y, chunk in itertools.product(pick(ntz), pick(nty), pick(nch)):
        z0, y0 = tile_z * TZ, tile_y * TY
        xb = chunk * xch
        xe = min(xb + xch, n0)
        zc = z0 + tz * VZ
        zin = zc < n2
        zload = h2 + zc + VZ <= s2
        # the extreme planes the queue touches (prologue and prefetches)
        if hasq:
            ql = queue_planes(xb, xe, qlo, qhi, pd)
            qplanes = (min(ql), max(ql))
        else:
            qplanes = None
        for k in range(KY):
            y = y0 + ty + k * NTY
            m = zload & (h1 + y < s1) & (ty + k * NTY < TY)
            if qplanes:
                for x in qplanes:
                    base = (h0 + x) * s1 * s2 + (h1 + y) * s2 + h2 + zc
                    gcheck(f"queue load plane {x}", base, m)
                    gcheck(f"queue load plane {x} (vector end)", base + VZ - 1, m)
        for bi, b in enumerate(t["bufs"]):
            r_, c_, m_ = rc[bi]
            gy = h1 + y0 - b["ry"] + r_
            gz = h2 + z0 - b["rzv"] + c_ * VZ
            ok = m_ & (gy >= 0) & (gy < s1) & (gz >= 0) & (gz + VZ <= s2)
            for x in (xb + b["dx"], xe - 1 + b["dx"]):
                p = (h0 + x) * s1 * s2
                gcheck(f"halo fetch plane {x}", p + gy * s2 + gz, ok)
                gcheck(f"halo fetch plane {x} (vector end)", p + gy * s2 + gz + VZ - 1, ok)
        for k in range(KY):
            lyk = ty + k * NTY
            y = y0 + lyk
            m = (lyk < TY) & (y < n1) & zin
            for (f, dx, dy, dz) in (accesses or []):
                if dy == 0 and dz == 0 and hasq and qlo <= dx <= qhi:
                    continue
                b = next(bb for bb in t["bufs"] if bb["f"] == f and bb["dx"] == dx)
                for v in range(VZ):
                    idx = b["off"] + (lyk + b["ry"] + dy) * b["cols"] + (tz * VZ + v + b["rzv"] + dz)
                    lcheck("LDS read", idx, m)
            for x in (xb, xe - 1):
                off = ((h0 + x) * s1 + (h1 + y)) * s2 + h2 + zc
                gcheck(f"output store plane {x}", off, m)
                gcheck(f"output store plane {x} (vector end)", off + VZ - 1, m & (zc + VZ <= n2))
                if wrap:
                    # images of cells near a face of a wrapped axis (st::sweep_tile store_wrapped)
                    wx = bool(wrap & 1) and (x < h0 or x >= n0 - h0)
                    xm = x + n0 if x < h0 else x - n0
                    wy = (wrap & 2) > 0
                    wy = wy & ((y < h1) | (y >= n1 - h1))
                    ym = np.where(y < h1, y + n1, y - n1)
                    for v in range(VZ):
                        z = zc + v
                        mz = m & (z < n2)
                        wz = bool(wrap & 4) & ((z < h2) | (z >= n2 - h2))
                        zm = np.where(z < h2, z + n2, z - n2)
                        for c in range(1, 8):
                            sel = mz.copy()
                            if c & 1:
                                if not wx:
                                    continue
                            if c & 2:
                                sel = sel & wy
                            if c & 4:
                                sel = sel & wz
                            X = xm if c & 1 else x
                            Y = ym if c & 2 else y
                            Z = zm if c & 4 else z
                            o = ((h0 + X) * s1 + (h1 + Y)) * s2 + h2 + Z
                            gcheck(f"wrapped store c={c} plane {x}", o, sel)
                            # an image must land in the halo, never in the interior
                            inside = ((X >= 0) & (X < n0)) & ((Y >= 0) & (Y < n1)) & ((Z >= 0) & (Z < n2))
                            if np.any(np.broadcast_to(inside, sel.shape) & sel):
                                errs.append(f"wrapped store c={c} plane {x} hits the interior")
        if len(errs) > 20:
            return errs
    return errs


def gvec_rc(b, i):
    """Python copy of ``Tab::gvec_rc``."""
    w = b["w"]
    if not b["fromq"]:
        return i // w, i % w
    band = b["ry"] * w
    if i < band:
        return i // w, i % w
    i -= band
    if i < band:
        return b["ry"] + b["rows"] - 2 * b["ry"] + i // w, i % w
    i -= band
    sv = b["sv"]
    r = b["ry"] + i // (2 * sv)
    k = i % (2 * sv)
    return r, (k if k < sv else w - 2 * sv + k)
