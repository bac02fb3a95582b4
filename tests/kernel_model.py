"""Host-side model of the sweep kernel's address arithmetic.

Mirrors, index for index, ``st::sweep_tile`` in ``csrc/stencil_core.h``
(register-queue loads, LDS halo-ring fetches, LDS stores and reads, output
stores) for a given variant, array shape and launch, and reports every
access that would leave its buffer.  Global accesses go through one buffer
resource per plane: the plane itself must lie in the array and a live
access must stay inside the plane (accesses the kernel marks out of range
are dropped by the hardware and are not checked).  ``wrap`` bit d maps reads
past the interior along axis d around it, as the kernel does.  Used by the
CPU tests to prove a launch configuration in bounds before it runs on a
device."""

from __future__ import annotations

import itertools

import numpy as np


def lds_pitch(tz, vz, rzv, elem_bytes):
    # row pitch of an LDS plane buffer: padded to tile width mod 256 B when a
    # row of tz / vz lanes does not end on a 32-lane boundary
    p = tz + 2 * rzv
    if (tz // vz) % 32:
        while ((p - tz) * elem_bytes) % 256:
            p += 1
    return p


def _tab(cfg, plan, elem_bytes=8):
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
        pitch = lds_pitch(TZ, VZ, rzv, elem_bytes)
        fromq = hasq and plan["qf"][f] and plan["qlo"] <= dx <= plan["qhi"]
        w = cols // VZ
        sv = rzv // VZ
        gvecs = 2 * ry * w + TY * 2 * sv if fromq else rows * w
        t["bufs"].append(dict(f=f, dx=dx, ry=ry, rz=rz, rzv=rzv, rows=rows, cols=cols, pitch=pitch, w=w, sv=sv,
                              fromq=fromq, gvecs=gvecs, nvb=-(-gvecs // NT), off=off))
        off += rows * pitch
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


def check_launch(cfg, plan, shape3, halo3, n3, xch, yslab, nxcd=8, accesses=None, tiles="edges", pd=1, wrap=0,
                 elem_bytes=8):
    """Return a list of out-of-bounds descriptions (empty when clean)."""
    s0, s1, s2 = shape3
    h0, h1, h2 = halo3
    n0, n1, n2 = n3
    t = _tab(cfg, plan, elem_bytes)
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

    def wrapi(i, n, on):
        i = np.asarray(i)
        return np.where(i < 0, i + n, np.where(i >= n, i - n, i)) if on else i

    def pcheck(what, x, rowcol, mask):
        """plane x (interior coords) inside the array, offsets inside the plane"""
        idx = np.asarray(rowcol)[np.broadcast_to(mask, np.shape(rowcol))]
        if idx.size and (idx.min() < 0 or idx.max() >= s1 * s2):
            errs.append(f"{what}: in-plane offset range [{idx.min()}, {idx.max()}] outside [0, {s1 * s2})")
        if np.any(mask) and not (0 <= h0 + x < s0):
            errs.append(f"{what}: plane {x} outside the array")

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
        idx = b["off"] + r_ * b["pitch"] + c_ * VZ
        lcheck("LDS halo store", idx, m_)
        lcheck("LDS halo store (end)", idx + VZ - 1, m_)
        if b["fromq"]:
            for k in range(KY):
                m = ty + k * NTY < TY
                d = b["off"] + (b["ry"] + ty + k * NTY) * b["pitch"] + b["rzv"] + tz * VZ
                lcheck("LDS interior store", d, m)
                lcheck("LDS interior store (end)", d + VZ - 1, m)

    for tile_z, tile_y, chunk in itertools.product(pick(ntz), pick(nty), pick(nch)):
        z0, y0 = tile_z * TZ, tile_y * TY
        xb = chunk * xch
        xe = min(xb + xch, n0)
        zc = z0 + tz * VZ
        zin = zc < n2
        zload = h2 + zc + VZ <= s2
        # every plane the queue loads (prologue and prefetches)
        qpl = queue_planes(xb, xe, qlo, qhi, pd) if hasq else []
        zq = wrapi(zc, n2, wrap & 4)
        zload_q = h2 + zq + VZ <= s2
        for k in range(KY):
            y = y0 + ty + k * NTY
            yq = wrapi(y, n1, wrap & 2)
            m = zload_q & (h1 + yq < s1) & (ty + k * NTY < TY)
            for x in (min(qpl), max(qpl)) if qpl else ():
                px = int(wrapi(x, n0, wrap & 1))
                pcheck(f"queue load plane {x}", px, (h1 + yq) * s2 + h2 + zq, m)
                pcheck(f"queue load plane {x} (vector end)", px, (h1 + yq) * s2 + h2 + zq + VZ - 1, m)
        for bi, b in enumerate(t["bufs"]):
            r_, c_, m_ = rc[bi]
            gy = h1 + wrapi(y0 - b["ry"] + r_, n1, wrap & 2)
            gz = h2 + wrapi(z0 - b["rzv"] + c_ * VZ, n2, wrap & 4)
            ok = m_ & (gy >= 0) & (gy < s1) & (gz >= 0) & (gz + VZ <= s2)
            for x in (xb + b["dx"], xe - 1 + b["dx"]):
                px = int(wrapi(x, n0, wrap & 1))
                pcheck(f"halo fetch plane {x}", px, gy * s2 + gz, ok)
                pcheck(f"halo fetch plane {x} (vector end)", px, gy * s2 + gz + VZ - 1, ok)
        for k in range(KY):
            lyk = ty + k * NTY
            y = y0 + lyk
            m = (lyk < TY) & (y < n1) & zin
            lyc = np.minimum(lyk, TY - 1)       # rows past the tile compute on the last row
            for (f, dx, dy, dz) in (accesses or []):
                if dy == 0 and dz == 0 and hasq and qlo <= dx <= qhi:
                    continue
                b = next(bb for bb in t["bufs"] if bb["f"] == f and bb["dx"] == dx)
                for v in range(VZ):
                    idx = b["off"] + (lyc + b["ry"] + dy) * b["pitch"] + (tz * VZ + v + b["rzv"] + dz)
                    lcheck("LDS read", idx, np.ones_like(m))
            for x in (xb, xe - 1):
                pcheck(f"output store plane {x}", x, (h1 + y) * s2 + h2 + zc, m)
                pcheck(f"output store plane {x} (vector end)", x, (h1 + y) * s2 + h2 + zc + VZ - 1,
                       m & (zc + VZ <= n2))
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
