"""
CPU restatement of the hit merge (test infrastructure: the checker for
mi_sim_merge_hits_device; never used by the product path).

Follows SPM::merge_partial_select (sysrap/SPM.cu:135-240) with the functors of
sphoton (sysrap/sphoton.h:292-320) and sphotonlite (sysrap/sphotonlite.h:68-97):
select flagmask & mask != 0 (all records for mask 0xffffffff), key =
(identity << 48) | uint32(time / window) taken modulo 2^64, stable sort by key,
then each run of equal keys folded left to right with reduce_op.
"""
import numpy as np

ALREADY_SELECTED = 0xFFFFFFFF


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _keys(ident, time, tw):
    bucket = (time.astype(np.float32) / np.float32(tw)).astype(np.float32)
    b = np.floor(bucket).astype(np.uint64) & np.uint64(0xFFFFFFFF)
    return ((ident.astype(np.uint64) << np.uint64(48)) & np.uint64(0xFFFFFFFFFFFFFFFF)) | b


def merge_photons(photons, mask, tw):
    """photons (n,4,4) float32 sphoton -> merged (m,4,4) float32"""
    p = np.ascontiguousarray(photons, np.float32).reshape(-1, 16)
    u = p.view(np.uint32)
    sel = np.arange(len(p)) if mask == ALREADY_SELECTED else np.nonzero(u[:, 15] & np.uint32(mask))[0]
    if tw == 0.0:
        return p[sel].reshape(-1, 4, 4).copy()
    ident = u[sel, 13] & np.uint32(0xFFFFFF)
    keys = _keys(ident, p[sel, 3], tw)
    order = sel[np.argsort(keys, kind="stable")]
    ks = np.sort(keys, kind="stable")
    out = []
    start = 0
    for i in range(1, len(ks) + 1):
        if i == len(ks) or ks[i] != ks[start]:
            run = order[start:i]
            best = run[0]
            fm = int(u[best, 15])
            hc = int(u[best, 7]) >> 16
            for b in run[1:]:
                if not (min(p[best, 3], p[b, 3]) == p[best, 3]):
                    best = b
                fm |= int(u[b, 15])
                hc = (hc + (int(u[b, 7]) >> 16)) & 0xFFFF
            r = u[best].copy()
            r[7] = (r[7] & np.uint32(0xFFFF)) | np.uint32(hc << 16)
            r[15] = np.uint32(fm)
            out.append(r)
            start = i
    return np.array(out, np.uint32).reshape(-1, 4, 4).view(np.float32)


def merge_lites(lites, mask, tw):
    """lites (n,4) uint32 words of sphotonlite -> merged (m,4) uint32"""
    u = np.ascontiguousarray(lites).view(np.uint32).reshape(-1, 4)
    t = u[:, 1].view(np.float32)
    sel = np.arange(len(u)) if mask == ALREADY_SELECTED else np.nonzero(u[:, 3] & np.uint32(mask))[0]
    if tw == 0.0:
        return u[sel].copy()
    keys = _keys(u[sel, 0] & np.uint32(0xFFFF), t[sel], tw)
    order = sel[np.argsort(keys, kind="stable")]
    ks = np.sort(keys, kind="stable")
    out = []
    start = 0
    for i in range(1, len(ks) + 1):
        if i == len(ks) or ks[i] != ks[start]:
            run = order[start:i]
            a = u[run[0]].copy()
            hci, tt, fm = int(a[0]), np.float32(t[run[0]]), int(a[3])
            for b in run[1:]:
                tt = np.float32(min(tt, t[b]))
                fm |= int(u[b, 3])
                hc = ((hci >> 16) + (int(u[b, 0]) >> 16)) & 0xFFFF
                hci = (hc << 16) | (hci & 0xFFFF)
            a[0] = np.uint32(hci)
            a[1] = np.array([tt], np.float32).view(np.uint32)[0]
            a[3] = np.uint32(fm)
            out.append(a)
            start = i
    return np.array(out, np.uint32).reshape(-1, 4)
