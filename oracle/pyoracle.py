"""
ctypes access to the CPU checker (oracle/liboracle.so) and, when it was built,
to the reference's own CSG intersection (oracle/_ref/libref_csg.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libref_csg.so")


class OracleGeom(C.Structure):
    _fields_ = [("node", C.c_void_p), ("n_node", C.c_int64),
                ("prim", C.c_void_p), ("n_prim", C.c_int64),
                ("solid", C.c_void_p), ("n_solid", C.c_int64),
                ("itra", C.c_void_p), ("n_tran", C.c_int64),
                ("inst", C.c_void_p), ("n_inst", C.c_int64),
                ("plan", C.c_void_p), ("n_plan", C.c_int64)]


def _geom(g):
    """OracleGeom over the (kept alive by the caller) contiguous arrays of g"""
    plan = g.get("plan")
    return OracleGeom(_p(g["node"]), len(g["node"]), _p(g["prim"]), len(g["prim"]),
                      _p(g["solid"]), len(g["solid"]), _p(g["itra"]), len(g["itra"]),
                      _p(g["inst"]), len(g["inst"]), _p(plan), 0 if plan is None else len(plan))


class OraclePhys(C.Structure):
    _fields_ = [("bnd", C.c_void_p), ("n_bnd", C.c_int32), ("n_wl", C.c_int32),
                ("wl_low", C.c_float), ("wl_step", C.c_float),
                ("optical", C.c_void_p),
                ("icdf", C.c_void_p), ("n_icdf_layer", C.c_int32), ("n_icdf", C.c_int32),
                ("icdf_hd_factor", C.c_int32)]


class OracleCfg(C.Structure):
    _fields_ = [("max_bounce", C.c_int32), ("max_time", C.c_float), ("tmin", C.c_float),
                ("tmin0", C.c_float), ("eps0_mask", C.c_uint32), ("tmax", C.c_float),
                ("seed", C.c_uint64), ("rng_offset", C.c_uint64), ("event_skipahead", C.c_uint64),
                ("rng_kind", C.c_int32), ("precooked", C.c_void_p), ("precooked_ni", C.c_int64),
                ("precooked_nk", C.c_int64), ("traverse", C.c_int32), ("nthreads", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_world_create.restype = C.c_void_p
        L.oracle_world_create.argtypes = [C.POINTER(OracleGeom)]
        L.oracle_world_destroy.argtypes = [C.c_void_p]
        L.oracle_simulate.restype = C.c_int64
        L.oracle_simulate.argtypes = [C.c_void_p, C.POINTER(OraclePhys), C.POINTER(OracleCfg), C.c_int32,
                                      C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_void_p, C.c_void_p]
        L.oracle_trace.restype = C.c_int
        L.oracle_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32]
        L.oracle_simulate_ex.argtypes = [C.c_void_p, C.POINTER(OraclePhys), C.POINTER(OracleCfg), C.c_int32,
                                         C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                         C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_simulate_ex.restype = C.c_int64
        L.oracle_simtrace.argtypes = [C.c_void_p, C.POINTER(OracleCfg), C.c_void_p, C.c_int64, C.c_void_p]
        L.oracle_simtrace.restype = C.c_int64
        L.oracle_atan2f_n.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_adjoint_rows.argtypes = [C.c_void_p, C.POINTER(OraclePhys), C.POINTER(OracleCfg), C.c_int32,
                                          C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64,
                                          C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.oracle_adjoint_rows.restype = C.c_int
        L.oracle_intersect_prim.restype = C.c_int
        L.oracle_intersect_prim.argtypes = [C.POINTER(OracleGeom), C.c_int32, C.c_float,
                                            C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_packed_lut.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_tree_counts.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int32]
        L.oracle_logf.restype = C.c_float
        L.oracle_logf.argtypes = [C.c_float]
        L.oracle_sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.oracle_logf_n.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_sincosf_n.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_philox_block.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        L.oracle_uniforms.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                      C.c_int32, C.c_void_p]
        L.oracle_xorwow_words.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                          C.c_int32, C.c_void_p, C.c_void_p]
        L.oracle_xorwow_jump.argtypes = [C.c_uint64, C.c_void_p]
        L.oracle_xorwow_from_state.argtypes = [C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p, C.c_void_p]
        L.oracle_step_boundary.restype = C.c_int
        L.oracle_step_boundary.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float,
                                           C.c_float, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_float)]
        L.oracle_rayleigh.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_int32,
                                      C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def reflib():
    """The reference's CSG intersection compiled from its own headers, or None."""
    global _ref
    if _ref is None and os.path.exists(REFLIB):
        L = C.CDLL(REFLIB)
        L.ref_intersect_prim.restype = C.c_int
        L.ref_intersect_prim.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                         C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_lut_lookup.restype = C.c_int
        L.ref_lut_lookup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        _ref = L
    return _ref


def _p(a):
    return a.ctypes.data if a is not None else None


class World:
    """Geometry + physics held for repeated oracle runs."""

    def __init__(self, geometry, physics):
        self._keep = {k: np.ascontiguousarray(v) for k, v in geometry.items()}
        g = self._keep
        self.geom = _geom(g)
        self._bnd = np.ascontiguousarray(physics["bnd"], np.float32)
        self._opt = np.ascontiguousarray(physics["optical"], np.int32)
        icdf = physics.get("icdf")
        self._icdf = np.ascontiguousarray(icdf, np.float32) if icdf is not None else None
        self.phys = OraclePhys(_p(self._bnd), self._bnd.shape[0], self._bnd.shape[3],
                               float(physics["wl_low"]), float(physics["wl_step"]), _p(self._opt),
                               _p(self._icdf), 0 if icdf is None else self._icdf.shape[0],
                               0 if icdf is None else self._icdf.shape[1],
                               int(physics.get("icdf_hd_factor", 0)))
        self.handle = lib().oracle_world_create(C.byref(self.geom))
        if not self.handle:
            raise RuntimeError("oracle_world_create failed")

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().oracle_world_destroy(self.handle)
        except Exception:
            pass

    def cfg(self, seed=0, max_bounce=31, traverse=1, nthreads=0, precooked=None, rng="philox", **kw):
        c = OracleCfg()
        c.max_bounce = max_bounce
        c.max_time = kw.get("max_time", 1e27)
        c.tmin = kw.get("tmin", 0.05)
        c.tmin0 = kw.get("tmin0", 0.05)
        c.eps0_mask = kw.get("eps0_mask", 0x4 | 0x1 | 0x2 | 0x20 | 0x10)
        c.tmax = kw.get("tmax", 1e6)
        c.seed = seed
        c.rng_offset = kw.get("rng_offset", 0)
        c.event_skipahead = kw.get("event_skipahead", 100000)
        c.rng_kind = {"philox": 0, "xorwow": 2}[rng]
        if precooked is not None:
            self._pre = np.ascontiguousarray(precooked, np.float32)
            c.rng_kind = 1
            c.precooked = self._pre.ctypes.data
            c.precooked_ni, c.precooked_nk = self._pre.shape
        c.traverse = traverse
        c.nthreads = nthreads
        return c

    def simulate(self, gensteps, event_idx=0, photon_offset=0, p_begin=0, p_end=None, **cfgkw):
        gs = np.ascontiguousarray(gensteps, np.float32).reshape(-1, 6, 4)
        total = int(gs.view(np.uint32)[:, 0, 3].sum())
        if p_end is None or p_end > total:
            p_end = total
        n = max(0, p_end - p_begin)
        photon = np.zeros((n, 4, 4), np.float32)
        seq = np.zeros((n, 2, 2), np.uint64)
        cfg = self.cfg(**cfgkw)
        rc = lib().oracle_simulate(self.handle, C.byref(self.phys), C.byref(cfg), event_idx,
                                   gs.ctypes.data, len(gs), photon_offset, p_begin, p_end,
                                   photon.ctypes.data, seq.ctypes.data)
        if rc < 0:
            raise RuntimeError(f"oracle_simulate failed rc={rc}")
        return photon, seq

    def simulate_lite(self, gensteps, event_idx=0, photon_offset=0, p_begin=0, p_end=None, **cfgkw):
        """-> photon (n,4,4) f32, seq (n,2,2) u64, lite (n,4) u32 sphotonlite"""
        gs = np.ascontiguousarray(gensteps, np.float32).reshape(-1, 6, 4)
        total = int(gs.view(np.uint32)[:, 0, 3].sum())
        if p_end is None or p_end > total:
            p_end = total
        n = max(0, p_end - p_begin)
        photon = np.zeros((n, 4, 4), np.float32)
        seq = np.zeros((n, 2, 2), np.uint64)
        lite = np.zeros((n, 4), np.uint32)
        cfg = self.cfg(**cfgkw)
        rc = lib().oracle_simulate_ex(self.handle, C.byref(self.phys), C.byref(cfg), event_idx, gs.ctypes.data,
                                      len(gs), photon_offset, p_begin, p_end, photon.ctypes.data, seq.ctypes.data,
                                      lite.ctypes.data)
        if rc < 0:
            raise RuntimeError(f"oracle_simulate_ex failed rc={rc}")
        return photon, seq, lite

    def adjoint_rows(self, gensteps, photons, mat_block, sample_bin, n_bin, event_idx=0, photon_offset=0,
                     **cfgkw):
        """Score rows (n, n_param) f32 of the in-event photons listed in `photons`."""
        gs = np.ascontiguousarray(gensteps, np.float32).reshape(-1, 6, 4)
        ph = np.ascontiguousarray(photons, np.int64)
        mb = np.ascontiguousarray(mat_block, np.int32)
        sb = np.ascontiguousarray(sample_bin, np.int32)
        n_param = int(max(mb.max(), 0) + 1) * 2 * int(n_bin)
        rows = np.zeros((len(ph), n_param), np.float32)
        cfg = self.cfg(**cfgkw)
        rc = lib().oracle_adjoint_rows(self.handle, C.byref(self.phys), C.byref(cfg), event_idx, gs.ctypes.data,
                                       len(gs), photon_offset, ph.ctypes.data, len(ph), mb.ctypes.data,
                                       len(mb) - 1, sb.ctypes.data, int(n_bin), n_param, rows.ctypes.data)
        if rc < 0:
            raise RuntimeError(f"oracle_adjoint_rows failed rc={rc}")
        return rows

    def simtrace(self, gensteps, traverse=1, **cfgkw):
        """(n,4,4) f32 simtrace records of OpticksGenstep_FRAME gensteps (oracle_simtrace)."""
        gs = np.ascontiguousarray(gensteps, np.float32).reshape(-1, 6, 4)
        total = int(gs.view(np.uint32)[:, 0, 3].sum())
        out = np.zeros((total, 4, 4), np.float32)
        if total == 0:
            return out
        cfg = self.cfg(traverse=traverse, **cfgkw)
        rc = lib().oracle_simtrace(self.handle, C.byref(cfg), gs.ctypes.data, len(gs), out.ctypes.data)
        if rc != total:
            raise RuntimeError(f"oracle_simtrace failed rc={rc}")
        return out

    def trace(self, rays, traverse=1):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 2, 4)
        out = np.zeros((len(rays), 2, 4), np.float32)
        rc = lib().oracle_trace(self.handle, rays.ctypes.data, len(rays), out.ctypes.data, traverse)
        if rc != 0:
            raise RuntimeError("oracle_trace failed")
        return out


def intersect_prim(geometry, prim_idx, t_min, origin, direction):
    g = {k: np.ascontiguousarray(v) for k, v in geometry.items()}
    geom = _geom(g)
    o = np.ascontiguousarray(origin, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    out = np.zeros(4, np.float32)
    rc = lib().oracle_intersect_prim(C.byref(geom), prim_idx, t_min, o.ctypes.data, d.ctypes.data,
                                     out.ctypes.data)
    if rc < 0:
        raise RuntimeError("oracle_intersect_prim failed")
    return bool(rc), out


def tree_counts(reset=True):
    """(CSG tree evaluations, loop decisions) of the checker since the last reset; the
    first reset switches counting on."""
    e, l = C.c_int64(0), C.c_int64(0)
    lib().oracle_tree_counts(C.byref(e), C.byref(l), 1 if reset else 0)
    return e.value, l.value


def packed_lut():
    a = np.zeros(3, np.uint32)
    b = np.zeros(3, np.uint32)
    lib().oracle_packed_lut(a.ctypes.data, b.ctypes.data)
    return a, b


def uniforms(seed, photon, n, offset=0, skip=100000, event_idx=0):
    out = np.zeros(n, np.float32)
    lib().oracle_uniforms(seed, offset, skip, event_idx, photon, n, out.ctypes.data)
    return out


def philox_block(ctr, k0, k1):
    c = np.asarray(ctr, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox_block(c.ctypes.data, k0, k1, out.ctypes.data)
    return out


def logf(x):
    return lib().oracle_logf(float(x))


def atan2f_n(y, x):
    y = np.ascontiguousarray(y, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(y)
    lib().oracle_atan2f_n(y.ctypes.data, x.ctypes.data, out.ctypes.data, len(y))
    return out


def logf_n(x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().oracle_logf_n(x.ctypes.data, y.ctypes.data, len(x))
    return y


def sincosf_n(x):
    x = np.ascontiguousarray(x, np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().oracle_sincosf_n(x.ctypes.data, s.ctypes.data, c.ctypes.data, len(x))
    return s, c


def sincosf(x):
    s, c = C.c_float(), C.c_float()
    lib().oracle_sincosf(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def step_boundary(mom, pol, normal, n1, n2, u_reflect):
    mom = np.ascontiguousarray(mom, np.float32)
    pol = np.ascontiguousarray(pol, np.float32)
    nrm = np.ascontiguousarray(normal, np.float32)
    mo = np.zeros(3, np.float32)
    po = np.zeros(3, np.float32)
    fl = C.c_uint32()
    tc = C.c_float()
    lib().oracle_step_boundary(mom.ctypes.data, pol.ctypes.data, nrm.ctypes.data, n1, n2, u_reflect,
                               mo.ctypes.data, po.ctypes.data, C.byref(fl), C.byref(tc))
    return mo, po, fl.value, tc.value


def rayleigh(mom, pol, seed, photon, n):
    mom = np.ascontiguousarray(mom, np.float32)
    pol = np.ascontiguousarray(pol, np.float32)
    mo = np.zeros((n, 3), np.float32)
    po = np.zeros((n, 3), np.float32)
    lib().oracle_rayleigh(mom.ctypes.data, pol.ctypes.data, seed, photon, n, mo.ctypes.data, po.ctypes.data)
    return mo, po


def ref_intersect_prim(geometry, prim_idx, t_min, origin, direction):
    """Reference CSG code on the same prim; None when _ref was not built."""
    L = reflib()
    if L is None:
        return None
    node = np.ascontiguousarray(geometry["node"], np.float32)
    itra = np.ascontiguousarray(geometry["itra"], np.float32)
    prim = np.ascontiguousarray(geometry["prim"], np.float32)
    off = int(prim.view(np.int32)[prim_idx, 0, 1])
    plan = geometry.get("plan")
    plan = np.ascontiguousarray(plan, np.float32) if plan is not None and len(plan) else np.zeros((1, 4), np.float32)
    o = np.ascontiguousarray(origin, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    out = np.zeros(4, np.float32)
    rc = L.ref_intersect_prim(node[off:].ctypes.data, itra.ctypes.data if len(itra) else plan.ctypes.data,
                              plan.ctypes.data, t_min, o.ctypes.data, d.ctypes.data, out.ctypes.data)
    return bool(rc), out
