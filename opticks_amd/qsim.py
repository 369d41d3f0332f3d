"""
Host-side mirror of the reference's simulate surface, over the C-ABI.

  SEventConfig   run configuration with the reference's names and defaults
                 (sysrap/SEventConfig.cc:41-110); env overrides are applied only
                 when `from_env()` is called (the reference reads them at static init)
  QSim           one handle per geometry: simulate(gs, eventID) -> hits, the
                 NP-in/NP-out service path of qudarap/QSim.cc:720-755; after a call
                 the event arrays (photon, seq, hit) stay on the object until
                 reset(eventID), like the SEvt fold kept by reset:false
                 (g4cx/G4CXOpticks.cc:491-511)
  QSim.simulate_device / simulate_batch
                 device-resident forms for torch callers (torch is only the
                 device-memory and stream plumbing here)
  QSim.simtrace / simtrace_device
                 simtrace of OpticksGenstep_FRAME grids (qudarap/QSim.cc:876-910)

Errors follow the C-ABI: RuntimeError with mi_last_error() text; zero gensteps
returns None like the reference's rc=1 "nothing to do" (qudarap/QEvt.cc:436).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields

import numpy as np

from . import _lib as L

# sysrap/OpticksPhoton.h:22-44
FLAG = dict(CK=0x1, SI=0x2, TO=0x4, AB=0x8, RE=0x10, SC=0x20, SD=0x40, SA=0x80, DR=0x100,
            SR=0x200, BR=0x400, BT=0x800, NA=0x1000, EC=0x2000, EX=0x4000, MI=0x8000)
ABBREV = {v.bit_length(): k for k, v in FLAG.items()}


def flag_mask(labels: str) -> int:
    """"TO,CK,SI" -> bitmask (OpticksPhoton::GetFlagMask)."""
    m = 0
    for t in labels.replace(" ", "").split(","):
        if t:
            m |= FLAG[t]
    return m


def seqhis_label(seqhis0: int, seqhis1: int = 0) -> str:
    """Nibble sequence -> "TO BT SD" (sysrap/sseq.h desc_seqhis)."""
    out = []
    for word in (int(seqhis0), int(seqhis1)):
        for k in range(16):
            n = (word >> (4 * k)) & 0xF
            if n == 0:
                continue
            out.append(ABBREV.get(n, f"?{n}"))
    return " ".join(out)


@dataclass
class SEventConfig:
    max_bounce: int = 31
    max_time: float = 1e27
    propagate_epsilon: float = 0.05
    propagate_epsilon0: float = 0.05
    propagate_epsilon0_mask: str = "TO,CK,SI,SC,RE"
    tmax: float = 1e6
    hit_mask: str = "SD"
    seed: int = 0
    rng_offset: int = 0
    event_skipahead: int = 100000
    rng: str = "philox"             # "philox" or "xorwow" (the reference's RNG_PHILOX / RNG_XORWOW, sysrap/srng.h)
    precooked: np.ndarray | None = None   # (ni, nk) uniforms; overrides rng
    output_mask: int = L.MI_OUT_PHOTON | L.MI_OUT_SEQ | L.MI_OUT_HIT
    max_slot: int = 0
    kernel: int = L.MI_KERNEL_WAVEFRONT
    block_size: int = 64
    refill_period: int = 1
    bvh_leaf_size: int = 1
    waves_per_cu: int = 0
    trace_regroup_pct: int = 0
    refill_batch: int = 16
    adjoint_tape: int = 0           # segment records per photon kept by forwards for a following adjoint (0..64)

    ENV = dict(max_bounce=("OPTICKS_MAX_BOUNCE", int), max_time=("OPTICKS_MAX_TIME", float),
               propagate_epsilon=("OPTICKS_PROPAGATE_EPSILON", float),
               propagate_epsilon0=("OPTICKS_PROPAGATE_EPSILON0", float),
               propagate_epsilon0_mask=("OPTICKS_PROPAGATE_EPSILON0_MASK", str),
               hit_mask=("OPTICKS_HIT_MASK", str), max_slot=("OPTICKS_MAX_SLOT", int),
               event_skipahead=("OPTICKS_EVENT_SKIPAHEAD", int))

    @classmethod
    def from_env(cls, **kw):
        c = cls(**kw)
        for name, (env, typ) in cls.ENV.items():
            if env in os.environ:
                setattr(c, name, typ(os.environ[env]))
        if "QRng__SEED_OFFSET" in os.environ:
            s, o = os.environ["QRng__SEED_OFFSET"].split(":")
            c.seed, c.rng_offset = int(s), int(o)
        return c

    def to_c(self) -> L.MiConfig:
        c = L.MiConfig()
        L.lib().mi_config_default(C.byref(c))
        c.max_bounce = self.max_bounce
        c.max_time = self.max_time
        c.tmin = self.propagate_epsilon
        c.tmin0 = self.propagate_epsilon0
        c.eps0_mask = flag_mask(self.propagate_epsilon0_mask)
        c.tmax = self.tmax
        c.hit_mask = flag_mask(self.hit_mask)
        c.seed = self.seed
        c.rng_offset = self.rng_offset
        c.event_skipahead = self.event_skipahead
        rng = {"philox": L.MI_RNG_PHILOX, "xorwow": L.MI_RNG_XORWOW}
        if self.rng not in rng:
            raise ValueError(f"rng must be one of {sorted(rng)}")
        c.rng_kind = rng[self.rng]
        if self.precooked is not None:
            self._pre = np.ascontiguousarray(self.precooked, np.float32)
            c.rng_kind = L.MI_RNG_PRECOOKED
            c.precooked = self._pre.ctypes.data
            c.precooked_ni, c.precooked_nk = self._pre.shape
        c.output_mask = self.output_mask
        c.max_slot = self.max_slot
        c.kernel = self.kernel
        c.block_size = self.block_size
        c.refill_period = self.refill_period
        c.bvh_leaf_size = self.bvh_leaf_size
        c.waves_per_cu = self.waves_per_cu
        c.trace_regroup_pct = self.trace_regroup_pct
        c.refill_batch = self.refill_batch
        c.adjoint_tape = self.adjoint_tape
        return c


def _ptr(a):
    return a.ctypes.data if a is not None and len(a) else None


class QSim:
    """One geometry + physics upload; events simulated against it."""

    def __init__(self, geometry: dict, physics: dict, config: SEventConfig | None = None):
        self.config = config or SEventConfig()
        g = {k: np.ascontiguousarray(v) for k, v in geometry.items()}
        self._g = g
        itra = g.get("itra")
        self._geo = L.MiGeometry(_ptr(g["node"]), len(g["node"]), _ptr(g["prim"]), len(g["prim"]),
                                 _ptr(g["solid"]), len(g["solid"]), _ptr(itra), 0 if itra is None else len(itra),
                                 _ptr(g.get("plan")), 0 if g.get("plan") is None else len(g["plan"]),
                                 _ptr(g["inst"]), len(g["inst"]))
        self._bnd = np.ascontiguousarray(physics["bnd"], np.float32)
        self._opt = np.ascontiguousarray(physics["optical"], np.int32)
        icdf = physics.get("icdf")
        self._icdf = None if icdf is None else np.ascontiguousarray(icdf, np.float32)
        self._phys = L.MiPhysics(self._bnd.ctypes.data, self._bnd.shape[0], self._bnd.shape[3],
                                 float(physics["wl_low"]), float(physics["wl_step"]), self._opt.ctypes.data,
                                 None if self._icdf is None else self._icdf.ctypes.data,
                                 0 if self._icdf is None else self._icdf.shape[0],
                                 0 if self._icdf is None else self._icdf.shape[1],
                                 int(physics.get("icdf_hd_factor", 0)))
        cfg = self.config.to_c()
        h = C.c_void_p()
        L.check(L.lib().mi_sim_create(C.byref(self._geo), C.byref(self._phys), C.byref(cfg), C.byref(h)),
                "mi_sim_create")
        self.handle = h
        self.photon = self.seq = self.hit = None
        self.device_ms = None

    @classmethod
    def load(cls, base, config: SEventConfig | None = None):
        """geometry and physics from a reference-layout folder (CSGFoundry::Load); see loader.py"""
        from .loader import load_csgfoundry
        geometry, physics = load_csgfoundry(base)
        return cls(geometry, physics, config)

    def close(self):
        if getattr(self, "handle", None):
            L.lib().mi_sim_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, config: SEventConfig):
        cfg = config.to_c()
        L.check(L.lib().mi_sim_set_config(self.handle, C.byref(cfg)), "mi_sim_set_config")
        self.config = config

    def get_config_c(self) -> L.MiConfig:
        """The handle's configuration as the C struct (precooked pointer null: the handle keeps its copy)."""
        cfg = L.MiConfig()
        L.check(L.lib().mi_sim_get_config(self.handle, C.byref(cfg)), "mi_sim_get_config")
        return cfg

    def set_config_c(self, cfg: L.MiConfig):
        L.check(L.lib().mi_sim_set_config(self.handle, C.byref(cfg)), "mi_sim_set_config")

    def simulate(self, gs: np.ndarray, eventID: int = 0):
        """gensteps (G,6,4) -> hits (M,4,4); photon/seq/hit kept on self."""
        gs = np.ascontiguousarray(gs, np.float32).reshape(-1, 6, 4)
        out = L.MiEventOut()
        rc = L.lib().mi_sim_simulate(self.handle, eventID, gs.ctypes.data if len(gs) else None, len(gs),
                                     C.byref(out))
        if rc == L.MI_NOTHING_TO_DO:
            self.photon = self.seq = self.hit = None
            return None
        L.check(rc, "mi_sim_simulate")
        n, m = out.n_photon, out.n_hit
        om = self.config.output_mask
        self.photon = (np.ctypeslib.as_array((C.c_float * (16 * n)).from_address(out.photon)).reshape(n, 4, 4).copy()
                       if out.photon else None)
        self.seq = (np.ctypeslib.as_array((C.c_uint64 * (4 * n)).from_address(out.seq)).reshape(n, 2, 2).copy()
                    if out.seq else None)
        self.hit = (np.ctypeslib.as_array((C.c_float * (16 * max(m, 1))).from_address(out.hit))
                    .reshape(-1, 4, 4)[:m].copy() if (out.hit and om & L.MI_OUT_HIT) else None)
        self.device_ms = out.device_ms
        return self.hit

    def reset(self, eventID: int = 0):
        L.check(L.lib().mi_sim_reset(self.handle, eventID), "mi_sim_reset")
        self.photon = self.seq = self.hit = None

    def last_timing(self):
        t = (C.c_double * 4)()
        L.check(L.lib().mi_sim_last_timing(self.handle, t), "mi_sim_last_timing")
        return dict(seed_ms=t[0], propagate_ms=t[1], compact_ms=t[2], total_ms=t[3])

    # ---------------------------------------------------------- device forms

    def simulate_device(self, gs_t, photon_t=None, seq_t=None, hit_t=None, eventID=0, n_photon=-1,
                        photon_offset=0, stream=None):
        """torch tensors on the GPU: gs (G,6,4) f32, photon (N,4,4) f32, seq (N,2,2) i64, hit (M,4,4)."""
        n_hit = C.c_int64(0)
        cap = 0 if hit_t is None else hit_t.shape[0]
        rc = L.lib().mi_sim_simulate_device(
            self.handle, eventID, gs_t.data_ptr(), gs_t.shape[0], n_photon, photon_offset,
            None if photon_t is None else photon_t.data_ptr(), None if seq_t is None else seq_t.data_ptr(),
            None if hit_t is None else hit_t.data_ptr(), cap, C.byref(n_hit),
            None if stream is None else C.c_void_p(stream.cuda_stream))
        if rc == L.MI_NOTHING_TO_DO:
            return 0
        L.check(rc, "mi_sim_simulate_device")
        return n_hit.value

    def simulate_lite_device(self, gs_t, lite_t, photon_t=None, seq_t=None, hit_t=None, eventID=0, n_photon=-1,
                             photon_offset=0, stream=None):
        """as simulate_device, also writing (N,4) sphotonlite records into lite_t (int32 or uint32 words)"""
        n_hit = C.c_int64(0)
        cap = 0 if hit_t is None else hit_t.shape[0]
        rc = L.lib().mi_sim_simulate_lite_device(
            self.handle, eventID, gs_t.data_ptr(), gs_t.shape[0], n_photon, photon_offset, lite_t.data_ptr(),
            None if photon_t is None else photon_t.data_ptr(), None if seq_t is None else seq_t.data_ptr(),
            None if hit_t is None else hit_t.data_ptr(), cap, C.byref(n_hit),
            None if stream is None else C.c_void_p(stream.cuda_stream))
        if rc == L.MI_NOTHING_TO_DO:
            return 0
        L.check(rc, "mi_sim_simulate_lite_device")
        return n_hit.value

    def simulate_batch(self, gs_t, gs_offsets, event_ids, photon_t, seq_t=None, hit_t=None, stream=None):
        """All events' gensteps concatenated in gs_t; returns (photon_offsets, hit_offsets)."""
        ne = len(event_ids)
        ev = np.ascontiguousarray(event_ids, np.int32)
        go = np.ascontiguousarray(gs_offsets, np.int64)
        po = np.zeros(ne + 1, np.int64)
        ho = np.zeros(ne + 1, np.int64)
        cap = 0 if hit_t is None else hit_t.shape[0]
        rc = L.lib().mi_sim_simulate_batch_device(
            self.handle, ne, ev.ctypes.data, go.ctypes.data, gs_t.data_ptr(), po.ctypes.data,
            photon_t.data_ptr(), None if seq_t is None else seq_t.data_ptr(),
            None if hit_t is None else hit_t.data_ptr(), cap, ho.ctypes.data,
            None if stream is None else C.c_void_p(stream.cuda_stream))
        if rc == L.MI_NOTHING_TO_DO:
            return po, ho
        L.check(rc, "mi_sim_simulate_batch_device")
        return po, ho

    def set_bnd(self, bnd):
        """Replace the boundary table values (same shape as at create)."""
        bnd = np.ascontiguousarray(bnd, np.float32)
        if bnd.shape != self._bnd.shape:
            raise ValueError(f"bnd shape {bnd.shape} != {self._bnd.shape}")
        L.check(L.lib().mi_sim_set_bnd(self.handle, bnd.ctypes.data), "mi_sim_set_bnd")
        self._bnd_current = bnd

    @property
    def base_bnd(self):
        """The boundary table given at create."""
        return self._bnd

    def adjoint_device(self, gs_t, hit_t, n_hit, adj, grad_t, eventID=0, n_photon=-1, photon_offset=0,
                       stream=None):
        """d E[hits per detector]/d theta into grad_t (n_det, n_param) f32 on the GPU.

        adj: MiAdjoint (see opticks_amd.adjoint.LengthParams.c_struct). The event must have
        been simulated with the same gensteps, eventID, n_photon, photon_offset and bnd.
        """
        rc = L.lib().mi_sim_adjoint_device(
            self.handle, eventID, gs_t.data_ptr(), gs_t.shape[0], n_photon, photon_offset,
            None if n_hit == 0 else hit_t.data_ptr(), int(n_hit), C.byref(adj), grad_t.data_ptr(),
            None if stream is None else C.c_void_p(stream.cuda_stream))
        L.check(rc, "mi_sim_adjoint_device")

    def last_adjoint_timing(self):
        t = (C.c_double * 3)()
        L.check(L.lib().mi_sim_last_adjoint_timing(self.handle, t), "mi_sim_last_adjoint_timing")
        return dict(replay_ms=t[0], reduce_ms=t[1], total_ms=t[2])

    def batch(self, gs_t, gs_offsets, event_ids, photon_t, seq_t=None, hit_t=None, adjoint=None, grad_t=None):
        """Captured batch (hipGraph) over fixed buffers; with adjoint (MiAdjoint) and grad_t
        (n_det, n_param) f32 the replay also runs the yield adjoint of all hits. See Batch."""
        return Batch(self, gs_t, gs_offsets, event_ids, photon_t, seq_t, hit_t, adjoint, grad_t)

    def autotune(self, gs_t, gs_offsets, event_ids, photon_t, seq_t=None, hit_t=None, reps=2):
        """Pick block size, refill batch, TLAS leaf size and trace regroup percent on these
        events; returns all candidates."""
        ne = len(event_ids)
        ev = np.ascontiguousarray(event_ids, np.int32)
        go = np.ascontiguousarray(gs_offsets, np.int64)
        res = (L.MiTuneResult * 128)()
        n = C.c_int32(0)
        cap = 0 if hit_t is None else hit_t.shape[0]
        L.check(L.lib().mi_sim_autotune(self.handle, ne, ev.ctypes.data, go.ctypes.data, gs_t.data_ptr(),
                                        photon_t.data_ptr(), None if seq_t is None else seq_t.data_ptr(),
                                        None if hit_t is None else hit_t.data_ptr(), cap, reps, res, 128,
                                        C.byref(n)), "mi_sim_autotune")
        cfg = L.MiConfig()
        L.check(L.lib().mi_sim_get_config(self.handle, C.byref(cfg)), "mi_sim_get_config")
        self.config.block_size, self.config.refill_period = cfg.block_size, cfg.refill_period
        self.config.bvh_leaf_size = cfg.bvh_leaf_size
        self.config.trace_regroup_pct = cfg.trace_regroup_pct
        self.config.refill_batch = cfg.refill_batch
        return [dict(block_size=r.block_size, refill_period=r.refill_period, bvh_leaf_size=r.bvh_leaf_size,
                     trace_regroup_pct=r.trace_regroup_pct, refill_batch=r.refill_batch, ms=r.ms)
                for r in res[:min(n.value, 128)]]

    def merge_hits_device(self, in_t, select_mask=None, time_window=1.0, out_t=None, stream=None):
        """SPM::merge_partial_select on the GPU. in_t: (n,4,4) f32 sphoton records or (n,4)
        sphotonlite records (int32/float32 words); select_mask None = the hit mask, or
        L.MI_MERGE_ALREADY_SELECTED; time_window in ns (0 = select only). Returns the
        merged records as a view of out_t (allocated like in_t when None)."""
        import torch
        record = L.MI_RECORD_PHOTON if in_t.dim() == 3 else L.MI_RECORD_PHOTONLITE
        mask = flag_mask(self.config.hit_mask) if select_mask is None else select_mask
        n = in_t.shape[0]
        if out_t is None:
            out_t = torch.empty_like(in_t)
        n_out = C.c_int64(0)
        rc = L.lib().mi_sim_merge_hits_device(self.handle, record, in_t.data_ptr() if n else None, n,
                                              C.c_uint32(mask & 0xFFFFFFFF), float(time_window),
                                              out_t.data_ptr() if out_t.shape[0] else None, out_t.shape[0],
                                              C.byref(n_out), None if stream is None else C.c_void_p(stream.cuda_stream))
        L.check(rc, "mi_sim_merge_hits_device")
        return out_t[:n_out.value]

    def trace_device(self, rays_t, isect_t):
        L.check(L.lib().mi_sim_trace_device(self.handle, rays_t.data_ptr(), rays_t.shape[0], isect_t.data_ptr(),
                                            None), "mi_sim_trace_device")

    def simtrace(self, gs: np.ndarray):
        """QSim::simtrace (qudarap/QSim.cc:876-910) over host gensteps: (n,4,4) f32 simtrace
        records, or None without gensteps."""
        gs = np.ascontiguousarray(gs, np.float32).reshape(-1, 6, 4)
        ptr = C.POINTER(C.c_float)()
        n = C.c_int64(0)
        ms = C.c_double(0.0)
        rc = L.lib().mi_sim_simtrace(self.handle, gs.ctypes.data, len(gs), C.byref(ptr), C.byref(n), C.byref(ms))
        if rc == L.MI_NOTHING_TO_DO:
            return None
        L.check(rc, "mi_sim_simtrace")
        self.simtrace_ms = ms.value
        return np.ctypeslib.as_array(ptr, shape=(n.value * 16,)).reshape(n.value, 4, 4).copy()

    def simtrace_device(self, gs_t, out_t=None, n_photon=-1):
        """Simtrace of OpticksGenstep_FRAME gensteps (CSGOptiX7.cu:563-604): (n_photon,4,4)
        f32 sevent::add_simtrace records, written into out_t (allocated when None)."""
        import torch
        if n_photon < 0:
            n_photon = int(gs_t[:, 0, 3].view(torch.int32).to(torch.int64).sum().item())
        if out_t is None:
            out_t = torch.empty((n_photon, 4, 4), dtype=torch.float32, device=gs_t.device)
        if n_photon == 0:
            return out_t
        L.check(L.lib().mi_sim_simtrace_device(self.handle, gs_t.data_ptr(), gs_t.shape[0], n_photon,
                                               out_t.data_ptr(), None), "mi_sim_simtrace_device")
        return out_t

    def refit_instances(self, inst):
        inst = np.ascontiguousarray(inst, np.float32).reshape(-1, 4, 4)
        L.check(L.lib().mi_sim_refit_instances(self.handle, inst.ctypes.data, len(inst)), "mi_sim_refit_instances")


class Batch:
    """Many events' launch sequence captured once into a hipGraph, replayed by launch().

    The tensors passed at creation are the batch's inputs/outputs for every
    replay; gensteps may change between replays as long as each genstep keeps
    its photon count.
    """

    def __init__(self, sim: QSim, gs_t, gs_offsets, event_ids, photon_t, seq_t=None, hit_t=None, adjoint=None,
                 grad_t=None):
        self.sim = sim
        self.n_event = len(event_ids)
        ev = np.ascontiguousarray(event_ids, np.int32)
        go = np.ascontiguousarray(gs_offsets, np.int64)
        self.photon_offsets = np.zeros(self.n_event + 1, np.int64)
        self._keep = (gs_t, photon_t, seq_t, hit_t, adjoint, grad_t)
        h = C.c_void_p()
        cap = 0 if hit_t is None else hit_t.shape[0]
        seq_p = None if seq_t is None else seq_t.data_ptr()
        hit_p = None if hit_t is None else hit_t.data_ptr()
        if adjoint is None:
            L.check(L.lib().mi_batch_create(sim.handle, self.n_event, ev.ctypes.data, go.ctypes.data, gs_t.data_ptr(),
                                            photon_t.data_ptr(), seq_p, hit_p, cap, self.photon_offsets.ctypes.data,
                                            C.byref(h)), "mi_batch_create")
        else:
            L.check(L.lib().mi_batch_create_adjoint(sim.handle, self.n_event, ev.ctypes.data, go.ctypes.data,
                                                    gs_t.data_ptr(), photon_t.data_ptr(), seq_p, hit_p, cap,
                                                    C.byref(adjoint), grad_t.data_ptr(),
                                                    self.photon_offsets.ctypes.data, C.byref(h)),
                    "mi_batch_create_adjoint")
        self.handle = h
        self.num_photon = int(L.lib().mi_batch_num_photon(h))

    def launch(self, stream=None):
        L.check(L.lib().mi_batch_launch(self.handle, None if stream is None else C.c_void_p(stream.cuda_stream)),
                "mi_batch_launch")

    def wait(self):
        """-> (hit_offsets (n_event+1,), device ms of the replay)"""
        ho = np.zeros(self.n_event + 1, np.int64)
        ms = C.c_double(0)
        L.check(L.lib().mi_batch_wait(self.handle, ho.ctypes.data, C.byref(ms)), "mi_batch_wait")
        return ho, ms.value

    def last_timing(self):
        """[forward ms, adjoint ms] of the last replay (adjoint 0 without one)"""
        t = (C.c_double * 2)()
        L.check(L.lib().mi_batch_last_timing(self.handle, t), "mi_batch_last_timing")
        return dict(forward_ms=t[0], adjoint_ms=t[1])

    def close(self):
        if getattr(self, "handle", None):
            L.lib().mi_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
