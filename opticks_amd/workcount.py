"""
Work counts of the propagation kernel: how many traces, TLAS / BLAS node visits,
instance ray transforms, prim tests, CSG tree evaluations, leaf evaluations inside
trees and propagation steps one launch performs, per photon, plus the wave-level
lane occupancy and the cycle split of each part. Read from the OA_STATS counters
build (libopticks_amd_stats.so, `python -m opticks_amd.build --stats`), which
counts with device atomics and so runs slower than the product library; counts
are properties of the algorithm and the data (BVH, geometry, photon histories),
not of the kernel's speed. These are the work units of the algorithmic-flop
roofline in bench.py.

  OPTICKS_AMD_LIB=opticks_amd/libopticks_amd_stats.so \\
      python -m opticks_amd.workcount batched64 profiles/work_k_sim_wave_batched64.json
  OPTICKS_AMD_LIB=opticks_amd/libopticks_amd_stats.so \\
      python -m opticks_amd.workcount csg10M profiles/work_k_sim_wave_csg10M.json
  OPTICKS_AMD_LIB=opticks_amd/libopticks_amd_stats.so \\
      python -m opticks_amd.workcount full256 profiles/work_k_sim_wave_full256.json
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

NAMES = ["trace", "tlas", "inst", "blas", "prim", "tree", "leaf", "step", "iter", "active",
         "w_trace", "l_trace", "w_p1", "l_p1", "w_p2", "l_p2", "w_prim", "l_prim", "w_phys", "l_phys",
         "c_iter", "c_refill", "c_trace", "c_p1", "c_p2", "c_prim", "w_tree", "l_tree", "c_tree", "c_tleaf",
         "tlas_dead", "blas_dead", "tree_general"]


def read_counters(lib, reset=False):
    out = (C.c_ulonglong * len(NAMES))()
    lib.oa_stats_read.argtypes = [C.c_void_p, C.c_int]
    if lib.oa_stats_read(out, 1 if reset else 0) != 0:
        raise RuntimeError("oa_stats_read failed")
    return dict(zip(NAMES, [int(v) for v in out]))


def summarize(c, n_photon):
    per = {k: c[k] / n_photon for k in ("trace", "tlas", "inst", "blas", "prim", "tree", "leaf", "step",
                                         "tlas_dead", "blas_dead", "tree_general")}
    lanes = {}
    for part in ("trace", "p1", "p2", "prim", "phys", "tree"):
        w, l = c["w_" + part], c["l_" + part]
        lanes[part] = {"wave_executions": w, "lanes_per_execution": (l / w) if w else None}
    cyc = {k: c["c_" + k] for k in ("iter", "refill", "trace", "p1", "p2", "prim", "tree", "tleaf")}
    tot = max(1, cyc["iter"])
    return {"per_photon": per, "lanes": lanes, "cycles": cyc,
            "cycle_share": {k: v / tot for k, v in cyc.items() if k != "iter"}}


def run(which):
    import torch
    from . import _lib as L
    from . import configs
    from .qsim import QSim, SEventConfig
    lib = L.lib()
    if not hasattr(lib, "oa_stats_read"):
        raise RuntimeError("not a library with counters")
    dev = torch.device("cuda:0")
    if which in ("batched64", "full256"):
        w = configs.config2() if which == "batched64" else configs.config4()
        gs_np = np.concatenate(w.events).astype(np.float32)
        offs = np.cumsum([0] + [len(e) for e in w.events])
        ev = list(range(len(w.events)))
        N = w.num_photon
        cfg = SEventConfig(seed=w.seed)
        sim = QSim(w.geometry, w.physics, cfg)
        gs = torch.from_numpy(gs_np).to(dev)
        photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
        read_counters(lib, reset=True)
        sim.simulate_batch(gs, offs, ev, photon)
        torch.cuda.synchronize()
        desc = ("batched-64-events (20000 detectors at 19.4 m around a 17.7 m LS sphere)" if which == "batched64" else
                "256 events (20000 detectors, binned LS/water lengths)") + ", library launch defaults"
    elif which == "csg10M":
        w = configs.config1()
        N = w.num_photon
        cfg = SEventConfig(seed=w.seed, trace_regroup_pct=25, refill_batch=32)
        sim = QSim(w.geometry, w.physics, cfg)
        gs = torch.from_numpy(w.gensteps).to(dev)
        photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
        read_counters(lib, reset=True)
        sim.simulate_device(gs, photon, n_photon=N)
        torch.cuda.synchronize()
        desc = "csg-boolean-10M (1000 torch gensteps x 10000 photons, 2000 detectors), trace_regroup_pct 25, refill_batch 32"
    else:
        raise ValueError(which)
    c = read_counters(lib)
    sim.close()
    out = {"workload": which, "description": desc, "kernel": "k_sim_wave", "n_photon": int(N),
           "n_photon_sampled": int(N), "counters": c}
    out.update(summarize(c, N))
    return out


if __name__ == "__main__":
    which, path = sys.argv[1], sys.argv[2]
    r = run(which)
    r["command"] = ("OPTICKS_AMD_LIB=opticks_amd/libopticks_amd_stats.so python -m opticks_amd.workcount "
                    f"{which} {path}")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(r, f, indent=1)
    print(json.dumps(r["per_photon"]), flush=True)
