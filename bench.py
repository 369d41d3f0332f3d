"""
Benchmark: photons propagated per second on one MI355X.

Workload (default): the CSG boolean geometry at 10M photons -- 1,000 torch
gensteps x 10,000 photons in a 5 m liquid-scintillator sphere, 120 mm acrylic
shell, water, 2,000 instanced detector solids (difference(union(z-cut
ellipsoid, cylinder), cylinder)) on a Fibonacci lattice at 5.5 m; wavelengths
uniform in 200-800 nm, FP32 physics, integer seqhis.

One step = one simulate call with gensteps already in HBM: genstep prefix
(seeding) -> propagation kernel -> stable hit compaction, outputs photon
(N,4,4), seq (N,2,2) and hits into preallocated device buffers.

Also reported: the dominant kernel's roofline (HIP events on the kernel's
stream), the CPU checker's rate on a bounded sample on this host, and the
batched 64-event workload when --batched is given.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def algorithmic_bytes(n_photon, n_gs, steps_per_photon, hit_frac):
    """Compulsory HBM bytes of one step of the workload (per launch):
    gensteps 96 B each + prefix 8 B each; per photon: photon record 64 B written,
    seq 32 B written, hit bitmask 1/8 B; per propagation step: four lerped
    boundary-table float4 pairs (4 x 32 B) and one optical word (4 B).
    The hit compaction (a separate kernel) is not included."""
    per_photon = 64 + 32 + 0.125
    per_step = 4 * 32 + 4
    return n_gs * (96 + 8) + n_photon * (per_photon + steps_per_photon * per_step)


def mean_steps(seq):
    """propagation steps per photon = nibbles of seqhis minus the generation nibble"""
    s = seq.reshape(-1, 4)[:, :2].astype(np.uint64)
    cnt = np.zeros(len(s), np.int64)
    for w in range(2):
        x = s[:, w]
        for k in range(16):
            cnt += ((x >> np.uint64(4 * k)) & np.uint64(0xF)) != 0
    return float(cnt.mean() - 1.0)


def cpu_baseline(w, sample_photons, threads):
    """The CPU checker (C restatement, OpenMP over photons) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    W = pyoracle.World(w.geometry, w.physics)
    W.simulate(w.gensteps, seed=w.seed, p_begin=0, p_end=2000, nthreads=threads)   # warm
    t0 = time.perf_counter()
    W.simulate(w.gensteps, seed=w.seed, p_begin=0, p_end=sample_photons, nthreads=threads)
    dt = time.perf_counter() - t0
    return sample_photons / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel", default="wavefront", choices=["wavefront", "plain"])
    ap.add_argument("--block", type=int, default=128)
    ap.add_argument("--refill", type=int, default=1)
    ap.add_argument("--leaf", type=int, default=1)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched", action="store_true", help="also time the 64-event batched workload")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "photon ranges would shard per GPU with no exchange until the hit gather", file=sys.stderr)
        sys.exit(2)

    import torch
    from opticks_amd import configs
    from opticks_amd import _lib as L
    from opticks_amd.qsim import QSim, SEventConfig

    dev = torch.device("cuda:0")
    w = configs.config1()
    N = w.num_photon
    G = len(w.gensteps)
    kernel = L.MI_KERNEL_WAVEFRONT if args.kernel == "wavefront" else L.MI_KERNEL_PLAIN
    cfg = SEventConfig(seed=w.seed, kernel=kernel, block_size=args.block, refill_period=args.refill,
                       bvh_leaf_size=args.leaf, waves_per_cu=args.waves_per_cu)
    sim = QSim(w.geometry, w.physics, cfg)
    gs = torch.from_numpy(w.gensteps).to(dev)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=dev)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)

    n_hit = 0
    for _ in range(args.warmup):
        n_hit = sim.simulate_device(gs, photon, seq, hit, n_photon=N)
    torch.cuda.synchronize()
    prop_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n_hit = sim.simulate_device(gs, photon, seq, hit, n_photon=N)
        prop_ms.append(sim.last_timing()["propagate_ms"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    value = N * args.steps / dt

    S = mean_steps(seq.cpu().numpy().view(np.uint64))
    kernel_ms = float(np.mean(prop_ms))
    B = algorithmic_bytes(N, G, S, n_hit / N)
    achieved = B / (kernel_ms * 1e-3) / 1e9
    # VALU side: counted flops per propagation step of this workload (counting rule:
    # add/sub/mul 1, div/sqrt/log/sin/cos as one each), per photon:
    flops_per_step = 520.0
    valu_tflops = N * S * flops_per_step / (kernel_ms * 1e-3) / 1e12

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_config1.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("n_photon") == N and t.get("kernel") == args.kernel:
            traffic = t.get("hbm_bytes_per_launch")

    line = {
        "metric": "photons propagated/sec on 1 MI355X vs CPU qsim; adjoint/forward time ratio",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic gensteps built in code (numpy default_rng(42) placement; Philox photon streams seed 42)",
        "config": {"workload": "csg-boolean-10M: 1000 torch gensteps x 10000 photons, LS sphere r5m + acrylic "
                               "+ water + 2000 instanced difference(union(zsphere,cylinder),cylinder) detectors",
                   "n_photon": N, "n_genstep": G, "kernel": args.kernel, "block": args.block,
                   "refill_period": args.refill, "bvh_leaf": args.leaf,
                   "outputs": "photon+seq+hit", "steps_per_photon": S, "hits": int(n_hit)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_sim_wave" if kernel == L.MI_KERNEL_WAVEFRONT else "k_sim_plain",
                     "kernel_ms": kernel_ms, "alg_bytes_per_launch": B,
                     "valu": {"achieved_tflops": valu_tflops, "peak_tflops": FP32_PEAK_TFLOPS,
                              "frac": valu_tflops / FP32_PEAK_TFLOPS, "flops_per_step": flops_per_step}},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rate, cdt = cpu_baseline(w, args.cpu_sample, threads)
        line["cpu_baseline"] = {"value": rate, "unit": "photons/s", "cores": threads, "kind": "port",
                                "sample": f"first {args.cpu_sample} photons of the same workload "
                                          f"({cdt:.1f} s, C restatement of qsim/CSG, OpenMP)"}
    if args.batched:
        line["batched"] = bench_batched(sim, torch, dev)
    print(json.dumps(line))
    sim.close()


def bench_batched(sim_unused, torch, dev):
    """64 events x Poisson(200) scintillation gensteps x Poisson(5000) photons in
    the 20,000-detector geometry, one persistent launch per step."""
    from opticks_amd import configs
    from opticks_amd.qsim import QSim, SEventConfig
    w = configs.config2()
    N = w.num_photon
    gs_all = np.concatenate(w.events)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs = torch.from_numpy(gs_all).to(dev)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=dev)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    ev = list(range(len(w.events)))
    sim.simulate_batch(gs, offs, ev, photon, seq, hit)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = 2
    for _ in range(steps):
        po_, ho = sim.simulate_batch(gs, offs, ev, photon, seq, hit)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    t = sim.last_timing()
    sim.close()
    return {"workload": "64 events, 20000 detectors at 19.4 m, LS r 17.7 m", "n_photon": int(N),
            "photons_per_s": N / dt, "ms_per_step": dt * 1e3, "propagate_ms": t["propagate_ms"],
            "hits": int(ho[-1])}


if __name__ == "__main__":
    main()
