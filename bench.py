"""
Benchmark: photons propagated per second on one MI355X.

Workload (default): the CSG boolean geometry at 10M photons -- 1,000 torch
gensteps x 10,000 photons in a 5 m liquid-scintillator sphere, 120 mm acrylic
shell, water, 2,000 instanced detector solids (difference(union(z-cut
ellipsoid, cylinder), cylinder)) on a Fibonacci lattice at 5.5 m; wavelengths
uniform in 200-800 nm, FP32 physics, integer seqhis.

One step = one simulate call with gensteps already in HBM: genstep prefix
(seeding) -> propagation kernel -> stable hit compaction, outputs photon
(N,4,4), seq (N,2,2) and hits into preallocated device buffers. The 10M legs
run with 64-thread blocks, trace regroup 25% and refill batches of 32 lanes
(--block, --regroup, --batch); none of these changes a result.

Also reported: the dominant kernel's roofline (HIP events on the kernel's
stream), the CPU checker's rate on a bounded sample on this host, the
adjoint/forward device-time ratio on the same geometry with 256 binned
LS/water absorption and Rayleigh lengths (10M photons, 2,000 detectors), and
the batched 64-event workload (autotuned, replayed from a captured hipGraph),
the 256-event forward + adjoint graph, and the rows around the core path (hit merge, sphotonlite output, Cerenkov
generation, simtrace), each timed on the GPU beside its CPU counterpart.

--copy-calibration-mib M appends one device copy of M MiB after the timed
region: a kernel of known traffic for scaling FETCH_SIZE / WRITE_SIZE counters.
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
# VALU issue peak: 256 CUs x 2.4 GHz x one wave64 VALU instruction per CU per
# clock (4 SIMDs, 16 lanes, 4 cycles per wave64 FP32 instruction)
VALU_PEAK_GINST = 256 * 2.4


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


def cpu_model():
    """model name and logical CPU count of this host"""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for l in f:
                if l.startswith("model name"):
                    name = l.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name}, {os.cpu_count()} logical CPUs visible, {len(os.sched_getaffinity(0))} usable by this process"


def profile_json(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def bench_adjoint(torch, dev, steps, regroup=0, batch=1):
    """Forward then path-replay adjoint on the binned-length workload: config1
    geometry and photons, LS and water absorption/Rayleigh lengths as 4 x 64
    bins (x(1+U(-0.2,0.2)), seed 42); gradient (2000 detectors, 256) f32."""
    from opticks_amd import configs
    from opticks_amd.adjoint import yield_jacobian
    from opticks_amd.qsim import QSim, SEventConfig
    w = configs.config3()
    params = w.meta["params"]
    N = w.num_photon
    n_det = w.meta["num_det"]
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, trace_regroup_pct=regroup, refill_batch=batch))
    gs = torch.from_numpy(w.gensteps).to(dev)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    n_hit = sim.simulate_device(gs, photon, None, hit, n_photon=N)
    G = yield_jacobian(sim, params, gs, hit, n_hit, n_det, n_photon=N)
    fwd, adj, rep, red = [], [], [], []
    for _ in range(steps):
        n_hit = sim.simulate_device(gs, photon, None, hit, n_photon=N)
        fwd.append(sim.last_timing()["total_ms"])
        sim.adjoint_device(gs, hit, n_hit, params.c_struct(n_det), G, n_photon=N)
        t = sim.last_adjoint_timing()
        adj.append(t["total_ms"]); rep.append(t["replay_ms"]); red.append(t["reduce_ms"])
    torch.cuda.synchronize()
    sim.close()
    f, a = float(np.mean(fwd)), float(np.mean(adj))
    return {"workload": "csg-boolean-10M with LS/water absorption+Rayleigh lengths as 4x64 bins "
                        "(256 parameters), 2000 detectors",
            "trace_regroup_pct": regroup, "refill_batch": batch, "n_photon": int(N), "hits": int(n_hit), "n_param": int(params.n_param), "n_det": int(n_det),
            "forward_ms": f, "adjoint_ms": a, "replay_ms": float(np.mean(rep)), "reduce_ms": float(np.mean(red)),
            "adjoint_forward_ratio": a / f}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel", default="wavefront", choices=["wavefront", "plain"])
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--refill", type=int, default=1)
    ap.add_argument("--leaf", type=int, default=1)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--regroup", type=int, default=25,
                    help="trace_regroup_pct of the 10M forward/adjoint legs (best of 0-60 there)")
    ap.add_argument("--batch", type=int, default=32,
                    help="refill_batch of the 10M forward/adjoint legs (best of 1-64 there)")
    ap.add_argument("--cpu-sample", type=int, default=0, help="photons for the CPU baseline (0 = all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-event batched workload")
    ap.add_argument("--no-full", action="store_true", help="skip the 256-event forward+adjoint graph")
    ap.add_argument("--no-next", action="store_true", help="skip the hit merge / lite / Cerenkov / simtrace rows")
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
                       bvh_leaf_size=args.leaf, waves_per_cu=args.waves_per_cu, trace_regroup_pct=args.regroup,
                       refill_batch=args.batch)
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
    kname = "k_sim_wave" if kernel == L.MI_KERNEL_WAVEFRONT else "k_sim_plain"
    # HBM bytes and VALU instructions per launch of this kernel on this workload,
    # from rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE scaled by a copy kernel;
    # SQ_INSTS_VALU, SQ_THREAD_CYCLES_VALU, SQ_ACTIVE_INST_VALU)
    traffic, valu = None, None
    pj = profile_json("pmc_k_sim_wave_csg10M.json")
    if pj and pj.get("n_photon") == N and str(pj.get("kernel", "")).split("<")[0] == kname:
        traffic = pj.get("hbm_bytes_per_launch")
        if pj.get("valu_insts_per_launch"):
            rate = pj["valu_insts_per_launch"] / (kernel_ms * 1e-3) / 1e9
            valu = {"achieved_ginst_s": rate, "peak_ginst_s": VALU_PEAK_GINST, "frac": rate / VALU_PEAK_GINST,
                    "insts_per_launch": pj["valu_insts_per_launch"],
                    "lane_utilisation": pj.get("valu_lane_utilisation")}

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
                   "refill_period": args.refill, "bvh_leaf": args.leaf, "trace_regroup_pct": args.regroup,
                   "refill_batch": args.batch,
                   "outputs": "photon+seq+hit", "steps_per_photon": S, "hits": int(n_hit)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "kernel_ms": kernel_ms, "alg_bytes_per_launch": B, "valu": valu},
        "cpu_baseline": None,
    }
    if args.copy_calibration_mib > 0:
        src = torch.ones(args.copy_calibration_mib * 1024 * 256, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if not args.no_adjoint:
        a = bench_adjoint(torch, dev, max(1, min(args.steps, 3)), regroup=args.regroup, batch=args.batch)
        line["adjoint"] = a
        line["adjoint_forward_ratio"] = a["adjoint_forward_ratio"]
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sample = args.cpu_sample if args.cpu_sample > 0 else N
        rate, cdt = cpu_baseline(w, sample, threads)
        sample1 = min(sample, 1_000_000)
        rate1, cdt1 = cpu_baseline(w, sample1, 1)
        line["cpu_baseline"] = {"value": rate, "unit": "photons/s", "cores": threads, "kind": "port",
                                "sample": f"first {sample} photons of the same workload "
                                          f"({cdt:.1f} s wall on {threads} threads; C restatement of "
                                          f"qsim/CSG, OpenMP over photons, own median BVH)",
                                "value_1_thread": rate1,
                                "sample_1_thread": f"first {sample1} photons, {cdt1:.1f} s on 1 thread",
                                "host_cpu": cpu_model()}
    if not args.no_next:
        line["next_rows"] = bench_next_rows(torch, dev, sim, gs, photon, max(1, min(args.steps, 3)), not args.no_cpu)
    if not args.no_batched:
        line["batched"] = bench_batched(torch, dev, max(1, min(args.steps, 3)))
    if not args.no_full:
        line["full_batched_adjoint"] = bench_full(torch, dev, max(1, min(args.steps, 2)))
    print(json.dumps(line), flush=True)
    sim.close()


def _timed(torch, fn, steps):
    """mean wall seconds of fn() bracketed by device synchronizations (fn is synchronous)"""
    fn()
    ts = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return float(np.mean(ts)), r


def bench_next_rows(torch, dev, sim, gs, photon, steps, cpu):
    """Hit merge (SPM) of the forward run's 10M sphoton records, the sphotonlite
    output form of the same forward, Cerenkov generation + propagation (10M photons
    in a water sphere), and simtrace of a 201 x 201 XZ grid (10.1M rays) over the
    2,000-detector geometry. GPU times are wall clock around synchronous calls
    (merge, simtrace) or HIP events (propagation); CPU counterparts run the
    checker's restatements on this host."""
    from opticks_amd import configs
    from opticks_amd.qsim import QSim, SEventConfig, flag_mask
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {}
    N = photon.shape[0]
    sim.simulate_device(gs, photon, None, None, n_photon=N)
    mask = flag_mask("SD")
    t, merged = _timed(torch, lambda: sim.merge_hits_device(photon, select_mask=mask, time_window=1.0), steps)
    n_sel = int(((photon[:, 3, 3].view(torch.int32) & mask) != 0).sum().item())
    row = {"workload": "SPM merge_partial_select of the 10M sphoton records of the forward run: select SD, "
                       "key (identity, time / 1 ns), stable sort, run reduce",
           "records": int(N), "selected": n_sel, "merged": int(merged.shape[0]), "ms": 1e3 * t,
           "records_per_s": N / t, "selected_per_s": n_sel / t,
           "input_gb_s": N * 64 / t / 1e9, "input_hbm_frac": N * 64 / t / 1e9 / HBM_PEAK_GBS}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import spm
        host = photon.cpu().numpy()
        t0 = time.perf_counter()
        ref = spm.merge_photons(host, mask, 1.0)
        tc = time.perf_counter() - t0
        row["cpu_baseline"] = {"value": N / tc, "unit": "records/s", "cores": 1, "kind": "port",
                               "sample": f"all {N} records, numpy restatement of SPM (oracle/spm.py), {tc:.2f} s",
                               "merged": int(len(ref))}
    out["hit_merge"] = row

    lite = torch.empty((N, 4), dtype=torch.int32, device=dev)
    full_ms, lite_ms = [], []
    for _ in range(steps + 1):
        sim.simulate_device(gs, photon, None, None, n_photon=N)
        full_ms.append(sim.last_timing()["total_ms"])
        sim.simulate_lite_device(gs, lite, n_photon=N)
        lite_ms.append(sim.last_timing()["total_ms"])
    lm, fm = float(np.mean(lite_ms[1:])), float(np.mean(full_ms[1:]))
    out["photonlite"] = {"workload": "the 10M forward with only sphotonlite (16 B) records out",
                         "ms": lm, "photons_per_s": N / (lm * 1e-3),
                         "photon_output_ms": fm, "photon_output_photons_per_s": N / (fm * 1e-3)}
    del lite

    wc = configs.cerenkov_water(num_photon=10_000_000, n_gs=1000)
    sc = QSim(wc.geometry, wc.physics, SEventConfig(seed=wc.seed))
    gsc = torch.from_numpy(wc.gensteps).to(dev)
    Nc = wc.num_photon
    pc = torch.empty((Nc, 4, 4), dtype=torch.float32, device=dev)
    ms = []
    for _ in range(steps + 1):
        sc.simulate_device(gsc, pc, None, None, n_photon=Nc)
        ms.append(sc.last_timing()["total_ms"])
    cm = float(np.mean(ms[1:]))
    row = {"workload": "Cerenkov gensteps (1000 x 10000 photons, 10 mm steps at beta 0.99 -> 0.985, 200-600 nm) "
                       "in the water sphere of the 1M-photon geometry (vacuum box with a detecting inner "
                       "surface): generation + propagation, photon records out",
           "n_photon": int(Nc), "ms": cm, "photons_per_s": Nc / (cm * 1e-3)}
    if cpu:
        import pyoracle
        W = pyoracle.World(wc.geometry, wc.physics)
        sample = Nc
        t0 = time.perf_counter()
        W.simulate(wc.gensteps, seed=wc.seed, p_end=sample, nthreads=threads)
        tc = time.perf_counter() - t0
        row["cpu_baseline"] = {"value": sample / tc, "unit": "photons/s", "cores": threads, "kind": "port",
                               "sample": f"first {sample} photons, {tc:.2f} s on {threads} threads"}
    out["cerenkov"] = row
    sc.close()
    del pc

    fg = configs.frame_gensteps((0.0, 0.0, 0.0, 6500.0), (-100, 100, 0, 0, -100, 100, 250, 1), gridscale=0.01)
    fgt = torch.from_numpy(fg).to(dev)
    n_ray = int(fg.view(np.uint32)[:, 0, 3].sum())
    rec = torch.empty((n_ray, 4, 4), dtype=torch.float32, device=dev)
    t, _ = _timed(torch, lambda: sim.simtrace_device(fgt, rec, n_photon=n_ray), steps)
    row = {"workload": "simtrace of a 201 x 201 XZ grid x 250 rays (65 mm pitch over +-6.5 m) through the "
                       "2000-detector geometry, sevent simtrace records out",
           "rays": n_ray, "ms": 1e3 * t, "rays_per_s": n_ray / t}
    if cpu:
        import pyoracle
        W = pyoracle.World(*_config1_arrays())
        sample = fg
        ns = int(sample.view(np.uint32)[:, 0, 3].sum())
        t0 = time.perf_counter()
        W.simtrace(sample, nthreads=threads, seed=42)
        tc = time.perf_counter() - t0
        row["cpu_baseline"] = {"value": ns / tc, "unit": "rays/s", "cores": threads, "kind": "port",
                               "sample": f"first {ns} rays, {tc:.2f} s on {threads} threads"}
    out["simtrace"] = row
    return out


def _config1_arrays():
    from opticks_amd import configs
    w = configs.config1(num_gs=1, per_gs=1)
    return w.geometry, w.physics


def bench_batched(torch, dev, steps, tune_events=8, check_events=4):
    """64 events x Poisson(200) scintillation gensteps x Poisson(5000) photons in
    the 20,000-detector geometry: autotune (block, refill batch, TLAS leaf,
    trace regroup percent) on
    the first events, capture the whole batch into a hipGraph, time replays.
    History agreement: seq of the first events against the plain kernel run
    per event (one launch per event, one lane per photon)."""
    from opticks_amd import configs
    from opticks_amd import _lib as L
    from opticks_amd.qsim import QSim, SEventConfig
    w = configs.config2()
    N = w.num_photon
    gs_all = np.concatenate(w.events)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    ev = list(range(len(w.events)))
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs = torch.from_numpy(gs_all).to(dev)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=dev)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    t0 = time.perf_counter()
    tune = sim.autotune(gs, offs[:tune_events + 1], ev[:tune_events], photon, seq, hit, reps=1)
    tune_s = time.perf_counter() - t0
    best = min(tune, key=lambda r: r["ms"])
    b = sim.batch(gs, offs, ev, photon, seq, hit)
    b.launch(); b.wait()
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch()
        ho, ms = b.wait()
        dev_ms.append(ms)
    dt = (time.perf_counter() - t0) / steps
    # agreement with the plain per-event form on the first events
    S = seq[: int(b.photon_offsets[check_events])].cpu().numpy()
    plain = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, kernel=L.MI_KERNEL_PLAIN))
    same = total = 0
    for e in range(check_events):
        a0, a1 = int(b.photon_offsets[e]), int(b.photon_offsets[e + 1])
        ge = torch.from_numpy(w.events[e]).to(dev)
        sq = torch.empty((a1 - a0, 2, 2), dtype=torch.int64, device=dev)
        ph = torch.empty((a1 - a0, 4, 4), dtype=torch.float32, device=dev)
        plain.simulate_device(ge, ph, sq, None, eventID=e, n_photon=a1 - a0)
        same += int(np.all(sq.cpu().numpy() == S[a0:a1], axis=(1, 2)).sum())
        total += a1 - a0
    plain.close()
    b.close()
    sim.close()
    mean_ms = float(np.mean(dev_ms))
    return {"workload": "64 events (Poisson(200) scintillation gensteps x Poisson(5000) photons), 20000 CSG "
                        "detectors at 19.4 m, LS r 17.7 m; one captured hipGraph replay per step",
            "n_photon": int(N), "n_event": len(ev), "photons_per_s": N / (mean_ms * 1e-3),
            "photons_per_s_wall": N / dt, "device_ms_per_step": mean_ms, "wall_ms_per_step": dt * 1e3,
            "hits": int(ho[-1]),
            "autotune": {"events": tune_events, "seconds": tune_s, "candidates": len(tune),
                         "chosen": {k: best[k] for k in ("block_size", "refill_period", "bvh_leaf_size",
                                                         "trace_regroup_pct", "refill_batch")},
                         "chosen_ms": best["ms"], "slowest_ms": max(r["ms"] for r in tune)},
            "seqhis_agreement_vs_plain_per_event": same / max(total, 1), "agreement_photons": total}


def bench_full(torch, dev, steps):
    """256 events (Poisson(200) scintillation gensteps x Poisson(5000) photons, about
    256M photons) in the 20,000-detector geometry with the 256 binned lengths:
    forward hits and the yield adjoint of every hit in one captured graph."""
    from opticks_amd import configs
    from opticks_amd.qsim import QSim, SEventConfig
    w = configs.config4()
    params = w.meta["params"]
    n_det = w.meta["num_det"]
    N = w.num_photon
    gs_all = np.concatenate(w.events)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    ev = list(range(len(w.events)))
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, block_size=64))
    gs = torch.from_numpy(gs_all).to(dev)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=dev)
    cap = int(0.2 * N)
    hit = torch.empty((cap, 4, 4), dtype=torch.float32, device=dev)
    G = torch.empty((n_det, params.n_param), dtype=torch.float32, device=dev)
    b = sim.batch(gs, offs, ev, photon, seq, hit, adjoint=params.c_struct(n_det), grad_t=G)
    b.launch(); b.wait()
    fwd, adj, tot = [], [], []
    for _ in range(steps):
        b.launch()
        ho, ms = b.wait()
        t = b.last_timing()
        fwd.append(t["forward_ms"]); adj.append(t["adjoint_ms"]); tot.append(ms)
    steps_per_photon = mean_steps(seq[: min(N, 20_000_000)].cpu().numpy().view(np.uint64))
    b.close()
    sim.close()
    f, a, T = float(np.mean(fwd)), float(np.mean(adj)), float(np.mean(tot))
    B = algorithmic_bytes(N, len(gs_all), steps_per_photon, 0.0)
    return {"workload": "256 events (Poisson(200) scintillation gensteps x Poisson(5000) photons), 20000 CSG "
                        "detectors, LS/water lengths as 4x64 bins; forward + adjoint in one captured graph",
            "n_photon": int(N), "n_event": len(ev), "hits": int(ho[-1]), "n_param": int(params.n_param),
            "n_det": int(n_det), "forward_ms": f, "adjoint_ms": a, "graph_ms": T,
            "photons_per_s_forward": N / (f * 1e-3), "photons_per_s_forward_plus_adjoint": N / (T * 1e-3),
            "adjoint_forward_ratio": a / f, "steps_per_photon": steps_per_photon,
            "forward_alg_hbm_gb_s": B / (f * 1e-3) / 1e9}


if __name__ == "__main__":
    main()
