"""
Benchmark: photons propagated per second on one MI355X.

Headline workload: the batched form, 64 events x Poisson(200) scintillation
gensteps x Poisson(5000) photons (about 64M photons, uneven) in the
20,000-detector geometry (LS sphere r 17.7 m, 120 mm acrylic shell, water,
difference(union(z-cut ellipsoid, cylinder), cylinder) detectors on a Fibonacci
lattice at 19.4 m), FP32 physics, integer seqhis. Setup (untimed): the
autotuner picks block size, refill batch, TLAS leaf size and trace regroup on
the first events, and the whole batch (genstep prefix, seeding, persistent
propagation, stable hit compaction, per-event hit offsets) is captured into a
hipGraph. One step = one replay over all events with gensteps resident in HBM,
outputs photon (N,4,4), seq (N,2,2) and hits into preallocated buffers.

roofline: the propagation kernel (k_sim_wave), timed with HIP events on its
stream in an eager launch of the same batch; bound "valu": algorithmic FP32
flops per launch (work counts of profiles/work_*.json x per-unit costs, FLOPS
below) over the FP32 vector peak, with the HBM view (compulsory bytes, bytes with
the table/BVH/CSG/transform fetches, PMC traffic) and the measured VALU issue
rate and lane utilisation (profiles/pmc_*.json) beside it. cpu_baseline: the
CPU checker on whole events of the same workload for a bounded time.

Extra legs: the 10M-photon single event (2,000 detectors) with its own roofline and
CPU rate, the adjoint/forward ratio with 256 binned LS/water lengths (10M
photons, 2,000 detectors), the 256-event forward + adjoint captured graph, and
the rows around the core path (hit merge, sphotonlite output, Cerenkov
generation, simtrace), each timed beside its CPU counterpart.

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
VALU_PEAK_TFLOPS = 157.3          # MI355X FP32 vector, published
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


# Algorithmic FP32 work of one unit, counted from the kernel source with the rule
# add/sub/mul = 1, FMA = 2, compare/select/min/max = 0; divisions, square roots,
# reciprocals, log/sin/cos counted apart as transcendentals (T). Units come from the
# work counters (opticks_amd.workcount): per photon, traces, TLAS and BLAS node
# visits, instance tests, prim tests, CSG tree evaluations, leaf evaluations in
# trees, propagation steps.
FLOPS = {
    "node_visit": (48, 0),      # 4-wide node: 4 boxes x 6 slab FMAs (box_enter)
    "trace": (62, 5),           # ray setup (slab_ray: 3 mul, 3 rcp) + hit record: normal to world (15),
                                # object ray again (33), local hit point (6), lposcost (5, sqrt, div)
    "inst": (36, 3),            # world->object ray (33) + its slab setup (3 mul, 3 rcp)
    "prim_leaf": (30, 3),       # single-leaf prim (sphere/box): quadratic or slabs, roots, normal
    "tree": (15, 0),            # tree evaluation: classification dot products of 3 leaves
    "tree_leaf": (71, 4),       # leaf in a tree, averaged over the detector's scaled zsphere
                                # (frame transform 48 + 45) and two cylinders (60 each)
    "step": (185, 23),          # per propagation step: normalize + cosTheta (13, 2T), fill_state
                                # (4 lerped lookups, 56, 4T), to_boundary (9, 3T), at_boundary
                                # Fresnel (107, 14T) counted for every step (an upper bound)
    "photon": (55, 10),         # scintillation generation (mom, pol, wavelength, position, time)
}


def algorithmic_flops(work, n_photon):
    """(flops, transcendentals) per launch from the per-photon work units"""
    u = work["per_photon"]
    units = {"node_visit": u["tlas"] + u["blas"], "trace": u["trace"], "inst": u["inst"],
             "prim_leaf": u["prim"] - u["tree"], "tree": u["tree"], "tree_leaf": u["leaf"], "step": u["step"],
             "photon": 1.0}
    f = sum(FLOPS[k][0] * v for k, v in units.items()) * n_photon
    t = sum(FLOPS[k][1] * v for k, v in units.items()) * n_photon
    return f, t


def algorithmic_bytes_full(work, n_photon, n_gs):
    """The fetch model of the measurement statement: compulsory bytes plus, per step,
    fill_state's 4 lookups x 2 texels x 16 B and 3 optical int4; 128 B per 4-wide BVH
    node visited, 64 B per instance transform, 16 B per prim record and 64 B per CSG
    node of a tree evaluation (7 per detector tree) or single-leaf prim (mostly
    L2-resident: the geometry and tables are a few MB)."""
    u = work["per_photon"]
    per_photon = (64 + 32 + 0.125 + u["step"] * (4 * 2 * 16 + 3 * 16) + (u["tlas"] + u["blas"]) * 128
                  + u["inst"] * 64 + u["prim"] * 16 + (u["tree"] * 7 + (u["prim"] - u["tree"])) * 64)
    return n_gs * (96 + 8) + n_photon * per_photon


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


def bench_adjoint(torch, dev, steps, regroup=0, batch=1, tape=32):
    """Forward then yield adjoint on the binned-length workload: config1 geometry and
    photons, LS and water absorption/Rayleigh lengths as 4 x 64 bins (x(1+U(-0.2,0.2)),
    seed 42); gradient (2000 detectors, 256) f32. Two ways: the forward keeps each
    photon's segment records (adjoint_tape records per photon; 32 holds every photon at
    max_bounce 31) and the adjoint scores hits from them, replaying only photons with
    more segments; and the path replay of every hit from a plain forward. adjoint_forward_ratio = (taped forward - plain
    forward + adjoint) / plain forward: all the time the gradient adds."""
    from opticks_amd import configs
    from opticks_amd.adjoint import yield_jacobian
    from opticks_amd.qsim import QSim, SEventConfig
    w = configs.config3()
    params = w.meta["params"]
    N = w.num_photon
    n_det = w.meta["num_det"]
    base = dict(seed=w.seed, trace_regroup_pct=regroup, refill_batch=batch)
    sim = QSim(w.geometry, w.physics, SEventConfig(**base))
    gs = torch.from_numpy(w.gensteps).to(dev)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    n_hit = sim.simulate_device(gs, photon, None, hit, n_photon=N)
    G = yield_jacobian(sim, params, gs, hit, n_hit, n_det, n_photon=N)
    fwd, rep_adj = [], []
    for _ in range(steps):
        n_hit = sim.simulate_device(gs, photon, None, hit, n_photon=N)
        fwd.append(sim.last_timing()["total_ms"])
        sim.adjoint_device(gs, hit, n_hit, params.c_struct(n_det), G, n_photon=N)
        rep_adj.append(sim.last_adjoint_timing()["total_ms"])
    sim.set_config(SEventConfig(adjoint_tape=tape, **base))
    n_hit = sim.simulate_device(gs, photon, None, hit, n_photon=N)
    sim.adjoint_device(gs, hit, n_hit, params.c_struct(n_det), G, n_photon=N)
    fwd_t, adj, sco, red = [], [], [], []
    for _ in range(steps):
        n_hit = sim.simulate_device(gs, photon, None, hit, n_photon=N)
        fwd_t.append(sim.last_timing()["total_ms"])
        sim.adjoint_device(gs, hit, n_hit, params.c_struct(n_det), G, n_photon=N)
        t = sim.last_adjoint_timing()
        adj.append(t["total_ms"]); sco.append(t["replay_ms"]); red.append(t["reduce_ms"])
    torch.cuda.synchronize()
    sim.close()
    f, ft, a, ra = float(np.mean(fwd)), float(np.mean(fwd_t)), float(np.mean(adj)), float(np.mean(rep_adj))
    return {"workload": "csg-boolean-10M with LS/water absorption+Rayleigh lengths as 4x64 bins "
                        "(256 parameters), 2000 detectors",
            "trace_regroup_pct": regroup, "refill_batch": batch, "n_photon": int(N), "hits": int(n_hit),
            "n_param": int(params.n_param), "n_det": int(n_det), "adjoint_tape": tape,
            "forward_ms": f, "forward_taped_ms": ft, "adjoint_ms": a, "score_ms": float(np.mean(sco)),
            "reduce_ms": float(np.mean(red)), "adjoint_forward_ratio": (ft - f + a) / f,
            "replay_adjoint_ms": ra, "replay_adjoint_forward_ratio": ra / f}


def batched_workload():
    """The batched workload: 64 events x Poisson(200) scintillation gensteps x Poisson(5000) photons,
    20,000 detectors at 19.4 m around a 17.7 m LS sphere (opticks_amd.configs.config2)"""
    from opticks_amd import configs
    w = configs.config2()
    gs_all = np.concatenate(w.events).astype(np.float32)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    return w, gs_all, offs, list(range(len(w.events)))


def pmc_summary(name, n_photon, kname):
    """HBM bytes and VALU instructions per launch of the dominant kernel on this
    workload, from rocprofv3 --pmc runs kept in profiles/ (FETCH_SIZE, WRITE_SIZE
    scaled by a copy kernel of known size; SQ_INSTS_VALU, SQ_THREAD_CYCLES_VALU,
    SQ_ACTIVE_INST_VALU)"""
    pj = profile_json(name)
    if not pj or pj.get("n_photon") != n_photon or str(pj.get("kernel", "")).split("<")[0] != kname:
        return None
    return pj


def roofline(kernel_ms, n_photon, n_gs, steps, work, pmc):
    """The dominant kernel's rooflines. It is bound by VALU issue under divergence,
    so `bound` is "valu": achieved = algorithmic FP32 flops per launch / kernel time
    against the FP32 vector peak. Beside it the HBM view: the compulsory bytes and the
    bytes with the table, BVH node, CSG node and instance-transform reads per step
    (algorithmic fetches, mostly L2-resident), and the PMC-measured traffic."""
    t = kernel_ms * 1e-3
    out = {"bound": "valu", "kernel": "k_sim_wave", "kernel_ms": kernel_ms, "peak": VALU_PEAK_TFLOPS,
           "unit": "TFLOP/s", "traffic": None}
    if work is not None:
        F, T = algorithmic_flops(work, n_photon)
        out.update(achieved=F / t / 1e12, frac=F / t / 1e12 / VALU_PEAK_TFLOPS, alg_flops_per_launch=F,
                   alg_transcendentals_per_launch=T, work_units=work["per_photon"])
    else:
        out.update(achieved=None, frac=None)
    Bc = algorithmic_bytes(n_photon, n_gs, steps, 0.0)
    Bf = algorithmic_bytes_full(work, n_photon, n_gs) if work is not None else None
    out["hbm"] = {"peak_gb_s": HBM_PEAK_GBS, "compulsory_bytes_per_launch": Bc,
                  "compulsory_gb_s": Bc / t / 1e9, "compulsory_frac": Bc / t / 1e9 / HBM_PEAK_GBS,
                  "with_fetches_bytes_per_launch": Bf,
                  "with_fetches_frac": (Bf / t / 1e9 / HBM_PEAK_GBS) if Bf else None}
    if pmc:
        out["traffic"] = pmc.get("hbm_bytes_per_launch")
        if pmc.get("valu_insts_per_launch"):
            rate = pmc["valu_insts_per_launch"] / t / 1e9
            out["valu_issue"] = {"achieved_ginst_s": rate, "peak_ginst_s": VALU_PEAK_GINST,
                                 "frac": rate / VALU_PEAK_GINST, "insts_per_launch": pmc["valu_insts_per_launch"],
                                 "lane_utilisation": pmc.get("valu_lane_utilisation"),
                                 "insts_per_photon_step": pmc["valu_insts_per_launch"] / (n_photon * steps)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tune-events", type=int, default=8, help="events the autotuner times (untimed setup)")
    ap.add_argument("--regroup", type=int, default=25,
                    help="trace_regroup_pct of the 10M forward/adjoint legs (best of 0-60 there)")
    ap.add_argument("--batch", type=int, default=32,
                    help="refill_batch of the 10M forward/adjoint legs (best of 1-64 there)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline: events of the workload until this much wall time on all threads")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="headline only (no 10M / adjoint / 256-event / next rows)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-full", action="store_true", help="skip the 256-event forward+adjoint graph")
    ap.add_argument("--no-next", action="store_true", help="skip the hit merge / lite / Cerenkov / simtrace rows")
    ap.add_argument("--copy-calibration-mib", type=int, default=0)
    ap.add_argument("--pmc-pass", choices=["batched64", "csg10M", "full256"], default=None,
                    help="counter-collection pass: two eager launches of one workload's propagation with fixed "
                         "launch settings (no autotuner, no graph), then the copy calibration; no JSON line")
    ap.add_argument("--block", type=int, default=128, help="--pmc-pass: block size")
    ap.add_argument("--pmc-regroup", type=int, default=0, help="--pmc-pass: trace_regroup_pct")
    ap.add_argument("--pmc-batch", type=int, default=16, help="--pmc-pass: refill_batch")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "events would shard per GPU with no exchange until the hit gather", file=sys.stderr)
        sys.exit(2)
    if args.pmc_pass:
        pmc_pass(args)
        return

    import torch
    from opticks_amd.qsim import QSim, SEventConfig

    dev = torch.device("cuda:0")
    w, gs_all, offs, ev = batched_workload()
    N = w.num_photon
    G = len(gs_all)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs = torch.from_numpy(gs_all).to(dev)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=dev)
    hit = torch.empty((N // 4, 4, 4), dtype=torch.float32, device=dev)

    # setup (untimed): autotune the launch on the first events, capture the whole batch
    t0 = time.perf_counter()
    tune = sim.autotune(gs, offs[:args.tune_events + 1], ev[:args.tune_events], photon, seq, hit, reps=1)
    tune_s = time.perf_counter() - t0
    best = min(tune, key=lambda r: r["ms"])
    b = sim.batch(gs, offs, ev, photon, seq, hit)
    for _ in range(args.warmup):
        b.launch(); b.wait()
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
        ho, ms = b.wait()
        dev_ms.append(ms)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    value = N * args.steps / dt
    n_hit = int(ho[-1])
    b.close()

    # the propagation kernel alone, HIP events on its stream, same batch launched eagerly
    prop = []
    for _ in range(3):
        sim.simulate_batch(gs, offs, ev, photon, seq, hit)
        prop.append(sim.last_timing()["propagate_ms"])
    kernel_ms = float(np.mean(prop[1:]))
    S = mean_steps(seq[:20_000_000].cpu().numpy().view(np.uint64))
    work = profile_json("work_k_sim_wave_batched64.json")
    pmc = pmc_summary("pmc_k_sim_wave_batched64.json", N, "k_sim_wave")

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
        "data": "synthetic scintillation gensteps built in code (numpy default_rng([42, event]) counts and "
                "placement; Philox photon streams, seed 42, event skipahead)",
        "config": {"workload": "batched-64-events: 64 events x Poisson(200) scintillation gensteps x "
                               "Poisson(5000) photons, LS sphere r17.7m + acrylic + water + 20000 instanced "
                               "difference(union(zsphere,cylinder),cylinder) detectors at r19.4m; one step = "
                               "one replay of the captured hipGraph over all events",
                   "n_photon": int(N), "n_event": len(ev), "n_genstep": int(G), "kernel": "wavefront",
                   "outputs": "photon+seq+hit", "steps_per_photon": S, "hits": n_hit,
                   "device_ms_per_step": float(np.mean(dev_ms)),
                   "autotune": {"events": args.tune_events, "seconds": tune_s, "candidates": len(tune),
                                "chosen": {k: best[k] for k in ("block_size", "refill_period", "bvh_leaf_size",
                                                                "trace_regroup_pct", "refill_batch")},
                                "chosen_ms": best["ms"], "slowest_ms": max(r["ms"] for r in tune)}},
        "roofline": roofline(kernel_ms, N, G, S, work, pmc),
        "cpu_baseline": None,
    }
    if args.copy_calibration_mib > 0:
        src = torch.ones(args.copy_calibration_mib * 1024 * 256, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    sim.close()
    del photon, seq, hit
    torch.cuda.empty_cache()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_events(w, args.cpu_seconds)
    if not args.no_legs:
        c1 = bench_config1(torch, dev, max(1, min(args.steps, 5)), args.regroup, args.batch, not args.no_cpu)
        line["config1"] = c1
        if not args.no_adjoint:
            a = bench_adjoint(torch, dev, max(1, min(args.steps, 3)), regroup=args.regroup, batch=args.batch)
            line["adjoint"] = a
            line["adjoint_forward_ratio"] = a["adjoint_forward_ratio"]
        if not args.no_next:
            line["next_rows"] = c1.pop("_next_rows", None)
        if not args.no_full:
            line["full_batched_adjoint"] = bench_full(torch, dev, max(1, min(args.steps, 2)))
    print(json.dumps(line), flush=True)


def pmc_pass(args):
    """Two eager launches of one workload's propagation with the given launch settings
    and, with --copy-calibration-mib, one device copy of known size: the dispatches a
    rocprofv3 --pmc run reads (opticks_amd.pmc summarizes them)."""
    import torch
    from opticks_amd import configs
    from opticks_amd.qsim import QSim, SEventConfig
    dev = torch.device("cuda:0")
    cfg = dict(block_size=args.block, trace_regroup_pct=args.pmc_regroup, refill_batch=args.pmc_batch)
    if args.pmc_pass == "csg10M":
        w = configs.config1()
        sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, **cfg))
        gs = torch.from_numpy(w.gensteps).to(dev)
        N = w.num_photon
        photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
        seq = torch.empty((N, 2, 2), dtype=torch.int64, device=dev)
        hit = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
        for _ in range(2):
            sim.simulate_device(gs, photon, seq, hit, n_photon=N)
    else:
        w = configs.config2() if args.pmc_pass == "batched64" else configs.config4()
        gs_all = np.concatenate(w.events).astype(np.float32)
        offs = np.cumsum([0] + [len(e) for e in w.events])
        ev = list(range(len(w.events)))
        N = w.num_photon
        sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, **cfg))
        gs = torch.from_numpy(gs_all).to(dev)
        photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
        seq = torch.empty((N, 2, 2), dtype=torch.int64, device=dev)
        hit = torch.empty((N // 4, 4, 4), dtype=torch.float32, device=dev)
        for _ in range(2):
            sim.simulate_batch(gs, offs, ev, photon, seq, hit)
    torch.cuda.synchronize()
    print(f"pmc pass {args.pmc_pass}: {N} photons, settings {cfg}, propagate "
          f"{sim.last_timing()['propagate_ms']:.3f} ms", flush=True)
    if args.copy_calibration_mib > 0:
        src = torch.ones(args.copy_calibration_mib * 1024 * 256, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    sim.close()


def cpu_baseline_events(w, seconds):
    """The CPU checker (C restatement of qsim/CSG, OpenMP over photons, own BVH) on
    whole events of the batched workload, in order, until `seconds` of wall time on
    all threads; then a 1-thread rate on part of the first event."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    W = pyoracle.World(w.geometry, w.physics)
    W.simulate(w.events[0], event_idx=0, seed=w.seed, p_end=2000, nthreads=threads)    # warm
    done = photons = 0
    t = 0.0
    while done < len(w.events) and t < seconds:
        n = int(w.events[done].view(np.uint32)[:, 0, 3].sum())
        t0 = time.perf_counter()
        W.simulate(w.events[done], event_idx=done, seed=w.seed, nthreads=threads)
        t += time.perf_counter() - t0
        photons += n
        done += 1
    n1 = 200_000
    t0 = time.perf_counter()
    W.simulate(w.events[0], event_idx=0, seed=w.seed, p_end=n1, nthreads=1)
    t1 = time.perf_counter() - t0
    return {"value": photons / t, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"events 0..{done - 1} of the same 64-event workload ({photons} photons, {t:.1f} s wall on "
                      f"{threads} threads; C restatement of qsim/CSG, OpenMP over photons, own median BVH)",
            "value_1_thread": n1 / t1, "sample_1_thread": f"first {n1} photons of event 0, {t1:.1f} s on 1 thread",
            "host_cpu": cpu_model()}


def bench_config1(torch, dev, steps, regroup, batch, cpu):
    """The 10M-photon single event (1,000 torch gensteps x 10,000 photons,
    2,000 detectors), one simulate call per step with gensteps resident: seeding,
    propagation, stable hit compaction into preallocated buffers."""
    from opticks_amd import configs
    from opticks_amd.qsim import QSim, SEventConfig
    w = configs.config1()
    N = w.num_photon
    G = len(w.gensteps)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, trace_regroup_pct=regroup, refill_batch=batch))
    gs = torch.from_numpy(w.gensteps).to(dev)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=dev)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=dev)
    n_hit = sim.simulate_device(gs, photon, seq, hit, n_photon=N)
    torch.cuda.synchronize()
    prop_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        n_hit = sim.simulate_device(gs, photon, seq, hit, n_photon=N)
        prop_ms.append(sim.last_timing()["propagate_ms"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    S = mean_steps(seq.cpu().numpy().view(np.uint64))
    kernel_ms = float(np.mean(prop_ms))
    out = {"workload": "csg-boolean-10M: 1000 torch gensteps x 10000 photons, LS sphere r5m + acrylic + water + "
                       "2000 instanced difference(union(zsphere,cylinder),cylinder) detectors",
           "n_photon": int(N), "photons_per_s": N * steps / dt, "ms_per_step": 1e3 * dt / steps,
           "trace_regroup_pct": regroup, "refill_batch": batch, "steps_per_photon": S, "hits": int(n_hit),
           "roofline": roofline(kernel_ms, N, G, S, profile_json("work_k_sim_wave_csg10M.json"),
                                pmc_summary("pmc_k_sim_wave_csg10M.json", N, "k_sim_wave"))}
    if cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rate, cdt = cpu_baseline(w, N, threads)
        out["cpu_baseline"] = {"value": rate, "unit": "photons/s", "cores": threads, "kind": "port",
                               "sample": f"all {N} photons ({cdt:.1f} s wall on {threads} threads)"}
    out["_next_rows"] = bench_next_rows(torch, dev, sim, gs, photon, max(1, min(steps, 3)), cpu)
    sim.close()
    return out


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


def bench_full(torch, dev, steps, tape=32):
    """256 events (Poisson(200) scintillation gensteps x Poisson(5000) photons, about
    256M photons) in the 20,000-detector geometry with the 256 binned lengths:
    forward hits and the yield adjoint of every hit in one captured graph, the
    forward keeping `tape` segment records per photon for the adjoint. The plain
    forward (no records) is timed eagerly beside it for the ratio."""
    from opticks_amd import configs
    from opticks_amd.qsim import QSim, SEventConfig
    torch.cuda.empty_cache()          # the segment tape of 255M photons is 98 GB at 32 records
    w = configs.config4()
    params = w.meta["params"]
    n_det = w.meta["num_det"]
    N = w.num_photon
    gs_all = np.concatenate(w.events)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    ev = list(range(len(w.events)))
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, block_size=64, adjoint_tape=tape))
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
    # the plain forward (no segment records): one eager launch of the same batch, HIP
    # events; its propagation kernel is the roofline's
    sim.set_config(SEventConfig(seed=w.seed, block_size=64))
    sim.simulate_batch(gs, offs, ev, photon, seq, hit)
    kernel_ms = sim.last_timing()["propagate_ms"]
    plain_ms = sim.last_timing()["total_ms"]
    sim.close()
    f, a, T = float(np.mean(fwd)), float(np.mean(adj)), float(np.mean(tot))
    B = algorithmic_bytes(N, len(gs_all), steps_per_photon, 0.0)
    return {"workload": "256 events (Poisson(200) scintillation gensteps x Poisson(5000) photons), 20000 CSG "
                        "detectors, LS/water lengths as 4x64 bins; forward + adjoint in one captured graph",
            "n_photon": int(N), "n_event": len(ev), "hits": int(ho[-1]), "n_param": int(params.n_param),
            "n_det": int(n_det), "adjoint_tape": tape, "forward_taped_ms": f, "adjoint_ms": a, "graph_ms": T,
            "forward_plain_ms": plain_ms, "photons_per_s_forward": N / (plain_ms * 1e-3),
            "photons_per_s_forward_plus_adjoint": N / (T * 1e-3),
            "adjoint_forward_ratio": (T - plain_ms) / plain_ms, "steps_per_photon": steps_per_photon,
            "forward_alg_hbm_gb_s": B / (plain_ms * 1e-3) / 1e9,
            "roofline": roofline(kernel_ms, N, len(gs_all), steps_per_photon, profile_json("work_k_sim_wave_full256.json"),
                                 pmc_summary("pmc_k_sim_wave_full256.json", N, "k_sim_wave"))}


if __name__ == "__main__":
    main()
