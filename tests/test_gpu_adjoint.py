"""Path-replay adjoint on the MI355X (mi_sim_adjoint_device) against

* the oracle's score rows for the same hit photons, summed per detector in
  float64 (the replay reproduces the forward path bit for bit, so only the
  FP32 Kahan summation differs);
* the closed-form yield derivative of the water sphere without scattering;
* central finite differences, step 1%, with common random numbers (same seed,
  same photon streams) on the 2,000-detector CSG workload with the binned
  LS/water lengths, along several parameter directions summed over detectors,
  with the FD error bar taken from the photons whose outcome changed and the
  adjoint's from independent photon subsets;
* the torch.autograd.Function wrapper (backward = Jacobian contraction).
"""
import numpy as np
import pytest

import pyoracle as po
from opticks_amd import configs
from opticks_amd.adjoint import LengthParams, expected_hits, yield_jacobian
from opticks_amd.qsim import QSim, SEventConfig

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

DEV = torch.device("cuda:0")
N_DET = 2000


def run_forward(sim, gs_t, n_photon, event_id=0):
    photon = torch.empty((n_photon, 4, 4), dtype=torch.float32, device=DEV)
    hit = torch.empty((n_photon, 4, 4), dtype=torch.float32, device=DEV)
    n_hit = sim.simulate_device(gs_t, photon, None, hit, eventID=event_id, n_photon=n_photon)
    del photon
    return hit, n_hit


def hit_photon_index(hit, n_hit):
    return hit[:n_hit, 3, 2].contiguous().view(torch.int32).long()


def test_adjoint_matches_oracle_rows():
    w = configs.config3(num_gs=20, per_gs=1000)
    params = w.meta["params"]
    N = w.num_photon
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    hit, n_hit = run_forward(sim, gs_t, N)
    assert n_hit > 100
    idx_all = hit_photon_index(hit, n_hit).cpu().numpy()
    world = po.World(w.geometry, w.physics)
    ref_photon, _ = world.simulate(w.gensteps, seed=w.seed)
    ref_hit_idx = np.nonzero(ref_photon[:, 3, 3].view(np.uint32) & 0x40)[0]
    common = np.intersect1d(idx_all, ref_hit_idx)
    assert len(common) >= 0.999 * max(len(idx_all), len(ref_hit_idx))
    keep = torch.from_numpy(np.isin(idx_all, common)).to(DEV)
    hit = hit[:n_hit][keep].contiguous()
    n_hit = hit.shape[0]
    G = yield_jacobian(sim, params, gs_t, hit, n_hit, N_DET, n_photon=N).cpu().numpy().astype(np.float64)
    h = hit.cpu().numpy()
    idx = h[:, 3, 2].view(np.uint32).astype(np.int64)
    det = (h[:, 3, 1].view(np.uint32) & 0xffff).astype(np.int64) - 1
    rows = world.adjoint_rows(w.gensteps, idx, params.mat_block, params.sample_bin, params.n_bin,
                              seed=w.seed).astype(np.float64)
    G64 = np.zeros((N_DET, params.n_param))
    A64 = np.zeros((N_DET, params.n_param))
    np.add.at(G64, det, rows)
    np.add.at(A64, det, np.abs(rows))
    err = np.abs(G - G64)
    assert np.all(err <= 1e-5 * A64 + 1e-30), err.max()
    # chunks of 37 hits (score rows 37 x 256 floats) give the same sums up to rounding
    Gc = yield_jacobian(sim, params, gs_t, hit, n_hit, N_DET, n_photon=N, max_row_floats=37 * 256)
    errc = np.abs(Gc.cpu().numpy().astype(np.float64) - G64)
    assert np.all(errc <= 1e-5 * A64 + 1e-30), errc.max()
    assert np.count_nonzero(G) > 1000
    # launch tuning does not change the replay: batched refills and early node-walk hand-over
    sim.set_config(SEventConfig(seed=w.seed, refill_batch=32, trace_regroup_pct=25))
    Gt = yield_jacobian(sim, params, gs_t, hit, n_hit, N_DET, n_photon=N).cpu().numpy().astype(np.float64)
    assert np.array_equal(Gt, G)
    # the 256 parameters in order: LS absorption, LS Rayleigh, water absorption, water Rayleigh
    assert np.abs(G).sum(axis=0).reshape(2, 2, 64).sum(axis=2).min() > 0
    sim.close()


def test_adjoint_closed_form_water_sphere():
    La, R, n = 20000.0, 1000.0, 1.333
    w = configs.config0(num_photon=4_000_000, water_absorption=La, water_rayleigh=1e12, sensor=True)
    w.gensteps[0, 3, 3] = 443.25   # samples 443 (bin 25, weight 0.75) and 444 nm (bin 26, 0.25)
    params = LengthParams(w.physics, ("Water",), n_bin=64)
    N = w.num_photon
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    hit, n_hit = run_forward(sim, gs_t, N)
    G = yield_jacobian(sim, params, gs_t, hit, n_hit, 1, n_photon=N).cpu().numpy()[0] / N
    rf = ((n - 1) / (n + 1)) ** 2
    x = np.exp(-R / La)
    q = rf * x * x
    dP = (1 - rf) * x * (1 + q) / (1 - q) ** 2 * R / La ** 2
    assert abs(G[25] / (0.75 * dP) - 1) < 2e-3
    assert abs(G[26] / (0.25 * dP) - 1) < 2e-3
    sim.close()


FD_EVENTS = 40          # 40 events x 40M photons = 1.6G photons per FD evaluation
N_GROUP = 20            # detector groups of 100 for per-detector-group directions


@pytest.fixture(scope="module")
def fd_setup():
    w = configs.config3(num_gs=4000, per_gs=10_000)
    params = w.meta["params"]
    theta = w.meta["theta"].astype(np.float64)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    yield w, params, theta, sim, gs_t
    sim.close()


def hit_group(hit, n_hit):
    """detector group (sensor identifier - 1) // (N_DET / N_GROUP) of each hit"""
    det = (hit[:n_hit, 3, 1].contiguous().view(torch.int32) & 0xffff).long() - 1
    return torch.div(det, N_DET // N_GROUP, rounding_mode="floor")


def test_adjoint_vs_central_fd(fd_setup):
    """Central differences (theta -> theta(1 + s v), s = +-1%) of the hits along
    parameter directions v, with common random numbers (same seeds and photon
    streams on both sides), over 1.6G photons as 40 events with their own streams:
    the aggregate directions, each property block, the 8 LS-absorption bins of
    largest yield sensitivity, a random-sign direction, and the hits of each of 20
    detector groups along the all-parameter direction. FD error from the photons
    whose outcome changed; adjoint error from 8 disjoint photon subsets. The
    tolerance is 1% wherever the FD's own error is below 0.3% of its value, and
    4 combined standard deviations elsewhere (that floor is printed)."""
    w, params, theta, sim, gs_t = fd_setup
    N = w.num_photon
    base_bnd = sim.base_bnd
    h = 0.01
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    Gdet = np.zeros((N_DET, params.n_param))
    parts = np.zeros((8,) + params.shape)     # adjoint error bar: 8 disjoint photon subsets
    for ev in range(FD_EVENTS):
        n_hit = sim.simulate_device(gs_t, photon, None, hit, eventID=ev, n_photon=N)
        G = yield_jacobian(sim, params, gs_t, hit, n_hit, N_DET, event_id=ev, n_photon=N)
        Gdet += G.double().cpu().numpy()
        idx = hit_photon_index(hit, n_hit)
        for k in range(8):
            sel = hit[:n_hit][(idx % 8) == k].contiguous()
            Gk = yield_jacobian(sim, params, gs_t, sel, sel.shape[0], N_DET, event_id=ev, n_photon=N)
            parts[k] += Gk.sum(dim=0).double().cpu().numpy().reshape(params.shape)
    Gtot = Gdet.sum(axis=0).reshape(params.shape)
    assert np.allclose(parts.sum(axis=0), Gtot, rtol=1e-4, atol=1e-6 * np.abs(Gtot).max())
    Ggrp = Gdet.reshape(N_GROUP, N_DET // N_GROUP, *params.shape).sum(axis=1)

    rng = np.random.default_rng(7)
    directions = {"all": np.ones(params.shape), "random signs": rng.choice([-1.0, 1.0], size=params.shape)}
    for name, (blk, prop) in {"LS abs": (0, 0), "LS ray": (0, 1), "water abs": (1, 0), "water ray": (1, 1)}.items():
        directions[name] = np.zeros(params.shape)
        directions[name][blk, prop] = 1
    sens = np.abs(Gtot[0, 0] * theta[0, 0])
    for b in np.argsort(sens)[::-1][:8]:
        directions[f"LS abs bin {b}"] = np.zeros(params.shape)
        directions[f"LS abs bin {b}"][0, 0, b] = 1
    report = {}
    for name, v in directions.items():
        dtheta = theta * v                      # d theta / d s along s -> theta (1 + s v)
        adj = float((Gtot * dtheta).sum())
        adj_parts = (parts * dtheta).sum(axis=(1, 2, 3))
        sig_adj = float(adj_parts.std(ddof=1) * np.sqrt(len(adj_parts)))
        dH, n_changed = 0, 0
        dG = np.zeros(N_GROUP)
        nG = np.zeros(N_GROUP)
        bnds = [params.apply(base_bnd, (theta * (1.0 + sgn * h * v)).astype(np.float32)) for sgn in (+1, -1)]
        for ev in range(FD_EVENTS):
            masks, groups = [], []
            for bnd in bnds:
                sim.set_bnd(bnd)
                nh = sim.simulate_device(gs_t, photon, None, hit, eventID=ev, n_photon=N)
                m = torch.zeros(N, dtype=torch.bool, device=DEV)
                m[hit_photon_index(hit, nh)] = True
                masks.append(m)
                if name == "all":
                    g = torch.full((N,), -1, dtype=torch.int64, device=DEV)
                    g[hit_photon_index(hit, nh)] = hit_group(hit, nh)
                    groups.append(g)
            dH += masks[0].sum().item() - masks[1].sum().item()
            n_changed += (masks[0] ^ masks[1]).sum().item()
            if name == "all":
                a, b = groups

                def count(x):
                    return torch.bincount(x[x >= 0], minlength=N_GROUP)[:N_GROUP].double().cpu().numpy()
                dG += count(a) - count(b)
                ch = a != b                      # photons whose group (or hit/no-hit) changed
                nG += count(a[ch]) + count(b[ch])
        fd = dH / (2 * h)
        sig_fd = np.sqrt(n_changed) / (2 * h)
        report[name] = (adj, fd, sig_adj, sig_fd)
        if name == "all":
            for g in range(N_GROUP):
                # adjoint error of a group taken as the all-detector error / sqrt(groups) (the FD error dominates)
                report[f"detector group {g}"] = (float((Ggrp[g] * dtheta).sum()), dG[g] / (2 * h),
                                                 sig_adj / np.sqrt(N_GROUP), np.sqrt(nG[g]) / (2 * h))
    sim.set_bnd(base_bnd)
    strict = []
    for name, (adj, fd, sa, sf) in report.items():
        rel = sf / abs(fd) if fd else np.inf
        tag = "1%" if rel < 0.003 else "4 sigma"
        print(f"FD check {name:20s} adjoint {adj:12.1f} +- {sa:8.1f}  fd {fd:12.1f} +- {sf:8.1f}  "
              f"rel diff {(adj - fd) / fd:+.4f}  FD error {rel:.4f}  tolerance {tag}", flush=True)
        if tag == "1%":
            strict.append(name)
            assert abs(adj - fd) <= 0.01 * abs(fd), (name, adj, fd, sa, sf)
        else:
            assert abs(adj - fd) < 4 * np.hypot(sa, sf) + 1e-9, (name, adj, fd, sa, sf)
    assert {"all", "LS abs", "water abs"} <= set(strict), strict


def test_autograd_wrapper_matches_jacobian():
    w = configs.config3(num_gs=50, per_gs=2000)
    params = w.meta["params"]
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    theta = torch.tensor(w.meta["theta"], dtype=torch.float32, requires_grad=True)
    counts = expected_hits(theta, sim, params, gs_t, N_DET)
    assert counts.shape == (N_DET,)
    weights = torch.linspace(-1.0, 1.0, N_DET, device=counts.device)
    (counts * weights).sum().backward()
    hit, n_hit = run_forward(sim, gs_t, w.num_photon)
    assert int(counts.sum().item()) == n_hit
    G = yield_jacobian(sim, params, gs_t, hit, n_hit, N_DET, n_photon=w.num_photon)
    ref = (weights[None, :] @ G).reshape(params.shape).cpu()
    assert torch.allclose(theta.grad, ref, rtol=1e-5, atol=1e-7 * float(ref.abs().max()))
    sim.close()


def test_adjoint_rejects_bad_arguments():
    from opticks_amd import _lib as L
    w = configs.config3(num_gs=2, per_gs=100)
    params = w.meta["params"]
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    hit, n_hit = run_forward(sim, gs_t, w.num_photon)
    G = torch.empty((N_DET, params.n_param), dtype=torch.float32, device=DEV)
    bad = LengthParams(w.physics, ("LS",), n_bin=64)
    bad.sample_bin = bad.sample_bin.copy()
    bad.sample_bin[0] = 64
    with pytest.raises(RuntimeError):
        sim.adjoint_device(gs_t, hit, n_hit, bad.c_struct(N_DET), G, n_photon=w.num_photon)
    with pytest.raises(RuntimeError):
        sim.adjoint_device(gs_t, hit, n_hit, params.c_struct(70000), G, n_photon=w.num_photon)
    # no hits: zero gradient
    sim.adjoint_device(gs_t, hit, 0, params.c_struct(N_DET), G, n_photon=w.num_photon)
    assert float(G.abs().sum()) == 0.0
    sim.close()


@pytest.mark.parametrize("K,n_bin,max_slot", [(12, 64, 0), (2, 64, 0), (32, 64, 0), (12, 100, 0), (12, 200, 0),
                                             (12, 300, 0), (12, 64, 7000)])
def test_tape_adjoint_equals_replay_bitwise(K, n_bin, max_slot):
    """Forwards made with adjoint_tape = K keep each photon's segment records; the adjoint
    scores hits from them inside the per-detector reduction (the same ScoreTape calls with
    the same arguments, rows summed in the same order) and replays only photons with more
    than K segments (K = 2 sends most hits down that path), so the gradient is the replay's
    bit for bit, and the forward's outputs do not change. 100 and 200 bins per property
    (400 and 800 parameters) run the reduction with two and four parameters per thread;
    300 bins (1200 parameters) are beyond it, so that adjoint replays every hit; max_slot
    splits the forward into launches that write their own part of the tape; K = 32, at least
    max_bounce, holds every photon, so the adjoint needs no score rows and takes all hits in
    one chunk unless the caller limits it."""
    w = configs.config3(num_gs=100, per_gs=1000, n_bin=n_bin)
    params = w.meta["params"]
    N = w.num_photon
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, max_slot=max_slot))
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    hit0, n0 = run_forward(sim, gs_t, N)
    G0 = yield_jacobian(sim, params, gs_t, hit0, n0, N_DET, n_photon=N).cpu().numpy()
    sim.set_config(SEventConfig(seed=w.seed, adjoint_tape=K, max_slot=max_slot))
    hit1, n1 = run_forward(sim, gs_t, N)
    assert n1 == n0 and torch.equal(hit1[:n1].view(torch.int32), hit0[:n0].view(torch.int32))
    G1 = yield_jacobian(sim, params, gs_t, hit1, n1, N_DET, n_photon=N).cpu().numpy()
    assert np.count_nonzero(G0) > 1000
    assert np.array_equal(G1.view(np.uint32), G0.view(np.uint32))
    # chunks of hits on the tape give the replay's chunked sums
    rows = 37 * params.n_param
    Gc0 = yield_jacobian(sim, params, gs_t, hit1, n1, N_DET, n_photon=N, max_row_floats=rows).cpu().numpy()
    sim.set_config(SEventConfig(seed=w.seed, max_slot=max_slot))
    run_forward(sim, gs_t, N)
    Gc1 = yield_jacobian(sim, params, gs_t, hit0, n0, N_DET, n_photon=N, max_row_floats=rows).cpu().numpy()
    assert np.array_equal(Gc0.view(np.uint32), Gc1.view(np.uint32))
    sim.close()


def test_tape_batch_adjoint_equals_replay_bitwise():
    """the captured forward + adjoint graph with the segment tape equals the replaying one,
    with a tape of 12 records (53-row chunks, replays of longer photons) and of 32 (at least
    max_bounce: no score rows; the explicit row limit still chunks the hits)"""
    w = configs.config3(num_gs=30, per_gs=1000)
    params = w.meta["params"]
    offs = np.array([0, 10, 20, 30])
    ev_ids = [4, 5, 6]
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    N = w.num_photon
    out = []
    for K in (0, 12, 32):
        sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, adjoint_tape=K))
        photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
        seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
        hit = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
        G = torch.full((N_DET, params.n_param), 7.0, dtype=torch.float32, device=DEV)
        b = sim.batch(gs_t, offs, ev_ids, photon, seq, hit, adjoint=params.c_struct(N_DET, 53 * params.n_param),
                      grad_t=G)
        for rep in range(2):
            b.launch()
            ho, ms = b.wait()
        out.append((G.cpu().numpy(), hit[:ho[-1]].cpu().numpy(), photon.cpu().numpy()))
        b.close()
        sim.close()
    assert np.count_nonzero(out[0][0]) > 1000
    for k in (1, 2):
        for a, b_ in zip(out[0], out[k]):
            assert np.array_equal(a.view(np.uint32), b_.view(np.uint32))
