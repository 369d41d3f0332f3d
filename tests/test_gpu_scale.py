"""The batched and differentiable workloads at their own geometry and scale (MI355X only).

* 256-event workload layout (20,000-detector geometry, LS/water lengths as 64 bins
  each, 256 parameters) at test size: forward + adjoint captured in one graph
  equals the per-event forward and adjoint, and the checker's score rows summed
  per detector; the graph's hits equal the checker's photons.
* 64-event batched workload at full size (about 64M photons, 20,000 detectors):
  one captured-graph replay equals the eager batched launch bit for bit, every
  event's photon indices and per-event hit ranges are consistent, and three
  sampled photon ranges of three events equal the checker bit for bit.
"""
import numpy as np
import pytest

import pyoracle as po
from opticks_amd import configs
from opticks_amd.adjoint import yield_jacobian
from opticks_amd.qsim import QSim, SEventConfig, FLAG

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

DEV = torch.device("cuda:0")
N_DET = 20000


def test_config4_graph_forward_adjoint_equals_per_event_and_checker():
    w = configs.config4(num_events=6, mean_gs=10, mean_photons=800)
    params = w.meta["params"]
    assert params.n_param == 256
    gs_np = np.concatenate(w.events).astype(np.float32)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    ev_ids = list(range(100, 100 + len(w.events)))
    N = w.num_photon
    print("config4 test size: events", len(w.events), "photons", N, flush=True)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs_t = torch.from_numpy(gs_np).to(DEV)
    # per-event forward + adjoint
    ref = np.zeros((N_DET, params.n_param))
    hits_ref = []
    for e in range(len(w.events)):
        ge = gs_t[offs[e]:offs[e + 1]].contiguous()
        ne = int(w.events[e].view(np.uint32)[:, 0, 3].sum())
        ph = torch.empty((ne, 4, 4), dtype=torch.float32, device=DEV)
        ht = torch.empty((ne, 4, 4), dtype=torch.float32, device=DEV)
        nh = sim.simulate_device(ge, ph, None, ht, eventID=ev_ids[e], n_photon=ne)
        hits_ref.append(ht[:nh].cpu().numpy())
        ref += yield_jacobian(sim, params, ge, ht, nh, N_DET, event_id=ev_ids[e],
                              n_photon=ne).double().cpu().numpy()
    assert sum(len(h) for h in hits_ref) > 50
    # one captured graph: forward of every event, then the adjoint of every hit
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    G = torch.zeros((N_DET, params.n_param), dtype=torch.float32, device=DEV)
    b = sim.batch(gs_t, offs, ev_ids, photon, seq, hit, adjoint=params.c_struct(N_DET, 0), grad_t=G)
    b.launch()
    ho, ms = b.wait()
    H = hit.cpu().numpy()
    for e in range(len(w.events)):
        assert np.array_equal(H[ho[e]:ho[e + 1]].view(np.uint32), hits_ref[e].view(np.uint32))
    Gg = G.double().cpu().numpy()
    scale = np.abs(ref).max()
    assert scale > 0 and np.count_nonzero(Gg) > 100
    assert np.allclose(Gg, ref, rtol=1e-4, atol=1e-6 * scale)
    # the checker: forward photons of every event and the score rows of its hits per detector
    W = po.World(w.geometry, w.physics)
    rows64 = np.zeros((N_DET, params.n_param))
    abs64 = np.zeros((N_DET, params.n_param))
    P = photon.cpu().numpy()
    S = seq.cpu().numpy().view(np.uint64)
    po_ = b.photon_offsets
    for e in range(len(w.events)):
        rp, rs = W.simulate(w.events[e], event_idx=ev_ids[e], seed=w.seed)
        assert np.array_equal(S[po_[e]:po_[e + 1]], rs)
        assert np.array_equal(P[po_[e]:po_[e + 1]].view(np.uint32), rp.view(np.uint32))
        he = hits_ref[e]
        idx = he[:, 3, 2].view(np.uint32).astype(np.int64)
        det = (he[:, 3, 1].view(np.uint32) & 0xffff).astype(np.int64) - 1
        rows = W.adjoint_rows(w.events[e], idx, params.mat_block, params.sample_bin, params.n_bin,
                              event_idx=ev_ids[e], seed=w.seed).astype(np.float64)
        np.add.at(rows64, det, rows)
        np.add.at(abs64, det, np.abs(rows))
    assert np.all(np.abs(Gg - rows64) <= 2e-5 * abs64 + 1e-30)
    b.close()
    # the same graph with a 32-record tape (at least max_bounce: no score rows, all hits in
    # one chunk, scored from the tape): the same gradient bit for bit
    sim.set_config(SEventConfig(seed=w.seed, adjoint_tape=32))
    G2 = torch.zeros_like(G)
    b2 = sim.batch(gs_t, offs, ev_ids, photon, seq, hit, adjoint=params.c_struct(N_DET, 0), grad_t=G2)
    b2.launch()
    b2.wait()
    assert np.array_equal(G2.cpu().numpy().view(np.uint32), G.cpu().numpy().view(np.uint32))
    b2.close()
    sim.close()


def test_config2_full_size_graph_and_samples():
    w = configs.config2()
    gs_np = np.concatenate(w.events).astype(np.float32)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    ev_ids = list(range(len(w.events)))
    N = w.num_photon
    print("config2 full size: photons", N, flush=True)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs_t = torch.from_numpy(gs_np).to(DEV)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
    hit = torch.empty((N // 4, 4, 4), dtype=torch.float32, device=DEV)
    po_e, ho_e = sim.simulate_batch(gs_t, offs, ev_ids, photon, seq, hit)
    assert po_e[-1] == N
    # properties of every event
    idx = photon[:, 3, 2].view(torch.int32).cpu().numpy().view(np.uint32)
    for e in range(len(ev_ids)):
        a, b = po_e[e], po_e[e + 1]
        assert np.array_equal(idx[a:b], np.arange(b - a, dtype=np.uint32))
    fm = photon[:, 3, 3].view(torch.int32).cpu().numpy().view(np.uint32)
    is_hit = (fm & FLAG["SD"]) != 0
    for e in range(len(ev_ids)):
        assert ho_e[e + 1] - ho_e[e] == is_hit[po_e[e]:po_e[e + 1]].sum()
    first = seq[:, 0, 0].cpu().numpy().view(np.uint64) & np.uint64(0xF)
    assert np.all(first == FLAG["SI"].bit_length())
    # a captured replay equals the eager launch
    seq_e = seq.clone()
    hit_e = hit[:ho_e[-1]].clone()
    ph_bits = photon.view(torch.int32).clone()
    b = sim.batch(gs_t, offs, ev_ids, photon, seq, hit)
    photon.zero_(); seq.zero_()
    torch.cuda.synchronize()
    b.launch()
    ho, ms = b.wait()
    assert np.array_equal(ho, ho_e)
    assert torch.equal(seq, seq_e)
    assert torch.equal(photon.view(torch.int32), ph_bits)
    assert torch.equal(hit[:ho[-1]].view(torch.int32), hit_e.view(torch.int32))
    b.close()
    del seq_e, hit_e, ph_bits
    # sampled ranges against the checker
    W = po.World(w.geometry, w.physics)
    S = seq.cpu().numpy().view(np.uint64)
    for e, off in ((0, 0), (31, 1234), (63, None)):
        a, n_e = po_e[e], po_e[e + 1] - po_e[e]
        lo = n_e - 3000 if off is None else off
        rp, rs = W.simulate(w.events[e], event_idx=ev_ids[e], seed=w.seed, p_begin=lo, p_end=lo + 3000)
        gp = photon[a + lo:a + lo + 3000].cpu().numpy()
        assert np.array_equal(S[a + lo:a + lo + 3000], rs)
        assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))
    sim.close()


@pytest.mark.parametrize("kernel", [0, 1])
def test_surface_zoo_matches_checker(kernel):
    """every propagate_at_surface outcome (SA, SD, DR, SR) and the
    Surface_zplus_sensor_A dispatch, histories up to 31 bounces"""
    w = configs.surface_zoo(num_photon=200_000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, kernel=kernel))
    hits = sim.simulate(w.gensteps, 0)
    rp, rs = po.World(w.geometry, w.physics).simulate(w.gensteps, seed=w.seed)
    assert np.array_equal(sim.seq, rs)
    assert np.array_equal(sim.photon.view(np.uint32), rp.view(np.uint32))
    fm = rp[:, 3, 3].view(np.uint32)
    for lab in ("SA", "SD", "DR", "SR"):
        assert np.count_nonzero(fm & FLAG[lab]) > 1000, lab
    assert np.array_equal(hits.view(np.uint32), rp[(fm & FLAG["SD"]) != 0].view(np.uint32))
    sim.close()
