"""Captured batches (hipGraph replay) and the autotuner (MI355X only).

Every replay of a captured batch must equal the eagerly launched batched form
bit for bit (photon, seq, hits, per-event hit offsets), with the program's own
eager work and synchronisations between replays; new genstep contents are
picked up by the next replay; a changed photon count is reported, not
silently propagated. Autotuned settings (block size, refill period, TLAS leaf
size) change speed only: results stay bit-identical.
"""
import numpy as np
import pytest

from opticks_amd import configs
from opticks_amd.qsim import QSim, SEventConfig

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def w2():
    return configs.config2(num_events=6, mean_gs=10, mean_photons=800)


def buffers(N):
    return (torch.empty((N, 4, 4), dtype=torch.float32, device=DEV),
            torch.empty((N, 2, 2), dtype=torch.int64, device=DEV),
            torch.empty((N, 4, 4), dtype=torch.float32, device=DEV))


def eager(sim, gs_np, offs, ev_ids, N):
    photon, seq, hit = buffers(N)
    po_, ho = sim.simulate_batch(torch.from_numpy(gs_np).to(DEV), offs, ev_ids, photon, seq, hit)
    return po_, ho, photon.cpu().numpy(), seq.cpu().numpy(), hit[:ho[-1]].cpu().numpy()


def test_graph_replays_equal_eager(w2):
    sim = QSim(w2.geometry, w2.physics, SEventConfig(seed=3))
    gs_np = np.concatenate(w2.events)
    offs = np.cumsum([0] + [len(e) for e in w2.events])
    ev_ids = list(range(20, 20 + len(w2.events)))
    N = w2.num_photon
    po_ref, ho_ref, P, S, H = eager(sim, gs_np, offs, ev_ids, N)
    gs_t = torch.from_numpy(gs_np).to(DEV)
    photon, seq, hit = buffers(N)
    b = sim.batch(gs_t, offs, ev_ids, photon, seq, hit)
    assert b.num_photon == N
    assert np.array_equal(b.photon_offsets, po_ref)
    for rep in range(3):
        photon.fill_(0.0)
        seq.fill_(0)
        torch.cuda.synchronize()
        b.launch()
        ho, ms = b.wait()
        assert ms > 0.0
        assert np.array_equal(ho, ho_ref), rep
        assert np.array_equal(photon.cpu().numpy().view(np.uint32), P.view(np.uint32)), rep
        assert np.array_equal(seq.cpu().numpy(), S), rep
        assert np.array_equal(hit[:ho[-1]].cpu().numpy().view(np.uint32), H.view(np.uint32)), rep
    # new contents, same photon counts: the replay follows the gensteps
    gs2 = gs_np.copy()
    gs2[:, 1, 3] += 5.0        # genstep time
    gs2[:, 1, 0] += 100.0      # genstep x position
    gs_t.copy_(torch.from_numpy(gs2).to(DEV))
    torch.cuda.synchronize()
    b.launch()
    ho2, _ = b.wait()
    _, ho2_ref, P2, S2, H2 = eager(sim, gs2, offs, ev_ids, N)
    assert np.array_equal(ho2, ho2_ref)
    assert np.array_equal(photon.cpu().numpy().view(np.uint32), P2.view(np.uint32))
    assert np.array_equal(seq.cpu().numpy(), S2)
    assert not np.array_equal(P2.view(np.uint32), P.view(np.uint32))
    # a changed photon count is reported by the device-side prefix check
    gs3 = gs2.copy()
    gs3.view(np.uint32)[0, 0, 3] -= 1
    gs_t.copy_(torch.from_numpy(gs3).to(DEV))
    torch.cuda.synchronize()
    b.launch()
    with pytest.raises(RuntimeError, match="photon counts"):
        b.wait()
    b.close()
    sim.close()


def test_batch_invalidated_by_leaf_change(w2):
    sim = QSim(w2.geometry, w2.physics, SEventConfig(seed=3))
    gs_np = np.concatenate(w2.events)
    offs = np.cumsum([0] + [len(e) for e in w2.events])
    N = w2.num_photon
    photon, seq, hit = buffers(N)
    b = sim.batch(torch.from_numpy(gs_np).to(DEV), offs, list(range(len(w2.events))), photon, seq, hit)
    cfg = sim.config
    cfg.bvh_leaf_size = 2
    sim.set_config(cfg)
    with pytest.raises(RuntimeError, match="recreate"):
        b.launch()
    b.close()
    sim.close()


def test_autotune_picks_a_candidate_and_keeps_results(w2):
    sim = QSim(w2.geometry, w2.physics, SEventConfig(seed=3))
    gs_np = np.concatenate(w2.events)
    offs = np.cumsum([0] + [len(e) for e in w2.events])
    ev_ids = list(range(len(w2.events)))
    N = w2.num_photon
    _, ho_ref, P, S, H = eager(sim, gs_np, offs, ev_ids, N)
    photon, seq, hit = buffers(N)
    res = sim.autotune(torch.from_numpy(gs_np).to(DEV), offs, ev_ids, photon, seq, hit, reps=1)
    assert len(res) == 54
    assert {r["trace_regroup_pct"] for r in res} == {0, 25} and {r["refill_batch"] for r in res} == {1, 16, 32}
    assert all(r["ms"] > 0 for r in res)
    best = min(res, key=lambda r: r["ms"])
    c = sim.config
    keys = ("block_size", "refill_period", "bvh_leaf_size", "trace_regroup_pct", "refill_batch")
    assert tuple(getattr(c, k) for k in keys) == tuple(best[k] for k in keys)
    _, ho, P2, S2, H2 = eager(sim, gs_np, offs, ev_ids, N)
    assert np.array_equal(ho, ho_ref)
    assert np.array_equal(P2.view(np.uint32), P.view(np.uint32))
    assert np.array_equal(S2, S)
    sim.close()


def test_batch_adjoint_graph_equals_per_event_adjoints():
    from opticks_amd.adjoint import yield_jacobian
    w = configs.config3(num_gs=30, per_gs=1000)
    params = w.meta["params"]
    n_det = 2000
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    offs = np.array([0, 10, 20, 30])
    ev_ids = [4, 5, 6]
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    N = w.num_photon
    # per-event reference: forward + adjoint of each event on its own
    ref = np.zeros((n_det, params.n_param))
    for e in range(3):
        ge = gs_t[offs[e]:offs[e + 1]].contiguous()
        ne = int(w.gensteps[offs[e]:offs[e + 1]].view(np.uint32)[:, 0, 3].sum())
        ph = torch.empty((ne, 4, 4), dtype=torch.float32, device=DEV)
        ht = torch.empty((ne, 4, 4), dtype=torch.float32, device=DEV)
        nh = sim.simulate_device(ge, ph, None, ht, eventID=ev_ids[e], n_photon=ne)
        ref += yield_jacobian(sim, params, ge, ht, nh, n_det, event_id=ev_ids[e], n_photon=ne).double().cpu().numpy()
    scale = np.abs(ref).max()
    assert scale > 0
    photon, seq, hit = buffers(N)
    for max_rows in (0, 53 * params.n_param):          # one chunk, then chunks of 53 hit slots
        G = torch.full((n_det, params.n_param), 7.0, dtype=torch.float32, device=DEV)
        b = sim.batch(gs_t, offs, ev_ids, photon, seq, hit, adjoint=params.c_struct(n_det, max_rows), grad_t=G)
        outs = []
        for rep in range(2):
            b.launch()
            ho, ms = b.wait()
            t = b.last_timing()
            assert t["forward_ms"] > 0 and t["adjoint_ms"] > 0
            assert abs(t["forward_ms"] + t["adjoint_ms"] - ms) < 1e-3 * ms + 1e-3
            outs.append(G.double().cpu().numpy())
            torch.cuda.synchronize()
        assert np.array_equal(outs[0], outs[1])            # replays are deterministic
        assert np.allclose(outs[0], ref, rtol=1e-4, atol=1e-6 * scale)
        b.close()
    sim.close()
