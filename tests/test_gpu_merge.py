"""GPU hit merge (mi_sim_merge_hits_device) against the CPU restatement
(oracle/spm.py), bit for bit: real hits of the CSG workload, synthetic records
with many (identity, bucket) collisions and time ties, both record types,
selection-only and already-selected modes, and the edge cases."""
import numpy as np
import pytest

import spm
from opticks_amd import configs
from opticks_amd import _lib as L
from opticks_amd.qsim import QSim, SEventConfig

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

DEV = torch.device("cuda:0")
SD, SA = 0x40, 0x80


@pytest.fixture(scope="module")
def sim_and_photons():
    w = configs.config1(num_gs=20, per_gs=5000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    N = w.num_photon
    gs = torch.from_numpy(w.gensteps).to(DEV)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    sim.simulate_device(gs, photon, None, None, n_photon=N)
    yield sim, photon
    sim.close()


def synthetic(n, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 4, 4), np.float32)
    u = p.view(np.uint32)
    p[:, 0, 0] = np.arange(n)
    p[:, 0, 3] = rng.integers(0, 40, n) * 0.25          # many exact time ties
    u[:, 1, 3] = (rng.integers(1, 4, n).astype(np.uint32) << 16) | 3
    u[:, 3, 1] = rng.integers(1, 30, n).astype(np.uint32) | (rng.integers(0, 2, n).astype(np.uint32) << 16)
    u[:, 3, 3] = np.where(rng.random(n) < 0.6, SD, SA).astype(np.uint32) | (rng.integers(0, 4, n).astype(np.uint32) << 2)
    return p


@pytest.mark.parametrize("tw", [1.0, 0.5, 7.0, 0.0])
def test_photon_merge_real_hits(sim_and_photons, tw):
    sim, photon = sim_and_photons
    out = sim.merge_hits_device(photon, SD, tw)
    ref = spm.merge_photons(photon.cpu().numpy(), SD, tw)
    assert out.shape[0] == len(ref) > 0
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("seed", [1, 2])
def test_photon_merge_synthetic_collisions(sim_and_photons, seed):
    sim, _ = sim_and_photons
    P = synthetic(50_000, seed)
    t = torch.from_numpy(P).to(DEV)
    for mask, tw in [(SD, 1.0), (SD | SA, 2.5), (L.MI_MERGE_ALREADY_SELECTED, 1.0), (SD, 0.0)]:
        out = sim.merge_hits_device(t, mask, tw).cpu().numpy()
        ref = spm.merge_photons(P, mask, tw)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), (mask, tw)


def test_lite_merge_synthetic(sim_and_photons):
    sim, _ = sim_and_photons
    P = synthetic(40_000, 9)
    u = P.view(np.uint32)
    lt = np.stack([(u[:, 1, 3] & 0xFFFF0000) | (u[:, 3, 1] & 0xFFFF), u[:, 0, 3], np.arange(len(P), dtype=np.uint32),
                   u[:, 3, 3]], axis=1).astype(np.uint32)
    t = torch.from_numpy(lt.view(np.int32)).to(DEV)
    for mask, tw in [(SD, 1.0), (L.MI_MERGE_ALREADY_SELECTED, 0.75), (SD, 0.0)]:
        out = sim.merge_hits_device(t, mask, tw).cpu().numpy().view(np.uint32)
        ref = spm.merge_lites(lt, mask, tw)
        assert np.array_equal(out, ref), (mask, tw)


def test_merge_edge_cases(sim_and_photons):
    sim, _ = sim_and_photons
    empty = torch.empty((0, 4, 4), dtype=torch.float32, device=DEV)
    assert sim.merge_hits_device(empty, SD, 1.0).shape[0] == 0
    P = synthetic(1000, 4)
    P.view(np.uint32)[:, 3, 3] = SA
    t = torch.from_numpy(P).to(DEV)
    assert sim.merge_hits_device(t, SD, 1.0).shape[0] == 0          # nothing selected
    small = torch.empty((3, 4, 4), dtype=torch.float32, device=DEV)
    with pytest.raises(RuntimeError, match="out_capacity"):
        sim.merge_hits_device(t, SA, 1.0, out_t=small)
    with pytest.raises(RuntimeError):
        sim.merge_hits_device(t, SA, -1.0)


def test_photonlite_output_matches_checker_and_merges():
    import pyoracle as po
    w = configs.config1(num_gs=20, per_gs=2000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    N = w.num_photon
    gs = torch.from_numpy(w.gensteps).to(DEV)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
    lite = torch.empty((N, 4), dtype=torch.int32, device=DEV)
    sim.simulate_lite_device(gs, lite, photon, seq, n_photon=N)
    G = lite.cpu().numpy().view(np.uint32)
    P = photon.cpu().numpy()
    rp, rs, rl = po.World(w.geometry, w.physics).simulate_lite(w.gensteps, seed=w.seed)
    same = np.all(seq.cpu().numpy().view(np.uint64) == rs, axis=(1, 2))
    assert same.mean() >= 0.999
    assert np.array_equal(G[same], rl[same])
    # lite fields are the photon's identity, time and flagmask, hitcount 1
    u = P.view(np.uint32)
    assert np.array_equal(G[:, 0] & 0xFFFF, u[:, 3, 1] & 0xFFFF)
    assert np.all(G[:, 0] >> 16 == 1)
    assert np.array_equal(G[:, 1], u[:, 0, 3])
    assert np.array_equal(G[:, 3], u[:, 3, 3])
    # most detected photons arrive from the centre onto the detector front (local z > 0),
    # some on the side or neck (lposcost <= 0, clamped to 0 by the 16-bit packing)
    sd = (u[:, 3, 3] & SD) != 0
    assert sd.sum() > 100 and ((G[sd, 2] >> 16) > 0).mean() > 0.5
    assert np.ptp((G[sd, 2] & 0xFFFF).astype(np.int64)) > 30000      # azimuths spread round the axis
    # same results with and without the lite instantiation
    p2 = torch.empty_like(photon)
    sim.simulate_device(gs, p2, None, None, n_photon=N)
    assert np.array_equal(p2.cpu().numpy().view(np.uint32), u)
    # lite hits selected and merged on the GPU equal the restatement on the checker's lites
    out = sim.merge_hits_device(lite, SD, 1.0).cpu().numpy().view(np.uint32)
    ref = spm.merge_lites(rl[same], SD, 1.0) if same.all() else None
    if ref is not None:
        assert np.array_equal(out, ref)
    assert np.array_equal(out, spm.merge_lites(G, SD, 1.0))
    sim.close()
