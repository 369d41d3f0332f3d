"""Random-stream kinds on the GPU against the checker (MI355X only).

xorwow (oa_portable.h; the reference's RNG_XORWOW generator, qudarap/qrng.h:68-106):
bitwise equal to the checker with plain and persistent kernels, with event
skipahead jumps, in batches, and the event-skipahead shift invariant of
qudarap/tests/QRngTest.py:66-72 (event e == event 0 with rng_offset e*skipahead).

Precooked uniforms (the reference's aligned-randoms mode, sysrap/s_seq.h:105-129):
rows shorter than a history recycle modulo their length like s_seq::flat, rows
that never let a rejection loop accept end at the loop bound instead of spinning,
and a launch whose photons lack a row fails on the host.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from opticks_amd import configs
from opticks_amd import _lib as L
from opticks_amd.qsim import QSim, SEventConfig

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

DEV = torch.device("cuda:0")
PLAIN, WAVE = L.MI_KERNEL_PLAIN, L.MI_KERNEL_WAVEFRONT


def same(sim, rp, rs):
    assert np.array_equal(sim.seq, rs)
    assert np.array_equal(sim.photon.view(np.uint32), rp.view(np.uint32))


@pytest.mark.parametrize("kernel", [PLAIN, WAVE])
def test_xorwow_config0_matches_checker(kernel):
    w = configs.config0(num_photon=100_000)
    print("xorwow config0 kernel", kernel, flush=True)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, kernel=kernel, rng="xorwow"))
    sim.simulate(w.gensteps, 0)
    rp, rs = po.World(w.geometry, w.physics).simulate(w.gensteps, seed=w.seed, rng="xorwow")
    same(sim, rp, rs)
    # a different stream from Philox on the same seed
    ph = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, kernel=kernel))
    ph.simulate(w.gensteps, 0)
    assert not np.array_equal(ph.seq, sim.seq)
    sim.close(); ph.close()


@pytest.mark.parametrize("event", [0, 3])
def test_xorwow_config1_subset_matches_checker(event):
    w = configs.config1(num_gs=30, per_gs=2000)
    print("xorwow config1 event", event, flush=True)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, rng="xorwow", rng_offset=11))
    sim.simulate(w.gensteps, event)
    rp, rs = po.World(w.geometry, w.physics).simulate(w.gensteps, event_idx=event, seed=w.seed, rng="xorwow",
                                                      rng_offset=11)
    same(sim, rp, rs)
    sim.close()


def test_xorwow_event_skipahead_shift_invariant():
    w = configs.config1(num_gs=10, per_gs=2000)
    skip = 100000
    a = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, rng="xorwow", event_skipahead=skip))
    a.simulate(w.gensteps, 2)
    b = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, rng="xorwow", event_skipahead=skip,
                                                 rng_offset=2 * skip))
    b.simulate(w.gensteps, 0)
    assert np.array_equal(a.seq, b.seq)
    assert np.array_equal(a.photon.view(np.uint32), b.photon.view(np.uint32))
    a.close(); b.close()


def test_xorwow_batch_equals_per_event_and_checker():
    w = configs.config2(num_events=3, mean_gs=4, mean_photons=800, seed=9)
    gs_all = np.concatenate(w.events).astype(np.float32)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    N = w.num_photon
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=5, rng="xorwow"))
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
    ev_ids = [4, 5, 6]
    po_, _ = sim.simulate_batch(torch.from_numpy(gs_all).to(DEV), offs, ev_ids, photon, seq)
    P = photon.cpu().numpy(); S = seq.cpu().numpy().view(np.uint64)
    W = po.World(w.geometry, w.physics)
    for e, gs in enumerate(w.events):
        a, b = po_[e], po_[e + 1]
        rp, rs = W.simulate(gs, event_idx=ev_ids[e], seed=5, rng="xorwow")
        assert np.array_equal(S[a:b], rs)
        assert np.array_equal(P[a:b].view(np.uint32), rp.view(np.uint32))
    sim.close()


def _mixed_gensteps(n):
    """torch sphere, Marsaglia torch sphere and Cerenkov gensteps in the water sphere"""
    wc = configs.cerenkov_water(num_photon=n, n_gs=1)
    t = configs.torch_genstep(n, configs.T_SPHERE)
    m = configs.torch_genstep(n, configs.T_SPHERE_MARSAGLIA, radius=10.0)
    return np.stack([t, m, wc.gensteps[0]]).astype(np.float32), wc


def test_precooked_short_rows_recycle_like_checker():
    n = 3000
    gs, w = _mixed_gensteps(n)
    pre = np.random.default_rng(3).random((3 * n, 5)).astype(np.float32) * 0.998 + 0.001
    sim = QSim(w.geometry, w.physics, SEventConfig(precooked=pre))
    sim.simulate(gs, 0)
    rp, rs = po.World(w.geometry, w.physics).simulate(gs, precooked=pre)
    same(sim, rp, rs)
    sim.close()


def test_precooked_rows_that_never_accept_end_at_the_bound():
    """all uniforms 0.999: Marsaglia's b = 2*0.998^2 > 1 forever and Rayleigh's
    acceptance (cos^2 >= 0.999) almost never holds; every loop stops at its bound"""
    n = 64
    gs, w = _mixed_gensteps(n)
    pre = np.full((3 * n, 7), 0.999, np.float32)
    sim = QSim(w.geometry, w.physics, SEventConfig(precooked=pre))
    sim.simulate(gs, 0)
    rp, rs = po.World(w.geometry, w.physics).simulate(gs, precooked=pre)
    same(sim, rp, rs)
    assert np.all(np.isfinite(sim.photon[:, :3]))
    sim.close()


def test_precooked_rows_missing_fail_on_the_host():
    w = configs.config0(num_photon=1000)
    pre = np.full((999, 16), 0.5, np.float32)
    sim = QSim(w.geometry, w.physics, SEventConfig(precooked=pre))
    with pytest.raises(RuntimeError, match="precooked"):
        sim.simulate(w.gensteps, 0)
    sim.close()


def test_config_round_trip_keeps_precooked_and_invalidates_batches():
    w = configs.config0(num_photon=2000)
    pre = np.random.default_rng(1).random((2000, 64)).astype(np.float32) * 0.998 + 0.001
    sim = QSim(w.geometry, w.physics, SEventConfig(precooked=pre))
    sim.simulate(w.gensteps, 0)
    first = sim.seq.copy()
    cfg = sim.get_config_c()
    assert not cfg.precooked and cfg.precooked_ni == 2000 and cfg.precooked_nk == 64
    sim.set_config_c(cfg)                       # null pointer, same shape: keeps the device copy
    sim.simulate(w.gensteps, 0)
    assert np.array_equal(first, sim.seq)
    cfg.precooked_nk = 32                       # null pointer, other shape: rejected, old config kept
    with pytest.raises(RuntimeError):
        sim.set_config_c(cfg)
    assert sim.get_config_c().precooked_nk == 64
    sim.close()
    # a seed change after capture makes the batch refuse to replay stale physics
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=1))
    N = w.num_photon
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    b = sim.batch(gs_t, [0, 1], [0], photon)
    b.launch(); b.wait()
    sim.set_config(SEventConfig(seed=2))
    with pytest.raises(RuntimeError, match="recreate"):
        b.launch()
    b.close()
    sim.close()


def test_n_photon_beyond_the_gensteps_is_rejected():
    w = configs.config0(num_photon=1000)
    sim = QSim(w.geometry, w.physics, SEventConfig())
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    photon = torch.empty((2000, 4, 4), dtype=torch.float32, device=DEV)
    with pytest.raises(RuntimeError, match="exceeds"):
        sim.simulate_device(gs_t, photon, n_photon=1001)
    assert sim.simulate_device(gs_t, photon, n_photon=1000) >= 0
    sim.close()
