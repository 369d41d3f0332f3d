"""HIP path vs the CPU checker on the same seeded inputs (MI355X only).

Bar: integer/bit outputs (seqhis, seqbnd, flags, indices) identical; photon
floats identical bit for bit (both sides share operand order, no FMA
contraction, correctly rounded sqrt/div and the same log/sin/cos), with a
floor of >= 99.9% seqhis agreement and 1e-3 mm / 1e-3 ns kept as the
asserted tolerance for the rare photons that could differ.
"""
import numpy as np
import pytest

import pyoracle as po
from opticks_amd import configs
from opticks_amd import _lib as L
from opticks_amd.qsim import QSim, SEventConfig, FLAG

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

DEV = torch.device("cuda:0")
PLAIN, WAVE = L.MI_KERNEL_PLAIN, L.MI_KERNEL_WAVEFRONT


def compare(gpu_photon, gpu_seq, ref_photon, ref_seq, floor=0.999):
    assert gpu_seq.shape == ref_seq.shape
    same_seq = np.all(gpu_seq == ref_seq, axis=(1, 2))
    assert same_seq.mean() >= floor
    bits = np.all(gpu_photon.view(np.uint32) == ref_photon.view(np.uint32), axis=(1, 2))
    assert bits.mean() >= floor
    # photons with identical history agree in position and time to 1e-3 mm / ns
    m = same_seq
    assert np.abs(gpu_photon[m, 0] - ref_photon[m, 0]).max() <= 1e-3
    # integer words identical on identical histories
    assert np.array_equal(gpu_photon[m, 3].view(np.uint32), ref_photon[m, 3].view(np.uint32))
    return same_seq.mean(), bits.mean()


def stable_hits(photon, mask=FLAG["SD"]):
    return photon[(photon[:, 3, 3].view(np.uint32) & mask) != 0]


@pytest.fixture(scope="module")
def w0():
    return configs.config0(num_photon=200_000)


@pytest.fixture(scope="module")
def w1():
    return configs.config1(num_gs=40, per_gs=2500)


@pytest.mark.parametrize("kernel", [PLAIN, WAVE])
def test_config0_bitwise(w0, kernel):
    sim = QSim(w0.geometry, w0.physics, SEventConfig(seed=w0.seed, kernel=kernel))
    hits = sim.simulate(w0.gensteps, 0)
    ref_p, ref_s = po.World(w0.geometry, w0.physics).simulate(w0.gensteps, seed=w0.seed)
    s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
    assert s == 1.0 and b == 1.0
    assert np.array_equal(hits.view(np.uint32), stable_hits(ref_p).view(np.uint32))
    sim.close()


@pytest.mark.parametrize("kernel", [PLAIN, WAVE])
@pytest.mark.parametrize("block", [64, 128, 256])
def test_config1_subset_bitwise(w1, kernel, block):
    sim = QSim(w1.geometry, w1.physics, SEventConfig(seed=w1.seed, kernel=kernel, block_size=block))
    hits = sim.simulate(w1.gensteps, 0)
    ref_p, ref_s = po.World(w1.geometry, w1.physics).simulate(w1.gensteps, seed=w1.seed)
    compare(sim.photon, sim.seq, ref_p, ref_s)
    assert np.array_equal(hits.view(np.uint32), stable_hits(ref_p).view(np.uint32))
    sim.close()


def test_config1_full_size_properties_and_samples():
    """10M photons on the GPU: size-independent properties over all of them,
    and bitwise parity with the checker on three sampled photon ranges."""
    w = configs.config1()
    N = w.num_photon
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    gs = torch.from_numpy(w.gensteps).to(DEV)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
    hit = torch.empty((N // 2, 4, 4), dtype=torch.float32, device=DEV)
    n_hit = sim.simulate_device(gs, photon, seq, hit, eventID=0, n_photon=N)
    p = photon.cpu().numpy()
    s = seq.cpu().numpy().view(np.uint64)
    h = hit[:n_hit].cpu().numpy()
    # properties of the whole event
    assert np.array_equal(p[:, 3, 2].view(np.uint32), np.arange(N, dtype=np.uint32))
    fm = p[:, 3, 3].view(np.uint32)
    is_hit = (fm & FLAG["SD"]) != 0
    assert n_hit == is_hit.sum()
    assert np.array_equal(h.view(np.uint32), p[is_hit].view(np.uint32))       # stable order
    final = p[:, 3, 0].view(np.uint32) & 0xFFFF
    terminal = (final & (FLAG["SD"] | FLAG["SA"] | FLAG["AB"])) != 0
    # a photon that did not end on SD/SA/AB was truncated at max_bounce: 31 steps,
    # i.e. its 32nd nibble (slot 31, top of seqhis[1]) is set
    truncated = (s[:, 0, 1] >> np.uint64(60)) != 0
    assert np.all(terminal | truncated)
    assert terminal.mean() > 0.95
    assert np.all(np.isfinite(p[:, 0]))
    assert np.abs(np.linalg.norm(p[:, 1, :3].astype(np.float64), axis=1) - 1).max() < 1e-4
    first = s[:, 0, 0] & 0xF
    assert np.all(first == 3)                                                   # TO
    # three sampled ranges against the checker
    W = po.World(w.geometry, w.physics)
    for b in (0, N // 2 + 123, N - 4000):
        rp, rs = W.simulate(w.gensteps, seed=w.seed, p_begin=b, p_end=b + 4000)
        compare(p[b:b + 4000], s[b:b + 4000], rp, rs)
    sim.close()


def test_plain_wavefront_and_regroup_identical_full_size():
    """One lane per photon, the persistent wave kernel, the wave kernel with its
    node walk handing over to the leaf phase early (25% and 100% regroup), and
    with lanes refilled in batches (32, and all 64 at once) give the same photons
    and histories bit for bit on the 10M workload."""
    w = configs.config1()
    N = w.num_photon
    gs = torch.from_numpy(w.gensteps).to(DEV)
    ref = None
    for k, rg, rb in ((PLAIN, 0, 1), (WAVE, 0, 1), (WAVE, 25, 1), (WAVE, 100, 1), (WAVE, 25, 32), (WAVE, 0, 64)):
        print("kernel", k, "trace_regroup_pct", rg, "refill_batch", rb, flush=True)
        sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, kernel=k, trace_regroup_pct=rg, refill_batch=rb))
        photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
        seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
        sim.simulate_device(gs, photon, seq, None, n_photon=N)
        sim.close()
        if ref is None:
            ref = (photon, seq)
            continue
        assert torch.equal(ref[1], seq)
        assert torch.equal(ref[0].view(torch.int32), photon.view(torch.int32))
        del photon, seq


@pytest.mark.parametrize("dscale", [1.0, 2.5, 0.4])
def test_trace_matches_checker_brute_force(dscale):
    """60k rays from anywhere in the world box (a third start inside the free ball, which
    skips or clips their instance walk), against the checker's brute force; directions of
    length 2.5 and 0.4 as well as unit ones (t is in units of the direction's length)"""
    w = configs.config1(num_gs=1, per_gs=1)
    sim = QSim(w.geometry, w.physics, SEventConfig())
    rng = np.random.default_rng(7)
    n = 60000
    rays = np.zeros((n, 2, 4), np.float32)
    rays[:, 0, :3] = rng.uniform(-6400, 6400, (n, 3))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 1, :3] = d * dscale
    rays[:, 0, 3] = rng.choice([0.0, 0.05, 10.0], n)
    rays[:, 1, 3] = 1e6
    rt = torch.from_numpy(rays).to(DEV)
    it = torch.zeros((n, 2, 4), dtype=torch.float32, device=DEV)
    sim.trace_device(rt, it)
    g = it.cpu().numpy()
    r = po.World(w.geometry, w.physics).trace(rays, traverse=0)
    assert np.array_equal(g[:, 0].view(np.uint32), r[:, 0].view(np.uint32))       # normal, t
    assert np.array_equal(g[:, 1, :3].view(np.uint32), r[:, 1, :3].view(np.uint32))  # ids, boundary, lposcost
    assert np.mean(g[:, 1, 1].view(np.uint32) == 0xFFFF) < 0.01
    sim.close()


def test_multi_slice_equals_single_launch(w1):
    a = QSim(w1.geometry, w1.physics, SEventConfig(seed=3))
    b = QSim(w1.geometry, w1.physics, SEventConfig(seed=3, max_slot=7000))
    ha = a.simulate(w1.gensteps, 2)
    hb = b.simulate(w1.gensteps, 2)
    assert np.array_equal(a.seq, b.seq)
    assert np.array_equal(a.photon.view(np.uint32), b.photon.view(np.uint32))
    assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
    a.close(); b.close()


def test_batch_equals_per_event_and_checker():
    w = configs.config2(num_events=4, mean_gs=6, mean_photons=800)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=5))
    gs_all = np.concatenate(w.events)
    offs = np.cumsum([0] + [len(e) for e in w.events])
    N = w.num_photon
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    ev_ids = [10, 11, 12, 13]
    po_, ho = sim.simulate_batch(torch.from_numpy(gs_all).to(DEV), offs, ev_ids, photon, seq, hit)
    P = photon.cpu().numpy(); S = seq.cpu().numpy().view(np.uint64); H = hit.cpu().numpy()
    W = po.World(w.geometry, w.physics)
    for e, gs in enumerate(w.events):
        a, b = po_[e], po_[e + 1]
        single = QSim(w.geometry, w.physics, SEventConfig(seed=5))
        hs = single.simulate(gs, ev_ids[e])
        assert np.array_equal(single.seq, S[a:b])
        assert np.array_equal(single.photon.view(np.uint32), P[a:b].view(np.uint32))
        assert np.array_equal(hs.view(np.uint32), H[ho[e]:ho[e + 1]].view(np.uint32))
        single.close()
        rp, rs = W.simulate(gs, event_idx=ev_ids[e], seed=5)
        compare(P[a:b], S[a:b], rp, rs)
    sim.close()


def test_precooked_uniforms(w0):
    n = 20000
    gs = w0.gensteps.copy()
    gs.view(np.uint32)[0, 0, 3] = n
    pre = np.random.default_rng(9).random((n, 96)).astype(np.float32) * 0.999 + 0.0005
    sim = QSim(w0.geometry, w0.physics, SEventConfig(precooked=pre))
    sim.simulate(gs, 0)
    rp, rs = po.World(w0.geometry, w0.physics).simulate(gs, precooked=pre)
    s, b = compare(sim.photon, sim.seq, rp, rs)
    assert s == 1.0 and b == 1.0
    sim.close()


def test_wavelengths_outside_the_table_domain_wrap_like_the_checker(w0):
    """Torch photons at 40, 900 and 1500 nm, outside the 60-820 nm table domain:
    the lookups take the wrap-addressing path for out-of-range texels, as the
    reference's wrap-mode textures do, and match the checker bit for bit."""
    gs = np.stack([configs.torch_genstep(20000, configs.T_SPHERE, wavelength=wl) for wl in (40.0, 900.0, 1500.0)])
    gs = gs.astype(np.float32)
    sim = QSim(w0.geometry, w0.physics, SEventConfig(seed=5))
    hits = sim.simulate(gs, 0)
    ref_p, ref_s = po.World(w0.geometry, w0.physics).simulate(gs, seed=5)
    s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
    assert s == 1.0 and b == 1.0
    assert np.array_equal(np.unique(sim.photon[:, 2, 3]), np.float32([40.0, 900.0, 1500.0]))
    assert np.array_equal(hits.view(np.uint32), stable_hits(ref_p).view(np.uint32))
    sim.close()


def test_zero_gensteps_and_empty_counts(w0):
    sim = QSim(w0.geometry, w0.physics, SEventConfig())
    assert sim.simulate(np.zeros((0, 6, 4), np.float32), 0) is None
    gs = w0.gensteps.copy()
    gs.view(np.uint32)[0, 0, 3] = 0
    assert sim.simulate(gs, 0) is None
    sim.close()


@pytest.mark.parametrize("kw", [dict(max_bounce=1), dict(max_bounce=0), dict(max_time=3.0)])
def test_truncation_matches_checker(w1, kw):
    cfg = SEventConfig(seed=8, **kw)
    sim = QSim(w1.geometry, w1.physics, cfg)
    sim.simulate(w1.gensteps, 0)
    rp, rs = po.World(w1.geometry, w1.physics).simulate(w1.gensteps, seed=8, **kw)
    s, b = compare(sim.photon, sim.seq, rp, rs)
    assert s == 1.0
    sim.close()


@pytest.mark.parametrize("scale", [1.03, 0.97])
def test_refit_after_moving_instances(scale):
    """Detectors moved out to 5665 mm, or in to 5335 mm: inward, the detector domes (5145 mm)
    lie inside the free ball of the first layout (about 5309 mm), so the refit must shrink
    it; a batch captured before the refit reads the refitted TLAS and ball in place."""
    w = configs.config1(num_gs=10, per_gs=1000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=4))
    N = w.num_photon
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    photon = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    seq = torch.empty((N, 2, 2), dtype=torch.int64, device=DEV)
    hit = torch.empty((N, 4, 4), dtype=torch.float32, device=DEV)
    b = sim.batch(gs_t, np.array([0, len(w.gensteps)]), [0], photon, seq, hit)
    moved = w.geometry["inst"].copy()
    moved[1:, 3, :3] *= np.float32(scale)
    sim.refit_instances(moved)
    g2 = dict(w.geometry); g2["inst"] = moved
    sim.simulate(w.gensteps, 0)
    rp, rs = po.World(g2, w.physics).simulate(w.gensteps, seed=4)
    compare(sim.photon, sim.seq, rp, rs)
    b.launch()
    b.wait()
    compare(photon.cpu().numpy(), seq.cpu().numpy().view(np.uint64), rp, rs)
    b.close()
    sim.close()


def test_tlas_rebuild_after_refit_changes_the_front_set():
    """Detectors moved to three times their radius, outside the world box: a TLAS rebuild
    (a new leaf size) after the refit no longer keeps the world out of the TLAS, and its
    instance record keeps the one-leaf link over its three prims that it got as a front
    instance, now reached through a TLAS leaf: trace and simulate match the checker."""
    w = configs.config1(num_gs=10, per_gs=1000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=4))
    moved = w.geometry["inst"].copy()
    moved[1:, 3, :3] *= np.float32(3.0)
    sim.refit_instances(moved)
    sim.set_config(SEventConfig(seed=4, bvh_leaf_size=2))
    g2 = dict(w.geometry); g2["inst"] = moved
    W = po.World(g2, w.physics)
    rng = np.random.default_rng(5)
    n = 40000
    rays = np.zeros((n, 2, 4), np.float32)
    rays[:, 0, :3] = rng.uniform(-20000, 20000, (n, 3))
    rays[: n // 2, 0, :3] = rng.uniform(-6000, 6000, (n // 2, 3))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 1, :3] = d
    rays[:, 1, 3] = 1e6
    rt = torch.from_numpy(rays).to(DEV)
    it = torch.zeros((n, 2, 4), dtype=torch.float32, device=DEV)
    sim.trace_device(rt, it)
    g = it.cpu().numpy()
    r = W.trace(rays, traverse=0)
    assert np.array_equal(g[:, 0].view(np.uint32), r[:, 0].view(np.uint32))
    assert np.array_equal(g[:, 1, :3].view(np.uint32), r[:, 1, :3].view(np.uint32))
    inst = g[:, 1, 0].view(np.uint32) >> 16
    hit = g[:, 1, 1].view(np.uint32) != 0xFFFF
    assert (hit & (inst == 0)).sum() > 1000 and (hit & (inst > 0)).sum() > 100
    sim.simulate(w.gensteps, 0)
    rp, rs = W.simulate(w.gensteps, seed=4)
    compare(sim.photon, sim.seq, rp, rs)
    sim.close()


def test_leaf_zoo_trace_and_simulate_match_checker():
    """cone, hyperboloid, convex polyhedra, phicut/halfspace trees and a transformed cone:
    nearest hits against the checker's brute force, then full propagation"""
    w = configs.leaf_zoo(num_photon=200_000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    rng = np.random.default_rng(12)
    n = 60000
    rays = np.zeros((n, 2, 4), np.float32)
    rays[:, 0, :3] = rng.uniform(-600, 600, (n, 3))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 1, :3] = d
    rays[:, 0, 3] = rng.choice([0.0, 0.05, 1.0], n)
    rays[:, 1, 3] = 1e6
    rt = torch.from_numpy(rays).to(DEV)
    it = torch.zeros((n, 2, 4), dtype=torch.float32, device=DEV)
    sim.trace_device(rt, it)
    g = it.cpu().numpy()
    W = po.World(w.geometry, w.physics)
    r = W.trace(rays, traverse=0)
    assert np.array_equal(g[:, 0].view(np.uint32), r[:, 0].view(np.uint32))
    assert np.array_equal(g[:, 1, :3].view(np.uint32), r[:, 1, :3].view(np.uint32))
    inst = g[:, 1, 0].view(np.uint32) >> 16
    assert len(np.unique(inst[g[:, 1, 1].view(np.uint32) != 0xFFFF])) == 10      # every instance is hit
    hits = sim.simulate(w.gensteps, 0)
    ref_p, ref_s = W.simulate(w.gensteps, seed=w.seed)
    s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
    assert s >= 0.999 and len(hits) > 1000
    sim.close()


def test_cerenkov_gensteps_match_checker():
    """Cerenkov generation (qcerenkov::generate) + propagation, bitwise with the checker,
    for both the plain and the persistent kernel"""
    w = configs.cerenkov_water(num_photon=200_000)
    ref_p, ref_s = po.World(w.geometry, w.physics).simulate(w.gensteps, seed=w.seed)
    for kernel in (PLAIN, WAVE):
        sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, kernel=kernel))
        sim.simulate(w.gensteps, 0)
        s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
        assert s >= 0.999
        assert np.all((sim.photon[:, 3, 3].view(np.uint32) & 0x1) != 0)       # CK in every flagmask
        sim.close()


@pytest.mark.parametrize("which", ["config0", "config1"])
def test_simtrace_matches_checker(which):
    """Simtrace records (frame gensteps -> grid rays -> nearest hit, sevent::add_simtrace
    layout) bitwise with the checker, planar and volumetric grids, including misses"""
    if which == "config0":
        w = configs.config0(num_photon=10)
        grids = [((0, 0, 0, 1000.0), (-8, 8, 0, 0, -8, 8, 40, 1), 0.3),
                 ((0, 0, 0, 1000.0), (-3, 3, -3, 3, -3, 3, 20, 1), 0.6)]
    else:
        w = configs.config1(num_gs=1, per_gs=10)
        grids = [((0, 0, 0, 20000.0), (-16, 16, -16, 16, 0, 0, 30, 1), 0.06)]
    W = po.World(w.geometry, w.physics)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    for ce, cegs, scale in grids:
        gs = configs.frame_gensteps(ce, cegs, gridscale=scale)
        ref = W.simtrace(gs, seed=w.seed)
        print(which, cegs, len(ref), flush=True)
        out = sim.simtrace_device(torch.from_numpy(gs).to(DEV)).cpu().numpy()
        assert out.shape == ref.shape
        same = np.all(out.view(np.uint32) == ref.view(np.uint32), axis=(1, 2))
        assert same.mean() >= 0.999, same.mean()
        miss = ref.view(np.uint32)[:, 2, 3] == 0xFFFFFFFF
        assert np.array_equal(out.view(np.uint32)[:, 2, 3] == 0xFFFFFFFF, miss)
        host = sim.simtrace(gs)                                    # QSim::simtrace host form
        assert np.array_equal(host.view(np.uint32), out.view(np.uint32))
    with pytest.raises(RuntimeError):
        sim.simtrace_device(torch.from_numpy(gs).to(DEV), n_photon=len(ref) + 1)
    assert sim.simtrace(gs[:0]) is None
    sim.close()


def test_boolean_zoo_trace_and_simulate_match_checker():
    """Boolean trees of height 1, every height-2 shape and height 3, with ray origins
    inside and outside the solids so that many evaluations loop: nearest hits bit
    for bit against the checker's brute force (the straight-line evaluation of
    small trees and its fall-back to the general algorithm), then full propagation."""
    w = configs.boolean_zoo(num_photon=200_000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    rng = np.random.default_rng(12)
    n = 60000
    rays = np.zeros((n, 2, 4), np.float32)
    rays[:, 0, :3] = rng.uniform(-600, 600, (n, 3))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 1, :3] = d
    rays[:, 0, 3] = rng.choice([0.0, 0.05, 1.0], n)
    rays[:, 1, 3] = 1e6
    rt = torch.from_numpy(rays).to(DEV)
    it = torch.zeros((n, 2, 4), dtype=torch.float32, device=DEV)
    sim.trace_device(rt, it)
    g = it.cpu().numpy()
    W = po.World(w.geometry, w.physics)
    po.tree_counts(reset=True)
    r = W.trace(rays, traverse=0)
    evals, loops = po.tree_counts(reset=True)
    assert loops > 0.1 * evals
    assert np.array_equal(g[:, 0].view(np.uint32), r[:, 0].view(np.uint32))
    assert np.array_equal(g[:, 1, :3].view(np.uint32), r[:, 1, :3].view(np.uint32))
    inst = g[:, 1, 0].view(np.uint32) >> 16
    assert len(np.unique(inst[g[:, 1, 1].view(np.uint32) != 0xFFFF])) == 10
    hits = sim.simulate(w.gensteps, 0)
    ref_p, ref_s = W.simulate(w.gensteps, seed=w.seed)
    s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
    assert s >= 0.999 and len(hits) > 1000
    sim.close()


def test_fifty_thousand_detectors_match_checker():
    """50,000 instanced CSG detectors (the deepest TLAS of the suite): the
    traversal stack bound holds at create, and 20,000 photons give the checker's
    photons, histories and hits bit for bit."""
    w = configs.detector_shell(num_det=50_000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    hits = sim.simulate(w.gensteps, 0)
    ref_p, ref_s = po.World(w.geometry, w.physics).simulate(w.gensteps, seed=w.seed)
    s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
    assert s == 1.0 and b == 1.0
    assert len(hits) > 1000 and np.array_equal(hits.view(np.uint32), stable_hits(ref_p).view(np.uint32))
    sim.close()


def test_list_zoo_trace_and_simulate_match_checker():
    """CSG_CONTIGUOUS / DISCONTIGUOUS / OVERLAP list nodes at prim roots and inside trees:
    nearest hits against the checker's brute force, then full propagation"""
    w = configs.list_zoo(num_photon=200_000)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    rng = np.random.default_rng(13)
    n = 60000
    rays = np.zeros((n, 2, 4), np.float32)
    rays[:, 0, :3] = rng.uniform(-600, 600, (n, 3))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 1, :3] = d
    rays[:, 0, 3] = rng.choice([0.0, 0.05, 1.0], n)
    rays[:, 1, 3] = 1e6
    print("list_zoo trace", n, flush=True)
    rt = torch.from_numpy(rays).to(DEV)
    it = torch.zeros((n, 2, 4), dtype=torch.float32, device=DEV)
    sim.trace_device(rt, it)
    g = it.cpu().numpy()
    W = po.World(w.geometry, w.physics)
    r = W.trace(rays, traverse=0)
    assert np.array_equal(g.view(np.uint32), r.view(np.uint32))
    inst = g[:, 1, 0].view(np.uint32) >> 16
    assert len(np.unique(inst[g[:, 1, 1].view(np.uint32) != 0xFFFF])) == 7      # world + six list solids
    print("list_zoo simulate", flush=True)
    hits = sim.simulate(w.gensteps, 0)
    ref_p, ref_s = W.simulate(w.gensteps, seed=w.seed)
    s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
    assert s >= 0.999 and len(hits) > 1000
    sim.close()


@pytest.mark.parametrize("rows", [(10, 7), (5, 4)])
def test_multi_prim_solids_trace_and_simulate_match_checker(rows):
    """Eight instances of a five-prim solid (their BLAS roots have several children)
    and a world solid of 141 prims (above the 127 one leaf link holds: the front solid
    is walked through its BLAS) or of 41 (tested as one leaf): nearest hits against
    the checker's brute force, then full propagation"""
    w = configs.prim_clusters(num_photon=200_000, rows=rows)
    assert w.meta["n_world_prims"] == 1 + 2 * rows[0] * rows[1]
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    rng = np.random.default_rng(12)
    n = 60000
    rays = np.zeros((n, 2, 4), np.float32)
    rays[:, 0, :3] = rng.uniform(-900, 900, (n, 3))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 1, :3] = d
    rays[:, 0, 3] = rng.choice([0.0, 0.05, 1.0], n)
    rays[:, 1, 3] = 1e6
    rt = torch.from_numpy(rays).to(DEV)
    it = torch.zeros((n, 2, 4), dtype=torch.float32, device=DEV)
    sim.trace_device(rt, it)
    g = it.cpu().numpy()
    W = po.World(w.geometry, w.physics)
    r = W.trace(rays, traverse=0)
    assert np.array_equal(g[:, 0].view(np.uint32), r[:, 0].view(np.uint32))
    assert np.array_equal(g[:, 1, :3].view(np.uint32), r[:, 1, :3].view(np.uint32))
    inst = g[:, 1, 0].view(np.uint32) >> 16
    bnd = g[:, 1, 1].view(np.uint32)
    assert len(np.unique(inst[bnd != 0xFFFF])) == 9                        # the world and every cluster
    assert ((inst == 0) & (bnd == w.meta["b_obj"])).sum() > 100           # beads of the world solid
    hits = sim.simulate(w.gensteps, 0)
    ref_p, ref_s = W.simulate(w.gensteps, seed=w.seed)
    s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
    assert s >= 0.999 and len(hits) > 10000
    sim.close()


def test_torch_types_match_checker():
    """every storch type (point, line, disc, circle, rectangle, sphere, Marsaglia sphere)
    generated and propagated in the config0 world, bitwise with the checker"""
    w = configs.config0(num_photon=10)
    T = configs
    # the rectangle's sides come from the absolute photon index (storch.h:484-487), so it goes first
    gs = np.stack([
        T.torch_genstep(20000, T.T_RECTANGLE, zenith=(-300.0, 500.0), azimuth=(-200.0, 400.0)),
        T.torch_genstep(20000, T.T_POINT, pos=(1.0, 2.0, 3.0), mom=(0.0, 0.6, 0.8)),
        T.torch_genstep(20000, T.T_LINE, pos=(0.0, 0.0, 5.0), mom=(0.0, 0.0, 1.0), radius=300.0),
        T.torch_genstep(20000, T.T_DISC, pos=(10.0, 0.0, 0.0), mom=(0.0, 0.6, 0.8), radius=500.0),
        T.torch_genstep(20000, T.T_CIRCLE, radius=-900.0),
        T.torch_genstep(20000, T.T_SPHERE, radius=-300.0),
        T.torch_genstep(20000, T.T_SPHERE_MARSAGLIA, radius=200.0),
    ]).astype(np.float32)
    ref_p, ref_s = po.World(w.geometry, w.physics).simulate(gs, seed=w.seed)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed))
    sim.simulate(gs, 0)
    s, b = compare(sim.photon, sim.seq, ref_p, ref_s)
    assert s >= 0.999 and b >= 0.999
    sim.close()


@pytest.mark.parametrize("kernel", [PLAIN, WAVE])
def test_partial_photon_count_is_a_prefix(w1, kernel):
    """a caller's n_photon below the gensteps' total simulates exactly the first n_photon
    photons of the full launch (the per-photon genstep map is written only that far)"""
    N = w1.num_photon
    n = N - 3777
    sim = QSim(w1.geometry, w1.physics, SEventConfig(seed=w1.seed, kernel=kernel))
    gs = torch.from_numpy(w1.gensteps).to(DEV)
    full_p = torch.zeros((N, 4, 4), dtype=torch.float32, device=DEV)
    full_s = torch.zeros((N, 2, 2), dtype=torch.int64, device=DEV)
    part_p = torch.zeros((n, 4, 4), dtype=torch.float32, device=DEV)
    part_s = torch.zeros((n, 2, 2), dtype=torch.int64, device=DEV)
    sim.simulate_device(gs, full_p, full_s, None, n_photon=N)
    sim.simulate_device(gs, part_p, part_s, None, n_photon=n)
    assert np.array_equal(part_p.cpu().numpy().view(np.uint32), full_p[:n].cpu().numpy().view(np.uint32))
    assert np.array_equal(part_s.cpu().numpy(), full_s[:n].cpu().numpy())
    sim.close()
