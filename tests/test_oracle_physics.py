"""The CPU checker's physics against closed forms, invariants and statistics."""
import numpy as np
import pytest

import pyoracle as po
from opticks_amd import configs
from opticks_amd.qsim import seqhis_label, FLAG

C_LIGHT = 299.792458


@pytest.fixture(scope="module")
def config0_run():
    w = configs.config0(num_photon=1_000_000)
    W = po.World(w.geometry, w.physics)
    photon, seq = W.simulate(w.gensteps, seed=w.seed)
    return w, photon, seq


def labels(seq):
    return np.array([seqhis_label(a, b) for a, b in seq[:, 0]])


def test_config0_first_interaction_probabilities(config0_run):
    """Radial photons from the centre of the water sphere: one attenuation
    length scale L=1/(1/20000+1/30000)=12000 mm over 1000 mm, absorption share
    0.6, normal-incidence Fresnel R=((n-1)/(n+1))^2 (qsim.h:1027 branch)."""
    w, photon, seq = config0_run
    N = len(seq)
    lab = labels(seq[: 400_000])
    n = len(lab)
    att = 1 - np.exp(-1000.0 / 12000.0)
    R = ((1.333 - 1.0) / 2.333) ** 2
    expect = {"TO AB": 0.6 * att, "TO BT SD": (1 - att) * (1 - R)}
    for k, p in expect.items():
        got = np.mean(lab == k)
        sig = np.sqrt(p * (1 - p) / n)
        assert abs(got - p) < 5 * sig, (k, got, p)
    sc = np.mean(np.char.startswith(lab, "TO SC"))
    assert abs(sc - 0.4 * att) < 5 * np.sqrt(0.4 * att / n)
    br = np.mean(np.char.startswith(lab, "TO BR"))
    assert abs(br - (1 - att) * R) < 5 * np.sqrt((1 - att) * R / n)


def test_config0_direct_detection_kinematics(config0_run):
    """TO BT SD photons: straight from the origin, time = 1000*n/c + (d_box-1000)/c."""
    w, photon, seq = config0_run
    lab = labels(seq[:200_000])
    m = lab == "TO BT SD"
    p = photon[:200_000][m]
    pos = p[:, 0, :3].astype(np.float64)
    u = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    assert np.allclose(np.abs(pos).max(axis=1), 2000.0, atol=1e-3)
    d_box = 2000.0 / np.abs(u).max(axis=1)
    t = 1000.0 * 1.333 / C_LIGHT + (d_box - 1000.0) / C_LIGHT
    assert np.abs(p[:, 0, 3] - t).max() < 1e-3
    # flags: final flag SD, flagmask TO|BT|SD, hitcount 1
    flag = p[:, 3, 0].view(np.uint32) & 0xFFFF
    assert np.all(flag == FLAG["SD"])
    assert np.all(p[:, 3, 3].view(np.uint32) == FLAG["TO"] | FLAG["BT"] | FLAG["SD"])


def test_config0_indices_and_norms(config0_run):
    w, photon, seq = config0_run
    assert np.array_equal(photon[:, 3, 2].view(np.uint32), np.arange(len(photon), dtype=np.uint32))
    mom = photon[:, 1, :3].astype(np.float64)
    pol = photon[:, 2, :3].astype(np.float64)
    assert np.abs(np.linalg.norm(mom, axis=1) - 1).max() < 1e-5
    assert np.abs(np.linalg.norm(pol, axis=1) - 1).max() < 1e-5
    # transversality after whole FP32 paths (single scatters are held to 1e-5 below)
    dots = np.abs(np.einsum("ij,ij->i", mom, pol))
    assert np.mean(dots < 1e-5) > 0.9999 and dots.max() < 1e-3


@pytest.mark.parametrize("theta_deg", [0.0, 10.0, 35.0, 60.0, 80.0])
@pytest.mark.parametrize("pol_kind", ["s", "p"])
def test_fresnel_coefficients(theta_deg, pol_kind):
    """TransCoeff against the Fresnel equations (Hecht, Optics 4.42-4.43):
    R_s=((n1c1-n2c2)/(n1c1+n2c2))^2, R_p=((n2c1-n1c2)/(n2c1+n1c2))^2."""
    n1, n2 = 1.0, 1.5
    th = np.radians(theta_deg)
    mom = np.array([np.sin(th), 0.0, np.cos(th)])
    normal = np.array([0.0, 0.0, -1.0])
    s_dir = np.array([0.0, 1.0, 0.0])
    pol = s_dir if pol_kind == "s" or theta_deg == 0 else np.cross(s_dir, mom)
    pol /= np.linalg.norm(pol)
    c1 = np.cos(th)
    c2 = np.sqrt(1 - (n1 / n2) ** 2 * (1 - c1 * c1))
    if pol_kind == "s":
        R = ((n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2)) ** 2
    else:
        R = ((n2 * c1 - n1 * c2) / (n2 * c1 + n1 * c2)) ** 2
    mo, po_, flag, T = po.step_boundary(mom, pol, normal, n1, n2, 0.999999)
    assert abs(T - (1 - R)) < 3e-6
    # transmit branch obeys Snell, reflect branch mirrors
    mo_t, pol_t, f_t, _ = po.step_boundary(mom, pol, normal, n1, n2, 0.0)
    assert f_t == FLAG["BT"]
    assert abs(np.linalg.norm(mo_t) - 1) < 1e-6
    assert abs(np.hypot(mo_t[0], mo_t[1]) - n1 / n2 * np.sin(th)) < 1e-6
    mo_r, pol_r, f_r, _ = po.step_boundary(mom, pol, normal, n1, n2, 1.0)
    if T < 1.0:
        assert f_r == FLAG["BR"]
        assert np.allclose(mo_r, [mom[0], mom[1], -mom[2]], atol=1e-6)
    for m_, p_ in ((mo_t, pol_t), (mo_r, pol_r)):
        assert abs(np.linalg.norm(p_) - 1) < 1e-5 and abs(np.dot(m_, p_)) < 1e-5


def test_total_internal_reflection_always_reflects():
    th = np.radians(50.0)                 # beyond asin(1/1.5) = 41.8 deg
    mom = np.array([np.sin(th), 0.0, np.cos(th)])
    for u in (2.0 ** -24, 0.5, 1.0):     # uniforms are in (0, 1]
        mo, pl, flag, T = po.step_boundary(mom, [0, 1, 0], [0, 0, -1], 1.5, 1.0, u)
        assert T == 0.0 and flag == FLAG["BR"]
        assert np.allclose(mo, [mom[0], 0, -mom[2]], atol=1e-6)


def test_rayleigh_invariants_and_dipole_distribution():
    """rayleigh_scatter_align.py:41-49 invariants plus the dipole law
    dP/dOmega ~ 1-(e.k')^2: E[(e.k')^2]=0.2, E[(k.k')^2]=0.4."""
    mom = np.array([0.0, 0.0, 1.0], np.float32)
    pol = np.array([1.0, 0.0, 0.0], np.float32)
    mo, pl = po.rayleigh(mom, pol, 123, 0, 200_000)
    mo = mo.astype(np.float64); pl = pl.astype(np.float64)
    assert np.abs(np.linalg.norm(mo, axis=1) - 1).max() < 1e-5
    assert np.abs(np.linalg.norm(pl, axis=1) - 1).max() < 1e-5
    assert np.abs(np.einsum("ij,ij->i", mo, pl)).max() < 1e-5
    assert abs(np.mean(mo[:, 0] ** 2) - 0.2) < 0.003
    assert abs(np.mean(mo[:, 2] ** 2) - 0.4) < 0.003


def test_bvh_traversal_equals_brute_force():
    w = configs.config1(num_gs=8, per_gs=500)
    W = po.World(w.geometry, w.physics)
    p0, s0 = W.simulate(w.gensteps, seed=w.seed, traverse=0)
    p1, s1 = W.simulate(w.gensteps, seed=w.seed, traverse=1)
    assert np.array_equal(s0, s1)
    assert np.array_equal(p0.view(np.uint32), p1.view(np.uint32))
    rng = np.random.default_rng(2)
    rays = np.zeros((20000, 2, 4), np.float32)
    rays[:, 0, :3] = rng.uniform(-6000, 6000, (20000, 3))
    d = rng.normal(size=(20000, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 1, :3] = d
    rays[:, 0, 3] = 0.05
    rays[:, 1, 3] = 1e6
    a = W.trace(rays, traverse=0)
    b = W.trace(rays, traverse=1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.mean(a[:, 1, 1].view(np.uint32) != 0xFFFF) > 0.99   # inside the world box


def test_config1_scintillator_physics_present():
    w = configs.config1(num_gs=20, per_gs=1000)
    W = po.World(w.geometry, w.physics)
    photon, seq = W.simulate(w.gensteps, seed=w.seed)
    fm = photon[:, 3, 3].view(np.uint32)
    for f in ("RE", "SC", "AB", "SD", "SA", "BT", "BR"):
        assert np.any(fm & FLAG[f]), f
    wl = photon[:, 2, 3]
    re = (fm & FLAG["RE"]) != 0
    # reemitted photons carry emission-spectrum wavelengths (430 +- 20 nm)
    assert abs(np.median(wl[re]) - 430) < 5
    hits = (fm & FLAG["SD"]) != 0
    ident = photon[hits, 3, 1].view(np.uint32) & 0xFFFFFF
    assert ident.min() >= 1 and ident.max() <= 2000    # sensor_identifier+1


def test_zero_and_truncation_edges():
    w = configs.config0(num_photon=2000)
    W = po.World(w.geometry, w.physics)
    p, s = W.simulate(w.gensteps, seed=1, max_bounce=1)
    nib = np.array([len(seqhis_label(a, b).split()) for a, b in s[:, 0]])
    assert nib.max() == 2
    p, s = W.simulate(w.gensteps, seed=1, max_bounce=0)
    assert np.all(s[:, 0, 0] == 3) and np.all(p[:, 0, :3] == 0)
    gs = w.gensteps.copy()
    gs.view(np.uint32)[0, 0, 3] = 0
    p, s = W.simulate(gs, seed=1)
    assert len(p) == 0
