"""
The oracle's adjoint score rows (oracle_adjoint_rows) against what the
mathematics fixes, on the water-sphere geometry (point source at the centre,
sphere r=1000 mm, perfect detector outside):

* no scattering: every photon meets the sphere at normal incidence, so
  P(hit) = T x / (1 - Rf x^2), x = exp(-R/L_abs), Rf = ((n-1)/(n+1))^2, T = 1-Rf,
  and dP/dL_abs = T x (1 + Rf x^2) / (1 - Rf x^2)^2 * R / L_abs^2, split over
  the two bins of the interpolated samples with weights (1-a, a);
* no absorption: every photon is detected whatever the Rayleigh length, so the
  Rayleigh columns sum to zero (the -1/L_ray term on scatter segments cancels
  the ell/L_ray^2 survival terms);
* reemission probability 1: every absorbed photon is reemitted and detected,
  so the absorption columns sum to zero (the -1/L_abs term on absorb segments).
A dropped term, a wrong sign or a wrong interpolation weight fails one of them.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from opticks_amd import configs
from opticks_amd.adjoint import LengthParams
from opticks_amd.physics import Material

R = 1000.0
N_WATER = 1.333
SD = 0x40   # SURFACE_DETECT


def _world(water_abs, water_ray, reemission=0.0, n=200_000, wavelength=440.0, seed=42):
    w = configs.config0(num_photon=n, seed=seed, water_absorption=water_abs, water_rayleigh=water_ray)
    if reemission:
        t = Material("Water", N_WATER, water_abs, water_ray, reemission=reemission).table()
        bnd = w.physics["bnd"]
        for b in range(bnd.shape[0]):
            for j in (0, 3):
                if w.physics["optical"][b, j, 0] == w.physics["material_names"].index("Water") + 1:
                    bnd[b, j] = t
    w.gensteps[0, 3, 3] = wavelength
    return w


def _scores(w, max_bounce=31):
    world = O.World(w.geometry, w.physics)
    photon, seq = world.simulate(w.gensteps, seed=w.seed, max_bounce=max_bounce)
    flagmask = photon.view(np.uint32)[:, 3, 3]
    hit = np.nonzero(flagmask & SD)[0]
    params = LengthParams(w.physics, ("Water",), n_bin=64)
    rows = world.adjoint_rows(w.gensteps, hit, params.mat_block, params.sample_bin, params.n_bin,
                              seed=w.seed, max_bounce=max_bounce)
    full = np.zeros((len(photon), params.n_param), np.float64)
    full[hit] = rows
    return params, full, len(hit)


def _mean_and_err(full):
    n = full.shape[0]
    return full.mean(axis=0), full.std(axis=0) / np.sqrt(n)


def test_closed_form_absorption_no_scatter():
    La = 20000.0
    # 443.25 nm: fx = 383.25, a = 0.25; samples 443 nm (bin 25) and 444 nm (bin 26)
    w = _world(La, 1e12, wavelength=443.25)
    params, full, n_hit = _scores(w)
    mean, err = _mean_and_err(full)
    rf = ((N_WATER - 1.0) / (N_WATER + 1.0)) ** 2
    T = 1.0 - rf
    x = np.exp(-R / La)
    q = rf * x * x
    p_hit = T * x / (1 - q)
    dP = T * x * (1 + q) / (1 - q) ** 2 * R / La ** 2
    assert abs(n_hit / full.shape[0] - p_hit) < 5 * np.sqrt(p_hit * (1 - p_hit) / full.shape[0])
    b25, b26 = 25, 26
    assert params.sample_bin[443 - 60] == b25 and params.sample_bin[444 - 60] == b26
    for b, wgt in ((b25, 0.75), (b26, 0.25)):
        assert abs(mean[b] - wgt * dP) < 5 * err[b] + 1e-6 * dP, (b, mean[b], wgt * dP, err[b])
        assert abs(mean[b] / (wgt * dP) - 1) < 5e-3
    others = np.delete(np.arange(64), [b25, b26])
    assert np.all(mean[others] == 0.0)
    # Rayleigh columns: only ell/L^2 with L = 1e12, negligible
    assert np.abs(mean[64:]).max() < 1e-15


def test_rayleigh_identity_no_absorption():
    Ls = 3000.0
    w = _world(1e12, Ls, n=100_000)
    params, full, n_hit = _scores(w, max_bounce=1000)
    assert n_hit >= full.shape[0] - 10   # a photon may stay trapped past max_bounce
    mean, err = _mean_and_err(full)
    b = params.sample_bin[440 - 60]
    # without the -1/L term the column would sit at E[ell]/Ls^2 >= R/Ls^2
    assert abs(mean[64 + b]) < 5 * err[64 + b]
    assert 20 * err[64 + b] < R / Ls ** 2
    assert np.abs(mean[:64]).max() < 1e-15


def test_reemission_identity():
    La = 3000.0
    w = _world(La, 1e12, reemission=1.0, n=100_000)
    params, full, n_hit = _scores(w, max_bounce=1000)
    assert n_hit >= full.shape[0] - 10   # a photon may stay trapped past max_bounce
    mean, err = _mean_and_err(full)
    cols = np.nonzero(np.abs(mean[:64]) > 0)[0]
    assert len(cols) > 3   # reemitted photons spread over the emission spectrum's bins
    tot = full[:, :64].sum(axis=1)
    assert abs(tot.mean()) < 5 * tot.std() / np.sqrt(len(tot))
    for c in cols:
        assert abs(mean[c]) < 5 * err[c] + 1e-12, (c, mean[c], err[c])


def test_rows_zero_outside_parameterised_materials():
    w = _world(20000.0, 30000.0, n=2000)
    world = O.World(w.geometry, w.physics)
    params = LengthParams(w.physics, ("Vacuum",), n_bin=64)
    rows = world.adjoint_rows(w.gensteps, np.arange(100), params.mat_block, params.sample_bin, params.n_bin,
                              seed=w.seed)
    # vacuum segments exist (water -> vacuum -> detector), water ones are not scored
    assert rows.shape == (100, 128)
    assert np.all(rows[:, :64] >= 0.0)


@pytest.mark.parametrize("bad", [[-1], [10**9]])
def test_rows_reject_bad_photon_index(bad):
    w = _world(20000.0, 30000.0, n=100)
    world = O.World(w.geometry, w.physics)
    params = LengthParams(w.physics, ("Water",), n_bin=64)
    with pytest.raises(RuntimeError):
        world.adjoint_rows(w.gensteps, np.array(bad), params.mat_block, params.sample_bin, params.n_bin)
