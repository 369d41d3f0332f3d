"""oa_portable.h transcendentals and Philox, pinned by mathematics and published vectors."""
import numpy as np
import pytest

import pyoracle as po


def ulp_err(got, ref64):
    """error of float32 `got` in ulps of the correctly rounded float32 of ref64"""
    want = ref64.astype(np.float32)
    spacing = np.spacing(np.abs(want)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref64) / spacing


def test_logf_uniform_grid_within_one_ulp():
    # every value the photon RNG can produce: k * 2^-24, k = 1..2^24 (strided)
    k = np.arange(1, 1 << 24, 7, dtype=np.float64)
    x = (k * 2.0 ** -24).astype(np.float32)
    err = ulp_err(po.logf_n(x), np.log(x.astype(np.float64)))
    assert err.max() <= 1.0


def test_logf_wide_range_within_one_ulp():
    bits = np.arange(0x00800000, 0x7F800000, 4099, dtype=np.uint32)
    x = bits.view(np.float32)
    err = ulp_err(po.logf_n(x), np.log(x.astype(np.float64)))
    assert err.max() <= 1.0
    assert po.logf(1.0) == 0.0


def test_sincos_photon_domain():
    # phi = 2*pi_f*u for u on the RNG grid: absolute error below 1.2e-7 (1 ulp of 1.0)
    u = (np.arange(1, 1 << 24, 5, dtype=np.float64) * 2.0 ** -24).astype(np.float32)
    phi = (np.float32(2.0) * np.float32(np.pi) * u).astype(np.float32)
    s, c = po.sincosf_n(phi)
    p64 = phi.astype(np.float64)
    assert np.abs(s - np.sin(p64)).max() < 1.2e-7
    assert np.abs(c - np.cos(p64)).max() < 1.2e-7
    # where the result is not tiny, relative accuracy is ulp level
    big = np.abs(np.sin(p64)) > 1e-3
    assert ulp_err(s[big], np.sin(p64[big])).max() <= 2.0


def test_sincos_consistency_identities():
    x = np.linspace(-20, 20, 200001).astype(np.float32)
    s, c = po.sincosf_n(x)
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 4e-7
    s2, c2 = po.sincosf_n(-x)
    assert np.array_equal(s2, -s) or np.abs(s2 + s).max() < 1e-7


# Random123 distribution kat_vectors (Salmon et al., SC'11), philox4x32_10
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(int(v) for v in po.philox_block(ctr, *key)) == out


def test_uniform_range_and_mean():
    u = np.concatenate([po.uniforms(42, p, 64) for p in range(2000)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.005
    # 24-bit grid: every value is k * 2^-24
    assert np.all((u.astype(np.float64) * 2 ** 24) % 1 == 0)


def test_event_skipahead_shift_invariant():
    # qudarap/tests/QRngTest.py:66-72: event i+1 is event i advanced by `offset` words
    offset = 7
    for photon in (0, 1, 12345, 2 ** 33 + 5):
        e0 = po.uniforms(9, photon, 40, skip=offset, event_idx=0)
        e1 = po.uniforms(9, photon, 40, skip=offset, event_idx=1)
        assert np.array_equal(e1[:-offset], e0[offset:])


def test_streams_are_distinct_per_photon_and_seed():
    a = po.uniforms(0, 0, 16)
    b = po.uniforms(0, 1, 16)
    c = po.uniforms(1, 0, 16)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)


def test_atan2_accuracy_and_quadrants():
    rng = np.random.default_rng(11)
    ang = rng.uniform(-np.pi, np.pi, 400000)
    r = 10.0 ** rng.uniform(-3, 4, len(ang))
    y = (r * np.sin(ang)).astype(np.float32)
    x = (r * np.cos(ang)).astype(np.float32)
    got = po.atan2f_n(y, x)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    # absolute error within 3e-7 rad everywhere (floats near pi are 2.4e-7 apart),
    # and ulp level away from the zero crossings
    assert np.abs(got - ref).max() < 3e-7
    big = np.abs(ref) > 1e-2
    assert ulp_err(got[big], ref[big]).max() <= 3.0
    # octant edges and axes
    e = np.array([[0, 1], [1, 1], [1, 0], [1, -1], [0, -1], [-1, -1], [-1, 0], [-1, 1], [0, 0]], np.float32)
    g = po.atan2f_n(e[:, 0], e[:, 1])
    assert np.allclose(g, np.arctan2(e[:, 0], e[:, 1]), atol=2.5e-7)
    # lposfphi = (atan2 + pi) / (2 pi) stays in [0, 1]
    f = (got + np.float32(np.pi)) / (np.float32(2.0) * np.float32(np.pi))
    assert f.min() >= 0.0 and f.max() <= 1.0
