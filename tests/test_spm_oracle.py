"""The hit-merge restatement (oracle/spm.py) on hand-built records whose merged
result is worked out here by hand from SPM::merge_partial_select
(sysrap/SPM.cu:135-240) and the sphoton / sphotonlite functors
(sysrap/sphoton.h:292-320, sysrap/sphotonlite.h:68-97)."""
import numpy as np

import spm

SD, SA, BT = 0x40, 0x80, 0x800


def photon(ident, time, flagmask, hitcount=1, iindex=7, x=0.0):
    p = np.zeros((4, 4), np.float32)
    u = p.view(np.uint32)
    p[0] = (x, 1.0, 2.0, time)
    u[1, 3] = (hitcount << 16) | iindex
    u[3, 1] = ident
    u[3, 3] = flagmask
    return p


def lite(ident, time, flagmask, hitcount=1, lpos=0):
    return np.array([(hitcount << 16) | ident, np.float32(time).view(np.uint32), lpos, flagmask], np.uint32)


def test_photon_merge_by_identity_and_bucket():
    P = np.stack([photon(1, 0.5, SD | BT, x=0), photon(1, 0.9, SD, x=1), photon(2, 0.2, SD, x=2),
                  photon(1, 1.5, SD, x=3), photon(3, 0.1, SA, x=4)])
    m = spm.merge_photons(P, SD, 1.0)
    u = m.view(np.uint32)
    # keys (1,0) <- records 0,1 ; (1,1) <- record 3 ; (2,0) <- record 2 ; record 4 not selected
    assert len(m) == 3
    assert list(m[:, 0, 0]) == [0.0, 3.0, 2.0]
    assert list(u[:, 1, 3] >> 16) == [2, 1, 1]             # hitcounts
    assert list(u[:, 1, 3] & 0xFFFF) == [7, 7, 7]          # iindex kept
    assert list(u[:, 3, 3]) == [SD | BT, SD, SD]           # OR of flagmasks
    assert list(m[:, 0, 3]) == [np.float32(0.5), np.float32(1.5), np.float32(0.2)]


def test_photon_earliest_record_wins_and_ties_keep_first():
    P = np.stack([photon(5, 0.7, SD, x=0), photon(5, 0.3, SD | BT, x=1), photon(5, 0.3, SD, x=2)])
    m = spm.merge_photons(P, SD, 1.0)
    assert len(m) == 1
    assert m[0, 0, 0] == 1.0                               # earliest, first of the tie
    assert (m.view(np.uint32)[0, 1, 3] >> 16) == 3
    assert m.view(np.uint32)[0, 3, 3] == SD | BT


def test_photon_hitcount_is_a_16_bit_field():
    P = np.stack([photon(9, 0.1, SD, hitcount=0xFFFF), photon(9, 0.2, SD, hitcount=2)])
    m = spm.merge_photons(P, SD, 1.0)
    assert (m.view(np.uint32)[0, 1, 3] >> 16) == 1         # (0xffff + 2) & 0xffff


def test_photon_key_keeps_low_16_bits_of_identity():
    # identity << 48 in 64 bits drops all but the low 16 bits of the 24-bit identity
    P = np.stack([photon(0x10001, 0.1, SD, x=0), photon(0x00001, 0.2, SD, x=1)])
    m = spm.merge_photons(P, SD, 1.0)
    assert len(m) == 1 and (m.view(np.uint32)[0, 1, 3] >> 16) == 2


def test_window_zero_selects_only_and_already_selected_skips_selection():
    P = np.stack([photon(1, 0.5, SD, x=0), photon(1, 0.6, SA, x=1), photon(1, 0.7, SD, x=2)])
    m = spm.merge_photons(P, SD, 0.0)
    assert list(m[:, 0, 0]) == [0.0, 2.0]
    m = spm.merge_photons(P, spm.ALREADY_SELECTED, 10.0)
    assert len(m) == 1 and (m.view(np.uint32)[0, 1, 3] >> 16) == 3
    assert m.view(np.uint32)[0, 3, 3] == SD | SA


def test_lite_merge_keeps_first_record_fields_and_min_time():
    L = np.stack([lite(4, 0.8, SD, lpos=111), lite(4, 0.2, SD | BT, lpos=222), lite(2, 3.5, SD, lpos=5)])
    m = spm.merge_lites(L, SD, 1.0)
    assert m.shape == (2, 4)
    # key order: (2, 3) < (4, 0)
    assert list(m[:, 0] & 0xFFFF) == [2, 4]
    assert list(m[:, 0] >> 16) == [1, 2]
    assert list(m[:, 2]) == [5, 111]                       # lpos of the run's first record
    assert m[1, 1].view(np.float32) == np.float32(0.2)     # earliest time
    assert m[1, 3] == SD | BT


def test_merge_invariants_random():
    rng = np.random.default_rng(3)
    n = 2000
    P = np.stack([photon(int(rng.integers(1, 40)), float(rng.integers(0, 50)) * 0.25,
                         SD if rng.random() < 0.7 else SA, x=i) for i in range(n)])
    m = spm.merge_photons(P, SD, 1.0)
    u = P.view(np.uint32)
    sel = (u[:, 3, 3] & SD) != 0
    # hitcounts are conserved and one record per distinct (identity, bucket)
    assert (m.view(np.uint32)[:, 1, 3] >> 16).sum() == sel.sum()
    pairs = {(int(u[i, 3, 1]), int(P[i, 0, 3] // 1.0)) for i in np.nonzero(sel)[0]}
    assert len(m) == len(pairs)
    k = (m.view(np.uint32)[:, 3, 1].astype(np.int64) << 32) | (m[:, 0, 3] // 1.0).astype(np.int64)
    assert np.all(np.diff(k) > 0)                          # key order, no duplicates
