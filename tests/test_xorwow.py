"""The xorwow generator of oa_portable.h (shared by the kernels and the checker).

Recurrence: G. Marsaglia, "Xorshift RNGs" (2003), the generator behind the
reference's qrng<XORWOW> (qudarap/qrng.h:68-106, sysrap/srng.h:30-36). Pinned by:
  * rocRAND's open-source xorwow engine (its next() and its discard() jump
    tables) on explicitly set states: words and jumped states must be equal;
  * brute force: a jump by n equals n single steps;
  * the reference's event-skipahead structure (qudarap/tests/QRngTest.py:66-72):
    event e's stream is event 0's stream shifted by e * skipahead words;
  * this project's per-photon seeding: distinct photons start in distinct
    states, and uniforms fill (0, 1] with the moments of U(0, 1).
curand's seeding constants and skipahead tables are not reproduced (parity
unpinned for curand's stream values)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
ROCRAND_H = "/opt/rocm/include/rocrand/rocrand_xorwow.h"


def words_from_state(state6, skip, n):
    s = np.ascontiguousarray(state6, np.uint32)
    out = np.zeros(n, np.uint32)
    st = np.zeros(6, np.uint32)
    pyoracle.lib().oracle_xorwow_from_state(s.ctypes.data, C.c_uint64(skip), n, out.ctypes.data, st.ctypes.data)
    return st, out


def stream(seed, offset, skip, event, photon, n):
    out = np.zeros(n, np.uint32)
    st = np.zeros(6, np.uint32)
    pyoracle.lib().oracle_xorwow_words(C.c_uint64(seed), C.c_uint64(offset), C.c_uint64(skip), C.c_uint64(event),
                                       C.c_uint64(photon), n, out.ctypes.data, st.ctypes.data)
    return st, out


def step_py(s):
    """one xorwow step in Python integers (Marsaglia 2003, section 3.1)"""
    x0, x1, x2, x3, x4, d = (int(v) for v in s)
    t = (x0 ^ (x0 >> 2)) & 0xFFFFFFFF
    x4n = ((x4 ^ (x4 << 4)) ^ (t ^ (t << 1))) & 0xFFFFFFFF
    d = (d + 362437) & 0xFFFFFFFF
    return [x1, x2, x3, x4, x4n, d], (d + x4n) & 0xFFFFFFFF


@pytest.fixture(scope="module")
def rocrand_probe(tmp_path_factory):
    if not (os.path.exists(HIPCC) and os.path.exists(ROCRAND_H)):
        pytest.skip("hipcc or rocRAND headers absent")
    d = tmp_path_factory.mktemp("xw")
    exe = str(d / "xorwow_rocrand")
    subprocess.run([HIPCC, "--offload-host-only", "-O1", os.path.join(ROOT, "tests", "cpp", "xorwow_rocrand.cpp"),
                    "-o", exe], check=True, capture_output=True)

    def run(state6, skip, n):
        r = subprocess.run([exe, *[str(int(v)) for v in state6], str(skip), str(n)], check=True,
                           capture_output=True, text=True)
        lines = r.stdout.split("\n")
        return np.array(lines[0].split(), np.uint32), np.array([int(v) for v in lines[1:1 + n]], np.uint32)
    return run


def test_recurrence_matches_python_and_rocrand(rocrand_probe):
    rng = np.random.default_rng(7)
    for _ in range(4):
        s = rng.integers(0, 2**32, 6, dtype=np.uint64).astype(np.uint32)
        _, ours = words_from_state(s, 0, 64)
        _, theirs = rocrand_probe(s, 0, 64)
        assert np.array_equal(ours, theirs)
        st = list(s)
        py = []
        for _ in range(64):
            st, w = step_py(st)
            py.append(w)
        assert np.array_equal(ours, np.array(py, np.uint32))


@pytest.mark.parametrize("skip", [1, 2, 3, 5, 37, 1000, 100000, 123456789, 2**40 + 17, 2**63 + 5])
def test_jump_matches_rocrand_discard(rocrand_probe, skip):
    rng = np.random.default_rng(skip % 1000)
    s = rng.integers(0, 2**32, 6, dtype=np.uint64).astype(np.uint32)
    st, ours = words_from_state(s, skip, 8)
    st_r, theirs = rocrand_probe(s, skip, 8)
    assert np.array_equal(st, st_r)
    assert np.array_equal(ours, theirs)


def test_jump_equals_single_steps():
    s = np.array([0x12345678, 0x9abcdef0, 0x0fedcba9, 0x87654321, 0x13579bdf, 0x2468ace0], np.uint32)
    _, long = words_from_state(s, 0, 5000)
    for n in (1, 4, 17, 255, 256, 1000, 4999):
        _, w = words_from_state(s, n, 1)
        assert w[0] == long[n]
    # the jump matrix of 0 words is the identity, of 1 word the step matrix (columns = images of bits)
    J = np.zeros(800, np.uint32)
    pyoracle.lib().oracle_xorwow_jump(C.c_uint64(0), J.ctypes.data)
    eye = np.zeros((160, 5), np.uint32)
    for j in range(160):
        eye[j, j // 32] = 1 << (j % 32)
    assert np.array_equal(J.reshape(160, 5), eye)


def test_event_skipahead_shift_invariant():
    """QRngTest.py:66-72: stream(event e) = stream(event 0) shifted by e * skipahead"""
    skip = 1000
    for photon in (0, 1, 99999, 2**33 + 5):
        _, e0 = stream(42, 0, skip, 0, photon, 3 * skip + 16)
        for e in (1, 2, 3):
            _, ee = stream(42, 0, skip, e, photon, 16)
            assert np.array_equal(ee, e0[e * skip:e * skip + 16])
        _, off = stream(42, 7, skip, 1, photon, 16)       # offset adds to the event skip
        assert np.array_equal(off, e0[skip + 7:skip + 23])


def test_seeding_distinct_and_uniform():
    starts = np.array([stream(42, 0, 100000, 0, p, 0)[0] for p in range(4000)])
    assert len({tuple(r) for r in starts[:, :5]}) == 4000
    assert not np.array_equal(stream(42, 0, 1, 0, 5, 8)[1], stream(43, 0, 1, 0, 5, 8)[1])
    w = np.concatenate([stream(42, 0, 1, 0, p, 256)[1] for p in range(200)])
    u = ((w >> 8).astype(np.float64) + 1.0) * 2.0**-24
    assert u.min() > 0.0 and u.max() <= 1.0
    n = len(u)
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / n)
    assert abs((u**2).mean() - 1 / 3) < 5 * np.sqrt(4 / 45 / n)
    # bits are balanced across the 32 positions
    bits = ((w[:, None] >> np.arange(32, dtype=np.uint32)) & 1).mean(axis=0)
    assert np.all(np.abs(bits - 0.5) < 5 * np.sqrt(0.25 / n))
