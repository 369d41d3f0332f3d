"""The no-allocation entry point mi_sim_simulate_device_noalloc (MI355X only),
called through the C-ABI: two calls in a row, the second with fewer gensteps and
photons, each equal bit for bit to mi_sim_simulate_device on the same inputs
(photons, seq, hits, hit count), with the Philox and xorwow streams and both
kernels; sizes beyond the scratch limits are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from opticks_amd import configs
from opticks_amd import _lib as L
from opticks_amd.qsim import QSim, SEventConfig

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

DEV = torch.device("cuda:0")


def call(fn, sim, gs_t, n, event_id, photon_offset=0):
    photon = torch.full((n, 4, 4), -1.0, dtype=torch.float32, device=DEV)
    seq = torch.full((n, 2, 2), -1, dtype=torch.int64, device=DEV)
    hit = torch.full((n, 4, 4), -1.0, dtype=torch.float32, device=DEV)
    n_hit = C.c_int64(-1)
    rc = fn(sim.handle, event_id, gs_t.data_ptr(), gs_t.shape[0], -1, photon_offset, photon.data_ptr(),
            seq.data_ptr(), hit.data_ptr(), n, C.byref(n_hit), None)
    L.check(rc, fn.__name__)
    return photon.cpu().numpy(), seq.cpu().numpy(), hit[:n_hit.value].cpu().numpy(), n_hit.value


@pytest.mark.parametrize("rng,kernel", [("philox", L.MI_KERNEL_WAVEFRONT), ("philox", L.MI_KERNEL_PLAIN),
                                        ("xorwow", L.MI_KERNEL_WAVEFRONT)])
def test_noalloc_equals_simulate_device_twice(rng, kernel):
    w = configs.config1(num_gs=40, per_gs=2500)
    sim = QSim(w.geometry, w.physics, SEventConfig(seed=w.seed, rng=rng, kernel=kernel))
    lib = L.lib()
    for gs, event_id, off in ((w.gensteps, 3, 0), (w.gensteps[5:17], 4, 1000)):
        gs_t = torch.from_numpy(np.ascontiguousarray(gs)).to(DEV)
        n = int(gs.view(np.uint32)[:, 0, 3].sum())
        print("noalloc", rng, kernel, "gensteps", len(gs), "photons", n, flush=True)
        a = call(lib.mi_sim_simulate_device_noalloc, sim, gs_t, n, event_id, off)
        b = call(lib.mi_sim_simulate_device, sim, gs_t, n, event_id, off)
        assert a[3] == b[3] > 100
        assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
    sim.close()


def test_noalloc_limits_rejected():
    w = configs.config0(num_photon=1000)
    sim = QSim(w.geometry, w.physics, SEventConfig())
    lib = L.lib()
    gs_t = torch.from_numpy(w.gensteps).to(DEV)
    n_hit = C.c_int64(0)
    rc = lib.mi_sim_simulate_device_noalloc(sim.handle, 0, gs_t.data_ptr(), (1 << 20) + 1, -1, 0, None, None,
                                            None, 0, C.byref(n_hit), None)
    assert rc == -4 and "2^20" in L.last_error()
    rc = lib.mi_sim_simulate_device_noalloc(sim.handle, 0, gs_t.data_ptr(), 1, (1 << 26) + 1, 0, None, None,
                                            None, 0, C.byref(n_hit), None)
    assert rc == -4 and "2^26" in L.last_error()
    sim.close()
