"""C-ABI library loads and exports every declared symbol; host-side logic."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from opticks_amd import _lib as L
from opticks_amd import configs
from opticks_amd.qsim import SEventConfig, seqhis_label, flag_mask, FLAG
from opticks_amd.foundry import CSG_DIFFERENCE, CSG_UNION, CSG_ZSPHERE, CSG_CYLINDER, CSG_ZERO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "opticks_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mi_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SYMBOLS)


def test_defaults_and_errors_without_compute():
    c = L.MiConfig()
    L.lib().mi_config_default(C.byref(c))
    assert c.max_bounce == 31 and abs(c.tmin - 0.05) < 1e-9 and c.hit_mask == FLAG["SD"]
    assert c.eps0_mask == flag_mask("TO,CK,SI,SC,RE") and c.event_skipahead == 100000
    assert L.lib().mi_version().decode().startswith("opticks_amd")
    # bad arguments are rejected before any device call
    assert L.lib().mi_sim_create(None, None, C.byref(c), None) < 0
    assert L.last_error()
    h = C.c_void_p()
    assert L.lib().mi_sim_create(None, None, C.byref(c), C.byref(h)) < 0 and not h.value
    bad = L.MiConfig(); L.lib().mi_config_default(C.byref(bad)); bad.block_size = 96
    assert L.lib().mi_sim_create(None, None, C.byref(bad), C.byref(h)) == -1
    assert "block_size" in L.last_error()
    assert c.trace_regroup_pct == 0 and c.refill_batch == 16
    bad.block_size = 64; bad.trace_regroup_pct = 101
    assert L.lib().mi_sim_create(None, None, C.byref(bad), C.byref(h)) == -1
    assert "trace_regroup_pct" in L.last_error()
    bad.trace_regroup_pct = 0; bad.refill_batch = 65
    assert L.lib().mi_sim_create(None, None, C.byref(bad), C.byref(h)) == -1
    assert "refill_batch" in L.last_error()


def test_seqhis_labels_and_masks():
    assert seqhis_label(0x7c3) == "TO BT SD"      # nibbles 3, 0xc, 7
    assert seqhis_label(0x4 << 0 | 0, 0) == "AB"
    assert flag_mask("SD") == 0x40 and flag_mask("TO,BT") == 0x804


def test_config_from_env(monkeypatch):
    monkeypatch.setenv("OPTICKS_MAX_BOUNCE", "9")
    monkeypatch.setenv("QRng__SEED_OFFSET", "7:3")
    c = SEventConfig.from_env()
    assert c.max_bounce == 9 and c.seed == 7 and c.rng_offset == 3


def test_foundry_layout_of_detector_tree():
    w = configs.config1(num_gs=1, per_gs=1)
    g = w.geometry
    prim = g["prim"].view(np.int32)
    num, off = prim[3, 0, 0], prim[3, 0, 1]
    assert num == 7
    nodes = g["node"][off:off + 7].view(np.uint32)
    assert list(nodes[:, 3, 2]) == [CSG_DIFFERENCE, CSG_UNION, CSG_CYLINDER, CSG_ZSPHERE, CSG_CYLINDER,
                                    CSG_ZERO, CSG_ZERO]
    assert nodes[0, 0, 0] == 7                                  # subNum on the root
    assert (nodes[3, 3, 3] & 0x7FFFFFFF) == 1                   # ellipsoid scale transform
    assert len(g["inst"]) == 2001
    ident = g["inst"].view(np.int32)[:, 2, 3]
    assert ident[0] == 0 and ident[1] == 1 and ident[-1] == 2000
    # detectors point their +z at the centre
    z_axis = g["inst"][1:, 2, :3]
    pos = g["inst"][1:, 3, :3]
    cosang = np.einsum("ij,ij->i", z_axis, -pos / np.linalg.norm(pos, axis=1, keepdims=True))
    assert np.allclose(cosang, 1.0, atol=1e-5)


def test_physics_tables():
    w = configs.config1(num_gs=1, per_gs=1)
    ph = w.physics
    assert ph["bnd"].shape == (4, 4, 2, 761, 4)
    assert ph["optical"].shape == (4, 4, 4)
    ls = ph["boundary_names"].index("Acrylic///LS")
    wl = 60 + np.arange(761)
    reem = ph["bnd"][ls, 3, 0, :, 3]
    assert np.all(reem[wl < 410] == np.float32(0.4)) and np.all(reem[wl >= 410] == 0)
    assert np.allclose(ph["bnd"][ls, 3, 1, :, 0], 299.792458 / 1.49)
    assert ph["optical"][ls, 1, 1] == 1 and ph["optical"][ls, 3, 1] == 0   # NoSurface / material
    det = ph["boundary_names"].index("Water/detSurf//Vacuum")
    assert ph["optical"][det, 1, 1] == 2 and np.allclose(ph["bnd"][det, 1, 0, :, :2], [0.25, 0.75])
    icdf = ph["icdf"]
    assert icdf.shape == (3, 4096) and np.all(np.diff(icdf[0]) >= 0)
    assert abs(np.median(icdf[0]) - 430) < 0.1


def test_cpp_adaptor_compiles_and_reports_errors(tmp_path):
    """MI355XSimulator.hh builds with g++ against the library; a bad geometry
    surfaces as a C++ exception carrying mi_last_error()."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "adaptor.cpp"
    src.write_text(r'''
#include <cstdio>
#include "MI355XSimulator.hh"
int main() {
    mi_geometry g = {}; mi_physics p = {};
    std::printf("%s\n", mi_version());
    try { MI355XSimulator s(g, p); std::printf("unexpected\n"); return 1; }
    catch (const std::exception& e) { std::printf("caught: %s\n", e.what()); }
    return 0;
}
''')
    exe = tmp_path / "adaptor"
    lib = os.path.join(ROOT, "opticks_amd")
    cmd = ["g++", "-std=c++17", "-I", os.path.join(ROOT, "opticks_amd", "csrc"), str(src), "-o", str(exe),
           "-L", lib, "-lopticks_amd", f"-Wl,-rpath,{lib}", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert out.startswith("opticks_amd") and "caught: mi_sim_create:" in out


@pytest.mark.parametrize("writer", ["numpy", "reference NP.hh"])
def test_cpp_npy_reader_and_folder_load(tmp_path, writer):
    """mi_npy.hh reads the reference-layout arrays as numpy wrote them and as the
    reference's own NP.hh wrote them (tests/golden/np_foundry, bnd in double
    precision, tests/golden/make_np_fixture.py), and MI355XSimulator::Load passes
    them to mi_sim_create: on a host without a GPU the geometry validates and
    creation stops at the device query."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    from opticks_amd.loader import save_csgfoundry
    w = configs.leaf_zoo(num_photon=10)
    if writer == "numpy":
        cf = save_csgfoundry(str(tmp_path / "geom"), w.geometry, w.physics)
        base = str(tmp_path / "geom")
    else:
        base = os.path.join(ROOT, "tests", "golden", "np_foundry")
        cf = os.path.join(base, "CSGFoundry")
    src = tmp_path / "npy.cpp"
    src.write_text(r'''
#include <cstdio>
#include "MI355XSimulator.hh"
int main(int argc, char** argv) {
    std::string cf = argv[1];
    const char* names[] = {"solid", "prim", "node", "plan", "inst", "SSim/stree/standard/bnd",
                           "SSim/stree/standard/optical", "SSim/stree/standard/icdf"};
    for (const char* n : names) {
        mi_npy::Array a = mi_npy::load(cf + "/" + n + ".npy");
        unsigned long long sum = 0;        /* checksum of the 32-bit words as read */
        if (a.kind == 'f') for (float v : a.as_f32()) { unsigned u; std::memcpy(&u, &v, 4); sum += u; }
        else for (int32_t v : a.as_i32()) sum += (unsigned)v;
        std::printf("%s %c%d %lld %llu\n", n, a.kind, a.itemsize, (long long)a.size(), sum);
    }
    std::printf("domain_low %g\n", mi_npy::meta_value(cf + "/SSim/stree/standard/bnd_meta.txt", "domain_low", -1));
    try { delete MI355XSimulator::Load(argv[2]); std::printf("created\n"); }
    catch (const std::exception& e) { std::printf("caught: %s\n", e.what()); }
    return 0;
}
''')
    exe = tmp_path / "npy"
    lib = os.path.join(ROOT, "opticks_amd")
    cmd = ["g++", "-std=c++17", "-I", os.path.join(ROOT, "opticks_amd", "csrc"), str(src), "-o", str(exe),
           "-L", lib, "-lopticks_amd", f"-Wl,-rpath,{lib}", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe), cf, base], capture_output=True, text=True, check=True).stdout
    lines = dict(l.split(" ", 1) for l in out.strip().splitlines())
    ref = {"solid": w.geometry["solid"].view(np.int32), "prim": w.geometry["prim"], "node": w.geometry["node"],
           "plan": w.geometry["plan"], "inst": w.geometry["inst"], "SSim/stree/standard/bnd": w.physics["bnd"],
           "SSim/stree/standard/optical": w.physics["optical"], "SSim/stree/standard/icdf": w.physics["icdf"]}
    for k, a in ref.items():
        kind, n, s = lines[k].split()
        assert int(n) == a.size, k
        assert int(s) == int(np.ascontiguousarray(a).view(np.uint32).astype(np.uint64).sum()), k
    assert lines["domain_low"] == "60"
    assert "created" in out or "no HIP device" in out
