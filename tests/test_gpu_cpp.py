"""The C++ SSimulator adaptor (opticks_amd/csrc/MI355XSimulator.hh) end to end on
the MI355X, used the way a Geant4-side caller uses the reference's CSGOptiX
simulator (g4cx/G4CXOpticks.cc:470-489): a CSGFoundry folder written in the
reference layout is loaded by MI355XSimulator::Load, gensteps are added,
simulate(eventID) runs, and the photons, seq and hits it returns equal the
checker's bit for bit; reset clears the event."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle as po
from opticks_amd import configs
from opticks_amd.loader import save_csgfoundry

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include <cstdio>
#include <cstdlib>
#include "MI355XSimulator.hh"
static void dump(const char* path, const void* p, size_t bytes)
{
    FILE* f = std::fopen(path, "wb");
    if (bytes) std::fwrite(p, 1, bytes, f);
    std::fclose(f);
}
int main(int argc, char** argv)
{
    mi_config cfg; mi_config_default(&cfg);
    cfg.seed = std::strtoull(argv[3], nullptr, 10);
    MI355XSimulator* sim = MI355XSimulator::Load(argv[1], &cfg);     /* the SSimulator method set */
    MI355XSimulator* s = sim;
    const std::vector<float> gs = mi_npy::load(argv[2]).as_f32();
    s->add_gensteps(gs.data(), (int64_t)gs.size() / 24);
    const double ms = sim->simulate(0, false);
    const std::string out = argv[4];
    dump((out + "/photon.bin").c_str(), s->photon(), (size_t)s->num_photon() * 64);
    dump((out + "/seq.bin").c_str(), s->seq(), (size_t)s->num_photon() * 32);
    dump((out + "/hit.bin").c_str(), s->hit(), (size_t)s->num_hit() * 64);
    std::printf("%lld %lld %d %s\n", (long long)s->num_photon(), (long long)s->num_hit(), ms > 0.0 ? 1 : 0, sim->desc());
    sim->reset(0);
    std::printf("after reset %lld %lld %lld\n", (long long)s->num_genstep(), (long long)s->num_photon(),
                (long long)s->num_hit());
    delete sim;
    return 0;
}
'''


def test_cpp_adaptor_simulates_like_the_checker(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    w = configs.config1(num_gs=4, per_gs=2500)
    save_csgfoundry(str(tmp_path / "geom"), w.geometry, w.physics)
    np.save(tmp_path / "gs.npy", w.gensteps)
    src = tmp_path / "run.cpp"
    src.write_text(SRC)
    exe = tmp_path / "run"
    lib = os.path.join(ROOT, "opticks_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "opticks_amd", "csrc"), str(src), "-o", str(exe),
                    "-L", lib, "-lopticks_amd", f"-Wl,-rpath,{lib}", "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    print("running the C++ adaptor", flush=True)
    r = subprocess.run([str(exe), str(tmp_path / "geom"), str(tmp_path / "gs.npy"), str(w.seed), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    first, second = r.stdout.strip().splitlines()[-2:]
    n_photon, n_hit, timed = (int(v) for v in first.split()[:3])
    assert n_photon == 10000 and timed == 1 and "MI355X" in first
    assert second.split()[2:] == ["0", "0", "0"]
    photon = np.fromfile(tmp_path / "photon.bin", np.float32).reshape(-1, 4, 4)
    seq = np.fromfile(tmp_path / "seq.bin", np.uint64).reshape(-1, 2, 2)
    hit = np.fromfile(tmp_path / "hit.bin", np.float32).reshape(-1, 4, 4)
    ref_p, ref_s = po.World(w.geometry, w.physics).simulate(w.gensteps, seed=w.seed)
    assert np.array_equal(seq, ref_s.view(np.uint64).reshape(seq.shape))
    assert np.array_equal(photon.view(np.uint32), ref_p.view(np.uint32))
    ref_hit = ref_p[(ref_p[:, 3, 3].view(np.uint32) & 0x40) != 0]
    assert n_hit == len(ref_hit) > 100 and np.array_equal(hit.view(np.uint32), ref_hit.view(np.uint32))


def _check_outputs(out_dir, w, seed, n_photon):
    photon = np.fromfile(out_dir / "photon.bin", np.float32).reshape(-1, 4, 4)
    seq = np.fromfile(out_dir / "seq.bin", np.uint64).reshape(-1, 2, 2)
    hit = np.fromfile(out_dir / "hit.bin", np.float32).reshape(-1, 4, 4)
    assert len(photon) == n_photon
    ref_p, ref_s = po.World(w.geometry, w.physics).simulate(w.gensteps, seed=seed)
    assert np.array_equal(seq, ref_s.view(np.uint64).reshape(seq.shape))
    assert np.array_equal(photon.view(np.uint32), ref_p.view(np.uint32))
    ref_hit = ref_p[(ref_p[:, 3, 3].view(np.uint32) & 0x40) != 0]
    assert np.array_equal(hit.view(np.uint32), ref_hit.view(np.uint32))


def test_ssimulator_protocol_caller_on_reference_written_folder(tmp_path):
    """oracle/_ref/ssimulator_check (built here from the reference's sysrap/SSimulator.h,
    tests/cpp/ssimulator_check.cpp) drives the path through an SSimulator pointer on
    the folder the reference's NP.hh wrote (tests/golden/np_foundry)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ssimulator_check")
    if not os.access(exe, os.X_OK):
        pytest.skip("oracle/_ref/ssimulator_check not built (needs the reference checkout at build time)")
    w = configs.leaf_zoo(num_photon=4000)
    np.save(tmp_path / "gs.npy", w.gensteps)
    print("running the SSimulator caller", flush=True)
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "np_foundry"), str(tmp_path / "gs.npy"),
                        str(w.seed), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    n_photon, n_hit, timed = (int(v) for v in lines[-3].split()[:3])
    assert n_photon == w.num_photon and timed == 1 and "MI355X" in lines[-3]
    assert lines[-2].split()[2:] == ["0", "0", "0"]
    assert lines[-1] == "render -1 -1"
    _check_outputs(tmp_path, w, w.seed, n_photon)
