"""The drop-in class SMI355XSimulator (opticks_amd/csrc/SMI355XSimulator.hh) against
the reference's own protocol header sysrap/SSimulator.h:16-34, when the reference
checkout is present (CPU only: compilation and linking).

* tests/cpp/ssimulator_check.cpp, the caller that drives the path only through an
  SSimulator pointer, compiles with -Wall -Wextra -Werror against the reference
  header: SMI355XSimulator derives from SSimulator, overrides every pure virtual
  (static_assert !is_abstract) with the reference's exact signatures (override);
* the type check bites: a class with one drifted signature (simulate(int) without
  the reset flag, or a non-const desc) fails to compile against the same header;
* oracle/Makefile's ref target links the caller into oracle/_ref/ssimulator_check,
  which the GPU tests run (tests/test_gpu_cpp.py).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("OPTICKS_REFERENCE", "/root/reference")
HDR = os.path.join(REF, "sysrap", "SSimulator.h")

pytestmark = pytest.mark.skipif(not os.path.exists(HDR) or shutil.which("g++") is None,
                                reason="reference checkout (sysrap/SSimulator.h) or g++ absent")


def compile_only(src_path):
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                           "-I", os.path.join(REF, "sysrap"), "-I", os.path.join(ROOT, "opticks_amd", "csrc"),
                           str(src_path)], capture_output=True, text=True)


def test_caller_compiles_against_the_reference_protocol():
    r = compile_only(os.path.join(ROOT, "tests", "cpp", "ssimulator_check.cpp"))
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("drift", ["double simulate(int eventID) override { return 0; }",
                                   "const char* desc() override { return \"\"; }"])
def test_signature_drift_fails_to_compile(tmp_path, drift):
    good = {"simulate": "double simulate(int, bool) override { return 0; }",
            "desc": "const char* desc() const override { return \"\"; }"}
    body = [good["simulate"] if "simulate" not in drift else drift,
            good["desc"] if "desc" not in drift else drift,
            "double render_launch() override { return 0; }", "double simtrace_launch() override { return 0; }",
            "double simulate_launch() override { return 0; }", "double launch() override { return 0; }",
            "double simtrace(int) override { return 0; }", "double render(const char*) override { return 0; }",
            "void reset(int) override {}"]
    src = tmp_path / "drift.cpp"
    src.write_text('#include "SSimulator.h"\nstruct D : SSimulator {\n' + "\n".join(body) + "\n};\n")
    assert compile_only(src).returncode != 0
    ok = tmp_path / "ok.cpp"
    ok.write_text('#include "SSimulator.h"\nstruct D : SSimulator {\n' + "\n".join(
        [good["simulate"], good["desc"]] + body[2:]) + "\n};\nint main() { D d; SSimulator* s = &d; "
        "return (int)s->launch(); }\n")
    assert compile_only(ok).returncode == 0


def test_reference_build_links_the_caller():
    lib = os.path.join(ROOT, "opticks_amd", "libopticks_amd.so")
    if not os.path.exists(lib):
        pytest.skip("libopticks_amd.so not built")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True, capture_output=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "ssimulator_check")
    assert os.access(exe, os.X_OK)
