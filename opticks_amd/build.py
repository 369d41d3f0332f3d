"""
Build libopticks_amd.so in-tree with hipcc for gfx950.

  python -m opticks_amd.build            # the library
  python -m opticks_amd.build --stats    # OA_STATS counters build (one translation unit)

The propagation kernels (csrc/oa_sim.hip) are compiled once per thread-block
size (64, 128, 256) and random generator (Philox, precooked, xorwow), each
compile holding only that pair's kernels; these and the other sources are
compiled in parallel, then linked.

Flags: -ffp-contract=off keeps a*b+c unfused so results match the CPU checker
bit for bit; packed FP32 instructions are disabled (-packed-fp32-ops); the
AMDGPU machine scheduler runs its iterative ILP strategy, which leaves the
propagation kernel fewer register spills under its 168-VGPR budget (56 instead
of 85 spilled VGPRs) and measures 2% faster; instruction order only, the
arithmetic and so the results are the same.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libopticks_amd.so")
OUT_STATS = os.path.join(HERE, "libopticks_amd_stats.so")   # OA_STATS counters build (opticks_amd.workcount)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-sched-strategy=iterative-ilp",
         "-Wall", "-Wno-unused-function"]

BLOCKS = (64, 128, 256)
RNGS = (0, 1, 2)    # Philox, precooked, xorwow


def units(stats=False):
    """(source, extra flags, object name) of every compile."""
    u = [("oa_merge.hip", [], "oa_merge.o"), ("oa_capi.cpp", [], "oa_capi.o")]
    if stats:
        u.append(("oa_kernels.hip", ["-DOA_STATS", "-DOA_SIM_ALL"], "oa_kernels.o"))
    else:
        u.append(("oa_kernels.hip", [], "oa_kernels.o"))
        u += [("oa_sim.hip", [f"-DOA_SIM_BLOCK={b}", f"-DOA_SIM_RNG={k}"], f"oa_sim_{b}_{k}.o")
              for b in BLOCKS for k in RNGS]
    return u


def _filter(stderr):
    return [l for l in (stderr or "").splitlines() if "packed-fp32-ops" not in l and l.strip()]


def build(verbose=False, stats=False, jobs=None):
    tmpdir = tempfile.mkdtemp(prefix="oa_build_")
    try:
        jobs = jobs or min(16, os.cpu_count() or 4)

        def compile_one(u):
            src, extra, obj = u
            cmd = [HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", os.path.join(tmpdir, obj)]
            r = subprocess.run(cmd, capture_output=True, text=True)
            return cmd, r

        objs = []
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for (cmd, r), u in zip(ex.map(compile_one, units(stats)), units(stats)):
                errs = _filter(r.stderr)
                if r.returncode != 0:
                    sys.stderr.write("\n".join(errs) + "\n")
                    raise RuntimeError("hipcc failed: " + " ".join(cmd))
                if verbose and errs:
                    sys.stderr.write("\n".join(errs) + "\n")
                objs.append(os.path.join(tmpdir, u[2]))
        out = OUT_STATS if stats else OUT
        tmp = out + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write("\n".join(_filter(r.stderr)) + "\n")
            raise RuntimeError("link failed: " + " ".join(cmd))
        os.replace(tmp, out)
        return out
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose=True, stats="--stats" in sys.argv))
