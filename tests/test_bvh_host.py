"""The host BVH builder (opticks_amd/csrc/oa_bvh.h) on the instance boxes of the
2,000- and 20,000-detector workloads, compiled with g++ (tests/cpp/bvh_check.cpp):
every instance in exactly one leaf, children inside their parent's slot box,
the traversal stack bound within the LDS stack, and a nearest-first walk with
the kernel's slab arithmetic reaching every box a brute-force scan says the
ray enters."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from opticks_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def instance_boxes(w):
    g = w.geometry
    prim, inst, solid = g["prim"], g["inst"], g["solid"]
    boxes = []
    for i in range(len(inst)):
        gas = int(inst[i].view(np.int32)[1, 3])
        po, npr = int(solid[gas]["primOffset"]), int(solid[gas]["numPrim"])
        pl = prim[po:po + npr].astype(np.float64)
        lo = np.array([pl[:, 2, 0].min(), pl[:, 2, 1].min(), pl[:, 2, 2].min()])
        hi = np.array([pl[:, 2, 3].max(), pl[:, 3, 0].max(), pl[:, 3, 1].max()])
        m = inst[i].astype(np.float64).copy()
        m[:, 3] = (0.0, 0.0, 0.0, 1.0)
        c = np.array([[x, y, z, 1.0] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])]) @ m
        boxes.append(np.r_[c[:, :3].min(0) - 1e-2, c[:, :3].max(0) + 1e-2])
    return np.array(boxes, np.float64)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("bvh")
    exe = d / "bvh_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "opticks_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "bvh_check.cpp"), "-o", str(exe)], check=True)
    return d, exe


@pytest.mark.parametrize("which, rays", [("config1", 3000), ("config2", 300)])
@pytest.mark.parametrize("leaf", [1, 4])
def test_bvh4_builder(checker, which, rays, leaf):
    d, exe = checker
    w = configs.config1(num_gs=1, per_gs=1) if which == "config1" else configs.config2(num_events=1)
    f = d / f"{which}.bin"
    if not f.exists():
        instance_boxes(w).tofile(f)
    out = subprocess.run([str(exe), str(f), str(leaf), str(rays)], capture_output=True, text=True, check=True).stdout
    v = dict(zip(out.split()[0::2], out.split()[1::2]))
    assert int(v["coverage"]) == 0 and int(v["bad_box"]) == 0 and int(v["missed"]) == 0, out
    assert int(v["stack_need"]) + 2 <= int(v["limit"]), out
    assert int(v["max_stack"]) <= int(v["stack_need"]), out
    assert int(v["items"]) == len(w.geometry["inst"])
