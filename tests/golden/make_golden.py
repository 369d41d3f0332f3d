"""
Writes the golden fixtures of tests/golden from the reference checkout:

  csg_golden.npz  rays against the prims of the 2,000-detector workload geometry
                  (world box, acrylic and scintillator spheres, and the detector's
                  difference(union(scaled zsphere, cylinder), cylinder) tree),
                  intersected by the reference's own CSG code
                  (CSG/csg_intersect_tree.h intersect_prim) compiled into
                  oracle/_ref/libref_csg.so by oracle/Makefile.
  csg_golden_leaves.npz  rays against the leaf-zoo prims (cone, hyperboloid, convex
                  polyhedra, phicut and halfspace inside Boolean trees, a
                  transformed cone), intersected by the same reference code.
  csg_golden_lists.npz  rays against the list-zoo prims (CSG_CONTIGUOUS,
                  CSG_DISCONTIGUOUS and CSG_OVERLAP list nodes at the prim root and
                  inside Boolean trees), intersected by the same reference code built
                  with WITH_CONTIGUOUS as the reference builds its CSG library.
  csg_golden_booleans.npz  rays against the Boolean-zoo prims (trees of height 1,
                  every shape of height 2, and height 3; many of these rays make
                  the tree evaluation loop), intersected by the same reference code.
  csg_lut.json    the packed Boolean LUT words parsed from CSG/csg_classify.h.

Run in the container that holds the reference (python tests/golden/make_golden.py
[name ...], default all of lut csg leaves lists booleans); the outputs are data only
and are what the CPU tests read.
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

REF = os.environ.get("OPTICKS_REFERENCE", "/root/reference")


def lut():
    src = open(os.path.join(REF, "CSG", "csg_classify.h")).read()
    a = re.search(r"packed_boolean_lut_ACloser\[4\]\s*=\s*\{([^}]*)\}", src).group(1)
    b = re.search(r"packed_boolean_lut_BCloser\[4\]\s*=\s*\{([^}]*)\}", src).group(1)
    to = lambda s: [int(x.strip(), 16) for x in s.split(",") if x.strip()]
    return dict(ACloser=to(a), BCloser=to(b), source="CSG/csg_classify.h")


def csg():
    import pyoracle as po
    from opticks_amd import configs
    po.build()
    assert po.reflib() is not None, "oracle/_ref/libref_csg.so missing"
    w = configs.config1(num_gs=1, per_gs=1)
    g = w.geometry
    rng = np.random.default_rng(20261202)
    rows = []
    for prim, n in ((3, 3000), (0, 300), (1, 300), (2, 300)):
        for k in range(n):
            if prim == 3:
                o = rng.uniform(-420, 420, 3)
                o[2] = rng.uniform(-420, 320)
                if k % 5 == 0:       # aim at the bore and the neck from below
                    o = np.array([rng.uniform(-150, 150), rng.uniform(-150, 150), -600.0])
            else:
                o = rng.uniform(-6600, 6600, 3)
            d = rng.normal(size=3)
            if prim == 3 and k % 5 == 0:
                d = np.array([rng.normal() * 0.05, rng.normal() * 0.05, 1.0])
            if k % 11 == 0:
                d = np.zeros(3); d[k % 3] = 1.0 if k % 2 else -1.0
            d /= np.linalg.norm(d)
            tmin = [0.0, 0.05, 1.0][k % 3]
            v, isect = po.ref_intersect_prim(g, prim, np.float32(tmin), o.astype(np.float32), d.astype(np.float32))
            rows.append((prim, tmin, *o.astype(np.float32), *d.astype(np.float32), int(v), *isect))
    a = np.array(rows, np.float64)
    return dict(prim=a[:, 0].astype(np.int32), tmin=a[:, 1].astype(np.float32),
                origin=a[:, 2:5].astype(np.float32), direction=a[:, 5:8].astype(np.float32),
                valid=a[:, 8].astype(np.int32), isect=a[:, 9:13].astype(np.float32))


def leaves():
    from opticks_amd import configs
    return zoo_vectors(configs.leaf_zoo(num_photon=4).geometry, 20261203)


def lists():
    from opticks_amd import configs
    return zoo_vectors(configs.list_zoo(num_photon=4).geometry, 20261204)


def booleans():
    from opticks_amd import configs
    return zoo_vectors(configs.boolean_zoo(num_photon=4).geometry, 20261205)


def zoo_vectors(g, seed):
    import pyoracle as po
    po.build()
    assert po.reflib() is not None, "oracle/_ref/libref_csg.so missing"
    # prim k+1 is the zoo's k-th node, in its own solid frame (instance transforms not applied)
    rng = np.random.default_rng(seed)
    rows = []
    for prim in range(1, len(g["prim"])):
        for k in range(700):
            o = rng.uniform(-260, 260, 3)
            if k % 7 == 0:
                o = rng.uniform(-30, 30, 3)              # from inside / near the axis
            d = rng.normal(size=3)
            if k % 11 == 0:
                d = np.zeros(3); d[k % 3] = 1.0 if k % 2 else -1.0
            if k % 13 == 0 and k % 11 != 0:
                d[2] = 0.0                               # parallel to cap planes
            d /= np.linalg.norm(d)
            tmin = [0.0, 0.05, 1.0][k % 3]
            v, isect = po.ref_intersect_prim(g, prim, np.float32(tmin), o.astype(np.float32), d.astype(np.float32))
            rows.append((prim, tmin, *o.astype(np.float32), *d.astype(np.float32), int(v), *isect))
    a = np.array(rows, np.float64)
    return dict(prim=a[:, 0].astype(np.int32), tmin=a[:, 1].astype(np.float32),
                origin=a[:, 2:5].astype(np.float32), direction=a[:, 5:8].astype(np.float32),
                valid=a[:, 8].astype(np.int32), isect=a[:, 9:13].astype(np.float32))


if __name__ == "__main__":
    which = sys.argv[1:] or ["lut", "csg", "leaves", "lists", "booleans"]
    if "lut" in which:
        with open(os.path.join(HERE, "csg_lut.json"), "w") as f:
            json.dump(lut(), f, indent=1)
    files = dict(csg="csg_golden.npz", leaves="csg_golden_leaves.npz", lists="csg_golden_lists.npz",
                 booleans="csg_golden_booleans.npz")
    for name, fn in files.items():
        if name in which:
            np.savez_compressed(os.path.join(HERE, fn), **globals()[name]())
    print("wrote", which)
