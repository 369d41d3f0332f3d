"""
Writes tests/golden/np_foundry: the leaf-zoo workload's geometry and physics in
the reference's CSGFoundry / SSim layout, each array written by the reference's
own NP.hh (tests/golden/np_writer.cpp compiled against the reference checkout's
sysrap/NP.hh), with the bnd table in double precision and its domain metadata as
sstandard.h:500-506 sets it (bnd_meta.txt written by NP::save_meta). The loader
tests read this folder as the reference would have left it on disk.

  python tests/golden/make_np_fixture.py       (needs the reference checkout and g++)

Arrays (CSG/CSGFoundry.cc:2769-2789, sysrap/sstandard.h:197-230):
  CSGFoundry/solid.npy (n,3,4) i4, prim/node/tran/itra/inst.npy (n,4,4) f4,
  plan.npy (n,1,4) f4, SSim/stree/standard/bnd.npy (nbnd,4,2,nwl,4) f8,
  optical.npy (nbnd,4,4) i4, icdf.npy (3,4096,1) f4.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.environ.get("OPTICKS_REFERENCE", "/root/reference")
OUT = os.path.join(HERE, "np_foundry")


def workload():
    from opticks_amd import configs
    return configs.leaf_zoo(num_photon=4000)


def arrays(w):
    g, p = w.geometry, w.physics
    cf = "CSGFoundry"
    sd = os.path.join(cf, "SSim", "stree", "standard")
    n_wl = p["bnd"].shape[3]
    lo, step = float(p["wl_low"]), float(p["wl_step"])
    meta = dict(domain_low=lo, domain_high=lo + step * (n_wl - 1), domain_step=step,
                domain_range=step * (n_wl - 1))
    out = [(os.path.join(cf, "solid.npy"), "i4", np.ascontiguousarray(g["solid"]).view(np.int32).reshape(-1, 3, 4), {})]
    for k in ("prim", "node", "tran", "itra", "inst"):
        out.append((os.path.join(cf, k + ".npy"), "f4", np.asarray(g[k], np.float32).reshape(-1, 4, 4), {}))
    out.append((os.path.join(cf, "plan.npy"), "f4", np.asarray(g["plan"], np.float32).reshape(-1, 1, 4), {}))
    out.append((os.path.join(sd, "bnd.npy"), "f8", np.asarray(p["bnd"], np.float64), meta))
    out.append((os.path.join(sd, "optical.npy"), "i4", np.asarray(p["optical"], np.int32), {}))
    out.append((os.path.join(sd, "icdf.npy"), "f4", np.asarray(p["icdf"], np.float32).reshape(-1, 4096, 1), {}))
    return out


def main():
    hdr = os.path.join(REF, "sysrap", "NP.hh")
    if not os.path.exists(hdr):
        sys.exit(f"reference NP.hh not found at {hdr}")
    tmp = tempfile.mkdtemp(prefix="np_fixture_")
    try:
        exe = os.path.join(tmp, "np_writer")
        subprocess.run(["g++", "-std=c++17", "-O1", "-w", "-I", os.path.join(REF, "sysrap"),
                        os.path.join(HERE, "np_writer.cpp"), "-o", exe], check=True)
        if os.path.isdir(OUT):
            shutil.rmtree(OUT)
        lines = []
        for i, (rel, dt, a, meta) in enumerate(arrays(workload())):
            raw = os.path.join(tmp, f"{i}.raw")
            np.ascontiguousarray(a).tofile(raw)
            dst = os.path.join(OUT, rel)
            os.makedirs(os.path.dirname(dst), exist_ok=True)
            m = (" -- " + " ".join(f"{k}={v:g}" for k, v in meta.items())) if meta else ""
            lines.append(f"{dst} {dt} {raw} " + " ".join(str(d) for d in a.shape) + m)
        man = os.path.join(tmp, "manifest.txt")
        with open(man, "w") as f:
            f.write("\n".join(lines) + "\n")
        subprocess.run([exe, man], check=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
