"""
Geometry and physics arrays from, and to, the reference's on-disk layout.

  <base>/CSGFoundry/solid.npy  (n,3,4) int32  CSGSolid (48 B)       CSG/CSGFoundry.cc:2783-2789
                    prim.npy   (n,4,4) f32    CSGPrim
                    node.npy   (n,4,4) f32    CSGNode
                    plan.npy   (n,1,4) f32    convex polyhedron planes (optional)
                    tran.npy   (n,4,4) f32    node transforms (optional)
                    itra.npy   (n,4,4) f32    their inverses (optional)
                    inst.npy   (n,4,4) f32    instance transforms + identity column
  <base>/CSGFoundry/SSim/stree/standard/        sysrap/sstandard.h:197-230, sysrap/snam.h:15-17
                    bnd.npy     (nbnd,4,2,nwl,4) f32 or f64 boundary table
                    bnd_meta.txt  "domain_low:60" ... (sysrap/sstandard.h:503-506, NP _meta.txt)
                    optical.npy (nbnd,4,4) int32
                    icdf.npy    (1 or 3, 4096, 1) scintillation inverse CDF (qudarap/QScint.cc:84-96)

load_csgfoundry(base) returns (geometry, physics) dicts that QSim accepts;
save_csgfoundry(base, geometry, physics) writes the same layout. Arrays are
read with numpy.load(allow_pickle=False).
"""
from __future__ import annotations

import os

import numpy as np

from .foundry import SOLID_DTYPE

GEOMETRY = ("solid", "prim", "node", "plan", "tran", "itra", "inst")
STANDARD = os.path.join("SSim", "stree", "standard")


def _cf_dir(base):
    d = os.path.join(base, "CSGFoundry")
    return d if os.path.isdir(d) else base


def _read_meta(path):
    meta = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if ":" in line:
                    k, v = line.strip().split(":", 1)
                    meta[k.strip()] = v.strip()
    return meta


def load_csgfoundry(base):
    cf = _cf_dir(base)
    g = {}
    for k in GEOMETRY:
        p = os.path.join(cf, k + ".npy")
        if os.path.exists(p):
            g[k] = np.load(p, allow_pickle=False)
    missing = [k for k in ("solid", "prim", "node", "inst") if k not in g]
    if missing:
        raise FileNotFoundError(f"{cf}: missing {missing}")
    solid = np.ascontiguousarray(g["solid"])
    if solid.dtype != SOLID_DTYPE:
        solid = np.ascontiguousarray(solid.astype(np.int32).reshape(-1, 12)).view(SOLID_DTYPE).reshape(-1)
    geometry = dict(solid=solid,
                    prim=np.ascontiguousarray(g["prim"], np.float32).reshape(-1, 4, 4),
                    node=np.ascontiguousarray(g["node"], np.float32).reshape(-1, 4, 4),
                    inst=np.ascontiguousarray(g["inst"], np.float32).reshape(-1, 4, 4),
                    tran=np.ascontiguousarray(g.get("tran", np.zeros((0, 4, 4))), np.float32).reshape(-1, 4, 4),
                    itra=np.ascontiguousarray(g.get("itra", np.zeros((0, 4, 4))), np.float32).reshape(-1, 4, 4),
                    plan=np.ascontiguousarray(g.get("plan", np.zeros((0, 4))), np.float32).reshape(-1, 4))
    sd = os.path.join(cf, STANDARD)
    bnd = np.load(os.path.join(sd, "bnd.npy"), allow_pickle=False)
    optical = np.load(os.path.join(sd, "optical.npy"), allow_pickle=False)
    meta = _read_meta(os.path.join(sd, "bnd_meta.txt"))
    physics = dict(bnd=np.ascontiguousarray(bnd, np.float32),
                   optical=np.ascontiguousarray(optical, np.int32).reshape(-1, 4, 4),
                   wl_low=float(meta.get("domain_low", 60.0)), wl_step=float(meta.get("domain_step", 1.0)),
                   n_wl=int(bnd.shape[3]))
    icdf_path = os.path.join(sd, "icdf.npy")
    if os.path.exists(icdf_path):
        icdf = np.load(icdf_path, allow_pickle=False).astype(np.float32)
        icdf = icdf.reshape(icdf.shape[0], -1)
        physics["icdf"] = np.ascontiguousarray(icdf)
        physics["icdf_hd_factor"] = 20 if icdf.shape[0] == 3 else 0    # QSim.cc:184, QScint layers
    return geometry, physics


def save_csgfoundry(base, geometry, physics):
    cf = os.path.join(base, "CSGFoundry")
    os.makedirs(os.path.join(cf, STANDARD), exist_ok=True)
    np.save(os.path.join(cf, "solid.npy"), np.ascontiguousarray(geometry["solid"]).view(np.int32).reshape(-1, 3, 4))
    for k in ("prim", "node", "tran", "itra", "inst"):
        a = geometry.get(k)
        if a is not None and len(a):
            np.save(os.path.join(cf, k + ".npy"), np.ascontiguousarray(a, np.float32).reshape(-1, 4, 4))
    plan = geometry.get("plan")
    if plan is not None and len(plan):
        np.save(os.path.join(cf, "plan.npy"), np.ascontiguousarray(plan, np.float32).reshape(-1, 1, 4))
    sd = os.path.join(cf, STANDARD)
    np.save(os.path.join(sd, "bnd.npy"), np.ascontiguousarray(physics["bnd"], np.float32))
    np.save(os.path.join(sd, "optical.npy"), np.ascontiguousarray(physics["optical"], np.int32))
    n_wl = physics["bnd"].shape[3]
    lo, step = float(physics["wl_low"]), float(physics["wl_step"])
    with open(os.path.join(sd, "bnd_meta.txt"), "w") as f:
        f.write(f"domain_low:{lo:g}\ndomain_high:{lo + step * (n_wl - 1):g}\ndomain_step:{step:g}\n"
                f"domain_range:{step * (n_wl - 1):g}\n")
    icdf = physics.get("icdf")
    if icdf is not None:
        np.save(os.path.join(sd, "icdf.npy"), np.ascontiguousarray(icdf, np.float32).reshape(icdf.shape[0], -1, 1))
    return cf
