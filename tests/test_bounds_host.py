"""World boxes of the TLAS leaves (opticks_amd/csrc/oa_bounds.h, compiled with g++ through
tests/cpp/bounds_check.cpp): exact support-function bounds of sphere, zsphere, cylinder and
box3 leaves under node and instance transforms, combined over CSG trees.

* conservative: every hit the CPU checker finds on an instance (its own traversal, brute force
  for the zoos) lies inside that instance's unpadded box, on randomly rotated detectors and on
  the leaf, Boolean and list zoos with random rotations;
* never looser than the transformed corners of the prim AABB (the box used before);
* the closed forms: a rotated sphere's box is its centre +- r on every axis, a cylinder tilted
  about x by 90 degrees swaps its height into y;
* the free ball (a trace whose segment lies inside it skips the TLAS walk) holds no hit the
  checker finds on a non-front instance, and on the detector-shell layouts holds the
  scintillator sphere (with upright detectors also the acrylic), where traces inside it end.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle as po
from opticks_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("bounds")
    out = d / "bounds_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "opticks_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "bounds_check.cpp"), "-o", str(out)], check=True)
    return d, out


def boxes_of(exe, g, tag):
    d, prog = exe
    solid = np.array([[int(s["primOffset"]), int(s["numPrim"])] for s in g["solid"]], np.int32).ravel()
    f = d / f"{tag}.in"
    with open(f, "wb") as fh:
        for a in (g["node"], g["itra"], g["prim"], g["inst"], solid):
            a = np.ascontiguousarray(a).ravel()
            np.int64(a.size).tofile(fh)
            a.tofile(fh)
    o = d / f"{tag}.out"
    subprocess.run([str(prog), str(f), str(o)], check=True)
    v = np.fromfile(o, np.float64)
    n = len(g["inst"])
    return v[:6 * n].reshape(-1, 6)


def front_and_ball(exe, g, tag):
    """(front mask, ball centre, ball r^2) as the TLAS build computes them from these boxes"""
    boxes_of(exe, g, tag)
    v = np.fromfile(exe[0] / f"{tag}.out", np.float64)
    n = len(g["inst"])
    return v[6 * n:7 * n].astype(bool), v[7 * n:7 * n + 3], v[7 * n + 3]


def corner_boxes(g):
    out = []
    for i in range(len(g["inst"])):
        gas = int(g["inst"][i].view(np.int32)[1, 3])
        s = g["solid"][gas]
        pl = g["prim"][int(s["primOffset"]):int(s["primOffset"]) + int(s["numPrim"])].astype(np.float64)
        lo = np.array([pl[:, 2, 0].min(), pl[:, 2, 1].min(), pl[:, 2, 2].min()])
        hi = np.array([pl[:, 2, 3].max(), pl[:, 3, 0].max(), pl[:, 3, 1].max()])
        m = g["inst"][i].astype(np.float64).copy()
        m[:, 3] = (0.0, 0.0, 0.0, 1.0)
        c = np.array([[x, y, z, 1.0] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])]) @ m
        out.append(np.r_[c[:, :3].min(0), c[:, :3].max(0)])
    return np.array(out)


def random_rotation(rng):
    q = rng.normal(size=4)
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def rotated(w, seed):
    """the workload's geometry with every instance but the world (instance 0) given a
    random rotation about its own position"""
    g = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in w.geometry.items()}
    rng = np.random.default_rng(seed)
    inst = g["inst"]
    for i in range(1, len(inst)):
        r = random_rotation(rng)
        inst[i, :3, :3] = (inst[i, :3, :3].astype(np.float64) @ r).astype(np.float32)
    return g


def hits_inside(g, boxes, physics, n_per_inst, seed, traverse, tol=1e-3):
    """rays at every instance but the world; returns (checked hits, hits outside their box)"""
    rng = np.random.default_rng(seed)
    cb = corner_boxes(g)
    rays = []
    for i in range(1, len(g["inst"])):
        lo, hi = cb[i, :3], cb[i, 3:]
        c, ext = 0.5 * (lo + hi), np.linalg.norm(hi - lo)
        o = c + 3.0 * ext * rng.normal(size=(n_per_inst, 3)) / np.linalg.norm(rng.normal(size=(n_per_inst, 3)), axis=1)[:, None]
        p = lo + (hi - lo) * rng.random((n_per_inst, 3))
        d = p - o
        d /= np.linalg.norm(d, axis=1)[:, None]
        r = np.zeros((n_per_inst, 2, 4), np.float32)
        r[:, 0, :3], r[:, 0, 3] = o, 0.0
        r[:, 1, :3], r[:, 1, 3] = d, 1e6
        rays.append(r)
    rays = np.concatenate(rays)
    out = po.World(g, physics).trace(rays, traverse=traverse)
    t = out[:, 0, 3].astype(np.float64)
    inst = out[:, 1, 3].view(np.uint32).astype(np.int64)
    hit = (out[:, 1, 1].view(np.uint32) != 0xffff) & (inst > 0)
    p = rays[hit, 0, :3].astype(np.float64) + t[hit, None] * rays[hit, 1, :3].astype(np.float64)
    b = boxes[inst[hit]]
    ext = b[:, 3:] - b[:, :3]
    slack = tol + 1e-5 * np.abs(p) + 1e-6 * ext
    outside = np.any((p < b[:, :3] - slack) | (p > b[:, 3:] + slack), axis=1)
    return int(hit.sum()), int(outside.sum())


def test_sphere_and_cylinder_closed_forms(exe):
    from opticks_amd.foundry import Foundry, Node
    fd = Foundry()
    s0 = fd.addSolid("a", [fd.addPrim(Node.sphere(50.0, (10.0, 0.0, 0.0)), 0)])
    s1 = fd.addSolid("b", [fd.addPrim(Node.cylinder(20.0, -100.0, 100.0), 0)])
    r = random_rotation(np.random.default_rng(1))
    t0 = np.eye(4)
    t0[:3, :3] = r
    t0[3, :3] = (1.0, 2.0, 3.0)
    fd.addInstance(t0, s0)
    t1 = np.eye(4)
    t1[:3, :3] = [[1, 0, 0], [0, 0, 1], [0, -1, 0]]       # z axis -> y axis (row vectors)
    fd.addInstance(t1, s1)
    g = fd.arrays()
    b = boxes_of(exe, g, "closed")
    c = np.array([10.0, 0.0, 0.0]) @ r + (1.0, 2.0, 3.0)
    assert np.allclose(b[0], np.r_[c - 50.0, c + 50.0], atol=1e-4)
    assert np.allclose(b[1], [-20.0, -100.0, -20.0, 20.0, 100.0, 20.0], atol=1e-4)


@pytest.mark.parametrize("which", ["config1", "leaf_zoo", "boolean_zoo", "list_zoo", "prim_clusters"])
def test_boxes_hold_every_hit_and_are_no_looser(exe, which):
    if which == "config1":
        w = configs.config1(num_gs=1, per_gs=1)
        g, n, traverse = dict(w.geometry), 60, 1
    else:
        w = getattr(configs, which)(num_photon=4)
        g, n, traverse = rotated(w, 7), 20000, 0
    b = boxes_of(exe, g, which)
    cb = corner_boxes(g)
    assert np.all(b[:, :3] >= cb[:, :3] - 1e-3) and np.all(b[:, 3:] <= cb[:, 3:] + 1e-3)
    checked, outside = hits_inside(g, b, w.physics, n, 11, traverse)
    assert checked >= (20000 if which == "config1" else 5000) and outside == 0, (checked, outside)
    if which == "config1":
        # the rotated detectors' boxes are much tighter than their corner boxes
        v = np.prod(b[1:, 3:] - b[1:, :3], axis=1) / np.prod(cb[1:, 3:] - cb[1:, :3], axis=1)
        assert np.median(v) < 0.8, np.median(v)


def hit_points(g, physics, n_per_inst, seed, traverse):
    """checker hits of rays aimed at every instance but the world: (points, instance)"""
    rng = np.random.default_rng(seed)
    cb = corner_boxes(g)
    rays = []
    for i in range(1, len(g["inst"])):
        lo, hi = cb[i, :3], cb[i, 3:]
        c, ext = 0.5 * (lo + hi), np.linalg.norm(hi - lo)
        o = c + 3.0 * ext * rng.normal(size=(n_per_inst, 3)) / np.linalg.norm(rng.normal(size=(n_per_inst, 3)), axis=1)[:, None]
        p = lo + (hi - lo) * rng.random((n_per_inst, 3))
        d = p - o
        d /= np.linalg.norm(d, axis=1)[:, None]
        r = np.zeros((n_per_inst, 2, 4), np.float32)
        r[:, 0, :3], r[:, 0, 3] = o, 0.0
        r[:, 1, :3], r[:, 1, 3] = d, 1e6
        rays.append(r)
    rays = np.concatenate(rays)
    out = po.World(g, physics).trace(rays, traverse=traverse)
    t = out[:, 0, 3].astype(np.float64)
    inst = out[:, 1, 3].view(np.uint32).astype(np.int64)
    hit = (out[:, 1, 1].view(np.uint32) != 0xffff) & (inst > 0)
    p = rays[hit, 0, :3].astype(np.float64) + t[hit, None] * rays[hit, 1, :3].astype(np.float64)
    return p, inst[hit]


@pytest.mark.parametrize("which", ["config1", "config2", "rotated_config1", "boolean_zoo", "prim_clusters"])
def test_free_ball_holds_no_hit(exe, which):
    """No hit the checker finds on a TLAS (non-front) instance lies inside the free ball,
    on the detector shells (upright and randomly rotated detectors), the rotated Boolean
    zoo and the rotated five-prim clusters; on the shells the ball holds the LS sphere (and with upright detectors the acrylic
    sphere), where a photon's traces inside the shell end; with upright detectors a 2% larger
    ball would hold hits."""
    if which == "config2":
        w = configs.config2(num_events=1, mean_gs=1, mean_photons=1)
        g, n, traverse = dict(w.geometry), 6, 1
    elif which in ("boolean_zoo", "prim_clusters"):
        w = getattr(configs, which)(num_photon=4)
        g, n, traverse = rotated(w, 7), 20000, 0
    else:
        w = configs.config1(num_gs=1, per_gs=1)
        g, n, traverse = (dict(w.geometry) if which == "config1" else rotated(w, 5)), 60, 1
    front, c, r2 = front_and_ball(exe, g, which)
    assert front[0] and front.sum() == 1             # the global solid (world box, acrylic, LS)
    p, inst = hit_points(g, w.physics, n, 13, traverse)
    assert len(p) > 5000
    dist = np.linalg.norm(p - c, axis=1)
    if r2 > 0:
        assert dist.min() > np.sqrt(r2), (dist.min(), np.sqrt(r2))
    if which not in ("boolean_zoo", "prim_clusters"):
        # upright detectors: the ball holds the acrylic sphere; randomly rotated ones
        # can point a side inward, and it holds the LS sphere
        inner = w.meta["ls_r"] + (120.0 if which != "rotated_config1" else 0.0)
        assert r2 > 0 and np.sqrt(r2) > inner, (np.sqrt(r2), inner)
        if which != "rotated_config1":
            assert dist.min() < 1.02 * np.sqrt(r2)     # the dome apex: a 2% larger ball holds hits


def test_free_ball_absent_without_tlas_instances(exe):
    """a world whose only instance is front (the water-sphere workload) has no ball"""
    w = configs.config0(num_photon=4)
    front, c, r2 = front_and_ball(exe, dict(w.geometry), "config0")
    assert front.all() and r2 <= 0


def test_moved_detectors_take_the_world_out_of_the_front(exe):
    """detectors moved to three times their radius (outside the world box): the union of
    the instance boxes grows so that the world's box is under a quarter of its area, and a
    TLAS build makes no instance front (the premise of the GPU test that rebuilds the TLAS
    after such a refit); the unmoved layout keeps the world front"""
    w = configs.config1(num_gs=1, per_gs=1)
    g = dict(w.geometry)
    front, _, _ = front_and_ball(exe, g, "config1_unmoved")
    assert front[0] and front.sum() == 1
    moved = g["inst"].copy()
    moved[1:, 3, :3] *= np.float32(3.0)
    g["inst"] = moved
    front, _, r2 = front_and_ball(exe, g, "config1_moved")
    assert front.sum() == 0 and r2 <= 0


def test_free_ball_with_scaled_detectors(exe):
    """detectors scaled by 1.2 about their own positions: the placed local box no longer
    bounds their distance (the transform is not rigid), so the ball falls back to the world
    boxes and still holds no hit the checker finds"""
    w = configs.config1(num_gs=1, per_gs=1)
    g = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in w.geometry.items()}
    g["inst"][1:, :3, :3] *= np.float32(1.2)
    front, c, r2 = front_and_ball(exe, g, "scaled")
    assert r2 > 0
    p, inst = hit_points(g, w.physics, 60, 17, 1)
    assert len(p) > 5000
    dist = np.linalg.norm(p - c, axis=1)
    assert dist.min() > np.sqrt(r2), (dist.min(), np.sqrt(r2))
    # the world-box distance is looser than the rigid case's placed-box distance
    _, _, r2_rigid = front_and_ball(exe, dict(w.geometry), "rigid")
    assert r2 < r2_rigid
