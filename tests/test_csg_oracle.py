"""CPU restatement of the CSG intersection, pinned against the reference's own code."""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from opticks_amd import configs
from opticks_amd.foundry import Foundry, Node, CSG_DIFFERENCE, CSG_UNION, CSG_INTERSECTION

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def geo1():
    return configs.config1(num_gs=1, per_gs=1).geometry


def test_golden_vectors_bitwise(geo1):
    """oracle intersect_prim == reference CSG code on committed (ray, prim) vectors"""
    z = np.load(os.path.join(GOLD, "csg_golden.npz"))
    bad = 0
    for i in range(len(z["prim"])):
        v, isect = po.intersect_prim(geo1, int(z["prim"][i]), z["tmin"][i], z["origin"][i], z["direction"][i])
        if int(v) != int(z["valid"][i]) or not np.array_equal(isect.view(np.uint32), z["isect"][i].view(np.uint32)):
            bad += 1
    assert bad == 0
    assert z["valid"].sum() > 1000      # the vectors exercise hits, not only misses


@pytest.mark.skipif(po.reflib() is None, reason="oracle/_ref not built (reference checkout absent)")
def test_random_rays_vs_reference_build(geo1):
    rng = np.random.default_rng(5)
    bad = 0
    for k in range(4000):
        prim = 3 if k % 2 else int(rng.integers(0, 3))
        o = rng.uniform(-450, 450, 3).astype(np.float32) if prim == 3 else rng.uniform(-7000, 7000, 3).astype(np.float32)
        d = rng.normal(size=3); d = (d / np.linalg.norm(d)).astype(np.float32)
        t = np.float32([0.0, 0.05, 2.0][k % 3])
        v1, i1 = po.intersect_prim(geo1, prim, t, o, d)
        v2, i2 = po.ref_intersect_prim(geo1, prim, t, o, d)
        bad += (v1 != v2) or not np.array_equal(i1.view(np.uint32), i2.view(np.uint32))
    assert bad == 0


def test_lut_matches_reference_constants():
    ref = json.load(open(os.path.join(GOLD, "csg_lut.json")))
    a, b = po.packed_lut()
    assert [int(x) for x in a] == ref["ACloser"][:3]
    assert [int(x) for x in b] == ref["BCloser"][:3]


def single_prim(node):
    fd = Foundry()
    p = fd.addPrim(node, 0)
    s = fd.addSolid("s", [p])
    fd.addInstance(np.eye(4), s)
    return fd.arrays()


def test_sphere_closed_form_fp64_budget():
    """FP32 intersect vs the FP64 closed form t = (-b -+ sqrt(b^2 - d c))/d.

    Budget (first-order rounding analysis of the robust-roots algorithm of
    CSG/csg_robust_quadratic_roots.h, eps = 2^-24):
        |t32 - t64| <= 8 eps (b^2 + |O|^2 + r^2) / (2 sqrt(disc)) + 8 eps t
    The first term is the cancellation in disc = b^2 - d c, which dominates for
    origins far from the sphere; it is a property of the reference algorithm,
    which the FP32 path reproduces bit for bit (test_golden_vectors_bitwise).
    Typical error is about one ulp of t (median checked too)."""
    g = single_prim(Node.sphere(1000.0))
    rng = np.random.default_rng(3)
    eps = 2.0 ** -24
    ulps = []
    for _ in range(6000):
        o = rng.uniform(-3000, 3000, 3)
        d = rng.normal(size=3); d /= np.linalg.norm(d)
        of, df = o.astype(np.float32), d.astype(np.float32)
        v, isect = po.intersect_prim(g, 0, np.float32(0.0), of, df)
        O, D = of.astype(np.float64), df.astype(np.float64)
        b, c, dd = O @ D, O @ O - 1e6, D @ D
        disc = b * b - dd * c
        tol_disc = 8 * eps * (b * b + O @ O + 1e6)
        if disc <= tol_disc:          # tangent within rounding: either answer is allowed
            continue
        r1, r2 = (-b - np.sqrt(disc)) / dd, (-b + np.sqrt(disc)) / dd
        t = r1 if r1 > 0 else r2
        if t <= 1e-3:
            assert not v or isect[3] < 1e-2
            continue
        assert v
        bound = tol_disc / (2 * np.sqrt(disc)) + 8 * eps * t
        assert abs(float(isect[3]) - t) <= bound, (o, d, isect, t, bound)
        ulps.append(abs(float(isect[3]) - t) / float(np.spacing(np.float32(t))))
        n = (O + t * D) / 1000.0
        assert np.abs(isect[:3] - n).max() < 1e-4
    assert len(ulps) > 300 and np.median(ulps) <= 2.0


def test_box3_slab_closed_form():
    g = single_prim(Node.box3(400.0, 600.0, 800.0))
    for o, d, t in [((0, 0, 0), (1, 0, 0), 200.0), ((0, 0, 0), (0, -1, 0), 300.0), ((0, 0, -1000), (0, 0, 1), 600.0),
                    ((-500, 0, 0), (1, 0, 0), 300.0)]:
        v, isect = po.intersect_prim(g, 0, np.float32(0.0), np.float32(o), np.float32(d))
        assert v and isect[3] == np.float32(t)
        p = np.array(o, float) + t * np.array(d, float)
        ax = int(np.argmax(np.abs(d)))
        n = np.zeros(3, np.float32)
        n[ax] = np.sign(p[ax])
        assert np.array_equal(isect[:3], n)


def test_cylinder_caps_and_sheet():
    g = single_prim(Node.cylinder(100.0, -50.0, 50.0))
    v, i = po.intersect_prim(g, 0, np.float32(0.0), np.float32((0, 0, -500)), np.float32((0, 0, 1)))
    assert v and i[3] == np.float32(450.0) and tuple(i[:3]) == (0.0, 0.0, -1.0)
    v, i = po.intersect_prim(g, 0, np.float32(0.0), np.float32((-500, 0, 0)), np.float32((1, 0, 0)))
    assert v and i[3] == np.float32(400.0) and tuple(i[:3]) == (-1.0, 0.0, 0.0)
    v, i = po.intersect_prim(g, 0, np.float32(0.0), np.float32((-500, 0, 200)), np.float32((1, 0, 0)))
    assert not v


def test_boolean_trees_against_point_membership():
    """A ray's first valid hit on union/intersection/difference lies on the
    boundary of the combined solid: just before it the ray is on one side of
    the point-membership function of the Boolean, just after on the other."""
    a, b = Node.sphere(100.0, (-40, 0, 0)), Node.sphere(100.0, (40, 0, 0))
    inside = {
        CSG_UNION: lambda p: (np.linalg.norm(p - (-40, 0, 0)) < 100) or (np.linalg.norm(p - (40, 0, 0)) < 100),
        CSG_INTERSECTION: lambda p: (np.linalg.norm(p - (-40, 0, 0)) < 100) and (np.linalg.norm(p - (40, 0, 0)) < 100),
        CSG_DIFFERENCE: lambda p: (np.linalg.norm(p - (-40, 0, 0)) < 100) and not (np.linalg.norm(p - (40, 0, 0)) < 100),
    }
    rng = np.random.default_rng(11)
    for op, fn in inside.items():
        g = single_prim(Node.op(op, a, b))
        checked = 0
        for _ in range(5000):
            o = rng.uniform(-250, 250, 3)
            d = rng.normal(size=3); d /= np.linalg.norm(d)
            v, i = po.intersect_prim(g, 0, np.float32(0.0), o.astype(np.float32), d.astype(np.float32))
            if not v:
                continue
            p = o + float(i[3]) * d
            before, after = fn(p - 0.05 * d), fn(p + 0.05 * d)
            assert before != after, (op, o, d, i)
            # normal points out of the solid: dot with the ray sign matches entering/exiting
            assert (np.dot(i[:3], d) < 0) == (not before)
            checked += 1
        assert checked > 100


@pytest.fixture(scope="module")
def zoo():
    return configs.leaf_zoo(num_photon=4).geometry


def test_leaf_golden_vectors_bitwise(zoo):
    """cone, hyperboloid, convex polyhedra, phicut and halfspace (in trees), transformed
    cone: oracle == reference CSG code on committed vectors"""
    z = np.load(os.path.join(GOLD, "csg_golden_leaves.npz"))
    bad = []
    for i in range(len(z["prim"])):
        v, isect = po.intersect_prim(zoo, int(z["prim"][i]), z["tmin"][i], z["origin"][i], z["direction"][i])
        if int(v) != int(z["valid"][i]) or not np.array_equal(isect.view(np.uint32), z["isect"][i].view(np.uint32)):
            bad.append(i)
    assert not bad, (len(bad), bad[:5])
    for p in range(1, 10):
        assert z["valid"][z["prim"] == p].sum() > 50          # every prim is hit


@pytest.mark.skipif(po.reflib() is None, reason="oracle/_ref not built (reference checkout absent)")
def test_leaf_random_rays_vs_reference_build(zoo):
    rng = np.random.default_rng(8)
    bad = 0
    for k in range(6000):
        prim = 1 + k % 9
        o = rng.uniform(-300, 300, 3).astype(np.float32)
        d = rng.normal(size=3); d = (d / np.linalg.norm(d)).astype(np.float32)
        t = np.float32([0.0, 0.05, 2.0][k % 3])
        v1, i1 = po.intersect_prim(zoo, prim, t, o, d)
        v2, i2 = po.ref_intersect_prim(zoo, prim, t, o, d)
        bad += (v1 != v2) or not np.array_equal(i1.view(np.uint32), i2.view(np.uint32))
    assert bad == 0


def test_leaf_closed_forms(zoo):
    """cone side / cap distances, the hyperboloid waist, a cube face, and the halfspace cut"""
    # cone prim 1: r 100 at z -50 to r 30 at z 80; a ray along +x at z = 15 meets the side at
    # r(z) = 100 - 70 (z + 50) / 130
    r15 = 100.0 - 70.0 * 65.0 / 130.0
    v, i = po.intersect_prim(zoo, 1, 0.0, (-500.0, 0.0, 15.0), (1.0, 0.0, 0.0))
    assert v and abs(i[3] - (500.0 - r15)) < 1e-3
    v, i = po.intersect_prim(zoo, 1, 0.0, (0.0, 0.0, -500.0), (0.0, 0.0, 1.0))
    assert v and abs(i[3] - 450.0) < 1e-3 and i[2] == -1.0           # bottom cap
    # hyperboloid prim 2: waist radius 60 at z = 0
    v, i = po.intersect_prim(zoo, 2, 0.0, (-500.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    assert v and abs(i[3] - 440.0) < 1e-3 and abs(i[0] + 1.0) < 1e-6
    # cube prim 3: faces at +-70
    v, i = po.intersect_prim(zoo, 3, 0.0, (0.0, -500.0, 3.0), (0.0, 1.0, 0.0))
    assert v and abs(i[3] - 430.0) < 1e-4 and tuple(i[:3]) == (0.0, -1.0, 0.0)
    # sphere r 140 minus the halfspace z <= 40 (prim 7): from below along +z the first surface is z = 40
    v, i = po.intersect_prim(zoo, 7, 0.0, (0.0, 0.0, -500.0), (0.0, 0.0, 1.0))
    assert v and abs(i[3] - 540.0) < 1e-3


@pytest.fixture(scope="module")
def lzoo():
    return configs.list_zoo(num_photon=4).geometry


def test_list_golden_vectors_bitwise(lzoo):
    """CSG_CONTIGUOUS / DISCONTIGUOUS / OVERLAP list nodes at the prim root and inside
    trees: oracle == reference CSG code (built WITH_CONTIGUOUS) on committed vectors"""
    z = np.load(os.path.join(GOLD, "csg_golden_lists.npz"))
    bad = []
    for i in range(len(z["prim"])):
        v, isect = po.intersect_prim(lzoo, int(z["prim"][i]), z["tmin"][i], z["origin"][i], z["direction"][i])
        if int(v) != int(z["valid"][i]) or not np.array_equal(isect.view(np.uint32), z["isect"][i].view(np.uint32)):
            bad.append(i)
    assert not bad, (len(bad), bad[:5])
    for p in range(1, 7):
        assert z["valid"][z["prim"] == p].sum() > 20


@pytest.mark.skipif(po.reflib() is None, reason="oracle/_ref not built (reference checkout absent)")
def test_list_random_rays_vs_reference_build(lzoo):
    rng = np.random.default_rng(9)
    bad = 0
    for k in range(6000):
        prim = 1 + k % 6
        o = rng.uniform(-250, 250, 3).astype(np.float32)
        d = rng.normal(size=3); d = (d / np.linalg.norm(d)).astype(np.float32)
        t = np.float32([0.0, 0.05, 2.0][k % 3])
        v1, i1 = po.intersect_prim(lzoo, prim, t, o, d)
        v2, i2 = po.ref_intersect_prim(lzoo, prim, t, o, d)
        bad += (v1 != v2) or not np.array_equal(i1.view(np.uint32), i2.view(np.uint32))
    assert bad == 0


def test_list_closed_forms(lzoo):
    """bead chain (contiguous spheres r 60 at x = -80, 0, 80): first surface from outside
    at x = -140, exit from inside the middle bead at the far end x = 140 (not at the bead's
    own surface x = 60); islands (discontiguous): nearest island only; lens (overlap of two
    r 120 spheres at z = +-70 and a cylinder r 90, |z| <= 60): z in [-50, 50] on the axis,
    and a ray through the cylinder side outside the spheres' overlap misses"""
    v, i = po.intersect_prim(lzoo, 1, 0.0, (-500.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    assert v and abs(i[3] - 360.0) < 1e-3 and abs(i[0] + 1.0) < 1e-6
    v, i = po.intersect_prim(lzoo, 1, 0.0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    assert v and abs(i[3] - 140.0) < 1e-3 and abs(i[0] - 1.0) < 1e-6
    v, i = po.intersect_prim(lzoo, 2, 0.0, (-500.0, -60.0, 0.0), (1.0, 0.0, 0.0))
    assert v and abs(i[3] - 365.0) < 1e-3
    v, i = po.intersect_prim(lzoo, 2, 0.0, (0.0, -60.0, 0.0), (1.0, 0.0, 0.0))
    assert v and abs(i[3] - 65.0) < 1e-3                   # to the right island, not back
    v, i = po.intersect_prim(lzoo, 3, 0.0, (0.0, 0.0, -500.0), (0.0, 0.0, 1.0))
    assert v and abs(i[3] - 450.0) < 1e-3 and abs(i[2] + 1.0) < 1e-6
    v, i = po.intersect_prim(lzoo, 3, 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    assert v and abs(i[3] - 50.0) < 1e-3 and abs(i[2] - 1.0) < 1e-6
    # the sphere overlap at z = 0 reaches x = sqrt(120^2 - 70^2) = 97.5 > 90: the cylinder side is the surface
    v, i = po.intersect_prim(lzoo, 3, 0.0, (-500.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    assert v and abs(i[3] - 410.0) < 1e-3
    # bead chain with a bore (difference with cylinder r 25 along z): along x at y = 0 the bore splits it
    v, i = po.intersect_prim(lzoo, 4, 0.0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    assert v and abs(i[3] - 25.0) < 1e-3


@pytest.fixture(scope="module")
def bzoo():
    return configs.boolean_zoo(num_photon=4).geometry


def test_boolean_golden_vectors_bitwise(bzoo):
    """Boolean trees of height 1 (each operator), every height-2 shape and height 3:
    oracle == reference CSG code on committed vectors, many of which loop"""
    z = np.load(os.path.join(GOLD, "csg_golden_booleans.npz"))
    po.tree_counts(reset=True)
    bad = []
    for i in range(len(z["prim"])):
        v, isect = po.intersect_prim(bzoo, int(z["prim"][i]), z["tmin"][i], z["origin"][i], z["direction"][i])
        if int(v) != int(z["valid"][i]) or not np.array_equal(isect.view(np.uint32), z["isect"][i].view(np.uint32)):
            bad.append(i)
    evals, loops = po.tree_counts(reset=True)
    assert not bad, (len(bad), bad[:5])
    assert evals == len(z["prim"]) and loops > 500        # the vectors take the loop decisions
    for p in range(1, 10):
        assert z["valid"][z["prim"] == p].sum() > 50


@pytest.mark.skipif(po.reflib() is None, reason="oracle/_ref not built (reference checkout absent)")
def test_boolean_random_rays_vs_reference_build(bzoo):
    rng = np.random.default_rng(10)
    bad = 0
    for k in range(6000):
        prim = 1 + k % 9
        o = rng.uniform(-250, 250, 3).astype(np.float32)
        d = rng.normal(size=3); d = (d / np.linalg.norm(d)).astype(np.float32)
        t = np.float32([0.0, 0.05, 2.0][k % 3])
        v1, i1 = po.intersect_prim(bzoo, prim, t, o, d)
        v2, i2 = po.ref_intersect_prim(bzoo, prim, t, o, d)
        bad += (v1 != v2) or not np.array_equal(i1.view(np.uint32), i2.view(np.uint32))
    assert bad == 0
