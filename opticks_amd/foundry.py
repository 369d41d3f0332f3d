"""
Geometry in the reference's CSGFoundry array layout, built in numpy.

The arrays are exactly the ones the reference uploads (CSG/CSGFoundry.h:458-470):

  solid  (n)      48 B CSGSolid   label[16], numPrim, primOffset, type, intent+pad,
                                  center_extent              (CSG/CSGSolid.h:37)
  prim   (n,4,4)  CSGPrim  q0.i=(numNode,nodeOffset,tranOffset,planOffset)
                           q1.u=(sbtIndexOffset,meshIdx,repeatIdx,primIdx)
                           q2.f=(mn.x,mn.y,mn.z,mx.x) q3.f=(mx.y,mx.z,-,-) q3.u.w=globalPrimIdx
                                                             (CSG/CSGPrim.h:75-135)
  node   (n,4,4)  CSGNode  q0/q1 params, q1.u.z boundary, q1.u.w index, q2/q3 AABB,
                           q3.u.z typecode, q3.u.w gtransformIdx|complement<<31
                                                             (CSG/CSGNode.h:84-345)
  tran/itra (n,4,4) qat4 row-vector transforms (p' = p M)    (sysrap/sqat4.h:45-104)
  inst   (n,4,4)  qat4 object->world with the identity in the w column:
                  q0.i.w ins_idx, q1.i.w gas_idx, q2.i.w sensor_identifier+1, q3.i.w sensor_index
                                                             (sysrap/sqat4.h:345-390)

Boolean trees are complete binary trees in level order padded with CSG_ZERO
(CSG/csg_intersect_tree.h:199-274); gtransformIdx is 1-based into tran/itra.
List nodes (CSG_CONTIGUOUS / DISCONTIGUOUS / OVERLAP) hold subNum leaves stored
after the tree at root + subOffset (CSG/CSGMaker.cc:805-870).
Method names follow CSGFoundry (addSolid, addPrim, addNode, addTran, addInstance).
"""
from __future__ import annotations

import numpy as np

# sysrap/OpticksCSG.h
CSG_ZERO = 0
CSG_UNION = 1
CSG_INTERSECTION = 2
CSG_DIFFERENCE = 3
CSG_CONTIGUOUS = 11
CSG_DISCONTIGUOUS = 12
CSG_OVERLAP = 13
CSG_SPHERE = 101
CSG_ZSPHERE = 103
CSG_CYLINDER = 105
CSG_CONE = 108
CSG_BOX3 = 110
CSG_CONVEXPOLYHEDRON = 112
CSG_HYPERBOLOID = 117
CSG_PHICUT = 121
CSG_HALFSPACE = 125
UNBOUNDED = 1.0e9     # AABB half-extent given to unbounded leaves (bounded by their Boolean partners)

SOLID_DTYPE = np.dtype([
    ("label", "S16"), ("numPrim", "<i4"), ("primOffset", "<i4"), ("type", "<i4"),
    ("intent", "S1"), ("pad", "S3"), ("center_extent", "<f4", (4,)),
])
assert SOLID_DTYPE.itemsize == 48


class Node:
    """One CSG node: a leaf with parameters or a Boolean operator."""

    def __init__(self, typecode, q0=(0, 0, 0, 0), q1=(0, 0), aabb=None, transform=None,
                 complement=False, left=None, right=None, planes=None, subs=None):
        self.typecode = typecode
        self.planes = None if planes is None else np.asarray(planes, np.float32).reshape(-1, 4)
        self.q0 = tuple(float(v) for v in q0)
        self.q1 = tuple(float(v) for v in q1)
        self.aabb = aabb
        self.transform = transform      # 4x4 row-vector local->solid, or None
        self.complement = complement
        self.left = left
        self.right = right
        self.subs = subs                # leaves of a list node

    # leaf constructors (parameters as CSGNode.h:84-100)
    @staticmethod
    def sphere(radius, center=(0.0, 0.0, 0.0)):
        c = np.asarray(center, float)
        return Node(CSG_SPHERE, (c[0], c[1], c[2], radius),
                    aabb=(c - radius, c + radius))

    @staticmethod
    def zsphere(radius, zmin, zmax, center=(0.0, 0.0, 0.0)):
        c = np.asarray(center, float)
        lo = np.array([c[0] - radius, c[1] - radius, c[2] + zmin])
        hi = np.array([c[0] + radius, c[1] + radius, c[2] + zmax])
        return Node(CSG_ZSPHERE, (c[0], c[1], c[2], radius), (zmin, zmax), aabb=(lo, hi))

    @staticmethod
    def cylinder(radius, z1, z2):
        return Node(CSG_CYLINDER, (0.0, 0.0, 0.0, radius), (z1, z2),
                    aabb=(np.array([-radius, -radius, z1]), np.array([radius, radius, z2])))

    @staticmethod
    def box3(fx, fy, fz):
        h = np.array([fx, fy, fz], float) / 2.0
        return Node(CSG_BOX3, (fx, fy, fz, 0.0), aabb=(-h, h))

    @staticmethod
    def cone(r1, z1, r2, z2):
        """truncated cone, radius r1 at z1, r2 at z2 (z2 > z1), with caps (CSG_CONE, q0 = r1,z1,r2,z2)"""
        r = max(r1, r2)
        return Node(CSG_CONE, (r1, z1, r2, z2), aabb=(np.array([-r, -r, z1]), np.array([r, r, z2])))

    @staticmethod
    def hyperboloid(r0, zf, z1, z2):
        """one-sheet hyperboloid x^2+y^2 = r0^2 (1 + (z/zf)^2) between z1 < z2, with caps"""
        r = r0 * np.sqrt(1.0 + max((z1 / zf) ** 2, (z2 / zf) ** 2))
        return Node(CSG_HYPERBOLOID, (r0, zf, z1, z2), aabb=(np.array([-r, -r, z1]), np.array([r, r, z2])))

    @staticmethod
    def phicut(phi0, phi1):
        """the unbounded wedge phi0 <= phi <= phi1 about the z axis (radians)"""
        return Node(CSG_PHICUT, (np.cos(phi0), np.sin(phi0), np.cos(phi1), np.sin(phi1)),
                    aabb=(np.full(3, -UNBOUNDED), np.full(3, UNBOUNDED)))

    @staticmethod
    def halfspace(normal, w):
        """the unbounded region n.x <= w (n unit)"""
        n = np.asarray(normal, float)
        n = n / np.linalg.norm(n)
        return Node(CSG_HALFSPACE, (n[0], n[1], n[2], w), aabb=(np.full(3, -UNBOUNDED), np.full(3, UNBOUNDED)))

    @staticmethod
    def convexpolyhedron(planes, aabb):
        """intersection of the half-spaces n.x <= d, planes (k,4) rows (nx,ny,nz,d), n unit;
        aabb (lo, hi) bounds it"""
        return Node(CSG_CONVEXPOLYHEDRON, aabb=(np.asarray(aabb[0], float), np.asarray(aabb[1], float)),
                    planes=planes)

    @staticmethod
    def op(typecode, left, right):
        return Node(typecode, left=left, right=right)

    @staticmethod
    def list(typecode, subs):
        """list node over leaves: CSG_CONTIGUOUS (touching union), CSG_DISCONTIGUOUS (disjoint
        union) or CSG_OVERLAP (intersection)"""
        assert typecode in (CSG_CONTIGUOUS, CSG_DISCONTIGUOUS, CSG_OVERLAP)
        assert subs and all(n.is_leaf() for n in subs)
        return Node(typecode, subs=list(subs))

    def is_leaf(self):
        return self.typecode >= 101

    def is_list(self):
        return CSG_CONTIGUOUS <= self.typecode <= CSG_OVERLAP

    def height(self):
        if self.is_leaf() or self.is_list():
            return 0
        return 1 + max(self.left.height(), self.right.height())

    def world_aabb(self):
        """Conservative AABB in the solid frame (transformed leaf boxes)."""
        if self.is_leaf():
            lo, hi = (np.asarray(self.aabb[0], float), np.asarray(self.aabb[1], float))
            if self.transform is None:
                return lo, hi
            corners = np.array([[x, y, z, 1.0] for x in (lo[0], hi[0]) for y in (lo[1], hi[1])
                                for z in (lo[2], hi[2])])
            w = corners @ self.transform
            return w[:, :3].min(0), w[:, :3].max(0)
        if self.is_list():
            boxes = [n.world_aabb() for n in self.subs]
            los, his = np.array([b[0] for b in boxes]), np.array([b[1] for b in boxes])
            if self.typecode == CSG_OVERLAP:
                return los.max(0), his.min(0)
            return los.min(0), his.max(0)
        llo, lhi = self.left.world_aabb()
        rlo, rhi = self.right.world_aabb()
        if self.typecode == CSG_DIFFERENCE:
            return llo, lhi
        if self.typecode == CSG_INTERSECTION:
            return np.maximum(llo, rlo), np.minimum(lhi, rhi)
        return np.minimum(llo, rlo), np.maximum(lhi, rhi)


def scale_transform(sx, sy, sz):
    t = np.eye(4)
    t[0, 0], t[1, 1], t[2, 2] = sx, sy, sz
    return t


class Foundry:
    """Accumulates solids/prims/nodes/transforms/instances (CSGFoundry shape)."""

    def __init__(self):
        self.solids = []
        self.prims = []
        self.nodes = []
        self.trans = []
        self.insts = []
        self.plans = []

    def addTran(self, t):
        t = np.asarray(t, float)
        self.trans.append(t)
        return len(self.trans)            # 1-based gtransformIdx

    def _node_record(self, nd, boundary, index):
        rec = np.zeros(16, np.float32)
        u = rec.view(np.uint32)
        if nd is None:
            u[14] = CSG_ZERO
            u[6] = boundary
            return rec
        rec[0:4] = nd.q0
        rec[4:6] = nd.q1
        if nd.planes is not None:
            u[0] = len(self.plans)            # planeIdx: offset into plan (CSGNode.h:204-207)
            u[1] = len(nd.planes)             # planeNum
            self.plans.extend(list(nd.planes))
        u[6] = boundary
        u[7] = index
        if nd.is_leaf() or nd.is_list():
            lo, hi = nd.world_aabb()
            rec[8:11] = lo
            rec[11] = hi[0]
            rec[12:14] = hi[1:3]
        u[14] = nd.typecode
        gt = 0
        if nd.transform is not None:
            gt = self.addTran(nd.transform)
        u[15] = (gt & 0x7FFFFFFF) | (0x80000000 if nd.complement else 0)
        return rec

    def addPrim(self, root: Node, boundary: int, mesh_idx=0, pad=1e-3):
        """Serialize a tree into level-order complete-binary-tree nodes."""
        h = root.height()
        num = (1 << (h + 1)) - 1
        slots = [None] * num

        def place(nd, i):
            slots[i - 1] = nd
            if not (nd.is_leaf() or nd.is_list()):
                place(nd.left, 2 * i)
                place(nd.right, 2 * i + 1)

        place(root, 1)
        node_offset = len(self.nodes)
        for k, nd in enumerate(slots):
            self.nodes.append(self._node_record(nd, boundary, k))
        if not (root.is_leaf() or root.is_list()):
            self.nodes[node_offset].view(np.uint32)[0] = num      # subNum
            self.nodes[node_offset].view(np.uint32)[1] = 0        # subOffset
        sub_offset = num                                          # list leaves follow the tree
        for k, nd in enumerate(slots):
            if nd is not None and nd.is_list():
                hdr = self.nodes[node_offset + k].view(np.uint32)
                hdr[0], hdr[1] = len(nd.subs), sub_offset
                for j, leaf in enumerate(nd.subs):
                    self.nodes.append(self._node_record(leaf, boundary, sub_offset + j))
                sub_offset += len(nd.subs)
        num = sub_offset
        lo, hi = root.world_aabb()
        ext = hi - lo
        lo = lo - pad * ext - 1e-2
        hi = hi + pad * ext + 1e-2
        rec = np.zeros(16, np.float32)
        i = rec.view(np.int32)
        u = rec.view(np.uint32)
        i[0], i[1], i[2], i[3] = num, node_offset, 0, 0
        u[4], u[5], u[6], u[7] = 0, mesh_idx, 0, 0
        rec[8:11] = lo
        rec[11] = hi[0]
        rec[12:14] = hi[1:3]
        u[15] = len(self.prims)
        self.prims.append(rec)
        return len(self.prims) - 1

    def addSolid(self, label: str, prim_indices):
        prim_indices = list(prim_indices)
        assert prim_indices == list(range(prim_indices[0], prim_indices[0] + len(prim_indices)))
        s = np.zeros((), SOLID_DTYPE)
        s["label"] = label.encode()[:16]
        s["numPrim"] = len(prim_indices)
        s["primOffset"] = prim_indices[0]
        los = np.array([self.prims[p][8:11] for p in prim_indices])
        his = np.array([np.r_[self.prims[p][11], self.prims[p][12:14]] for p in prim_indices])
        lo, hi = los.min(0), his.max(0)
        c = 0.5 * (lo + hi)
        s["center_extent"] = (c[0], c[1], c[2], float((hi - lo).max() / 2))
        self.solids.append(s)
        return len(self.solids) - 1

    def addInstance(self, transform, gas_idx, sensor_identifier=-1, sensor_index=-1):
        """transform: 4x4 row-vector object->world. sensor_identifier -1 = not a sensor."""
        q = np.asarray(transform, np.float32).copy().reshape(4, 4)
        iv = q.view(np.int32)
        iv[0, 3] = len(self.insts)
        iv[1, 3] = gas_idx
        iv[2, 3] = sensor_identifier + 1
        iv[3, 3] = sensor_index
        self.insts.append(q.reshape(16))
        return len(self.insts) - 1

    def arrays(self):
        node = np.array(self.nodes, np.float32).reshape(-1, 4, 4)
        prim = np.array(self.prims, np.float32).reshape(-1, 4, 4)
        solid = np.array(self.solids, SOLID_DTYPE)
        tran = (np.array(self.trans, np.float32).reshape(-1, 4, 4) if self.trans
                else np.zeros((0, 4, 4), np.float32))
        itra = (np.array([np.linalg.inv(t) for t in self.trans], np.float32).reshape(-1, 4, 4)
                if self.trans else np.zeros((0, 4, 4), np.float32))
        inst = np.array(self.insts, np.float32).reshape(-1, 4, 4)
        plan = np.array(self.plans, np.float32).reshape(-1, 4)
        return dict(node=node, prim=prim, solid=solid, tran=tran, itra=itra, inst=inst, plan=plan)


def rotation_z_to(direction):
    """Row-vector rotation whose local +z maps to `direction` (unit)."""
    d = np.asarray(direction, float)
    d = d / np.linalg.norm(d)
    a = np.array([1.0, 0.0, 0.0]) if abs(d[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    x = np.cross(a, d)
    x /= np.linalg.norm(x)
    y = np.cross(d, x)
    r = np.eye(4)
    r[0, :3], r[1, :3], r[2, :3] = x, y, d
    return r


def fibonacci_sphere(n):
    """Fibonacci lattice: z_i = 1-(2i+1)/n, phi_i = i*pi*(3-sqrt(5))."""
    i = np.arange(n, dtype=float)
    z = 1.0 - (2.0 * i + 1.0) / n
    r = np.sqrt(1.0 - z * z)
    phi = i * np.pi * (3.0 - np.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
