"""
The benchmark workloads, built in code (the reference ships no data fixtures).

  config0  1M photons, one torch genstep (T_SPHERE, radius 0: isotropic from the
           origin), 440 nm, seed 42. Water sphere r=1000 mm (n 1.333, absorption
           20 m, Rayleigh 30 m) in a vacuum box of half-side 2000 mm whose inner
           surface detects with efficiency 1.
  config1  10M photons from 1,000 torch gensteps placed uniformly in the r<5 m
           liquid-scintillator ball, isotropic, wavelength uniform in 200-800 nm.
           LS sphere r=5000 (n 1.49, absorption 25 m, Rayleigh 40 m, reemission 0.4
           below 410 nm, Gaussian emission 430+-20 nm) inside a 120 mm acrylic shell
           (n 1.50) in water (n 1.33) with 2,000 instanced detectors on a Fibonacci
           lattice at r=5.5 m, local +z towards the centre. Each detector is
           difference(union(z-cut ellipsoid 254/254/190 mm z>-50, cylinder r127
           h200), cylinder r120) with a 0.25 detect / 0.75 absorb surface.
  config2  64 events x Poisson(200) scintillation gensteps x Poisson(5000)
           photons in the config1 geometry scaled to 20,000 detectors at r=19.4 m
           around a 17.7 m LS sphere.
  config3  config1 with LS and water absorption/Rayleigh lengths as 64 bins over
           200-800 nm each (256 parameters), each bin x(1+U(-0.2,0.2)) (seed 42).
  config4  256 events as config2 (about 256M photons) with the config3 parameterisation.

Choices the workload statements leave open (documented, fixed here):
  * world box half-side 6500 mm (config1) / 21000 mm (config2), inner surface absorbs;
  * water absorption 20 m and Rayleigh 30 m (as config0), acrylic absorption and
    Rayleigh 1e6 mm; detector solids are vacuum inside;
  * detector neck z in [-240,-40] mm (h 200 mm as stated; its top is 10 mm above
    the ellipsoid's z=-50 cut, so the union has no coincident faces), inner bore
    z in [-260,-60] mm (solid frame);
  * torch subtype T_POINT_ISO=100: T_SPHERE directions about gs.pos, wavelength
    uniform in [gs.wavelength, gs.wavelength+gs.weight] (sysrap/storch.h has no
    isotropic point source at gs.pos);
  * genstep placement and Poisson counts from numpy.random.default_rng(seed).
"""
from __future__ import annotations

import numpy as np

from .foundry import (Foundry, Node, CSG_UNION, CSG_DIFFERENCE, scale_transform,
                      rotation_z_to, fibonacci_sphere)
from .physics import BoundaryLib, Material, Surface, gaussian_icdf, EMS_SURFACE, EMS_ZPLUS_SENSOR_A

# sysrap/OpticksGenstep.h, sysrap/storchtype.h
OpticksGenstep_TORCH = 6
OpticksGenstep_SCINTILLATION = 16
T_DISC, T_LINE, T_POINT, T_CIRCLE, T_RECTANGLE, T_SPHERE_MARSAGLIA, T_SPHERE = 1, 2, 3, 4, 5, 6, 7
T_POINT_ISO = 100


def torch_genstep(numphoton, type_=T_SPHERE, pos=(0, 0, 0), time=0.0, mom=(0, 0, 1), weight=0.0,
                  pol=(1, 0, 0), wavelength=440.0, zenith=(0.0, 1.0), azimuth=(0.0, 1.0),
                  radius=0.0, distance=0.0, mode=0):
    """quad6 genstep with the storch layout (sysrap/storch.h:42-74)."""
    g = np.zeros(24, np.float32)
    u = g.view(np.uint32)
    u[0] = OpticksGenstep_TORCH
    u[3] = numphoton
    g[4:7] = pos
    g[7] = time
    g[8:11] = mom
    g[11] = weight
    g[12:15] = pol
    g[15] = wavelength
    g[16:18] = zenith
    g[18:20] = azimuth
    g[20] = radius
    g[21] = distance
    u[22] = mode
    u[23] = type_
    return g.reshape(6, 4)


def scint_genstep(numphoton, pos, time=0.0, delta=(0, 0, 0), step_length=0.0, charge=0.0,
                  weight=1.0, mean_velocity=1.0, scint_time=0.0):
    """quad6 genstep with the sscint layout (sysrap/sscint.h:34-62)."""
    g = np.zeros(24, np.float32)
    u = g.view(np.uint32)
    u[0] = OpticksGenstep_SCINTILLATION
    u[3] = numphoton
    g[4:7] = pos
    g[7] = time
    g[8:11] = delta
    g[11] = step_length
    g[13] = charge
    g[14] = weight
    g[15] = mean_velocity
    g[20] = scint_time
    return g.reshape(6, 4)


OpticksGenstep_CERENKOV = 15


def cerenkov_genstep(numphoton, matline, pos, time, delta, beta, wmin=200.0, wmax=600.0, n_max=1.333,
                     n1=None, n2=None, charge=-1.0, weight=1.0, post_beta=None):
    """quad6 genstep with the scerenkov layout (sysrap/scerenkov.h:29-58) for a charged
    track step of displacement `delta` (mm) at speed beta (post-step speed post_beta);
    matline is the bnd line of the medium (boundary*4 + 3 for a boundary's inner material)."""
    g = np.zeros(24, np.float32)
    u = g.view(np.uint32)
    u[0] = OpticksGenstep_CERENKOV
    u[1] = 0
    u[2] = matline
    u[3] = numphoton
    g[4:7] = pos
    g[7] = time
    g[8:11] = delta
    step_length = float(np.linalg.norm(delta))
    g[11] = step_length
    g[12] = 0.0
    g[13] = charge
    g[14] = weight
    c = 299.792458
    g[15] = beta * c
    bi = 1.0 / beta
    max_cos = bi / n_max
    g[16] = bi
    g[17] = wmin
    g[18] = wmax
    g[19] = max_cos
    g[20] = (1.0 - max_cos) * (1.0 + max_cos)
    g[21] = numphoton if n1 is None else n1
    g[22] = numphoton if n2 is None else n2
    g[23] = (beta if post_beta is None else post_beta) * c
    return g.reshape(6, 4)


OpticksGenstep_FRAME = 17
GRID_XYZ, GRID_YZ, GRID_XZ, GRID_XY = 0, 1, 2, 3


def grid_axes(nx, ny, nz):
    """SGenstep::GridAxes (sysrap/SGenstep.h:138-154): planar grids have one zero extent."""
    if nx == 0 and ny > 0 and nz > 0:
        return GRID_YZ
    if nx > 0 and ny == 0 and nz > 0:
        return GRID_XZ
    if nx > 0 and ny > 0 and nz == 0:
        return GRID_XY
    return GRID_XYZ


def frame_gensteps(ce, cegs, gridscale=0.1, geotran=None, ce_offset=((0.0, 0.0, 0.0),), ce_scale=True):
    """OpticksGenstep_FRAME gensteps of a simtrace grid, as SFrameGenstep::MakeCenterExtentGenstep
    (sysrap/SFrameGenstep.cc:566-662) lays them out:
      ce      (cx, cy, cz, extent) of the frame
      cegs    (ix0, ix1, iy0, iy1, iz0, iz1, photons_per_genstep, high): grid index ranges
              (inclusive), photons per grid point and a resolution multiplier 1..8
      geotran 4x4 local->world transform in row-vector form (translation in the last row),
              identity when None
    One genstep per (offset, ix, iy, iz); q0 = (17, gridaxes, gsid, photons), q1 = (offset, 1),
    q2..q5 = translate(ix, iy, iz scaled) followed by geotran, as a qat4."""
    ix0, ix1, iy0, iy1, iz0, iz1, ppg, high = (int(v) for v in cegs)
    assert 1 <= high <= 8
    scale = float(gridscale) / high
    ix0, ix1, iy0, iy1, iz0, iz1 = (v * high for v in (ix0, ix1, iy0, iy1, iz0, iz1))
    axes = grid_axes((ix1 - ix0) // 2, (iy1 - iy0) // 2, (iz1 - iz0) // 2)
    local_scale = scale * float(ce[3]) if ce_scale else scale
    G = np.eye(4) if geotran is None else np.asarray(geotran, np.float64).reshape(4, 4)
    out = []
    for ip, off in enumerate(ce_offset):
        for ix in range(ix0, ix1 + 1):
            for iy in range(iy0, iy1 + 1):
                for iz in range(iz0, iz1 + 1):
                    S = np.eye(4)
                    S[3, :3] = (ix * local_scale, iy * local_scale, iz * local_scale)
                    T = (S @ G).astype(np.float32)
                    g = np.zeros(24, np.float32)
                    u = g.view(np.uint32)
                    u[0] = OpticksGenstep_FRAME
                    u[1] = axes
                    u[2] = np.array([ix, iy, iz, ip], np.int8).view(np.uint32)[0]   # SGenstep::GenstepID
                    u[3] = abs(ppg)
                    g[4:7] = off
                    g[7] = 1.0
                    g[8:24] = T.reshape(-1)
                    out.append(g.reshape(6, 4))
    return np.stack(out) if out else np.zeros((0, 6, 4), np.float32)


class Workload:
    def __init__(self, name, geometry, physics, gensteps, seed=42, events=None, meta=None):
        self.name = name
        self.geometry = geometry      # dict of foundry arrays
        self.physics = physics        # dict bnd/optical/... + icdf
        self.gensteps = gensteps      # (G,6,4) float32 (single event)
        self.events = events          # list of (G_e,6,4) for batched workloads
        self.seed = seed
        self.meta = meta or {}

    @property
    def num_photon(self):
        if self.events is not None:
            return int(sum(e.view(np.uint32)[:, 0, 3].sum() for e in self.events))
        return int(self.gensteps.view(np.uint32)[:, 0, 3].sum())


def _scint_materials(lib, water_n=1.33):
    lib.add_material(Material("Vacuum", 1.0, 1e9, 1e9))
    lib.add_material(Material("Water", water_n, 20000.0, 30000.0))
    lib.add_material(Material("Acrylic", 1.50, 1e6, 1e6))
    lib.add_material(Material("LS", 1.49, 25000.0, 40000.0,
                              reemission=lambda wl: np.where(wl < 410.0, 0.4, 0.0)))
    lib.add_material(Material("Rock", 1.5, 1.0, 1e9))
    lib.add_surface(Surface("absSurf", detect=0.0, absorb=1.0))
    lib.add_surface(Surface("detSurf", detect=0.25, absorb=0.75))


def detector_tree():
    """difference(union(z-cut ellipsoid, neck cylinder), inner bore cylinder)."""
    ell = Node.zsphere(190.0, -50.0, 190.0)
    ell.transform = scale_transform(254.0 / 190.0, 254.0 / 190.0, 1.0)
    neck = Node.cylinder(127.0, -240.0, -40.0)
    bore = Node.cylinder(120.0, -260.0, -60.0)
    return Node.op(CSG_DIFFERENCE, Node.op(CSG_UNION, ell, neck), bore)


def cerenkov_water(num_photon=200_000, seed=42, n_gs=20):
    """config0's water sphere with Cerenkov gensteps of 10 mm track steps inside the
    water (beta 0.99 -> 0.985, 200-600 nm, mean photon yield falling 100 -> 80 along
    each step), from numpy.random.default_rng(seed) positions and directions."""
    w = config0(num_photon=1, seed=seed)
    water_line = w.physics["boundary_names"].index("Vacuum///Water") * 4 + 3
    rng = np.random.default_rng(seed)
    per = num_photon // n_gs
    gs = []
    for k in range(n_gs):
        pos = rng.uniform(-400.0, 400.0, 3)
        d = rng.normal(size=3)
        d = 10.0 * d / np.linalg.norm(d)
        gs.append(cerenkov_genstep(per, water_line, pos, float(rng.uniform(0, 1)), d, 0.99,
                                   n1=100.0, n2=80.0, post_beta=0.985))
    w.gensteps = np.stack(gs).astype(np.float32)
    w.name = "cerenkov_water"
    w.meta.update(water_line=water_line)
    return w


def config0(num_photon=1_000_000, seed=42, water_absorption=20000.0, water_rayleigh=30000.0, sensor=False):
    """sensor=True gives the world instance sensor_identifier 0, so its detector
    surface counts as detector 0 for per-detector yields."""
    lib = BoundaryLib()
    lib.add_material(Material("Vacuum", 1.0, 1e9, 1e9))
    lib.add_material(Material("Water", 1.333, water_absorption, water_rayleigh))
    lib.add_surface(Surface("detSurf", detect=1.0, absorb=0.0))
    b_world = lib.boundary("Vacuum/detSurf/detSurf/Vacuum")
    b_sphere = lib.boundary("Vacuum///Water")
    fd = Foundry()
    p0 = fd.addPrim(Node.box3(4000.0, 4000.0, 4000.0), b_world)
    p1 = fd.addPrim(Node.sphere(1000.0), b_sphere)
    s0 = fd.addSolid("world", [p0, p1])
    if sensor:
        fd.addInstance(np.eye(4), s0, sensor_identifier=0, sensor_index=0)
    else:
        fd.addInstance(np.eye(4), s0)
    phys = lib.arrays()
    phys["icdf"], phys["icdf_hd_factor"] = gaussian_icdf()
    phys["boundary_names"] = lib.boundaries
    gs = torch_genstep(num_photon, T_SPHERE, wavelength=440.0, radius=0.0, distance=0.0)[None]
    return Workload("config0", fd.arrays(), phys, gs, seed=seed,
                    meta=dict(world_half=2000.0, sphere_r=1000.0))


def _scint_world(num_det, det_r, ls_r, shell, world_half):
    lib = BoundaryLib()
    _scint_materials(lib)
    b_world = lib.boundary("Rock//absSurf/Water")
    b_acr = lib.boundary("Water///Acrylic")
    b_ls = lib.boundary("Acrylic///LS")
    b_det = lib.boundary("Water/detSurf//Vacuum")
    fd = Foundry()
    p0 = fd.addPrim(Node.box3(2 * world_half, 2 * world_half, 2 * world_half), b_world)
    p1 = fd.addPrim(Node.sphere(ls_r + shell), b_acr)
    p2 = fd.addPrim(Node.sphere(ls_r), b_ls)
    s0 = fd.addSolid("global", [p0, p1, p2])
    pd = fd.addPrim(detector_tree(), b_det, mesh_idx=1)
    s1 = fd.addSolid("detector", [pd])
    fd.addInstance(np.eye(4), s0)
    dirs = fibonacci_sphere(num_det)
    for i, d in enumerate(dirs):
        t = rotation_z_to(-d)
        t[3, :3] = det_r * d
        fd.addInstance(t, s1, sensor_identifier=i, sensor_index=i)
    phys = lib.arrays()
    phys["icdf"], phys["icdf_hd_factor"] = gaussian_icdf()
    phys["boundary_names"] = lib.boundaries
    return fd.arrays(), phys


def leaf_zoo_nodes():
    """One prim per entry: the cone, hyperboloid, convex polyhedra, and phicut /
    halfspace leaves inside Boolean trees, with one transformed cone."""
    from .foundry import Node, CSG_INTERSECTION, CSG_DIFFERENCE
    cube = np.array([[1, 0, 0, 70], [-1, 0, 0, 70], [0, 1, 0, 70], [0, -1, 0, 70], [0, 0, 1, 70], [0, 0, -1, 70]],
                    np.float32)
    tn = np.array([[1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1]], float) / np.sqrt(3.0)
    tetra = np.concatenate([tn, np.full((4, 1), 60.0)], axis=1).astype(np.float32)
    rot = rotation_z_to([0.3, -0.5, 0.81])
    rot[3, :3] = (15.0, -20.0, 10.0)
    tc = Node.cone(90.0, -60.0, 20.0, 70.0)
    tc.transform = rot
    return [
        Node.cone(100.0, -50.0, 30.0, 80.0),
        Node.hyperboloid(60.0, 80.0, -100.0, 120.0),
        Node.convexpolyhedron(cube, (np.full(3, -70.0), np.full(3, 70.0))),
        Node.convexpolyhedron(tetra, (np.full(3, -110.0), np.full(3, 110.0))),
        Node.op(CSG_INTERSECTION, Node.cylinder(120.0, -80.0, 80.0), Node.phicut(0.3, 2.2)),
        Node.op(CSG_INTERSECTION, Node.sphere(150.0), Node.phicut(0.5, 4.5)),
        Node.op(CSG_DIFFERENCE, Node.sphere(140.0), Node.halfspace((0.0, 0.0, 1.0), 40.0)),
        Node.op(CSG_INTERSECTION, Node.box3(200.0, 200.0, 200.0), Node.halfspace((1.0, 1.0, 0.0), 30.0)),
        tc,
    ]


def list_zoo_nodes():
    """One prim per entry: list nodes at the prim root and inside Boolean trees
    (the subOffset of each list counts from the prim's root node)."""
    from .foundry import Node, CSG_CONTIGUOUS, CSG_DISCONTIGUOUS, CSG_OVERLAP, CSG_DIFFERENCE, CSG_INTERSECTION, CSG_UNION
    beads = [Node.sphere(60.0, (x, 0.0, 0.0)) for x in (-80.0, 0.0, 80.0)]
    bx = Node.box3(50.0, 50.0, 50.0)
    bx.transform = np.eye(4)
    bx.transform[3, :3] = (0.0, 90.0, 30.0)
    islands = [Node.sphere(35.0, (-100.0, -60.0, 0.0)), Node.sphere(35.0, (100.0, -60.0, 0.0)), bx]
    lens = [Node.sphere(120.0, (0.0, 0.0, -70.0)), Node.sphere(120.0, (0.0, 0.0, 70.0)), Node.cylinder(90.0, -60.0, 60.0)]
    cross = [Node.box3(200.0, 60.0, 60.0), Node.box3(60.0, 200.0, 60.0), Node.box3(60.0, 60.0, 200.0)]
    return [
        Node.list(CSG_CONTIGUOUS, beads),
        Node.list(CSG_DISCONTIGUOUS, islands),
        Node.list(CSG_OVERLAP, lens),
        Node.op(CSG_DIFFERENCE, Node.list(CSG_CONTIGUOUS, [Node.sphere(60.0, (x, 0.0, 0.0)) for x in (-80.0, 0.0, 80.0)]),
                Node.cylinder(25.0, -200.0, 200.0)),
        Node.op(CSG_INTERSECTION, Node.list(CSG_CONTIGUOUS, cross), Node.sphere(90.0)),
        Node.op(CSG_UNION, Node.list(CSG_DISCONTIGUOUS, [Node.sphere(30.0, (-90.0, 0.0, 0.0)), Node.sphere(30.0, (90.0, 0.0, 0.0))]),
                Node.list(CSG_OVERLAP, [Node.sphere(70.0, (0.0, 0.0, -30.0)), Node.sphere(70.0, (0.0, 0.0, 30.0))])),
    ]


def leaf_zoo(num_photon=200_000, seed=42, spacing=400.0, nodes=None, name="leaf_zoo"):
    """The leaf_zoo_nodes solids (or `nodes`) as acrylic pieces on a 3x3 grid in a
    water box with an absorbing, partly detecting wall; photons from a torch point
    source at the centre of each cell's corner region."""
    lib = BoundaryLib()
    lib.add_material(Material("Water", 1.33, 20000.0, 30000.0))
    lib.add_material(Material("Acrylic", 1.50, 1e6, 1e6))
    lib.add_surface(Surface("wallSurf", detect=0.5, absorb=0.5))
    b_world = lib.boundary("Water/wallSurf/wallSurf/Water")
    b_obj = lib.boundary("Water///Acrylic")
    fd = Foundry()
    half = 1.6 * spacing
    p0 = fd.addPrim(Node.box3(2 * half, 2 * half, 2 * half), b_world)
    s0 = fd.addSolid("world", [p0])
    fd.addInstance(np.eye(4), s0)
    nodes = leaf_zoo_nodes() if nodes is None else nodes
    for k, nd in enumerate(nodes):
        pk = fd.addPrim(nd, b_obj)
        sk = fd.addSolid(f"leaf{k}", [pk])
        t = np.eye(4)
        t[3, :3] = ((k % 3) - 1) * spacing, ((k // 3) - 1) * spacing, 0.0
        fd.addInstance(t, sk)
    phys = lib.arrays()
    phys["icdf"], phys["icdf_hd_factor"] = gaussian_icdf()
    phys["boundary_names"] = lib.boundaries
    gs = np.stack([torch_genstep(num_photon // 4, T_POINT_ISO, pos=p, wavelength=300.0, weight=300.0)
                   for p in ((200.0, 200.0, 0.0), (-200.0, 200.0, 50.0), (200.0, -200.0, -50.0), (-200.0, -200.0, 0.0))]
                  ).astype(np.float32)
    return Workload(name, fd.arrays(), phys, gs, seed=seed, meta=dict(n_leaf_prims=len(nodes)))


def boolean_zoo_nodes():
    """One prim per entry: Boolean trees of height 1 (each operator), every shape of
    a height-2 tree (operators on both sides, a leaf on the right, a leaf on the
    left, nested differences) and one tree of height 3."""
    from .foundry import Node, CSG_DIFFERENCE, CSG_INTERSECTION, CSG_UNION
    U, I, D = CSG_UNION, CSG_INTERSECTION, CSG_DIFFERENCE
    op = Node.op
    return [
        op(D, Node.sphere(150.0), Node.sphere(100.0)),
        op(I, Node.box3(240.0, 240.0, 240.0), Node.sphere(150.0)),
        op(U, Node.sphere(100.0, (-50.0, 0.0, 0.0)), Node.sphere(100.0, (50.0, 0.0, 0.0))),
        op(D, op(U, Node.box3(200.0, 200.0, 200.0), Node.sphere(100.0, (0.0, 0.0, 80.0))),
           op(I, Node.cylinder(60.0, -150.0, 150.0), Node.sphere(130.0))),
        op(D, op(U, Node.zsphere(140.0, -40.0, 140.0), Node.cylinder(70.0, -180.0, -30.0)),
           Node.cylinder(60.0, -190.0, -50.0)),
        op(I, Node.sphere(160.0), op(D, Node.box3(260.0, 260.0, 260.0), Node.cylinder(70.0, -200.0, 200.0))),
        op(U, op(D, Node.sphere(150.0), Node.box3(150.0, 150.0, 150.0)), Node.sphere(60.0)),
        op(D, op(D, Node.box3(250.0, 250.0, 250.0), Node.sphere(150.0)), Node.cylinder(50.0, -200.0, 200.0)),
        op(U, op(D, op(I, Node.box3(200.0, 200.0, 200.0), Node.sphere(130.0)), Node.cylinder(40.0, -150.0, 150.0)),
           Node.sphere(50.0)),
    ]


def boolean_zoo(num_photon=200_000, seed=42, spacing=400.0):
    """leaf_zoo world with the boolean_zoo_nodes solids"""
    return leaf_zoo(num_photon, seed, spacing, nodes=boolean_zoo_nodes(), name="boolean_zoo")


def list_zoo(num_photon=200_000, seed=42, spacing=400.0):
    """leaf_zoo world with the list_zoo_nodes solids"""
    return leaf_zoo(num_photon, seed, spacing, nodes=list_zoo_nodes(), name="list_zoo")


SURFACE_ZOO_SOURCES = {"centre_A": (0.0, 0.0, 0.0), "outside_A": (-1500.0, 0.0, 0.0),
                       "centre_B": (1400.0, 0.0, 0.0), "above_B": (1400.0, 0.0, 900.0),
                       "below_B": (1400.0, 0.0, -900.0)}


def surface_zoo(num_photon=200_000, seed=42, mix=(0.2, 0.3, 0.3, 0.2), sens=(0.1, 0.4, 0.3, 0.2)):
    """Surfaces with every propagate_at_surface outcome and the zplus_sensor_A dispatch
    (qudarap/qsim.h:1680-1758, 2296-2320). mix / sens = (detect, absorb, specular, diffuse).

      sphere A  r 800 at the origin, boundary Vacuum/mixSurf/mixSurf/Clear: an ordinary
                surface (ems Surface) on both sides: SA, SD, DR (Lambertian), SR (mirror)
      sphere B  r 300, its own instance at (1400, 0, 0) (sensor 0), boundary
                Vacuum/sensA/sensA/Clear with ems Surface_zplus_sensor_A: hits with
                lposcost >= 0 (upper hemisphere of its frame) detect after one burned
                uniform, hits with lposcost < 0 take the surface fractions
      world     box half-side 2000 mm, Vacuum/detSurf/detSurf/Vacuum (detect 1)
      Clear     n 1.333, absorption and Rayleigh lengths 1e6 mm

    One genstep of num_photon // 5 photons per SURFACE_ZOO_SOURCES entry, 440 nm:
    T_SPHERE at A's centre, T_POINT_ISO elsewhere (outside A, centre of B, above
    and below B)."""
    lib = BoundaryLib()
    lib.add_material(Material("Vacuum", 1.0, 1e9, 1e9))
    lib.add_material(Material("Clear", 1.333, 1e6, 1e6))
    lib.add_surface(Surface("detSurf", detect=1.0, absorb=0.0))
    d, a, sp, df = mix
    lib.add_surface(Surface("mixSurf", detect=d, absorb=a, specular=sp, diffuse=df))
    d, a, sp, df = sens
    lib.add_surface(Surface("sensA", detect=d, absorb=a, specular=sp, diffuse=df, ems=EMS_ZPLUS_SENSOR_A))
    b_world = lib.boundary("Vacuum/detSurf/detSurf/Vacuum")
    b_a = lib.boundary("Vacuum/mixSurf/mixSurf/Clear")
    b_b = lib.boundary("Vacuum/sensA/sensA/Clear")
    fd = Foundry()
    p0 = fd.addPrim(Node.box3(4000.0, 4000.0, 4000.0), b_world)
    p1 = fd.addPrim(Node.sphere(800.0), b_a)
    s0 = fd.addSolid("world", [p0, p1])
    pb = fd.addPrim(Node.sphere(300.0), b_b)
    s1 = fd.addSolid("sensor", [pb])
    fd.addInstance(np.eye(4), s0)
    t = np.eye(4)
    t[3, :3] = SURFACE_ZOO_SOURCES["centre_B"]
    fd.addInstance(t, s1, sensor_identifier=0, sensor_index=0)
    phys = lib.arrays()
    phys["icdf"], phys["icdf_hd_factor"] = gaussian_icdf()
    phys["boundary_names"] = lib.boundaries
    per = num_photon // len(SURFACE_ZOO_SOURCES)
    gs = [torch_genstep(per, T_SPHERE, wavelength=440.0)]
    gs += [torch_genstep(per, T_POINT_ISO, pos=p, wavelength=440.0)
           for k, p in SURFACE_ZOO_SOURCES.items() if k != "centre_A"]
    return Workload("surface_zoo", fd.arrays(), phys, np.stack(gs).astype(np.float32), seed=seed,
                    meta=dict(mix=mix, sens=sens, per=per, r_a=800.0, r_b=300.0,
                              sources=list(SURFACE_ZOO_SOURCES.values())))


def prim_clusters(num_photon=200_000, seed=42, rows=(10, 7)):
    """Solids of several prims, on both sides of the front/TLAS split: a water box
    whose world solid also holds two planes of rows[0] x rows[1] small acrylic beads
    (z = +-700 mm; more prims than one BLAS leaf link holds, so the front solid is
    walked through its BLAS), and eight rotated instances of a five-prim acrylic
    cluster (sphere, translated box, cylinder and z-sphere, offset sphere) at
    (+-400, +-400, +-250) mm; a torch point source at the centre."""
    lib = BoundaryLib()
    lib.add_material(Material("Water", 1.33, 20000.0, 30000.0))
    lib.add_material(Material("Acrylic", 1.50, 1e6, 1e6))
    lib.add_surface(Surface("wallSurf", detect=0.5, absorb=0.5))
    b_world = lib.boundary("Water/wallSurf/wallSurf/Water")
    b_obj = lib.boundary("Water///Acrylic")
    fd = Foundry()
    world = [fd.addPrim(Node.box3(2000.0, 2000.0, 2000.0), b_world)]
    nx, ny = rows
    for z in (-700.0, 700.0):
        for i in range(nx):
            for j in range(ny):
                c = (-810.0 + 1620.0 * i / (nx - 1), -810.0 + 1620.0 * j / (ny - 1), z)
                world.append(fd.addPrim(Node.sphere(15.0, c), b_obj))
    s0 = fd.addSolid("world", world)

    def moved(node, xyz):
        node.transform = np.eye(4)
        node.transform[3, :3] = xyz
        return node
    cluster = [fd.addPrim(Node.sphere(60.0), b_obj),
               fd.addPrim(moved(Node.box3(80.0, 80.0, 80.0), (150.0, 0.0, 0.0)), b_obj),
               fd.addPrim(moved(Node.cylinder(40.0, -50.0, 50.0), (-150.0, 0.0, 0.0)), b_obj),
               fd.addPrim(moved(Node.zsphere(50.0, -20.0, 50.0), (0.0, 150.0, 0.0)), b_obj),
               fd.addPrim(Node.sphere(30.0, (0.0, -150.0, 0.0)), b_obj)]
    s1 = fd.addSolid("cluster", cluster)
    fd.addInstance(np.eye(4), s0)
    for k in range(8):
        c = np.array([400.0 if k & 1 else -400.0, 400.0 if k & 2 else -400.0, 250.0 if k & 4 else -250.0])
        t = rotation_z_to(c + np.array([30.0 * k, -20.0, 45.0]))
        t[3, :3] = c
        fd.addInstance(t, s1)
    phys = lib.arrays()
    phys["icdf"], phys["icdf_hd_factor"] = gaussian_icdf()
    phys["boundary_names"] = lib.boundaries
    gs = torch_genstep(num_photon, T_POINT_ISO, pos=(0.0, 0.0, 0.0), wavelength=440.0)[None]
    return Workload("prim_clusters", fd.arrays(), phys, gs, seed=seed,
                    meta=dict(n_world_prims=len(world), n_cluster_prims=len(cluster), b_obj=b_obj, b_world=b_world))


def uniform_ball(rng, n, r):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * (r * rng.random(n) ** (1.0 / 3.0))[:, None]


def config1(num_gs=1000, per_gs=10_000, seed=42):
    geo, phys = _scint_world(2000, 5500.0, 5000.0, 120.0, 6500.0)
    rng = np.random.default_rng(seed)
    pos = uniform_ball(rng, num_gs, 5000.0)
    gs = np.stack([torch_genstep(per_gs, T_POINT_ISO, pos=p, wavelength=200.0, weight=600.0)
                   for p in pos]).astype(np.float32)
    return Workload("config1", geo, phys, gs, seed=seed,
                    meta=dict(ls_r=5000.0, det_r=5500.0, num_det=2000, world_half=6500.0))


def config2(num_events=64, mean_gs=200, mean_photons=5000, seed=42):
    geo, phys = _scint_world(20000, 19400.0, 17700.0, 120.0, 21000.0)
    events = []
    for e in range(num_events):
        rng = np.random.default_rng([seed, e])
        ng = max(1, int(rng.poisson(mean_gs)))
        counts = rng.poisson(mean_photons, size=ng).astype(np.int64)
        counts[counts == 0] = 1
        pos = uniform_ball(rng, ng, 17700.0)
        dirs = rng.normal(size=(ng, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        gs = np.stack([scint_genstep(int(c), p, time=float(rng.random() * 10.0),
                                     delta=1.0 * d, step_length=1.0, charge=1.0,
                                     mean_velocity=0.9 * 299.792458, scint_time=1.5)
                       for c, p, d in zip(counts, pos, dirs)]).astype(np.float32)
        events.append(gs)
    return Workload("config2", geo, phys, None, seed=seed, events=events,
                    meta=dict(ls_r=17700.0, det_r=19400.0, num_det=20000, world_half=21000.0))


CONFIG3_BASE_LENGTHS = {"LS": (25000.0, 40000.0), "Water": (20000.0, 30000.0)}


def detector_shell(num_det=50_000, num_gs=8, per_gs=2500, seed=3):
    """The config1 layout at the instance count of a large detector: num_det
    detectors on a Fibonacci lattice at the radius that keeps config2's area per
    detector (about 0.236 m^2; 30.6 m for 50,000), LS sphere 28 m, world half-side
    32 m, and num_gs isotropic torch points in the LS (numpy default_rng(seed))."""
    det_r = float(np.sqrt(num_det * 0.2365e6 / (4.0 * np.pi)))
    ls_r, world_half = det_r - 2600.0, det_r + 1400.0
    geo, phys = _scint_world(num_det, det_r, ls_r, 120.0, world_half)
    rng = np.random.default_rng(seed)
    pos = uniform_ball(rng, num_gs, ls_r)
    gs = np.stack([torch_genstep(per_gs, T_POINT_ISO, pos=p, wavelength=200.0, weight=600.0)
                   for p in pos]).astype(np.float32)
    return Workload("detector_shell", geo, phys, gs, seed=seed,
                    meta=dict(ls_r=ls_r, det_r=det_r, num_det=num_det, world_half=world_half))


def config3(num_gs=1000, per_gs=10_000, seed=42, n_bin=64):
    """config1 geometry and photons; LS and water absorption and Rayleigh lengths
    as n_bin bins over 200-800 nm (theta, mm), each the config1 value times
    (1 + U(-0.2, 0.2)) from numpy.random.default_rng(seed)."""
    from .adjoint import LengthParams, perturbed_theta
    w = config1(num_gs=num_gs, per_gs=per_gs, seed=seed)
    params = LengthParams(w.physics, ("LS", "Water"), n_bin=n_bin, lo=200.0, hi=800.0)
    theta = perturbed_theta(params, CONFIG3_BASE_LENGTHS, seed=seed, frac=0.2)
    w.physics["bnd"] = params.apply(w.physics["bnd"], theta)
    w.name = "config3"
    w.meta.update(params=params, theta=theta)
    return w


def config4(num_events=256, mean_gs=200, mean_photons=5000, seed=42, n_bin=64):
    """config2 geometry (20,000 detectors at 19.4 m, LS r 17.7 m) with num_events
    Poisson(mean_gs) x Poisson(mean_photons) scintillation events and the LS/water
    absorption and Rayleigh lengths as n_bin bins each (x(1+U(-0.2,0.2)), seed)."""
    from .adjoint import LengthParams, perturbed_theta
    w = config2(num_events=num_events, mean_gs=mean_gs, mean_photons=mean_photons, seed=seed)
    params = LengthParams(w.physics, ("LS", "Water"), n_bin=n_bin, lo=200.0, hi=800.0)
    theta = perturbed_theta(params, CONFIG3_BASE_LENGTHS, seed=seed, frac=0.2)
    w.physics["bnd"] = params.apply(w.physics["bnd"], theta)
    w.name = "config4"
    w.meta.update(params=params, theta=theta)
    return w


def scaled_config1(num_gs, per_gs, seed=42):
    """config1 geometry with fewer photons (parity-test sizes)."""
    return config1(num_gs=num_gs, per_gs=per_gs, seed=seed)
