"""
Boundary property tables in the reference's standard layout.

  bnd     (nbnd, 4, 2, nwl, 4) float32   lines (omat, osur, isur, imat) x 2 float4 groups
          material group0 = (RINDEX, ABSLENGTH, RAYLEIGH, REEMISSIONPROB)
          material group1 = (GROUPVEL, 0, 0, 0)
          surface  group0 = (detect, absorb, reflect_specular, reflect_diffuse)
                                                   (qudarap/QBnd.cc:130-175, qudarap/qbnd.h:108-223)
  optical (nbnd, 4, 4) int32  (idx+1, ems, 0, 0) per line; materials have ems 0,
          a missing surface gives (0, NoSurface=1), a present one (idx+1, Surface=2)
                                                   (sysrap/sstandard.h make_optical)
  icdf    (3, 4096) float32 inverse CDF of the scintillation emission spectrum with
          the hd_factor=20 layers: full range, u in [0,0.05), u in (0.95,1]
                                                   (qudarap/QScint.cc:25-52, qscint.h:160-230)

The wavelength domain is 60..820 nm at 1 nm (sysrap/sdomain.h:26-31).
Group velocity is c/n_g with n_g = n - lambda dn/dlambda (constant n gives c/n).
Boundaries are named "omat/osur/isur/imat" as in CSGOptiX/tests/CXRaindropTest.cc:24-29.
"""
from __future__ import annotations

import numpy as np

C_LIGHT_MM_PER_NS = 299.792458
WL_LOW, WL_STEP, N_WL = 60.0, 1.0, 761

# smatsur.h
EMS_MATERIAL, EMS_NOSURFACE, EMS_SURFACE, EMS_ZPLUS_SENSOR_A = 0, 1, 2, 3


def wavelength_domain():
    return WL_LOW + WL_STEP * np.arange(N_WL, dtype=np.float64)


def _as_table(v, wl):
    if callable(v):
        return np.asarray(v(wl), np.float64)
    return np.full_like(wl, float(v))


class Material:
    def __init__(self, name, rindex, abslength, rayleigh, reemission=0.0):
        self.name = name
        self.rindex, self.abslength, self.rayleigh, self.reemission = rindex, abslength, rayleigh, reemission

    def table(self):
        wl = wavelength_domain()
        n = _as_table(self.rindex, wl)
        dn = np.gradient(n, wl)
        ng = n - wl * dn
        t = np.zeros((2, N_WL, 4), np.float32)
        t[0, :, 0] = n
        t[0, :, 1] = _as_table(self.abslength, wl)
        t[0, :, 2] = _as_table(self.rayleigh, wl)
        t[0, :, 3] = _as_table(self.reemission, wl)
        t[1, :, 0] = C_LIGHT_MM_PER_NS / ng
        return t


class Surface:
    def __init__(self, name, detect=0.0, absorb=0.0, specular=0.0, diffuse=0.0, ems=EMS_SURFACE):
        self.name = name
        self.detect, self.absorb, self.specular, self.diffuse = detect, absorb, specular, diffuse
        self.ems = ems

    def table(self):
        wl = wavelength_domain()
        t = np.zeros((2, N_WL, 4), np.float32)
        t[0, :, 0] = _as_table(self.detect, wl)
        t[0, :, 1] = _as_table(self.absorb, wl)
        t[0, :, 2] = _as_table(self.specular, wl)
        t[0, :, 3] = _as_table(self.diffuse, wl)
        return t


class BoundaryLib:
    """Collects materials, surfaces and "omat/osur/isur/imat" boundaries."""

    def __init__(self):
        self.materials = {}
        self.surfaces = {}
        self.mat_order = []
        self.sur_order = []
        self.boundaries = []

    def add_material(self, m: Material):
        self.materials[m.name] = m
        self.mat_order.append(m.name)

    def add_surface(self, s: Surface):
        self.surfaces[s.name] = s
        self.sur_order.append(s.name)

    def boundary(self, spec: str) -> int:
        if spec in self.boundaries:
            return self.boundaries.index(spec)
        omat, osur, isur, imat = spec.split("/")
        assert omat in self.materials and imat in self.materials, spec
        assert (not osur or osur in self.surfaces) and (not isur or isur in self.surfaces), spec
        self.boundaries.append(spec)
        return len(self.boundaries) - 1

    def arrays(self):
        nb = len(self.boundaries)
        bnd = np.zeros((nb, 4, 2, N_WL, 4), np.float32)
        optical = np.zeros((nb, 4, 4), np.int32)
        for b, spec in enumerate(self.boundaries):
            omat, osur, isur, imat = spec.split("/")
            for j, (name, is_mat) in enumerate(((omat, True), (osur, False), (isur, False), (imat, True))):
                if is_mat:
                    bnd[b, j] = self.materials[name].table()
                    optical[b, j] = (self.mat_order.index(name) + 1, EMS_MATERIAL, 0, 0)
                elif name:
                    s = self.surfaces[name]
                    bnd[b, j] = s.table()
                    optical[b, j] = (self.sur_order.index(name) + 1, s.ems, 0, 0)
                else:
                    bnd[b, j] = -1.0
                    optical[b, j] = (0, EMS_NOSURFACE, 0, 0)
        return dict(bnd=bnd, optical=optical, wl_low=WL_LOW, wl_step=WL_STEP, n_wl=N_WL,
                    material_names=list(self.mat_order))


def gaussian_icdf(mean=430.0, sigma=20.0, lo=200.0, hi=800.0, n=4096, hd_factor=20):
    """Inverse CDF of a Gaussian emission spectrum truncated to [lo, hi].

    Texel i of each layer holds F^-1 at its centre (i+0.5)/n, so that the
    linearly filtered lookup at u returns F^-1(u) between texel centres.
    """
    from scipy.special import ndtr, ndtri
    a, b = ndtr((lo - mean) / sigma), ndtr((hi - mean) / sigma)

    def inv(p):
        return mean + sigma * ndtri(a + p * (b - a))

    x = (np.arange(n) + 0.5) / n
    if hd_factor == 0:
        return np.asarray([inv(x)], np.float32), 0
    edge = 1.0 / hd_factor
    layers = [inv(x), inv(x * edge), inv(1.0 - edge + x * edge)]
    return np.asarray(layers, np.float32), hd_factor
