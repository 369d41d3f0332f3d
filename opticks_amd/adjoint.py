"""
Yield gradients with respect to binned absorption and Rayleigh lengths.

There is no counterpart in the reference (it has no differentiation); the
forward path it differentiates is qsim::propagate_to_boundary
(qudarap/qsim.h:721-866), where a photon in material m travels
  absorption_distance = -L_abs(lambda) log u,  scattering_distance = -L_ray(lambda) log u'
and the lengths come from the bnd table by FP32 interpolation between two
1 nm samples (qudarap/qbnd.h:108-133).

Parameters theta (n_block, 2, n_bin) float32, in mm: block b is one material,
theta[b, 0] its absorption lengths and theta[b, 1] its Rayleigh lengths in
n_bin wavelength bins of width w = (hi - lo) / n_bin starting at lo; table
samples outside [lo, hi) take the edge bins. Table sample i of a block
material holds theta[b, p, sample_bin[i]], so the interpolated length is
  L(lambda) = (1 - a) theta[b, p, bin(i0)] + a theta[b, p, bin(i1)].

Estimator (score function / likelihood ratio): for detector d
  d E[hits_d] / d theta = E[ 1{photon detected in d} * d log p(path) / d theta ],
  d log p / d L = sum over segments in the material of ell / L^2
                  - 1/L on a segment ending in that process,
chained through the interpolation weights (1 - a, a). The HIP adjoint
(mi_sim_adjoint_device) replays each hit photon from its RNG stream, so
segment lengths are recomputed rather than stored, and sums the scores per
detector in photon order with Kahan compensation.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L

PROP_ABSORPTION, PROP_RAYLEIGH = 0, 1
_BND_COLUMN = {PROP_ABSORPTION: 1, PROP_RAYLEIGH: 2}   # material group 0: (n, L_abs, L_ray, reemission)


class LengthParams:
    """Binned absorption/Rayleigh lengths of the listed materials."""

    def __init__(self, physics: dict, materials=("LS", "Water"), n_bin=64, lo=200.0, hi=800.0):
        names = list(physics["material_names"])
        self.materials = tuple(materials)
        self.n_block = len(self.materials)
        self.n_bin = int(n_bin)
        self.lo, self.hi = float(lo), float(hi)
        self.width = (self.hi - self.lo) / self.n_bin
        self.n_mat = len(names)
        self.mat_block = np.full(self.n_mat + 1, -1, np.int32)
        for b, m in enumerate(self.materials):
            self.mat_block[names.index(m) + 1] = b
        n_wl = int(physics["n_wl"])
        wl = float(physics["wl_low"]) + float(physics["wl_step"]) * np.arange(n_wl, dtype=np.float64)
        self.wavelength = wl
        self.sample_bin = np.clip(np.floor((wl - self.lo) / self.width), 0, self.n_bin - 1).astype(np.int32)
        self.optical = np.asarray(physics["optical"], np.int32)

    @property
    def n_param(self):
        return self.n_block * 2 * self.n_bin

    @property
    def shape(self):
        return (self.n_block, 2, self.n_bin)

    def bin_centres(self):
        return self.lo + self.width * (np.arange(self.n_bin) + 0.5)

    def _lines(self):
        """(boundary, line j, block) for every material line of a parameterised material."""
        out = []
        for b in range(self.optical.shape[0]):
            for j in (0, 3):
                blk = self.mat_block[self.optical[b, j, 0]] if 0 <= self.optical[b, j, 0] <= self.n_mat else -1
                if blk >= 0:
                    out.append((b, j, int(blk)))
        return out

    def theta_from(self, bnd):
        """theta sampled from a table at the bin centres (nearest 1 nm sample)."""
        theta = np.zeros(self.shape, np.float32)
        idx = np.clip(np.rint((self.bin_centres() - self.wavelength[0]) / (self.wavelength[1] - self.wavelength[0])),
                      0, len(self.wavelength) - 1).astype(int)
        seen = set()
        for b, j, blk in self._lines():
            if blk in seen:
                continue
            seen.add(blk)
            for p, col in _BND_COLUMN.items():
                theta[blk, p] = bnd[b, j, 0, idx, col]
        return theta

    def apply(self, bnd, theta):
        """Copy of bnd with the parameterised lengths replaced by theta."""
        theta = np.asarray(theta, np.float32).reshape(self.shape)
        out = np.array(bnd, np.float32, copy=True)
        for b, j, blk in self._lines():
            for p, col in _BND_COLUMN.items():
                out[b, j, 0, :, col] = theta[blk, p, self.sample_bin]
        return out

    def c_struct(self, n_det: int, max_row_floats: int = 0) -> L.MiAdjoint:
        a = L.MiAdjoint(self.mat_block.ctypes.data, self.n_mat, self.sample_bin.ctypes.data, self.n_bin,
                        self.n_block, int(n_det), int(max_row_floats))
        a._keep = (self.mat_block, self.sample_bin)
        return a


def perturbed_theta(params: LengthParams, base: dict, seed=42, frac=0.2):
    """theta = base length of each (material, property) x (1 + U(-frac, frac)) per bin.

    base: {material: (L_abs, L_ray)} in mm.
    """
    rng = np.random.default_rng(seed)
    theta = np.zeros(params.shape, np.float32)
    for b, m in enumerate(params.materials):
        for p in (PROP_ABSORPTION, PROP_RAYLEIGH):
            theta[b, p] = base[m][p] * (1.0 + rng.uniform(-frac, frac, params.n_bin))
    return theta


def hits_per_detector(hit_t, n_hit, n_det):
    """Counts of hits per detector (sensor identifier 1..n_det) from a (M,4,4) hit tensor."""
    import torch
    if n_hit == 0:
        return torch.zeros(n_det, dtype=torch.float32, device=hit_t.device)
    ident = hit_t[:n_hit, 3, 1].contiguous().view(torch.int32) & 0xffff
    ident = ident.long()
    ok = (ident >= 1) & (ident <= n_det)
    return torch.bincount(ident[ok] - 1, minlength=n_det).to(torch.float32)


def _autograd_function():
    import torch

    class ExpectedHits(torch.autograd.Function):
        """theta (n_block,2,n_bin) -> hits per detector (n_det,), differentiable.

        forward propagates the event on the GPU with the lengths theta; backward
        runs the path-replay adjoint for those hits and contracts the
        (n_det, n_param) yield Jacobian with the incoming gradient.
        """

        @staticmethod
        def forward(ctx, theta, sim, params, gs_t, n_det, event_id, n_photon):
            bnd = params.apply(sim.base_bnd, theta.detach().cpu().numpy())
            sim.set_bnd(bnd)
            dev = gs_t.device
            photon_t = torch.empty((n_photon, 4, 4), dtype=torch.float32, device=dev)
            hit_t = torch.empty((n_photon, 4, 4), dtype=torch.float32, device=dev)
            n_hit = sim.simulate_device(gs_t, photon_t, None, hit_t, eventID=event_id, n_photon=n_photon)
            del photon_t
            counts = hits_per_detector(hit_t, n_hit, n_det)
            ctx.state = (sim, params, gs_t, hit_t, int(n_hit), n_det, event_id, n_photon, bnd)
            ctx.theta_shape = theta.shape
            ctx.theta_device, ctx.theta_dtype = theta.device, theta.dtype
            return counts

        @staticmethod
        def backward(ctx, grad_counts):
            sim, params, gs_t, hit_t, n_hit, n_det, event_id, n_photon, bnd = ctx.state
            sim.set_bnd(bnd)
            G = torch.empty((n_det, params.n_param), dtype=torch.float32, device=gs_t.device)
            sim.adjoint_device(gs_t, hit_t, n_hit, params.c_struct(n_det), G, eventID=event_id,
                               n_photon=n_photon)
            grad_theta = (grad_counts.to(torch.float32).reshape(1, n_det) @ G).reshape(ctx.theta_shape)
            grad_theta = grad_theta.to(device=ctx.theta_device, dtype=ctx.theta_dtype)
            return grad_theta, None, None, None, None, None, None

    return ExpectedHits


_EH = None


def expected_hits(theta, sim, params: LengthParams, gs_t, n_det: int, event_id: int = 0, n_photon: int = -1):
    """Differentiable hits per detector; theta is a torch tensor of params.shape."""
    global _EH
    if _EH is None:
        _EH = _autograd_function()
    if n_photon < 0:
        n_photon = int(gs_t[:, 0, 3].cpu().numpy().view(np.uint32).astype(np.int64).sum())
    return _EH.apply(theta, sim, params, gs_t, int(n_det), int(event_id), int(n_photon))


def yield_jacobian(sim, params: LengthParams, gs_t, hit_t, n_hit, n_det, event_id=0, n_photon=-1, stream=None,
                   max_row_floats=0):
    """(n_det, n_param) d E[hits_d]/d theta for an event already simulated with sim."""
    import torch
    G = torch.empty((n_det, params.n_param), dtype=torch.float32, device=gs_t.device)
    sim.adjoint_device(gs_t, hit_t, n_hit, params.c_struct(n_det, max_row_floats), G, eventID=event_id,
                       n_photon=n_photon, stream=stream)
    return G
