"""Simtrace of the checker (qsim::generate_photon_simtrace_frame + the simtrace raygen +
sevent::add_simtrace restated) against what fixes it independently:
  - the grid layout of SFrameGenstep::MakeCenterExtentGenstep (gridaxes, gsid, qat4);
  - directions drawn from the photon's own RNG stream (event 0), sin/cos within 4e-7;
  - the nearest hit equal to the checker's trace() of the same rays (trace is pinned by the
    golden vectors of the reference build), and for a sphere at the grid origin t = R exactly;
  - misses carrying the midgrey background colour, t = 1, u32 words 0xffffffff."""
import numpy as np

import pyoracle as po
from opticks_amd import configs


def test_grid_layout_matches_reference_rules():
    gs = configs.frame_gensteps((10.0, 20.0, 30.0, 100.0), (-2, 2, 0, 0, -1, 1, 7, 1), gridscale=0.5)
    u = gs.view(np.uint32)
    assert gs.shape == (5 * 3, 6, 4)
    assert np.all(u[:, 0, 0] == 17) and np.all(u[:, 0, 1] == configs.GRID_XZ) and np.all(u[:, 0, 3] == 7)
    ids = u[:, 0, 2].copy().view(np.int8).reshape(-1, 4)
    assert ids[0].tolist() == [-2, 0, -1, 0] and ids[-1].tolist() == [2, 0, 1, 0]
    # local_scale = gridscale / high * extent; translation in q5
    assert np.allclose(gs[:, 5, :3], ids[:, :3] * 50.0)
    assert np.allclose(gs[:, 2:5, :3].reshape(-1, 3, 3), np.eye(3))
    assert configs.grid_axes(0, 3, 3) == configs.GRID_YZ
    assert configs.grid_axes(3, 3, 0) == configs.GRID_XY
    assert configs.grid_axes(3, 3, 3) == configs.GRID_XYZ
    hi = configs.frame_gensteps((0, 0, 0, 100.0), (-1, 1, -1, 1, 0, 0, 1, 2), gridscale=0.5)
    assert len(hi) == 5 * 5 and np.allclose(np.abs(hi[:, 5, :2]).max(), 2 * 25.0)
    # geotran applied after the grid shift (Tran::product(geotran, shift))
    G = np.eye(4)
    G[:3, :3] = [[0, 1, 0], [-1, 0, 0], [0, 0, 1]]
    G[3, :3] = (5.0, 6.0, 7.0)
    gt = configs.frame_gensteps((0, 0, 0, 1.0), (1, 1, 0, 0, 0, 0, 1, 1), gridscale=1.0, geotran=G)
    p = np.array([0.0, 0.0, 0.0, 1.0]) @ gt[0, 2:6].astype(np.float64)
    assert np.allclose(p[:3], np.array([1.0, 0.0, 0.0, 1.0]) @ G[:, :3])


def _dirs_from_rng(seed, n, axes, skip):
    u = np.stack([po.uniforms(seed, i, 2, skip=skip, event_idx=0) for i in range(n)])
    phi = (np.float32(2.0) * np.float32(np.pi) * u[:, 0]).astype(np.float64)
    s, c = np.sin(phi), np.cos(phi)
    cost = (np.float32(2.0) * u[:, 1] - np.float32(1.0)).astype(np.float64)
    sint = np.sqrt(1.0 - cost * cost)
    z = np.zeros_like(s)
    return {configs.GRID_YZ: np.stack([z, c, s], 1), configs.GRID_XZ: np.stack([c, z, s], 1),
            configs.GRID_XY: np.stack([c, s, z], 1),
            configs.GRID_XYZ: np.stack([sint * c, sint * s, cost], 1)}[axes]


def test_directions_follow_the_photon_rng_stream():
    w = configs.config0(num_photon=10)
    W = po.World(w.geometry, w.physics)
    for cegs, axes in (((-2, 2, 0, 0, -2, 2, 20, 1), configs.GRID_XZ), ((-1, 1, -1, 1, -1, 1, 10, 1), configs.GRID_XYZ),
                       ((0, 0, -2, 2, -2, 2, 20, 1), configs.GRID_YZ), ((-2, 2, -2, 2, 0, 0, 20, 1), configs.GRID_XY)):
        gs = configs.frame_gensteps((0, 0, 0, 1000.0), cegs, gridscale=0.2)
        r = W.simtrace(gs, seed=w.seed)
        n = len(r)
        want = _dirs_from_rng(w.seed, n, axes, 100000)
        assert np.allclose(r[:, 3, :3], want, atol=4e-7, rtol=0)
        per = gs.view(np.uint32)[0, 0, 3]
        assert np.array_equal(r[:, 2, :3], np.repeat(gs[:, 5, :3], per, axis=0))


def test_sphere_at_grid_origin_closed_form_and_trace_agreement():
    w = configs.config0(num_photon=10)
    W = po.World(w.geometry, w.physics)
    gs = configs.frame_gensteps((0, 0, 0, 1.0), (0, 0, 0, 0, 0, 0, 2000, 1), gridscale=1.0)
    r = W.simtrace(gs, seed=3)
    u = r.view(np.uint32)
    assert np.all((u[:, 2, 3] & 0xFFFF) != 0xFFFF)
    assert np.abs(r[:, 0, 3] - 1000.0).max() < 1e-3                       # t = R from the centre
    assert np.abs(np.linalg.norm(r[:, 1, :3], axis=1) - 1000.0).max() < 1e-3
    assert np.all(r[:, 1, 3] == np.float32(0.05))                          # q1.w = tmin
    # every record equals the checker's trace() of its ray
    gs2 = configs.frame_gensteps((0, 0, 0, 1000.0), (-3, 3, 0, 0, -3, 3, 30, 1), gridscale=0.6)
    r2 = W.simtrace(gs2, seed=5)
    rays = np.zeros((len(r2), 2, 4), np.float32)
    rays[:, 0, :3] = r2[:, 2, :3]
    rays[:, 0, 3] = 0.05
    rays[:, 1, :3] = r2[:, 3, :3]
    rays[:, 1, 3] = 1e6
    t = W.trace(rays)
    tu = t.view(np.uint32)
    u2 = r2.view(np.uint32)
    assert np.array_equal(r2[:, 0, 3], t[:, 0, 3])
    assert np.array_equal(u2[:, 2, 3] & 0xFFFF, tu[:, 1, 1])
    assert np.array_equal(u2[:, 3, 3], tu[:, 1, 0])
    nn = r2[:, 0, :3] / np.linalg.norm(r2[:, 0, :3], axis=1, keepdims=True)
    assert np.allclose(nn, t[:, 0, :3], atol=1e-6)
    # the sphere (prim 1 of the solid) carries globalPrimIdx 1 in the high half
    inside = np.linalg.norm(r2[:, 2, :3], axis=1) < 1000.0
    assert np.all(u2[inside, 2, 3] >> 16 == 1)


def test_misses_carry_background_and_sentinels():
    w = configs.config0(num_photon=10)
    W = po.World(w.geometry, w.physics)
    G = np.eye(4)
    G[3, :3] = (0.0, 0.0, 9000.0)                       # grid far above the 4 m world box
    gs = configs.frame_gensteps((0, 0, 0, 1.0), (0, 0, -1, 1, -1, 1, 50, 1), gridscale=10.0, geotran=G)
    r = W.simtrace(gs, seed=1)
    u = r.view(np.uint32)
    away = r[:, 3, 2] > 0
    assert away.sum() > 50
    assert np.all(r[away, 0] == np.array([0.6, 0.6, 0.6, 1.0], np.float32))
    assert np.all(u[away, 2, 3] == 0xFFFFFFFF) and np.all(u[away, 3, 3] == 0xFFFFFFFF)
    assert np.allclose(r[away, 1, :3], r[away, 2, :3] + r[away, 3, :3])
    assert len(W.simtrace(gs[:0])) == 0
