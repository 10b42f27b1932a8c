"""Pins of the oracle's OpenGL warp producer (target depth by depth splatting,
scene.cpp:1506-1544 with depthSplatting.*; tau, deformation weight and dtau/dz
by tauWarp.frag / tauPartial.frag), none of which the reference's tests hold:
closed forms of fronto-parallel planes, the import path's warps where the two
producers must agree, and a brute-force rasteriser written independently of
the oracle's edge functions."""
import numpy as np

import oracle

RNG = np.random.default_rng(2026)


def _P(f, cx, cy, C):
    K = np.array([[f, 0, cx], [0, f, cy], [0, 0, 1.0]])
    return K @ np.hstack([np.eye(3), -np.asarray(C, np.float64).reshape(3, 1)])


def test_identity_camera_reproduces_the_view():
    h, w, z0 = 9, 13, 3.0
    P = _P(20.0, w / 2, h / 2, [0, 0, 0])
    depth = np.full((1, h, w), z0, np.float32)
    ud = oracle.gl_target_depth(depth, P[None], P, w, h)
    np.testing.assert_allclose(ud, z0, rtol=1e-12)
    tx, ty, wd, dx, dy = oracle.gl_warps(depth, P[None], P, ud)
    X, Y = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    np.testing.assert_allclose(tx[0], X, atol=1e-5)
    np.testing.assert_allclose(ty[0], Y, atol=1e-5)
    np.testing.assert_allclose(wd, 1.0, atol=1e-6)
    np.testing.assert_allclose(dx, 0.0, atol=1e-6)
    np.testing.assert_allclose(dy, 0.0, atol=1e-6)


def test_fronto_parallel_plane_shift_and_field_of_view():
    """x_u = x + f (C_v - C_u) / z on a plane at depth z; pixels landing outside
    (0, W) are -2; inside, the producer agrees with import_warps_from_depth."""
    h, w, f, z0 = 12, 16, 24.0, 4.0
    Pu = _P(f, w / 2, h / 2, [0, 0, 0])
    b = np.array([-0.5, 0.0, 0.7])
    Pv = np.stack([_P(f, w / 2, h / 2, [bx, by, 0]) for bx, by in zip(b, [0.3, -0.4, 0.0])])
    depth = np.full((3, h, w), z0, np.float32)
    ud = oracle.gl_target_depth(depth, Pv, Pu, w, h)
    assert np.all(ud == np.float32(z0)) or np.allclose(ud[ud > 0], z0)
    tx, ty, wd, _, _ = oracle.gl_warps(depth, Pv, Pu, ud)
    X, Y = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    for v, (bx, by) in enumerate(zip(b, [0.3, -0.4, 0.0])):
        ex, ey = X + f * bx / z0, Y + f * by / z0
        inside = (ex > 0) & (ex < w) & (ey > 0) & (ey < h)
        np.testing.assert_allclose(tx[v][inside], ex[inside], atol=1e-4)
        np.testing.assert_allclose(ty[v][inside], ey[inside], atol=1e-4)
        assert np.all(tx[v][~inside] == -2) and np.all(wd[v][~inside] == 0)
        np.testing.assert_allclose(wd[v][inside], 1.0, atol=1e-5)
    # the import producer on the same plane (no occlusion) gives the same warps
    itx, ity, _, _, _ = oracle.warps_from_depth(depth, Pv, Pu, 1)
    both = (tx >= 0) & (itx >= 0)
    assert both.mean() > 0.5
    np.testing.assert_allclose(tx[both], itx[both], atol=1e-4)
    np.testing.assert_allclose(ty[both], ity[both], atol=1e-4)


def test_invalid_depth_is_minus_two_and_splats_nothing():
    h, w = 8, 8
    P = _P(10.0, 4, 4, [0, 0, 0])
    depth = np.full((1, h, w), 2.0, np.float32)
    depth[0, 2:5, 3:6] = 1000.0
    ud = oracle.gl_target_depth(depth, P[None], P, w, h)
    assert np.all(ud[2:5, 3:6] == 0.0) and np.all(ud[depth[0] < 500] == np.float32(2.0))
    tx, ty, wd, _, _ = oracle.gl_warps(depth, P[None], P, ud)
    assert np.all(tx[0, 2:5, 3:6] == -2) and np.all(ty[0, 2:5, 3:6] == -2)


def test_occluded_background_is_minus_one():
    """A near square in front of a far wall, seen from a shifted view: wall
    pixels whose target position lies well inside the square's image are
    occluded; wall pixels well away from it are not."""
    h, w, f, zf, zn = 24, 32, 30.0, 8.0, 2.0
    Pu = _P(f, w / 2, h / 2, [0, 0, 0])
    bx = 0.4
    Pv = np.stack([Pu, _P(f, w / 2, h / 2, [bx, 0, 0])])
    depth = np.full((2, h, w), zf, np.float32)
    depth[0, 8:16, 10:18] = zn            # the square as the target's own camera sees it
    sh = f * bx / zn                      # its shift into view 1: 6 px
    depth[1, 8:16, int(10 - sh):int(18 - sh)] = zn
    ud = oracle.gl_target_depth(depth, Pv, Pu, w, h, skip=0)
    assert np.allclose(ud[9:15, 11:17], zn) and np.allclose(ud[:, 24:], zf)
    tx, ty, _, _, _ = oracle.gl_warps(depth, Pv, Pu, ud)
    X = np.arange(w) + 0.5
    xu = X + f * bx / zf                  # where view 1's wall pixels land
    wall = depth[1] == zf
    deep = (xu > 11.5) & (xu < 16.5)      # well inside the square's target image
    rows = slice(9, 15)
    assert np.all(tx[1][rows][:, deep][wall[rows][:, deep]] == -1)
    free = (xu > 19.5) & (xu < w - 0.5)
    assert np.all(tx[1][rows][:, free] > 0)


def _brute_target_depth(depth, shifts, f, z_of, W, H):
    """Independent rasteriser for fronto-parallel quads under translating
    cameras: source pixel (x, y) at depth z covers the target square
    [x + s, x + 1 + s] x [y + t, y + 1 + t], s = f bx / z; the nearest square
    containing a pixel centre wins.  Centres within 1e-6 of a square's edge
    are marked ambiguous."""
    V, h, w = depth.shape
    best = np.full((H, W), np.inf)
    amb = np.zeros((H, W), bool)
    cx = np.arange(W) + 0.5
    cy = np.arange(H) + 0.5
    for v in range(V):
        for y in range(h):
            for x in range(w):
                z = float(depth[v, y, x])
                s, t = f * shifts[v][0] / z, f * shifts[v][1] / z
                x0, x1, y0, y1 = x + s, x + 1 + s, y + t, y + 1 + t
                inx = (cx > x0 + 1e-6) & (cx < x1 - 1e-6)
                iny = (cy > y0 + 1e-6) & (cy < y1 - 1e-6)
                onx = (np.abs(cx - x0) <= 1e-6) | (np.abs(cx - x1) <= 1e-6)
                ony = (np.abs(cy - y0) <= 1e-6) | (np.abs(cy - y1) <= 1e-6)
                cover = iny[:, None] & inx[None, :]
                edge = (iny[:, None] | ony[:, None]) & (inx[None, :] | onx[None, :]) & ~cover
                amb |= edge
                zz = z_of(z)
                upd = cover & (zz < best)
                best[upd] = zz
    return np.where(np.isinf(best), 0.0, best), amb


def test_target_depth_matches_a_brute_force_rasteriser():
    h, w, f = 10, 14, 12.0
    Pu = _P(f, w / 2, h / 2, [0, 0, 0])
    shifts = [(0.37, 0.11), (-0.29, 0.23), (0.13, -0.41)]
    Pv = np.stack([_P(f, w / 2, h / 2, [sx, sy, 0]) for sx, sy in shifts])
    depth = RNG.choice([1.7, 3.1, 5.3], size=(3, h, w)).astype(np.float32)
    ud = oracle.gl_target_depth(depth, Pv, Pu, w, h)
    want, amb = _brute_target_depth(depth, shifts, f, lambda z: z, w, h)
    ok = ~amb
    assert ok.mean() > 0.8
    np.testing.assert_allclose(ud[ok], want[ok], rtol=1e-6)


def test_big_quads_are_rasterised():
    """A target camera that magnifies 80x: every source pixel's quad spans 80
    target pixels, beyond any per-thread walk, and is still rasterised (GL
    clips by the viewport, not by size).  Brute force: source pixel (x, y) at
    depth z covers [k (x - cxv) + cxu + s, k (x + 1 - cxv) + cxu + s], k = F / f,
    s = F b / z."""
    h, w, f, k = 4, 5, 4.0, 80.0
    F = k * f
    W, H = 200, 160
    Pu = _P(F, W / 2, H / 2, [0, 0, 0])
    shifts = [(0.0, 0.0), (0.004, -0.003)]
    Pv = np.stack([_P(f, w / 2, h / 2, [sx, sy, 0]) for sx, sy in shifts])
    depth = np.stack([np.full((h, w), 2.0), RNG.choice([1.5, 2.5], size=(h, w))]).astype(np.float32)
    ud = oracle.gl_target_depth(depth, Pv, Pu, W, H)
    best = np.full((H, W), np.inf)
    amb = np.zeros((H, W), bool)
    cx, cy = np.arange(W) + 0.5, np.arange(H) + 0.5
    for v, (bx, by) in enumerate(shifts):
        for y in range(h):
            for x in range(w):
                z = float(depth[v, y, x])
                x0 = k * (x - w / 2) + W / 2 + F * bx / z
                y0 = k * (y - h / 2) + H / 2 + F * by / z
                x1, y1 = x0 + k, y0 + k
                inx = (cx > x0 + 1e-6) & (cx < x1 - 1e-6)
                iny = (cy > y0 + 1e-6) & (cy < y1 - 1e-6)
                onx = (np.abs(cx - x0) <= 1e-6) | (np.abs(cx - x1) <= 1e-6)
                ony = (np.abs(cy - y0) <= 1e-6) | (np.abs(cy - y1) <= 1e-6)
                cover = iny[:, None] & inx[None, :]
                amb |= (iny[:, None] | ony[:, None]) & (inx[None, :] | onx[None, :]) & ~cover
                upd = cover & (z < best)
                best[upd] = z
    want = np.where(np.isinf(best), 0.0, best)
    ok = ~amb
    assert (want[ok] > 0).sum() > 0.5 * W * H      # most of the target is covered by big quads
    np.testing.assert_allclose(ud[ok], want[ok], rtol=1e-6)
