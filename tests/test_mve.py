"""The MVE front end of the OpenGL warp path on the CPU: meta.ini cameras
(view.cpp:214-270), .mvei images (view.cpp:279-326) and the depth maps as
Scene::importViews leaves them (view.cpp:607-634, texture.cpp:218-278,
addDepthScale.frag, view.cpp:735-741), read from files written by the tests
from the format itself; and the oracle of the radial-to-orthographic depth
filter (fromRadial2Ortho.frag) against closed forms."""
import numpy as np
import pytest

import oracle
from lfrv import io as lio
from mve_files import radial_from_ortho, view_dir, write_meta, write_mvei

RNG = np.random.default_rng(31)


def _rotation(a, b, c):
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


@pytest.mark.parametrize("W,H", [(64, 48), (40, 60)])
def test_mve_camera(tmp_path, W, H):
    """focal length in units of the larger side, fy = fx / pixel aspect, the
    principal point in units of each side, rotation listed row by row, and
    P = K [R | t]: a world point projects where K (R X + t) puts it."""
    R = _rotation(0.1, -0.2, 0.3)
    t = np.array([0.3, -0.1, 2.0])
    d = view_dir(tmp_path, 0)
    write_meta(f"{d}/meta.ini", 1.25, 0.8, (0.45, 0.55), R, t)
    K, R2, t2, P = lio.mve_camera(f"{d}/meta.ini", W, H)
    fx = 1.25 * max(W, H)
    assert np.allclose(K, [[fx, 0, 0.45 * W], [0, fx / 0.8, 0.55 * H], [0, 0, 1]], rtol=1e-15)
    assert np.allclose(R2, R, rtol=0, atol=1e-15) and np.allclose(t2, t, rtol=0, atol=1e-15)
    X = np.array([0.2, -0.3, 1.5])
    c = R @ X + t
    want = np.array([fx * c[0] / c[2] + 0.45 * W, fx / 0.8 * c[1] / c[2] + 0.55 * H])
    h = P @ np.append(X, 1.0)
    assert np.allclose(h[:2] / h[2], want, rtol=1e-12)


def test_mve_camera_missing(tmp_path):
    """A meta.ini whose [view] section comes first has no camera; so has a
    missing file."""
    d = view_dir(tmp_path, 0)
    write_meta(f"{d}/meta.ini", 1.0, 1.0, (0.5, 0.5), np.eye(3), np.zeros(3), with_camera=False)
    assert lio.mve_camera(f"{d}/meta.ini", 32, 32) is None
    assert lio.mve_camera(f"{tmp_path}/nowhere/meta.ini", 32, 32) is None


def test_mve_depth_last_scale_nearest_and_invalid(tmp_path):
    """Two scales: only the last one read survives (the accumulated map is
    cleared before each), upsampled by pixel repetition, its derivatives
    divided by 2^scale; pixels whose depth and derivatives are all 0 become
    (1000, 0, 0)."""
    W, H = 16, 12
    d = view_dir(tmp_path, 3)
    z0 = RNG.uniform(2, 5, (H, W)).astype(np.float32)
    write_mvei(f"{d}/depth-L0.mvei", z0[None])
    write_mvei(f"{d}/dz-L0.mvei", RNG.normal(size=(2, H, W)))
    z1 = RNG.uniform(2, 5, (H // 2, W // 2)).astype(np.float32)
    g1 = RNG.normal(size=(2, H // 2, W // 2)).astype(np.float32)
    z1[1, 2] = 0.0
    g1[:, 1, 2] = 0.0      # all three zero: invalid
    z1[3, 3] = 0.0         # depth zero but a derivative not: kept as read
    write_mvei(f"{d}/depth-L1.mvei", z1[None])
    write_mvei(f"{d}/dz-L1.mvei", g1)
    z, gx, gy = lio.mve_depth(str(tmp_path), 3, 0, 1, W, H)
    up = lambda a: np.repeat(np.repeat(a, 2, 0), 2, 1)  # noqa: E731
    want_z, want_gx, want_gy = up(z1), up(g1[0]) / 2, up(g1[1]) / 2
    bad = up(np.arange(z1.size).reshape(z1.shape) == 1 * z1.shape[1] + 2)
    want_z[bad], want_gx[bad], want_gy[bad] = 1000.0, 0.0, 0.0
    assert np.array_equal(z, want_z) and np.array_equal(gx, want_gx) and np.array_equal(gy, want_gy)
    assert z[6, 6] == 0.0   # the kept zero depth
    # one scale: the file itself
    z, gx, _ = lio.mve_depth(str(tmp_path), 3, 0, 0, W, H)
    assert np.array_equal(z, z0)
    # a missing scale or a size that is not a multiple fails
    assert lio.mve_depth(str(tmp_path), 3, 0, 2, W, H) is None
    assert lio.mve_depth(str(tmp_path), 3, 1, 1, W + 1, H) is None


def test_radial_to_ortho_oracle_plane_and_pixel():
    """A fronto-parallel plane at depth Z seen radially (Z |K^-1 x|) comes back
    as Z with zero derivatives when the radial derivatives are the exact ones;
    one pixel by hand: at the principal point the ray is the axis (d = 1)."""
    K = np.array([[50.0, 0.0, 16.0], [0.0, 40.0, 12.0], [0.0, 0.0, 1.0]])
    H, W = 24, 32
    Z = 3.5
    zr = radial_from_ortho(np.full((H, W), Z), K)
    # d/dx of Z |u| = Z (u . Ki e_x) / |u|, exactly, with u = Ki (x + 0.5, y + 0.5, 1)
    Ki = np.linalg.inv(K)
    X, Y = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    u = np.einsum("ij,jhw->ihw", Ki, np.stack([X, Y, np.ones_like(X)]))
    n = np.sqrt((u * u).sum(0))
    gxr = Z * np.einsum("ihw,i->hw", u, Ki[:, 0]) / n
    gyr = Z * np.einsum("ihw,i->hw", u, Ki[:, 1]) / n
    z, gx, gy = oracle.radial_to_ortho(zr, K, gxr, gyr)
    assert np.allclose(z, Z, rtol=1e-6)
    assert np.max(np.abs(gx)) < 1e-5 and np.max(np.abs(gy)) < 1e-5
    # the pixel whose centre (x + 0.5, y + 0.5) is the principal point: the ray is the axis, d = 1
    K2 = np.array([[50.0, 0.0, 15.5], [0.0, 40.0, 11.5], [0.0, 0.0, 1.0]])
    zz = np.zeros((H, W), np.float32)
    zz[11, 15] = 7.0
    gg = np.zeros((H, W), np.float32)
    gg[11, 15] = 0.25
    z, gx, gy = oracle.radial_to_ortho(zz, K2, gg, gg)
    assert z[11, 15] == np.float32(7.0) and gx[11, 15] == np.float32(0.25) and gy[11, 15] == np.float32(0.25)
    # one step right: u = (1/50, 0, 1), d^2 = 1 + 1/2500; dz/dx -= z (1/50)(1/50) / d^3, dz/dy unchanged
    zz[11, 16] = 7.0
    z, gx, gy = oracle.radial_to_ortho(zz, K2, np.zeros_like(zz), np.zeros_like(zz))
    d = np.sqrt(1 + 1 / 2500)
    assert np.isclose(z[11, 16], 7.0 / d, rtol=1e-7)
    assert np.isclose(gx[11, 16], -7.0 * (1 / 50) * (1 / 50) / d**3, rtol=1e-6)
    assert gy[11, 16] == 0.0
