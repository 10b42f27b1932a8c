"""Pins of the FP64 oracle (CPU only).

The reference ships no golden vectors for this path, so every oracle function
is checked against something it does not share code with: a brute-force
construction from the definition (box footprint integrated over pixel cells,
dense warp matrices, dense forward-difference matrices, dense linear solves),
closed forms (ramps, constants, fronto-parallel disparity, identity cameras,
affine zooms, Gaussian taps) and invariants (adjoint identities, projection
idempotence, fixed points).  A plausible slip in the oracle (a dropped edge
term, a swapped tap index, a sign, a transposed camera) fails at least one of
these.
"""
import sys

import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------------------
# Independent constructions
# ---------------------------------------------------------------------------

def box_weights_1d(t, n, dsf):
    """Footprint of one warped coordinate along one axis, from the definition:
    a box of width dsf centred at t - 0.5 (pixel-centre coordinates)
    integrated over the cells [x - .5, x + .5] of the n in-image pixels,
    normalised to sum 1 (gradientIBR_kernels.cuh:447-507 describes the same
    box by its first tap and fractional edge overlaps)."""
    c = float(np.float32(t)) - 0.5
    lo, hi = c - dsf / 2.0, c + dsf / 2.0
    x = np.arange(n, dtype=np.float64)
    ov = np.clip(np.minimum(hi, x + 0.5) - np.maximum(lo, x - 0.5), 0.0, None)
    s = ov.sum()
    return ov / s if s > 0 else ov


def dense_warp(tx, ty, W, H, dsf, rule):
    """Dense S (n x N) with S[m, y*W + x] = wy * wx for the rows admitted by
    `rule` ('gather': tau inside [0, W] x [0, H]; 'cell': additionally
    floor(tau - .5) inside the image)."""
    tx, ty = np.asarray(tx, np.float32).ravel(), np.asarray(ty, np.float32).ravel()
    S = np.zeros((tx.size, W * H))
    for m in range(tx.size):
        a, b = tx[m], ty[m]
        if a < 0 or b < 0 or a > W or b > H:
            continue
        if rule == "cell":
            if not (0 <= np.floor(np.float32(a - np.float32(0.5))) <= W - 1 and
                    0 <= np.floor(np.float32(b - np.float32(0.5))) <= H - 1):
                continue
        S[m] = np.outer(box_weights_1d(b, H, dsf), box_weights_1d(a, W, dsf)).ravel()
    return S


def dense_D(W, H):
    """Forward differences with Neumann boundary: rows for x < W-1 and y < H-1
    (the last column / row difference is absent), stacked [Dx; Dy]."""
    N = W * H
    Dx = np.zeros((N, N))
    Dy = np.zeros((N, N))
    for y in range(H):
        for x in range(W):
            o = y * W + x
            if x < W - 1:
                Dx[o, o], Dx[o, o + 1] = -1.0, 1.0
            if y < H - 1:
                Dy[o, o], Dy[o, o + W] = -1.0, 1.0
    return Dx, Dy


def random_taus(V, h, w, W, H, frac_invalid=0.15, margin=1.5):
    tx = RNG.uniform(-margin, W + margin, size=(V, h, w)).astype(np.float32)
    ty = RNG.uniform(-margin, H + margin, size=(V, h, w)).astype(np.float32)
    bad = RNG.random((V, h, w)) < frac_invalid
    tx[bad] = -float(w)
    ty[bad] = -float(h)
    return tx, ty


# ---------------------------------------------------------------------------
# Footprint (set_B_bilinear)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dsf", [1, 2, 3, 4])
def test_footprint_matches_box_integral(dsf):
    W = H = 12
    for _ in range(400):
        tx, ty = RNG.uniform(0, W), RNG.uniform(0, H)
        r, x0, y0, wx, wy = oracle.footprint(tx, ty, W, H, dsf)
        bx, by = box_weights_1d(tx, W, dsf), box_weights_1d(ty, H, dsf)
        fx, fy = np.zeros(W), np.zeros(H)
        for j in range(dsf + 1):
            if 0 <= x0 + j < W:
                fx[x0 + j] += wx[j]
        for i in range(dsf + 1):
            if 0 <= y0 + i < H:
                fy[y0 + i] += wy[i]
        assert r == dsf + 1
        np.testing.assert_allclose(fx, bx, atol=1e-12)
        np.testing.assert_allclose(fy, by, atol=1e-12)
        assert abs(fx.sum() - 1) < 1e-12 and abs(fy.sum() - 1) < 1e-12


def test_footprint_dsf1_is_bilinear():
    W = H = 16
    for _ in range(200):
        tx, ty = RNG.uniform(1, W - 1), RNG.uniform(1, H - 1)
        _, x0, y0, wx, wy = oracle.footprint(tx, ty, W, H, 1)
        px, py = float(np.float32(tx)) - 0.5, float(np.float32(ty)) - 0.5
        assert x0 == int(np.floor(px)) and y0 == int(np.floor(py))
        fx, fy = px - x0, py - y0
        np.testing.assert_allclose(wx, [1 - fx, fx], atol=1e-12)
        np.testing.assert_allclose(wy, [1 - fy, fy], atol=1e-12)


def test_footprint_pixel_centre_is_single_tap():
    for x in range(1, 7):
        _, x0, _, wx, _ = oracle.footprint(x + 0.5, 3.5, 8, 8, 1)
        assert x0 == x and wx[0] == 1.0 and wx[1] == 0.0
        # dsf = 2 at a pixel centre: half, full, half
        _, x0, _, wx, _ = oracle.footprint(x + 0.5, 3.5, 8, 8, 2)
        np.testing.assert_allclose(wx, [0.25, 0.5, 0.25], atol=1e-15)
        assert x0 == x - 1


def test_footprint_outside_taps_are_dropped_and_renormalised():
    # left edge: the tap at x = -1 carries no weight, the rest sum to one
    r, x0, _, wx, _ = oracle.footprint(0.2, 4.0, 8, 8, 2)
    assert x0 == -1 and wx[0] == 0.0 and abs(wx.sum() - 1) < 1e-15


# ---------------------------------------------------------------------------
# Gather / splat
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dsf", [1, 2, 4])
def test_gather_equals_dense_warp(dsf):
    w = h = 6
    W, H = w * dsf, h * dsf
    tx, ty = random_taus(1, h, w, W, H)
    u = RNG.random((3, H, W))
    S = dense_warp(tx, ty, W, H, dsf, "gather")
    got = oracle.gather(W, H, dsf, tx[0], ty[0], u)
    want = (S @ u.reshape(3, -1).T).T.reshape(3, h, w)
    np.testing.assert_allclose(got, want, atol=1e-12)


@pytest.mark.parametrize("dsf", [1, 2, 3])
def test_splat_equals_dense_transpose(dsf):
    w = h = 6
    W, H = w * dsf, h * dsf
    tx, ty = random_taus(1, h, w, W, H)
    r = RNG.random((3, h, w))
    wg = RNG.random((h, w))
    S = dense_warp(tx, ty, W, H, dsf, "cell")
    out, wsum = oracle.splat(W, H, dsf, tx[0], ty[0], r, wgt=wg)
    want = (S.T @ (r.reshape(3, -1) * wg.ravel()).T).T.reshape(3, H, W)
    np.testing.assert_allclose(out, want, atol=1e-12)
    np.testing.assert_allclose(wsum.ravel(), S.T @ wg.ravel(), atol=1e-12)


def test_gather_splat_adjoint_on_cell_valid_sources():
    w = h = 10
    W = H = 20
    tx = RNG.uniform(0.5, W - 0.5, size=(h, w)).astype(np.float32)
    ty = RNG.uniform(0.5, H - 0.5, size=(h, w)).astype(np.float32)
    u = RNG.standard_normal((1, H, W))
    r = RNG.standard_normal((1, h, w))
    Su = oracle.gather(W, H, 2, tx, ty, u)
    STr, _ = oracle.splat(W, H, 2, tx, ty, r)
    assert abs(np.vdot(Su, r) - np.vdot(u, STr)) < 1e-10 * abs(np.vdot(Su, r))


def test_integer_shift_gather_is_exact_translation():
    W = H = 9
    u = RNG.random((3, H, W))
    ys, xs = np.mgrid[0:H, 0:W]
    tx = (xs + 2 + 0.5).astype(np.float32)
    ty = (ys + 1 + 0.5).astype(np.float32)
    v = oracle.gather(W, H, 1, tx, ty, u)
    np.testing.assert_array_equal(v[:, :H - 1, :W - 2], u[:, 1:, 2:])


def test_cell_validity_edges():
    W = H = 8
    assert oracle.cell_valid(0.5, 0.5, W, H)
    assert not oracle.cell_valid(0.49, 4.0, W, H)   # gather-valid, floor(tau-.5) = -1
    assert oracle.cell_valid(W - 0.01, 4.0, W, H)
    assert not oracle.cell_valid(W + 0.01, 4.0, W, H)
    assert not oracle.cell_valid(-8.0, -8.0, W, H)


def test_coverage_sums_to_number_of_cell_valid_sources():
    w = h = 16
    W = H = 32
    tx, ty = random_taus(3, h, w, W, H)
    cov = oracle.coverage(W, H, 2, tx, ty)
    nvalid = sum(oracle.cell_valid(a, b, W, H) for a, b in zip(tx.ravel(), ty.ravel()))
    assert abs(cov.sum() - nvalid) < 1e-9


def test_target_visibility_single_tap():
    W = H = 8
    tx = np.array([[3.5]], np.float32)
    ty = np.array([[5.5]], np.float32)
    vis = oracle.target_visibility(W, H, 1, [(tx, ty)])
    want = np.zeros((H, W), np.uint8)
    want[5, 3] = 1
    np.testing.assert_array_equal(vis, want)


# ---------------------------------------------------------------------------
# Stencils
# ---------------------------------------------------------------------------

def test_sobel_ramp_constant_and_luminance():
    H, W = 7, 9
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    a, b = 0.3, -0.7
    ramp = a * xs + b * ys
    gx, gy = oracle.sobel(np.stack([ramp, ramp, ramp]))
    inner = (slice(1, H - 1), slice(1, W - 1))
    np.testing.assert_allclose(gx[inner], a, atol=1e-14)
    np.testing.assert_allclose(gy[inner], b, atol=1e-14)
    assert np.all(gx[0] == 0) and np.all(gy[:, 0] == 0)
    gx, gy = oracle.sobel(np.full((3, H, W), 0.42))
    assert np.abs(gx).max() < 1e-16 and np.abs(gy).max() < 1e-16
    # ramp in one channel only: the Rec.709 weight of that channel
    for c, wgt in enumerate((0.2126, 0.7152, 0.0722)):
        img = np.zeros((3, H, W))
        img[c] = xs
        gx, _ = oracle.sobel(img)
        np.testing.assert_allclose(gx[inner], wgt, atol=1e-14)


def test_laplacian_closed_forms():
    H, W = 6, 7
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    lap = oracle.laplacian(2.0 * xs - ys)
    np.testing.assert_allclose(lap[1:-1, 1:-1], 0.0, atol=1e-14)
    lap = oracle.laplacian(xs ** 2)
    np.testing.assert_allclose(lap[1:-1, 1:-1], 0.5, atol=1e-13)
    # corner: mean of the two neighbours
    u = RNG.random((H, W))
    assert abs(oracle.laplacian(u)[0, 0] - (0.5 * (u[0, 1] + u[1, 0]) - u[0, 0])) < 1e-15


def test_dtd_and_dtg_equal_dense_forward_differences():
    H, W = 5, 6
    Dx, Dy = dense_D(W, H)
    u = RNG.standard_normal((H, W))
    gx, gy = RNG.standard_normal((H, W)), RNG.standard_normal((H, W))
    np.testing.assert_allclose(oracle.dtd(u).ravel(), (Dx.T @ Dx + Dy.T @ Dy) @ u.ravel(), atol=1e-12)
    # only the rows that exist in D carry g (last column of gx, last row of gy are unused)
    gxm, gym = gx.copy(), gy.copy()
    gxm[:, -1] = 0
    gym[-1, :] = 0
    np.testing.assert_allclose(oracle.dtg(gx, gy).ravel(), Dx.T @ gxm.ravel() + Dy.T @ gym.ravel(), atol=1e-12)


def test_pd_dual_primal_adjoint_and_projection():
    H, W = 6, 5
    u = RNG.standard_normal((3, H, W))
    zero = np.zeros((3, H, W))
    # grad+ from the dual step with a huge radius, div- from the primal step
    px, py = oracle.pd_dual(u, zero, zero, 1.0, g=np.full((H, W), 1e30))
    p = RNG.standard_normal((2, H, W))
    p[0][:, -1] = 0.0  # the dual step never writes the last column / row (grad+ is 0 there)
    p[1][-1, :] = 0.0
    div = (oracle.pd_primal(np.zeros((H, W)), np.zeros((H, W)), p[0], p[1], 1.0, 1e300)
           - np.zeros((H, W)))
    # <grad+ u, p> = -<u, div- p> (div from u + tau*div with lambda -> inf)
    lhs = np.vdot(px[0], p[0]) + np.vdot(py[0], p[1])
    assert abs(lhs + np.vdot(u[0], div)) < 1e-9 * max(1.0, abs(lhs))
    # projection onto the Frobenius ball of radius g, idempotent
    g = RNG.uniform(0.1, 1.0, (H, W))
    qx, qy = oracle.pd_dual(u, 3 * RNG.standard_normal((3, H, W)), RNG.standard_normal((3, H, W)), 2.0, g)
    nrm = np.sqrt((qx ** 2 + qy ** 2).sum(0))
    assert np.all(nrm <= g * (1 + 1e-12))
    rx, ry = oracle.pd_dual(np.zeros((3, H, W)), qx, qy, 0.0, g)
    np.testing.assert_allclose(rx, qx, atol=1e-15)
    np.testing.assert_allclose(ry, qy, atol=1e-15)


def test_pd_primal_limits_and_constant_fixed_point():
    H, W = 5, 5
    u, f = RNG.random((H, W)), RNG.random((H, W))
    zp = np.zeros((H, W))
    np.testing.assert_allclose(oracle.pd_primal(u, f, zp, zp, 0.3, 0.0), f, atol=1e-15)
    np.testing.assert_allclose(oracle.pd_primal(u, f, zp, zp, 0.3, 1e300), u, atol=1e-12)
    c = np.full((3, H, W), 0.37)
    px, py = oracle.pd_dual(c, np.zeros((3, H, W)), np.zeros((3, H, W)), 0.5)
    assert np.abs(px).max() == 0 and np.abs(py).max() == 0
    np.testing.assert_allclose(oracle.pd_primal(c[0], c[0], px[0], py[0], 0.2, 0.7), c[0], atol=1e-15)


def test_gauss_taps_and_clamped_constant():
    H, W = 25, 27
    img = np.zeros((H, W))
    img[12, 13] = 1.0
    out = oracle.gauss_smooth(img)
    k = np.exp(-0.5 * np.arange(-5, 6) ** 2 / 4.0)
    k /= k.sum()
    np.testing.assert_allclose(out[7:18, 8:19], np.outer(k, k), atol=1e-15)
    assert abs(out.sum() - 1) < 1e-14
    np.testing.assert_allclose(oracle.gauss_smooth(np.full((H, W), 2.5)), 2.5, atol=1e-14)


# ---------------------------------------------------------------------------
# Weights
# ---------------------------------------------------------------------------

def test_angular_weights_closed_forms():
    sig = 1.0 / 255
    z = np.zeros((1, 3))
    dp = np.array([[0.1, 0.1, 4.0]], np.float32)
    gx = np.array([[0.0, 5.0, 1.0]])
    om = oracle.angular_weights(gx, z, sig, dp, np.zeros((1, 3), np.float32), 1.0)
    # zero gradient -> 1/sigma^2; |g| = 5 clipped to thr = 1 -> sigma_geom = 0.1;
    # sigma_geom = 4 capped at 0.5
    np.testing.assert_allclose(om[0], [1 / sig ** 2, 1 / (0.01 + sig ** 2), 1 / (0.25 + sig ** 2)], rtol=1e-6)


@pytest.mark.parametrize("zoom,want", [(2.0, 0.25), (0.5, 1.0), (1.0, 1.0)])
def test_deform_weights_affine_zoom(zoom, want):
    w = h = 8
    W = H = 64
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    tx = (zoom * xs + 10).astype(np.float32)
    ty = (zoom * ys + 10).astype(np.float32)
    om = oracle.deform_weights(W, H, tx, ty)
    np.testing.assert_allclose(om[1:-1, 1:-1], want, rtol=1e-6)
    assert np.all(om[0] == 1.0) and np.all(om[:, -1] == 1.0)
    tx[3, 3] = -8
    om = oracle.deform_weights(W, H, tx, ty)
    assert om[3, 3] == 0.0 and om[3, 4] == 1.0 and om[2, 3] == 1.0


# ---------------------------------------------------------------------------
# Cameras and warps from depth
# ---------------------------------------------------------------------------

def random_rotation():
    q, _ = np.linalg.qr(RNG.standard_normal((3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_decompose_P_round_trip():
    for _ in range(50):
        K = np.array([[RNG.uniform(300, 2000), RNG.uniform(-5, 5), RNG.uniform(100, 900)],
                      [0, RNG.uniform(300, 2000), RNG.uniform(100, 900)], [0, 0, 1]])
        R = random_rotation()
        t = RNG.standard_normal(3)
        lam = RNG.uniform(0.1, 10) * RNG.choice([-1, 1])
        P = lam * K @ np.hstack([R, t[:, None]])
        Ko, Ro, to, Co = oracle.decompose_P(P)
        np.testing.assert_allclose(Ko, K, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(Ro, R, atol=1e-9)
        np.testing.assert_allclose(to, t, atol=1e-9)
        np.testing.assert_allclose(Co, -R.T @ t, atol=1e-9)


def cam(f, cx, cy, C):
    K = np.array([[f, 0, cx], [0, f, cy], [0, 0, 1.0]])
    return K @ np.hstack([np.eye(3), -np.asarray(C, float)[:, None]])


def test_warp_identity_camera():
    h, w = 12, 16
    P = cam(20.0, 8, 6, [0, 0, 0])
    depth = RNG.uniform(2, 5, (1, h, w)).astype(np.float32)
    tx, ty, dx, dy, _ = oracle.warps_from_depth(depth, P[None], P, 1)
    ys, xs = np.mgrid[0:h, 0:w]
    np.testing.assert_allclose(tx[0], xs + 0.5, atol=1e-5)
    np.testing.assert_allclose(ty[0], ys + 0.5, atol=1e-5)
    assert np.abs(dx).max() < 1e-5 and np.abs(dy).max() < 1e-5


def test_warp_fronto_parallel_disparity_and_partial():
    h, w = 20, 40
    f, b, Z, sz = 100.0, 0.05, 2.5, 10.0
    Pv = cam(f, 20, 10, [0, 0, 0])
    Pu = cam(f, 20, 10, [b, 0, 0])
    depth = np.full((1, h, w), Z, np.float32)
    tx, ty, dx, dy, _ = oracle.warps_from_depth(depth, Pv[None], Pu, 1, sigma_z=sz)
    ys, xs = np.mgrid[0:h, 0:w]
    ok = tx[0] >= 0
    assert ok.sum() > 0.8 * ok.size
    np.testing.assert_allclose(tx[0][ok], (xs + 0.5 - f * b / Z)[ok], atol=1e-4)
    np.testing.assert_allclose(ty[0][ok], (ys + 0.5)[ok], atol=1e-4)
    # d tau_x / dZ = f b / Z^2, scaled by sigma_z
    np.testing.assert_allclose(dx[0][ok], sz * f * b / Z ** 2, rtol=1e-4)
    np.testing.assert_allclose(dy[0][ok], 0.0, atol=1e-5)


def test_warp_invalid_depth_and_failed_round_trip():
    h, w = 6, 8
    P = cam(10.0, 4, 3, [0, 0, 0])
    depth = np.full((1, h, w), 3.0, np.float32)
    depth[0, 2, 3] = 0.0
    depth[0, 4, 5] = 5000.0
    tx, ty, dx, _, _ = oracle.warps_from_depth(depth, P[None], P, 1)
    assert tx[0, 2, 3] == -w and ty[0, 2, 3] == -h and dx[0, 2, 3] == 0
    assert tx[0, 4, 5] == -w
    # a negative round-trip tolerance rejects every pixel
    tx, ty, _, _, _ = oracle.warps_from_depth(depth, P[None], P, 1, occl_tol=-1.0)
    assert np.all((tx == -w) | (tx == -2 * w))


# ---------------------------------------------------------------------------
# Solver pieces against dense linear algebra
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dsf", [1, 2])
def test_normal_equations_and_pcg_against_dense_solve(dsf):
    V, h, w = 2, 5, 5
    W, H = w * dsf, h * dsf
    tx, ty = random_taus(V, h, w, W, H, frac_invalid=0.1, margin=0.3)
    wd = RNG.uniform(0.2, 2.0, (V, h, w))
    views = RNG.random((V, 3, h, w)).astype(np.float32)
    gx, gy = RNG.standard_normal((H, W)), RNG.standard_normal((H, W))
    alpha, c0, beta = 1.3, 0.05, 0.5
    u0 = RNG.random((3, H, W))
    Dx, Dy = dense_D(W, H)
    A = c0 * np.eye(W * H) + beta * (Dx.T @ Dx + Dy.T @ Dy)
    b = np.zeros((3, W * H))
    gxm, gym = gx.copy(), gy.copy()
    gxm[:, -1] = 0
    gym[-1, :] = 0
    for v in range(V):
        S = dense_warp(tx[v], ty[v], W, H, dsf, "cell")
        Wm = np.diag(wd[v].ravel())
        A += alpha * S.T @ Wm @ S
        b += alpha * (S.T @ Wm @ views[v].reshape(3, -1).astype(np.float64).T).T
    b += c0 * u0.reshape(3, -1) + beta * (Dx.T @ gxm.ravel() + Dy.T @ gym.ravel())
    d = RNG.standard_normal((H, W))
    np.testing.assert_allclose(oracle.apply_A(W, H, dsf, tx, ty, wd, alpha, c0, beta, d).ravel(),
                               A @ d.ravel(), atol=1e-11)
    np.testing.assert_allclose(oracle.diag_A(W, H, dsf, tx, ty, wd, alpha, c0, beta).ravel(), np.diag(A),
                               atol=1e-12)
    ob = oracle.rhs(W, H, dsf, tx, ty, views, wd, alpha, c0, u0, beta, gx, gy)
    np.testing.assert_allclose(ob.reshape(3, -1), b, atol=1e-11)
    X, st = oracle.pcg(W, H, dsf, tx, ty, wd, alpha, c0, beta, ob, np.zeros((3, H, W)), 1e-13, 500)
    assert st["status"] == 0 and all(st["converged"])
    want = np.linalg.solve(A, b.T).T
    np.testing.assert_allclose(X.reshape(3, -1), want, atol=1e-9 * np.abs(want).max())
    # the solution minimises the quadratic energy: E(X) <= E(X + small perturbation)
    E0 = oracle.gibr_energy(W, H, dsf, tx, ty, views, wd, alpha, c0, u0, beta, gx, gy, X)
    E1 = oracle.gibr_energy(W, H, dsf, tx, ty, views, wd, alpha, c0, u0, beta, gx, gy,
                            X + 1e-3 * RNG.standard_normal(X.shape))
    assert E0 < E1


def test_gibr_energy_matches_dense_quadratic():
    V, h, w, dsf = 2, 4, 4, 1
    W, H = 4, 4
    tx, ty = random_taus(V, h, w, W, H, margin=0.2)
    wd = RNG.uniform(0.5, 1.5, (V, h, w))
    views = RNG.random((V, 3, h, w)).astype(np.float32)
    gx, gy = RNG.standard_normal((H, W)), RNG.standard_normal((H, W))
    U = RNG.random((3, H, W))
    alpha, beta = 0.7, 0.4
    Dx, Dy = dense_D(W, H)
    e = 0.0
    for c in range(3):
        for v in range(V):
            S = dense_warp(tx[v], ty[v], W, H, dsf, "cell")
            keep = S.sum(1) > 0
            r = (S @ U[c].ravel() - views[v, c].ravel().astype(np.float64))[keep]
            e += alpha * np.sum(wd[v].ravel()[keep] * r * r)
        rx = (Dx @ U[c].ravel())[(np.arange(W * H) % W) < W - 1] - gx[:, :-1].ravel()
        ry = (Dy @ U[c].ravel())[np.arange(W * H) < W * (H - 1)] - gy[:-1, :].ravel()
        e += beta * (np.sum(rx * rx) + np.sum(ry * ry))
    got = oracle.gibr_energy(W, H, dsf, tx, ty, views, wd, alpha, 0.0, None, beta, gx, gy, U)
    assert abs(got - 0.5 * e) < 1e-12 * max(1.0, e)


# ---------------------------------------------------------------------------
# Initial image, hole filling, push-pull, u-gradient, PSNR
# ---------------------------------------------------------------------------

def test_initial_image_reproduces_constant_views():
    V, h, w, dsf = 3, 8, 8, 2
    W, H = 16, 16
    tx, ty = random_taus(V, h, w, W, H)
    views = np.empty((V, 3, h, w), np.float32)
    views[:, 0], views[:, 1], views[:, 2] = 0.2, 0.5, 0.9
    om = RNG.uniform(0.1, 3.0, (V, h, w))
    U, nm = oracle.initial_image(W, H, dsf, tx, ty, views, om)
    hit = nm > 0
    for c, val in enumerate((0.2, 0.5, 0.9)):
        np.testing.assert_allclose(U[c][hit], val, rtol=1e-6)
        assert np.all(U[c][~hit] == 0)


def test_hole_filling_constant_and_harmonic_ramp():
    H = W = 12
    vis = np.ones((H, W), np.uint8)
    vis[3:8, 4:9] = 0
    c = np.full((3, H, W), 0.6)
    c[:, ~vis.astype(bool)] = 0.0
    U, k = oracle.hole_filling(c, vis, iters=3000, eps=0.0)
    np.testing.assert_allclose(U, 0.6, atol=1e-6)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    ramp = np.stack([0.1 * xs + 0.05 * ys] * 3)
    U0 = ramp.copy()
    U0[:, ~vis.astype(bool)] = 0.0
    U, _ = oracle.hole_filling(U0, vis, iters=3000, eps=0.0)
    np.testing.assert_allclose(U, ramp, atol=1e-6)
    # visible pixels are never touched
    assert np.array_equal(U[:, vis.astype(bool)], ramp[:, vis.astype(bool)])


def test_push_pull_fills_constant_holes_and_returns_the_difference():
    # coco_vtv_push_pull ends with cuda_subtract_from_device (optimization.cu:
    # 629-646): the result is filled - input, i.e. the fill value in the holes
    # and 0 on visible pixels (reproduced quirk)
    H, W = 17, 23
    vis = (RNG.random((H, W)) > 0.4).astype(np.uint8)
    U0 = np.where(vis[None].astype(bool), 0.3, 0.0) * np.ones((3, 1, 1))
    U = oracle.push_pull(U0, vis, levels=10)
    hole = ~vis.astype(bool)
    np.testing.assert_allclose(U[:, hole], 0.3, atol=1e-12)
    np.testing.assert_allclose(U[:, ~hole], 0.0, atol=1e-15)


def test_u_gradient_of_linear_ramp_views():
    # identical views of a luminance ramp warped by a pure translation: the
    # splatted Sobel gradient equals the ramp slope wherever the source
    # Sobel is interior
    V, h, w, dsf = 2, 12, 12, 1
    W = H = 12
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    img = 0.02 * xs + 0.01 * ys
    views = np.stack([np.stack([img] * 3)] * V).astype(np.float32)
    tx = np.stack([xs + 0.5] * V).astype(np.float32)
    ty = np.stack([ys + 0.5] * V).astype(np.float32)
    om = np.ones((V, h, w))
    gx, gy, mask, nm = oracle.u_gradient(W, H, dsf, tx, ty, views, om, 0.5)
    np.testing.assert_allclose(gx[1:-1, 1:-1], 0.02, rtol=1e-6)
    np.testing.assert_allclose(gy[1:-1, 1:-1], 0.01, rtol=1e-6)
    np.testing.assert_allclose(mask, 0.25, rtol=1e-12)
    np.testing.assert_allclose(nm, V, rtol=1e-12)


def test_psnr_hand_cases():
    H, W = 10, 12
    A = np.zeros((3, H, W))
    B = np.full((3, H, W), 0.1)
    assert abs(oracle.psnr(A, B) - 20.0) < 1e-12
    C = A.copy()
    C[:, :2, :] = 1.0  # differences only inside the 2-px border
    C[:, :, -2:] = 1.0
    assert oracle.psnr(A, C) == sys.float_info.max  # coco::psnr returns DBL_MAX
    D = A.copy()
    D[0, 5, 5] = 1.0  # one pixel of one channel: mse = 1 / (3 * 48)
    assert abs(oracle.psnr(A, D) - 10 * np.log10(3 * 6 * 8)) < 1e-12


# ---------------------------------------------------------------------------
# IBR_color pieces: regulariser weight, energy, one FISTA step
# ---------------------------------------------------------------------------

def identity_warps(V, H, W):
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    return (np.stack([xs + 0.5] * V).astype(np.float32), np.stack([ys + 0.5] * V).astype(np.float32))


def test_regularizer_weight_closed_form_and_accumulating_norm():
    V, H, W = 3, 14, 15
    tx, ty = identity_warps(V, H, W)
    sig, om, lmax, lmin = 0.1, 40.0, 1.0, 0.05
    nm0 = np.full((H, W), 5.0)  # the reference does not clear its normalisation buffer
    g, nm = oracle.regularizer_weight(W, H, 1, tx, ty, np.full((V, H, W), om), sig, lmax, lmin, nm0)
    np.testing.assert_allclose(nm, 5.0 + V, rtol=1e-14)
    m = V * om * sig * sig / (5.0 + V)
    np.testing.assert_allclose(g, max(lmin, lmax - m / V), rtol=1e-13)


def test_sr_energy_closed_forms():
    V, H, W = 2, 9, 11
    tx, ty = identity_warps(V, H, W)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    a, b = 0.0625, -0.03125  # exact in float32
    U = np.stack([0.25 + a * xs + b * ys] * 3)
    views = np.stack([U] * V).astype(np.float32)
    om = np.full((V, H, W), 2.0)
    vis = np.ones((H, W), np.uint8)
    sig, lam = 0.5, 0.3
    e, parts = oracle.sr_energy(W, H, 1, tx, ty, views, om, sig, lam, 1.0, 1.0, 0, vis, U)
    np.testing.assert_allclose(parts, [(W - 2) * (H - 2) * (abs(a) + abs(b)), 0.0, 0.0], atol=1e-12)
    assert abs(e - lam * parts[0]) < 1e-12
    d = 0.125
    views2 = views.copy()
    views2[:, 0] += d  # a constant offset leaves every Sobel unchanged
    _, parts = oracle.sr_energy(W, H, 1, tx, ty, views2, om, sig, lam, 1.0, 1.0, 0, vis, U)
    np.testing.assert_allclose(parts[1], V * H * W * d * d * 2.0 * sig * sig, rtol=1e-12)
    assert abs(parts[2]) < 1e-20
    vis[4, 5] = 0  # the TV term skips the 3x3 neighbourhood of an invisible pixel
    _, parts = oracle.sr_energy(W, H, 1, tx, ty, views, om, sig, lam, 1.0, 1.0, 0, vis, U)
    assert abs(parts[0] - ((W - 2) * (H - 2) - 9) * (abs(a) + abs(b))) < 1e-12


def test_fista_step_constant_fixed_point():
    V, H, W = 2, 8, 8
    tx, ty = identity_warps(V, H, W)
    U0 = np.full((3, H, W), 0.25)
    views = np.full((V, 3, H, W), 0.25, np.float32)
    st = oracle.SrState(W, H, U0)
    for _ in range(3):
        oracle.fista_step(W, H, 1, tx, ty, views, np.ones((V, H, W)), 0.5, 1.0, 0.5, 1.0, 1, 0.05, 4, None, st)
    np.testing.assert_allclose(st.U, 0.25, atol=1e-15)


def test_fista_step_closed_form_with_zero_tv_radius():
    # g = 0 projects the dual to 0, so each inner step is u <- (tau F + lam u)/(tau + lam)
    # and k steps give F + (u0 - F) (lam / (tau + lam))^k; one identity view,
    # omega = 1, coverage 1: G = sigma^2 (U - v), F = clamp(U - G, 0, 1).
    H, W, k, lam, sig = 6, 7, 5, 0.05, 0.5
    tx, ty = identity_warps(1, H, W)
    U0 = RNG.uniform(0.2, 0.8, (3, H, W))
    v = RNG.uniform(0.0, 1.0, (1, 3, H, W)).astype(np.float32)
    st = oracle.SrState(W, H, U0)
    oracle.fista_step(W, H, 1, tx, ty, v, np.ones((1, H, W)), sig, 1.0, 0.0, 1.0, 0, lam, k,
                      np.zeros((H, W)), st)
    F = np.clip(U0 - sig * sig * (U0 - v[0].astype(np.float64)), 0, 1)
    tau = 0.3 / np.sqrt(8.0)
    want = F + (U0 - F) * (lam / (tau + lam)) ** k
    np.testing.assert_allclose(st.U, want, atol=1e-13)
    np.testing.assert_allclose(st.G, F, atol=1e-15)
    # first step: a_1 = (1 + sqrt 5) / 2 and theta = 0, so Uq holds the previous U
    assert abs(st.fista_alpha.value - 0.5 * (1 + np.sqrt(5))) < 1e-15
    np.testing.assert_allclose(st.Uq, U0, atol=1e-15)


# ---------------------------------------------------------------------------
# Gradient-constraint data term (vtv_sr_dataterm_gradients,
# optimization.cu:1969-2116) and its energy (E_grad, :1615-1717)
# ---------------------------------------------------------------------------

def mean_laplacian_of_quadratic(a, b):
    """cuda_laplacian_device (vtv_sr_kernels_unstructured.cu:720-760) away from
    the border: the mean of the 4 neighbours minus the centre.  For
    u = a x^2 + b y^2 + (linear) this is (2a + 2b) / 4."""
    return 0.5 * (a + b)


def quadratic(H, W, a, b, c0, cx, cy):
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    return c0 + a * (xs - cx) ** 2 + b * (ys - cy) ** 2


def block_warps(V, h, w, dsf, shifts=None):
    """Source pixel (x, y) of view v warped to the centre of the target block
    [dsf x, dsf x + dsf) (shifted by an integer number of blocks): the
    footprint is exactly that block."""
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    shifts = shifts or [(0, 0)] * V
    tx = np.stack([dsf * (xs + sx) + 0.5 * dsf for sx, _ in shifts]).astype(np.float32)
    ty = np.stack([dsf * (ys + sy) + 0.5 * dsf for _, sy in shifts]).astype(np.float32)
    return tx, ty


@pytest.mark.parametrize("dsf,gw_type,shifted", [(1, 0, False), (2, 0, False), (1, 1, False), (2, 1, True)])
def test_gradient_term_closed_form(dsf, gw_type, shifted):
    """Quadratic U (constant mean-Laplacian c_u per channel) and quadratic views
    (constant c_v): wherever every footprint involved is interior, the warped
    residual is r = dsf^2 c_u - c_v (the dsf^2 warp-Laplacian approximation,
    :2035-2037), the splat normalised by the footprint weights returns r, and
    G = -dt_beta * gradient_step * r, times omega sigma^2 when gw_type != 0
    (:2078-2113).  F = clamp(U - clamp(G, -1, 1), 0, 1) (:1785-1811).  A wrong
    sign of gradient_step, a lost dsf^2 or a lost omega sigma^2 changes G."""
    V, h, w = 3, 14, 16
    W, H = w * dsf, h * dsf
    shifts = [(0, 0), (1, 0), (0, -1)] if shifted else None
    tx, ty = block_warps(V, h, w, dsf, shifts)
    coeff_u = [(0.001, 0.0005), (-0.0008, 0.0004), (0.0006, 0.0009)]
    coeff_v = [(0.03, -0.01), (0.008, 0.016), (-0.02, -0.006)]
    U = np.stack([quadratic(H, W, a, b, 0.5, W / 2.3, H / 1.7) for a, b in coeff_u])
    views = np.stack([np.stack([quadratic(h, w, a, b, 0.4, w / 2.1, h / 2.6) for a, b in coeff_v])] * V)
    views = views.astype(np.float32)
    om_v = np.array([0.7, 1.9, 1.3])
    om = np.stack([np.full((h, w), o) for o in om_v])
    sig, beta, gstep = 0.8, 0.35, 0.9
    st = oracle.SrState(W, H, U)
    oracle.fista_step(W, H, dsf, tx, ty, views, om, sig, 0.0, beta, gstep, gw_type, 0.05, 0, None, st)
    F = st.G
    # target pixels all of whose contributions come from view pixels at least
    # 2 away from the view border, whose footprints are at least 2 away from
    # the target border
    ok = np.zeros((H, W), bool)
    for v in range(V):
        sx, sy = shifts[v] if shifted else (0, 0)
        for y in range(2, h - 2):
            for x in range(2, w - 2):
                X, Y = dsf * (x + sx), dsf * (y + sy)
                if 2 <= X and X + dsf <= W - 2 and 2 <= Y and Y + dsf <= H - 2:
                    ok[Y:Y + dsf, X:X + dsf] = True
    # ... and no contribution from a view pixel outside that interior
    bad = np.zeros((H, W), bool)
    for v in range(V):
        sx, sy = shifts[v] if shifted else (0, 0)
        for y in range(h):
            for x in range(w):
                if 2 <= y < h - 2 and 2 <= x < w - 2:
                    continue
                X, Y = dsf * (x + sx), dsf * (y + sy)
                bad[max(Y, 0):max(Y + dsf, 0), max(X, 0):max(X + dsf, 0)] = True
    ok &= ~bad
    assert ok.sum() > 40
    scale = sig * sig * om_v.mean() if gw_type else 1.0
    for c in range(3):
        r = dsf * dsf * mean_laplacian_of_quadratic(*coeff_u[c]) - mean_laplacian_of_quadratic(*coeff_v[c])
        G = -beta * gstep * scale * r
        assert 1e-3 < abs(G) < 1.0
        np.testing.assert_allclose(F[c][ok], np.clip(U[c][ok] - G, 0, 1), atol=1e-12)


def test_gradient_term_sign_and_clamp():
    """The clamp of G to [-1, 1] (:1785-1789) and of F to [0, 1] (:1809-1811),
    with the closed form of the test above at dsf 1: a large curvature mismatch
    saturates G, and the sign of G follows -gradient_step."""
    V, h, w = 1, 10, 10
    tx, ty = block_warps(V, h, w, 1)
    U = np.stack([quadratic(h, w, 0.0, 0.0, 0.5, 5, 5)] * 3)
    views = np.stack([np.stack([quadratic(h, w, a, a, 0.5, 5, 5) for a in (4.0, -4.0, 0.5)])]).astype(np.float32)
    st = oracle.SrState(w, h, U)
    oracle.fista_step(w, h, 1, tx, ty, views, np.ones((V, h, w)), 1.0, 0.0, 1.0, 1.0, 0, 0.05, 0, None, st)
    inner = (slice(3, 7), slice(3, 7))
    # r = -c_v: G = +c_v * gradient_step; c_v = 4 -> G = 1 (clamped) -> F = 0 (clamped);
    # c_v = -4 -> G = -1 -> F = 1; c_v = 0.5 -> G = 0.5 -> F = 0
    assert np.all(st.G[0][inner] == 0.0) and np.all(st.G[1][inner] == 1.0)
    np.testing.assert_allclose(st.G[2][inner], 0.0, atol=1e-12)
    st = oracle.SrState(w, h, U)
    oracle.fista_step(w, h, 1, tx, ty, views, np.ones((V, h, w)), 1.0, 0.0, 1.0, 0.25, 0, 0.05, 0, None, st)
    np.testing.assert_allclose(st.G[2][inner], 0.5 - 0.25 * 0.5, atol=1e-12)


def luminance_ramp(H, W, slopes_x, slopes_y, base=(0.3, 0.4, 0.5)):
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    return np.stack([base[c] + slopes_x[c] * xs + slopes_y[c] * ys for c in range(3)])


REC709 = np.array([0.2126, 0.7152, 0.0722])


@pytest.mark.parametrize("gw_type", [0, 1])
def test_gradient_energy_closed_form(gw_type):
    """E_grad = sum_i sum_m |S_i Sobel(Y U) - Sobel(Y v_i)|^2 (x omega sigma^2):
    with identity warps and colour ramps, Sobel/8 of the Rec. 709 luminance
    is the luminance slope inside the 1-px border and 0 on it (both images
    share the border), so each view contributes (w-2)(h-2) |dU - dv|^2."""
    V, H, W = 3, 9, 12
    tx, ty = identity_warps(V, H, W)
    ux, uy = (0.01, -0.02, 0.03), (0.005, 0.015, -0.01)
    U = luminance_ramp(H, W, ux, uy)
    vslopes = [((0.02, 0.0, 0.01), (0.0, 0.01, 0.0)), ((-0.01, 0.01, 0.0), (0.01, 0.0, 0.02)),
               ((0.0, -0.02, 0.05), (0.02, 0.02, -0.03))]
    views = np.stack([luminance_ramp(H, W, sx, sy) for sx, sy in vslopes]).astype(np.float32)
    om_v = np.array([0.5, 2.0, 1.25])
    om = np.stack([np.full((H, W), o) for o in om_v])
    sig = 0.6
    gUx, gUy = REC709 @ np.array(ux), REC709 @ np.array(uy)
    want = 0.0
    for v, (sx, sy) in enumerate(vslopes):
        d2 = (gUx - REC709 @ np.array(sx)) ** 2 + (gUy - REC709 @ np.array(sy)) ** 2
        want += (W - 2) * (H - 2) * d2 * (om_v[v] * sig * sig if gw_type else 1.0)
    vis = np.ones((H, W), np.uint8)
    e, parts = oracle.sr_energy(W, H, 1, tx, ty, views, om, sig, 0.0, 0.0, 1.7, gw_type, vis, U)
    assert want > 1e-6
    # the views are float32 copies of the ramps: 1e-6 relative
    np.testing.assert_allclose(parts[2], want, rtol=1e-6)
    np.testing.assert_allclose(e, 1.7 * parts[2], rtol=1e-14)


def test_gradient_energy_with_a_translation():
    """A view shifted by t target pixels: its pixels whose tau leaves [0, W]
    gather 0, so there the term is |Sobel(Y v)|^2; the rest as above."""
    H, W, t = 8, 14, 3
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    tx = (xs + 0.5 + t)[None].astype(np.float32)
    ty = (ys + 0.5)[None].astype(np.float32)
    ux, uy = (0.02, 0.01, 0.0), (0.0, -0.01, 0.04)
    vx, vy = (0.01, 0.0, 0.02), (0.03, 0.0, 0.0)
    U = luminance_ramp(H, W, ux, uy)
    views = luminance_ramp(H, W, vx, vy)[None].astype(np.float32)
    gU = np.array([REC709 @ np.array(ux), REC709 @ np.array(uy)])
    gv = np.array([REC709 @ np.array(vx), REC709 @ np.array(vy)])
    want = 0.0
    for y in range(1, H - 1):
        for x in range(1, W - 1):  # the view's own Sobel is zero on its border
            X = x + t  # the target pixel the view pixel lands on (tau = X + .5)
            if X + 0.5 > W:
                want += np.sum(gv ** 2)              # gather invalid: S u = 0
            elif 1 <= X <= W - 2:
                want += np.sum((gU - gv) ** 2)       # interior target Sobel
            else:
                want += np.sum(gv ** 2)              # the target's border Sobel is 0
    # border rows / columns of the view: its Sobel is 0, the gathered one is not
    for y in range(H):
        for x in range(W):
            if 1 <= y <= H - 2 and 1 <= x <= W - 2:
                continue
            X = x + t
            if X + 0.5 <= W and 1 <= X <= W - 2 and 1 <= y <= H - 2:
                want += np.sum(gU ** 2)
    _, parts = oracle.sr_energy(W, H, 1, tx, ty, views, np.ones((1, H, W)), 1.0, 0.0, 0.0, 1.0, 0,
                                np.ones((H, W), np.uint8), U)
    np.testing.assert_allclose(parts[2], want, rtol=1e-6)


# ---------------------------------------------------------------------------
# Push-pull (optimization.cu:210-651) against dense pyramid matrices
# ---------------------------------------------------------------------------

PP_K = np.array([0.25 - 0.2, 0.25, 0.4, 0.25, 0.25 - 0.2])


def pp_reduce_matrix(n):
    """n -> n // 2 along one axis: row i takes taps 2i-2 .. 2i+2 with weights
    PP_K, indices below 0 read 0 and above n-1 read n-1 (reduce, :209-352)."""
    m = n // 2
    R = np.zeros((m, n))
    for i in range(m):
        for q in range(5):
            R[i, min(max(2 * i + q - 2, 0), n - 1)] += PP_K[q]
    return R


def pp_expand_matrix(n, m):
    """m -> n along one axis (expand, :401-517): even i from coarse rows
    i/2 - 1, i/2, i/2 + 1 with 2 PP_K[0], 2 PP_K[2], 2 PP_K[4]; odd i from
    (i-1)/2 and (i+1)/2 with 2 PP_K[1], 2 PP_K[3]; coarse indices clamped to
    [0, m-1] (the centre tap included, where the reference reads past it)."""
    E = np.zeros((n, m))
    c = lambda k: min(max(k, 0), m - 1)  # noqa: E731
    for i in range(n):
        if i % 2 == 0:
            for k, q in ((i // 2 - 1, 0), (i // 2, 2), (i // 2 + 1, 4)):
                E[i, c(k)] += 2 * PP_K[q]
        else:
            for k, q in (((i - 1) // 2, 1), ((i + 1) // 2, 3)):
                E[i, c(k)] += 2 * PP_K[q]
    return E


def pp_reduce_dense(img, R_r, R_c):
    """The zero-skipping normalised reduce as masked matrix products: rows
    first, then columns; a sample equal to 0 carries no weight."""
    M = (img != 0).astype(np.float64)
    num, den = R_r @ (img * M), R_r @ M
    T = np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)
    M2 = (T != 0).astype(np.float64)
    num, den = (T * M2) @ R_c.T, M2 @ R_c.T
    return np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)


def pp_dense(U, vis, levels):
    out = np.empty_like(U)
    H, W = vis.shape
    for ch in range(U.shape[0]):
        imgs, viss, sizes = [U[ch].copy()], [vis.astype(bool)], [(H, W)]
        for s in range(1, levels + 1):
            Hs, Ws = H >> (s - 1), W >> (s - 1)
            hs, ws = H >> s, W >> s
            if hs == 0 or ws == 0:
                break
            Rr, Rc = pp_reduce_matrix(Hs), pp_reduce_matrix(Ws)
            imgs.append(pp_reduce_dense(imgs[-1], Rr, Rc))
            viss.append(((Rr > 0).astype(int) @ viss[-1].astype(int) @ (Rc > 0).astype(int).T) > 0)
            sizes.append((hs, ws))
        for s in range(len(imgs) - 1, 0, -1):
            (Hs, Ws), (hs, ws) = sizes[s - 1], sizes[s]
            up = pp_expand_matrix(Hs, hs) @ imgs[s] @ pp_expand_matrix(Ws, ws).T
            imgs[s - 1] = np.where(viss[s - 1], imgs[s - 1], up)
        out[ch] = imgs[0] - U[ch]
    return out


@pytest.mark.parametrize("shape,levels", [((20, 24), 1), ((20, 24), 3), ((37, 29), 10), ((16, 16), 10)])
def test_push_pull_matches_dense_pyramid(shape, levels):
    """Random values (zeros included, which the reduce skips), random
    visibility with a large hole block, odd and even sizes."""
    H, W = shape
    U = RNG.uniform(0.1, 1.0, (3, H, W))
    U[RNG.random((3, H, W)) < 0.1] = 0.0
    vis = (RNG.random((H, W)) > 0.3).astype(np.uint8)
    vis[4:14, 5:17] = 0
    U[:, ~vis.astype(bool)] = 0.0
    got = oracle.push_pull(U, vis, levels)
    want = pp_dense(U, vis, levels)
    np.testing.assert_allclose(got, want, atol=1e-13)
    assert np.all(got[:, vis.astype(bool)] == 0.0)


def test_push_pull_fills_a_hole_in_a_ramp():
    """A hole in a linear ramp comes back close to the ramp: the symmetric
    5-tap kernels reproduce linear functions wherever no sample is skipped,
    so what is left is the bias of the normalised zero-skip near the hole;
    a swapped tap or even/odd rule leaves errors of the order of the slope
    times the hole size."""
    H, W = 32, 32
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    ramp = 0.2 + 0.015 * xs + 0.01 * ys
    vis = np.ones((H, W), np.uint8)
    vis[12:20, 10:18] = 0
    U = np.stack([ramp] * 3)
    U[:, ~vis.astype(bool)] = 0.0
    filled = oracle.push_pull(U, vis, 10) + U
    hole = ~vis.astype(bool)
    err = np.abs(filled[0][hole] - ramp[hole])
    assert err.max() < 0.03 and err.mean() < 0.012
    np.testing.assert_allclose(filled[:, ~hole], np.stack([ramp] * 3)[:, ~hole], atol=1e-15)


# ---------------------------------------------------------------------------
# Hole filling stop rule (gradientIBR.cu:87-152)
# ---------------------------------------------------------------------------

def test_hole_filling_stop_rule_single_hole():
    """One hole pixel with four visible neighbours of mean m: sweep 1 sets it to
    m with energy E1 = (m - u0) m (only the last channel's energy is kept,
    the variable is overwritten per channel), sweep 2 has E2 = 0, sweep 3 E3 =
    0.  The loop stops after the first sweep whose |E_prev - E| < eps (E_prev
    = 0 before the first)."""
    H = W = 5
    vis = np.ones((H, W), np.uint8)
    vis[2, 2] = 0
    U = np.zeros((3, H, W))
    U[:, 1, 2], U[:, 3, 2], U[:, 2, 1], U[:, 2, 3] = 0.2, 0.4, 0.6, 0.8  # m = 0.5
    U[0] += 0.3  # channel 0: m = 0.8, large energy, not looked at
    out, k = oracle.hole_filling(U, vis, iters=500, eps=1e-3)
    assert k == 3                      # E1 = 0.25 >= eps, |E1 - E2| = 0.25, |E2 - E3| = 0
    np.testing.assert_allclose(out[2, 2, 2], 0.5, atol=1e-15)
    np.testing.assert_allclose(out[0, 2, 2], 0.8, atol=1e-15)
    # u0 such that E1 = (m - u0) m < eps on the last channel: one sweep only
    U[2, 2, 2] = 0.5 - 0.5e-3          # E1 = 2.5e-4
    _, k = oracle.hole_filling(U, vis, iters=500, eps=1e-3)
    assert k == 1


def test_hole_filling_stop_rule_geometric_decay():
    """A 1 x 4 row with visible ends at 1 and two holes starting at 0: each
    Jacobi sweep halves the error, u_k = 1 - 2^-k, the sweep's laplacian is
    2^-k, so E_k = 2 * 2^-k (1 - 2^-k).  With eps = 1e-3 the first k with
    |E_{k-1} - E_k| < eps (E_0 = 0) is the number of sweeps run."""
    vis = np.array([[1, 0, 0, 1]], np.uint8)
    U = np.zeros((3, 1, 4))
    U[:, 0, 0] = U[:, 0, 3] = 1.0
    E = lambda k: 0.0 if k == 0 else 2.0 * 2.0 ** -k * (1 - 2.0 ** -k)  # noqa: E731
    want = next(k for k in range(1, 100) if abs(E(k - 1) - E(k)) < 1e-3)
    assert want == 11
    out, k = oracle.hole_filling(U, vis, iters=500, eps=1e-3)
    assert k == want
    np.testing.assert_allclose(out[:, 0, 1:3], 1 - 2.0 ** -want, atol=1e-15)
    # an iteration cap below it: every sweep runs
    _, k = oracle.hole_filling(U, vis, iters=7, eps=1e-3)
    assert k == 7


# ---------------------------------------------------------------------------
# Occlusion round trip of import_warps_from_depth (import.cpp:363-442)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dsf", [1, 2])
def test_occlusion_round_trip_foreground_square(dsf):
    """A fronto-parallel background at Zb with a foreground square at Zf in
    front of it, views on a horizontal line, the target between them.  A
    background source pixel that lands inside the target's projection of the
    square reads the square's depth there; back-projected at that depth
    through the target camera and projected into its own view it moves by
    f dB (1/Zb - 1/Zf) px, beyond the 10 px tolerance for every view here,
    so it is rejected (-2w, -2h).  Background pixels landing outside the
    square, and every foreground pixel, round-trip to themselves and keep
    their tau.  Pixels within 3 px of the square's edge are not judged."""
    h, w = 40, 160
    W, H = w * dsf, h * dsf
    f, cx, cy = 100.0, w / 2.0, h / 2.0
    Zb, Zf = 10.0, 2.0
    X0, X1, Y0, Y1 = 0.15, 0.45, -0.12, 0.12            # the square in world units at Zf
    bu = 0.3
    bvs = [-0.45, -0.1, 0.7, 1.05]                       # |dB| >= 0.4: the round trip moves >= 16 px
    def P(fx, c_x, c_y, b):
        return np.array([[fx, 0, c_x, -fx * b], [0, fx, c_y, 0], [0, 0, 1, 0]], np.float64)
    Pv = np.stack([P(f, cx, cy, b) for b in bvs])
    Pu = P(f * dsf, cx * dsf, cy * dsf, bu)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    depths = np.empty((len(bvs), h, w), np.float32)
    fg = []
    for v, b in enumerate(bvs):
        Xw = (xs + 0.5 - cx) * Zf / f + b
        Yw = (ys + 0.5 - cy) * Zf / f
        m = (Xw >= X0) & (Xw <= X1) & (Yw >= Y0) & (Yw <= Y1)
        depths[v] = np.where(m, Zf, Zb)
        fg.append(m)
    tx, ty, _, _, _ = oracle.warps_from_depth(depths, Pv, Pu, dsf)
    # the square's projection in the target (target pixel units)
    sx0, sx1 = dsf * (f * (X0 - bu) / Zf + cx), dsf * (f * (X1 - bu) / Zf + cx)
    sy0, sy1 = dsf * (f * Y0 / Zf + cy), dsf * (f * Y1 / Zf + cy)
    margin = 3.0 + dsf
    judged = rejected = kept = 0
    for v, b in enumerate(bvs):
        assert abs(f * (b - bu) * (1 / Zb - 1 / Zf)) > 12      # the round trip moves it beyond 10 px
        for y in range(h):
            for x in range(w):
                d = depths[v, y, x]
                ex = dsf * (x + 0.5) + dsf * f * (b - bu) / d
                ey = dsf * (y + 0.5)
                if not (margin <= ex <= W - margin and margin <= ey <= H - margin):
                    continue
                if fg[v][y, x]:
                    # foreground pixels near the square's own edge mix depths too
                    if not (sx0 + margin <= ex <= sx1 - margin and sy0 + margin <= ey <= sy1 - margin):
                        continue
                    assert tx[v, y, x] >= 0, (v, y, x)
                    kept += 1
                    judged += 1
                    continue
                inside = sx0 + margin <= ex <= sx1 - margin and sy0 + margin <= ey <= sy1 - margin
                outside = ex < sx0 - margin or ex > sx1 + margin or ey < sy0 - margin or ey > sy1 + margin
                if inside:
                    assert tx[v, y, x] == -2 * w and ty[v, y, x] == -2 * h, (v, y, x)
                    rejected += 1
                    judged += 1
                elif outside:
                    assert abs(tx[v, y, x] - ex) < 1e-3 and abs(ty[v, y, x] - ey) < 1e-3, (v, y, x)
                    kept += 1
                    judged += 1
    assert rejected > 50 and kept > 500
