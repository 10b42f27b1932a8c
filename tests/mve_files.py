"""Writers of MVE scene files for the tests: a view directory
views/view_%04i.mve/ holding meta.ini (the [camera] section: focal length in
units of the larger image side, pixel aspect, principal point in units of the
sides, row-major rotation, translation), .mvei images (the signature line,
width / height / channels / type as uint32, interleaved float32 rows top to
bottom) and undistorted.png.  Written from the format itself with numpy, not
with the library under test."""
import os
import struct

import numpy as np

MVEI_FLOAT = 9


def write_mvei(path, planes):
    planes = np.asarray(planes, np.float32)
    C, h, w = planes.shape
    with open(path, "wb") as f:
        f.write(b"\x89MVE_IMAGE\n")
        f.write(struct.pack("<4I", w, h, C, MVEI_FLOAT))
        f.write(np.ascontiguousarray(np.moveaxis(planes, 0, -1)).tobytes())


def write_meta(path, focal, aspect, pp, R, t, with_camera=True):
    lines = ["# MVE view meta data is stored in INI-like syntax.", ""]
    if with_camera:
        lines += ["[camera]", f"focal_length = {focal:.17g}", f"pixel_aspect = {aspect:.17g}",
                  f"principal_point = {pp[0]:.17g} {pp[1]:.17g}",
                  "rotation = " + " ".join(f"{x:.17g}" for x in np.asarray(R).ravel()),
                  "translation = " + " ".join(f"{x:.17g}" for x in np.asarray(t).ravel()), ""]
    lines += ["[view]", "id = 0", "name = test", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def view_dir(root, s):
    d = os.path.join(root, "views", f"view_{s:04d}.mve")
    os.makedirs(d, exist_ok=True)
    return d


def decompose(P):
    """K (upper triangular, positive diagonal, K[2,2] = 1), R (a rotation), t
    and the scale s with P = s K [R | t]."""
    from scipy.linalg import rq
    P = np.asarray(P, np.float64)
    if np.linalg.det(P[:, :3]) < 0:  # det R = sign det P[:, :3] once diag K > 0
        P = -P
    K, R = rq(P[:, :3])
    D = np.diag(np.sign(np.diag(K)))
    K, R = K @ D, D @ R
    t = np.linalg.solve(K, P[:, 3])
    s = K[2, 2]
    return K / s, R, t, s


def mve_params(K, W, H):
    """meta.ini's focal length, pixel aspect and principal point of an
    intrinsic matrix with zero skew for views of W x H."""
    f = K[0, 0] / max(W, H)
    return f, K[0, 0] / K[1, 1], (K[0, 2] / W, K[1, 2] / H)


def radial_from_ortho(z, K):
    """Radial depth (distance along the ray of each pixel centre) from
    orthographic depth: z |K^-1 (x + 0.5, y + 0.5, 1)|."""
    h, w = z.shape
    Ki = np.linalg.inv(K)
    X, Y = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    u = np.einsum("ij,jhw->ihw", Ki, np.stack([X, Y, np.ones_like(X)]))
    return z * np.sqrt((u * u).sum(0))
