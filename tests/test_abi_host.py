"""CPU checks of the C-ABI library: it loads, exports every function the
public headers declare, and its host-only entry points (defaults, PSNR,
synthetic scenes, camera algebra) behave as documented.  No device compute
is called here."""
import ctypes
import glob
import os
import re
import sys

import numpy as np
import pytest

import lfrv
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(headers=None):
    names = set()
    for h in headers or glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"typedef[^;]*;", "", text)
        names |= set(re.findall(r"\b(lfrv_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared_functions()
    for must in ("lfrv_create", "lfrv_destroy", "lfrv_warps_from_depth", "lfrv_prepare", "lfrv_cg_solve",
                 "lfrv_sr_iteration_fista", "lfrv_get_solution", "lfrv_last_error"):
        assert must in names
    assert len(names) >= 40


def test_library_exports_every_declared_function():
    io_h = os.path.join(ROOT, "include", "lfrv_io.h")
    core = [h for h in glob.glob(os.path.join(ROOT, "include", "*.h")) if h != io_h]
    L = lfrv.lib()
    missing = [n for n in declared_functions(core) if not hasattr(L, n)]
    assert not missing, missing
    from lfrv import io as lio
    H = lio.lib()
    missing = [n for n in declared_functions([io_h]) if not hasattr(H, n)]
    assert not missing, missing


@pytest.mark.parametrize("cname,pyname", [("lfrv_cg_stats", "CgStats"), ("lfrv_sr_stats", "SrStats"),
                                           ("lfrv_params", "Params")])
def test_struct_layouts_match_the_header(tmp_path, cname, pyname):
    """The ctypes mirrors of the C-ABI structs have the header's size and
    field offsets (compiled with gcc against include/lfrv.h)."""
    import subprocess
    S = getattr(lfrv, pyname)
    fields = [f[0] for f in S._fields_]
    src = tmp_path / "layout.c"
    cfield = {"lambda_": "lambda"}  # a Python keyword on the ctypes side
    body = "".join(f'    printf("%zu\\n", offsetof({cname}, {cfield.get(f, f)}));\n' for f in fields)
    src.write_text(f"#include <stddef.h>\n#include <stdio.h>\n#include \"lfrv.h\"\nint main(void) {{\n"
                   f'    printf("%zu\\n", sizeof({cname}));\n{body}    return 0;\n}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got[0] == ctypes.sizeof(S)
    assert got[1:] == [getattr(S, f).offset for f in fields]


def test_host_library_exports_the_reference_named_interface():
    """Every function include/lfrv_ref.hpp declares is defined in liblfrv_host.so."""
    import subprocess
    from lfrv import io as lio
    text = open(os.path.join(ROOT, "include", "lfrv_ref.hpp")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = set(re.findall(r"^[A-Za-z_][\w:<>,* ]*?[ *&]([A-Za-z_][A-Za-z0-9_]*)\(", text, flags=re.M))
    assert {"init_data", "possion_conj_grad", "IBR_gradient", "coco_vtv_sr_iteration_fista_unstructured",
            "sr_synthesize_view", "import_cam"} <= declared
    out = subprocess.run(["nm", "-DC", "--defined-only", lio.HOST_LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    defined = set(re.findall(r"lfrv_ref::(\w+)\(", out))
    assert not (declared - defined), sorted(declared - defined)


def test_library_is_built_for_gfx950_only():
    so = os.path.join(ROOT, "light-field-remote-vision_amd", "liblfrv.so")
    blob = open(so, "rb").read()
    assert b"gfx950" in blob
    for other in (b"gfx942", b"gfx90a", b"gfx1100"):
        assert other not in blob


def test_default_params_follow_reference_defaults():
    p = lfrv.Params()
    lfrv.lib().lfrv_default_params(ctypes.byref(p), lfrv.MODE_GRADIENT_IBR, 9, 64, 48, 2)
    assert (p.nviews, p.w, p.h, p.dsf) == (9, 64, 48, 2)
    assert abs(p.sigma_sensor - 1.0 / 255.0) < 1e-9     # config.cpp:118-119 divides by 255
    assert p.ugrad_threshold == 1.0
    assert abs(p.gradient_step - 0.1) < 1e-7              # config.cpp default
    assert p.dt_alpha == 1.0 and p.dt_beta == 0.0
    assert p.occlusion_tol_px == 10.0 and p.sigma_z == 10.0


def test_psnr_host_entry_matches_oracle():
    rng = np.random.default_rng(5)
    a = rng.random((3, 20, 30)).astype(np.float32)
    b = (a + rng.normal(0, 0.02, a.shape)).astype(np.float32)
    assert abs(lfrv.psnr(a, b) - oracle.psnr(a, b)) < 1e-9
    assert lfrv.psnr(a, a) == sys.float_info.max   # DBL_MAX, as coco::psnr


@pytest.mark.parametrize("config", [0, 1, 2])
def test_synth_presets_are_deterministic_and_well_formed(config):
    s1 = lfrv.synth(config, seed=11, w=40, h=24)
    s2 = lfrv.synth(config, seed=11, w=40, h=24)
    s3 = lfrv.synth(config, seed=12, w=40, h=24)
    for k in ("views", "depths", "P_views", "P_target", "target"):
        assert np.array_equal(s1[k], s2[k])
    assert not np.array_equal(s1["views"], s3["views"])
    v = s1["views"]
    assert v.dtype == np.float32 and np.isfinite(v).all()
    assert v.min() >= -0.2 and v.max() <= 1.2
    assert s1["W"] == s1["w"] * s1["dsf"] and s1["H"] == s1["h"] * s1["dsf"]


def test_synth_config2_cameras_and_depths():
    s = lfrv.synth(2, seed=7, w=96, h=64)
    d = s["desc"]
    z = s["depths"]
    assert z.min() >= d.zmin - 1e-4 and z.max() <= d.zmax + 1e-4
    centres = []
    for P in s["P_views"]:
        K, R, t, C = oracle.decompose_P(P)
        np.testing.assert_allclose(R, np.eye(3), atol=1e-9)
        assert abs(K[0, 0] - d.focal) < 1e-6 * d.focal and abs(K[1, 1] - d.focal) < 1e-6 * d.focal
        assert abs(C[2]) < 1e-9
        centres.append(C[:2])
    centres = np.array(centres) / d.baseline
    grid = sorted(map(tuple, np.round(centres).astype(int)))
    assert grid == sorted((sx, sy) for sx in (-1, 0, 1) for sy in (-1, 0, 1))
    np.testing.assert_allclose(centres, np.round(centres), atol=1e-9)
    _, _, _, Ct = oracle.decompose_P(s["P_target"])
    assert np.all(np.abs(Ct[:2]) <= d.baseline + 1e-12)


def test_product_path_fails_loudly_without_a_device():
    # no GPU in this container: creating a context must raise, never fall back
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU device node is present")
    s = lfrv.synth(0, seed=1, w=32, h=32)
    from lfrv import pipelines
    with pytest.raises(lfrv.LfrvError):
        lfrv.Context(pipelines.gradient_ibr_params(s))


BAD_PARAMS = [
    ("nviews", 0), ("w", 2), ("h", 2), ("dsf", 0), ("dsf", 5), ("mode", 7),
    ("dw_type", 1),            # the reference asserts on the other deformation weights (gradientIBR.cu:636-641)
    ("pull_lds_kb", 200),      # more LDS than a CU has
    ("tile_w", 1),             # narrower than a footprint
    ("tile_h", 64),            # window rows beyond one wave
]


@pytest.mark.parametrize("field,value", BAD_PARAMS)
def test_create_rejects_bad_parameters(field, value):
    """lfrv_create checks its parameters before it touches the device: each
    bad field alone gives LFRV_E_ARG and no context."""
    p = lfrv.default_params(lfrv.MODE_GRADIENT_IBR, 9, 64, 48, 1)
    setattr(p, field, value)
    h = ctypes.c_void_p(1)
    assert lfrv.lib().lfrv_create(ctypes.byref(p), ctypes.byref(h)) == lfrv.E_ARG
    assert not h.value


def test_create_rejects_a_colour_problem_without_tv_weight():
    p = lfrv.default_params(lfrv.MODE_COLOR_SR, 9, 64, 48, 2, lambda_=0.0)
    h = ctypes.c_void_p(1)
    assert lfrv.lib().lfrv_create(ctypes.byref(p), ctypes.byref(h)) == lfrv.E_ARG
    assert not h.value
    assert lfrv.lib().lfrv_create(None, ctypes.byref(h)) == lfrv.E_ARG


class _RecordingLib:
    """Stands in for liblfrv.so: each input entry point reads the floats at
    the addresses it is given, at call time, as the library's copy would."""

    def __init__(self, nfloats):
        self.n = nfloats
        self.seen = {}

    def _read(self, name, *ptrs):
        self.seen[name] = [np.frombuffer(ctypes.string_at(p.value, 4 * self.n), np.float32).copy()
                           for p in ptrs if p is not None]
        return lfrv.OK

    def __getattr__(self, name):
        if name.startswith("lfrv_set_") or name in ("lfrv_compute_initial_image", "lfrv_sr_count_moving"):
            def f(h, *args):
                ptrs = [a for a in args if isinstance(a, ctypes.c_void_p)]
                return self._read(name, *ptrs)
            return f
        raise AttributeError(name)


@pytest.mark.parametrize("kind", ["float64", "strided", "float32"])
def test_converted_inputs_stay_alive_during_the_call(monkeypatch, kind):
    """Inputs that need a conversion to contiguous float32 (float64, strided
    views) are converted into a temporary; that temporary must still hold the
    data when the C function reads it.  4 MiB arrays: glibc returns such
    blocks to the system on free, so a dangling address would not read back
    the values."""
    n = 1 << 20
    rng = np.random.default_rng(1)
    base = rng.random(2 * n)
    x = {"float64": base[:n], "strided": base.astype(np.float32)[::2],
         "float32": base[:n].astype(np.float32)}[kind]
    assert x.size == n and (kind == "float32") == (x.flags["C_CONTIGUOUS"] and x.dtype == np.float32)
    want = np.asarray(x, np.float32).ravel()
    fake = _RecordingLib(n)
    monkeypatch.setattr(lfrv, "_lib", fake)
    ctx = object.__new__(lfrv.Context)
    ctx._h = ctypes.c_void_p(1)
    ctx.set_view(0, x)
    ctx.set_tau(0, x, x)
    ctx.set_partial_tau(0, x, x, x)
    ctx.set_depth(0, x, np.zeros(12))
    ctx.set_depth_gradient(0, x, x)
    ctx.compute_initial_image(x)
    ctx.set_solution(x)
    ctx.set_buffer("U", x)
    for name, arrays in fake.seen.items():
        assert arrays, name
        for a in arrays:
            assert np.array_equal(a, want), name
    ctx._h = ctypes.c_void_p()


# ---------------------------------------------------------------------------
# Quality harness: SSIM (gsl_matrix_ssim) and the psnr tool's line
# ---------------------------------------------------------------------------

def _ssim_lib(a, b, rng=1.0):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    L = lfrv.lib()
    L.lfrv_ssim.restype = ctypes.c_double
    L.lfrv_ssim.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
    return L.lfrv_ssim(a.shape[2], a.shape[1], a.ctypes.data, b.ctypes.data, rng)


def _normalised_gauss(img, sigma=1.5, taps=11):
    """Separable Gaussian as normalised convolution: (g * (x, 1)) / (g * 1)
    along x, then along y, on zero-padded data."""
    o = taps // 2
    k = np.exp(-0.5 * (np.arange(-o, o + 1) / sigma) ** 2)
    k /= k.sum()

    def along(x, axis):
        num = np.apply_along_axis(lambda r: np.convolve(r, k, mode="same"), axis, x)
        den = np.apply_along_axis(lambda r: np.convolve(r, k, mode="same"), axis, np.ones_like(x))
        return num / den
    return along(along(img, 1), 0)


def test_ssim_identical_and_constant_images():
    rng = np.random.default_rng(3)
    a = rng.random((3, 20, 23)).astype(np.float32)
    assert abs(_ssim_lib(a, a) - 1.0) < 1e-12
    # constant images: mu = the constants, every sigma 0
    ca, cb = np.float32(0.25), np.float32(0.75)
    A, B = np.full((3, 9, 12), ca, np.float32), np.full((3, 9, 12), cb, np.float32)
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    want = ((2 * 0.25 * 0.75 + C1) + C2) / ((0.25 ** 2 + 0.75 ** 2 + C1) + C2)
    assert abs(_ssim_lib(A, B) - want) < 1e-12
    # the dynamic range scales C1 and C2
    C1, C2 = 2.55 ** 2, 7.65 ** 2
    want = ((2 * 0.25 * 0.75 + C1) + C2) / ((0.25 ** 2 + 0.75 ** 2 + C1) + C2)
    assert abs(_ssim_lib(A, B, 255.0) - want) < 1e-12


def test_ssim_matches_a_normalised_convolution_restatement():
    """Random images against the reference's formula written with normalised
    convolutions: the factors of numerator and denominator added (as
    gsl_matrix_add_scaled adds them), the map averaged over every pixel and
    the three channels.  The standard (multiplied) SSIM differs."""
    rng = np.random.default_rng(4)
    a = rng.random((3, 17, 26)).astype(np.float32)
    b = np.clip(a + 0.1 * rng.standard_normal(a.shape), 0, 1).astype(np.float32)
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    s_add = s_mul = 0.0
    for c in range(3):
        A, B = a[c].astype(np.float64), b[c].astype(np.float64)
        m1, m2 = _normalised_gauss(A), _normalised_gauss(B)
        v1 = _normalised_gauss(A * A) - m1 * m1
        v2 = _normalised_gauss(B * B) - m2 * m2
        v12 = _normalised_gauss(A * B) - m1 * m2
        s_add += np.mean(((2 * m1 * m2 + C1) + (2 * v12 + C2)) / ((m1 * m1 + m2 * m2 + C1) + (v1 + v2 + C2)))
        s_mul += np.mean(((2 * m1 * m2 + C1) * (2 * v12 + C2)) / ((m1 * m1 + m2 * m2 + C1) * (v1 + v2 + C2)))
    got = _ssim_lib(a, b)
    assert abs(got - s_add / 3) < 1e-12
    assert abs(got - s_mul / 3) > 1e-3


def test_psnr_tool_prints_the_reference_line(tmp_path):
    """lfrv_psnr_tool a b prints '& PSNR & 10000 (1 - SSIM) / 2' (psnr_main.cpp:17-20)."""
    import subprocess
    from lfrv import io as lio
    rng = np.random.default_rng(5)
    a = rng.random((3, 24, 30)).astype(np.float32)
    b = np.clip(a + 0.05 * rng.standard_normal(a.shape), 0, 1).astype(np.float32)
    lio.save_image(str(tmp_path / "a.pfm"), a)
    lio.save_image(str(tmp_path / "b.pfm"), b)
    tool = os.path.join(ROOT, "light-field-remote-vision_amd", "lfrv_psnr_tool")
    r = subprocess.run([tool, str(tmp_path / "a.pfm"), str(tmp_path / "b.pfm")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    p = lfrv.psnr(a, b)
    s = _ssim_lib(a, b)
    assert r.stdout == "& %.4g & %d\n" % (p, int(10000 * (1 - s) / 2))
    assert abs(p - oracle.psnr(a, b)) < 1e-9
