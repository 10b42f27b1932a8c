"""The reference-format front end (liblfrv_host.so, lfrv_cli) on the CPU.

Config semantics follow cocolib/common/parse_config.cpp:43-196 and
sources/config.cpp:67-128; PFM follows gsl_image.cpp:416-502 and PNG the
QImage conversions of gsl_image.cpp:539-637.  The independent checks are
numpy's own PFM bytes and PIL's PNG codec.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from lfrv import io as lio


def _write_pfm_numpy(path, planes, little=True, gray=False):
    h, w = planes.shape[-2:]
    if gray:
        data = planes[0][::-1]
    else:
        data = np.stack([planes[0], planes[1], planes[2]], axis=-1)[::-1]
    dt = "<f4" if little else ">f4"
    with open(path, "wb") as f:
        f.write(f"{'Pf' if gray else 'PF'}\n{w} {h}\n{-1.0 if little else 1.0}\n".encode())
        f.write(np.ascontiguousarray(data, dtype=dt).tobytes())


def _read_pfm_numpy(path):
    with open(path, "rb") as f:
        t = f.readline().strip()
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        nc = 3 if t == b"PF" else 1
        a = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4").reshape(h, w, nc)[::-1]
    return t, scale, np.moveaxis(a, -1, 0).astype(np.float32)


@pytest.mark.parametrize("little", [True, False])
@pytest.mark.parametrize("gray", [False, True])
def test_pfm_load_matches_numpy(tmp_path, little, gray):
    rng = np.random.default_rng(1)
    img = rng.standard_normal((3, 7, 11)).astype(np.float32) * 100
    p = str(tmp_path / "a.pfm")
    _write_pfm_numpy(p, img, little=little, gray=gray)
    got = lio.load_image(p)
    want = np.repeat(img[:1], 3, 0) if gray else img
    assert np.array_equal(got, want)    # bit-exact, rows flipped top-down


def test_pfm_save_is_little_endian_bottom_up(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.random((3, 5, 9)).astype(np.float32)
    p = str(tmp_path / "b.pfm")
    lio.save_image(p, img)
    t, scale, back = _read_pfm_numpy(p)
    assert t == b"PF" and scale == -1.0
    assert np.array_equal(back, img)
    assert np.array_equal(lio.load_image(p), img)


def test_pfm_header_with_comment_lines(tmp_path):
    img = np.arange(3 * 2 * 3, dtype=np.float32).reshape(3, 2, 3)
    data = np.ascontiguousarray(np.stack(list(img), -1)[::-1], "<f4").tobytes()
    p = str(tmp_path / "c.pfm")
    with open(p, "wb") as f:
        f.write(b"PF\n# a comment\n3 2\n-1.0\n" + data)
    assert np.array_equal(lio.load_image(p), img)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA"])
def test_png_load_matches_pil(tmp_path, mode):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(3)
    nc = {"RGB": 3, "RGBA": 4, "L": 1, "LA": 2}[mode]
    a = rng.integers(0, 256, (13, 17, nc), dtype=np.uint8)
    # smooth ramps as well as noise so PIL's adaptive filter picks several filter types
    a[:6] = (np.arange(17)[None, :, None] * 7 + np.arange(6)[:, None, None] * 3) % 256
    p = str(tmp_path / "v.png")
    Image.fromarray(a[..., 0] if nc == 1 else a, mode).save(p, optimize=True)
    got = lio.load_image(p)
    rgb = np.repeat(a[..., :1], 3, -1) if nc <= 2 else a[..., :3]
    want = (np.moveaxis(rgb, -1, 0).astype(np.float64) / 255.0).astype(np.float32)
    assert np.array_equal(got, want)


def test_png_save_matches_reference_quantisation(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(4)
    img = rng.uniform(-0.2, 1.2, (3, 9, 10)).astype(np.float32)
    # exact multiples of 1/255 and values just below them
    img[0, 0, :5] = np.array([0, 1, 2, 128, 255], np.float32) / np.float32(255)
    img[1, 0, :5] = np.nextafter(img[0, 0, :5], np.float32(-1))
    p = str(tmp_path / "o.png")
    lio.save_image(p, img)
    got = np.asarray(Image.open(p).convert("RGB"))
    want = np.clip(np.trunc(img.astype(np.float64) * 255.0), 0, 255).astype(np.uint8)
    assert np.array_equal(np.moveaxis(got, -1, 0), want)


def test_png_round_trip_is_exact_on_8bit_values(tmp_path):
    a = (np.arange(3 * 16 * 16) % 256).astype(np.float64).reshape(3, 16, 16)
    img = (a / 255.0).astype(np.float32)
    p = str(tmp_path / "r.png")
    lio.save_image(p, img)
    assert np.array_equal(lio.load_image(p), img)


def test_unreadable_files_raise(tmp_path):
    p = tmp_path / "bad.pfm"
    p.write_bytes(b"P6\n1 1\n255\n\0\0\0")
    with pytest.raises(IOError):
        lio.load_image(str(p))
    with pytest.raises(IOError):
        lio.load_image(str(tmp_path / "missing.png"))


def test_config_file_and_command_line_order(tmp_path):
    out = tmp_path / "out"
    cfg = tmp_path / "scene.cfg"
    cfg.write_text(
        "# a comment line\n"
        "lf_name   views/v_%02d.png\n"
        f"outdir {out}\n"
        "sigma_sensor 2.55\n"
        "dsf 2\n"
        "s_min 0\n"
        "window_height_1 1\n"
        "lonely_word\n"
        "dt_beta 0,5\n"
        "tau_name some dir/tau_%02d.pfm\n")
    d = lio.config_data(["prog", "-s_min", "3", "-config", str(cfg), "-dsf", "3", "+nchannels", "-s_max", "8"])
    assert d["s_min"] == "0"            # the included file is read where "-config" stands
    assert d["dsf"] == "3"              # a later switch overrides it
    assert d["s_max"] == "8"
    assert d["tau_name"] == "some dir/tau_%02d.pfm"   # the value is the rest of the line
    assert float(d["sigma_sensor"]) == pytest.approx(2.55 / 255.0, rel=1e-6)  # always divided by 255
    assert d["outdir"] == str(out)     # read a second time, without the added '/'
    assert d["outfile"] == str(out) + "/"   # outdir + '/' + (no outfile)
    assert float(d["dt_beta"]) == 0.0  # C-locale parse stops at ','
    assert d["t_min"] == "1" and d["nmeta_iter"] == "1" and d["gradient_step"] == "0.100000001"
    assert os.path.isdir(out)           # outdir is created


def test_ibr_system_switch():
    """possion_conj_grad's system: the reference's screened Poisson system
    unless the extension switch asks for the normal equations."""
    assert lio.config_data(["prog"])["ibr_system"] == "poisson"
    d = lio.config_data(["prog", "-ibr_system", "normal_equations"])
    assert d["ibr_system"] == "normal_equations"


def test_config_defaults_match_config_data():
    d = lio.config_data(["prog"])
    want = dict(s_min="-1", s_max="-1", s_rmv="-1", t_min="1", t_max="1", t_rmv="-1", dsf="0", width="0",
                height="0", dw_type="0", gw_type="0", niter="0", nmeta_iter="1")
    for k, v in want.items():
        assert d[k] == v, k
    assert float(d["sigma_sensor"]) == pytest.approx(1.0 / 255.0 / 255.0, rel=1e-6)
    assert float(d["ugrad_threshold"]) == 1.0 and float(d["dt_alpha"]) == 1.0 and float(d["dt_beta"]) == 0.0


def test_cli_dumps_config_and_rejects_unknown_algorithm(tmp_path):
    out = tmp_path / "o"
    r = subprocess.run([lio.CLI_PATH, "-outdir", str(out), "-algorithm", "IBR_direct", "+verbose", "x"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "IBR_direct" in r.stderr
    lines = (out / "config").read_text().splitlines()
    assert lines == ["algorithm IBR_direct", f"outdir {out}", "verbose 1"]   # sorted switches


def test_cli_reports_missing_inputs(tmp_path):
    out = tmp_path / "o"
    r = subprocess.run([lio.CLI_PATH, "-outdir", str(out), "-algorithm", "IBR_gradient", "-nchannels", "3",
                        "-lf_name", str(tmp_path / "none_%02d.png"), "-s_min", "0", "-s_max", "1", "-dsf", "1"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot read view" in r.stderr


def test_cli_requires_three_channels(tmp_path):
    r = subprocess.run([lio.CLI_PATH, "-outdir", str(tmp_path), "-algorithm", "IBR_color"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "nchannels" in r.stderr


@pytest.mark.parametrize("algo", ["no_such_algorithm", "IBR_colour"])
def test_cli_runs_ibr_color_for_other_algorithm_names(tmp_path, algo):
    """main.cpp:430-434: a name none of the branches knows runs
    sr_synthesize_view (seen here by its channel check, before any GPU work)."""
    r = subprocess.run([lio.CLI_PATH, "-outdir", str(tmp_path), "-algorithm", algo],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "nchannels" in r.stderr


def test_cli_without_algorithm_is_the_opengl_viewer(tmp_path):
    """No algorithm switch starts the reference's OpenGL scene viewer (main.cpp:405-428), not provided here."""
    r = subprocess.run([lio.CLI_PATH, "-outdir", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "OpenGL" in r.stderr
