"""Host-side checks of libvissws.so (CPU only: no device compute).

* the library loads and exports every function include/vissws.h declares;
* Python defaults equal vw_config_init;
* without a GPU vw_create fails loudly (no CPU fallback);
* synthetic skies, the reference file formats (common.hpp:89-153,
  main_recon_online.cu:358-362, 516-538) and the CLI's argument errors.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import viss_optimize_ws_amd as V
from viss_optimize_ws_amd import _abi as A
from viss_optimize_ws_amd import quality

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vissws.h")
CLI = os.path.join(ROOT, "viss_optimize_ws_amd", "bin", "vissws_recon_online")


def _declared():
    txt = open(HEADER).read() + open(os.path.join(ROOT, "include", "vw_unet.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = V.load_library()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.vw_version().startswith(b"vissws")


@pytest.mark.parametrize("geom", [A.VW_GEOM_LUNAR8, A.VW_GEOM_CIRCULAR])
def test_python_defaults_equal_c_defaults(geom):
    lib = V.load_library()
    c = A.vw_config()
    assert lib.vw_config_init(C.byref(c), geom) == 0
    p = V.default_config(geom)
    for name, _ in A.vw_config._fields_:
        if name == "seg_tlen":
            continue
        assert getattr(c, name) == getattr(p, name), name
    assert lib.vw_config_init(C.byref(c), 7) == A.VW_ERR_ARG


def _has_gpu():
    lib = V.load_library()
    ctx = C.c_void_p()
    rc = lib.vw_create(C.byref(V.default_config(1, nside=1, n_epochs=2, segs=1, img_n=0)), C.byref(ctx))
    if rc == 0:
        lib.vw_destroy(ctx)
        return True
    return False


def test_create_fails_loudly_without_gpu():
    if _has_gpu():
        pytest.skip("a GPU is present")
    with pytest.raises(V.VissError) as e:
        V.Pipeline(V.default_config(1, nside=4, n_epochs=4, segs=1, img_n=0))
    assert e.value.code == A.VW_ERR_DEVICE
    assert "device" in str(e.value).lower()


def test_create_rejects_bad_configs():
    lib = V.load_library()
    ctx = C.c_void_p()
    for bad in (dict(tile_pix=300), dict(nside=0), dict(n_epochs=0), dict(blockage=2), dict(precision=5),
                dict(recon_mode=3), dict(n_ant=1),
                # batched skies: mixed precision only, one member per group, at most 4096
                dict(nsky=4), dict(nsky=4, precision=1, occ_mode=0), dict(nsky=4, precision=1, occ_mode=3),
                dict(nsky=5000, precision=1)):
        cfg = V.default_config(1, **{**dict(nside=4, n_epochs=4, segs=1, img_n=0), **bad})
        assert lib.vw_create(C.byref(cfg), C.byref(ctx)) == A.VW_ERR_ARG, bad
        assert lib.vw_last_error(None)
    lun = V.default_config(0, nside=4, n_ant=6)
    assert lib.vw_create(C.byref(lun), C.byref(ctx)) == A.VW_ERR_ARG
    over = V.default_config(1, nside=4, n_epochs=10, seg_tlen=[6, 6], img_n=0)
    assert lib.vw_create(C.byref(over), C.byref(ctx)) == A.VW_ERR_ARG


def test_make_sky():
    a = V.make_sky(7, 200000, 0)
    b = V.make_sky(7, 200000, 0)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, V.make_sky(8, 200000, 0))
    lg = np.log(a)
    assert abs(lg.mean()) < 0.01 and abs(lg.std() - 1.0) < 0.01
    c = V.make_sky(7, 200000, 1)
    d = c - a
    nsrc = np.count_nonzero(d)
    assert 50 <= nsrc <= 200
    assert d.min() >= 0 and d[d > 0].min() >= 1.0 - 1e-12  # Pareto x_m = 1


def test_make_sky_smooth_band():
    """kind 2: (1 + 9 exp(-|lat| / 8 deg)) times a field that is constant
    over the NESTED children of each Nside-16 pixel and log-normal(-1/8,
    1/2) across those pixels"""
    nside = 64
    npix = 12 * nside * nside
    s = V.make_sky(3, npix, 2)
    assert np.array_equal(s, V.make_sky(3, npix, 2))
    th, _ = quality.pix2ang_nest(nside, np.arange(npix))
    lat = np.pi / 2 - th
    profile = 1.0 + 9.0 * np.exp(-np.abs(lat) * 180.0 / (np.pi * 8.0))
    field = (s / profile).reshape(12 * 16 * 16, (nside // 16) ** 2)
    assert np.allclose(field, field[:, :1], rtol=1e-12)  # constant within each Nside-16 pixel
    g = np.log(field[:, 0])
    assert abs(g.mean() + 0.125) < 0.05 and abs(g.std() - 0.5) < 0.05
    with pytest.raises(V.VissError):
        V.make_sky(3, 12 * 8 * 8, 2)  # below Nside 16
    with pytest.raises(V.VissError):
        V.make_sky(3, 1000, 2)  # not a HEALPix size


def test_map_and_viss_writers(tmp_path):
    lib = V.load_library()
    x = np.array([1.5, -2.25e-7, 3.0e5, 0.0, 1.0 / 3.0])
    pb = str(tmp_path / "C.bin").encode()
    pt = str(tmp_path / "C.txt").encode()
    ptr = x.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.vw_io_write_map(pb, ptr, x.size, 0) == 0
    assert np.array_equal(np.fromfile(pb.decode(), np.float32), x.astype(np.float32))
    assert lib.vw_io_write_map(pt, ptr, x.size, 1) == 0
    assert open(pt.decode()).read().split("\n")[:5] == ["1.5", "-2.25e-07", "300000", "0", "0.333333"]
    v = np.array([1.0 + 2.0j, -0.1234567 + 1e-9j])
    pv = str(tmp_path / "V.txt").encode()
    assert lib.vw_io_write_viss(pv, v.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), 2, 0) == 0
    assert lib.vw_io_write_viss(pv, v.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), 1, 1) == 0
    assert open(pv.decode()).read() == "1 2\n-0.123457 1e-09\n1 2\n"
    u = np.array([1.0, -2.0]); w = np.array([3.0, 4.0]); z = np.array([1e-20, 5.5])
    p3 = str(tmp_path / "uvw.txt").encode()
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.vw_io_write_cols3(p3, f(u), f(w), f(z), 2, 0) == 0
    assert open(p3.decode()).read().splitlines() == [f"{a:.15e} {b:.15e} {c:.15e}" for a, b, c in zip(u, w, z)]
    assert lib.vw_io_write_map(b"/nonexistent/dir/C.bin", ptr, x.size, 0) == A.VW_ERR_IO


def test_sky_loader_formats(tmp_path):
    lib = V.load_library()
    npix = 12
    B = np.arange(npix, dtype=np.float64) * 0.5 + 0.25
    out = np.zeros(npix)
    used = C.create_string_buffer(512)
    optr = out.ctypes.data_as(C.POINTER(C.c_double))
    d = str(tmp_path).encode()
    # nothing there yet
    assert lib.vw_io_load_sky(d, b"10M", b"auto", npix, optr, used, 512) == A.VW_ERR_IO
    # text with the optional npix header line
    (tmp_path / "B.txt").write_text(f"{npix}\n" + "".join(f"{b}\n" for b in B))
    assert lib.vw_io_load_sky(d, b"10M", b"auto", npix, optr, used, 512) == 0
    assert np.array_equal(out, B) and used.value.endswith(b"B.txt")
    # headerless text, band-tagged name wins over B.txt
    (tmp_path / "B_10M.txt").write_text("".join(f"{b + 1}\n" for b in B))
    assert lib.vw_io_load_sky(d, b"10M", b"txt", npix, optr, used, 512) == 0
    assert np.array_equal(out, B + 1) and used.value.endswith(b"B_10M.txt")
    # binary float32 takes precedence in auto mode and must have the exact size
    (B.astype(np.float32) * 2).tofile(str(tmp_path / "B_10M.bin"))
    assert lib.vw_io_load_sky(d, b"10M", b"auto", npix, optr, used, 512) == 0
    assert np.array_equal(out, (B * 2).astype(np.float32).astype(np.float64))
    assert lib.vw_io_load_sky(d, b"10M", b"bin", npix + 1, optr, used, 512) == A.VW_ERR_IO
    assert lib.vw_io_load_sky(d, b"10M", b"xyz", npix, optr, used, 512) == A.VW_ERR_ARG


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_argument_errors(tmp_path):
    r = subprocess.run([CLI, "--btag=5M"], capture_output=True, text=True)
    assert r.returncode == 1 and "ERROR --btag" in r.stderr
    r = subprocess.run([CLI, "--precision=half"], capture_output=True, text=True)
    assert r.returncode == 1 and "ERROR --precision" in r.stderr
    r = subprocess.run([CLI, "--viss_tile_pix=100"], capture_output=True, text=True)
    assert r.returncode == 1
    r = subprocess.run([CLI, f"--sky_dir={tmp_path}", "--nside=4"], capture_output=True, text=True)
    assert r.returncode == 1 and "ERROR reading B" in r.stderr
    if not _has_gpu():
        r = subprocess.run([CLI, "--geometry=circular", "--sky=lognormal", "--nside=4", "--n_epochs=4",
                            f"--out_dir={tmp_path}/o"], capture_output=True, text=True)
        assert r.returncode == 1 and "vw_create" in r.stderr


def test_combine_tool_sums_days_in_order(tmp_path):
    # offline accumulation of per-day float32 maps (combine.py:42-126): day
    # order by the number in C<n>day..., float32 adds, optional previous sum
    import subprocess

    tool = os.path.join(ROOT, "viss_optimize_ws_amd", "bin", "vissws_combine")
    rng = np.random.default_rng(4)
    maps = {n: rng.normal(0, 1e3, 500).astype(np.float32) for n in (1, 2, 10)}
    for n, m in maps.items():
        m.tofile(tmp_path / f"C{n}day30M.bin")
    r = subprocess.run([tool, f"--dir={tmp_path}", "--pattern=C*day30M.bin"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exp = np.zeros(500, np.float32)
    for n in (1, 2, 10):
        exp += maps[n]
    out = np.fromfile(tmp_path / "accumulated_sum.bin", np.float32)
    assert np.array_equal(out, exp)
    meta = (tmp_path / "accumulated_sum_metadata.txt").read_text()
    assert "Number of files processed: 3" in meta
    assert meta.index("C1day") < meta.index("C2day") < meta.index("C10day")
    prev = (np.arange(500) * 0.5).astype(np.float32)
    prev.tofile(tmp_path / "prev.bin")
    r = subprocess.run([tool, f"--dir={tmp_path}", "--pattern=C*day30M.bin", f"--previous_sum={tmp_path}/prev.bin",
                        f"--output={tmp_path}/acc2.bin"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exp2 = prev.copy()
    for n in (1, 2, 10):
        exp2 += maps[n]
    assert np.array_equal(np.fromfile(tmp_path / "acc2.bin", np.float32), exp2)
    np.zeros(3, np.float32).tofile(tmp_path / "C3day30M.bin")
    r = subprocess.run([tool, f"--dir={tmp_path}", "--pattern=C*day30M.bin"], capture_output=True, text=True)
    assert r.returncode == 1 and "size mismatch" in r.stderr
