"""The drop-in CLI (vissws_recon_online) end to end on the GPU: flags, the
segment loop, and the reference's output files (main_recon_online.cu:238-267,
358-362, 516-538; mirror variant --day_start/--day_count/--C_mode)."""
import os
import subprocess

import numpy as np
import pytest

from viss_optimize_ws_amd import Pipeline, default_config, make_sky

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "viss_optimize_ws_amd", "bin", "vissws_recon_online")


def _run(args, tmp):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_cli_circular_matches_library(tmp_path):
    out = tmp_path / "out"
    args = ["--geometry=circular", "--sky=lognormal", "--sky_seed=5", "--nside=16", "--n_ant=6", "--n_epochs=12",
            "--segs=3", "--btag=10M", "--recon=mirror", "--C_mode=txt", "--write_viss=1", "--write_baseline_txt=1",
            f"--out_dir={out}", "--img_n=32"]
    log = _run(args, tmp_path)
    assert log.count("[day 1 seg ") == 3 and "Recon done" in log
    C = np.loadtxt(out / "C1day10M.txt")
    V = np.loadtxt(out / "Viss1day10M.txt")
    uvw = np.loadtxt(out / "uvw1day10M.txt")
    xa = np.loadtxt(out / "xyza1day10M.txt")
    img = np.fromfile(out / "img1day10M.bin", np.float64).reshape(32, 32)
    P = 15
    assert V.shape == (2 * P * 12, 2) and uvw.shape == (2 * P * 12, 3) and xa.shape == uvw.shape
    cfg = default_config(1, nside=16, n_ant=6, n_epochs=12, segs=3, img_n=32, freq_hz=1e7)
    with Pipeline(cfg) as p:
        p.set_sky(make_sky(5, p.plan.npix, 0))
        full_v, geo = [], []
        for t0, tl in p.schedule():
            geo.append(p.gen_segment(t0, tl, geometry=True))
            p.viss_half()
            full_v.append(p.conj_fill()[1])
            p.recon_segment()
        Cm, W, Im, Wi = p.finalize()
    full_v = np.concatenate(full_v)
    # text formats: 6 significant digits of float32 values (the reference prints floats)
    assert np.allclose(C, Cm.astype(np.float32), rtol=1e-5, atol=1e-6 * np.abs(Cm).max())
    assert np.allclose(V[:, 0], full_v.real, rtol=1e-5, atol=1e-6 * np.abs(full_v).max())
    assert np.allclose(V[:, 1], full_v.imag, rtol=1e-5, atol=1e-6 * np.abs(full_v).max())
    u = np.concatenate([g["u"] for g in geo])
    assert np.allclose(uvw[:, 0], u, rtol=1e-14, atol=0)
    assert np.array_equal(img, Im)
    # conjugate completion visible in the file: row P+j of each epoch is the conjugate of row j
    Vr = V.reshape(12, 2 * P, 2)
    assert np.array_equal(Vr[:, :P, 0], Vr[:, P:, 0]) and np.array_equal(Vr[:, :P, 1], -Vr[:, P:, 1])


def test_cli_lunar_multiday_bin(tmp_path):
    out = tmp_path / "o"
    sky_dir = tmp_path / "sky"
    sky_dir.mkdir()
    npix = 12 * 4 * 4
    B = np.random.default_rng(3).lognormal(0, 1, npix).astype(np.float32)
    B.tofile(sky_dir / "B_1M.bin")
    args = ["--btag=1M", "--nside=4", "--segs=20", "--day_start=2", "--day_count=2", "--C_mode=bin",
            "--recon=mirror", f"--sky_dir={sky_dir}", f"--out_dir={out}", "--print_stats=0"]
    log = _run(args, tmp_path)
    assert "Loaded B from" in log and "===== Day 3 / 3 =====" in log
    c2 = np.fromfile(out / "C2day1M.bin", np.float32)
    c3 = np.fromfile(out / "C3day1M.bin", np.float32)
    assert c2.shape == (npix,) and c3.shape == (npix,)
    assert np.isfinite(c2).all() and not np.array_equal(c2, c3)
    # day 2 alone gives the same day-2 map (accumulators reset per day)
    out2 = tmp_path / "o2"
    _run(["--btag=1M", "--nside=4", "--segs=20", "--day=2", "--C_mode=bin", "--recon=mirror",
          f"--sky_dir={sky_dir}", f"--out_dir={out2}", "--print_stats=0"], tmp_path)
    assert np.array_equal(np.fromfile(out2 / "C2day1M.bin", np.float32), c2)


def test_cli_checkpoint_resume(tmp_path):
    # an interrupted run resumed from its segment-level checkpoint writes the
    # same files, byte for byte, as an uninterrupted one
    base = ["--geometry=circular", "--sky=lognormal", "--sky_seed=7", "--nside=16", "--n_ant=6", "--n_epochs=12",
            "--segs=3", "--btag=10M", "--recon=mirror", "--C_mode=txt", "--write_viss=1", "--write_baseline_txt=1",
            "--img_n=16", "--day_start=1", "--day_count=2", "--print_stats=0"]
    ref, part = tmp_path / "ref", tmp_path / "part"
    _run(base + [f"--out_dir={ref}"], tmp_path)
    ck = tmp_path / "ck.bin"
    log1 = _run(base + [f"--out_dir={part}", f"--checkpoint={ck}", "--stop_after_segs=2"], tmp_path)
    assert "Stopped after 2 segments" in log1 and ck.exists()
    # rows written after the checkpoint (a crash between segments) are dropped on resume
    with open(part / "Viss1day10M.txt", "a") as f:
        f.write("1 2\n")
    log2 = _run(base + [f"--out_dir={part}", f"--checkpoint={ck}", "--resume=1", "--stop_after_segs=2"], tmp_path)
    assert "Resuming from" in log2 and "day 1 segment 3" in log2 and "Stopped after 2 segments" in log2
    log3 = _run(base + [f"--out_dir={part}", f"--checkpoint={ck}", "--resume=1"], tmp_path)
    assert "day 2 segment 2" in log3
    for name in ("C1day10M.txt", "C2day10M.txt", "Viss1day10M.txt", "Viss2day10M.txt", "uvw2day10M.txt",
                 "xyzb1day10M.txt", "img1day10M.bin", "img2day10M.bin"):
        assert (ref / name).read_bytes() == (part / name).read_bytes(), name
    # a checkpoint from other parameters is refused
    other = [a.replace("--sky_seed=7", "--sky_seed=8") for a in base]
    r = subprocess.run([CLI] + other + [f"--out_dir={part}", f"--checkpoint={ck}", "--resume=1"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 1 and "other parameters" in r.stderr


def test_accumulators_roundtrip():
    cfg = default_config(1, nside=16, n_ant=6, n_epochs=12, segs=3, img_n=16, freq_hz=1e7, blockage=0)
    with Pipeline(cfg) as p:
        p.set_sky(make_sky(3, p.plan.npix, 0))
        sched = p.schedule()
        for t0, tl in sched[:2]:
            p.gen_segment(t0, tl); p.viss_half(); p.conj_fill(); p.recon_segment()
        acc = p.accumulators()
        assert acc[4] > 0  # blockage=0 divides by the running key total
        t0, tl = sched[2]
        p.gen_segment(t0, tl); p.viss_half(); p.conj_fill(); p.recon_segment()
        full = p.finalize()
        p.reset_accumulators()
        p.set_accumulators(*acc)
        assert all(np.array_equal(a, b) for a, b in zip(p.accumulators()[:4], acc[:4]))
        p.gen_segment(t0, tl); p.viss_half(); p.conj_fill(); p.recon_segment()
        again = p.finalize()
    for a, b in zip(full, again):
        assert np.array_equal(a, b)
