"""The HIP path at the sizes of the stated workloads, through the C-ABI,
against the CPU FP64 oracle (full segments where the oracle finishes in a few
seconds, bounded samples above that) and, for the sky batch, against the
one-sky plain FP64 path (needs a GPU).

Workloads:
  * Nside 128, 16 antennas, 512 epochs, 30 MHz, 16 segments, 768 tiles,
    512^2 image: plain FP64 form, two full segments against the oracle;
  * Nside 256, 32 antennas, 1,024 epochs, 10 MHz, 16 segments, 1024^2 image:
    plain and mixed forms on one segment, the oracle on 2,048 of its 31,744
    half rows and on 64 random HEALPix tiles + 64 random image tiles
    (16,384 + 16,384 points) of its reconstruction;
  * 64 skies (log-normal + point sources) at Nside 128 on an uneven schedule
    of segments drawn from [16, 256] epochs: the batched form against the
    one-sky plain form on four skies and against the oracle on 1,024 rows;
    and with every sky on its own such schedule (per-sky weights);
  * the deconvolver at 256 x 256 in FP16: forward and gradients against the
    oracle with the FP16 storage of the kernels;
  * the reference's lunar formation at 10 MHz and Nside 512: one segment of
    195,468 half rows in two tile passes, classes bit-exact over every
    (row, tile), Viss and recon on bounded samples.
Tolerances as in _lockstep: plain 1e-9, mixed 1e-4 (max |x - ref| / max |ref|
per output); masks, classes, keys, groups and weights bit-exact.
"""
import numpy as np
import pytest

from _lockstep import MIXED_TOL, PLAIN_TOL, config, rel_err, rms_rel_err, run_both
from oracle import unet_ref as U
from oracle.oracle import Oracle
from viss_optimize_ws_amd import Pipeline, make_sky
from viss_optimize_ws_amd.deconv import UNet

pytestmark = pytest.mark.gpu

NSIDE128 = dict(nside=128, n_ant=16, n_epochs=512, freq_hz=3e7, segs=16, img_n=512, tile_pix=256)
NSIDE256 = dict(nside=256, n_ant=32, n_epochs=1024, freq_hz=1e7, segs=16, img_n=1024, tile_pix=256)
SEG_KEYS = ("fa", "fb", "denom", "du", "res", "half", "nuniq", "nkeys", "n_half")


def _sky(npix, seed):
    return np.random.default_rng(seed).lognormal(0.0, 1.0, npix)


def _same_groups(a, b):
    ia, ka, oa, sa, ra = a.groups()
    ib, kb, ob, sb, rb = b.groups()
    for k in SEG_KEYS:
        assert ia[k] == ib[k], (k, ia[k], ib[k])
    assert np.array_equal(ka, kb) and np.array_equal(oa, ob)
    assert np.array_equal(sa, sb) and np.array_equal(ra, rb)


def test_nside128_plain_two_full_segments():
    """plain FP64 form, segments 1 and 10 of 16 in full: classes of both tile
    sets, masks of the first 64 rows, keys and groups bit-exact; Viss, the
    conjugate-completed rows and both maps within 1e-9; weights bit-exact"""
    cfg = config(**NSIDE128)
    errs = []

    def chk(g, o, k, d):
        assert np.array_equal(g.tile_classes(), o.tile_classes())
        assert np.array_equal(g.tile_classes(img=True), o.tile_classes(img=True))
        assert np.array_equal(g.masks(0, 64), o.masks(0, 64))
        for key in ("vg", "hg", "fg"):
            e = rel_err(d[key], d[key[0] + "o"])
            errs.append(e)
            assert e <= PLAIN_TOL, (k, key, e)
        _same_groups(g, o)

    (Cg, Wg, Ig, WIg), (Co, Wo, Io, WIo) = run_both(cfg, _sky(12 * 128 ** 2, 31), chk, segments={0, 9})
    assert np.array_equal(Wg, Wo) and np.array_equal(WIg, WIo)
    ec, ei = rel_err(Cg, Co), rel_err(Ig, Io)
    assert ec <= PLAIN_TOL and ei <= PLAIN_TOL
    print(f"Nside 128 plain, 2 segments: Viss max {max(errs):.2e}, C {ec:.2e}, image {ei:.2e}")


def _tile_points(ntile, npts, tp, nsel, rng):
    """all points of nsel random tiles (sorted), as indices into the point set"""
    tiles = np.sort(rng.choice(ntile, nsel, replace=False))
    pts = (tiles[:, None] * tp + np.arange(tp)[None, :]).ravel()
    return pts[pts < npts]


def test_nside256_plain_and_mixed_segment_sample():
    """one segment (31,744 half rows) in the plain and mixed forms.  Bit-exact:
    the per-(row, tile) classes of both tile sets, masks of 32 rows, fa/fb/RES,
    keys, groups and representatives, the weights of the sampled points.
    Against the oracle's Viss of 2,048 rows (the first and the last 1,024):
    plain <= 1e-9, mixed <= 1e-4.  Reconstruction: the oracle groups the
    plain device visibilities and evaluates 64 random HEALPix tiles and 64
    random image tiles; the device accumulators of the plain (<= 1e-9) and
    the mixed form (<= 1e-4, its own mixed visibilities) at those points."""
    seg = 1
    sky = _sky(12 * 256 ** 2, 8)
    cfg_p = config(**NSIDE256)
    cfg_m = config(precision=1, **NSIDE256)
    rng = np.random.default_rng(5)
    with Pipeline(cfg_p) as gp, Pipeline(cfg_m) as gm, Oracle(cfg_p) as o:
        for x in (gp, gm, o):
            x.set_sky(sky)
        t0, tl = gp.schedule()[seg]
        for x in (gp, gm, o):
            x.gen_segment(t0, tl)
        nh = gp.nh
        assert nh == 31744
        for img in (False, True):
            co = o.tile_classes(img=img)
            assert np.array_equal(gp.tile_classes(img=img), co)
            assert np.array_equal(gm.tile_classes(img=img), co)
            del co
        r32 = rng.integers(0, nh - 32)
        assert np.array_equal(gp.masks(r32, 32), o.masks(r32, 32))
        vp, vm = gp.viss_half(), gm.viss_half()
        rows = ((0, 1024), (nh - 1024, 1024))
        vo = np.concatenate([o.viss_half(r0, nr) for r0, nr in rows])
        sel = np.concatenate([np.arange(r0, r0 + nr) for r0, nr in rows])
        ev_p, ev_m = rel_err(vp[sel], vo), rel_err(vm[sel], vo)
        assert ev_p <= PLAIN_TOL, ev_p
        assert ev_m <= MIXED_TOL, ev_m
        o.set_viss_raw(vp)
        hp, _ = gp.conj_fill()
        hm, _ = gm.conj_fill()
        ho, _ = o.conj_fill()
        assert rel_err(hp, ho) <= PLAIN_TOL
        o.group_segment()
        gp.recon_segment()
        gm.recon_segment()
        _same_groups(gp, o)
        _same_groups(gm, o)
        Cp, Wp, Gp, GWp, _ = gp.accumulators()
        Cm, Wm, Gm, GWm, _ = gm.accumulators()
        out = {}
        for img, (accp, wp, accm, wm) in ((0, (Cp, Wp, Cm, Wm)), (1, (Gp, GWp, Gm, GWm))):
            n_pts = gp.plan.n_img_pts if img else gp.plan.npix
            ntile = gp.plan.n_img_tiles if img else gp.plan.ntile
            pts = _tile_points(ntile, n_pts, cfg_p.tile_pix, 64, rng)
            cs, ws = o.recon_points(img, pts)
            assert np.array_equal(wp[pts], ws) and np.array_equal(wm[pts], ws)
            ep, em = rel_err(accp[pts], cs), rel_err(accm[pts], cs)
            assert ep <= PLAIN_TOL, (img, ep)
            assert em <= MIXED_TOL, (img, em)
            out["image" if img else "healpix"] = (pts.size, ep, em, rms_rel_err(accm[pts], cs))
    print(f"Nside 256 segment {seg}: Viss plain {ev_p:.2e} mixed {ev_m:.2e}; recon (points, plain, mixed max, "
          f"mixed rms) {out}")


def _uneven_schedule(n_epochs, seed, lo=16, hi=256):
    rng = np.random.default_rng(seed)
    out, tot = [], 0
    while tot < n_epochs:
        n = min(int(rng.integers(lo, hi + 1)), n_epochs - tot)
        out.append(n)
        tot += n
    return out


def test_nside128_64_skies_uneven_schedule():
    """64 skies (log-normal + Poisson point sources) at Nside 128 on one
    uneven schedule: per-segment Viss of skies 0 and 63 against the oracle
    (first 1,024 rows of each segment), the maps of skies 0, 1, 31 and 63
    against the one-sky plain FP64 path (<= 1e-4), weights bit-exact"""
    W = dict(NSIDE128)
    W.pop("segs")
    sched = _uneven_schedule(W["n_epochs"], 4242)
    assert all(16 <= t <= 256 for t in sched[:-1])
    npix = 12 * 128 ** 2
    nsky = 64
    skies = np.stack([make_sky(7000 + s, npix, 1) for s in range(nsky)])
    check = (0, 1, 31, 63)
    ref = {}
    with Pipeline(config(precision=0, seg_tlen=sched, **W)) as p:
        for s in check:
            p.set_sky(skies[s])
            p.reset_accumulators()
            p.run_all()
            ref[s] = p.finalize()
    verr = []
    with Pipeline(config(precision=1, nsky=nsky, seg_tlen=sched, **W)) as b, \
            Oracle(config(precision=0, seg_tlen=sched, **W)) as o0, \
            Oracle(config(precision=0, seg_tlen=sched, **W)) as o63:
        b.set_sky(skies)
        o0.set_sky(skies[0])
        o63.set_sky(skies[63])
        for t0, tl in b.schedule():
            b.gen_segment(t0, tl)
            vb = b.viss_half()
            for s, o in ((0, o0), (63, o63)):
                o.gen_segment(t0, tl)
                vo = o.viss_half(0, 1024)
                verr.append(rel_err(vb[s, :1024], vo))
                assert verr[-1] <= MIXED_TOL, (s, t0, verr[-1])
            b.conj_fill()
            b.recon_segment()
        Cb, Wb, Ib, WIb = b.finalize()
    errs = []
    for s in check:
        Cr, Wr, Ir, WIr = ref[s]
        assert np.array_equal(Wb, Wr) and np.array_equal(WIb, WIr)
        errs.append((rel_err(Cb[s], Cr), rel_err(Ib[s], Ir)))
        assert max(errs[-1]) <= MIXED_TOL, (s, errs[-1])
    print(f"64 skies, schedule {sched}: Viss vs oracle max {max(verr):.2e}; maps vs plain (C, image) {errs}")


def test_nside128_64_skies_per_sky_schedules():
    """64 skies at Nside 128, each on its own uneven schedule (segments
    uniform in [16, 256] epochs, one seeded draw per sky): per-sky weights
    bit-exact and maps within 1e-4 of the one-sky plain FP64 path run on that
    sky's schedule, for skies 0, 1, 31 and 63"""
    W = dict(NSIDE128)
    W.pop("segs")
    nsky = 64
    scheds = [_uneven_schedule(W["n_epochs"], 4242 + s) for s in range(nsky)]
    npix = 12 * 128 ** 2
    skies = np.stack([make_sky(7000 + s, npix, 1) for s in range(nsky)])
    with Pipeline(config(precision=1, nsky=nsky, seg_tlen=[256, 256], **W)) as b:
        b.set_sky(skies)
        b.set_sky_schedules(scheds)
        b.run_all()
        Cb, Wb, Ib, WIb = b.finalize()
    errs = []
    for s in (0, 1, 31, 63):
        with Pipeline(config(seg_tlen=scheds[s], **W)) as p:
            p.set_sky(skies[s])
            p.run_all()
            Cp, Wp, Ip, WIp = p.finalize()
        assert np.array_equal(Wb[s], Wp) and np.array_equal(WIb[s], WIp), s
        errs.append((rel_err(Cb[s], Cp), rel_err(Ib[s], Ip)))
        assert max(errs[-1]) <= MIXED_TOL, (s, errs[-1])
    print(f"64 skies on their own schedules: maps vs plain (C, image) {errs}")


def _unet_params(seed):
    th = U.init_params(seed)
    rng = np.random.default_rng(seed + 1)
    P = U.unpack(th)
    for Wt, b in P:
        b[:] = rng.normal(0.0, 0.05, b.shape)
    P[14][0][:] = rng.normal(0.0, 0.5, P[14][0].shape)
    return np.concatenate([np.concatenate([Wt.ravel(), b.ravel()]) for Wt, b in P])


# FP16 storage, FP32 sums on the device against FP64 sums in the oracle: a
# stored value may round to the neighbouring FP16 value (2^-11 relative);
# the bound is per output and per layer, relative to its largest entry
UNET_FP16_TOL = 2e-3


def test_unet_256px_fp16_vs_oracle():
    """the deconvolver at its training size (256 x 256, 2 images, FP16
    activations on the matrix-core kernels): output, loss and the gradient of
    every layer against the oracle that rounds to FP16 where the kernels do"""
    N, n = 256, 2
    rng = np.random.default_rng(41)
    x = rng.normal(size=(n, N, N)).astype(np.float32)
    y = rng.normal(size=(n, N, N)).astype(np.float32)
    with UNet(N, n, fp16=True, seed=1) as net:
        net.set_params(_unet_params(42).astype(np.float32))
        th = net.params().astype(np.float64)
        out = net.predict(x)
        loss, g = net.loss_grad(x, y)
    ref, _ = U.forward(th, x, fp16=True)
    lref, gref = U.loss_grad(th, x, y, fp16=True)
    eo = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
    el = abs(loss - lref) / lref
    assert eo <= UNET_FP16_TOL, eo
    assert el <= UNET_FP16_TOL, el
    off, worst = 0, []
    for k, (ci, co, _) in enumerate(U.SPEC):
        m = co * ci * (9 if k < 14 else 1) + co
        gl, rl = g[off:off + m], gref[off:off + m]
        worst.append(float(np.max(np.abs(gl - rl)) / max(np.max(np.abs(rl)), 1e-30)))
        off += m
    print(f"U-Net 256^2 FP16: output {eo:.2e}, loss {el:.2e}, per-layer gradient max {max(worst):.2e} {worst}")
    assert max(worst) <= UNET_FP16_TOL, worst


def test_lunar_nside512_segment_sample():
    """the reference formation at 10 MHz and Nside 512 (3,145,728 pixels,
    12,288 tiles), its first segment of 195,468 half rows, whose (tile, row)
    class tables (2.4 GB) exceed the 2 GiB budget and run in two tile passes:
    every class bit-exact, masks of 16 rows bit-exact, keys, groups and
    representatives bit-exact, the plain Viss of 256 rows (first and last
    128) and the recon of 64 random tiles (16,384 pixels) within 1e-9 of the
    oracle, their weights bit-exact"""
    cfg = config(geometry=0, nside=512, freq_hz=1e7, segs=10, img_n=0)
    rng = np.random.default_rng(9)
    sky = _sky(12 * 512 ** 2, 19)
    with Pipeline(cfg) as g, Oracle(cfg) as o:
        g.set_sky(sky)
        o.set_sky(sky)
        t0, tl = g.schedule()[0]
        g.gen_segment(t0, tl)
        o.gen_segment(t0, tl)
        nh = g.nh
        assert nh == 195468
        assert np.array_equal(g.tile_classes(), o.tile_classes())
        r16 = int(rng.integers(0, nh - 16))
        assert np.array_equal(g.masks(r16, 16), o.masks(r16, 16))
        vg = g.viss_half()
        rows = ((0, 128), (nh - 128, 128))
        vo = np.concatenate([o.viss_half(r0, nr) for r0, nr in rows])
        sel = np.concatenate([np.arange(r0, r0 + nr) for r0, nr in rows])
        ev = rel_err(vg[sel], vo)
        assert ev <= PLAIN_TOL, ev
        o.set_viss_raw(vg)
        hg, _ = g.conj_fill()
        ho, _ = o.conj_fill()
        assert rel_err(hg, ho) <= PLAIN_TOL
        o.group_segment()
        g.recon_segment()
        _same_groups(g, o)
        Cg, Wg, _, _, _ = g.accumulators()
        pts = _tile_points(g.plan.ntile, g.plan.npix, cfg.tile_pix, 64, rng)
        cs, ws = o.recon_points(0, pts)
        assert np.array_equal(Wg[pts], ws)
        er = rel_err(Cg[pts], cs)
        assert er <= PLAIN_TOL, er
    print(f"lunar Nside 512, segment 1: Viss (256 rows) {ev:.2e}, recon (16,384 pixels) {er:.2e}")
