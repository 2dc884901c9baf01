"""gfx950 path vs the CPU FP64 oracle through the C-ABI (needs a GPU)."""
import numpy as np
import pytest

from _lockstep import MIXED_TOL, PLAIN_TOL, config, rel_err, rms_rel_err, run_both
from oracle.oracle import Oracle
from viss_optimize_ws_amd import Pipeline

pytestmark = pytest.mark.gpu


def _sky(npix, seed=3):
    return np.random.default_rng(seed).lognormal(0.0, 1.0, npix)


# the CPU-sized workload: Nside 32, 8 antennas, 64 epochs, 10 MHz,
# 4 segments, blockage=1, occ_mode=2, 256^2 image
C1 = dict(nside=32, n_ant=8, n_epochs=64, freq_hz=1e7, segs=4, img_n=256)


@pytest.mark.parametrize("nside,tp", [(4, 256), (32, 256), (32, 512), (64, 256)])
def test_resident_geometry_bitexact(nside, tp):
    cfg = config(nside=nside, tile_pix=tp, n_epochs=8, segs=1, img_n=48)
    with Pipeline(cfg) as g, Oracle(cfg) as o:
        for a, b in zip(g.lmn(), o.lmn()):
            assert np.array_equal(a, b)
        assert np.array_equal(g.tiles(), o.tiles())
        assert np.array_equal(g.tiles(img=True), o.tiles(img=True))
        for a, b in zip(g.img_points(), o.img_points()):
            assert np.array_equal(a, b)
        assert np.array_equal(g.positions(), o.positions())
        assert g.plan.bl_max_m == o.plan.bl_max_m


@pytest.mark.parametrize("geom", ["circular", "lunar"])
def test_segment_geometry_bitexact(geom):
    if geom == "lunar":
        cfg = config(geometry=0, nside=8, freq_hz=1e6, segs=10, img_n=0)
    else:
        cfg = config(**C1)
    with Pipeline(cfg) as g, Oracle(cfg) as o:
        assert np.array_equal(g.positions(), o.positions())
        for t0, tl in g.schedule()[:2]:
            gg = g.gen_segment(t0, tl, geometry=True)
            go = o.gen_segment(t0, tl, geometry=True)
            for k in gg:
                assert np.array_equal(gg[k], go[k]), k
            for a, b in zip(g.rows(), o.rows()):
                assert np.array_equal(a, b)
            assert np.array_equal(g.tile_classes(), o.tile_classes())
            if cfg.img_n:
                assert np.array_equal(g.tile_classes(img=True), o.tile_classes(img=True))
            nr = min(g.nh, 64)
            assert np.array_equal(g.masks(0, nr), o.masks(0, nr))


def test_viss_counts_and_masks_c1():
    cfg = config(**C1)
    sky = _sky(12 * 32 * 32)
    with Pipeline(cfg) as g, Oracle(cfg) as o:
        g.set_sky(sky)
        o.set_sky(sky)
        t0, tl = g.schedule()[1]
        g.gen_segment(t0, tl)
        o.gen_segment(t0, tl)
        vo, co = o.viss_half(counts=True)
        cg = g.viss_counts()
        assert np.array_equal(cg, co)
        # the effective mask equals the pure per-pixel predicate (cone invariant)
        assert np.array_equal(o.masks(0, o.nh), o.predicate(0, o.nh))
        vg = g.viss_half()
        assert rel_err(vg, vo) <= PLAIN_TOL


def _seg_checks(tol, exact_groups=True):
    def chk(g, o, k, d):
        assert rel_err(d["vg"], d["vo"]) <= tol, ("viss", k, rel_err(d["vg"], d["vo"]))
        assert rel_err(d["hg"], d["ho"]) <= tol
        assert rel_err(d["fg"], d["fo"]) <= tol
        ig, io = d["ig"], d["io"]
        for key in ("fa", "fb", "denom", "du", "res", "half", "nuniq", "nkeys"):
            assert ig[key] == io[key], (key, ig[key], io[key])
        if exact_groups:
            _, kg, og, sg, rg = g.groups()
            _, ko, oo, so, ro = o.groups()
            assert np.array_equal(kg, ko) and np.array_equal(og, oo)
            assert np.array_equal(sg, so) and np.array_equal(rg, ro)

    return chk


@pytest.mark.parametrize(
    "variant",
    [
        dict(),
        dict(tile_pix=512),
        dict(recon_mode=1),
        dict(occ_mode=0),
        dict(occ_mode=1),
        dict(occ_mode=3),
        dict(occ_mode=4),
        dict(recon_mode=1, occ_mode=3),
        dict(blockage=0),
        dict(recon_mode=1, blockage=0),
    ],
    ids=lambda v: ",".join(f"{k}={x}" for k, x in v.items()) or "default",
)
def test_pipeline_plain_c1(variant):
    kw = dict(C1)
    kw.update(variant)
    cfg = config(**kw)
    (Cg, Wg, Ig, WIg), (Co, Wo, Io, WIo) = run_both(cfg, _sky(12 * 32 * 32), _seg_checks(PLAIN_TOL))
    assert np.array_equal(Wg, Wo)
    assert np.array_equal(WIg, WIo)
    assert rel_err(Cg, Co) <= PLAIN_TOL
    assert rel_err(Ig, Io) <= PLAIN_TOL


def test_pipeline_lunar_reference_geometry():
    # reference formation (1 MHz band), small sky, 3 of its segments
    cfg = config(geometry=0, nside=8, freq_hz=1e6, segs=20, img_n=32)
    sky = _sky(12 * 8 * 8, seed=5)
    (Cg, Wg, Ig, WIg), (Co, Wo, Io, WIo) = run_both(cfg, sky, _seg_checks(PLAIN_TOL), segments={0, 7, 19})
    assert np.array_equal(Wg, Wo)
    assert rel_err(Cg, Co) <= PLAIN_TOL
    assert rel_err(Ig, Io) <= PLAIN_TOL


def test_pipeline_mixed_c1():
    cfg = config(precision=1, **C1)
    errs = {}

    def chk(g, o, k, d):
        errs.setdefault("viss", []).append(rel_err(d["vg"], d["vo"]))
        assert d["ig"]["nuniq"] == d["io"]["nuniq"]

    (Cg, Wg, Ig, WIg), (Co, Wo, Io, WIo) = run_both(cfg, _sky(12 * 32 * 32), chk)
    assert max(errs["viss"]) <= MIXED_TOL, errs
    assert np.array_equal(Wg, Wo)
    assert rel_err(Cg, Co) <= MIXED_TOL, rel_err(Cg, Co)
    assert rel_err(Ig, Io) <= MIXED_TOL, rel_err(Ig, Io)
    print("mixed C1 errors: viss max", max(errs["viss"]), "C", rel_err(Cg, Co), rms_rel_err(Cg, Co),
          "img", rel_err(Ig, Io), rms_rel_err(Ig, Io))


@pytest.mark.parametrize("recon_mode", [0, 1])
def test_large_segment_groups(recon_mode):
    # 16 antennas x 40 epochs = 4,800 half rows (9,600 full rows) in one segment:
    # a multi-block stable ranking with many equal keys
    cfg = config(nside=8, n_ant=16, n_epochs=40, segs=1, img_n=0, recon_mode=recon_mode, freq_hz=3e7)
    (Cg, Wg, _, _), (Co, Wo, _, _) = run_both(cfg, _sky(12 * 8 * 8, seed=9), _seg_checks(PLAIN_TOL))
    assert np.array_equal(Wg, Wo)
    assert rel_err(Cg, Co) <= PLAIN_TOL


def test_long_segment_frame_rows():
    # 32 antennas x 1,040 epochs in one segment: 33,280 antenna frames per
    # tile, a cone-bit row longer than the 32 KB the group classes stage in
    # LDS, so they gather the endpoint bits from global memory instead
    cfg = config(nside=8, n_ant=32, n_epochs=1040, segs=1, img_n=64, freq_hz=1e7)
    (Cg, Wg, Ig, IWg), (Co, Wo, Io, IWo) = run_both(cfg, _sky(12 * 8 * 8, seed=11), _seg_checks(PLAIN_TOL))
    assert np.array_equal(Wg, Wo) and np.array_equal(IWg, IWo)
    assert rel_err(Cg, Co) <= PLAIN_TOL
    assert rel_err(Ig, Io) <= PLAIN_TOL


@pytest.mark.parametrize("budget,variant", [(1, dict()), (7000, dict(recon_mode=1, occ_mode=3)),
                                           (7000, dict(precision=1))])
def test_tile_passes(budget, variant):
    # a class budget far below the full per-(tile, row) tables splits the Viss
    # and recon stages into many tile passes (budget 1: one tile per pass);
    # results must match the oracle as with a single pass
    kw = dict(C1)
    kw.update(variant)
    cfg = config(class_budget_bytes=budget, **kw)
    tol = MIXED_TOL if variant.get("precision") else PLAIN_TOL
    chk = _seg_checks(tol, exact_groups=not variant.get("precision"))
    (Cg, Wg, Ig, WIg), (Co, Wo, Io, WIo) = run_both(cfg, _sky(12 * 32 * 32), chk, segments={0, 2})
    assert np.array_equal(Wg, Wo)
    assert np.array_equal(WIg, WIo)
    assert rel_err(Cg, Co) <= tol
    assert rel_err(Ig, Io) <= tol


@pytest.mark.parametrize("recon_mode,precision", [(0, 0), (1, 0), (0, 1)])
def test_uneven_segments_and_row_padding(recon_mode, precision):
    # explicit uneven segments whose half-row counts (21 pairs x 17, 5, 31, 11
    # epochs) are multiples of neither the rows per wave nor the 64-row padding,
    # so the padded rows and partial row blocks are exercised
    cfg = config(nside=32, n_ant=7, seg_tlen=[17, 5, 31, 11], img_n=64, recon_mode=recon_mode, precision=precision,
                 freq_hz=2e7)
    tol = MIXED_TOL if precision else PLAIN_TOL
    chk = _seg_checks(tol, exact_groups=True) if not precision else None
    (Cg, Wg, Ig, WIg), (Co, Wo, Io, WIo) = run_both(cfg, _sky(12 * 32 * 32, seed=11), chk)
    assert np.array_equal(Wg, Wo) and np.array_equal(WIg, WIo)
    assert rel_err(Cg, Co) <= tol and rel_err(Ig, Io) <= tol


@pytest.mark.parametrize("form,precision", [(1, 0), (2, 0), (0, 1)], ids=["rows_fp64", "factorized_fp64", "mixed"])
def test_segments_off_schedule(form, precision):
    # vw_gen_segment accepts any segment up to the schedule's longest; shorter
    # ones ask for more tile chunks than the partial buffers were sized for
    # (sized on the schedule's 64-epoch segments: 35 chunks of the per-row
    # form, 128 x 64 epochs of the factorized one) and must run with the
    # chunk count capped.  8 epochs wants 192 per-row chunks; 50 epochs wants
    # 164 x 50 > 128 x 64 factorized chunk-epochs.
    cfg = config(nside=64, n_ant=16, n_epochs=128, segs=2, img_n=48, viss_form=form, precision=precision,
                 freq_hz=2e7)
    tol = MIXED_TOL if precision else (PLAIN_TOL if form == 1 else ANT_PLAIN_TOL)
    sky = _sky(12 * 64 * 64, seed=31)
    with Pipeline(cfg) as g, Oracle(cfg) as o:
        assert [tl for _, tl in g.schedule()] == [64, 64]
        g.set_sky(sky)
        o.set_sky(sky)
        for t0, tl in ((3, 8), (70, 50), (127, 1)):
            g.gen_segment(t0, tl)
            o.gen_segment(t0, tl)
            err = rel_err(g.viss_half(), o.viss_half())
            assert err <= tol, (t0, tl, err)
            g.conj_fill()
            o.conj_fill()
            ig, io = g.recon_segment(), o.recon_segment()
            assert ig["nkeys"] == io["nkeys"]
        Cg, Wg, Ig, WIg = g.finalize()
        Co, Wo, Io, WIo = o.finalize()
    assert np.array_equal(Wg, Wo) and np.array_equal(WIg, WIo)
    assert rel_err(Cg, Co) <= tol and rel_err(Ig, Io) <= tol


def test_lunar_30mhz_band():
    # the reference formation at its 30 MHz band (bl_max 100 km, RES ~ 46,000)
    cfg = config(geometry=0, nside=8, freq_hz=3e7, segs=20, img_n=32)
    (Cg, Wg, Ig, WIg), (Co, Wo, Io, WIo) = run_both(cfg, _sky(12 * 8 * 8, seed=4), _seg_checks(PLAIN_TOL),
                                                   segments={0, 13})
    assert np.array_equal(Wg, Wo)
    assert rel_err(Cg, Co) <= PLAIN_TOL
    assert rel_err(Ig, Io) <= PLAIN_TOL


# the antenna-factorized FP64 form sums exp(-2 pi j (psi_b - psi_a)) with
# per-antenna phases psi, so its phases differ from the per-row phase
# u l + v m + w (n-1) by FP64 rounding, about 2^-52 |u| turns per term (|u| up
# to ~1e6 wavelengths here): a looser bound than the per-row form's 1e-9
ANT_PLAIN_TOL = 1e-7


@pytest.mark.parametrize(
    "n_ant,form,precision",
    [(8, 1, 0), (8, 3, 0), (13, 2, 0), (16, 0, 1), (24, 2, 0), (24, 3, 1), (32, 2, 1), (32, 0, 1), (40, 0, 1)],
    ids=lambda v: str(v),
)
def test_viss_forms(n_ant, form, precision):
    # the per-row form (viss_form 1), the antenna-factorized form (2; the mixed
    # form's default, 0, up to 32 antennas) with one (n_ant <= 16) or two
    # (17..32) groups of 16 antennas, its correction pass on every tile with a
    # partly visible antenna (3), and the per-row form the mixed default takes
    # above 32 antennas
    cfg = config(nside=32, n_ant=n_ant, n_epochs=12, segs=2, img_n=0, viss_form=form, precision=precision,
                 freq_hz=2e7)
    tol = MIXED_TOL if precision else (PLAIN_TOL if form == 1 else ANT_PLAIN_TOL)
    sky = _sky(12 * 32 * 32, seed=n_ant)
    errs = []
    with Pipeline(cfg) as g, Oracle(cfg) as o:
        g.set_sky(sky)
        o.set_sky(sky)
        for t0, tl in g.schedule():
            g.gen_segment(t0, tl)
            o.gen_segment(t0, tl)
            vg = g.viss_half()
            vo = o.viss_half()
            errs.append(rel_err(vg, vo))
            assert errs[-1] <= tol, (t0, errs[-1])
            hg, _ = g.conj_fill()
            ho, _ = o.conj_fill()
            assert rel_err(hg, ho) <= tol
    print(f"viss_form {form} n_ant {n_ant} precision {precision}: max rel err {max(errs):.3e}")


@pytest.mark.parametrize("variant", [dict(tile_pix=512), dict(class_budget_bytes=1), dict(seg_tlen=[5, 13, 3])],
                         ids=["tp512", "one_tile_per_pass", "uneven_segments"])
def test_viss_factorized_shapes(variant):
    # the factorized form (FP64 and mixed) on 512-pixel tiles, with one tile
    # per pass (partials folded across passes), and on uneven segments whose
    # epoch counts leave workgroups with idle waves
    kw = dict(nside=32, n_ant=12, img_n=0, freq_hz=2e7)
    if "seg_tlen" not in variant:
        kw.update(n_epochs=12, segs=2)
    kw.update(variant)
    for precision, form, tol in ((0, 2, ANT_PLAIN_TOL), (1, 0, MIXED_TOL)):
        cfg = config(viss_form=form, precision=precision, **kw)
        sky = _sky(12 * 32 * 32, seed=21)
        with Pipeline(cfg) as g, Oracle(cfg) as o:
            g.set_sky(sky)
            o.set_sky(sky)
            for t0, tl in g.schedule():
                g.gen_segment(t0, tl)
                o.gen_segment(t0, tl)
                err = rel_err(g.viss_half(), o.viss_half())
                assert err <= tol, (variant, precision, t0, err)


def test_mixed_lunar_factorized():
    # the reference lunar formation (8 antennas, positions ~2,000 km from the
    # Moon's centre, global occultation cone) in the mixed form, whose Viss is
    # the antenna-factorized kernel with 8 of 16 antenna rows in use
    cfg = config(geometry=0, nside=8, freq_hz=1e7, segs=20, img_n=32, precision=1)
    assert cfg.viss_form == 0
    errs = []

    def chk(g, o, k, d):
        errs.append(rel_err(d["vg"], d["vo"]))

    (Cg, Wg, Ig, WIg), (Co, Wo, Io, WIo) = run_both(cfg, _sky(12 * 8 * 8, seed=6), chk, segments={0, 9, 19})
    assert max(errs) <= MIXED_TOL, errs
    assert np.array_equal(Wg, Wo)
    assert rel_err(Cg, Co) <= MIXED_TOL and rel_err(Ig, Io) <= MIXED_TOL


@pytest.mark.parametrize(
    "variant",
    [dict(), dict(tile_pix=512), dict(class_budget_bytes=7000),
     dict(geometry=0, nside=8, freq_hz=1e6, segs=20, img_n=0)],
    ids=["c1", "tp512", "tile_passes", "lunar"],
)
def test_viss_adjoint_vs_oracle(variant):
    """the adjoint of the half-visibility sum (vw_viss_adjoint) against the
    oracle's within 1e-9, and the identity <V(B), r> = <B, g(r)> with the
    device's own forward sum"""
    kw = dict(C1)
    kw.update(variant)
    geom = kw.pop("geometry", 1)
    cfg = config(geometry=geom, **kw)
    rng = np.random.default_rng(23)
    with Pipeline(cfg) as g, Oracle(cfg) as o:
        B = rng.lognormal(0.0, 1.0, g.plan.npix)
        g.set_sky(B)
        for t0, tl in g.schedule()[:2]:
            g.gen_segment(t0, tl)
            o.gen_segment(t0, tl)
            r = rng.normal(size=g.nh) + 1j * rng.normal(size=g.nh)
            gg, go = g.viss_adjoint(r), o.viss_adjoint(r)
            assert rel_err(gg, go) <= PLAIN_TOL, rel_err(gg, go)
            V = g.viss_half()
            lhs = float(np.sum(V.real * r.real + V.imag * r.imag))
            assert abs(lhs - float(np.dot(B, gg))) <= 1e-12 * float(np.sum(np.abs(V) * np.abs(r)))
