"""Batched skies (nsky > 1) through the C-ABI, against the CPU FP64 oracle
and against the one-sky plain FP64 path on the GPU (needs a GPU).

The batch shares geometry, schedule and groups, so the weight maps (which
depend only on the geometry) must equal the one-sky weights bit for bit;
visibilities and maps of every sky are held to the mixed-form bar
max|x - ref| / max|ref| <= 1e-4 (MIXED_TOL).
"""
import numpy as np
import pytest

from _lockstep import MIXED_TOL, config, rel_err
from oracle.oracle import Oracle
from viss_optimize_ws_amd import Pipeline, VissError, make_sky

pytestmark = pytest.mark.gpu


def _skies(nsky, npix, seed):
    """alternating log-normal and log-normal + point-source skies"""
    return np.stack([make_sky(seed + s, npix, s % 2) for s in range(nsky)])


def test_batch_lockstep_oracle():
    """every stage of every segment, per sky, against one oracle per sky"""
    nsky = 20  # two 16-sky blocks per wave, 12 padding skies
    kw = dict(nside=16, n_ant=8, n_epochs=16, segs=2, img_n=48)
    cfg = config(precision=1, nsky=nsky, **kw)
    npix = 12 * 16 * 16
    skies = _skies(nsky, npix, 100)
    oracles = [Oracle(config(**kw)) for _ in range(nsky)]
    try:
        with Pipeline(cfg) as g:
            g.set_sky(skies)
            for s, o in enumerate(oracles):
                o.set_sky(skies[s])
            for t0, tl in g.schedule():
                g.gen_segment(t0, tl)
                vg = g.viss_half()
                hg, fg = g.conj_fill()
                ig = g.recon_segment()
                assert vg.shape == (nsky, g.nh) and fg.shape == (nsky, 2 * g.nh)
                for s, o in enumerate(oracles):
                    o.gen_segment(t0, tl)
                    vo = o.viss_half()
                    ho, fo = o.conj_fill()
                    io = o.recon_segment()
                    assert rel_err(vg[s], vo) <= MIXED_TOL, (s, rel_err(vg[s], vo))
                    assert rel_err(hg[s], ho) <= MIXED_TOL
                    assert rel_err(fg[s], fo) <= MIXED_TOL
                    assert ig["nuniq"] == io["nuniq"] and ig["nkeys"] == io["nkeys"]
            Cg, Wg, Ig, WIg = g.finalize()
        assert Cg.shape == (nsky, npix) and Ig.shape == (nsky, 48, 48)
        for s, o in enumerate(oracles):
            Co, Wo, Io, WIo = o.finalize()
            assert np.array_equal(Wg, Wo) and np.array_equal(WIg, WIo)
            assert rel_err(Cg[s], Co) <= MIXED_TOL, (s, rel_err(Cg[s], Co))
            assert rel_err(Ig[s], Io) <= MIXED_TOL, (s, rel_err(Ig[s], Io))
    finally:
        for o in oracles:
            o.close()


CASES = {
    # 64 skies: four 16-sky blocks per wave
    "nsky64": dict(nside=32, n_ant=16, n_epochs=32, segs=2, img_n=64, nsky=64),
    # three blocks per wave, full uv grid, first-member rule
    "full_occ1": dict(nside=32, n_ant=8, n_epochs=24, segs=3, img_n=40, nsky=33, recon_mode=1, occ_mode=1),
    # blockage 0: group means, every group counts, W_total normalisation
    "blockage0": dict(nside=16, n_ant=8, n_epochs=16, segs=2, img_n=32, nsky=17, blockage=0),
    # 512-pixel tiles and uneven explicit segments
    "tp512_uneven": dict(nside=32, n_ant=12, n_epochs=40, img_n=48, nsky=40, tile_pix=512, seg_tlen=[5, 19, 16]),
    # a small class budget: several tile passes per stage, one block per wave
    "tile_passes": dict(nside=32, n_ant=8, n_epochs=16, segs=2, img_n=48, nsky=16, class_budget_bytes=64 * 1024),
    # a segment of 8,400 rows: grouping by the batched contexts' own radix sort
    # (equal weights mean equal groups)
    "long_segment": dict(nside=16, n_ant=16, n_epochs=72, img_n=32, nsky=20, seg_tlen=[70, 2]),
    # 70 skies: two sky groups of four 16-sky blocks (128 padded columns)
    "two_sky_groups": dict(nside=16, n_ant=8, n_epochs=16, segs=2, img_n=32, nsky=70),
    # the reference's lunar formation (global occlusion threshold), 1 MHz band
    "lunar": dict(geometry=0, nside=8, freq_hz=1e6, segs=10, img_n=16, nsky=18),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_batch_vs_one_sky_plain(case):
    kw = dict(CASES[case])
    nsky = kw.pop("nsky")
    npix = 12 * kw["nside"] ** 2
    if kw.get("geometry") == 0:
        with Pipeline(config(**kw)) as p:  # the lunar plan's run is long: compare its first two segments
            kw = dict(kw, segs=0, seg_tlen=[t for _, t in p.schedule()[:2]])
    skies = _skies(nsky, npix, 7)
    with Pipeline(config(precision=1, nsky=nsky, **kw)) as g:
        g.set_sky(skies)
        t0, tl = g.schedule()[0]
        g.gen_segment(t0, tl)
        v0 = g.viss_half()
        g.reset_accumulators()
        g.run_all()
        Cg, Wg, Ig, WIg = g.finalize()
    # the one-sky plain FP64 path on a subset of the skies (first, last, two inside)
    for s in sorted({0, 1, nsky // 2, nsky - 1}):
        with Pipeline(config(**kw)) as p:
            p.set_sky(skies[s])
            p.gen_segment(t0, tl)
            vp = p.viss_half()
            assert rel_err(v0[s], vp) <= MIXED_TOL, (case, s, rel_err(v0[s], vp))
            p.reset_accumulators()
            p.run_all()
            Cp, Wp, Ip, WIp = p.finalize()
        assert np.array_equal(Wg, Wp) and np.array_equal(WIg, WIp), case
        assert rel_err(Cg[s], Cp) <= MIXED_TOL, (case, s, rel_err(Cg[s], Cp))
        assert rel_err(Ig[s], Ip) <= MIXED_TOL, (case, s, rel_err(Ig[s], Ip))


def test_batch_linearity_and_permutation():
    """the batch is linear in the skies and each sky's result does not depend
    on its position in the batch or on its neighbours"""
    kw = dict(nside=16, n_ant=8, n_epochs=16, segs=2, img_n=32, precision=1, nsky=18)
    npix = 12 * 16 * 16
    skies = _skies(18, npix, 55)
    mix = skies.copy()
    mix[17] = 2.0 * skies[3] + 0.5 * skies[9]
    perm = np.arange(18)[::-1].copy()
    out = {}
    for name, S in (("base", skies), ("mix", mix), ("perm", skies[perm])):
        with Pipeline(config(**kw)) as g:
            g.set_sky(S)
            g.run_all()
            out[name] = g.finalize()
    C, _, I, _ = out["base"]
    Cm, _, Im, _ = out["mix"]
    Cq, _, Iq, _ = out["perm"]
    assert rel_err(Cm[17], 2.0 * C[3] + 0.5 * C[9]) <= MIXED_TOL
    assert rel_err(Im[17], 2.0 * I[3] + 0.5 * I[9]) <= MIXED_TOL
    assert np.array_equal(Cm[:17], C[:17])  # other skies are untouched bit for bit
    assert rel_err(Cq, C[perm]) <= MIXED_TOL and rel_err(Iq, I[perm]) <= MIXED_TOL


def test_batch_sky_errors():
    kw = dict(nside=8, n_ant=4, n_epochs=4, segs=1, img_n=0, precision=1, nsky=3)
    with Pipeline(config(**kw)) as g:
        sk = np.ones((2, 12 * 64))
        with pytest.raises(VissError):
            g.set_sky(sk)  # two skies for a three-sky context
        sk = np.ones((3, 12 * 64))
        sk[1, 5] = np.nan
        with pytest.raises(VissError):
            g.set_sky(sk)
        with pytest.raises(VissError):
            g.viss_counts()


TARGETS = (2048, 4096, 8192, 16384, 32768, 65536)


def test_batch_autotune_and_launch_targets():
    """the autotuner picks one of the launch targets and clears the
    accumulators; every launch shape gives the same maps to the mixed bar
    (only the chunking of the FP64 sums changes) and the same weights"""
    kw = dict(nside=32, n_ant=16, n_epochs=32, segs=2, img_n=48, precision=1, nsky=32)
    skies = _skies(32, 12 * 32 * 32, 11)
    with Pipeline(config(**kw)) as g:
        g.set_sky(skies)
        assert g.batch_targets() == (8192, 16384)
        g.run_all()
        C0, W0, I0, WI0 = g.finalize()
        v, r = g.autotune(reps=2)
        assert v in TARGETS and r in TARGETS and g.batch_targets() == (v, r)
        g.run_all()  # the autotuner left cleared accumulators
        C1, W1, I1, WI1 = g.finalize()
        for t in (2048, 65536):
            g.set_batch_targets(t, t)
            g.reset_accumulators()
            g.run_all()
            C2, W2, I2, WI2 = g.finalize()
            assert np.array_equal(W2, W0) and np.array_equal(WI2, WI0)
            assert rel_err(C2, C0) <= MIXED_TOL and rel_err(I2, I0) <= MIXED_TOL
        assert np.array_equal(W1, W0) and np.array_equal(WI1, WI0)
        assert rel_err(C1, C0) <= MIXED_TOL and rel_err(I1, I0) <= MIXED_TOL
        with pytest.raises(VissError):
            g.set_batch_targets(1000, 2048)
    with Pipeline(config(nside=8, n_ant=4, n_epochs=4, segs=1, img_n=0)) as one:
        with pytest.raises(VissError):
            one.autotune()


def test_batch_graph_replays():
    """a captured step replayed three times, with eager steps on another sky
    in between, equals the eager step bit for bit every time; a change of
    launch targets recaptures"""
    kw = dict(nside=16, n_ant=16, n_epochs=72, seg_tlen=[70, 2], img_n=32, precision=1, nsky=20)
    npix = 12 * 16 * 16
    A = _skies(20, npix, 61)
    B = _skies(20, npix, 62)
    with Pipeline(config(**kw)) as g:
        g.set_sky(B)
        g.reset_accumulators()
        g.run_all()
        refB = g.finalize()
        g.set_sky(A)
        g.reset_accumulators()
        g.run_all()
        refA = g.finalize()
        for _ in range(3):
            g.set_sky(A)
            g.run_all_graph()
            out = g.finalize()
            for a, b in zip(out, refA):
                assert np.array_equal(a, b)
            g.set_sky(B)
            g.reset_accumulators()
            g.run_all()
            for a, b in zip(g.finalize(), refB):
                assert np.array_equal(a, b)
        g.set_batch_targets(4096, 8192)
        g.set_sky(A)
        g.run_all_graph()
        out = g.finalize()
        assert np.array_equal(out[1], refA[1]) and rel_err(out[0], refA[0]) <= MIXED_TOL
    with Pipeline(config(nside=8, n_ant=4, n_epochs=4, segs=1, img_n=0)) as one:
        one.set_sky(np.ones(12 * 64))
        with pytest.raises(VissError):
            one.run_all_graph()


# per-sky schedules: (context schedule, one schedule per sky); two skies
# share their first segment, two share every segment
SKY_SCHED_CASES = {
    "small": (dict(nside=16, n_ant=8, n_epochs=40, img_n=32), [20, 20],
              [[20, 20], [7, 13, 20], [20, 20], [1, 19, 20], [13, 7, 12, 8], [9, 9, 9, 9, 4], [7, 20, 13]]),
    # segments above 8,192 rows (16 antennas x 70 epochs): the own radix sort
    "long_rows": (dict(nside=8, n_ant=16, n_epochs=72, img_n=16), [70, 2], [[70, 2], [2, 70], [36, 36], [70, 2]]),
}


@pytest.mark.parametrize("case", sorted(SKY_SCHED_CASES))
def test_batch_per_sky_schedules(case):
    """every sky on its own segment schedule: the maps and the per-sky weights
    of the batched context equal the one-sky plain FP64 path run on that
    sky's schedule (weights bit-exact, maps within 1e-4), from eager and
    graph steps"""
    kw, base, scheds = SKY_SCHED_CASES[case]
    nsky = len(scheds)
    npix = 12 * kw["nside"] ** 2
    skies = _skies(nsky, npix, 71)
    with Pipeline(config(precision=1, nsky=nsky, seg_tlen=base, **kw)) as g:
        g.set_sky(skies)
        g.set_sky_schedules(scheds)
        g.run_all()
        Cg, Wg, Ig, WIg = g.finalize()
        assert Wg.shape == (nsky, npix) and WIg.shape == (nsky, kw["img_n"], kw["img_n"])
        g.run_all_graph()
        for a, b in zip(g.finalize(), (Cg, Wg, Ig, WIg)):
            assert np.array_equal(a, b)
        g.set_sky_schedules(None)  # back to the shared schedule
        g.reset_accumulators()
        g.run_all()
        _, Wshared, _, _ = g.finalize()
    for s in range(nsky):
        with Pipeline(config(seg_tlen=scheds[s], **kw)) as p:
            p.set_sky(skies[s])
            p.run_all()
            Cp, Wp, Ip, WIp = p.finalize()
        assert np.array_equal(Wg[s], Wp) and np.array_equal(WIg[s], WIp), (case, s)
        assert rel_err(Cg[s], Cp) <= MIXED_TOL, (case, s, rel_err(Cg[s], Cp))
        assert rel_err(Ig[s], Ip) <= MIXED_TOL, (case, s, rel_err(Ig[s], Ip))
        if scheds[s] == base:
            assert np.array_equal(Wshared, Wp)


def test_batch_per_sky_schedule_errors():
    kw = dict(nside=8, n_ant=4, n_epochs=12, seg_tlen=[6, 6], img_n=8, precision=1, nsky=2)
    with Pipeline(config(**kw)) as g:
        g.set_sky(np.ones((2, 12 * 64)))
        for bad, code in (([[6, 6], [12]], 4),       # a segment longer than the context's longest
                          ([[6, 6], [6, 5]], 1),       # does not cover the run
                          ([[6, 6], [6, 6, 1]], 1),    # beyond the run
                          ([[6, 6], [0, 6, 6]], 4)):   # empty segment
            with pytest.raises(VissError) as e:
                g.set_sky_schedules(bad)
            assert e.value.code == code, bad
        g.set_sky_schedules([[6, 6], [3, 3, 6]])
        g.gen_segment(0, 6)
        g.viss_half()
        g.conj_fill()
        with pytest.raises(VissError):
            g.recon_segment()  # per-sky recon runs in vw_run_all
        with pytest.raises(VissError):
            g.accumulators()
    with Pipeline(config(nside=8, n_ant=4, n_epochs=4, segs=1, img_n=0)) as one:
        with pytest.raises(VissError):
            one.set_sky_schedules([[4]])
    with Pipeline(config(nside=8, n_ant=4, n_epochs=12, seg_tlen=[6, 6], img_n=8, precision=1, nsky=2,
                         blockage=0)) as b0:
        with pytest.raises(VissError):
            b0.set_sky_schedules([[6, 6], [6, 6]])
