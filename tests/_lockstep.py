"""Lockstep driver: the gfx950 Pipeline and the CPU oracle on the same inputs.

Tolerances (plain FP64 form, stated per output):
  * masks, tile classes, counts, weights, keys, groups: bit-exact;
  * lmn, tile cones, positions, u/v/w, fa/fb/du: bit-exact (shared FP64
    formulas, correctly rounded operations only);
  * Viss and images: max|x - ref| <= 1e-9 * max|ref| (sums are reordered on
    the device, so last-bit agreement is not expected).
Mixed form: max|x - ref| / max|ref| <= 1e-4 per output.
"""
import numpy as np

from oracle.oracle import Oracle
from viss_optimize_ws_amd import Pipeline, default_config

PLAIN_TOL = 1e-9
MIXED_TOL = 1e-4


def rel_err(x, ref):
    den = np.max(np.abs(ref))
    if den == 0:
        return float(np.max(np.abs(x)))
    return float(np.max(np.abs(np.asarray(x) - np.asarray(ref))) / den)


def rms_rel_err(x, ref):
    den = np.sqrt(np.mean(np.abs(ref) ** 2))
    return float(np.sqrt(np.mean(np.abs(np.asarray(x) - np.asarray(ref)) ** 2)) / den) if den else 0.0


def run_both(cfg, sky, check_seg=None, segments=None):
    """Run every segment on both sides; check_seg(g, o, k) is called after each
    segment's recon.  Returns (gpu finalize, oracle finalize)."""
    with Pipeline(cfg) as g, Oracle(cfg) as o:
        g.set_sky(sky)
        o.set_sky(sky)
        sched = g.schedule()
        assert sched == o.schedule()
        for k, (t0, tl) in enumerate(sched):
            if segments is not None and k not in segments:
                continue
            g.gen_segment(t0, tl)
            o.gen_segment(t0, tl)
            vg = g.viss_half()
            vo = o.viss_half()
            hg, fg = g.conj_fill()
            ho, fo = o.conj_fill()
            ig = g.recon_segment()
            io = o.recon_segment()
            if check_seg:
                check_seg(g, o, k, dict(vg=vg, vo=vo, hg=hg, ho=ho, fg=fg, fo=fo, ig=ig, io=io))
        return g.finalize(), o.finalize()


def config(**kw):
    geom = kw.pop("geometry", 1)
    return default_config(geom, **kw)
