"""Throughput of the visibility + reconstruction path on one MI355X.

Workload (one "step"): the plain FP64 form over the whole segment schedule of
HEALPix NESTED Nside=128 (196,608 pixels, 768 static 256-pixel tiles),
16 antennas on seeded circular Earth orbits (120 baselines), 512 epochs 10 s
apart, 30 MHz, 16 equal segments, tile-cone culling, blockage=1, occ_mode=2,
uv-mirror reconstruction into the HEALPix map and a 512x512 lm image.  A step
is: clear accumulators, for every segment generate geometry, Viss, phase
correction + conjugate fill, grouping and reconstruction, then normalise on
the device.  Inputs (sky, resident directions) are in HBM before timing.

Metric: pixel-baseline Gterms/s = (sum_seg npix*segN_half +
sum_seg (npix + n_img)*nkeys_seg) / t / 1e9, counting culled terms too.
The mixed form (antenna-factorized Viss on the FP32 matrix cores with FP64
antenna phases; recon with FP64 phase, FP32 hardware sincos, FP32 block sums
folded in FP64) is timed on the same workload and its error is stated
against the plain FP64 result (which matches the CPU FP64 oracle to 1e-9).
So is the FP64 antenna-factorized Viss option (viss_form 2, "plain_ant"),
whose phases differ from the per-row ones by FP64 rounding.

The batched leg ("batched") runs 64 sky realisations (log-normal(0,1)
diffuse field plus Poisson(100) point sources, capped at 200, with
Pareto(1.5) flux) through the same geometry on one uneven segment schedule
(lengths drawn uniformly from [16, 256] epochs): Viss and recon as BF16
hi/lo matrix-core products of on-the-fly phasor tiles against the sky /
group-visibility matrix.  Its rate counts every sky's terms; its error is
stated against the plain FP64 path for four of the skies.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pixel-baseline Gterms/s Viss+recon, plain FP64 vs mixed at stated FP64 err"
WORKLOADS = {
    # the headline configuration (default)
    "nside128": (dict(nside=128, n_ant=16, n_epochs=512, freq_hz=3e7, segs=16, img_n=512, tile_pix=256),
                 "HEALPix NESTED Nside=128 (196,608 px, 768 tiles), 16 antennas / 120 baselines on seeded "
                 "circular orbits r in [6900,7100] km, 512 epochs x 10 s, 30 MHz, 16 segments, "
                 "blockage=1 occ_mode=2 uv-mirror recon, HEALPix map + 512x512 lm image"),
    # the mixed-precision configuration: 507,904 half visibilities, 1024^2 image
    "nside256": (dict(nside=256, n_ant=32, n_epochs=1024, freq_hz=1e7, segs=16, img_n=1024, tile_pix=256),
                 "HEALPix NESTED Nside=256 (786,432 px, 3,072 tiles), 32 antennas / 496 baselines on seeded "
                 "circular orbits r in [6900,7100] km, 1,024 epochs x 10 s, 10 MHz, 16 segments, "
                 "blockage=1 occ_mode=2 uv-mirror recon, HEALPix map + 1024x1024 lm image"),
}
WORKLOAD, WORKLOAD_NAME = WORKLOADS["nside128"]
FP64_PEAK_TFLOPS = 78.6
# FP64 flops of one evaluated Viss term in the plain form as executed (fma =
# 2; the table index is not counted): phase u*l + v*m + w*(n-1) on operands
# pre-scaled by 1024, 1 mul + 2 fma = 5; split t + M - M, t - k = 3;
# z = f^2 1; sin f*(a + z*(b + z*c)) 2 fma + mul = 5; cos 2 fma = 4;
# rotation by the table entry 2 mul + 2 fma = 6; accumulation 2 fma = 4.
VISS_TERM_FLOPS_PLAIN = 28
VISS_TERM_FP64_INSTR = 18
# recon term, plain form: phase 1 mul + 1 fma = 3, then as Viss from the split
# on (3 + 1 + 5 + 4 + 6 + 4) = 26 flops, 17 FP64 instructions
RECON_TERM_FLOPS_PLAIN = 26
RECON_TERM_FP64_INSTR = 17
FP64_INSTR_PEAK_T = FP64_PEAK_TFLOPS / 2  # one FP64 lane instruction per FMA slot


TRAFFIC_FILES = {"nside128": os.path.join(ROOT, "profiles", "pmc_traffic_nside128_16ant_512ep.json"),
                 "nside256": os.path.join(ROOT, "profiles", "pmc_traffic_nside256_32ant_1024ep.json")}


def _traffic(workload, kernel):
    """HBM bytes per launch of a kernel at this workload from the committed
    FETCH_SIZE / WRITE_SIZE runs (scaled by the copy-kernel calibration), or
    None when that summary is absent."""
    try:
        with open(TRAFFIC_FILES[workload]) as f:
            return int(json.load(f)["kernels"][kernel]["bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


SERIAL = False


def _cfg(precision, viss_form=0):
    from viss_optimize_ws_amd import default_config

    return default_config(1, precision=precision, serial_stages=int(SERIAL), viss_form=viss_form, **WORKLOAD)


def _cfg_serial(precision):
    from viss_optimize_ws_amd import default_config

    return default_config(1, precision=precision, serial_stages=1, **WORKLOAD)


def _plan_counts(pipe):
    """One instrumented pass over the schedule: per-segment nkeys, evaluated
    (non-culled) Viss terms, the final maps for error statements, and the
    stage timings of the pass."""
    per_seg = []
    TP = pipe.cfg.tile_pix

    def sizes(n, nt):
        t = np.full(nt, TP, np.int64)
        t[-1] = n - TP * (nt - 1)
        return t

    tile_sizes = sizes(pipe.plan.npix, pipe.plan.ntile)
    img_sizes = sizes(pipe.plan.n_img_pts, pipe.plan.n_img_tiles) if pipe.plan.n_img_pts else None
    pipe.reset_accumulators()
    pipe.timings()  # clear
    for t0, tl in pipe.schedule():
        pipe.gen_segment(t0, tl)
        cls = pipe.tile_classes()
        evaluated = int(((cls != 2).astype(np.int64) * tile_sizes[None, :]).sum())
        vh = pipe.viss_half()
        pipe.conj_fill()
        info = pipe.recon_segment()
        # recon terms evaluated: (output tile, group) pairs whose representative
        # half row is not hidden from the tile (uv-mirror recon, occ_mode 2)
        _, keys, _, _, rep = pipe.groups()
        reps = rep[keys != 0]
        rec_eval = int(((cls[reps] != 2).astype(np.int64) * tile_sizes[None, :]).sum())
        if img_sizes is not None:
            ci = pipe.tile_classes(img=True)
            rec_eval += int(((ci[reps] != 2).astype(np.int64) * img_sizes[None, :]).sum())
        per_seg.append(dict(n_half=info["n_half"], nkeys=info["nkeys"], evaluated=evaluated, rec_evaluated=rec_eval,
                            vraw=vh))
    maps = pipe.finalize()
    return per_seg, maps, pipe.timings()


def _time_steps(pipe, steps, warmup):
    lib = pipe._lib
    for _ in range(warmup):
        pipe.reset_accumulators()
        pipe.run_all()
        lib.vw_finalize(pipe._ctx, None, None, None, None)
    pipe.timings()  # clear
    t0 = time.perf_counter()
    for _ in range(steps):
        pipe.reset_accumulators()
        pipe.run_all()
        rc = lib.vw_finalize(pipe._ctx, None, None, None, None)
        assert rc == 0
    t1 = time.perf_counter()
    return (t1 - t0) / steps, pipe.timings()


def _cpu_baseline(sky, threads, vraw0, full_segment):
    """CPU FP64 oracle on a bounded sample of the first segment.  full_segment:
    Viss of all its half rows and reconstruction of every HEALPix pixel and
    image point.  Otherwise (large workloads): Viss of the first 2,048 half
    rows, then, from the segment's GPU half visibilities, grouping and the
    reconstruction of the first 16,384 HEALPix pixels and image points."""
    from oracle.oracle import Oracle

    cfg = _cfg(0)
    with Oracle(cfg) as o:
        o.set_sky(sky)
        t0, tl = o.schedule()[0]
        o.gen_segment(t0, tl)
        if full_segment:
            s = time.perf_counter()
            o.viss_half()
            o.conj_fill()
            info = o.recon_segment()
            dt = time.perf_counter() - s
            terms = o.plan.npix * info["n_half"] + (o.plan.npix + o.plan.n_img_pts) * info["nkeys"]
            sample = f"segment 1 of 16 in full: {terms:.3e} terms in {dt:.2f} s (oracle, OpenMP FP64)"
        else:
            nr, npt = min(2048, o.nh), 16384
            s = time.perf_counter()
            o.viss_half(0, nr)
            dv = time.perf_counter() - s
            o.set_viss_raw(vraw0)
            o.conj_fill()
            info = o.group_segment()
            s = time.perf_counter()
            o.recon_points(0, np.arange(min(npt, o.plan.npix)))
            o.recon_points(1, np.arange(min(npt, o.plan.n_img_pts)))
            dr = time.perf_counter() - s
            dt = dv + dr
            terms = o.plan.npix * nr + (min(npt, o.plan.npix) + min(npt, o.plan.n_img_pts)) * info["nkeys"]
            sample = (f"segment 1 of 16: Viss of {nr} half rows ({dv:.2f} s) + recon of {npt} HEALPix pixels and "
                      f"{npt} image points over its {info['nkeys']} keys ({dr:.2f} s): {terms:.3e} terms "
                      "(oracle, OpenMP FP64)")
    return dict(value=terms / dt / 1e9, unit="Gterms/s", cores=threads, kind="port", sample=sample)


BATCH_NSKY = 64
BF16_PEAK_TFLOPS = 2500.0


def _uneven_schedule(n_epochs, seed=4242, lo=16, hi=256):
    """segment lengths drawn uniformly from [lo, hi] epochs (numpy's seeded
    PCG64 generator) until the run's epochs are covered; the last segment
    takes the remainder"""
    rng = np.random.default_rng(seed)
    out, tot = [], 0
    while tot < n_epochs:
        n = int(rng.integers(lo, hi + 1))
        n = min(n, n_epochs - tot)
        out.append(n)
        tot += n
    return out


def _batched_leg(steps, warmup, workload):
    from viss_optimize_ws_amd import Pipeline, default_config, make_sky

    W = dict(WORKLOAD)
    W.pop("segs", None)
    sched = _uneven_schedule(W["n_epochs"])
    npix = 12 * W["nside"] ** 2
    skies = np.stack([make_sky(7000 + s, npix, 1) for s in range(BATCH_NSKY)])
    cfg_b = default_config(1, precision=1, nsky=BATCH_NSKY, seg_tlen=sched, serial_stages=int(SERIAL), **W)
    cfg_p = default_config(1, precision=0, seg_tlen=sched, **W)
    # plain FP64 reference for four skies, and the per-segment key counts
    ref, nkeys = [], []
    with Pipeline(cfg_p) as p:
        nimg = p.plan.n_img_pts
        for s in range(4):
            p.set_sky(skies[s])
            p.reset_accumulators()
            for t0, tl in p.schedule():
                p.gen_segment(t0, tl)
                p.viss_half()
                p.conj_fill()
                info = p.recon_segment()
                if s == 0:
                    nkeys.append(info["nkeys"])
            ref.append(p.finalize())
    terms_sky = sum(npix * tl * (W["n_ant"] * (W["n_ant"] - 1) // 2) for tl in sched) + (npix + nimg) * sum(nkeys)
    with Pipeline(cfg_b) as b:
        b.set_sky(skies)
        targets = b.autotune(reps=2)  # launch shapes of both batched kernels
        b.run_all()
        Cb, Wb, Ib, _ = b.finalize()
        sec_eager, tim = _time_steps(b, steps, warmup)
        # the same steps from a captured graph (reset + every segment)
        for _ in range(max(warmup, 1)):
            b.run_all_graph()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run_all_graph()
            assert b._lib.vw_finalize(b._ctx, None, None, None, None) == 0
        sec_graph = (time.perf_counter() - t0) / steps
        Cg, Wg, Ig, _ = b.finalize()
        graph_equal = bool(np.array_equal(Cg, Cb) and np.array_equal(Ig, Ib) and np.array_equal(Wg, Wb))
        sec = min(sec_eager, sec_graph)
    if not SERIAL:
        with Pipeline(default_config(1, precision=1, nsky=BATCH_NSKY, seg_tlen=sched, serial_stages=1, **W)) as b2:
            b2.set_sky(skies)
            b2.set_batch_targets(*targets)
            _, tim = _time_steps(b2, steps, warmup)
    err = {}
    for name, k in (("C_healpix", 0), ("lm_image", 2)):
        got = Cb if k == 0 else Ib
        mx = [float(np.max(np.abs(got[s] - ref[s][k])) / np.max(np.abs(ref[s][k]))) for s in range(4)]
        rm = [float(np.sqrt(np.mean((got[s] - ref[s][k]) ** 2)) / np.sqrt(np.mean(ref[s][k] ** 2))) for s in range(4)]
        err[name] = {"max_rel": max(mx), "rms_rel": max(rm)}
    weights_equal = bool(all(np.array_equal(Wb, r[1]) for r in ref))
    nh_tot = sum(tl * (W["n_ant"] * (W["n_ant"] - 1) // 2) for tl in sched)
    viss_ms = tim["viss_kernel_ms"] / steps
    recon_ms = tim["recon_kernel_ms"] / steps
    # matrix-core work as formulated: every (row, pixel) and (point, group)
    # term of every sky takes 2 real multiply-adds, each as 3 BF16 products
    viss_flops = 12.0 * nh_tot * npix * BATCH_NSKY
    recon_flops = 12.0 * (npix + nimg) * sum(nkeys) * BATCH_NSKY
    return {
        "value": round(terms_sky * BATCH_NSKY / sec / 1e9, 2),
        "unit": "Gterms/s",
        "ms_per_step": round(sec * 1e3, 4),
        "ms_per_step_eager": round(sec_eager * 1e3, 4),
        "ms_per_step_graph": round(sec_graph * 1e3, 4),
        "graph_equals_eager": graph_equal,
        "nsky": BATCH_NSKY,
        "terms_per_step": terms_sky * BATCH_NSKY,
        "schedule_epochs": sched,
        "autotuned_waves_per_launch": {"viss": targets[0], "recon": targets[1]},
        "dtype": "f64 phase, f32 hardware sincos, bf16 hi/lo matrix-core products with f32 accumulation, f64 sums",
        "data": "synthetic: 64 skies log-normal(0,1) + Poisson(100) (cap 200) Pareto(1.5) point sources, seeds 7000-7063",
        "err_vs_fp64_4_skies": err,
        "weights_bitexact": weights_equal,
        "roofline_viss": {"kernel": "k_vissb", "bound": "mfma", "avg_launch_ms": round(viss_ms / len(sched), 4),
                          "achieved": round(viss_flops / (viss_ms * 1e-3) / 1e12, 2), "peak": BF16_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": round(viss_flops / (viss_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
                          "flops": "12 per (half row, pixel, sky), culled terms included"},
        "roofline_recon": {"kernel": "k_reconb", "bound": "mfma", "stage_kernel_ms_per_step": round(recon_ms, 4),
                           "achieved": round(recon_flops / (recon_ms * 1e-3) / 1e12, 2), "peak": BF16_PEAK_TFLOPS,
                           "unit": "TFLOP/s",
                           "frac": round(recon_flops / (recon_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
                           "flops": "12 per (output point, group, sky)"},
    }


def _copy_calibration(mib):
    """A plain copy kernel of known size: n bytes read and n bytes written per
    launch.  Its FETCH_SIZE / WRITE_SIZE readings give the factor that turns
    those counters into bytes for the kernels of the path."""
    import torch

    n = mib << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel, not a DMA copy
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-ant", action="store_true", help="skip the FP64 antenna-factorized Viss leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-skies leg")
    ap.add_argument("--serial", action="store_true",
                    help="run every stage on one stream (no overlap of a segment's recon with the next Viss)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="nside128",
                    help="nside128 (default, the headline configuration) or nside256")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calib:
        _copy_calibration(args.copy_calib)
        return
    global WORKLOAD, WORKLOAD_NAME, SERIAL
    WORKLOAD, WORKLOAD_NAME = WORKLOADS[args.workload]
    SERIAL = args.serial
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (every box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)

    from viss_optimize_ws_amd import Pipeline, make_sky

    sky_seed = 20240
    with Pipeline(_cfg(0)) as plain:
        npix, nimg = plain.plan.npix, plain.plan.n_img_pts
        sky = make_sky(sky_seed, npix, 0)
        plain.set_sky(sky)
        per_seg, maps_plain, tim_iso = _plan_counts(plain)
        t_viss = sum(npix * s["n_half"] for s in per_seg)
        t_recon = sum((npix + nimg) * s["nkeys"] for s in per_seg)
        terms = t_viss + t_recon
        evaluated = sum(s["evaluated"] for s in per_seg)
        sec, tim = _time_steps(plain, args.steps, args.warmup)
    if SERIAL:
        tim_ser = tim
    else:
        # kernel-level figures from the same steps with every stage on one
        # stream, where each kernel runs alone
        with Pipeline(_cfg_serial(0)) as ser:
            ser.set_sky(sky)
            _, tim_ser = _time_steps(ser, args.steps, args.warmup)

    value = terms / sec / 1e9
    nl = max(tim_ser["n_viss_launch"], 1)
    viss_avg_ms = tim_ser["viss_kernel_ms"] / nl
    eval_per_launch = evaluated / len(per_seg)
    achieved = VISS_TERM_FLOPS_PLAIN * eval_per_launch / (viss_avg_ms * 1e-3) / 1e12
    rec_eval_step = sum(s_["rec_evaluated"] for s_ in per_seg)
    # recon stage of the serial steps (the HEALPix and image sets run
    # concurrently on two streams): its wall time, kernels and class/reduce
    # launches, per step
    rec_ms = tim_ser["recon_ms"] / args.steps
    rec_achieved = RECON_TERM_FLOPS_PLAIN * rec_eval_step / (rec_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gterms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: log-normal(0,1) sky seed {sky_seed}, circular orbits seed 42",
        "config": {"workload": WORKLOAD_NAME, **WORKLOAD, "form": "plain FP64",
                   "stages": "serial" if SERIAL else "recon of segment s beside the Viss of segment s+1",
                   "terms_per_step": terms, "viss_terms": t_viss, "recon_terms": t_recon},
        "stage_ms_per_step": {k: round(v / args.steps, 4) for k, v in tim.items() if k.endswith("_ms")},
        "roofline": {
            "bound": "mfma",
            "pipe": "FP64 vector ALU (VALU-bound direct sum; published FP64 peak)",
            "kernel": "k_viss (plain)",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": _traffic(args.workload, "k_viss_plain"),
            "traffic_source": os.path.relpath(TRAFFIC_FILES[args.workload], ROOT),
            "flops_per_term": VISS_TERM_FLOPS_PLAIN,
            "evaluated_terms_per_launch": int(eval_per_launch),
            "avg_launch_ms": round(viss_avg_ms, 4),
            "timing": "HIP events around each k_viss launch of the timed steps run with serial stages (kernel alone)",
            "fp64_instr_issue_frac": round(VISS_TERM_FP64_INSTR * eval_per_launch / (viss_avg_ms * 1e-3) / 1e12
                                           / FP64_INSTR_PEAK_T, 4),
        },
        "roofline_recon": {
            "bound": "mfma",
            "pipe": "FP64 vector ALU",
            "kernel": "recon stage (k_recon plain, uv mirror, representative classes; HEALPix and image sets concurrent)",
            "achieved": round(rec_achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(rec_achieved / FP64_PEAK_TFLOPS, 4),
            "flops_per_term": RECON_TERM_FLOPS_PLAIN,
            "traffic": _traffic(args.workload, "k_recon_plain"),
            "evaluated_terms_per_step": rec_eval_step,
            "stage_ms_per_step": round(rec_ms, 4),
            "fp64_instr_issue_frac": round(RECON_TERM_FP64_INSTR * rec_eval_step / (rec_ms * 1e-3) / 1e12
                                           / FP64_INSTR_PEAK_T, 4),
        },
    }
    if not args.no_mixed:
        with Pipeline(_cfg(1)) as mixed:
            mixed.set_sky(sky)
            per_seg_m, maps_m, tim_m_iso = _plan_counts(mixed)
            sec_m, tim_m = _time_steps(mixed, args.steps, args.warmup)
        from_err = {}
        vr_p = np.concatenate([s["vraw"] for s in per_seg])
        vr_m = np.concatenate([s["vraw"] for s in per_seg_m])
        for name, a, b in (("viss_half", vr_m, vr_p), ("C_healpix", maps_m[0], maps_plain[0]),
                           ("lm_image", maps_m[2], maps_plain[2])):
            den = np.max(np.abs(b))
            from_err[name] = {"max_rel": float(np.max(np.abs(a - b)) / den),
                              "rms_rel": float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / np.sqrt(np.mean(np.abs(b) ** 2)))}
        out["mixed"] = {"value": round(terms / sec_m / 1e9, 4), "unit": "Gterms/s", "ms_per_step": round(sec_m * 1e3, 4),
                        "speedup_vs_plain": round(sec / sec_m, 3),
                        "dtype": "f64 phase + f32 sincos/sum (Viss: f32 matrix-core products of antenna phasors)",
                        "err_vs_fp64": from_err,
                        "viss_kernel_avg_ms": round(tim_m_iso["viss_kernel_ms"] / max(tim_m_iso["n_viss_launch"], 1), 4)}
    if not args.no_ant:
        # FP64 throughout, Viss by the antenna-factorized form on the FP64 matrix cores
        with Pipeline(_cfg(0, viss_form=2)) as ant:
            ant.set_sky(sky)
            per_seg_a, _, tim_a_iso = _plan_counts(ant)
            sec_a, _ = _time_steps(ant, args.steps, args.warmup)
        vr_a = np.concatenate([s["vraw"] for s in per_seg_a])
        vr_p = np.concatenate([s["vraw"] for s in per_seg])
        out["plain_ant"] = {"value": round(terms / sec_a / 1e9, 4), "unit": "Gterms/s",
                            "ms_per_step": round(sec_a * 1e3, 4), "dtype": "f64",
                            "viss_kernel_avg_ms": round(tim_a_iso["viss_kernel_ms"] / max(tim_a_iso["n_viss_launch"], 1), 4),
                            "viss_max_rel_err_vs_per_row": float(np.max(np.abs(vr_a - vr_p)) / np.max(np.abs(vr_p)))}
    if not args.no_batch:
        out["batched"] = _batched_leg(args.steps, args.warmup, args.workload)
        out["batched"]["speedup_vs_plain"] = round(out["batched"]["value"] / value, 3)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = _cpu_baseline(sky, threads, per_seg[0]["vraw"], args.workload == "nside128")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
