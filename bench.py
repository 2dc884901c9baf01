"""Throughput of the visibility + reconstruction path on one MI355X.

One "step" = one pass of the path over a workload's whole segment schedule:
clear the accumulators, then for every segment generate the geometry, sum the
half visibilities, phase-correct + conjugate-complete them, group the rows on
the uv grid and reconstruct the HEALPix map and the lm image, then normalise
on the device.  Inputs (sky, resident directions) are in HBM before timing.

Metric: pixel-baseline Gterms/s = (sum_seg npix*segN_half +
sum_seg (npix + n_img)*nkeys_seg) / t / 1e9, counting culled terms too.

Headline (`value`): the plain FP64 form on HEALPix NESTED Nside=128 (196,608
pixels, 768 static 256-pixel tiles), 16 antennas on seeded circular Earth
orbits (120 baselines), 512 epochs 10 s apart, 30 MHz, 16 equal segments,
tile-cone culling, blockage=1, occ_mode=2, uv-mirror reconstruction into the
HEALPix map and a 512x512 lm image, with the CPU FP64 oracle timed beside it
on a bounded sample.  Beside it on the same line:
  * `mixed`: the mixed form at the same workload (antenna-factorized Viss on
    the FP32 matrix cores with FP64 antenna phases; recon with FP64 phase,
    FP32 hardware sin/cos, FP32 block sums folded in FP64), its error stated
    per output against the plain FP64 result (which matches the oracle to
    1e-9 in the GPU tests);
  * `nside256`: the plain and mixed forms at Nside=256 (786,432 pixels), 32
    antennas (496 baselines), 1,024 epochs, 10 MHz, 16 segments, 1024x1024
    image, and 64 sky realisations through the batched form;
  * `batched`: 64 sky realisations (log-normal(0,1) diffuse field plus
    Poisson(100) point sources, capped at 200, Pareto(1.5) flux) at Nside=128
    on an uneven segment schedule (lengths uniform in [16, 256] epochs):
    Viss and recon as BF16 hi/lo matrix-core products of on-the-fly phasor
    tiles against the sky / group-visibility matrix.
Rooflines: kernel averages come from steps run with every stage on one stream
(each kernel alone), HIP events recorded on the library's stream around each
launch; `--legs` and `--serial` reproduce one leg for a rocprofv3 run.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pixel-baseline Gterms/s Viss+recon, plain FP64 vs mixed at stated FP64 err"
WORKLOADS = {
    # the headline configuration
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
LEGS = ("nside128", "nside256", "batched128", "batched256", "skysched128", "cpu")
SKY_SEED = 20240
BATCH_NSKY = 64

# published MI355X peaks
FP64_PEAK_TFLOPS = 78.6    # vector FP64: 16 lanes per SIMD, one FMA per lane per cycle
FP32_PEAK_TFLOPS = 157.3   # vector FP32: 32 lanes per SIMD (a wave64 FP32 FMA issues in 2 cycles)
BF16_PEAK_TFLOPS = 2500.0  # dense BF16 matrix cores
FP64_INSTR_PEAK_T = FP64_PEAK_TFLOPS / 2  # FP64 lane-instructions per second (1e12)

# FP64 flops of one evaluated term in the plain forms as executed (fma = 2,
# table index not counted).  Viss: phase u*l + v*m + w*(n-1) on operands
# pre-scaled by 1024, 1 mul + 2 fma = 5; split t + M - M, t - k = 3; z = f^2
# 1; the pixel weight folded into the coefficients: w cos = w0 + z*w2, 1 fma
# = 2, w sin = f*(w1 + z*w3), fma + mul = 3; rotation by the table entry
# accumulated directly, 4 fma = 8: 22 flops, 14 instructions.  Recon: phase
# 1 mul + 1 fma = 3; split 3; z 1; sin fma + mul = 3; cos 1 fma = 2;
# rotation 2 mul + 2 fma = 6; accumulation 2 fma = 4: 22 flops, 15
# instructions.
VISS_TERM_FLOPS_PLAIN, VISS_TERM_INSTR_PLAIN = 22, 14
RECON_TERM_FLOPS_PLAIN, RECON_TERM_INSTR_PLAIN = 22, 15
# instructions of one evaluated term of the mixed recon as written
# (ReconAccF32::term): phase mul_f64 + fma_f64, fract_f64, cvt_f32_f64,
# sin_f32, cos_f32, 2 fma_f32; weighted by their measured issue costs
MIXED_RECON_TERM = {"mul_f64": 1, "fma_f64": 1, "fract_f64": 1, "cvt_f32_f64": 1, "sin_f32": 1, "cos_f32": 1,
                    "fma_f32": 2}
BATCH_TERM_FLOPS = 12  # per (term, sky): 2 real multiply-adds, each 3 BF16 products (hi*hi, hi*lo, lo*hi)

TRAFFIC_FILES = {"nside128": os.path.join(ROOT, "profiles", "pmc_traffic_nside128_16ant_512ep.json"),
                 "nside256": os.path.join(ROOT, "profiles", "pmc_traffic_nside256_32ant_1024ep.json")}


T_START = time.perf_counter()


def _log(msg):
    """progress on stderr (the JSON line alone goes to stdout)"""
    print(f"[bench {time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def _traffic(workload, kernel):
    """HBM bytes per launch of a kernel at this workload from the committed
    FETCH_SIZE / WRITE_SIZE runs (scaled by the copy-kernel calibration), or
    None when that summary is absent."""
    try:
        with open(TRAFFIC_FILES[workload]) as f:
            return int(json.load(f)["kernels"][kernel]["bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def _cfg(workload, serial, **kw):
    from viss_optimize_ws_amd import default_config

    W = dict(WORKLOADS[workload][0])
    if "seg_tlen" in kw:
        W.pop("segs")
    W.update(kw)
    return default_config(1, serial_stages=int(serial), **W)


def _tile_sizes(n, nt, TP):
    t = np.full(nt, TP, np.int64)
    t[-1] = n - TP * (nt - 1)
    return t


def _plan_counts(pipe, classes=True):
    """One instrumented pass over the schedule: per-segment sizes, raw half
    visibilities and (classes=True) the evaluated terms: Viss (row, pixel)
    pairs whose (row, tile) class is not hidden, recon (point, group) pairs
    whose representative row is not hidden from the point's tile (uv-mirror,
    occ_mode 2).  Returns (per-segment dicts, final maps)."""
    TP = pipe.cfg.tile_pix
    tsz = _tile_sizes(pipe.plan.npix, pipe.plan.ntile, TP)
    isz = _tile_sizes(pipe.plan.n_img_pts, pipe.plan.n_img_tiles, TP) if pipe.plan.n_img_pts else None
    per_seg = []
    pipe.reset_accumulators()
    for t0, tl in pipe.schedule():
        pipe.gen_segment(t0, tl)
        d = {}
        if classes:
            cls = pipe.tile_classes()
            d["evaluated"] = int(np.dot((cls != 2).sum(axis=0, dtype=np.int64), tsz))
        d["vraw"] = pipe.viss_half()
        pipe.conj_fill()
        info = pipe.recon_segment()
        d.update(n_half=info["n_half"], nkeys=info["nkeys"])
        if classes:
            _, keys, _, _, rep = pipe.groups()
            reps = rep[keys != 0]
            rec = int(np.dot((cls[reps] != 2).sum(axis=0, dtype=np.int64), tsz))
            del cls
            if isz is not None:
                ci = pipe.tile_classes(img=True)
                rec += int(np.dot((ci[reps] != 2).sum(axis=0, dtype=np.int64), isz))
                del ci
            d["rec_evaluated"] = rec
        per_seg.append(d)
    return per_seg, pipe.finalize()


def _time_steps(pipe, steps, warmup, graph=False):
    lib = pipe._lib
    run = pipe.run_all_graph if graph else pipe.run_all
    for _ in range(warmup):
        if not graph:
            pipe.reset_accumulators()
        run()
        assert lib.vw_finalize(pipe._ctx, None, None, None, None) == 0
    pipe.timings()  # clear
    t0 = time.perf_counter()
    for _ in range(steps):
        if not graph:
            pipe.reset_accumulators()
        run()
        assert lib.vw_finalize(pipe._ctx, None, None, None, None) == 0
    t1 = time.perf_counter()
    return (t1 - t0) / steps, pipe.timings()


def _err(a, b):
    den = np.max(np.abs(b))
    return {"max_rel": float(np.max(np.abs(a - b)) / den),
            "rms_rel": float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / np.sqrt(np.mean(np.abs(b) ** 2)))}


def _kernel_ms(tim, steps):
    """per-step kernel time of a serial-stage run (one stream: sums are wall time)"""
    return {"viss_kernel": round(tim["viss_kernel_ms"] / steps, 4),
            "recon_kernel": round(tim["recon_kernel_ms"] / steps, 4),
            "gen": round(tim["gen_ms"] / steps, 4), "viss_stage": round(tim["viss_ms"] / steps, 4),
            "conj": round(tim["conj_ms"] / steps, 4), "recon_stage": round(tim["recon_ms"] / steps, 4),
            "finalize": round(tim["finalize_ms"] / steps, 4)}


def _roof(kernel, bound, pipe_desc, work_per_launch, avg_ms, peak, unit, basis, traffic=None, **extra):
    achieved = work_per_launch / (avg_ms * 1e-3) / 1e12
    d = {"kernel": kernel, "bound": bound, "pipe": pipe_desc, "achieved": round(achieved, 3), "peak": peak,
         "unit": unit, "frac": round(achieved / peak, 4), "traffic": traffic, "avg_launch_ms": round(avg_ms, 5),
         "work_per_launch": work_per_launch, "basis": basis}
    d.update(extra)
    return d


def _slots_per_term(rates, term):
    s = rates["slots"]
    cost = {"mul_f64": s["mul_f64"], "fma_f64": s["fma_f64"], "fract_f64": s["fract_f64"],
            "cvt_f32_f64": s["cvt_f32_f64_pair"] / 2, "sin_f32": s["sin_f32"], "cos_f32": s["cos_f32"],
            "fma_f32": s["fma_f32"]}
    return sum(cost[k] * n for k, n in term.items())


def issue_rates():
    """issue costs of the VALU instruction classes, measured now on this GPU
    by viss_optimize_ws_amd/bin/vw_issue_rates (one JSON line)"""
    exe = os.path.join(ROOT, "viss_optimize_ws_amd", "bin", "vw_issue_rates")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built; run __graft_entry__.build()")
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
    return json.loads(out.strip().splitlines()[-1])


def one_sky_leg(workload, steps, warmup, serial, rates, mixed=True, plain_ant=False):
    """plain (and mixed) form on one sky: rates of the step, kernel figures
    from serial-stage steps, rooflines of the dominant kernels, errors of the
    mixed form against the plain FP64 result per output"""
    from viss_optimize_ws_amd import Pipeline, make_sky

    W, name = WORKLOADS[workload]
    out = {"workload": name, **W}
    with Pipeline(_cfg(workload, serial)) as plain:
        npix, nimg = plain.plan.npix, plain.plan.n_img_pts
        sky = make_sky(SKY_SEED, npix, 0)
        plain.set_sky(sky)
        per_seg, maps_p = _plan_counts(plain)
        sec, _ = _time_steps(plain, steps, warmup)
    t_viss = sum(npix * s["n_half"] for s in per_seg)
    t_recon = sum((npix + nimg) * s["nkeys"] for s in per_seg)
    terms = t_viss + t_recon
    with Pipeline(_cfg(workload, True)) as ser:
        ser.set_sky(sky)
        _, tim = _time_steps(ser, steps, warmup)
    nv, nr = tim["n_viss_launch"] // steps, tim["n_recon_launch"] // steps
    viss_avg = tim["viss_kernel_ms"] / (nv * steps)
    recon_avg = tim["recon_kernel_ms"] / (nr * steps)
    ev_viss = sum(s["evaluated"] for s in per_seg) / nv
    ev_recon = sum(s["rec_evaluated"] for s in per_seg) / nr
    out["plain"] = {
        "value": round(terms / sec / 1e9, 4), "unit": "Gterms/s", "ms_per_step": round(sec * 1e3, 4),
        "terms_per_step": terms, "viss_terms": t_viss, "recon_terms": t_recon,
        "stages": "serial" if serial else "recon of segment s beside the Viss of segment s+1",
        "kernel_ms_per_step_serial": _kernel_ms(tim, steps),
        "roofline_recon": _roof(
            "k_recon<4,false,0> (plain recon, uv mirror, representative classes)", "valu_fp64",
            "FP64 vector ALU (published FP64 peak)", RECON_TERM_FLOPS_PLAIN * ev_recon, recon_avg, FP64_PEAK_TFLOPS,
            "TFLOP/s", f"{RECON_TERM_FLOPS_PLAIN} FP64 flops x {ev_recon:.6g} evaluated (point, group) terms per launch "
            f"({nr} launches per step: HEALPix map and lm image per segment)",
            _traffic(workload, "k_recon_plain"),
            fp64_instr_issue_frac=round(RECON_TERM_INSTR_PLAIN * ev_recon / (recon_avg * 1e-3) / 1e12 / FP64_INSTR_PEAK_T, 4),
            launches_per_step=nr),
        "roofline_viss": _roof(
            "k_viss_plain<256,8,false> (plain per-row Viss)", "valu_fp64", "FP64 vector ALU (published FP64 peak)",
            VISS_TERM_FLOPS_PLAIN * ev_viss, viss_avg, FP64_PEAK_TFLOPS, "TFLOP/s",
            f"{VISS_TERM_FLOPS_PLAIN} FP64 flops x {ev_viss:.6g} evaluated (half row, pixel) terms per launch",
            _traffic(workload, "k_viss_plain"),
            fp64_instr_issue_frac=round(VISS_TERM_INSTR_PLAIN * ev_viss / (viss_avg * 1e-3) / 1e12 / FP64_INSTR_PEAK_T, 4),
            launches_per_step=nv),
    }
    dom = "roofline_recon" if tim["recon_kernel_ms"] >= tim["viss_kernel_ms"] else "roofline_viss"
    out["plain"]["dominant"] = dom
    if mixed:
        with Pipeline(_cfg(workload, serial, precision=1)) as mx:
            mx.set_sky(sky)
            per_seg_m, maps_m = _plan_counts(mx, classes=False)
            sec_m, _ = _time_steps(mx, steps, warmup)
        with Pipeline(_cfg(workload, True, precision=1)) as mser:
            mser.set_sky(sky)
            _, tim_m = _time_steps(mser, steps, warmup)
        nvm, nrm = tim_m["n_viss_launch"] // steps, tim_m["n_recon_launch"] // steps
        rec_avg_m = tim_m["recon_kernel_ms"] / (nrm * steps)
        slots = _slots_per_term(rates, MIXED_RECON_TERM)
        vr_p = np.concatenate([s["vraw"] for s in per_seg])
        vr_m = np.concatenate([s["vraw"] for s in per_seg_m])
        out["mixed"] = {
            "value": round(terms / sec_m / 1e9, 4), "unit": "Gterms/s", "ms_per_step": round(sec_m * 1e3, 4),
            "speedup_vs_plain": round(sec / sec_m, 3),
            "dtype": "f64 phase + f32 sincos/sum (Viss: f32 matrix-core products of antenna phasors)",
            "err_vs_fp64": {"viss_half": _err(vr_m, vr_p), "C_healpix": _err(maps_m[0], maps_p[0]),
                            "lm_image": _err(maps_m[2], maps_p[2])},
            "kernel_ms_per_step_serial": _kernel_ms(tim_m, steps),
            "viss_kernel_avg_ms": round(tim_m["viss_kernel_ms"] / (nvm * steps), 5),
            "roofline_recon": _roof(
                "k_recon<4,true,0> (mixed recon)", "valu", "vector ALU issue in FP32-FMA slots (2 flops each) "
                "against the published 157.3 TFLOP/s vector FP32 rate",
                2 * slots * ev_recon, rec_avg_m, FP32_PEAK_TFLOPS, "TFLOP/s",
                f"{slots:.3f} issue slots per evaluated term (mul_f64 + fma_f64 + fract_f64 + cvt + sin_f32 + cos_f32 + "
                f"2 fma_f32, weighted by the issue costs measured by vw_issue_rates) x {ev_recon:.6g} terms per "
                "launch, 2 flops per slot", launches_per_step=nrm),
        }
    if plain_ant:
        with Pipeline(_cfg(workload, serial, viss_form=2)) as ant:
            ant.set_sky(sky)
            per_seg_a, _ = _plan_counts(ant, classes=False)
            sec_a, _ = _time_steps(ant, steps, warmup)
        vr_a = np.concatenate([s["vraw"] for s in per_seg_a])
        vr_p = np.concatenate([s["vraw"] for s in per_seg])
        out["plain_ant"] = {"value": round(terms / sec_a / 1e9, 4), "unit": "Gterms/s",
                            "ms_per_step": round(sec_a * 1e3, 4), "dtype": "f64",
                            "viss_max_rel_err_vs_per_row": _err(vr_a, vr_p)["max_rel"]}
    out["_sky"] = sky
    out["_vraw0"] = per_seg[0]["vraw"]
    return out


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


def batched_leg(workload, steps, warmup, serial, targets=None):
    """64 skies through the batched form on an uneven schedule, against the
    one-sky plain FP64 path for four skies; rooflines of the two matrix-core
    kernels on the evaluated (non-culled) terms"""
    from viss_optimize_ws_amd import Pipeline, make_sky

    W, name = WORKLOADS[workload]
    sched = _uneven_schedule(W["n_epochs"])
    npix = 12 * W["nside"] ** 2
    skies = np.stack([make_sky(7000 + s, npix, 1) for s in range(BATCH_NSKY)])
    ref = []
    with Pipeline(_cfg(workload, serial, seg_tlen=sched)) as p:
        nimg = p.plan.n_img_pts
        p.set_sky(skies[0])
        per_seg, m0 = _plan_counts(p)
        ref.append(m0)
        for s in range(1, 4):
            p.set_sky(skies[s])
            p.reset_accumulators()
            p.run_all()
            ref.append(p.finalize())
    nkeys = [s["nkeys"] for s in per_seg]
    P = W["n_ant"] * (W["n_ant"] - 1) // 2
    terms_sky = sum(npix * tl * P for tl in sched) + (npix + nimg) * sum(nkeys)
    with Pipeline(_cfg(workload, serial, precision=1, nsky=BATCH_NSKY, seg_tlen=sched)) as b:
        b.set_sky(skies)
        if targets is None:
            targets = b.autotune(reps=2)  # launch shapes of both batched kernels
            tuned = True
        else:
            b.set_batch_targets(*targets)
            tuned = False
        b.run_all()
        Cb, Wb, Ib, _ = b.finalize()
        sec_eager, _ = _time_steps(b, steps, warmup)
        sec_graph, _ = _time_steps(b, steps, max(warmup, 1), graph=True)
        Cg, Wg, Ig, _ = b.finalize()
        graph_equal = bool(np.array_equal(Cg, Cb) and np.array_equal(Ig, Ib) and np.array_equal(Wg, Wb))
    sec = min(sec_eager, sec_graph)
    with Pipeline(_cfg(workload, True, precision=1, nsky=BATCH_NSKY, seg_tlen=sched)) as b2:
        b2.set_sky(skies)
        b2.set_batch_targets(*targets)
        _, tim = _time_steps(b2, steps, warmup)
    err = {}
    for key, k, got in (("C_healpix", 0, Cb), ("lm_image", 2, Ib)):
        mx = [float(np.max(np.abs(got[s] - ref[s][k])) / np.max(np.abs(ref[s][k]))) for s in range(4)]
        rm = [float(np.sqrt(np.mean((got[s] - ref[s][k]) ** 2)) / np.sqrt(np.mean(ref[s][k] ** 2))) for s in range(4)]
        err[key] = {"max_rel": max(mx), "rms_rel": max(rm)}
    weights_equal = bool(all(np.array_equal(Wb, r[1]) for r in ref))
    nv, nr = tim["n_viss_launch"] // steps, tim["n_recon_launch"] // steps
    viss_avg = tim["viss_kernel_ms"] / (nv * steps)
    recon_avg = tim["recon_kernel_ms"] / (nr * steps)
    ev_viss = sum(s["evaluated"] for s in per_seg) / nv
    ev_recon = sum(s["rec_evaluated"] for s in per_seg) / nr
    all_viss = sum(npix * tl * P for tl in sched) / nv
    all_recon = (npix + nimg) * sum(nkeys) / nr
    return {
        "workload": name + f"; {BATCH_NSKY} skies, schedule {sched} epochs",
        "value": round(terms_sky * BATCH_NSKY / sec / 1e9, 2), "unit": "Gterms/s", "ms_per_step": round(sec * 1e3, 4),
        "ms_per_step_eager": round(sec_eager * 1e3, 4), "ms_per_step_graph": round(sec_graph * 1e3, 4),
        "graph_equals_eager": graph_equal, "nsky": BATCH_NSKY, "terms_per_step": terms_sky * BATCH_NSKY,
        "schedule_epochs": sched, "launch_targets": {"viss": targets[0], "recon": targets[1], "autotuned": tuned},
        "dtype": "f64 phase, f32 hardware sincos, bf16 hi/lo matrix-core products with f32 accumulation, f64 sums",
        "data": f"synthetic: {BATCH_NSKY} skies log-normal(0,1) + Poisson(100) (cap 200) Pareto(1.5) point sources, "
                "seeds 7000-7063",
        "err_vs_fp64_4_skies": err, "weights_bitexact": weights_equal,
        "kernel_ms_per_step_serial": _kernel_ms(tim, steps),
        "roofline_viss": _roof(
            "k_vissb<4>", "mfma", "BF16 matrix cores (dense peak)", BATCH_TERM_FLOPS * BATCH_NSKY * ev_viss, viss_avg,
            BF16_PEAK_TFLOPS, "TFLOP/s",
            f"{BATCH_TERM_FLOPS} flops per (half row, pixel, sky) on the {ev_viss:.6g} evaluated (non-culled) "
            f"(row, pixel) pairs per launch x {BATCH_NSKY} skies", launches_per_step=nv,
            frac_all_terms=round(BATCH_TERM_FLOPS * BATCH_NSKY * all_viss / (viss_avg * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
                                 4)),
        "roofline_recon": _roof(
            "k_reconb<4>", "mfma", "BF16 matrix cores (dense peak)", BATCH_TERM_FLOPS * BATCH_NSKY * ev_recon,
            recon_avg, BF16_PEAK_TFLOPS, "TFLOP/s",
            f"{BATCH_TERM_FLOPS} flops per (output point, group, sky) on the {ev_recon:.6g} evaluated (point, group) "
            f"pairs per launch x {BATCH_NSKY} skies", launches_per_step=nr,
            frac_all_terms=round(BATCH_TERM_FLOPS * BATCH_NSKY * all_recon / (recon_avg * 1e-3) / 1e12
                                 / BF16_PEAK_TFLOPS, 4)),
    }


def sky_schedule_leg(workload, steps, warmup, serial):
    """64 skies, each on its own uneven schedule (segment lengths uniform in
    [16, 256] epochs, one seeded draw per sky): the batched Viss over the
    context's 256-epoch segments shared by every sky, then per sky its own
    groups, one-sky mixed recon and weights; maps of four skies against the
    one-sky plain FP64 path on their schedules"""
    from viss_optimize_ws_amd import Pipeline, make_sky

    W, name = WORKLOADS[workload]
    T = W["n_epochs"]
    scheds = [_uneven_schedule(T, seed=4242 + s) for s in range(BATCH_NSKY)]
    base = [256] * (T // 256) + ([T % 256] if T % 256 else [])
    npix = 12 * W["nside"] ** 2
    P = W["n_ant"] * (W["n_ant"] - 1) // 2
    skies = np.stack([make_sky(7000 + s, npix, 1) for s in range(BATCH_NSKY)])
    # keys per distinct segment (they depend on the geometry only)
    segs = sorted({(t0, tl) for sc in scheds for t0, tl in zip(np.cumsum([0] + sc[:-1]).tolist(), sc)})
    nkeys = {}
    with Pipeline(_cfg(workload, True, seg_tlen=[256] * (T // 256) + ([T % 256] if T % 256 else []))) as p:
        nimg = p.plan.n_img_pts
        p.set_sky(skies[0])
        for t0, tl in segs:
            p.gen_segment(t0, tl)
            p.viss_half()
            p.conj_fill()
            nkeys[(t0, tl)] = p.recon_segment()["nkeys"]
        p.reset_accumulators()
    terms = 0
    for sc in scheds:
        t0s = np.cumsum([0] + sc[:-1]).tolist()
        terms += sum(npix * tl * P + (npix + nimg) * nkeys[(t0, tl)] for t0, tl in zip(t0s, sc))
    ref = []
    for s in range(4):
        with Pipeline(_cfg(workload, serial, seg_tlen=scheds[s])) as p:
            p.set_sky(skies[s])
            p.run_all()
            ref.append(p.finalize())
    with Pipeline(_cfg(workload, serial, precision=1, nsky=BATCH_NSKY, seg_tlen=base)) as b:
        b.set_sky(skies)
        b.set_sky_schedules(scheds)
        b.run_all()
        Cb, Wb, Ib, WIb = b.finalize()
        sec_eager, _ = _time_steps(b, steps, warmup)
        sec_graph, _ = _time_steps(b, steps, max(warmup, 1), graph=True)
        Cg, Wg, Ig, _ = b.finalize()
        graph_equal = bool(np.array_equal(Cg, Cb) and np.array_equal(Ig, Ib) and np.array_equal(Wg, Wb))
    sec = min(sec_eager, sec_graph)
    with Pipeline(_cfg(workload, True, precision=1, nsky=BATCH_NSKY, seg_tlen=base)) as b2:
        b2.set_sky(skies)
        b2.set_sky_schedules(scheds)
        _, tim = _time_steps(b2, steps, warmup)
    err = {}
    for key, k, got in (("C_healpix", 0, Cb), ("lm_image", 2, Ib)):
        err[key] = {"max_rel": max(float(np.max(np.abs(got[s] - ref[s][k])) / np.max(np.abs(ref[s][k])))
                                   for s in range(4)),
                    "rms_rel": max(float(np.sqrt(np.mean((got[s] - ref[s][k]) ** 2)) / np.sqrt(np.mean(ref[s][k] ** 2)))
                                   for s in range(4))}
    weights_equal = bool(all(np.array_equal(Wb[s], ref[s][1]) and np.array_equal(WIb[s], ref[s][3]) for s in range(4)))
    return {
        "workload": name + f"; {BATCH_NSKY} skies, each on its own schedule of segments uniform in [16, 256] epochs",
        "value": round(terms / sec / 1e9, 2), "unit": "Gterms/s", "ms_per_step": round(sec * 1e3, 4),
        "ms_per_step_eager": round(sec_eager * 1e3, 4), "ms_per_step_graph": round(sec_graph * 1e3, 4),
        "graph_equals_eager": graph_equal, "nsky": BATCH_NSKY, "terms_per_step": terms,
        "segments": sum(len(x) for x in scheds), "distinct_segments": len(segs), "viss_schedule_epochs": base,
        "schedules_first_4_skies": scheds[:4],
        "dtype": "Viss: f64 phase, f32 hardware sincos, bf16 hi/lo matrix-core products; recon per sky: f64 phase, "
                 "f32 hardware sincos, f32 block sums folded in f64",
        "err_vs_fp64_4_skies": err, "weights_bitexact_4_skies": weights_equal,
        "kernel_ms_per_step_serial": _kernel_ms(tim, steps),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, sky, vraw0, threads):
    """CPU FP64 oracle on a bounded sample of the first segment.  Nside 128:
    Viss of all its half rows and reconstruction of every HEALPix pixel and
    image point.  Larger workloads: Viss of the first 2,048 half rows, then,
    from the segment's GPU half visibilities, grouping and the
    reconstruction of the first 16,384 HEALPix pixels and image points."""
    from oracle.oracle import Oracle

    full = workload == "nside128"
    with Oracle(_cfg(workload, False)) as o:
        o.set_sky(sky)
        t0, tl = o.schedule()[0]
        o.gen_segment(t0, tl)
        if full:
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
    return dict(value=terms / dt / 1e9, unit="Gterms/s", cores=threads, cpu_model=_cpu_model(), kind="port",
                sample=sample)


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


def kernel_table(trace_csv, top=20):
    """per-kernel dispatch statistics of a rocprofv3 kernel_trace.csv: all
    dispatches, and those that overlapped no other dispatch (ran alone)"""
    import csv

    rows = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            rows.append((r["Kernel_Name"], int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    rows.sort(key=lambda r: r[1])
    # sorted by start: dispatch i overlaps an earlier one iff it starts before
    # the latest end so far, and a later one iff it ends after the next start
    alone, end_max = [], -1
    for i, (_, s, e) in enumerate(rows):
        nxt = rows[i + 1][1] if i + 1 < len(rows) else e
        alone.append(s >= end_max and e <= nxt)
        end_max = max(end_max, e)
    stats = {}
    for (k, s, e), a in zip(rows, alone):
        d = stats.setdefault(k, [0, 0.0, 0, 0.0])
        d[0] += 1
        d[1] += (e - s) / 1e3
        if a:
            d[2] += 1
            d[3] += (e - s) / 1e3
    tot = sum(d[1] for d in stats.values())
    print("| kernel | calls | total ms | avg us | % | alone calls | alone avg us |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for k, d in sorted(stats.items(), key=lambda kv: -kv[1][1])[:top]:
        al = f"{d[3] / d[2]:.1f}" if d[2] else "-"
        print(f"| `{k[:80]}` | {d[0]} | {d[1] / 1e3:.3f} | {d[1] / d[0]:.1f} | {100 * d[1] / tot:.2f} | {d[2]} | {al} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed steps of the Nside 128 legs")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--steps256", type=int, default=3, help="timed steps of the Nside 256 legs")
    ap.add_argument("--warmup256", type=int, default=1)
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma-separated subset of " + ",".join(LEGS) + " (default: all)")
    ap.add_argument("--serial", action="store_true",
                    help="run every stage on one stream (no overlap of a segment's recon with the next Viss)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision form in the one-sky legs")
    ap.add_argument("--plain-ant", action="store_true", help="add the FP64 antenna-factorized Viss form (Nside 128)")
    ap.add_argument("--batch-targets", default="", metavar="V,R",
                    help="launch targets of the batched kernels instead of the autotuner")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (calibrates HBM byte counters) and exit")
    ap.add_argument("--issue-rates", action="store_true", help="only measure the VALU issue costs and exit")
    ap.add_argument("--kernel-table", default="", metavar="CSV",
                    help="only summarise a rocprofv3 kernel_trace.csv (no GPU) and exit")
    args = ap.parse_args()
    if args.kernel_table:
        kernel_table(args.kernel_table)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (every box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        _copy_calibration(args.copy_calib)
        return
    rates = issue_rates()
    _log(f"VALU issue costs measured: {rates['slots']}")
    if args.issue_rates:
        print(json.dumps(rates), flush=True)
        return
    legs = [x for x in args.legs.split(",") if x]
    bad = [x for x in legs if x not in LEGS]
    if bad:
        sys.exit(f"unknown legs {bad}")
    targets = tuple(int(x) for x in args.batch_targets.split(",")) if args.batch_targets else None

    out = {"metric": METRIC, "value": None, "unit": "Gterms/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": f"synthetic: log-normal(0,1) sky seed {SKY_SEED}, "
           "circular orbits seed 42"}
    issue = {"effective_clock_ghz": rates["effective_clock_ghz"], "slots": rates["slots"]}
    if "nside128" in legs or "cpu" in legs:
        r = one_sky_leg("nside128", args.steps, args.warmup, args.serial, rates, mixed=not args.no_mixed,
                        plain_ant=args.plain_ant)
        _log("leg nside128 done")
        sky, vraw0 = r.pop("_sky"), r.pop("_vraw0")
        p = r["plain"]
        out["value"] = p["value"]
        out["ms_per_step"] = p["ms_per_step"]
        out["config"] = {"workload": r["workload"], **WORKLOADS["nside128"][0], "form": "plain FP64",
                         "stages": p["stages"], "terms_per_step": p["terms_per_step"],
                         "viss_terms": p["viss_terms"], "recon_terms": p["recon_terms"]}
        out["kernel_ms_per_step_serial"] = p["kernel_ms_per_step_serial"]
        out["roofline"] = p[p["dominant"]]
        out["roofline_viss"] = p["roofline_viss"]
        out["roofline_recon"] = p["roofline_recon"]
        if "mixed" in r:
            out["mixed"] = r["mixed"]
        if "plain_ant" in r:
            out["plain_ant"] = r["plain_ant"]
        if "cpu" in legs:
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            out["cpu_baseline"] = cpu_baseline("nside128", sky, vraw0, threads)
            _log("leg cpu done")
    out["valu_issue_costs"] = issue
    if "nside256" in legs:
        r = one_sky_leg("nside256", args.steps256, args.warmup256, args.serial, rates, mixed=not args.no_mixed)
        _log("leg nside256 done")
        r.pop("_sky"), r.pop("_vraw0")
        out["nside256"] = r
        if out["value"] is None:
            out["value"], out["ms_per_step"] = r["plain"]["value"], r["plain"]["ms_per_step"]
    if "batched128" in legs:
        out["batched"] = batched_leg("nside128", args.steps, args.warmup, args.serial, targets)
        _log("leg batched128 done")
        if "nside128" in legs:
            out["batched"]["speedup_vs_plain"] = round(out["batched"]["value"] / out["value"], 3)
    if "skysched128" in legs:
        out["batched_sky_schedules"] = sky_schedule_leg("nside128", args.steps256, args.warmup256, args.serial)
        _log("leg skysched128 done")
        if out.get("batched"):
            out["batched_sky_schedules"]["vs_shared_schedule"] = round(
                out["batched_sky_schedules"]["value"] / out["batched"]["value"], 3)
    if "batched256" in legs:
        b = batched_leg("nside256", args.steps256, args.warmup256, args.serial, targets)
        _log("leg batched256 done")
        if "nside256" in out:
            b["speedup_vs_plain"] = round(b["value"] / out["nside256"]["plain"]["value"], 3)
            if "mixed" in out["nside256"]:
                b["speedup_vs_mixed"] = round(b["value"] / out["nside256"]["mixed"]["value"], 3)
            out["nside256"]["batched"] = b
        else:
            out["batched256"] = b
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
