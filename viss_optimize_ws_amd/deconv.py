"""Learned deconvolution of segment-summed dirty images (include/vw_unet.h).

The reference stops at the dirty map; this is the learned stage beside it:

* ``UNet`` drives the U-Net-lite of libvissws.so (forward, backward and Adam
  in hand-written gfx950 kernels) through its C-ABI.
* ``make_pairs`` produces (dirty image, sky) pairs with the plain FP64
  pipeline: seeded skies (log-normal diffuse field plus Poisson(100) point
  sources, capped at 200, Pareto(1.5) flux) restricted to the hemisphere the
  lm image sees (n > 0), the circular-orbit geometry of 8 antennas and 64
  epochs in 4 segments at 10 MHz, and the recon's lm-grid image; the target
  is the sky sampled on the same grid (the HEALPix NESTED pixel of each grid
  direction, 0 outside the unit disk).
* ``normalise`` scales each pair by its dirty image's RMS inside the disk
  (and one global gain for the target), so the network sees O(1) values;
  ``denormalise`` undoes it.
* ``resimulate_gate`` puts an estimate back on the HEALPix pixels of the
  visible hemisphere, optionally refines it towards the measured half
  visibilities (``refine``: conjugate-gradient least squares with the plain
  Viss kernel and its adjoint, ``vw_viss_adjoint``, restricted to the
  visible hemisphere), re-simulates its half visibilities with the plain Viss
  kernel and compares them with the measured ones; an estimate whose
  relative residual exceeds the threshold (1e-2) is replaced by the dirty
  image under the training set's gain (information available at inference).

Command line (one MI355X): ``python -m viss_optimize_ws_amd.deconv --pairs
20000 --heldout 512`` prints one JSON line with the training time, the
held-out errors of the network and of the dirty image, and the gate's
acceptance.
"""
import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import quality
from .pipeline import Pipeline, VissError, default_config, load_library, make_sky

GATE_THRESHOLD = 1e-2


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class UNet:
    """The U-Net-lite of libvissws.so (img_n a multiple of 8)."""

    def __init__(self, img_n, batch_max, fp16=True, seed=1):
        self._lib = load_library()
        self.img_n, self.batch_max = int(img_n), int(batch_max)
        self._ctx = C.c_void_p()
        rc = self._lib.vw_unet_create(self.img_n, self.batch_max, 1 if fp16 else 0, int(seed), C.byref(self._ctx))
        if rc:
            raise VissError(rc, self._lib.vw_unet_last_error(None).decode())
        self.nparam = int(self._lib.vw_unet_param_count(self._ctx))

    def _check(self, rc):
        if rc:
            raise VissError(rc, self._lib.vw_unet_last_error(self._ctx).decode())

    def close(self):
        if self._ctx:
            self._lib.vw_unet_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        self._check(self._lib.vw_unet_get_params(self._ctx, _fp(out)))
        return out

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, np.float32)
        assert theta.size == self.nparam
        self._check(self._lib.vw_unet_set_params(self._ctx, _fp(theta)))

    def loss_grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        loss = C.c_double(0.0)
        g = np.zeros(self.nparam, np.float32)
        self._check(self._lib.vw_unet_loss_grad(self._ctx, _fp(X), _fp(Y), X.shape[0], C.byref(loss), _fp(g)))
        return loss.value, g

    def set_data(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        self._check(self._lib.vw_unet_set_data(self._ctx, _fp(X), _fp(Y), X.shape[0]))

    def train(self, steps, batch, lr, seed=0):
        losses = np.zeros(int(steps), np.float32)
        self._check(self._lib.vw_unet_train(self._ctx, int(steps), int(batch), float(lr), int(seed), _fp(losses)))
        return losses

    def train_ms(self):
        ms = C.c_double(0.0)
        self._check(self._lib.vw_unet_train_ms(self._ctx, C.byref(ms)))
        return ms.value

    def predict(self, X, noalloc=False):
        """network output for X [n][img_n][img_n]; noalloc=True stages the
        images through the library's static device array (no allocation)"""
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros_like(X)
        f = self._lib.vw_unet_predict_noalloc if noalloc else self._lib.vw_unet_predict
        self._check(f(self._ctx, _fp(X), X.shape[0], _fp(out)))
        return out


def geometry_config(nside=64, img_n=256, **kw):
    """the pairs' geometry: 8 antennas on seeded circular orbits, 64 epochs
    10 s apart, 4 equal segments, 10 MHz, plain FP64"""
    base = dict(nside=nside, n_ant=8, n_epochs=64, freq_hz=1e7, segs=4, img_n=img_n)
    base.update(kw)
    return default_config(1, **base)


def grid_pixel_index(nside, img_n):
    """HEALPix NESTED pixel of each lm-grid pixel centre (l = (2j + 1 - N) / N,
    m = (2i + 1 - N) / N, n = +sqrt(1 - l^2 - m^2); the pipeline's directions
    are l = sin(t) cos(p), m = -sin(t) sin(p), n = cos(t)); -1 outside the disk"""
    j = np.arange(img_n)
    lg = ((2 * j + 1 - img_n) / img_n)[None, :].repeat(img_n, 0)
    mg = ((2 * j + 1 - img_n) / img_n)[:, None].repeat(img_n, 1)
    r2 = lg * lg + mg * mg
    inside = r2 < 1.0
    ng = np.sqrt(np.clip(1.0 - r2, 0.0, 1.0))
    theta = np.arccos(np.clip(ng, -1.0, 1.0))
    phi = np.mod(np.arctan2(-mg, lg), 2 * np.pi)
    idx = np.asarray(quality.ang2pix_nest(nside, theta.ravel(), phi.ravel()), np.int64).reshape(img_n, img_n)
    return np.where(inside, idx, -1)


def front_sky(seed, npix, n_dir):
    """seeded sky (diffuse + point sources) on the hemisphere n > 0"""
    return np.where(n_dir > 0.0, make_sky(seed, npix, 1), 0.0)


def make_pairs(seeds, nside=64, img_n=256, cfg=None, log=None):
    """(dirty images, target skies on the grid) for the seeds, through the
    plain FP64 pipeline"""
    cfg = cfg or geometry_config(nside, img_n)
    idx = grid_pixel_index(nside, img_n)
    X = np.zeros((len(seeds), img_n, img_n), np.float32)
    Y = np.zeros((len(seeds), img_n, img_n), np.float32)
    with Pipeline(cfg) as p:
        n_dir = p.lmn()[2]
        npix = p.plan.npix
        for k, s in enumerate(seeds):
            sky = front_sky(int(s), npix, n_dir)
            p.set_sky(sky)
            p.reset_accumulators()
            p.run_all()
            X[k] = p.finalize()[2]
            Y[k] = np.where(idx >= 0, sky[np.maximum(idx, 0)], 0.0)
            if log and (k + 1) % 2000 == 0:
                log(f"{k + 1} pairs")
    return X, Y


def normalise(X, Y=None, gain=None):
    """per-pair scale s = RMS of the dirty image inside the disk; x = X / s,
    y = Y / (s * gain) with one global gain (median of RMS(Y) / RMS(X) when
    not given).  Returns (x, y, s, gain)."""
    n = X.shape[-1]
    j = np.arange(n)
    disk = ((2 * j + 1 - n)[None, :] ** 2 + (2 * j + 1 - n)[:, None] ** 2) < n * n
    s = np.sqrt(np.mean(X[:, disk].astype(np.float64) ** 2, axis=1))
    s = np.where(s > 0, s, 1.0)
    x = (X / s[:, None, None]).astype(np.float32)
    if Y is None:
        return x, None, s, gain
    if gain is None:
        sy = np.sqrt(np.mean(Y[:, disk].astype(np.float64) ** 2, axis=1))
        gain = float(np.median(sy / s))
    y = (Y / (s[:, None, None] * gain)).astype(np.float32)
    return x, y, s, gain


def denormalise(out, s, gain):
    return out * (s[:, None, None] * gain)


def disk_rel_err(est, ref):
    """relative L2 error inside the disk, per image"""
    n = ref.shape[-1]
    j = np.arange(n)
    disk = ((2 * j + 1 - n)[None, :] ** 2 + (2 * j + 1 - n)[:, None] ** 2) < n * n
    e = est[:, disk].astype(np.float64) - ref[:, disk]
    return np.sqrt((e * e).sum(axis=1) / np.maximum((ref[:, disk].astype(np.float64) ** 2).sum(axis=1), 1e-300))


def best_scaled(X, Y):
    """the dirty image under the least-squares scalar gain to each target (the
    method without the network, given its best calibration)"""
    n = X.shape[-1]
    j = np.arange(n)
    disk = ((2 * j + 1 - n)[None, :] ** 2 + (2 * j + 1 - n)[:, None] ** 2) < n * n
    xx = X[:, disk].astype(np.float64)
    yy = Y[:, disk].astype(np.float64)
    a = (xx * yy).sum(axis=1) / np.maximum((xx * xx).sum(axis=1), 1e-300)
    return X * a[:, None, None]


def grid_to_healpix(img, idx_pix, npix):
    """grid estimate back onto HEALPix pixels: each pixel of the visible
    hemisphere takes the mean of the grid pixels that fall in it (0 where
    none does); the far hemisphere is 0"""
    out = np.zeros(npix)
    cnt = np.zeros(npix)
    sel = idx_pix >= 0
    np.add.at(out, idx_pix[sel], img[sel])
    np.add.at(cnt, idx_pix[sel], 1.0)
    return np.where(cnt > 0, out / np.maximum(cnt, 1.0), 0.0)


def healpix_to_grid(sky, idx_pix):
    """HEALPix values sampled on the lm grid (the pixel of each grid
    direction; 0 outside the unit disk), the targets' convention"""
    return np.where(idx_pix >= 0, sky[np.maximum(idx_pix, 0)], 0.0)


class Measurement:
    """the plain FP64 half-visibility sums of the pairs' geometry as a linear
    map of the sky, on one pipeline context: ``forward`` (all segments' raw
    half visibilities) and ``adjoint`` (vw_viss_adjoint summed over the
    segments)"""

    def __init__(self, pipe):
        self.p = pipe
        self.sched = pipe.schedule()

    def forward(self, x):
        self.p.set_sky(x)
        out = []
        for t0, tl in self.sched:
            self.p.gen_segment(t0, tl)
            out.append(self.p.viss_half())
        return out

    def adjoint(self, rs):
        g = np.zeros(self.p.plan.npix)
        for (t0, tl), r in zip(self.sched, rs):
            self.p.gen_segment(t0, tl)
            g += self.p.viss_adjoint(r)
        return g


def _dot(a, b):
    return float(sum(np.sum(x.real * y.real + x.imag * y.imag) for x, y in zip(a, b)))


def refine(meas, x0, v, support, iters=50, tol=GATE_THRESHOLD * 0.5):
    """data-consistency refinement: conjugate-gradient least squares (CGLS)
    on min |A x - v|^2 over skies x supported on `support` (A = the plain
    half-visibility sums of every segment), started from x0; stops once the
    relative residual |A x - v| / |v| is at most tol.  Returns (x, relative
    residual after each iteration)."""
    x = np.where(support, x0, 0.0)
    vn = np.sqrt(_dot(v, v))
    r = [a - b for a, b in zip(v, meas.forward(x))]
    hist = [np.sqrt(_dot(r, r)) / vn]
    s = meas.adjoint(r) * support
    d = s.copy()
    gam = float(np.dot(s, s))
    for _ in range(iters):
        if hist[-1] <= tol or gam == 0.0:
            break
        q = meas.forward(d)
        alpha = gam / _dot(q, q)
        x = x + alpha * d
        r = [a - alpha * b for a, b in zip(r, q)]
        hist.append(np.sqrt(_dot(r, r)) / vn)
        s = meas.adjoint(r) * support
        g2 = float(np.dot(s, s))
        d = s + (g2 / gam) * d
        gam = g2
    return x, hist


def resimulate_gate(estimates, fallback, skies, cfg, threshold=GATE_THRESHOLD, refine_iters=0):
    """re-simulation gate over grid estimates: each goes back onto the HEALPix
    pixels of the visible hemisphere (grid_to_healpix), is refined towards
    the measured visibilities when refine_iters > 0, and is re-simulated by
    the plain Viss kernel; its relative residual against the visibilities of
    the true sky decides.  Returns (residuals, accepted mask, gated grid
    output = the (refined) estimate or the fallback image)."""
    idx = grid_pixel_index(cfg.nside, cfg.img_n)
    res = np.zeros(len(estimates))
    out = np.array(estimates, dtype=np.float64)
    with Pipeline(cfg) as p:
        npix = p.plan.npix
        support = p.lmn()[2] > 0.0
        meas = Measurement(p)
        for k in range(len(estimates)):
            v = meas.forward(skies[k])
            x = grid_to_healpix(estimates[k], idx, npix)
            if refine_iters:
                x, _ = refine(meas, x, v, support, refine_iters, 0.5 * threshold)
                out[k] = healpix_to_grid(x, idx)
            ve = meas.forward(x)
            r = [a - b for a, b in zip(ve, v)]
            res[k] = np.sqrt(_dot(r, r) / max(_dot(v, v), 1e-300))
    ok = res <= threshold
    gated = np.where(ok[:, None, None], out, fallback)
    return res, ok, gated


def run(pairs=20000, heldout=512, nside=64, img_n=256, batch=16, epochs=2.0, lr=1e-3, fp16=True, seed=1,
        gate_samples=64, refine_iters=50, log=None):
    """the full experiment; returns a dict of figures"""
    t0 = time.perf_counter()
    seeds = np.arange(100000, 100000 + pairs)
    hseeds = np.arange(900000, 900000 + heldout)
    X, Y = make_pairs(seeds, nside, img_n, log=log)
    Xh, Yh = make_pairs(hseeds, nside, img_n)
    t_data = time.perf_counter() - t0
    if log:
        log(f"pairs made in {t_data:.1f} s")
    x, y, _, gain = normalise(X, Y)
    xh, _, sh, _ = normalise(Xh)
    steps = int(np.ceil(epochs * pairs / batch))
    with UNet(img_n, batch, fp16=fp16, seed=seed) as net:
        net.set_data(x, y)
        del x, y
        # ten slices of the step budget (Adam state carries over), one log line each
        losses, train_ms = [], 0.0
        for i in range(10):
            k = steps // 10 + (1 if i < steps % 10 else 0)
            if k == 0:
                continue
            losses.append(net.train(k, batch, lr, seed=seed * 1000 + i))
            train_ms += net.train_ms()
            if log:
                log(f"train slice {i + 1}/10: loss {float(losses[-1][-1]):.4g}")
        losses = np.concatenate(losses)
        est = denormalise(net.predict(xh), sh, gain)
    e_net = disk_rel_err(est, Yh)
    e_dirty = disk_rel_err(best_scaled(Xh, Yh), Yh)
    # the dirty image as inference sees it: its own RMS scale and the training gain
    dirty_inf = denormalise(xh, sh, gain)
    e_dirty_inf = disk_rel_err(dirty_inf, Yh)
    cfg = geometry_config(nside, img_n)
    ng = min(gate_samples, heldout)
    with Pipeline(cfg) as p:
        n_dir = p.lmn()[2]
        skies = np.stack([front_sky(int(s), p.plan.npix, n_dir) for s in hseeds[:ng]])
    # floor of the grid representation: the true targets put back on the sphere
    res_floor, _, _ = resimulate_gate(Yh[:ng], dirty_inf[:ng], skies, cfg)
    # the gate on the network's estimates as they are, then after refinement
    res_raw, ok_raw, _ = resimulate_gate(est[:ng], dirty_inf[:ng], skies, cfg)
    res, ok, gated = resimulate_gate(est[:ng], dirty_inf[:ng], skies, cfg, refine_iters=refine_iters)
    # the same refinement without the network: from the dirty image, and from zero
    res_d, ok_d, gated_d = resimulate_gate(dirty_inf[:ng], dirty_inf[:ng], skies, cfg, refine_iters=refine_iters)
    res_z, ok_z, gated_z = resimulate_gate(np.zeros_like(est[:ng]), dirty_inf[:ng], skies, cfg,
                                           refine_iters=refine_iters)

    def stat(e):
        return {"mean": float(np.mean(e)), "median": float(np.median(e))}

    e_g = disk_rel_err(gated, Yh[:ng])
    return {
        "pairs": pairs, "heldout": heldout, "nside": nside, "img_n": img_n, "batch": batch, "steps": steps,
        "activations": "fp16" if fp16 else "fp32", "lr": lr,
        "data_s": round(t_data, 2), "train_ms": round(train_ms, 1),
        "train_images_per_s": round(steps * batch / (train_ms * 1e-3), 1),
        "loss_first": float(losses[: max(1, steps // 50)].mean()), "loss_last": float(losses[-max(1, steps // 50):].mean()),
        "heldout_rel_err_network": stat(e_net),
        "heldout_rel_err_dirty_best_gain": stat(e_dirty),
        "heldout_rel_err_dirty_training_gain": stat(e_dirty_inf),
        "network_better_fraction": float(np.mean(e_net < e_dirty)),
        "gate": {"threshold": GATE_THRESHOLD, "samples": ng, "refine_iters": refine_iters,
                 "floor_residual_targets": stat(res_floor),
                 "network_unrefined": {"accepted": int(ok_raw.sum()), "residual": stat(res_raw)},
                 "network_refined": {"accepted": int(ok.sum()), "residual": stat(res),
                                     "gated_rel_err": stat(e_g)},
                 "dirty_refined": {"accepted": int(ok_d.sum()), "residual": stat(res_d),
                                   "gated_rel_err": stat(disk_rel_err(gated_d, Yh[:ng]))},
                 "zero_refined": {"accepted": int(ok_z.sum()), "residual": stat(res_z),
                                  "gated_rel_err": stat(disk_rel_err(gated_z, Yh[:ng]))},
                 "network_refined_better_fraction_vs_dirty_refined": float(
                     np.mean(e_g < disk_rel_err(gated_d, Yh[:ng])))},
    }


def time_training(img_n=256, batch=16, steps=200, fp16=True, seed=1):
    """training throughput alone: a resident set of 16 batches of random
    images, 20 untimed steps, then `steps` timed Adam steps (wall clock around
    vw_unet_train, which returns after its stream is drained)"""
    rng = np.random.default_rng(seed)
    n = 16 * batch
    X = rng.standard_normal((n, img_n, img_n)).astype(np.float32)
    Y = rng.standard_normal((n, img_n, img_n)).astype(np.float32)
    with UNet(img_n, batch, fp16=fp16, seed=seed) as net:
        net.set_data(X, Y)
        net.train(20, batch, 1e-3, seed=0)
        t = time.perf_counter()
        losses = net.train(steps, batch, 1e-3, seed=1)
        dt = time.perf_counter() - t
    return {"img_n": img_n, "batch": batch, "steps": steps, "activations": "fp16" if fp16 else "fp32",
            "ms_per_step": round(1e3 * dt / steps, 4), "images_per_s": round(steps * batch / dt, 1),
            "loss_last": float(losses[-1])}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--heldout", type=int, default=512)
    ap.add_argument("--nside", type=int, default=64)
    ap.add_argument("--img_n", type=int, default=256)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--epochs", type=float, default=2.0)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--fp32", action="store_true", help="FP32 activations (default FP16)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--gate-samples", type=int, default=64)
    ap.add_argument("--refine-iters", type=int, default=50, help="CGLS iterations of the data-consistency refinement")
    ap.add_argument("--time-steps", type=int, default=0,
                    help="only time this many training steps on random images (no pairs, no gate) and exit")
    a = ap.parse_args(argv)
    if a.time_steps > 0:
        print(json.dumps(time_training(a.img_n, a.batch, a.time_steps, not a.fp32, a.seed)), flush=True)
        return
    out = run(a.pairs, a.heldout, a.nside, a.img_n, a.batch, a.epochs, a.lr, not a.fp32, a.seed,
              gate_samples=a.gate_samples, refine_iters=a.refine_iters,
              log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
