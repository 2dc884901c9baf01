"""Image-quality figures of a reconstructed HEALPix NESTED map against the
original sky (the evaluation of healpix_ws_10Mhz.ipynb cells 4-6, written
with numpy only).

Steps, as the notebook does them:
  * valid = finite in both maps; band = valid and |lat| <= 30 deg;
  * the reconstruction is linearly calibrated on the band,
    original ~ a * recon + b (least squares);
  * Pearson r, NRMSE (RMSE over the reference's max - min), cosine
    similarity, signed and positive-only flux ratios, on the full sky and
    on the band;
  * SSIM of a plate-carree projection (2000 x 600 pixels over
    lon [-180, 180], the region's latitude range) of both maps after a
    common [0, 1] normalisation between the joint 0.5 / 99.5 percentiles;
  * angular distance between the brightest pixels (full sky and band).

SSIM follows Wang et al. (2004) with the defaults of the notebook's library
call: 7 x 7 uniform window, K1 = 0.01, K2 = 0.03, data range 1, sample
(N - 1) covariances, mean over the window centres at least 3 pixels from the
border.  The projection samples the map at each output pixel centre through
a NESTED ang2pix of our own (Gorski et al. 2005).

Command line:
  python -m viss_optimize_ws_amd.quality --original B.bin --recon C.bin
(both float32 NESTED maps of the same size).
"""
import argparse
import json
import sys

import numpy as np

BAND_LAT_DEG = 30.0
SSIM_XSIZE, SSIM_YSIZE = 2000, 600


# ---------------------------------------------------------------- HEALPix
def _spread_bits(v):
    """Interleave the bits of v (< 2^30) with zeros: bit i -> bit 2i."""
    v = v.astype(np.int64)
    out = np.zeros_like(v)
    for i in range(30):
        out |= ((v >> i) & 1) << (2 * i)
    return out


def _compress_bits(v):
    v = v.astype(np.int64)
    out = np.zeros_like(v)
    for i in range(30):
        out |= ((v >> (2 * i)) & 1) << i
    return out


_JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4])
_JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7])


def pix2ang_nest(nside, ipix):
    """(theta, phi) of NESTED pixel centres."""
    ipix = np.asarray(ipix, np.int64)
    npface = nside * nside
    face = ipix // npface
    ip = ipix % npface
    ix = _compress_bits(ip)
    iy = _compress_bits(ip >> 1)
    jr = _JRLL[face] * nside - ix - iy - 1
    nl4 = 4 * nside
    z = np.empty(ipix.shape)
    nr = np.empty(ipix.shape, np.int64)
    kshift = np.zeros(ipix.shape, np.int64)
    north = jr < nside
    south = jr > 3 * nside
    eq = ~north & ~south
    nr[north] = jr[north]
    z[north] = 1.0 - nr[north] ** 2 / (3.0 * npface)
    nr[south] = nl4 - jr[south]
    z[south] = nr[south] ** 2 / (3.0 * npface) - 1.0
    nr[eq] = nside
    z[eq] = (2 * nside - jr[eq]) * 2.0 / (3.0 * nside)
    kshift[eq] = (jr[eq] - nside) & 1
    jp = (_JPLL[face] * nr + ix - iy + 1 + kshift) // 2
    jp = np.where(jp > nl4, jp - nl4, jp)
    jp = np.where(jp < 1, jp + nl4, jp)
    phi = (jp - (kshift + 1) * 0.5) * (np.pi / 2) / nr
    return np.arccos(np.clip(z, -1, 1)), phi


def ang2pix_nest(nside, theta, phi):
    """NESTED pixel containing each direction."""
    theta = np.asarray(theta, np.float64)
    phi = np.asarray(phi, np.float64)
    z = np.cos(theta)
    za = np.abs(z)
    tt = np.mod(phi, 2 * np.pi) / (np.pi / 2)  # in [0, 4)
    face = np.empty(z.shape, np.int64)
    ix = np.empty(z.shape, np.int64)
    iy = np.empty(z.shape, np.int64)
    eq = za <= 2.0 / 3.0
    # equatorial region
    temp1 = nside * (0.5 + tt[eq])
    temp2 = nside * (z[eq] * 0.75)
    jp = (temp1 - temp2).astype(np.int64)
    jm = (temp1 + temp2).astype(np.int64)
    ifp = jp // nside
    ifm = jm // nside
    f = np.where(ifp == ifm, (ifp % 4) + 4, np.where(ifp < ifm, ifp % 4, (ifm % 4) + 8))
    face[eq] = f
    ix[eq] = jm & (nside - 1)
    iy[eq] = nside - (jp & (nside - 1)) - 1
    # polar caps
    pc = ~eq
    ntt = np.minimum(tt[pc].astype(np.int64), 3)
    tp = tt[pc] - ntt
    tmp = nside * np.sqrt(3 * (1 - za[pc]))
    jp = np.minimum((tp * tmp).astype(np.int64), nside - 1)
    jm = np.minimum(((1.0 - tp) * tmp).astype(np.int64), nside - 1)
    north = z[pc] >= 0
    face[pc] = np.where(north, ntt, ntt + 8)
    ix[pc] = np.where(north, nside - jm - 1, jp)
    iy[pc] = np.where(north, nside - jp - 1, jm)
    return face * nside * nside + _spread_bits(ix) + (_spread_bits(iy) << 1)


def cart_project(m, lat_min, lat_max, xsize=SSIM_XSIZE, ysize=SSIM_YSIZE):
    """Plate-carree image of a NESTED map over lon [-180, 180] and the
    latitude range, sampled at the output pixel centres (row 0 = lat_min,
    longitude increasing to the left as in sky maps)."""
    nside = int(round(np.sqrt(m.size / 12)))
    lon = 180.0 - (np.arange(xsize) + 0.5) * (360.0 / xsize)
    lat = lat_min + (np.arange(ysize) + 0.5) * ((lat_max - lat_min) / ysize)
    LON, LAT = np.meshgrid(np.radians(lon), np.radians(lat))
    pix = ang2pix_nest(nside, np.pi / 2 - LAT, LON)
    return m[pix]


# ---------------------------------------------------------------- SSIM
def _uniform7(a):
    """Mean over 7 x 7 windows (valid region only)."""
    c = np.cumsum(np.cumsum(np.pad(a, ((1, 0), (1, 0))), 0), 1)
    s = c[7:, 7:] - c[:-7, 7:] - c[7:, :-7] + c[:-7, :-7]
    return s / 49.0


def ssim(x, y, data_range=1.0, k1=0.01, k2=0.03):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = 49
    ux, uy = _uniform7(x), _uniform7(y)
    uxx, uyy, uxy = _uniform7(x * x), _uniform7(y * y), _uniform7(x * y)
    cov = n / (n - 1.0)
    vx = cov * (uxx - ux * ux)
    vy = cov * (uyy - uy * uy)
    vxy = cov * (uxy - ux * uy)
    c1 = (k1 * data_range) ** 2
    c2 = (k2 * data_range) ** 2
    s = ((2 * ux * uy + c1) * (2 * vxy + c2)) / ((ux * ux + uy * uy + c1) * (vx + vy + c2))
    return float(s.mean())


# ---------------------------------------------------------------- figures
def calibrate(original, recon, mask):
    x = recon[mask].astype(np.float64)
    y = original[mask].astype(np.float64)
    a, b = np.linalg.lstsq(np.vstack([x, np.ones_like(x)]).T, y, rcond=None)[0]
    return float(a), float(b), a * recon.astype(np.float64) + b


def pearson(x, y, mask):
    x = x[mask].astype(np.float64)
    y = y[mask].astype(np.float64)
    x = x - x.mean()
    y = y - y.mean()
    return float(np.dot(x, y) / (np.sqrt(np.dot(x, x) * np.dot(y, y)) + 1e-12))


def nrmse(ref, pred, mask):
    r = ref[mask].astype(np.float64)
    p = pred[mask].astype(np.float64)
    rmse = np.sqrt(np.mean((r - p) ** 2))
    return float(rmse / (r.max() - r.min() + 1e-12)), float(rmse)


def cosine(x, y, mask):
    x = x[mask].astype(np.float64)
    y = y[mask].astype(np.float64)
    return float(np.dot(x, y) / (np.sqrt(np.dot(x, x) * np.dot(y, y)) + 1e-12))


def flux_ratio(ref, pred, mask, positive_only=False):
    r = ref[mask].astype(np.float64)
    p = pred[mask].astype(np.float64)
    if positive_only:
        r = np.clip(r, 0, None)
        p = np.clip(p, 0, None)
    return float(p.sum() / (r.sum() + 1e-12))


def common_norm(ref, pred, mask):
    v = np.concatenate([ref[mask].astype(np.float64), pred[mask].astype(np.float64)])
    lo, hi = np.percentile(v, 0.5), np.percentile(v, 99.5)
    return (np.clip((ref - lo) / (hi - lo + 1e-12), 0, 1), np.clip((pred - lo) / (hi - lo + 1e-12), 0, 1))


def peak_error_deg(ref, pred, mask):
    nside = int(round(np.sqrt(ref.size / 12)))
    p1 = int(np.argmax(np.where(mask, ref, -np.inf)))
    p2 = int(np.argmax(np.where(mask, pred, -np.inf)))
    th, ph = pix2ang_nest(nside, np.array([p1, p2]))
    v = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    return float(np.degrees(np.arccos(np.clip(np.dot(v[0], v[1]), -1, 1))))


def evaluate(original, recon, band_lat_deg=BAND_LAT_DEG, xsize=SSIM_XSIZE, ysize=SSIM_YSIZE):
    original = np.asarray(original)
    recon = np.asarray(recon)
    if original.shape != recon.shape:
        raise ValueError(f"shape mismatch: original {original.shape}, recon {recon.shape}")
    nside = int(round(np.sqrt(original.size / 12)))
    if 12 * nside * nside != original.size:
        raise ValueError(f"{original.size} is not a HEALPix pixel count")
    valid = np.isfinite(original) & np.isfinite(recon)
    theta, _ = pix2ang_nest(nside, np.arange(original.size))
    lat = 90.0 - np.degrees(theta)
    band = valid & (np.abs(lat) <= band_lat_deg)
    a, b, rec = calibrate(original, recon, band)
    out = {"nside": nside, "calibration": {"a": a, "b": b}}
    for name, mask, (l0, l1) in (("full", valid, (-90.0, 90.0)), ("band", band, (-band_lat_deg, band_lat_deg))):
        rn, pn = common_norm(original, rec, mask)
        nr, rmse = nrmse(original, rec, mask)
        out[name] = {
            "pearson": pearson(original, rec, mask),
            "nrmse": nr,
            "rmse": rmse,
            "ssim": ssim(np.nan_to_num(cart_project(rn, l0, l1, xsize, ysize), nan=0.0),
                         np.nan_to_num(cart_project(pn, l0, l1, xsize, ysize), nan=0.0)),
            "peak_error_deg": peak_error_deg(original, rec, mask),
            "cosine": cosine(original, rec, mask),
            "flux_ratio": flux_ratio(original, rec, mask),
            "flux_ratio_positive": flux_ratio(original, rec, mask, True),
        }
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--original", required=True, help="float32 NESTED sky map")
    ap.add_argument("--recon", required=True, help="float32 NESTED reconstruction")
    ap.add_argument("--band_lat_deg", type=float, default=BAND_LAT_DEG)
    args = ap.parse_args(argv)
    res = evaluate(np.fromfile(args.original, np.float32), np.fromfile(args.recon, np.float32), args.band_lat_deg)
    print(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
