"""Image-quality harness (healpix_ws_10Mhz.ipynb cells 4-6 without healpy):
HEALPix NESTED ang2pix/pix2ang, the plate-carree projection, SSIM and the
masked figures, pinned by closed forms and brute force."""
import numpy as np
import pytest

from viss_optimize_ws_amd import quality as Q

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64])
def test_ang2pix_inverts_pix2ang(nside):
    p = np.arange(12 * nside * nside)
    th, ph = Q.pix2ang_nest(nside, p)
    assert np.array_equal(Q.ang2pix_nest(nside, th, ph), p)


def test_pix2ang_matches_library_directions():
    # the path's own pix2lmn (l = sin t cos p, m = -sin t sin p, n = cos t)
    from oracle.oracle import Oracle
    from viss_optimize_ws_amd import default_config

    cfg = default_config(1, nside=16, n_epochs=8, segs=1, img_n=0)
    with Oracle(cfg) as o:
        l, m, n, _ = o.lmn()
    th, ph = Q.pix2ang_nest(16, np.arange(l.size))
    assert np.abs(np.cos(th) - n).max() < 1e-14
    assert np.abs(np.sin(th) * np.cos(ph) - l).max() < 1e-14
    assert np.abs(-np.sin(th) * np.sin(ph) - m).max() < 1e-14
    th0, ph0 = Q.pix2ang_nest(1, np.array([0]))
    assert abs(np.cos(th0[0]) - 2 / 3) < 1e-15 and abs(ph0[0] - np.pi / 4) < 1e-15


def test_ang2pix_returns_containing_pixel():
    # random directions land in a pixel whose centre is closer than the
    # largest pixel radius, and most often in the nearest centre's pixel
    nside = 32
    rng = np.random.default_rng(2)
    v = rng.normal(size=(20000, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    th, ph = np.arccos(v[:, 2]), np.arctan2(v[:, 1], v[:, 0])
    pix = Q.ang2pix_nest(nside, th, ph)
    cth, cph = Q.pix2ang_nest(nside, pix)
    c = np.stack([np.sin(cth) * np.cos(cph), np.sin(cth) * np.sin(cph), np.cos(cth)], 1)
    ang = np.arccos(np.clip((c * v).sum(1), -1, 1))
    assert ang.max() < 1.4 * np.sqrt(4 * np.pi / (12 * nside * nside))
    allc = np.stack(Q.pix2ang_nest(nside, np.arange(12 * nside * nside)), 1)
    C = np.stack([np.sin(allc[:, 0]) * np.cos(allc[:, 1]), np.sin(allc[:, 0]) * np.sin(allc[:, 1]),
                  np.cos(allc[:, 0])], 1)
    nearest = np.argmax(v[:2000] @ C.T, axis=1)
    assert np.mean(nearest == pix[:2000]) > 0.85


def _ssim_brute(x, y, k1=0.01, k2=0.03, L=1.0):
    vals = []
    for i in range(3, x.shape[0] - 3):
        for j in range(3, x.shape[1] - 3):
            a = x[i - 3:i + 4, j - 3:j + 4].ravel()
            b = y[i - 3:i + 4, j - 3:j + 4].ravel()
            ma, mb = a.mean(), b.mean()
            va, vb = a.var(ddof=1), b.var(ddof=1)
            cab = ((a - ma) * (b - mb)).sum() / 48
            c1, c2 = (k1 * L) ** 2, (k2 * L) ** 2
            vals.append(((2 * ma * mb + c1) * (2 * cab + c2)) / ((ma * ma + mb * mb + c1) * (va + vb + c2)))
    return np.mean(vals)


def test_ssim_brute_force_and_closed_forms():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, (20, 23))
    y = np.clip(x + rng.normal(0, 0.2, x.shape), 0, 1)
    assert abs(Q.ssim(x, y) - _ssim_brute(x, y)) < 1e-12
    assert abs(Q.ssim(x, x) - 1.0) < 1e-12
    assert abs(Q.ssim(x, y) - Q.ssim(y, x)) < 1e-14
    # constant images: only the luminance term remains
    a, b = np.full((10, 10), 0.3), np.full((10, 10), 0.6)
    c1 = 1e-4
    assert abs(Q.ssim(a, b) - (2 * 0.3 * 0.6 + c1) / (0.09 + 0.36 + c1)) < 1e-12


def test_figures_closed_forms():
    nside = 16
    rng = np.random.default_rng(8)
    orig = rng.lognormal(0, 1, 12 * nside * nside).astype(np.float32)
    rec = (0.5 * orig - 2.0).astype(np.float32)  # an exact affine image
    r = Q.evaluate(orig, rec, xsize=200, ysize=60)
    assert abs(r["calibration"]["a"] - 2.0) < 1e-5 and abs(r["calibration"]["b"] - 4.0) < 1e-4
    for k in ("full", "band"):
        assert abs(r[k]["pearson"] - 1) < 1e-9 and r[k]["nrmse"] < 1e-6 and r[k]["peak_error_deg"] < 1e-5
        assert abs(r[k]["ssim"] - 1) < 1e-9 and abs(r[k]["cosine"] - 1) < 1e-9
        assert abs(r[k]["flux_ratio"] - 1) < 1e-6
    noisy = orig + rng.normal(0, 1.0, orig.size).astype(np.float32)
    q = Q.evaluate(orig, noisy, xsize=200, ysize=60)
    assert 0.3 < q["full"]["pearson"] < 1 and 0 < q["full"]["ssim"] < 1
    mask = np.ones(orig.size, bool)
    x, y = orig.astype(np.float64), noisy.astype(np.float64)
    assert abs(Q.pearson(x, y, mask) - np.corrcoef(x, y)[0, 1]) < 1e-9


def test_projection_samples_map():
    # a map equal to its own latitude projects to an image whose rows are the
    # pixel-centre latitudes of the rows' sample points
    nside = 64
    th, _ = Q.pix2ang_nest(nside, np.arange(12 * nside * nside))
    lat = 90 - np.degrees(th)
    img = Q.cart_project(lat, -30, 30, 400, 120)
    row_lat = -30 + (np.arange(120) + 0.5) * 0.5
    assert np.abs(img - row_lat[:, None]).max() < 1.0  # within a pixel size (0.9 deg at Nside 64)
    assert np.all(np.diff(img.mean(1)) > 0)
