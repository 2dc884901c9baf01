"""The CPU FP64 U-Net oracle (oracle/unet_ref.py) pinned by the mathematics:
central finite differences of its loss against its backward pass, adjoint
pairs for pooling and upsampling, the identity of the untrained network, and
the grid <-> HEALPix pixel map of the training pairs (no GPU)."""
import numpy as np
import pytest

from oracle import unet_ref as U
from viss_optimize_ws_amd import deconv, quality


def _rand_params(seed, scale_out=0.5):
    th = U.init_params(seed)
    rng = np.random.default_rng(seed + 1)
    # non-zero output layer and biases so every layer receives gradient
    P = U.unpack(th)
    for W, b in P:
        b[:] = rng.normal(0.0, 0.05, b.shape)
    P[14][0][:] = rng.normal(0.0, scale_out, P[14][0].shape)
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in P])


def test_param_layout():
    assert U.param_count() == 14 * 0 + sum(co * ci * 9 + co for ci, co, _ in U.SPEC[:14]) + 16 + 1
    th = np.arange(U.param_count(), dtype=np.float64)
    back = np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in U.unpack(th)])
    assert np.array_equal(back, th)


def test_untrained_network_is_identity():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 16, 16))
    out, _ = U.forward(U.init_params(5), x)
    assert np.array_equal(out, x)


def test_pool_up_adjoint():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, 3, 8, 8))
    y = rng.normal(size=(2, 3, 4, 4))
    assert np.isclose(np.sum(U._pool(x) * y), np.sum(x * U._pool_bwd(y)))
    assert np.isclose(np.sum(U._up(y) * x), np.sum(y * U._up_bwd(x)))


def test_conv_backward_adjoint():
    """<conv(x), dz> - bias term = <x, dx> and <W, dW> for the linear map"""
    rng = np.random.default_rng(6)
    x = rng.normal(size=(2, 3, 6, 6))
    W = rng.normal(size=(4, 3, 3, 3))
    dz = rng.normal(size=(2, 4, 6, 6))
    y = U._conv(x, W, np.zeros(4))
    dW, db, dx = U._conv_bwd(x, W, dz)
    assert np.isclose(np.sum(y * dz), np.sum(x * dx))
    assert np.isclose(np.sum(y * dz), np.sum(W * dW))
    assert np.allclose(db, dz.sum(axis=(0, 2, 3)))


def test_loss_gradient_finite_differences():
    th = _rand_params(11)
    rng = np.random.default_rng(12)
    x = rng.normal(size=(2, 8, 8))
    y = rng.normal(size=(2, 8, 8))
    loss, g = U.loss_grad(th, x, y)
    # a few parameters of every layer
    picks = []
    off = 0
    for k, (ci, co, _) in enumerate(U.SPEC):
        n = co * ci * (9 if k < 14 else 1) + co
        picks += list(rng.integers(off, off + n, 3))
        off += n
    eps = 1e-6
    for i in picks:
        tp, tm = th.copy(), th.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (U.loss_grad(tp, x, y)[0] - U.loss_grad(tm, x, y)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_grid_pixel_map():
    """every in-disk grid direction lies in the pixel ang2pix returns: its
    angle to that pixel's centre is within the pixel's size"""
    nside, n = 8, 32
    idx = deconv.grid_pixel_index(nside, n)
    j = np.arange(n)
    lg = ((2 * j + 1 - n) / n)[None, :].repeat(n, 0)
    mg = ((2 * j + 1 - n) / n)[:, None].repeat(n, 1)
    inside = lg ** 2 + mg ** 2 < 1
    assert np.array_equal(idx >= 0, inside)
    th, ph = quality.pix2ang_nest(nside, idx[inside])
    # centre direction in the pipeline's convention: l = sin t cos p, m = -sin t sin p, n = cos t
    cl, cm, cn = np.sin(th) * np.cos(ph), -np.sin(th) * np.sin(ph), np.cos(th)
    ng = np.sqrt(1 - lg[inside] ** 2 - mg[inside] ** 2)
    ang = np.arccos(np.clip(cl * lg[inside] + cm * mg[inside] + cn * ng, -1, 1))
    assert ang.max() < 2.0 * np.sqrt(4 * np.pi / (12 * nside * nside))


def test_fp16_storage_mode():
    """the oracle's FP16 storage mode: every stored activation and the output
    are FP16 values, the first layer's FP32 weights and the FP16 weights of
    the others are the ones applied (a forward pass with those weights
    rounded by hand agrees), results stay near the FP64 network (a few
    percent of a layer's largest gradient at 16 x 16: FP16 rounding
    propagated through 30 layers and their ReLU masks), and the rounding is
    really applied"""
    th = _rand_params(13)
    rng = np.random.default_rng(14)
    x = rng.normal(size=(2, 16, 16))
    y = rng.normal(size=(2, 16, 16))
    out16, a16 = U.forward(th, x, fp16=True)
    for k, v in a16.items():
        assert np.array_equal(v, v.astype(np.float16).astype(np.float64)), k
    assert np.array_equal(out16, out16.astype(np.float16).astype(np.float64))
    # FP16 weights for layers 1..13 only: rounding them by hand and running
    # the FP16 mode again changes nothing, rounding layer 0 or 14 does
    P = U.unpack(th.copy())
    th_r = np.concatenate([np.concatenate([(U._f16(W) if 0 < k < 14 else W).ravel(), b.ravel()])
                           for k, (W, b) in enumerate(P)])
    assert np.array_equal(U.forward(th_r, x, fp16=True)[0], out16)
    out64, _ = U.forward(th, x)
    assert 0 < np.max(np.abs(out16 - out64)) <= 2e-2 * np.max(np.abs(out64))
    l64, g64 = U.loss_grad(th, x, y)
    l16, g16 = U.loss_grad(th, x, y, fp16=True)
    assert abs(l16 - l64) <= 1e-2 * l64
    off = 0
    for k, (ci, co, _) in enumerate(U.SPEC):
        m = co * ci * (9 if k < 14 else 1) + co
        a, b = g16[off:off + m], g64[off:off + m]
        assert 0 < np.max(np.abs(a - b)) <= 1e-1 * np.max(np.abs(b)), k
        off += m
