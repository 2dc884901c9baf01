"""The learned deconvolver on the GPU (include/vw_unet.h): forward and
gradients of the hand-written kernels against the CPU FP64 oracle
(oracle/unet_ref.py), deterministic training, and what the trained network
gains over the dirty image on held-out skies."""
import numpy as np
import pytest

from oracle import unet_ref as U
from viss_optimize_ws_amd import VissError, deconv
from viss_optimize_ws_amd.deconv import UNet

pytestmark = pytest.mark.gpu


def _params(seed):
    th = U.init_params(seed)
    rng = np.random.default_rng(seed + 1)
    P = U.unpack(th)
    for W, b in P:
        b[:] = rng.normal(0.0, 0.05, b.shape)
    P[14][0][:] = rng.normal(0.0, 0.5, P[14][0].shape)
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in P])


# 16 x 16 images: FP32 and FP16 kernels at levels of 16, 8, 4 and 2 pixels;
# 64 x 64: every 3x3 layer but the first on the matrix cores (sides 64 .. 8),
# and with 6 images the level-0 weight gradients fold 48 chunk partials in
# two groups.  The FP16 kernels are compared with the oracle's FP16 storage
# mode (activations, gradients and the matrix-core weights rounded to FP16
# where the kernels store them; sums FP64 there and FP32 here), so what is
# left is FP32 summation order plus the occasional stored value that rounds
# to the neighbouring FP16 value: measured up to 9e-4 (output) and 4.1e-3
# (the level-3 weight gradient at 64 x 64, sums of 384 terms) of the largest
# entry.  FP32 activations against the FP64 oracle: 2e-5 (gradients 1e-4).
@pytest.mark.parametrize("fp16,tol,N,n", [(False, 2e-5, 16, 3), (True, 5e-3, 16, 3), (True, 5e-3, 64, 6)])
def test_forward_and_gradient_vs_oracle(fp16, tol, N, n):
    rng = np.random.default_rng(21)
    x = rng.normal(size=(n, N, N))
    y = rng.normal(size=(n, N, N))
    th = _params(22)
    with UNet(N, n, fp16=fp16, seed=1) as net:
        assert net.nparam == U.param_count()
        net.set_params(th.astype(np.float32))
        th32 = net.params().astype(np.float64)  # the FP32 parameters the kernels use
        out = net.predict(x.astype(np.float32))
        loss, g = net.loss_grad(x.astype(np.float32), y.astype(np.float32))
    ref, _ = U.forward(th32, x.astype(np.float32), fp16=fp16)
    lref, gref = U.loss_grad(th32, x.astype(np.float32), y.astype(np.float32), fp16=fp16)
    assert np.max(np.abs(out - ref)) <= tol * np.max(np.abs(ref))
    assert abs(loss - lref) <= tol * lref
    # per layer, relative to the layer's largest gradient (FP32: 1e-4)
    gtol = tol if fp16 else 5 * tol
    off = 0
    for k, (ci, co, _) in enumerate(U.SPEC):
        m = co * ci * (9 if k < 14 else 1) + co
        gl, rl = g[off:off + m], gref[off:off + m]
        assert np.max(np.abs(gl - rl)) <= gtol * max(np.max(np.abs(rl)), 1e-30), (k, fp16)
        off += m


def test_training_is_deterministic_and_descends():
    rng = np.random.default_rng(31)
    N = 32
    X = rng.normal(size=(40, N, N)).astype(np.float32)
    Y = (0.5 * X + 0.1 * rng.normal(size=X.shape)).astype(np.float32)
    runs = []
    for _ in range(2):
        with UNet(N, 8, fp16=True, seed=3) as net:
            net.set_data(X, Y)
            losses = net.train(30, 8, 1e-3, seed=5)
            runs.append((net.params(), losses))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    l = runs[0][1]
    assert l[-5:].mean() < 0.5 * l[:5].mean()


def test_errors():
    with pytest.raises(VissError):
        UNet(12, 4)  # not a multiple of 8
    with UNet(16, 2) as net:
        with pytest.raises(VissError):
            net.train(1, 2, 1e-3)  # no data
        with pytest.raises(VissError):
            net.loss_grad(np.zeros((3, 16, 16)), np.zeros((3, 16, 16)))  # above batch_max


def test_network_gains_over_dirty_image():
    """a short training run on pairs from the plain pipeline (Nside 16, 32^2
    images) beats the dirty image under its best scalar gain on held-out
    skies, and the re-simulation gate reports a residual per estimate"""
    r = deconv.run(pairs=768, heldout=64, nside=16, img_n=32, batch=16, epochs=6.0, lr=2e-3, fp16=True, seed=2,
                   gate_samples=8, refine_iters=40)
    print(r)
    assert r["heldout_rel_err_network"]["mean"] < 0.9 * r["heldout_rel_err_dirty_best_gain"]["mean"], r
    assert r["loss_last"] < r["loss_first"]
    g = r["gate"]
    assert g["samples"] == 8
    # refinement towards the measured visibilities lowers the re-simulated residual
    assert g["network_refined"]["residual"]["median"] < g["network_unrefined"]["residual"]["median"]
