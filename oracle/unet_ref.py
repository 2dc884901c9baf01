"""CPU FP64 restatement of the U-Net-lite deconvolver (test infrastructure
only: the checker of the gfx950 kernels in viss_optimize_ws_amd/csrc/vw_unet.hip,
never called by the product path).

Same network, parameter layout and loss as include/vw_unet.h: fourteen 3x3
convolutions with ReLU (levels of 16, 32, 64, 128 channels, 2x2 average
pooling down, nearest upsampling and an additive encoder skip up), a 1x1
output convolution added to the input; loss = mean of (out - y)^2 over the
pixels inside the unit disk (pixel centres (2j + 1 - N) / N) and the batch.
The reference repository has no learned part, so nothing here follows a
reference file; the backward pass is pinned by central finite differences
(tests/test_unet_oracle.py).
"""
import numpy as np

# (ci, co, level) of the 14 3x3 layers in forward order, then the 1x1 output
SPEC = [(1, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (64, 128, 3), (128, 128, 3),
        (128, 64, 2), (64, 64, 2), (64, 32, 1), (32, 32, 1), (32, 16, 0), (16, 16, 0), (16, 1, 0)]


def param_count():
    return sum(co * ci * (9 if k < 14 else 1) + co for k, (ci, co, _) in enumerate(SPEC))


def unpack(theta):
    """parameter vector -> [(W, b)]: W [co][ci][3][3] (1x1: [1][16])"""
    out, off = [], 0
    for k, (ci, co, _) in enumerate(SPEC):
        ks = 9 if k < 14 else 1
        W = theta[off:off + co * ci * ks].reshape((co, ci, 3, 3) if ks == 9 else (co, ci))
        off += co * ci * ks
        b = theta[off:off + co]
        off += co
        out.append((W, b))
    return out


def disk_mask(n):
    j = np.arange(n)
    a = (2 * j + 1 - n)[None, :]
    b = (2 * j + 1 - n)[:, None]
    return (a * a + b * b < n * n).astype(np.float64)


def _conv(x, W, b):
    n, ci, H, _ = x.shape
    p = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    out = np.zeros((n, W.shape[0], H, H))
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("oi,nihw->nohw", W[:, :, ky, kx], p[:, :, ky:ky + H, kx:kx + H])
    return out + b[None, :, None, None]


def _conv_bwd(x, W, dz):
    """gradients of a 3x3 layer: (dW, db, dx) from the pre-activation gradient"""
    H = x.shape[2]
    p = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    dW = np.zeros_like(W)
    dp = np.zeros_like(p)
    for ky in range(3):
        for kx in range(3):
            dW[:, :, ky, kx] = np.einsum("nohw,nihw->oi", dz, p[:, :, ky:ky + H, kx:kx + H])
            dp[:, :, ky:ky + H, kx:kx + H] += np.einsum("oi,nohw->nihw", W[:, :, ky, kx], dz)
    return dW, dz.sum(axis=(0, 2, 3)), dp[:, :, 1:-1, 1:-1]


def _pool(x):
    return 0.25 * (x[:, :, 0::2, 0::2] + x[:, :, 0::2, 1::2] + x[:, :, 1::2, 0::2] + x[:, :, 1::2, 1::2])


def _pool_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=2), 2, axis=3)


def _up(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def _up_bwd(dy):
    return dy[:, :, 0::2, 0::2] + dy[:, :, 0::2, 1::2] + dy[:, :, 1::2, 0::2] + dy[:, :, 1::2, 1::2]


def forward(theta, x):
    """x [n][N][N] -> (out [n][N][N], activations)"""
    P = unpack(np.asarray(theta, np.float64))
    a = {"x0": np.asarray(x, np.float64)[:, None]}

    def cr(k, src, dst):
        a[dst + "_z"] = _conv(a[src], *P[k])
        a[dst] = np.maximum(a[dst + "_z"], 0.0)

    cr(0, "x0", "e0a"); cr(1, "e0a", "e0"); a["p1"] = _pool(a["e0"])
    cr(2, "p1", "e1a"); cr(3, "e1a", "e1"); a["p2"] = _pool(a["e1"])
    cr(4, "p2", "e2a"); cr(5, "e2a", "e2"); a["p3"] = _pool(a["e2"])
    cr(6, "p3", "e3a"); cr(7, "e3a", "e3")
    a["u2"] = _up(a["e3"]); cr(8, "u2", "d2a"); a["s2"] = a["d2a"] + a["e2"]; cr(9, "s2", "d2")
    a["u1"] = _up(a["d2"]); cr(10, "u1", "d1a"); a["s1"] = a["d1a"] + a["e1"]; cr(11, "s1", "d1")
    a["u0"] = _up(a["d1"]); cr(12, "u0", "d0a"); a["s0"] = a["d0a"] + a["e0"]; cr(13, "s0", "d0")
    W, b = P[14]
    out = np.einsum("c,nchw->nhw", W[0], a["d0"]) + b[0] + a["x0"][:, 0]
    return out, a


def loss_grad(theta, x, y):
    """(loss, gradient vector) of the masked mean squared error"""
    theta = np.asarray(theta, np.float64)
    P = unpack(theta)
    out, a = forward(theta, x)
    n, N = out.shape[0], out.shape[1]
    m = disk_mask(N)
    cnt = n * m.sum()
    e = (out - np.asarray(y, np.float64)) * m
    loss = float((e * e).sum() / cnt)
    dout = 2.0 * e / cnt
    grads = [None] * 15
    W, b = P[14]
    grads[14] = (np.einsum("nhw,nchw->c", dout, a["d0"])[None, :], np.array([dout.sum()]))
    d = {"d0": W[0][None, :, None, None] * dout[:, None]}

    def cb(k, src, dst, need_dx=True):
        dz = d[dst] * (a[dst + "_z"] > 0)
        dW, db, dx = _conv_bwd(a[src], P[k][0], dz)
        grads[k] = (dW, db)
        return dx

    d["s0"] = cb(13, "s0", "d0"); d["d0a"] = d["s0"]
    d["u0"] = cb(12, "u0", "d0a"); d["d1"] = _up_bwd(d["u0"])
    d["s1"] = cb(11, "s1", "d1"); d["d1a"] = d["s1"]
    d["u1"] = cb(10, "u1", "d1a"); d["d2"] = _up_bwd(d["u1"])
    d["s2"] = cb(9, "s2", "d2"); d["d2a"] = d["s2"]
    d["u2"] = cb(8, "u2", "d2a"); d["e3"] = _up_bwd(d["u2"])
    d["e3a"] = cb(7, "e3a", "e3")
    d["p3"] = cb(6, "p3", "e3a"); d["e2"] = d["s2"] + _pool_bwd(d["p3"])
    d["e2a"] = cb(5, "e2a", "e2")
    d["p2"] = cb(4, "p2", "e2a"); d["e1"] = d["s1"] + _pool_bwd(d["p2"])
    d["e1a"] = cb(3, "e1a", "e1")
    d["p1"] = cb(2, "p1", "e1a"); d["e0"] = d["s0"] + _pool_bwd(d["p1"])
    d["e0a"] = cb(1, "e0a", "e0")
    cb(0, "x0", "e0a")
    g = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])
    return loss, g


def init_params(seed):
    """He-normal 3x3 weights, zero biases and a zero output layer (numpy's
    generator; the device initialisation uses mt19937_64 and differs)"""
    rng = np.random.default_rng(seed)
    parts = []
    for k, (ci, co, _) in enumerate(SPEC):
        if k < 14:
            parts.append(rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), co * ci * 9))
        else:
            parts.append(np.zeros(co * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts)
