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


def _tap(p, ky, kx, H):
    """shifted view of a padded tensor as a [ci][n*H*H] matrix"""
    n, ci = p.shape[:2]
    return np.ascontiguousarray(p[:, :, ky:ky + H, kx:kx + H].transpose(1, 0, 2, 3)).reshape(ci, n * H * H)


def _conv(x, W, b):
    n, ci, H, _ = x.shape
    p = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    out = np.zeros((W.shape[0], n * H * H))
    for ky in range(3):
        for kx in range(3):
            out += np.ascontiguousarray(W[:, :, ky, kx]) @ _tap(p, ky, kx, H)
    return out.reshape(W.shape[0], n, H, H).transpose(1, 0, 2, 3) + b[None, :, None, None]


def _conv_bwd(x, W, dz, Wd=None):
    """gradients of a 3x3 layer: (dW, db, dx) from the pre-activation
    gradient; the data gradient uses the weights Wd (default W)"""
    Wd = W if Wd is None else Wd
    n, co, H, _ = dz.shape
    p = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    dzm = np.ascontiguousarray(dz.transpose(1, 0, 2, 3)).reshape(co, n * H * H)
    dW = np.zeros_like(W)
    dp = np.zeros((x.shape[1], n, H + 2, H + 2))
    for ky in range(3):
        for kx in range(3):
            dW[:, :, ky, kx] = dzm @ _tap(p, ky, kx, H).T
            dp[:, :, ky:ky + H, kx:kx + H] += (np.ascontiguousarray(Wd[:, :, ky, kx].T) @ dzm).reshape(-1, n, H, H)
    return dW, dz.sum(axis=(0, 2, 3)), dp[:, :, 1:-1, 1:-1].transpose(1, 0, 2, 3)


def _pool(x):
    return 0.25 * (x[:, :, 0::2, 0::2] + x[:, :, 0::2, 1::2] + x[:, :, 1::2, 0::2] + x[:, :, 1::2, 1::2])


def _pool_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=2), 2, axis=3)


def _up(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def _up_bwd(dy):
    return dy[:, :, 0::2, 0::2] + dy[:, :, 0::2, 1::2] + dy[:, :, 1::2, 0::2] + dy[:, :, 1::2, 1::2]


def _f16(v):
    """round to the nearest FP16 value (ties to even), returned as FP64"""
    return np.asarray(v).astype(np.float16).astype(np.float64)


def _ident(v):
    return v


def forward(theta, x, fp16=False):
    """x [n][N][N] -> (out [n][N][N], activations).

    fp16=True follows the storage of the FP16 kernels of vw_unet.hip: the
    input, every activation and the output are rounded to FP16 where the
    kernels store them (conv output after bias and ReLU, pooling average, skip
    sum, output layer), and the 3x3 layers with a multiple of 16 input
    channels (all but the first, which the device runs on the vector ALU with
    FP32 weights) use FP16 weights; sums are exact (FP64) here and FP32 on the
    device."""
    q = _f16 if fp16 else _ident
    P = unpack(np.asarray(theta, np.float64))
    a = {"x0": q(np.asarray(x, np.float64)[:, None])}

    def cr(k, src, dst):
        W, b = P[k]
        a[dst] = q(np.maximum(_conv(a[src], q(W) if k > 0 else W, b), 0.0))

    cr(0, "x0", "e0a"); cr(1, "e0a", "e0"); a["p1"] = q(_pool(a["e0"]))
    cr(2, "p1", "e1a"); cr(3, "e1a", "e1"); a["p2"] = q(_pool(a["e1"]))
    cr(4, "p2", "e2a"); cr(5, "e2a", "e2"); a["p3"] = q(_pool(a["e2"]))
    cr(6, "p3", "e3a"); cr(7, "e3a", "e3")
    a["u2"] = _up(a["e3"]); cr(8, "u2", "d2a"); a["s2"] = q(a["d2a"] + a["e2"]); cr(9, "s2", "d2")
    a["u1"] = _up(a["d2"]); cr(10, "u1", "d1a"); a["s1"] = q(a["d1a"] + a["e1"]); cr(11, "s1", "d1")
    a["u0"] = _up(a["d1"]); cr(12, "u0", "d0a"); a["s0"] = q(a["d0a"] + a["e0"]); cr(13, "s0", "d0")
    W, b = P[14]
    out = q(np.einsum("c,nchw->nhw", W[0], a["d0"]) + b[0] + a["x0"][:, 0])
    return out, a


def loss_grad(theta, x, y, fp16=False):
    """(loss, gradient vector) of the masked mean squared error.

    The output gradient is formed per image, dout = 2 (out - y) mask / n, and
    every parameter gradient divided by the disk's pixel count at the end
    (the same value as 2 (out - y) mask / (n * disk pixels) throughout).
    fp16=True rounds the target, dout and every stored gradient tensor to
    FP16 where vw_unet.hip stores them, with FP16 weights in every data
    gradient; the ReLU masks are those of the stored activations (y > 0)."""
    q = _f16 if fp16 else _ident
    theta = np.asarray(theta, np.float64)
    P = unpack(theta)
    out, a = forward(theta, x, fp16)
    n, N = out.shape[0], out.shape[1]
    m = disk_mask(N)
    nmask = m.sum()
    e = (out - q(np.asarray(y, np.float64))) * m
    loss = float((e * e).sum() / (n * nmask))
    dout = q(2.0 * e / n)
    grads = [None] * 15
    W, b = P[14]
    grads[14] = (np.einsum("nhw,nchw->c", dout, a["d0"])[None, :], np.array([dout.sum()]))
    pos = {k: (v > 0) for k, v in a.items()}
    d = {"d0": q(W[0][None, :, None, None] * dout[:, None]) * pos["d0"]}

    def cb(k, src, dst, mask_dx=None):
        dz = d[dst] * pos[dst]
        dW, db, dx = _conv_bwd(a[src], P[k][0], dz, q(P[k][0]))
        grads[k] = (dW, db)
        dx = q(dx)
        return dx * pos[mask_dx] if mask_dx else dx

    d["s0"] = cb(13, "s0", "d0"); d["d0a"] = d["s0"]
    d["u0"] = cb(12, "u0", "d0a"); d["d1"] = q(_up_bwd(d["u0"])) * pos["d1"]
    d["s1"] = cb(11, "s1", "d1"); d["d1a"] = d["s1"]
    d["u1"] = cb(10, "u1", "d1a"); d["d2"] = q(_up_bwd(d["u1"])) * pos["d2"]
    d["s2"] = cb(9, "s2", "d2"); d["d2a"] = d["s2"]
    d["u2"] = cb(8, "u2", "d2a"); d["e3"] = q(_up_bwd(d["u2"])) * pos["e3"]
    d["e3a"] = cb(7, "e3a", "e3", "e3a")
    d["p3"] = cb(6, "p3", "e3a"); d["e2"] = q(d["s2"] + _pool_bwd(d["p3"])) * pos["e2"]
    d["e2a"] = cb(5, "e2a", "e2", "e2a")
    d["p2"] = cb(4, "p2", "e2a"); d["e1"] = q(d["s1"] + _pool_bwd(d["p2"])) * pos["e1"]
    d["e1a"] = cb(3, "e1a", "e1", "e1a")
    d["p1"] = cb(2, "p1", "e1a"); d["e0"] = q(d["s0"] + _pool_bwd(d["p1"])) * pos["e0"]
    d["e0a"] = cb(1, "e0a", "e0", "e0a")
    cb(0, "x0", "e0a")
    g = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads]) / nmask
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
