"""vw_unet_predict_noalloc (include/vw_unet.h) against vw_unet_predict
through the C-ABI: the same kernels with the staging in the library's static
device array instead of per-call allocations, so the outputs must be equal
bit for bit, on two calls in a row (the second with fewer images, after the
first left its values in the array), and a batch beyond the array's size is
refused (needs a GPU)."""
import numpy as np
import pytest

from viss_optimize_ws_amd import VissError
from viss_optimize_ws_amd.deconv import UNet

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp16", [True, False])
def test_predict_noalloc_equals_predict(fp16):
    rng = np.random.default_rng(51)
    N = 64
    with UNet(N, 8, fp16=fp16, seed=4) as net:
        th = net.params()
        th += rng.normal(0.0, 0.02, th.shape).astype(np.float32)  # a non-identity network
        net.set_params(th)
        for n in (8, 3):  # two calls in a row, the second smaller
            X = rng.normal(size=(n, N, N)).astype(np.float32)
            ref = net.predict(X)
            got = net.predict(X, noalloc=True)
            assert np.array_equal(got, ref), n
            assert not np.array_equal(ref, X)
        # more images than batch_max: processed in batches like predict
        X = rng.normal(size=(11, N, N)).astype(np.float32)
        assert np.array_equal(net.predict(X, noalloc=True), net.predict(X))


def test_predict_noalloc_rejects_batches_beyond_the_array():
    # 80 images of 256 x 256 = 5,242,880 values > 4,194,304
    with UNet(256, 80, fp16=True, seed=1) as net:
        X = np.zeros((80, 256, 256), np.float32)
        with pytest.raises(VissError) as e:
            net.predict(X, noalloc=True)
        assert e.value.code == 4  # VW_ERR_RANGE
        assert np.array_equal(net.predict(X[:64], noalloc=True), net.predict(X[:64]))
