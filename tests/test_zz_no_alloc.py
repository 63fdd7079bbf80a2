"""The no-allocation forward pass (ptx_mlp_forward_noalloc) against the
allocating one (ptx_mlp_forward) with the same parameters: equal bit for bit
on two calls in a row, the second with a smaller batch (the fixed device
array keeps the first call's contents, which must not leak into the second),
and a clear error for a batch beyond the array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = [516, 256, 256, 256, 64]


def say(*a):
    print(*a, flush=True)


def test_noalloc_forward_equals_forward(ptx):
    net = ptx.Mlp(DIMS, max_batch=256, seed=11)
    rng = np.random.default_rng(5)
    params = (rng.standard_normal(net.nparams) * 0.05).astype(np.float32)
    net.set_params(params)
    for batch in (200, 37):
        x = rng.standard_normal((batch, DIMS[0])).astype(np.float32)
        say("batch", batch)
        want = net.forward(x)
        got = ptx.mlp_forward_noalloc(DIMS, params, x)
        assert got.shape == want.shape
        assert np.array_equal(got, want)
    net.close()


def test_noalloc_forward_rejects_a_batch_beyond_the_array(ptx):
    n = sum(DIMS[i] * DIMS[i + 1] + DIMS[i + 1] for i in range(len(DIMS) - 1))
    params = np.zeros(n, np.float32)
    with pytest.raises(ptx.PtxError) as ex:
        ptx.mlp_forward_noalloc(DIMS, params, np.zeros((289, DIMS[0]), np.float32))
    assert ex.value.code == -3 and "scratch" in str(ex.value)
    y = ptx.mlp_forward_noalloc(DIMS, params, np.zeros((288, DIMS[0]), np.float32))
    assert y.shape == (288, DIMS[-1]) and np.allclose(y, np.log(2.0), rtol=1e-6, atol=0)
