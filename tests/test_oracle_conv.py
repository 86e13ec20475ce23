"""Pins the network oracle.  cuDNN 5.1 (the reference's conv arithmetic,
core/marvin.hpp:4639-4662, 5486-5491, 5580-5585) is closed source and absent,
so the oracle is the definition: NCDHW cross-correlation + bias, ReLU, 2^3/2
max pool.  Pinned by a brute-force numpy loop (tiny shapes) for the float64
torch routine, and the float32 C loop nest against float64."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from tdm import netdef


def brute_conv(x, w, b):
    nb, cin, d = x.shape[0], x.shape[1], x.shape[2]
    cout, k = w.shape[0], w.shape[2]
    od = d - k + 1
    out = np.zeros((nb, cout, od, od, od))
    for n in range(nb):
        for o in range(cout):
            for z in range(od):
                for y in range(od):
                    for xx in range(od):
                        out[n, o, z, y, xx] = b[o] + np.sum(w[o] * x[n, :, z:z + k, y:y + k, xx:xx + k])
    return out


def test_fp64_oracle_is_cross_correlation():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 3, 6, 6, 6))
    w = rng.normal(size=(4, 3, 3, 3, 3))
    b = rng.normal(size=4)
    shapes = [("c", 3, 4, 3, 6, 4, None)]
    got = oracle.net_forward_fp64(x[:, :1], {"c.weight": w[:, :1], "c.bias": b}, shapes)[-1]
    want = brute_conv(x[:, :1], w[:, :1], b)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    # an unflipped kernel is not a convolution: flipping it changes the answer
    flipped = brute_conv(x[:, :1], w[:, :1, ::-1, ::-1, ::-1], b)
    assert not np.allclose(got, flipped)


@pytest.mark.parametrize("cin,cout,k,d", [(1, 8, 3, 7), (5, 4, 3, 6), (6, 3, 1, 2)])
def test_c_conv_matches_float64(cin, cout, k, d):
    rng = np.random.default_rng(cin * 100 + cout)
    x = rng.normal(size=(2, cin, d, d, d)).astype(np.float32)
    w = rng.normal(size=(cout, cin, k, k, k)).astype(np.float32)
    b = rng.normal(size=cout).astype(np.float32)
    got = oracle.conv3d(x, w, b, relu=False)
    want = brute_conv(x.astype(np.float64), w.astype(np.float64), b.astype(np.float64))
    assert rel_l2(got, want) < 1e-6
    got_relu = oracle.conv3d(x, w, b, relu=True)
    assert np.array_equal(got_relu, np.maximum(got, 0))


def test_maxpool_matches_reshape_max():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 26, 26, 26)).astype(np.float32)
    got = oracle.maxpool3d(x)
    want = x.reshape(2, 3, 13, 2, 13, 2, 13, 2).max(axis=(3, 5, 7))
    assert np.array_equal(got, want)
    # odd extent: 1 + ceil((5-2)/2) = 3 windows, the last one clipped
    y = rng.normal(size=(1, 1, 5, 5, 5)).astype(np.float32)
    assert oracle.maxpool3d(y).shape == (1, 1, 3, 3, 3)
    assert oracle.maxpool3d(y)[0, 0, 2, 2, 2] == y[0, 0, 4, 4, 4]


def test_full_net_fp32_loop_vs_fp64():
    """One patch through the whole 3DMatch net: float32 loop nest vs float64."""
    w = dict(netdef.he_normal_weights(5))
    rng = np.random.default_rng(2)
    p = (rng.random((1, 30, 30, 30)) < 0.1).astype(np.float32)
    shapes = netdef.layer_shapes()
    f32 = oracle.net_forward_fp32(p, w, shapes)
    f64 = oracle.net_forward_fp64(p, w, shapes)
    assert f32.shape == (1, 512, 1, 1, 1)
    assert rel_l2(f32, f64[-1]) < 1e-5
    assert [o.shape[1:] for o in f64] == [(64, 28, 28, 28), (64, 13, 13, 13), (128, 11, 11, 11),
                                         (128, 9, 9, 9), (256, 7, 7, 7), (256, 5, 5, 5),
                                         (512, 3, 3, 3), (512, 1, 1, 1), (512, 1, 1, 1)]
