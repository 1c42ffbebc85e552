"""Pins of the CNN oracle (oracle/cnn.py): tap orientation against np.roll,
linearity, bfloat16 rounding, and the gradients against central differences."""
import numpy as np

from oracle import cnn


def test_single_tap_is_a_periodic_shift():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(6, 6, 6, 2))
    for tap in (0, 5, 13, 22, 26):
        W = np.zeros((27, 2, 2))
        W[tap] = np.eye(2)
        off = cnn.OFFSETS[tap]
        ref = np.roll(np.roll(np.roll(a, -off[0], 0), -off[1], 1), -off[2], 2)   # a[n + off]
        assert np.array_equal(cnn.conv(a, W), ref)
    assert cnn.OFFSETS[13] == (0, 0, 0)


def test_conv_is_linear_and_bias_adds():
    rng = np.random.default_rng(1)
    a, b = rng.normal(size=(2, 4, 4, 4, 3))
    W = rng.normal(size=(27, 5, 3))
    bias = rng.normal(size=5)
    assert np.allclose(cnn.conv(2 * a - b, W), 2 * cnn.conv(a, W) - cnn.conv(b, W), atol=1e-12)
    assert np.allclose(cnn.conv(a, W, bias) - cnn.conv(a, W), bias, atol=1e-12)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.140625, 0.0], dtype=np.float32)
    r = cnn.to_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0                 # tie rounds to even
    assert r[2] == 1.0 + 2 ** -7                       # above the tie rounds up
    assert r[3] == -3.140625 and r[4] == 0.0


def test_gradients_against_central_differences():
    rng = np.random.default_rng(2)
    ch = (2, 4, 3, 2)
    layers = [(rng.normal(0, 0.4, (27, co, ci)), rng.normal(0, 0.1, co)) for ci, co in zip(ch[:-1], ch[1:])]
    x = rng.normal(size=(4, 4, 4, ch[0]))
    t = rng.normal(size=(4, 4, 4, ch[-1]))
    loss, grads = cnn.loss_and_grads(layers, x, t, bf16=False)
    eps = 1e-6
    for l, (W, b) in enumerate(layers):
        for idx in [(0, 0, 0), (13, 1, 0), (26, W.shape[1] - 1, W.shape[2] - 1), (7, 0, W.shape[2] - 1)]:
            Wp = W.copy(); Wp[idx] += eps
            Wm = W.copy(); Wm[idx] -= eps
            lp = cnn.loss_and_grads([(Wp, b) if k == l else layers[k] for k in range(len(layers))], x, t, False)[0]
            lm = cnn.loss_and_grads([(Wm, b) if k == l else layers[k] for k in range(len(layers))], x, t, False)[0]
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grads[l][0][idx]) < 1e-6 * max(1.0, abs(fd)), (l, idx)
        bp = b.copy(); bp[0] += eps
        bm = b.copy(); bm[0] -= eps
        lp = cnn.loss_and_grads([(W, bp) if k == l else layers[k] for k in range(len(layers))], x, t, False)[0]
        lm = cnn.loss_and_grads([(W, bm) if k == l else layers[k] for k in range(len(layers))], x, t, False)[0]
        assert abs((lp - lm) / (2 * eps) - grads[l][1][0]) < 1e-6


def test_box_filter_of_constant_and_linear_mode():
    u = np.full((8, 8, 8, 1), 3.0)
    assert np.allclose(cnn.box_filter(u), 3.0)
    # a single Fourier mode is multiplied by the filter's transfer function
    k = 1
    x = np.arange(8) * 2 * np.pi / 8
    v = np.cos(k * x)[:, None, None, None] * np.ones((8, 8, 8, 1))
    got = cnn.box_filter(v)
    h = 0.5 + 0.5 * np.cos(k * 2 * np.pi / 8)
    assert np.allclose(got, h * v[::2, ::2, ::2], atol=1e-12)


def test_forward_at_equals_full_forward():
    """The node-subset forward (used at grids where the full CPU forward is too
    slow) equals the full periodic forward at those nodes, including nodes
    whose neighbourhood wraps around the box."""
    rng = np.random.default_rng(3)
    N, ch = 12, (3, 8, 8, 2)
    layers = [(rng.normal(0, 0.3, (27, co, ci)), rng.normal(0, 0.1, co)) for ci, co in zip(ch[:-1], ch[1:])]
    x = rng.normal(size=(N, N, N, ch[0]))
    y, _ = cnn.forward(layers, x, bf16=True)
    nodes = np.array([[0, 0, 0], [N - 1, 5, 0], [3, 7, 11], [6, 6, 6], [11, 11, 11]])
    ya = cnn.forward_at(layers, x, nodes, bf16=True)
    assert np.max(np.abs(ya - y[nodes[:, 0], nodes[:, 1], nodes[:, 2]])) < 1e-12
