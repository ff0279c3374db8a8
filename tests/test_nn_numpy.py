"""Pins the NumPy restatement of the warm-start network (tests/nn_numpy.py)
that the GPU network is checked against: its analytic gradients equal
central finite differences of its own forward pass, and its pieces obey
closed forms (a 3x3 conv with a centre-tap identity kernel is the
identity; pooling an upsampled field gives it back)."""
import numpy as np
import pytest

import nn_numpy as R


def _net(seed=0, cin=2, cout=2, chans=(2, 3), H=8, W=8, N=2):
    rng = np.random.default_rng(seed)
    params = [(rng.normal(0, 0.5, (co, ci, 3, 3)), rng.normal(0, 0.1, co))
              for co, ci in R.layer_shapes(cin, cout, list(chans))]
    x = rng.normal(0, 1.0, (N, cin, H, W))
    return params, x, list(chans)


def test_conv_identity_and_pool_of_upsample():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 6, 4))
    w = np.zeros((3, 3, 3, 3))
    for c in range(3):
        w[c, c, 1, 1] = 1.0
    assert np.allclose(R.conv3x3(x, w, None, False), x)
    assert np.allclose(R.pool2(R.up2(x)), x)


def test_gradients_match_finite_differences():
    params, x, chans = _net()
    rng = np.random.default_rng(7)
    out, cache = R.forward(x, params, chans)
    G = rng.normal(size=out.shape)  # loss = sum(G * out)
    grads = R.backward(G, params, chans, cache)
    eps = 1e-6
    checked = 0
    for li in (0, 3, 5, len(params) - 1):
        w, b = params[li]
        for idx in [(0, 0, 1, 1), (w.shape[0] - 1, w.shape[1] - 1, 0, 2)]:
            keep = w[idx]
            w[idx] = keep + eps
            lp = np.sum(G * R.forward(x, params, chans)[0])
            w[idx] = keep - eps
            lm = np.sum(G * R.forward(x, params, chans)[0])
            w[idx] = keep
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grads[li][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (li, idx, fd, grads[li][0][idx])
            checked += 1
        keep = b[0]
        b[0] = keep + eps
        lp = np.sum(G * R.forward(x, params, chans)[0])
        b[0] = keep - eps
        lm = np.sum(G * R.forward(x, params, chans)[0])
        b[0] = keep
        assert abs((lp - lm) / (2 * eps) - grads[li][1][0]) <= 1e-6 * max(1.0, abs(lp - lm) / (2 * eps))
    assert checked == 8


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-30, 500), rng.normal(0, 1e30, 500),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 1.001953125, 3.0e38]]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = R.bf16(x)
    assert np.array_equal(got, ref)
    # ties go to the even mantissa: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert R.bf16(np.float32(1.00390625)) == 1.0
    assert R.bf16(np.float32(1.01171875)) == 1.015625


def test_bf16_pieces():
    rng = np.random.default_rng(12)
    x = R.bf16(rng.normal(size=(2, 4, 6, 8)))
    assert np.array_equal(R.pool2_bf16(R.up2(x)), x)
    w = rng.normal(size=(5, 12, 3, 3))
    wp = R.pack_weights_bf16(w, 16, 32).ravel()
    for co, ci, t in [(0, 0, 0), (4, 11, 8), (2, 7, 4), (3, 8, 5)]:
        g, j = divmod(ci, 8)
        assert wp[((g * 9 + t) * 32 + co) * 8 + j] == R.bf16(w[co, ci, t // 3, t % 3])
    assert wp.size == 16 * 9 * 32 and np.count_nonzero(wp) <= 5 * 12 * 9


def test_bf16_forward_is_close_to_float64():
    params, x, chans = _net(chans=(3, 4))
    ref, _ = R.forward(x, params, chans)
    got = R.forward_bf16(x, params, chans)
    err = np.sqrt(np.mean((got - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
    assert 0 < err < 2e-2
