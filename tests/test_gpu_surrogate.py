"""The warm-start network's HIP kernels (include/imexlb_nn.h) against the
NumPy restatement (tests/nn_numpy.py): forward output and every weight and
bias gradient of a small encoder-decoder, and a short training run whose
loss falls."""
import numpy as np
import pytest

import nn_numpy as R

pytestmark = pytest.mark.gpu


def _setup(chans=(4, 8, 8), cin=4, cout=3, N=2, H=32, W=64, seed=3):
    import torch
    from imexlb.surrogate import UNet
    net = UNet(cin, cout, list(chans), N, H, W, seed=seed)
    rng = np.random.default_rng(seed + 1)
    for b in net.b:  # non-zero biases exercise the bias paths
        b.copy_(torch.tensor(rng.normal(0, 0.1, b.shape).astype(np.float32), device="cuda"))
    x = rng.normal(0, 1.0, (N, cin, H, W)).astype(np.float32)
    return net, x, list(chans)


def test_forward_and_gradients_match_numpy():
    import torch
    net, x, chans = _setup()
    print("[surrogate] forward/backward vs numpy", flush=True)
    params = net.params_numpy()
    xd = torch.tensor(x, device="cuda")
    out = net.forward(xd).cpu().numpy()
    ref, cache = R.forward(x.astype(np.float64), params, chans)
    assert np.abs(out - ref).max() <= 1e-4 * np.abs(ref).max()
    rng = np.random.default_rng(9)
    G = rng.normal(0, 1.0, ref.shape).astype(np.float32)
    net.backward(torch.tensor(G, device="cuda"))
    grads = R.backward(G.astype(np.float64), params, chans, cache)
    for k, (dw, db) in enumerate(grads):
        gw = net.dw[k].cpu().numpy()
        gb = net.db[k].cpu().numpy()
        assert np.abs(gw - dw).max() <= 1e-3 * max(np.abs(dw).max(), 1e-6), k
        assert np.abs(gb - db).max() <= 1e-3 * max(np.abs(db).max(), 1e-6), k


def test_training_reduces_the_loss():
    import torch
    net, x, chans = _setup(chans=(8, 16), N=4, H=16, W=32)
    rng = np.random.default_rng(4)
    y = np.tanh(x[:, :3] + 0.5 * x[:, 1:2]).astype(np.float32)  # a smooth target of the input
    xd, yd = torch.tensor(x, device="cuda"), torch.tensor(y, device="cuda")
    first = None
    for step in range(150):
        out = net.forward(xd)
        loss = float(net.loss(out, yd).item())
        first = loss if first is None else first
        net.backward(net.grad_out)
        net.adam(3e-3)
    assert loss < 0.3 * first, (first, loss)
