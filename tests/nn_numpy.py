"""NumPy restatement of the warm-start network (test infrastructure): the
same encoder-decoder as imexlb.surrogate.UNet, forward and backward, in
float64, used to check the HIP kernels' outputs and gradients, and its BF16
inference form (imexlb.surrogate.UNetBF16).  Its own gradients are pinned
by finite differences and its bf16 rounding by torch's conversion
(tests/test_nn_numpy.py)."""
import numpy as np


def _tap(xp, ky, kx, H, W):
    """The input window of tap (ky, kx) as [N][CI][H*W] (a copy)."""
    N, CI = xp.shape[:2]
    return np.ascontiguousarray(xp[:, :, ky:ky + H, kx:kx + W]).reshape(N, CI, H * W)


def conv3x3(x, w, b, relu):
    """x [N][CI][H][W], w [CO][CI][3][3], b [CO]; padding 1.  Each tap is a
    matrix product w[:, :, ky, kx] @ window (BLAS)."""
    N, CI, H, W = x.shape
    CO = w.shape[0]
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    out = np.zeros((N, CO, H * W))
    for ky in range(3):
        for kx in range(3):
            out += np.matmul(np.ascontiguousarray(w[:, :, ky, kx]), _tap(xp, ky, kx, H, W))
    out = out.reshape(N, CO, H, W)
    if b is not None:
        out += b[None, :, None, None]
    return np.maximum(out, 0.0) if relu else out


def conv3x3_backward(x, w, out, dout, relu):
    """Gradients of conv3x3 (out is its output): (dx, dw, db)."""
    N, CI, H, W = x.shape
    CO = w.shape[0]
    dz = dout * (out > 0) if relu else dout
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    dw = np.zeros_like(w)
    dxp = np.zeros_like(xp)
    dzm = np.ascontiguousarray(dz).reshape(N, CO, H * W)
    dz2 = dzm.transpose(1, 0, 2).reshape(CO, N * H * W)
    for ky in range(3):
        for kx in range(3):
            xt = _tap(xp, ky, kx, H, W)
            dw[:, :, ky, kx] = dz2 @ np.ascontiguousarray(xt.transpose(1, 0, 2)).reshape(CI, N * H * W).T
            wt = np.ascontiguousarray(w[:, :, ky, kx].T)
            dxp[:, :, ky:ky + H, kx:kx + W] += np.matmul(wt, dzm).reshape(N, CI, H, W)
    return dxp[:, :, 1:-1, 1:-1], dw, dz.sum(axis=(0, 2, 3))


def pool2(x):
    N, C, H, W = x.shape
    return x.reshape(N, C, H // 2, 2, W // 2, 2).mean(axis=(3, 5))


def pool2_backward(dout):
    return np.repeat(np.repeat(dout, 2, axis=2), 2, axis=3) * 0.25


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def up2_backward(dout):
    N, C, H, W = dout.shape
    return dout.reshape(N, C, H // 2, 2, W // 2, 2).sum(axis=(3, 5))


def layer_shapes(cin, cout, chans):
    """(CO, CI) of every conv in order: encoder pairs, decoder pairs, output."""
    shapes, prev = [], cin
    for c in chans:
        shapes += [(c, prev), (c, c)]
        prev = c
    for l in range(len(chans) - 2, -1, -1):
        shapes += [(chans[l], chans[l + 1] + chans[l]), (chans[l], chans[l])]
    shapes.append((cout, chans[0]))
    return shapes


def forward(x, params, chans):
    """params: list of (w, b) in layer_shapes order.  Returns (out, cache)."""
    acts = []  # (input, output, relu) of every conv, in order
    it = iter(params)

    def conv(inp, relu=True):
        w, b = next(it)
        o = conv3x3(inp, w, b, relu)
        acts.append((inp, o, relu))
        return o

    enc, h = [], x
    for l in range(len(chans)):
        if l > 0:
            h = pool2(h)
        h = conv(conv(h))
        enc.append(h)
    d = enc[-1]
    for l in range(len(chans) - 2, -1, -1):
        cat = np.concatenate([up2(d), enc[l]], axis=1)
        d = conv(conv(cat))
    out = conv(d, relu=False)
    return out, (acts, enc)


def backward(dout, params, chans, cache, abs_sums=None):
    """Gradients [(dw, db)] in layer order for the loss gradient dout.  With
    a list as abs_sums, it is filled with the sums of the magnitudes of the
    terms of every gradient element [(sum |dz| |in|, sum |dz|)], the scale
    of its rounding error in any summation order."""
    acts, enc = cache
    L = len(chans)
    grads = [None] * len(params)
    if abs_sums is not None:
        abs_sums[:] = [None] * len(params)
    k = len(params) - 1

    def conv_b(g):
        nonlocal k
        inp, o, relu = acts[k]
        dx, dw, db = conv3x3_backward(inp, params[k][0], o, g, relu)
        grads[k] = (dw, db)
        if abs_sums is not None:
            dz = np.abs(g * (o > 0) if relu else g)
            _, dwa, dba = conv3x3_backward(np.abs(inp), params[k][0], None, dz, False)
            abs_sums[k] = (dwa, dba)
        k -= 1
        return dx

    g = conv_b(dout)
    denc = [np.zeros_like(e) for e in enc]
    for l in range(0, L - 1):
        g = conv_b(conv_b(g))  # the decoder pair at level l, innermost last
        cu = chans[l + 1]
        denc[l] += g[:, cu:]
        g = up2_backward(g[:, :cu])
    g = g + denc[L - 1]
    for l in range(L - 1, -1, -1):
        g = conv_b(conv_b(g))
        if l > 0:
            g = pool2_backward(g) + denc[l - 1]
    return grads


# ------------------------------------------------------------- BF16 inference
def bf16(x):
    """Round to bfloat16 through float32, to nearest even (finite values);
    returns float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def pool2_bf16(x):
    """2x2 average of bf16 values in float32, ((a + b) + (c + d)) / 4, rounded."""
    f = np.asarray(x, np.float32)
    s = (f[:, :, 0::2, 0::2] + f[:, :, 0::2, 1::2]) + (f[:, :, 1::2, 0::2] + f[:, :, 1::2, 1::2])
    return bf16(np.float32(0.25) * s)


def pack_weights_bf16(w, cip, cop):
    """[CO][CI][3][3] -> [CIP/8][9][COP][8] of bf16 values (zero-padded)."""
    co, ci = w.shape[:2]
    full = np.zeros((cop, cip, 9))
    full[:co, :ci] = bf16(w).reshape(co, ci, 9)
    return full.reshape(cop, cip // 8, 8, 9).transpose(1, 3, 0, 2).copy()


def forward_bf16(x, params, chans):
    """The network as its BF16 inference computes it: input, weights and
    every hidden activation rounded to bf16; sums, bias and ReLU exact
    (float64 here, FP32 on the GPU); the output not rounded."""
    it = iter(params)

    def conv(inp, relu=True, last=False):
        w, b = next(it)
        o = conv3x3(inp, bf16(w), np.asarray(b, np.float32).astype(np.float64), relu)
        return o if last else bf16(o)

    enc, h = [], bf16(x)
    for l in range(len(chans)):
        if l > 0:
            h = pool2_bf16(h)
        h = conv(conv(h))
        enc.append(h)
    d = enc[-1]
    for l in range(len(chans) - 2, -1, -1):
        d = conv(conv(np.concatenate([up2(d), enc[l]], axis=1)))
    return conv(d, relu=False, last=True)
