"""Learned warm start for D2Q9 cylinder/ellipse cases.

    python -m imexlb.surrogate [--train 2000] [--test 200] [--epochs 30]
    python -m imexlb.surrogate --inference-only

1. Cases: 512x128 channel, a body of the cylinder's area (radius Ny/10) at
   x = 0.2 Nx, a circle or an ellipse of axis ratio U[0.5, 2] turned by
   U[0, 90] degrees; Re log-uniform in [40, 300]; inlet velocity 0.05.
2. Data: every case run by the batched solver (FP16 storage) to a developed
   wake; the target is its (p, u, v) field at one phase of the shedding
   cycle (the lift's upward zero crossing), so the network predicts a state
   that already sheds at the developed amplitude.
3. Network: a 4-level convolutional encoder-decoder (3x3 convolutions with
   ReLU, 2x2 average pooling, nearest upsampling, skip connections), FP32,
   trained with Adam by the HIP kernels of include/imexlb_nn.h and run for
   the held-out cases in BF16 on the matrix cores (UNetBF16).  Input: body
   mask, normalised log Re, and the x and y coordinates.
4. Evaluation on held-out cases: the solver started cold (InitUP) and
   warm (the predicted field through imexlb_batch_init_fields), and warm from
   the training set's mean field (the method without the network); for each
   the number of steps until the drag's running mean stays within 1 % of the
   converged cold-start value, and whether the converged drag and lift match
   the cold run's.

Torch supplies device memory only; every arithmetic operation of the
network is one of the library's kernels.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _lib
from ._lib import lib
from .inputs import Deck, derive

NX, NY = 512, 128           # lattice (nodes 0..NX x 0..NY); the network sees NY x NX
UU = 0.05
RE_RANGE = (40.0, 300.0)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def layer_shapes(cin, cout, chans):
    """(CO, CI) of every convolution: encoder pairs, decoder pairs, output."""
    shapes, prev = [], cin
    for c in chans:
        shapes += [(c, prev), (c, c)]
        prev = c
    for l in range(len(chans) - 2, -1, -1):
        shapes += [(chans[l], chans[l + 1] + chans[l]), (chans[l], chans[l])]
    shapes.append((cout, chans[0]))
    return shapes


class UNet:
    """Encoder-decoder on device buffers; batch N of H x W fields."""

    def __init__(self, cin, cout, chans, N, H, W, seed=0, params=None):
        import torch
        self.t = torch
        self.cin, self.cout, self.chans, self.N, self.H, self.W = cin, cout, list(chans), N, H, W
        L = len(self.chans)
        if H % (1 << (L - 1)) or W % (1 << (L - 1)):
            raise ValueError("H and W must be divisible by 2^(levels-1)")
        self.shapes = layer_shapes(cin, cout, self.chans)
        rng = np.random.default_rng(seed)
        dev = "cuda"
        f32 = torch.float32
        self.w, self.b = [], []
        for k, (co, ci) in enumerate(self.shapes):
            if params is not None:
                w0, b0 = params[k]
            else:
                w0 = rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), (co, ci, 3, 3))
                b0 = np.zeros(co)
            self.w.append(torch.tensor(np.asarray(w0, np.float32), device=dev))
            self.b.append(torch.tensor(np.asarray(b0, np.float32), device=dev))
        z = lambda like: torch.zeros_like(like)
        self.dw = [z(w) for w in self.w]
        self.db = [z(b) for b in self.b]
        self.mw = [z(w) for w in self.w]
        self.vw = [z(w) for w in self.w]
        self.mb = [z(b) for b in self.b]
        self.vb = [z(b) for b in self.b]
        self.wt = [torch.empty((ci, co, 3, 3), dtype=f32, device=dev) for co, ci in self.shapes]
        self.step_count = 0
        # per-level sizes and buffers
        self.hw = [(H >> l, W >> l) for l in range(L)]
        e = lambda c, l: torch.empty((N, c, *self.hw[l]), dtype=f32, device=dev)
        self.level_of = []  # spatial level of every conv
        for l in range(L):
            self.level_of += [l, l]
        for l in range(L - 2, -1, -1):
            self.level_of += [l, l]
        self.level_of.append(0)
        self.out = [e(co, self.level_of[k]) for k, (co, ci) in enumerate(self.shapes)]
        self.dz = [e(co, self.level_of[k]) for k, (co, ci) in enumerate(self.shapes)]
        self.dx = [e(ci, self.level_of[k]) for k, (co, ci) in enumerate(self.shapes)]
        self.pooled = [None] + [e(self.chans[l - 1], l) for l in range(1, L)]
        self.cat = [e(self.chans[l + 1] + self.chans[l], l) for l in range(L - 1)]
        self.denc = [e(self.chans[l], l) for l in range(L - 1)]
        self.dpool = [None] + [e(self.chans[l - 1], l - 1) for l in range(1, L)]
        self.dup = [e(self.chans[l + 1], l + 1) for l in range(L - 1)]
        self.inp = [None] * len(self.shapes)
        self.loss_buf = torch.zeros(1, dtype=f32, device=dev)
        # per-tile weight-gradient sums, added in a fixed order (training repeats bit for bit)
        work = max(lib().imexlb_nn_conv3x3_dw_work(ci, co, *self.hw[self.level_of[k]])
                   for k, (co, ci) in enumerate(self.shapes))
        self.dw_work = torch.empty(int(work), dtype=f32, device=dev)
        self.grad_out = e(cout, 0)

    # ---------------------------------------------------------------- layers
    def _conv(self, k, x, relu=True):
        co, ci = self.shapes[k]
        h, w = self.hw[self.level_of[k]]
        self.inp[k] = x
        _ok(lib().imexlb_nn_conv3x3(_ptr(x), _ptr(self.w[k]), _ptr(self.b[k]), _ptr(self.out[k]), self.N, ci, co, h,
                                    w, 1 if relu else 0, None), "conv3x3")
        return self.out[k]

    def _conv_back(self, k, g, relu=True, need_dx=True):
        co, ci = self.shapes[k]
        h, w = self.hw[self.level_of[k]]
        n = self.N * co * h * w
        if relu:
            _ok(lib().imexlb_nn_relu_bwd(_ptr(self.out[k]), _ptr(g), _ptr(self.dz[k]), n, None), "relu_bwd")
            dz = self.dz[k]
        else:
            dz = g
        _ok(lib().imexlb_nn_conv3x3_dw_ordered(_ptr(self.inp[k]), _ptr(dz), _ptr(self.dw[k]), _ptr(self.dw_work),
                                               self.dw_work.numel(), self.N, ci, co, h, w, None), "conv3x3_dw")
        _ok(lib().imexlb_nn_bias_grad(_ptr(dz), _ptr(self.db[k]), self.N, co, h, w, None), "bias_grad")
        if not need_dx:
            return None
        _ok(lib().imexlb_nn_flip_weights(_ptr(self.w[k]), _ptr(self.wt[k]), ci, co, None), "flip")
        _ok(lib().imexlb_nn_conv3x3(_ptr(dz), _ptr(self.wt[k]), None, _ptr(self.dx[k]), self.N, co, ci, h, w, 0,
                                    None), "conv3x3 (data gradient)")
        return self.dx[k]

    # ------------------------------------------------------------------ passes
    def forward(self, x):
        """x: (N, cin, H, W) float32 device tensor; returns the output buffer."""
        L, ch, N = len(self.chans), self.chans, self.N
        k, h, enc = 0, x, []
        for l in range(L):
            if l > 0:
                hh, ww = self.hw[l]
                _ok(lib().imexlb_nn_pool2(_ptr(enc[l - 1]), _ptr(self.pooled[l]), N, ch[l - 1], hh, ww, None), "pool2")
                h = self.pooled[l]
            h = self._conv(k, h)
            h = self._conv(k + 1, h)
            k += 2
            enc.append(h)
        d = enc[-1]
        for l in range(L - 2, -1, -1):
            hh, ww = self.hw[l + 1]
            ctot = ch[l + 1] + ch[l]
            _ok(lib().imexlb_nn_up2(_ptr(d), _ptr(self.cat[l]), N, ch[l + 1], hh, ww, ctot, 0, None), "up2")
            H0, W0 = self.hw[l]
            _ok(lib().imexlb_nn_copy_channels(_ptr(enc[l]), ch[l], 0, _ptr(self.cat[l]), ctot, ch[l + 1], N, ch[l], H0,
                                              W0, 0, None), "concat")
            d = self._conv(k, self.cat[l])
            d = self._conv(k + 1, d)
            k += 2
        return self._conv(k, d, relu=False)

    def backward(self, g):
        """g: gradient of the loss w.r.t. the output; fills dw, db."""
        L, ch, N = len(self.chans), self.chans, self.N
        for t in self.dw:
            t.zero_()
        k = len(self.shapes) - 1
        g = self._conv_back(k, g, relu=False)
        k -= 1
        for l in range(0, L - 1):
            g = self._conv_back(k, g)
            g = self._conv_back(k - 1, g)
            k -= 2
            ctot, cu = ch[l + 1] + ch[l], ch[l + 1]
            H0, W0 = self.hw[l]
            _ok(lib().imexlb_nn_copy_channels(_ptr(g), ctot, cu, _ptr(self.denc[l]), ch[l], 0, N, ch[l], H0, W0, 0,
                                              None), "split")
            hh, ww = self.hw[l + 1]
            _ok(lib().imexlb_nn_up2_bwd(_ptr(g), _ptr(self.dup[l]), N, cu, hh, ww, ctot, 0, None), "up2_bwd")
            g = self.dup[l]
        for l in range(L - 1, -1, -1):
            g = self._conv_back(k, g)
            g = self._conv_back(k - 1, g, need_dx=(k - 1) > 0)
            k -= 2
            if l > 0:
                hh, ww = self.hw[l]
                _ok(lib().imexlb_nn_pool2_bwd(_ptr(g), _ptr(self.dpool[l]), N, ch[l - 1], hh, ww, None), "pool2_bwd")
                H0, W0 = self.hw[l - 1]
                _ok(lib().imexlb_nn_copy_channels(_ptr(self.denc[l - 1]), ch[l - 1], 0, _ptr(self.dpool[l]), ch[l - 1],
                                                  0, N, ch[l - 1], H0, W0, 1, None), "skip gradient")
                g = self.dpool[l]

    def loss(self, pred, target, with_grad=True):
        self.loss_buf.zero_()
        _ok(lib().imexlb_nn_mse(_ptr(pred), _ptr(target), _ptr(self.grad_out) if with_grad else None, pred.numel(),
                                _ptr(self.loss_buf), None), "mse")
        return self.loss_buf

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for group in ((self.w, self.dw, self.mw, self.vw), (self.b, self.db, self.mb, self.vb)):
            for p, g, m, v in zip(*group):
                _ok(lib().imexlb_nn_adam(_ptr(p), _ptr(g), _ptr(m), _ptr(v), p.numel(), lr, b1, b2, eps,
                                         self.step_count, None), "adam")

    def params_numpy(self):
        return [(w.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)) for w, b in zip(self.w, self.b)]

    def n_params(self):
        return int(sum(w.numel() + b.numel() for w, b in zip(self.w, self.b)))


class UNetBF16:
    """Inference with the weights of a UNet in BF16 on the matrix cores
    (imexlb_nn_conv3x3_bf16): activations NHWC bf16, every encoder output
    written straight into its slice of the decoder's concatenated input.
    Channel counts must be multiples of 16; the input is padded to 16
    channels.  forward() returns (N, cout, H, W) FP32."""

    def __init__(self, net, N):
        import torch
        self.t = torch
        self.net, self.N = net, N
        self.chans, self.cout, self.H, self.W = net.chans, net.cout, net.H, net.W
        if any(c % 16 for c in self.chans):
            raise ValueError("BF16 inference needs channel counts that are multiples of 16")
        L, ch = len(self.chans), self.chans
        self.cinp = 16 * ((net.cin + 15) // 16)
        bf, dev = torch.bfloat16, "cuda"
        self.cip = [self.cinp] + [ci for co, ci in net.shapes[1:]]
        self.cop = [32 * ((co + 31) // 32) for co, ci in net.shapes]
        self.wp = [torch.empty(self.cip[k] * 9 * self.cop[k], dtype=bf, device=dev) for k in range(len(net.shapes))]
        self.pack_weights()
        hw = [(self.H >> l, self.W >> l) for l in range(L)]
        self.hw = hw
        e = lambda c, l: torch.empty((N, *hw[l], c), dtype=bf, device=dev)
        self.x = e(self.cinp, 0)
        self.t1 = [e(ch[l], l) for l in range(L)]
        self.cat = [e(ch[l + 1] + ch[l], l) for l in range(L - 1)]
        self.bottom = e(ch[L - 1], L - 1)
        self.pooled = [None] + [e(ch[l - 1], l) for l in range(1, L)]
        self.dec = [e(ch[l], l) for l in range(L - 1)]
        self.out = torch.empty((N, self.cout, self.H, self.W), dtype=torch.float32, device=dev)

    def pack_weights(self):
        """(Re)pack the network's current FP32 weights."""
        for k, (co, ci) in enumerate(self.net.shapes):
            _ok(lib().imexlb_nn_pack_weights_bf16(_ptr(self.net.w[k]), _ptr(self.wp[k]), ci, co, self.cip[k],
                                                  self.cop[k], None), "pack weights")

    def _conv(self, k, x, out, level, ostride, ooff, relu=True, f32=False):
        co, ci = self.net.shapes[k]
        h, w = self.hw[level]
        _ok(lib().imexlb_nn_conv3x3_bf16(_ptr(x), _ptr(self.wp[k]), _ptr(self.net.b[k]), _ptr(out), self.N,
                                         self.cip[k], co, self.cop[k], h, w, ostride, ooff, 1 if relu else 0,
                                         1 if f32 else 0, None), "conv3x3_bf16")

    def forward(self, x):
        """x: (N, cin, H, W) contiguous FP32 device tensor."""
        if tuple(x.shape) != (self.N, self.net.cin, self.H, self.W) or x.dtype != self.t.float32 \
                or not x.is_contiguous():
            raise ValueError(f"input must be a contiguous float32 ({self.N}, {self.net.cin}, {self.H}, {self.W})")
        L, ch, N = len(self.chans), self.chans, self.N
        _ok(lib().imexlb_nn_pack_input_bf16(_ptr(x), _ptr(self.x), N, self.net.cin, self.cinp, self.H, self.W, None),
            "pack input")
        k = 0
        for l in range(L):
            src = self.x
            if l > 0:
                hh, ww = self.hw[l]
                _ok(lib().imexlb_nn_pool2_bf16(_ptr(self.cat[l - 1]), _ptr(self.pooled[l]), N, ch[l - 1], hh, ww,
                                               ch[l] + ch[l - 1], ch[l], None), "pool2_bf16")
                src = self.pooled[l]
            self._conv(k, src, self.t1[l], l, ch[l], 0)
            if l < L - 1:
                self._conv(k + 1, self.t1[l], self.cat[l], l, ch[l + 1] + ch[l], ch[l + 1])
            else:
                self._conv(k + 1, self.t1[l], self.bottom, l, ch[l], 0)
            k += 2
        d = self.bottom
        for l in range(L - 2, -1, -1):
            hh, ww = self.hw[l + 1]
            _ok(lib().imexlb_nn_up2_bf16(_ptr(d), _ptr(self.cat[l]), N, ch[l + 1], hh, ww, ch[l + 1] + ch[l], 0,
                                         None), "up2_bf16")
            self._conv(k, self.cat[l], self.t1[l], l, ch[l], 0)
            self._conv(k + 1, self.t1[l], self.dec[l], l, ch[l], 0)
            d = self.dec[l]
            k += 2
        self._conv(k, d, self.out, 0, 0, 0, relu=False, f32=True)
        return self.out


# --------------------------------------------------------------------- cases
def make_cases(n, seed):
    """Seeded cases: list of (Deck, aspect, angle_deg)."""
    rng = np.random.default_rng(seed)
    out = []
    radius = NY / 10.0
    D = 2.0 * radius
    for _ in range(n):
        re = float(np.exp(rng.uniform(np.log(RE_RANGE[0]), np.log(RE_RANGE[1]))))
        nu = UU * D / re
        ome = 1.0 / (3.0 * nu + 0.5)
        ellipse = rng.random() < 0.5
        aspect = float(rng.uniform(0.5, 2.0)) if ellipse else 1.0
        angle = float(rng.uniform(0.0, 90.0)) if ellipse else 0.0
        out.append((Deck(2, (NX, NY, 0), D / NX, 0.2, re, ome, 1.0, 1.0, 1.0, 0.01), aspect, angle))
    return out


def _batch(cases, precision):
    from .ensemble import Batch
    return Batch([c[0] for c in cases], precision=precision, shapes=[(c[1], c[2]) for c in cases])


def solid_masks(cases):
    """Body masks (1 = solid) on the NY x NX network grid, from the library's geometry."""
    from .lattice import Lattice
    out = np.zeros((len(cases), NY, NX), np.float32)
    for i, (dk, a, ang) in enumerate(cases):
        d = derive(dk)
        lat = Lattice(2, NX, NY, 0, device=-1)
        lat.set_params(dk.ome, dk.rho0, d.uu)
        lat.set_geometry(d.pos, d.radius, a, ang)
        geo = lat.geometry_geo().reshape(NY + 7, NX + 7)[3:3 + NY, 3:3 + NX]
        out[i] = (geo == 0).astype(np.float32)
        lat.close()
    return out


def norm_log_re(cases):
    """log Re of every case mapped to [-1, 1] over RE_RANGE (float32)"""
    lre = np.log([c[0].Re for c in cases])
    mid, half = np.log(RE_RANGE).mean(), np.ptp(np.log(RE_RANGE)) / 2
    return ((lre - mid) / half).astype(np.float32)


def coords():
    """x and y input channels (2, NY, NX), each 0..1 across the grid"""
    c = np.zeros((2, NY, NX), np.float32)
    c[0] = np.linspace(0.0, 1.0, NX, dtype=np.float32)[None, :]
    c[1] = np.linspace(0.0, 1.0, NY, dtype=np.float32)[:, None]
    return c


def inputs(cases, masks):
    """Network input (M, 4, NY, NX): mask, normalised log Re, x, y."""
    M = len(cases)
    x = np.zeros((M, 4, NY, NX), np.float32)
    x[:, 0] = masks
    x[:, 1] = norm_log_re(cases)[:, None, None]
    x[:, 2:4] = coords()[None]
    return x


def fill_inputs(src, cases, out, coords_dev):
    """The network input of `cases` built on the GPU in out (M, 4, NY, NX):
    the body masks from the solver's own geometry (src: a Batch of these
    cases or, for one case, a Lattice), the normalised log Re and the
    coordinates (coords_dev: coords() on the device)."""
    import torch
    if hasattr(src, "ncase"):
        src.body_mask_f32(out)
    else:
        src.body_mask_f32(out[0, 0])
    lre = torch.tensor(norm_log_re(cases), device=out.device)
    out[:, 1].copy_(lre[:, None, None].expand(len(cases), NY, NX))
    out[:, 2:4].copy_(coords_dev[None].expand(len(cases), 2, NY, NX))
    src.synchronize()  # the solver's stream wrote the masks


def developed_fields(cases, steps, r_vec, precision, probe_every=20, max_probe_steps=6000, log=None):
    """Run every case (batched) for `steps`, then capture each case's field
    at the same phase of its shedding cycle: the first probe (every
    `probe_every` steps) at which its lift turns from negative to
    non-negative.  Cases that do not shed within `max_probe_steps` are taken
    at the end.  Returns (M, 3, NY, NX) float32 of (p / uu^2, u / uu,
    v / uu) and the capture step of every case."""
    b = _batch(cases, precision)
    b.init(r_vec)
    b.advance(steps)
    M = len(cases)
    out = np.zeros((M, 3, NY, NX), np.float32)
    when = np.full(M, -1)
    n = (NX + 1) * (NY + 1)

    def grab(i, t):
        p, u = b.download_fields(i)
        out[i, 0] = p.reshape(NY + 1, NX + 1)[:NY, :NX] / UU ** 2
        out[i, 1] = u[:n].reshape(NY + 1, NX + 1)[:NY, :NX] / UU
        out[i, 2] = u[n:].reshape(NY + 1, NX + 1)[:NY, :NX] / UU
        when[i] = t

    prev = b.forces()[:, 1]
    t = steps
    while t < steps + max_probe_steps and (when < 0).any():
        b.advance(probe_every)
        t += probe_every
        lift = b.forces()[:, 1]
        for i in np.nonzero((when < 0) & (prev < 0) & (lift >= 0))[0]:
            grab(i, t)
        prev = lift
    if log:
        log(f"[surrogate] {int((when >= 0).sum())} of {M} cases captured at the lift's upward zero crossing")
    for i in np.nonzero(when < 0)[0]:
        grab(i, t)
    b.close()
    return out, when


def to_lattice_fields(pred, r_vec):
    """Network output (3, NY, NX) -> p, u of the (NY+1) x (NX+1) lattice
    (interior layout), the last row and column repeated.  u_x carries the
    same perturbation InitUP gives a cold start (factor 1 + 0.4 (r - 0.5),
    r = r_vec(mod(id, 100)), initialization.f90:273), so both starts seed
    vortex shedding alike."""
    full = np.pad(pred.astype(np.float64), ((0, 0), (0, 1), (0, 1)), mode="edge")
    y, x = np.mgrid[0:NY + 1, 0:NX + 1]
    ref_id = (x + 3) + (NX + 7) * (y + 3)
    r = 1.0 + ((np.asarray(r_vec)[ref_id % 100] - 0.5) * 2.0) * 0.2
    p = full[0] * UU ** 2
    u = np.concatenate([(full[1] * UU * r).ravel(), (full[2] * UU).ravel()])
    return p.ravel(), u


def convergence_step(t, drag, ref, window, tol=0.01):
    """First logged time after which the running mean of drag (over `window`
    steps) stays within tol of ref; None if never."""
    w = max(1, int(round(window / (t[1] - t[0]))))
    if len(drag) < w:
        return None
    rm = np.convolve(drag, np.ones(w) / w, mode="valid")
    tt = t[w - 1:]
    ok = np.abs(rm - ref) <= tol * abs(ref)
    bad = np.nonzero(~ok)[0]
    if bad.size == 0:
        return int(tt[0])
    if bad[-1] + 1 >= len(ok):
        return None
    return int(tt[bad[-1] + 1])


def warm_start(b, cases, engine, coords_dev, r_vec):
    """Start the batch `b` of `cases` from the network's prediction, all on
    the GPU: inputs from the solver's geometry, BF16 inference in batches
    of engine.N, then imexlb_batch_init_fields_f32 (u_x times InitUP's
    perturbation with r_vec, unperturbed with None).  Returns the
    predictions (M, 3, NY, NX) as a device tensor."""
    import torch
    M, NB = len(cases), engine.N
    xin = torch.zeros((-(-M // NB) * NB, 4, NY, NX), device="cuda")
    fill_inputs(b, cases, xin[:M], coords_dev)
    preds = torch.empty((M, 3, NY, NX), device="cuda")
    for s in range(0, M, NB):
        out = engine.forward(xin[s:s + NB])
        k = min(NB, M - s)
        preds[s:s + k].copy_(out[:k])
    torch.cuda.synchronize()  # the batch's stream reads the predictions
    b.init_fields_f32(preds, UU ** 2, UU, r_vec)
    return preds


def run_history(cases, steps, every, precision, start):
    """Batched run of `cases` begun by start(batch) (timed separately);
    drag/lift every `every` steps.  Returns (t, F[steps/every, M, 2],
    seconds stepping, seconds starting, what start returned)."""
    import torch
    b = _batch(cases, precision)
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    got = start(b)
    b.synchronize()
    t_start = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist, ts = [], []
    for s in range(every, steps + 1, every):
        b.advance(every)
        hist.append(b.forces())
        ts.append(s)
    sec = time.perf_counter() - t0
    b.close()
    return np.array(ts), np.array(hist), sec, t_start, got


def conv_flops(shapes, H, W, levels):
    """Multiply-add FLOPs (2 per MAC) of one forward pass of one case."""
    L = len(levels)
    return sum(2 * 9 * co * ci * (H >> levels[k]) * (W >> levels[k]) for k, (co, ci) in enumerate(shapes))


def bench_inference(chans, sizes=(40, 8, 1), reps=10):
    """Time FP32 and BF16 inference of the default network (random
    weights) per batch size; one JSON line."""
    import torch
    out = {"grid": [NY, NX], "chans": chans, "runs": []}
    for N in sizes:
        net = UNet(4, 3, chans, N, NY, NX, seed=1)
        eng = UNetBF16(net, N)
        flops = conv_flops(net.shapes, NY, NX, net.level_of)
        x = torch.randn((N, 4, NY, NX), device="cuda")
        row = {"batch": N, "gflop_per_case": round(flops / 1e9, 3)}
        for name, e, r in (("fp32", net, max(1, reps // 3)), ("bf16", eng, reps)):
            e.forward(x)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(r):
                e.forward(x)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / r
            row[name] = {"ms_per_batch": round(dt * 1e3, 4), "ms_per_case": round(dt / N * 1e3, 4),
                         "tflops": round(flops * N / dt / 1e12, 1)}
        a32, b16 = net.forward(x).cpu().numpy(), eng.forward(x).cpu().numpy()
        row["bf16_rel_rms_vs_fp32"] = float(np.sqrt(np.mean((a32 - b16) ** 2)) / np.sqrt(np.mean(a32 ** 2)))
        out["runs"].append(row)
        print(f"[surrogate] inference batch {N}: {row}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def build_parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--data-steps", type=int, default=24000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--chans", default="32,64,128,256")
    ap.add_argument("--infer-batch", type=int, default=40)
    ap.add_argument("--eval-steps", type=int, default=40000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--warm-perturbation", choices=("initup", "none"), default="initup",
                    help="multiply the predicted u_x by InitUP's perturbation (initup) or start from it as is")
    ap.add_argument("--inference-only", action="store_true",
                    help="time FP32 and BF16 inference of the network (random weights) and stop")
    return ap


def pipeline(a, log=lambda m: print(m, file=sys.stderr, flush=True)):
    """Data, training, inference and the cold / warm / mean-field
    evaluation for the options `a` (build_parser()); returns the result
    record (a dict) with every phase's wall time."""
    import torch
    t_all = time.perf_counter()
    prec = _lib.F16_F32
    r_vec = np.random.default_rng(a.seed).random(100)
    warm_rv = r_vec if a.warm_perturbation == "initup" else None
    cases = make_cases(a.train + a.test, a.seed)
    train, test = cases[:a.train], cases[a.train:]
    masks = solid_masks(cases)
    X = inputs(cases, masks)
    log(f"[surrogate] data: {len(cases)} cases, {a.data_steps} steps")
    t0 = time.perf_counter()
    Y, _ = developed_fields(cases, a.data_steps, r_vec, prec, log=log)
    t_data = time.perf_counter() - t0
    finite = np.isfinite(Y).all(axis=(1, 2, 3))
    tr_idx = np.nonzero(finite[:a.train])[0]
    log(f"[surrogate] data in {t_data:.1f} s; {int((~finite).sum())} diverged cases left out")
    Xd = torch.tensor(X, device="cuda")
    Yd = torch.tensor(np.nan_to_num(Y), device="cuda")
    net = UNet(4, 3, [int(c) for c in a.chans.split(",")], a.batch, NY, NX, seed=a.seed)
    xb = torch.empty((a.batch, 4, NY, NX), device="cuda")
    yb = torch.empty((a.batch, 3, NY, NX), device="cuda")
    rng = np.random.default_rng(a.seed + 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps_per_epoch = len(tr_idx) // a.batch
    curve = []
    for ep in range(a.epochs):
        perm = rng.permutation(tr_idx)
        lr = a.lr * (0.5 ** (ep // 5))
        tot = 0.0
        for s in range(steps_per_epoch):
            idx = torch.tensor(perm[s * a.batch:(s + 1) * a.batch], device="cuda")
            xb.copy_(Xd.index_select(0, idx))
            yb.copy_(Yd.index_select(0, idx))
            out = net.forward(xb)
            net.loss(out, yb)
            net.backward(net.grad_out)
            net.adam(lr)
            if s % 25 == 0:
                tot = float(net.loss_buf.item())
        curve.append(tot)
        log(f"[surrogate] epoch {ep + 1}/{a.epochs} loss {tot:.5f}")
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0

    # FP32 and BF16 inference of the held-out cases from the same inputs
    # (time and error only; the warm starts below run BF16 from inputs
    # built on the GPU)
    te_idx = np.arange(a.train, a.train + a.test)
    NB = min(a.infer_batch, a.test)
    coords_dev = torch.tensor(coords(), device="cuda")

    def predict(engine):
        dev = torch.empty((a.test, 3, NY, NX), device="cuda")
        xi = torch.empty((NB, 4, NY, NX), device="cuda")
        engine.forward(xi.zero_())  # first launch outside the timing
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(0, a.test, NB):
            sel = te_idx[s:s + NB]
            sel = np.concatenate([sel, np.full(NB - len(sel), sel[-1])])
            xi.copy_(Xd.index_select(0, torch.tensor(sel, device="cuda")))
            out = engine.forward(xi)
            k = min(NB, a.test - s)
            dev[s:s + k].copy_(out[:k])
        torch.cuda.synchronize()
        return dev.cpu().numpy(), time.perf_counter() - t0

    preds32, t_infer32 = predict(UNet(4, 3, net.chans, NB, NY, NX, params=net.params_numpy()))
    engine = UNetBF16(net, NB)
    _, t_infer = predict(engine)

    # one case alone: batch-1 inference, cold and warm starts (the warm one
    # builds the input from the lattice, infers and sets the fields), and
    # the single-lattice step time
    def time_one(eng):
        x1 = torch.empty((1, 4, NY, NX), device="cuda")
        x1.copy_(Xd[te_idx[0]:te_idx[0] + 1])
        eng.forward(x1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            eng.forward(x1)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / 5

    t_infer_one32 = time_one(UNet(4, 3, net.chans, 1, NY, NX, params=net.params_numpy()))
    one = UNetBF16(net, 1)
    t_infer_one = time_one(one)
    from .lattice import Lattice
    dk, asp, ang = test[0]
    dd = derive(dk)
    lat = Lattice(2, NX, NY, 0, device=0, precision=prec)
    lat.set_params(dk.ome, dk.rho0, dd.uu)
    lat.set_geometry(dd.pos, dd.radius, asp, ang)
    x1 = torch.empty((1, 4, NY, NX), device="cuda")
    starts = {}
    for kind in ("cold", "warm", "cold", "warm"):  # the second of each is timed
        lat.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if kind == "cold":
            lat.init_up(r_vec)
        else:
            fill_inputs(lat, test[:1], x1, coords_dev)
            pred = one.forward(x1)
            torch.cuda.synchronize()
            lat.init_fields_f32(pred[0], UU ** 2, UU, warm_rv)
        lat.init_pdf()
        lat.synchronize()
        starts[kind] = time.perf_counter() - t0
    lat.advance(64)
    lat.synchronize()
    t0 = time.perf_counter()
    lat.advance(4096)
    lat.synchronize()
    t_step_one = (time.perf_counter() - t0) / 4096
    lat.close()

    # cold, warm (network), warm (mean field)
    mean_field = Y[tr_idx].mean(axis=0)
    mean_dev = torch.tensor(np.repeat(mean_field[None], a.test, axis=0), device="cuda")
    every = 100
    tc, Fc, sc, ic, _ = run_history(test, a.eval_steps, every, prec, lambda b: b.init(r_vec))
    tw, Fw, sw, iw, preds_dev = run_history(test, a.eval_steps, every, prec,
                                            lambda b: warm_start(b, test, engine, coords_dev, warm_rv))
    tm, Fm, sm, im, _ = run_history(test, a.eval_steps, every, prec,
                                    lambda b: b.init_fields_f32(mean_dev, UU ** 2, UU, r_vec))
    preds = preds_dev.cpu().numpy()
    inputs_dev = torch.zeros((a.test, 4, NY, NX), device="cuda")
    bchk = _batch(test, prec)
    fill_inputs(bchk, test, inputs_dev, coords_dev)
    bchk.close()
    inputs_match = bool(np.array_equal(inputs_dev.cpu().numpy(), X[te_idx]))
    held = Y[te_idx]
    ok = finite[te_idx]
    err = float(np.sqrt(np.mean((preds[ok] - held[ok]) ** 2)) / np.sqrt(np.mean(held[ok] ** 2)))
    err32 = float(np.sqrt(np.mean((preds32[ok] - held[ok]) ** 2)) / np.sqrt(np.mean(held[ok] ** 2)))
    err_mean = float(np.sqrt(np.mean((mean_field[None] - held[ok]) ** 2)) / np.sqrt(np.mean(held[ok] ** 2)))
    log(f"[surrogate] held-out rel. RMS error: network {err:.3f} (BF16), {err32:.3f} (FP32), "
        f"mean field {err_mean:.3f}")

    window = 2560  # about one shedding period at St 0.2
    tail = max(1, len(tc) // 4)
    res = []
    for i in range(a.test):
        if not ok[i] or not np.isfinite(Fc[:, i]).all():
            continue
        ref = Fc[-tail:, i, 0].mean()
        lift_c = Fc[-tail:, i, 1].std()
        r = {"re": test[i][0].Re, "cold": convergence_step(tc, Fc[:, i, 0], ref, window),
             "warm": convergence_step(tw, Fw[:, i, 0], ref, window),
             "mean": convergence_step(tm, Fm[:, i, 0], ref, window),
             "drag_match": bool(abs(Fw[-tail:, i, 0].mean() - ref) <= 0.01 * abs(ref)),
             # lift amplitude within 10 %, or within 1 % of the drag for cases that barely shed
             "lift_match": bool(abs(Fw[-tail:, i, 1].std() - lift_c) <= max(0.1 * lift_c, 0.01 * abs(ref)))}
        res.append(r)
    valid = [r for r in res if r["cold"] and r["warm"] and r["drag_match"] and r["lift_match"]]
    ratio = [r["warm"] / r["cold"] for r in valid]
    # wall time of one case: its steps at the batched rate per case plus its
    # share of the batched inference and start; and run alone:
    # single-lattice steps plus a batch-1 inference and its start
    step_case = sc / a.eval_steps / a.test
    wall = [(r["warm"] * step_case + iw / a.test) / (r["cold"] * step_case + ic / a.test) for r in valid]
    wall_one = [(r["warm"] * t_step_one + starts["warm"]) / (r["cold"] * t_step_one + starts["cold"])
                for r in valid]
    ratio_m = [r["mean"] / r["cold"] for r in valid if r["mean"]]
    out = {"train_cases": int(len(tr_idx)), "seed": a.seed, "warm_perturbation": a.warm_perturbation,
           "inputs_built_on_gpu_match": inputs_match, "test_cases": a.test, "params": net.n_params(), "chans": a.chans,
           "infer_batch": NB,
           "epochs": a.epochs, "data_s": round(t_data, 1), "train_s": round(t_train, 1),
           "infer_s_per_case": round(t_infer / a.test, 6), "infer_fp32_s_per_case": round(t_infer32 / a.test, 6),
           "loss_curve": [round(c, 5) for c in curve],
           "heldout_rel_rms": round(err, 4), "heldout_rel_rms_fp32": round(err32, 4),
           "mean_field_rel_rms": round(err_mean, 4),
           "cases_scored": len(valid), "cases_evaluated": len(res),
           "drag_match": int(sum(r["drag_match"] for r in res)), "lift_match": int(sum(r["lift_match"] for r in res)),
           "median_steps_cold": float(np.median([r["cold"] for r in valid])) if valid else None,
           "median_steps_warm": float(np.median([r["warm"] for r in valid])) if valid else None,
           "median_ratio_warm_over_cold": float(np.median(ratio)) if ratio else None,
           "frac_within_one_third": float(np.mean(np.array(ratio) <= 1.0 / 3.0)) if ratio else None,
           "median_ratio_meanfield_over_cold": float(np.median(ratio_m)) if ratio_m else None,
           "median_wall_ratio_warm_over_cold": float(np.median(wall)) if wall else None,
           "frac_wall_within_one_third": float(np.mean(np.array(wall) <= 1.0 / 3.0)) if wall else None,
           "median_wall_ratio_single_case": float(np.median(wall_one)) if wall_one else None,
           "single_case": {"step_ms": round(t_step_one * 1e3, 5), "infer_ms": round(t_infer_one * 1e3, 3),
                           "infer_fp32_ms": round(t_infer_one32 * 1e3, 3)},
           "batch_seconds": {"cold": round(sc, 2), "warm": round(sw, 2), "mean_field": round(sm, 2)},
           "start_seconds": {"batch_cold": round(ic, 4), "batch_warm": round(iw, 4),
                             "single_cold": round(starts["cold"], 5), "single_warm": round(starts["warm"], 5)},
           "re_range": list(RE_RANGE), "grid": [NX, NY], "eval_steps": a.eval_steps,
           "pipeline_s": round(time.perf_counter() - t_all, 1)}
    return out


def main(argv=None):
    a = build_parser().parse_args(argv)
    if a.inference_only:
        return bench_inference([int(c) for c in a.chans.split(",")])
    print(json.dumps(pipeline(a)), flush=True)


if __name__ == "__main__":
    main()
