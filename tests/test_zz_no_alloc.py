"""imexlb_nn_conv3x3_dw_ordered_noalloc (include/imexlb_nn.h): the ordered
weight gradient with its partial sums in the library's own device array
instead of a caller workspace.  Through the C ABI, two calls in a row (the
second on a smaller shape) equal imexlb_nn_conv3x3_dw_ordered with a
workspace bit for bit, device memory in use does not change across a call,
and a shape needing more partial sums than the array holds is rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def _gradients(L, torch, N, CI, CO, H, W, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn((N, CI, H, W), device="cuda", generator=g)
    dout = torch.randn((N, CO, H, W), device="cuda", generator=g)
    dw0 = torch.randn((CO, CI, 3, 3), device="cuda", generator=g)  # both accumulate into dw
    work_floats = L.imexlb_nn_conv3x3_dw_work(CI, CO, H, W)
    work = torch.empty(work_floats, device="cuda")
    dw_ws, dw_na = dw0.clone(), dw0.clone()
    torch.cuda.synchronize()
    assert L.imexlb_nn_conv3x3_dw_ordered(_p(x), _p(dout), _p(dw_ws), _p(work), work_floats, N, CI, CO, H, W,
                                          None) == 0
    assert L.imexlb_nn_conv3x3_dw_ordered_noalloc(_p(x), _p(dout), _p(dw_na), N, CI, CO, H, W, None) == 0
    torch.cuda.synchronize()
    return dw_ws, dw_na, (x, dout)


def test_noalloc_weight_gradient_equals_workspace_version():
    import torch
    from imexlb._lib import lib
    L = lib()
    # the largest layer of the full-size network (decoder level 0: 96 -> 32
    # channels on 128 x 512, 7.08 M partial sums), then a smaller ragged one
    for k, (N, CI, CO, H, W) in enumerate(((2, 96, 32, 128, 512), (3, 20, 24, 37, 45))):
        print(f"[no_alloc] call {k}: N={N} CI={CI} CO={CO} H={H} W={W}", flush=True)
        dw_ws, dw_na, _ = _gradients(L, torch, N, CI, CO, H, W, seed=100 + k)
        assert torch.isfinite(dw_na).all()
        assert torch.equal(dw_ws, dw_na), (dw_ws - dw_na).abs().max().item()


def test_noalloc_allocates_nothing_and_rejects_oversize():
    import torch
    from imexlb._lib import lib
    L = lib()
    N, CI, CO, H, W = 1, 64, 64, 64, 64
    _, _, (x, dout) = _gradients(L, torch, N, CI, CO, H, W, seed=7)  # first call on this device done
    dw = torch.zeros((CO, CI, 3, 3), device="cuda")
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    assert L.imexlb_nn_conv3x3_dw_ordered_noalloc(_p(x), _p(dout), _p(dw), N, CI, CO, H, W, None) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    # 256 -> 256 channels on 128 x 512: 256 tiles x 589,824 partial sums
    need = L.imexlb_nn_conv3x3_dw_work(256, 256, 128, 512)
    assert need > 8 * 2 ** 20
    big = torch.zeros(1, device="cuda")
    assert L.imexlb_nn_conv3x3_dw_ordered_noalloc(_p(big), _p(big), _p(big), 1, 256, 256, 128, 512, None) == -1
    msg = L.imexlb_last_error().decode()
    assert str(need) in msg and str(8 * 2 ** 20) in msg, msg
    assert L.imexlb_nn_conv3x3_dw_ordered_noalloc(None, _p(big), _p(big), 1, 4, 4, 8, 8, None) == -1
