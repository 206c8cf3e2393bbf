"""Surrogate kernels against torch on the CPU in FP32: each convolution's
forward, data gradient and weight gradient (BF16 operands, FP32
accumulation), the whole network's loss and every parameter gradient, and
training that lowers the loss."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from xbeach_gpu_amd import lib
from xbeach_gpu_amd.surrogate import Surrogate, param_layout

pytestmark = pytest.mark.gpu


def bf(a):
    """round to BF16 (values exactly representable in the kernels)"""
    return torch.as_tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float32)


def dev(t_nhwc_f32):
    return t_nhwc_f32.to(torch.bfloat16).contiguous().cuda()


def close(a, b, rel):
    a = a.double()
    b = b.double()
    err = (a - b).abs().max().item()
    scale = b.abs().max().item()
    return err <= rel * scale, err, scale


@pytest.mark.parametrize("n,H,W,cin,cout,taps,relu", [(2, 8, 64, 64, 32, 9, 1), (1, 6, 128, 32, 64, 9, 0),
                                                      (2, 4, 32, 128, 256, 9, 1), (2, 8, 64, 256, 128, 1, 0),
                                                      (3, 10, 32, 64, 32, 1, 1), (2, 16, 32, 64, 64, 9, 1),
                                                      (1, 12, 64, 32, 32, 9, 0)])
def test_conv_forward_dgrad_wgrad(need_gpu, n, H, W, cin, cout, taps, relu):
    L = lib()
    g = torch.Generator().manual_seed(n * 1000 + H + cin + cout)
    k = 3 if taps == 9 else 1
    X = bf(torch.randn(n, cin, H, W, generator=g))
    Wt = bf(torch.randn(cout, cin, k, k, generator=g) / (k * k * cin) ** 0.5)
    b = torch.randn(cout, generator=g) * 0.1
    dZ = bf(torch.randn(n, cout, H, W, generator=g))
    # device layouts: NHWC activations, weights [taps][cout][cin]
    Xd = dev(X.permute(0, 2, 3, 1))
    Wd = dev(Wt.permute(2, 3, 0, 1).reshape(taps, cout, cin))
    bd = b.float().contiguous().cuda()
    Yd = torch.empty(n, H, W, cout, dtype=torch.bfloat16, device="cuda")
    dZd = dev(dZ.permute(0, 2, 3, 1))
    dXd = torch.empty(n, H, W, cin, dtype=torch.bfloat16, device="cuda")
    dWd = torch.empty(taps, cout, cin, dtype=torch.float32, device="cuda")
    dbd = torch.empty(cout, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert L.xbg_sur_conv_fwd(Xd.data_ptr(), Wd.data_ptr(), bd.data_ptr(), Yd.data_ptr(), n, H, W, cin, cout, taps,
                              relu) == 0
    assert L.xbg_sur_conv_dgrad(dZd.data_ptr(), Wd.data_ptr(), dXd.data_ptr(), n, H, W, cin, cout, taps) == 0
    assert L.xbg_sur_conv_wgrad(Xd.data_ptr(), dZd.data_ptr(), dWd.data_ptr(), dbd.data_ptr(), n, H, W, cin, cout,
                                taps) == 0
    # torch CPU FP32 reference on the same BF16 values
    Xr = X.clone().requires_grad_(True)
    Wr = Wt.clone().requires_grad_(True)
    br = b.float().clone().requires_grad_(True)
    Z = F.conv2d(Xr, Wr, br, padding=k // 2)
    Y = F.relu(Z) if relu else Z
    Z.backward(dZ)
    ok, e, s = close(Yd.float().cpu().permute(0, 3, 1, 2), Y.detach(), 1e-2)   # one BF16 rounding of the output
    assert ok, ("forward", e, s)
    ok, e, s = close(dXd.float().cpu().permute(0, 3, 1, 2), Xr.grad, 1e-2)
    assert ok, ("dgrad", e, s)
    ok, e, s = close(dWd.cpu().reshape(k, k, cout, cin).permute(2, 3, 0, 1), Wr.grad, 1e-5)   # FP32 all the way
    assert ok, ("wgrad", e, s)
    ok, e, s = close(dbd.cpu(), br.grad, 1e-5)
    assert ok, ("bias grad", e, s)


def make_batch(n, H, W, seed):
    """smooth synthetic inputs and targets (NCHW, FP32, BF16-exact)"""
    g = torch.Generator().manual_seed(seed)
    yy, xx = torch.meshgrid(torch.linspace(0, 1, H), torch.linspace(0, 1, W), indexing="ij")
    x = torch.zeros(n, 8, H, W)
    for c in range(7):
        a = torch.randn(n, 1, 1, generator=g)
        f = torch.rand(n, 1, 1, generator=g) * 6
        x[:, c] = a * torch.sin(f * xx + c) * torch.cos(f * yy)
    x = bf(x)
    t = torch.stack([x[:, 0] * x[:, 1], torch.tanh(x[:, 2]), x[:, 3] ** 2 - 0.5, x[:, 4] + x[:, 5], x[:, 6]], 1)
    return x, bf(t)


def test_network_loss_and_gradients_match_cpu(need_gpu):
    from sur_reference import forward, unpack

    n, H, W = 2, 16, 256
    s = Surrogate(n, H, W, seed=3)
    P = s.get_params()
    x, t = make_batch(n, H, W, 11)
    xd = dev(x.permute(0, 2, 3, 1))
    td = dev(t.permute(0, 2, 3, 1))
    torch.cuda.synchronize()
    loss = s.backward(xd.data_ptr(), td.data_ptr(), n)
    G = s.grads()
    # CPU FP32 autograd with the BF16 weights the kernels multiply by
    lay, total = param_layout()
    P16 = bf(P).numpy()
    layers = unpack(P16)
    params = []
    for W_, b_, _, _ in layers:
        W_.requires_grad_(True)
        b_.requires_grad_(True)
        params += [W_, b_]
    y = forward(layers, x)
    ref = F.mse_loss(y, t)
    ref.backward()
    assert loss == pytest.approx(ref.item(), rel=2e-2)
    for i, (w, b, cin, cout, taps, relu) in enumerate(lay):
        k = 3 if taps == 9 else 1
        gw = torch.as_tensor(G[w:w + taps * cout * cin]).reshape(k, k, cout, cin).permute(2, 3, 0, 1)
        rw = params[2 * i].grad
        cos = F.cosine_similarity(gw.flatten().double(), rw.flatten().double(), dim=0).item()
        rel = ((gw.double() - rw.double()).norm() / rw.double().norm()).item()
        assert cos > 0.99 and rel < 0.15, (i, cos, rel)
    s.close()


def test_training_lowers_the_loss(need_gpu):
    n, H, W = 8, 32, 256
    s = Surrogate(n, H, W, seed=5)
    x, t = make_batch(n, H, W, 21)
    xd = dev(x.permute(0, 2, 3, 1))
    td = dev(t.permute(0, 2, 3, 1))
    torch.cuda.synchronize()
    first = s.loss(xd.data_ptr(), td.data_ptr(), n)
    for _ in range(60):
        s.train_step(xd.data_ptr(), td.data_ptr(), n, lr=1e-3)
    s.sync()
    last = s.loss(xd.data_ptr(), td.data_ptr(), n)
    assert last < 0.3 * first, (first, last)
    s.close()
