"""The surrogate network of xbeach_gpu_amd.surrogate restated in torch on
the CPU in FP32 (test infrastructure: the checker of the HIP kernels)."""
import torch
import torch.nn.functional as F

from xbeach_gpu_amd.surrogate import param_layout


def unpack(params):
    """torch weights [cout, cin, k, k] and biases from the flat vector
    (weights stored [taps][cout][cin])."""
    lay, _ = param_layout()
    p = torch.as_tensor(params, dtype=torch.float32)
    out = []
    for w, b, cin, cout, taps, relu in lay:
        k = 3 if taps == 9 else 1
        W = p[w:w + taps * cout * cin].reshape(k, k, cout, cin).permute(2, 3, 0, 1).contiguous()
        out.append((W, p[b:b + cout].clone(), taps, relu))
    return out


def forward(layers, x):
    """x [n, 8, H, W] (NCHW) -> y [n, 5, H, W]."""
    def conv(i, t):
        W, b, taps, relu = layers[i]
        y = F.conv2d(t, W, b, padding=1 if taps == 9 else 0)
        return F.relu(y) if relu else y

    x32 = F.pad(x, (0, 0, 0, 0, 0, 32 - x.shape[1]))
    e0 = conv(1, conv(0, x32))
    e1 = conv(3, conv(2, F.avg_pool2d(e0, 2)))
    e2 = conv(5, conv(4, F.avg_pool2d(e1, 2)))
    b3 = conv(7, conv(6, F.avg_pool2d(e2, 2)))
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")  # noqa: E731
    d2 = conv(10, conv(9, torch.cat([up(conv(8, b3)), e2], 1)))
    d1 = conv(13, conv(12, torch.cat([up(conv(11, d2)), e1], 1)))
    d0 = conv(16, conv(15, torch.cat([up(conv(14, d1)), e0], 1)))
    return conv(17, d0)
