"""CPU reference of the closure network (test infrastructure).

Plain torch float64 on the CPU: periodic 3x3 convolution as circular padding
+ conv2d, GELU (tanh form) and its derivative from torch autograd, and bf16 rounding at
the points where the GPU path stores bf16 (inputs, weights, hidden
activations, gradient operands of the MFMA products).  Gradients of the
convolutions come from torch autograd, independent of the transposed-weight
formula the GPU path uses.
"""

import torch
import torch.nn.functional as F


def bf(x):
    """Round to bf16 and back (round to nearest even)."""
    return x.to(torch.bfloat16).to(torch.float64)


def conv_periodic(x, w):
    """x [B, Ci, H, W], w [Co, Ci, 3, 3] -> [B, Co, H, W], periodic in H and W."""
    return F.conv2d(F.pad(x, (1, 1, 1, 1), mode="circular"), w)


def conv_grads(x, w, dz):
    """(dx, dw) of sum(conv_periodic(x, w) * dz) by autograd."""
    xr = x.clone().requires_grad_(True)
    wr = w.clone().requires_grad_(True)
    conv_periodic(xr, wr).backward(dz)
    return xr.grad, wr.grad


def gelu(z):
    """GELU, tanh form (the network's activation)."""
    return F.gelu(z, approximate="tanh")


def gelu_grad(z):
    zr = z.clone().requires_grad_(True)
    F.gelu(zr, approximate="tanh").sum().backward()
    return zr.grad


def nhwc_to_nchw(a):
    return a.permute(0, 3, 1, 2).contiguous()


def nchw_to_nhwc(a):
    return a.permute(0, 2, 3, 1).contiguous()


def net_forward(x, ws, bs, emulate=True):
    """x [B, H, W, cin] -> (out [B, H, W, cout], acts, zs) in float64.

    With emulate, values are rounded to bf16 where the GPU path stores bf16."""
    r = bf if emulate else (lambda v: v)
    a = r(nhwc_to_nchw(x.double()))
    acts, zs = [a], []
    nl = len(ws)
    out = None
    for l in range(nl):
        z = conv_periodic(a, r(ws[l].double())) + bs[l].double().view(1, -1, 1, 1)
        if l < nl - 1:
            zs.append(z)
            a = r(gelu(z))
            acts.append(a)
        else:
            out = z
    return nchw_to_nhwc(out), acts, zs


def net_backward(g, acts, zs, ws, emulate=True):
    """Weight and bias gradients for output gradient g [B, H, W, cout]."""
    r = bf if emulate else (lambda v: v)
    nl = len(ws)
    dz = nhwc_to_nchw(g.double())
    gws, gbs = [None] * nl, [None] * nl
    for l in range(nl - 1, -1, -1):
        dzb = r(dz)
        dx, dw = conv_grads(acts[l], r(ws[l].double()), dzb)
        gws[l] = dw
        gbs[l] = dz.sum(dim=(0, 2, 3))
        if l > 0:
            dz = dx * gelu_grad(zs[l - 1])
    return gws, gbs


def brute_conv(x, w):
    """Loop form of the periodic convolution (numpy arrays, tiny sizes)."""
    import numpy as np
    B, Ci, H, W = x.shape
    Co = w.shape[0]
    out = np.zeros((B, Co, H, W))
    for b in range(B):
        for co in range(Co):
            for y in range(H):
                for xx in range(W):
                    s = 0.0
                    for ci in range(Ci):
                        for ky in range(3):
                            for kx in range(3):
                                s += w[co, ci, ky, kx] * x[b, ci, (y + ky - 1) % H, (xx + kx - 1) % W]
                    out[b, co, y, xx] = s
    return out
