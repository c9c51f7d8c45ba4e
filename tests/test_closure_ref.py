"""Pins of the closure network's CPU reference (no GPU).

The GPU closure tests compare the MFMA kernels with tests/closure_ref.py;
these tests check that reference against loop forms, finite differences and
torch autograd, so a wrong tap order, a missing flip or a dropped GELU'
factor fails here first.
"""

import math

import numpy as np
import pytest
import torch

import closure_ref as R


def test_periodic_conv_matches_loops():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 5, 4))
    w = rng.standard_normal((2, 3, 3, 3))
    got = R.conv_periodic(torch.tensor(x), torch.tensor(w)).numpy()
    assert np.allclose(got, R.brute_conv(x, w), rtol=1e-13, atol=1e-13)


def wt_from_taps(w):
    """The GPU weight preparation's transposed layout, wt[ci][8 - t][co] =
    w[co][ci][t] with t = 3 ky + kx, returned as a (ci, co, 3, 3) kernel."""
    co, ci = w.shape[:2]
    taps = w.reshape(co, ci, 9)
    wt = np.zeros((ci, 9, co))
    for t in range(9):
        wt[:, 8 - t, :] = taps[:, :, t].T
    return wt.transpose(0, 2, 1).reshape(ci, co, 3, 3)


def test_input_gradient_is_conv_with_flipped_transposed_weights():
    rng = np.random.default_rng(1)
    x = torch.tensor(rng.standard_normal((2, 3, 6, 5)))
    w = torch.tensor(rng.standard_normal((4, 3, 3, 3)))
    dz = torch.tensor(rng.standard_normal((2, 4, 6, 5)))
    dx, dw = R.conv_grads(x, w, dz)
    alt = R.conv_periodic(dz, torch.tensor(wt_from_taps(w.numpy())))
    assert torch.allclose(dx, alt, rtol=1e-12, atol=1e-12)
    # weight gradient: dw[co][ci][ky][kx] = sum dz[co](p) x[ci](p + (ky-1, kx-1))
    xn, dzn = x.numpy(), dz.numpy()
    ref = np.zeros(w.shape)
    for ky in range(3):
        for kx in range(3):
            xs = np.roll(xn, shift=(-(ky - 1), -(kx - 1)), axis=(2, 3))
            ref[:, :, ky, kx] = np.einsum('bohw,bihw->oi', dzn, xs)
    assert np.allclose(dw.numpy(), ref, rtol=1e-12, atol=1e-12)


def test_gelu_and_derivative():
    """The tanh-form GELU equals z sigma(s), s = 2 sqrt(2/pi) (z + 0.044715 z^3)
    (the form the kernels evaluate), stays within 1e-3 of the erf form, and
    its derivative matches finite differences and the closed form."""
    z = torch.linspace(-8, 8, 321, dtype=torch.float64)
    s = 2 * math.sqrt(2 / math.pi) * (z + 0.044715 * z ** 3)
    sig = torch.sigmoid(s)
    assert torch.allclose(R.gelu(z), z * sig, rtol=1e-13, atol=1e-15)
    phi = 0.5 * (1 + torch.erf(z / math.sqrt(2)))
    assert float((R.gelu(z) - z * phi).abs().max()) < 1e-3
    h = 1e-6
    fd = (R.gelu(z + h) - R.gelu(z - h)) / (2 * h)
    assert torch.allclose(R.gelu_grad(z), fd, atol=1e-8)
    ds = 2 * math.sqrt(2 / math.pi) * (1 + 3 * 0.044715 * z ** 2)
    closed = sig + z * sig * (1 - sig) * ds
    assert torch.allclose(R.gelu_grad(z), closed, atol=1e-13)


def test_bf16_round_to_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7
    v = torch.tensor([one + ulp / 2, one + 3 * ulp / 2, one + ulp / 2 + 1e-6, -(one + ulp / 2)],
                     dtype=torch.float64)
    assert R.bf(v).tolist() == [one, one + 2 * ulp, one + ulp, -one]


@pytest.mark.parametrize("emulate", [False])
def test_backward_chain_matches_autograd(emulate):
    """Explicit backward of the reference equals autograd of its forward (unrounded)."""
    rng = np.random.default_rng(2)
    chans = [2, 5, 4, 2]
    ws = [torch.tensor(rng.standard_normal((chans[l + 1], chans[l], 3, 3)) * 0.4) for l in range(3)]
    bs = [torch.tensor(rng.standard_normal(chans[l + 1]) * 0.1) for l in range(3)]
    x = torch.tensor(rng.standard_normal((2, 6, 7, 2)))
    g = torch.tensor(rng.standard_normal((2, 6, 7, 2)))
    out, acts, zs = R.net_forward(x, ws, bs, emulate=emulate)
    gws, gbs = R.net_backward(g, acts, zs, ws, emulate=emulate)
    wr = [w.clone().requires_grad_(True) for w in ws]
    br = [b.clone().requires_grad_(True) for b in bs]
    out2, _, _ = R.net_forward(x, wr, br, emulate=False)
    (out2 * g).sum().backward()
    assert torch.allclose(out, out2.detach(), rtol=1e-13)
    for l in range(3):
        assert torch.allclose(gws[l], wr[l].grad, rtol=1e-11, atol=1e-12)
        assert torch.allclose(gbs[l], br[l].grad, rtol=1e-11, atol=1e-12)
