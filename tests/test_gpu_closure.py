"""Closure kernels (BF16 MFMA convolutions, packing, box filter) against the
CPU reference of tests/closure_ref.py (GPU).

Tolerances: float32 outputs of a convolution over bf16 operands agree with
the float64 product of the same bf16 values to float32 accumulation error
(relative L2 <= 1e-5).  Through a whole network, a float32 difference of
~1e-6 flips the bf16 rounding of about one stored value in a thousand by one
bf16 ulp (2^-8 relative) per layer; measured output differences are
1e-4 to 3e-4.  The backward pass re-rounds the gradient at every layer, so
the first layer's weight gradient carries the flips of all later layers
(measured 1e-4 to 1.2e-3).  Network outputs are held to relative
L2 <= 1e-3 and gradients to <= 5e-3 against the bf16-emulating reference,
and both to <= 3e-2 against the unrounded float64 network.
"""

import numpy as np
import pytest
import torch

import closure_ref as R
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def case(*args):
    print("case:", *args, flush=True)


def dev():
    from dedalus import _hip
    _hip.init()
    return torch.device("cuda", 0)


def C():
    from dedalus.closure import net
    return net


def rl2(a, b):
    a = a.detach().double().cpu().numpy().ravel()
    b = b.detach().double().cpu().numpy().ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("k", [16, 32])
def test_mfma_tile_layout(k):
    d = dev()
    case("mfma probe", k)
    from dedalus import _hip
    g = torch.Generator().manual_seed(0)
    A = torch.randn(16, k, generator=g).to(torch.bfloat16)
    Bm = torch.randn(k, 16, generator=g).to(torch.bfloat16)
    D = torch.zeros(16, 16, dtype=torch.float32, device=d)
    Ad, Bd = A.to(d), Bm.to(d)
    import ctypes
    _hip._check(_hip._lib.ddx_mfma_probe(None, k, ctypes.c_void_p(Ad.data_ptr()),
                                         ctypes.c_void_p(Bd.data_ptr()), ctypes.c_void_p(D.data_ptr())))
    torch.cuda.synchronize()
    ref = A.double() @ Bm.double()
    assert rl2(D, ref) < 1e-6


# shapes with W % 64 == 0 and H % 4 == 0 and 16/32 channels take the LDS-tiled kernel
CONV_CASES = [(1, 8, 16, 16, 16), (2, 8, 16, 16, 32), (1, 5, 16, 32, 48), (2, 64, 64, 32, 32),
              (1, 16, 8, 48, 64), (3, 4, 4, 16, 16), (2, 8, 64, 16, 32), (1, 4, 128, 32, 16),
              (3, 12, 64, 16, 16), (1, 64, 192, 32, 32)]


def random_conv(B, H, W, Ci, Co, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(B, H, W, Ci, generator=g).to(torch.bfloat16)
    w = (torch.randn(Co, Ci, 3, 3, generator=g) * 0.2).to(torch.bfloat16)
    b = torch.randn(Co, generator=g) * 0.1
    return x, w, b


def taps(w):
    """(co, ci, 3, 3) -> [co, 9, ci] (the kernels' forward weight layout)."""
    return w.reshape(w.shape[0], w.shape[1], 9).permute(0, 2, 1).contiguous()


@pytest.mark.parametrize("B,H,W,Ci,Co", CONV_CASES, ids=str)
@pytest.mark.parametrize("mode", ["linear", "gelu"])
def test_conv_forward(B, H, W, Ci, Co, mode):
    d = dev()
    net = C()
    from dedalus import _hip
    case("conv forward", B, H, W, Ci, Co, mode)
    x, w, b = random_conv(B, H, W, Ci, Co, 1)
    m = _hip.CONV_GELU if mode == "gelu" else _hip.CONV_LINEAR
    of, ob = net.conv3x3(x.reshape(-1, Ci).to(d), taps(w).to(d), B, H, W, bias=b.to(d), mode=m,
                         out_f32=True, out_bf16=True)
    torch.cuda.synchronize()
    z = R.conv_periodic(R.nhwc_to_nchw(x.double()), w.double()) + b.double().view(1, -1, 1, 1)
    z = R.nchw_to_nhwc(z).reshape(-1, Co)
    assert rl2(of, z) < 1e-5
    # bf16 output = round-to-nearest-even of the kernel's own float32 value
    # (linear), or of GELU of it (float32 GELU may differ by a few float32
    # ulps from torch's, so a rare one-ulp bf16 flip is allowed there)
    of_c = of.cpu()
    got_b = ob.cpu()
    if mode == "linear":
        assert torch.equal(got_b, of_c.to(torch.bfloat16))
    else:
        exact = torch.nn.functional.gelu(of_c.double(), approximate="tanh")
        err = (got_b.double() - exact).abs()
        # bf16 ulp of |exact|: 2^(floor(log2|x|) - 7)
        ulp = torch.exp2(torch.floor(torch.log2(exact.abs().clamp_min(1e-30))) - 7)
        big = exact.abs() > 1e-6
        excess = ((err - 0.5 * ulp) / ulp)[big].max().item()
        print("GELU bf16 rounding: max error beyond half an ulp = %.3g ulp" % excess)
        assert bool((err <= 0.51 * ulp + 1e-7).all())


@pytest.mark.parametrize("B,H,W,Ci,Co", CONV_CASES, ids=str)
def test_conv_input_gradient(B, H, W, Ci, Co):
    d = dev()
    net = C()
    from dedalus import _hip
    case("conv input gradient", B, H, W, Ci, Co)
    x, w, _ = random_conv(B, H, W, Ci, Co, 2)
    g = torch.Generator().manual_seed(3)
    dz = torch.randn(B, H, W, Co, generator=g).to(torch.bfloat16)
    zprev = torch.randn(B, H, W, Ci, generator=g)
    wf, wt = net.wprep(w.float().to(d), Co, Ci)
    assert torch.equal(wf.cpu(), taps(w))
    of, _ = net.conv3x3(dz.reshape(-1, Co).to(d), wt, B, H, W, aux=zprev.reshape(-1, Ci).to(d),
                        mode=_hip.CONV_GELU_GRAD)
    torch.cuda.synchronize()
    dx, _ = R.conv_grads(R.nhwc_to_nchw(x.double()), w.double(), R.nhwc_to_nchw(dz.double()))
    ref = R.nchw_to_nhwc(dx * R.gelu_grad(R.nhwc_to_nchw(zprev.double()))).reshape(-1, Ci)
    assert rl2(of, ref) < 1e-5


@pytest.mark.parametrize("B,H,W,Ci,Co", CONV_CASES, ids=str)
@pytest.mark.parametrize("nchunk", [1, 7, 64, 100])
def test_conv_weight_gradient(B, H, W, Ci, Co, nchunk):
    """Weight gradient against the float64 one, and the same bits on a second
    call (the chunk sum adds its partials in a fixed order)."""
    d = dev()
    net = C()
    case("conv weight gradient", B, H, W, Ci, Co, nchunk)
    x, w, _ = random_conv(B, H, W, Ci, Co, 4)
    g = torch.Generator().manual_seed(5)
    dz = torch.randn(B, H, W, Co, generator=g).to(torch.bfloat16)
    xd, dzd = x.reshape(-1, Ci).to(d), dz.reshape(-1, Co).to(d)
    dwp = net.wgrad(xd, dzd, B, H, W, nchunk=nchunk)
    dw = net.wunpack(dwp, Co, Ci)
    again = net.wgrad(xd, dzd, B, H, W, nchunk=nchunk)
    torch.cuda.synchronize()
    _, ref = R.conv_grads(R.nhwc_to_nchw(x.double()), w.double(), R.nhwc_to_nchw(dz.double()))
    assert rl2(dw, ref) < 1e-5
    assert torch.equal(dwp, again)


def test_colsum_pack_unpack():
    d = dev()
    net = C()
    case("colsum / pack / unpack")
    g = torch.Generator().manual_seed(6)
    a = torch.randn(4096 + 16, 48, generator=g)
    ad = a.to(d)
    for nchunk in (None, 1, 7, 4112):
        s = net.colsum(ad, nchunk)
        assert rl2(s, a.double().sum(0)) < 1e-6, nchunk
    b = torch.randn(1 << 16, 32, generator=g)
    assert rl2(net.colsum(b.to(d)), b.double().sum(0)) < 1e-6
    # planar float64 sources with scales -> padded bf16 + f32
    srcs = [torch.randn(32, 48, dtype=torch.float64, generator=g).to(d) for _ in range(3)]
    scale = [0.5, 2.0, -1.0]
    ob, of = net.pack(srcs, 32 * 48, 16, scale=scale, bf16=True, f32=True)
    ref = torch.zeros(32 * 48, 16, dtype=torch.float64)
    for c in range(3):
        ref[:, c] = srcs[c].cpu().reshape(-1) * scale[c]
    assert torch.equal(of.cpu(), ref.float())
    assert torch.equal(ob.cpu(), ref.float().to(torch.bfloat16))
    # interleaved float32 source (pixel stride = channel count)
    inter = torch.randn(100 * 16, 3, generator=g).to(d)
    ob2, of2 = net.pack([(inter.data_ptr() + 4 * c, False) for c in range(3)], 1600, 16,
                        pstride=3, bf16=False, f32=True)
    assert torch.equal(of2[:, :3].cpu(), inter.cpu()) and not of2[:, 3:].any()
    # unpack back to float64 planes with a scale
    outs = [torch.empty(32 * 48, dtype=torch.float64, device=d) for _ in range(3)]
    net.unpack(of, 3, [o.data_ptr() for o in outs], f64=True, scale=[2.0, 0.5, -1.0])
    torch.cuda.synchronize()
    for c in range(3):
        assert torch.equal(outs[c].cpu(), srcs[c].cpu().reshape(-1).float().double())
    # unpack with addends: dst = scale * src + add, float64 and float32
    adds = [torch.randn(32 * 48, dtype=torch.float64, generator=g).to(d) for _ in range(3)]
    net.unpack(of, 3, [o.data_ptr() for o in outs], f64=True, scale=[2.0, 0.5, -1.0],
               add=[x.data_ptr() for x in adds])
    outs32 = [torch.empty(32 * 48, dtype=torch.float32, device=d) for _ in range(2)]
    adds32 = [x.float() for x in adds[:2]]
    net.unpack(of, 2, [o.data_ptr() for o in outs32], f64=False, scale=[2.0, 0.5],
               add=[x.data_ptr() for x in adds32])
    torch.cuda.synchronize()
    for c in range(3):
        expect = (of[:, c].cpu() * [2.0, 0.5, -1.0][c]).double() + adds[c].cpu()
        assert torch.equal(outs[c].cpu(), expect)
    for c in range(2):
        expect = of[:, c].cpu() * [2.0, 0.5][c] + adds32[c].cpu()
        assert torch.equal(outs32[c].cpu(), expect)


@pytest.mark.parametrize("f64", [True, False])
def test_box_filter(f64):
    d = dev()
    from dedalus import _hip
    import ctypes
    case("box filter", f64)
    dt = torch.float64 if f64 else torch.float32
    g = torch.Generator().manual_seed(7)
    src = torch.randn(512, 256, dtype=dt, generator=g)
    dst = torch.empty(64, 32, dtype=dt, device=d)
    sd = src.to(d)
    _hip._check(_hip._lib.ddx_box_filter(None, 512, 256, 8, 1 if f64 else 0,
                                         ctypes.c_void_p(sd.data_ptr()), ctypes.c_void_p(dst.data_ptr())))
    torch.cuda.synchronize()
    ref = src.double().reshape(64, 8, 32, 8).mean(axis=(1, 3))
    assert rl2(dst, ref) < (1e-15 if f64 else 1e-7)


NET_CASES = [((2, 16, 16), (2, 32, 32, 32, 32, 2)), ((1, 64, 64), (3, 32, 32, 32, 32, 2)),
             ((4, 8, 32), (2, 16, 48, 1))]


@pytest.mark.parametrize("shape,chans", NET_CASES, ids=str)
def test_network_forward_backward(shape, chans):
    d = dev()
    from dedalus.closure import ClosureNet
    B, H, W = shape
    case("closure net", shape, chans)
    net = ClosureNet(cin=chans[0], cout=chans[-1], width=chans[1], depth=len(chans) - 1, seed=3,
                     out_scale=1.0)
    if len(set(chans[1:-1])) > 1:  # uneven widths
        net.chans = tuple(chans)
        gen = torch.Generator().manual_seed(9)
        for l in range(len(chans) - 1):
            net.weights[l] = torch.nn.Parameter(torch.randn(chans[l + 1], chans[l], 3, 3, generator=gen) * 0.3)
            net.biases[l] = torch.nn.Parameter(torch.randn(chans[l + 1], generator=gen) * 0.1)
    gb = torch.Generator().manual_seed(10)
    with torch.no_grad():
        for b in net.biases:
            b.copy_(torch.randn(b.shape, generator=gb) * 0.1)
    g = torch.Generator().manual_seed(11)
    x = torch.randn(B, H, W, chans[0], generator=g)
    gout = torch.randn(B, H, W, chans[-1], generator=g)
    ws = [w.detach().clone() for w in net.weights]
    bs = [b.detach().clone() for b in net.biases]
    net = net.to(d)
    out = net(x.to(d))
    out.backward(gout.to(d))
    torch.cuda.synchronize()
    ref, acts, zs = R.net_forward(x, ws, bs, emulate=True)
    gws, gbs = R.net_backward(gout, acts, zs, ws, emulate=True)
    print("net vs emulated reference: out %.2e" % rl2(out, ref),
          " dW", ["%.1e" % rl2(net.weights[l].grad, gws[l]) for l in range(len(ws))],
          " db", ["%.1e" % rl2(net.biases[l].grad, gbs[l]) for l in range(len(ws))])
    assert rl2(out, ref) < 1e-3
    for l in range(len(ws)):
        assert rl2(net.weights[l].grad, gws[l]) < 5e-3, l
        assert rl2(net.biases[l].grad, gbs[l]) < 5e-3, l
    # against the unrounded float64 network: the bf16 error is bounded
    ref64, acts64, zs64 = R.net_forward(x, ws, bs, emulate=False)
    gws64, gbs64 = R.net_backward(gout, acts64, zs64, ws, emulate=False)
    assert rl2(out, ref64) < 3e-2
    for l in range(len(ws)):
        assert rl2(net.weights[l].grad, gws64[l]) < 3e-2, l
        assert rl2(net.biases[l].grad, gbs64[l]) < 3e-2, l


def test_network_trains():
    """A few Adam steps on a fixed batch reduce the loss of a smooth target."""
    d = dev()
    from dedalus.closure import ClosureNet
    case("closure net training")
    torch.manual_seed(0)
    B, H, W = 8, 32, 32
    yy, xx = torch.meshgrid(torch.arange(H) * 2 * np.pi / H, torch.arange(W) * 2 * np.pi / W, indexing="ij")
    g = torch.Generator().manual_seed(12)
    x = torch.randn(B, H, W, 2, generator=g)
    x[..., 0] += torch.sin(yy)
    tgt = torch.stack([torch.roll(x[..., 0], 1, dims=1) - x[..., 1], x[..., 0] * 0.5], -1)
    net = ClosureNet(seed=1).to(d)
    opt = torch.optim.Adam(net.parameters(), lr=2e-3)
    xd, td = x.to(d), tgt.to(d)
    losses = []
    for it in range(60):
        opt.zero_grad()
        loss = ((net(xd) - td) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(loss.item())
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.3 * losses[0]


def test_inference_matches_network_forward():
    """The preallocated inference pass (used inside captured coarse steps)
    gives the training network's forward output bit for bit."""
    d = dev()
    from dedalus.closure import ClosureNet
    from dedalus.closure.net import Inference, pack
    case("inference vs forward")
    net = ClosureNet(cin=3, cout=2, width=32, depth=5, seed=5).to(d)
    g = torch.Generator().manual_seed(13)
    with torch.no_grad():
        for b in net.biases:
            b.copy_((torch.randn(b.shape, generator=g) * 0.1).to(d))
    B, H, W = 2, 64, 64
    x = torch.randn(B, H, W, 3, generator=g).to(d)
    with torch.no_grad():
        ref = net(x)
    inf = Inference(net, B, H, W)
    pack([(x.data_ptr() + 4 * c, False) for c in range(3)], B * H * W, inf.x.shape[1],
         pstride=3, bf16=inf.x, f32=False)
    out = inf.run()
    torch.cuda.synchronize()
    assert torch.equal(out[:, :2].reshape(B, H, W, 2), ref)


@pytest.mark.parametrize("B,H,W,width,depth", [(1, 64, 64, 32, 5), (3, 64, 64, 32, 5),
                                                (256, 64, 64, 32, 5), (2, 32, 32, 16, 3),
                                                (1, 16, 48, 32, 5)], ids=str)
def test_fused_network_matches_layers(B, H, W, width, depth):
    """ddx_conv_net (the whole network in one launch, activations in LDS,
    large or small tiles with recomputed halos) equals the layer-by-layer
    ddx_conv3x3 calls bit for bit."""
    d = dev()
    from dedalus.closure import ClosureNet
    from dedalus.closure.net import Inference, pack
    case("fused network", B, H, W, width, depth)
    net = ClosureNet(cin=3, cout=2, width=width, depth=depth, seed=7).to(d)
    g = torch.Generator().manual_seed(17)
    with torch.no_grad():
        for b in net.biases:
            b.copy_((torch.randn(b.shape, generator=g) * 0.1).to(d))
    x = torch.randn(B, H, W, 3, generator=g).to(d)
    outs = []
    for fused in (True, False):
        inf = Inference(net, B, H, W, fused=fused)
        assert inf.fused == fused
        pack([(x.data_ptr() + 4 * c, False) for c in range(3)], B * H * W, inf.x.shape[1],
             pstride=3, bf16=inf.x, f32=False)
        outs.append(inf.run().clone())
    torch.cuda.synchronize()
    assert torch.isfinite(outs[0]).all()
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("B", [1, 4])
def test_fused_network_from_fields(B):
    """run_fields (solver fields in, scaled tendency plus forcing out, one
    launch) equals pack -> layer-by-layer network -> unpack bit for bit."""
    d = dev()
    from dedalus.closure import ClosureNet
    from dedalus.closure.net import Inference, pack, unpack
    case("fused network from fields", B)
    H = W = 64
    net = ClosureNet(cin=3, cout=2, width=32, depth=5, seed=8).to(d)
    g = torch.Generator().manual_seed(19)
    with torch.no_grad():
        for b in net.biases:
            b.copy_((torch.randn(b.shape, generator=g) * 0.1).to(d))
    src = [torch.randn(B, H, W, generator=g, dtype=torch.float64).to(d) for _ in range(3)]
    add = [torch.randn(B, H, W, generator=g, dtype=torch.float64).to(d) for _ in range(2)]
    isc, osc = [0.7, 1.3, 0.25], [0.05, 0.02]
    f = Inference(net, B, H, W, fused=True)
    dst_f = [torch.empty(B, H, W, dtype=torch.float64, device=d) for _ in range(2)]
    f.run_fields([t.data_ptr() for t in src], isc, [t.data_ptr() for t in dst_f], osc,
                 add=[t.data_ptr() for t in add])
    l = Inference(net, B, H, W, fused=False)
    pack([(t.data_ptr(), True) for t in src], B * H * W, l.x.shape[1], scale=isc, bf16=l.x, f32=False)
    out = l.run()
    dst_l = [torch.empty(B, H, W, dtype=torch.float64, device=d) for _ in range(2)]
    unpack(out, 2, [t.data_ptr() for t in dst_l], True, scale=osc, add=[t.data_ptr() for t in add])
    torch.cuda.synchronize()
    for a, b in zip(dst_f, dst_l):
        assert torch.isfinite(a).all()
        assert torch.equal(a, b)


@pytest.mark.parametrize("B,H,W,Ci,Co", [(2, 8, 64, 16, 32), (1, 4, 128, 32, 16), (2, 64, 64, 32, 32),
                                          (3, 12, 64, 16, 16)], ids=str)
@pytest.mark.parametrize("nchunk", [1, 7, 64])
def test_conv_grads_one_pass(B, H, W, Ci, Co, nchunk):
    """ddx_conv3x3_grads (weight and bias gradients in one pass over dz):
    the weight gradient equals ddx_conv3x3_wgrad's bit for bit, the bias
    gradient equals the float64 column sum of the fp32 dz to 1e-12 relative
    (it is summed in double) and ddx_conv_colsum's to float rounding."""
    d = dev()
    net = C()
    case("conv grads one pass", B, H, W, Ci, Co, nchunk)
    x, _, _ = random_conv(B, H, W, Ci, Co, 12)
    g = torch.Generator().manual_seed(13)
    dzf = torch.randn(B * H * W, Co, generator=g)
    dzb = dzf.to(torch.bfloat16)
    xd, dzbd, dzfd = x.reshape(-1, Ci).to(d), dzb.to(d), dzf.to(d)
    dw, db = net.grads(xd, dzbd, dzfd, B, H, W, nchunk=nchunk)
    dw_ref = net.wgrad(xd, dzbd, B, H, W, nchunk=nchunk)
    db_ref = net.colsum(dzfd)
    torch.cuda.synchronize()
    assert torch.equal(dw, dw_ref)
    exact = dzf.double().sum(dim=0)
    assert rl2(db, exact) < 1e-7
    assert rl2(db, db_ref) < 1e-7
