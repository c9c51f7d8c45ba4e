"""The closure network on libddx's BF16 MFMA convolutions.

Activations are pixel-major [B*H*W, C] (b, y, x order) with C padded to a
multiple of 16; the network's own input and output are [B, H, W, C] float32
tensors on the GPU.  Every layer is a 3x3 periodic convolution; hidden
layers apply GELU (tanh form, z sigma(2 sqrt(2/pi) (z + 0.044715 z^3))).  Weights and biases are float32
torch Parameters in torch's (co, ci, ky, kx) layout; each forward pass
rounds them to bf16 once (``ddx_conv_wprep``) and every product
accumulates in float32.

Forward of layer l:   z_l = conv(a_{l-1}, w_l) + b_l,  a_l = bf16(GELU(z_l))
Backward of layer l:  dW_l = sum_p bf16(dz_l) x shift(a_{l-1})   (ddx_conv3x3_wgrad)
                      db_l = sum_p dz_l                          (ddx_conv_colsum)
                      dz_{l-1} = conv(bf16(dz_l), w_l^T) * GELU'(z_{l-1})
                                                    (one ddx_conv3x3 call, fused)
"""

import ctypes
import math

import torch

from .. import _hip


def pad16(c):
    return (int(c) + 15) // 16 * 16


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream or None)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check(rc):
    _hip._check(rc)


def _lib():
    _hip.init()
    return _hip._lib


def conv3x3(x_bf16, w_bf16, B, H, W, bias=None, aux=None, mode=_hip.CONV_LINEAR,
            out_f32=True, out_bf16=False):
    """One ddx_conv3x3 call: x [npix, Ci] bf16, w [Co, 9, Ci] bf16 -> (f32, bf16) outputs.

    out_f32 / out_bf16: True (allocate), False (skip) or a preallocated tensor."""
    lib = _lib()
    Ci = x_bf16.shape[1]
    Co = w_bf16.shape[0]
    npix = B * H * W
    assert x_bf16.shape[0] == npix and w_bf16.shape == (Co, 9, Ci)
    assert x_bf16.dtype == torch.bfloat16 and w_bf16.dtype == torch.bfloat16
    if isinstance(out_f32, torch.Tensor):
        of = out_f32
        assert of.shape == (npix, Co) and of.dtype == torch.float32 and of.is_contiguous()
    else:
        of = torch.empty((npix, Co), dtype=torch.float32, device=x_bf16.device) if out_f32 else None
    if isinstance(out_bf16, torch.Tensor):
        ob = out_bf16
        assert ob.shape == (npix, Co) and ob.dtype == torch.bfloat16 and ob.is_contiguous()
    else:
        ob = torch.empty((npix, Co), dtype=torch.bfloat16, device=x_bf16.device) if out_bf16 else None
    if aux is not None:
        assert aux.shape == (npix, Co) and aux.dtype == torch.float32
    _check(lib.ddx_conv3x3(_stream(), B, H, W, Ci, Co, _ptr(x_bf16), _ptr(w_bf16), _ptr(bias),
                           _ptr(aux), int(mode), _ptr(of), _ptr(ob)))
    return of, ob


def wgrad(x_bf16, dz_bf16, B, H, W, nchunk=None):
    """Padded weight gradient [Co, 9, Ci] float32 of one layer."""
    lib = _lib()
    Ci = x_bf16.shape[1]
    Co = dz_bf16.shape[1]
    npix = B * H * W
    if nchunk is None:
        nchunk = max(1, min(512, npix // 256))
    partial = torch.empty((nchunk, Co, 9, Ci), dtype=torch.float32, device=x_bf16.device)
    dw = torch.empty((Co, 9, Ci), dtype=torch.float32, device=x_bf16.device)
    _check(lib.ddx_conv3x3_wgrad(_stream(), B, H, W, Ci, Co, _ptr(x_bf16), _ptr(dz_bf16),
                                 _ptr(partial), int(nchunk), _ptr(dw)))
    return dw


def grads(x_bf16, dz_bf16, dz_f32, B, H, W, nchunk=None):
    """(padded weight gradient [Co, 9, Ci], bias gradient [Co]) of one layer in
    one pass over dz (ddx_conv3x3_grads), or None where that kernel does not
    apply."""
    Ci = x_bf16.shape[1]
    Co = dz_bf16.shape[1]
    if not (W % 64 == 0 and H % 4 == 0 and Ci in (16, 32) and Co in (16, 32)):
        return None
    lib = _lib()
    npix = B * H * W
    if nchunk is None:
        nchunk = max(1, min(512, npix // 256))
    dev = x_bf16.device
    partial = torch.empty((nchunk, Co, 9, Ci), dtype=torch.float32, device=dev)
    bpartial = torch.empty((nchunk, Co), dtype=torch.float64, device=dev)
    dw = torch.empty((Co, 9, Ci), dtype=torch.float32, device=dev)
    db = torch.empty(Co, dtype=torch.float32, device=dev)
    _check(lib.ddx_conv3x3_grads(_stream(), B, H, W, Ci, Co, _ptr(x_bf16), _ptr(dz_bf16), _ptr(dz_f32),
                                 _ptr(partial), _ptr(bpartial), int(nchunk), _ptr(dw), _ptr(db)))
    return dw, db


def colsum(a, nchunk=None):
    """Column sums of a [npix, C] float32 tensor (deterministic)."""
    lib = _lib()
    npix, C = a.shape
    if nchunk is None:
        nchunk = max(1, min(1024, npix // 1024))
    partial = torch.empty((nchunk, C), dtype=torch.float64, device=a.device)
    out = torch.empty(C, dtype=torch.float32, device=a.device)
    _check(lib.ddx_conv_colsum(_stream(), _ptr(a), npix, C, _ptr(partial), int(nchunk), _ptr(out)))
    return out


def wprep(w, cop, cip, forward=True, transposed=True):
    """fp32 (co, ci, 3, 3) weights -> padded bf16 [cop, 9, cip] and [cip, 9, cop]."""
    lib = _lib()
    co, ci = w.shape[:2]
    w = w.detach().contiguous()
    wf = torch.empty((cop, 9, cip), dtype=torch.bfloat16, device=w.device) if forward else None
    wt = torch.empty((cip, 9, cop), dtype=torch.bfloat16, device=w.device) if transposed else None
    _check(lib.ddx_conv_wprep(_stream(), _ptr(w), co, ci, cop, cip, _ptr(wf), _ptr(wt)))
    return wf, wt


def wunpack(dwp, co, ci):
    lib = _lib()
    cop, _, cip = dwp.shape
    dw = torch.empty((co, ci, 3, 3), dtype=torch.float32, device=dwp.device)
    _check(lib.ddx_conv_wunpack(_stream(), _ptr(dwp), co, ci, cop, cip, _ptr(dw)))
    return dw


def pack(srcs, npix, cp, pstride=1, scale=None, bf16=True, f32=False):
    """Gather channel sources (tensors, or (device pointer, is_float64) pairs;
    all of one precision) into padded pixel-major activations.  bf16 / f32:
    True (allocate), False (skip) or a preallocated [npix, cp] tensor."""
    lib = _lib()
    C = len(srcs)
    ptrs = []
    f64 = None
    dev = None
    for s in srcs:
        if isinstance(s, torch.Tensor):
            ptrs.append(s.data_ptr())
            is64 = s.dtype == torch.float64
            dev = s.device
        else:
            ptr, is64 = s
            ptrs.append(int(ptr))
        if f64 is None:
            f64 = is64
        assert f64 == is64, "channel sources must share one precision"
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    if isinstance(bf16, torch.Tensor):
        ob = bf16
        assert ob.shape == (npix, cp) and ob.dtype == torch.bfloat16 and ob.is_contiguous()
    else:
        ob = torch.empty((npix, cp), dtype=torch.bfloat16, device=dev) if bf16 else None
    if isinstance(f32, torch.Tensor):
        of = f32
        assert of.shape == (npix, cp) and of.dtype == torch.float32 and of.is_contiguous()
    else:
        of = torch.empty((npix, cp), dtype=torch.float32, device=dev) if f32 else None
    sc = None
    if scale is not None:
        sc = (ctypes.c_float * C)(*[float(v) for v in scale])
    _check(lib.ddx_conv_pack(_stream(), npix, C, cp, _hip.ptr_array(ptrs), int(pstride),
                             1 if f64 else 0, sc, _ptr(ob), _ptr(of)))
    return ob, of


def unpack(src, C, dsts, f64, pstride=1, scale=None, add=None):
    """Scatter the first C channels of src [npix, Cp] float32 to channel
    destinations (device pointers, float64 if f64), optionally adding the
    arrays ``add`` (device pointers of the destination type)."""
    lib = _lib()
    sc = None
    if scale is not None:
        sc = (ctypes.c_float * C)(*[float(v) for v in scale])
    ad = _hip.ptr_array([int(d) for d in add]) if add is not None else None
    _check(lib.ddx_conv_unpack(_stream(), src.shape[0], C, src.shape[1], _ptr(src),
                               _hip.ptr_array([int(d) for d in dsts]), int(pstride),
                               1 if f64 else 0, sc, ad))


class _NetFn(torch.autograd.Function):
    """The whole network as one autograd node, so the GELU derivative of each
    layer fuses into the next input-gradient convolution."""

    @staticmethod
    def forward(ctx, x, chans, *params):
        nl = len(chans) - 1
        ws, bs = params[:nl], params[nl:]
        B, H, W, cin = x.shape
        npix = B * H * W
        assert cin == chans[0] and npix % 16 == 0
        xc = x.detach().contiguous()
        a, _ = pack([(xc.data_ptr() + 4 * c, False) for c in range(cin)], npix, pad16(cin),
                    pstride=cin)
        acts, zs, wts = [a], [], []
        out = None
        for l in range(nl):
            ci, co = chans[l], chans[l + 1]
            cip, cop = pad16(ci), pad16(co)
            wf, wt = wprep(ws[l], cop, cip)
            bp = torch.zeros(cop, dtype=torch.float32, device=x.device)
            bp[:co] = bs[l].detach()
            wts.append(wt)
            if l < nl - 1:
                z, a = conv3x3(acts[-1], wf, B, H, W, bias=bp, mode=_hip.CONV_GELU,
                               out_f32=True, out_bf16=True)
                zs.append(z)
                acts.append(a)
            else:
                out, _ = conv3x3(acts[-1], wf, B, H, W, bias=bp, mode=_hip.CONV_LINEAR)
        ctx.chans = chans
        ctx.shape = (B, H, W)
        ctx.save_for_backward(*acts, *zs, *wts)
        return out[:, :chans[-1]].reshape(B, H, W, chans[-1])

    @staticmethod
    def backward(ctx, g):
        chans = ctx.chans
        nl = len(chans) - 1
        B, H, W = ctx.shape
        npix = B * H * W
        saved = ctx.saved_tensors
        acts, zs, wts = saved[:nl], saved[nl:2 * nl - 1], saved[2 * nl - 1:]
        gc = g.detach().contiguous()
        cout = chans[-1]
        dzb, dzf = pack([(gc.data_ptr() + 4 * c, False) for c in range(cout)], npix,
                        pad16(cout), pstride=cout, bf16=True, f32=True)
        gws, gbs = [None] * nl, [None] * nl
        for l in range(nl - 1, -1, -1):
            ci, co = chans[l], chans[l + 1]
            g2 = grads(acts[l], dzb, dzf, B, H, W)
            if g2 is None:
                gws[l] = wunpack(wgrad(acts[l], dzb, B, H, W), co, ci)
                gbs[l] = colsum(dzf)[:co].contiguous()
            else:
                gws[l] = wunpack(g2[0], co, ci)
                gbs[l] = g2[1][:co].contiguous()
            if l > 0:
                dzf, dzb = conv3x3(dzb, wts[l], B, H, W, aux=zs[l - 1],
                                   mode=_hip.CONV_GELU_GRAD, out_f32=True, out_bf16=True)
        return (None, None, *gws, *gbs)


class ClosureNet(torch.nn.Module):
    """Periodic CNN: chans = (in, hidden, ..., out), 3x3 layers, GELU between."""

    def __init__(self, cin=2, cout=2, width=32, depth=5, seed=0, out_scale=0.1):
        super().__init__()
        self.chans = tuple([cin] + [width] * (depth - 1) + [cout])
        gen = torch.Generator().manual_seed(int(seed))
        ws, bs = [], []
        for l in range(depth):
            ci, co = self.chans[l], self.chans[l + 1]
            std = math.sqrt(2.0 / (9 * ci))
            if l == depth - 1:
                std *= out_scale
            ws.append(torch.nn.Parameter(torch.randn(co, ci, 3, 3, generator=gen) * std))
            bs.append(torch.nn.Parameter(torch.zeros(co)))
        self.weights = torch.nn.ParameterList(ws)
        self.biases = torch.nn.ParameterList(bs)

    def forward(self, x):
        """x [B, H, W, cin] float32 on the GPU -> [B, H, W, cout] float32."""
        if not x.is_cuda:
            raise _hip.DDXError(-5, "ClosureNet runs on the HIP kernels only (input is on the CPU)")
        if x.dtype != torch.float32 or x.dim() != 4 or x.shape[-1] != self.chans[0]:
            raise ValueError("ClosureNet input must be float32 [B, H, W, %d]" % self.chans[0])
        return _NetFn.apply(x, self.chans, *self.weights, *self.biases)

    def flops_per_pixel(self):
        """Multiply-adds x 2 per pixel of one forward pass, padded channels."""
        return sum(2 * 9 * pad16(self.chans[l]) * pad16(self.chans[l + 1])
                   for l in range(len(self.chans) - 1))


class Inference:
    """Forward pass of a ClosureNet on fixed-size batches with preallocated
    buffers: ``run()`` launches only the convolutions (no allocation, no host
    synchronisation), so it can be captured in a graph.  Call ``refresh()``
    after the weights change.

    fused: the whole network as one launch (``ddx_conv_net``: activations
    stay on chip) where the kernel supports the architecture and image size
    (None: whenever it does), else one ``ddx_conv3x3`` launch per layer; both
    give the same values."""

    def __init__(self, net, B, H, W, fused=None):
        self.net = net
        self.shape = (int(B), int(H), int(W))
        self.npix = B * H * W
        if self.npix % 16:
            raise ValueError("B*H*W must be a multiple of 16")
        dev = net.weights[0].device
        ch = net.chans
        self.x = torch.zeros((self.npix, pad16(ch[0])), dtype=torch.bfloat16, device=dev)
        self.out = torch.empty((self.npix, pad16(ch[-1])), dtype=torch.float32, device=dev)
        self.wf = [torch.empty((pad16(ch[l + 1]), 9, pad16(ch[l])), dtype=torch.bfloat16, device=dev)
                   for l in range(len(ch) - 1)]
        self.bp = [torch.zeros(pad16(ch[l + 1]), dtype=torch.float32, device=dev)
                   for l in range(len(ch) - 1)]
        nl = len(ch) - 1
        self._chans = (ctypes.c_int * (nl + 1))(*ch)
        self._wptr = _hip.ptr_array([t.data_ptr() for t in self.wf])
        self._bptr = _hip.ptr_array([t.data_ptr() for t in self.bp])
        self.fused = fused
        if self.fused is None:
            self.fused = self._fused_supported()
        self.acts = None if self.fused else [
            torch.empty((self.npix, pad16(c)), dtype=torch.bfloat16, device=dev) for c in ch[1:-1]]
        self.refresh()

    def _fused_supported(self):
        """The fused kernel's shape rule (include/ddx_closure.h, ddx_conv_net)."""
        ch = self.net.chans
        nl = len(ch) - 1
        B, H, W = self.shape
        hid = set(ch[1:-1])
        return ((nl, hid) in ((5, {32}), (3, {16})) and ch[0] <= 16 and ch[-1] <= 16
                and H % 8 == 0 and W % 16 == 0)

    def refresh(self):
        lib = _lib()
        ch = self.net.chans
        for l in range(len(ch) - 1):
            w = self.net.weights[l].detach().contiguous()
            _check(lib.ddx_conv_wprep(_stream(), _ptr(w), ch[l + 1], ch[l], pad16(ch[l + 1]),
                                      pad16(ch[l]), _ptr(self.wf[l]), None))
            self.bp[l].zero_()
            self.bp[l][:ch[l + 1]].copy_(self.net.biases[l].detach())

    def run(self):
        B, H, W = self.shape
        nl = len(self.net.chans) - 1
        if self.fused:
            _check(_lib().ddx_conv_net(_stream(), B, H, W, nl, self._chans, self._wptr, self._bptr,
                                       _ptr(self.x), _ptr(self.out), None, None, None, None, None))
            return self.out
        a = self.x
        for l in range(nl):
            if l < nl - 1:
                conv3x3(a, self.wf[l], B, H, W, bias=self.bp[l], mode=_hip.CONV_GELU,
                        out_f32=False, out_bf16=self.acts[l])
                a = self.acts[l]
            else:
                conv3x3(a, self.wf[l], B, H, W, bias=self.bp[l], mode=_hip.CONV_LINEAR,
                        out_f32=self.out, out_bf16=False)
        return self.out

    def run_fields(self, src, in_scale, dst, out_scale, add=None):
        """The network from solver fields to solver fields in one launch:
        input channel c = bf16(in_scale[c] * src[c]) (float64 x-space arrays,
        as ``pack`` rounds them), output dst[c] = out_scale[c] * y_c (+ add[c])
        (as ``unpack``).  Needs the fused kernel."""
        if not self.fused:
            raise _hip.DDXError(-3, "run_fields needs the fused network kernel")
        B, H, W = self.shape
        nl = len(self.net.chans) - 1
        isc = (ctypes.c_float * len(in_scale))(*[float(v) for v in in_scale])
        osc = (ctypes.c_float * len(out_scale))(*[float(v) for v in out_scale])
        ad = _hip.ptr_array([int(d) for d in add]) if add is not None else None
        _check(_lib().ddx_conv_net(_stream(), B, H, W, nl, self._chans, self._wptr, self._bptr,
                                   None, None, _hip.ptr_array([int(d) for d in src]), isc,
                                   _hip.ptr_array([int(d) for d in dst]), ad, osc))
