"""3-D residual U-Net for sparse-view FDK correction (no reference counterpart).

Architecture: 3 levels, base 32 channels, two 3x3x3 conv + ReLU per level,
2x2x2 average pooling down, nearest 2x upsampling up with channel-concatenated
skips, a 1x1 output head, and a residual connection (output = input + net).

With ``s2d=f`` (2 or 4) the network runs on a coarse grid: the input's f^3
voxel blocks become f^3 channels (space-to-depth; 8 channels are zero-padded to
32 for the first convolution) and an f^3-output head writes each block back
(depth-to-space).  Nothing of the input is lost, and every convolution runs on
1/f^3 of the voxels.

Every operation runs in the library's HIP kernels (fdk_nn_*): BF16 MFMA
convolutions with FP32 accumulation, BF16 channels-last activations, FP32
master weights and Adam.  This module only sequences them; torch tensors serve
as device buffers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import FdkError, _check, lib

I16 = None  # torch.int16, set lazily (BF16 bit patterns)


def _torch():
    import torch
    global I16
    I16 = torch.int16
    return torch


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _st():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _sig():
    L = lib()
    vp, i, i64, f = C.c_void_p, C.c_int, C.c_int64, C.c_float
    L.fdk_nn_conv3d_fwd.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, i, i, i, i, vp]
    L.fdk_nn_conv3d_fwd_cat.argtypes = [vp, i, vp, vp, vp, vp, vp, i, i, i, i, i, i, i, i, vp]
    L.fdk_nn_conv3d_wgrad.argtypes = [vp, vp, vp, i, i, i, i, i, i, vp]
    L.fdk_nn_conv3d_wgrad_cat.argtypes = [vp, i, vp, vp, vp, i, i, i, i, i, i, vp]
    L.fdk_nn_flip_weights.argtypes = [vp, vp, i, i, vp]
    L.fdk_nn_bias_grad.argtypes = [vp, vp, i64, i, vp]
    L.fdk_nn_relu_bwd.argtypes = [vp, vp, vp, i64, i, vp]
    L.fdk_nn_avgpool2.argtypes = [vp, vp, i, i, i, i, i, i, vp]
    L.fdk_nn_upsample2.argtypes = [vp, vp, i, i, i, i, i, i, i, i, vp]
    L.fdk_nn_channel_copy.argtypes = [vp, i, i, vp, i, i, i, i64, vp]
    L.fdk_nn_add.argtypes = [vp, vp, vp, i64, vp]
    L.fdk_nn_head.argtypes = [vp, vp, vp, vp, vp, i64, i, vp]
    L.fdk_nn_head_bwd.argtypes = [vp, vp, vp, vp, vp, vp, i64, i, vp]
    L.fdk_nn_unshuffle.argtypes = [vp, vp, i, i, i, i, i, i, vp]
    L.fdk_nn_head_shuffle.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, i, i, vp]
    L.fdk_nn_head_shuffle_bwd.argtypes = [vp, vp, vp, vp, vp, vp, i, i, i, i, i, i, vp]
    L.fdk_nn_mse.argtypes = [vp, vp, vp, vp, i64, vp]
    L.fdk_nn_adam.argtypes = [vp, vp, vp, vp, vp, i64, f, f, f, f, i, i, i, vp]
    L.fdk_nn_convert.argtypes = [vp, vp, i64, i, vp]
    L.fdk_nn_box_copy.argtypes = [vp, vp, vp, vp, vp, vp, vp, f, vp]
    L.fdk_poisson_noise.argtypes = [vp, i64, C.c_double, C.c_double, C.c_uint32, vp]
    L.fdk_nn_mfma_probe.argtypes = [vp, vp, vp, vp]
    ip = C.POINTER(C.c_int)
    L.fdk_unet_inspect.argtypes = [C.c_char_p, ip, ip, ip]
    L.fdk_unet_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.fdk_unet_info.argtypes = [vp, ip, ip, ip]
    L.fdk_unet_forward.argtypes = [vp, vp, vp, i, i, i, i, vp]
    L.fdk_unet_infer.argtypes = [vp, vp, vp, i, i, i, i, i, vp]
    L.fdk_unet_destroy.argtypes = [vp]
    return L


_L = None


def L():
    global _L
    if _L is None:
        _L = _sig()
    return _L


# ----------------------------------------------------------- kernel wrappers
def conv3d(x, w, bias, y, shape, cin, cout, relu=True, out_f32=False, resid=None):
    n, d, h, wd = shape
    _check(L().fdk_nn_conv3d_fwd(_p(x), _p(w), _p(bias), _p(resid), _p(y), n, d, h, wd, cin, cout,
                                 int(relu), int(out_f32), _st()), "conv3d")
    return y


def conv3d_cat(up, cup, x, w, bias, y, shape, cin, cout, relu=True):
    """conv3d on the concatenation [nearest 2x upsampling of up (cup channels), x]
    without forming it; shape is the full-resolution (n, d, h, w)."""
    n, d, h, wd = shape
    _check(L().fdk_nn_conv3d_fwd_cat(_p(up), int(cup), _p(x), _p(w), _p(bias), None, _p(y), n, d,
                                     h, wd, cin, cout, int(relu), 0, _st()), "conv3d_cat")
    return y


def conv3d_wgrad(x, dy, dw, shape, cin, cout):
    n, d, h, wd = shape
    _check(L().fdk_nn_conv3d_wgrad(_p(x), _p(dy), _p(dw), n, d, h, wd, cin, cout, _st()), "wgrad")


def conv3d_wgrad_cat(up, cup, x, dy, dw, shape, cin, cout):
    """conv3d_wgrad of the input [nearest 2x upsampling of up (cup channels), x]
    read without forming it; shape is the full-resolution (n, d, h, w)."""
    n, d, h, wd = shape
    _check(L().fdk_nn_conv3d_wgrad_cat(_p(up), int(cup), _p(x), _p(dy), _p(dw), n, d, h, wd, cin,
                                       cout, _st()), "wgrad_cat")


def flip_weights(w, wb, cin, cout):
    _check(L().fdk_nn_flip_weights(_p(w), _p(wb), cin, cout, _st()), "flip_weights")


def bias_grad(dy, db, nvox, c):
    _check(L().fdk_nn_bias_grad(_p(dy), _p(db), nvox, c, _st()), "bias_grad")


def relu_bwd(dy, y, dz, mode=1):
    _check(L().fdk_nn_relu_bwd(_p(dy), _p(y), _p(dz), dy.numel(), int(mode), _st()), "relu_bwd")
    return dz


def avgpool(x, y, shape, c, adjoint=False):
    n, d, h, w = shape
    _check(L().fdk_nn_avgpool2(_p(x), _p(y), n, d, h, w, c, int(adjoint), _st()), "avgpool")
    return y


def upsample(x, y, small_shape, cx, cy, c0, adjoint=False):
    n, d, h, w = small_shape
    _check(L().fdk_nn_upsample2(_p(x), _p(y), n, d, h, w, cx, cy, c0, int(adjoint), _st()),
           "upsample")
    return y


def channel_copy(src, cs, s0, dst, cd, d0, cc, nvox):
    _check(L().fdk_nn_channel_copy(_p(src), cs, s0, _p(dst), cd, d0, cc, nvox, _st()), "copy")


def add(a, b, y):
    _check(L().fdk_nn_add(_p(a), _p(b), _p(y), a.numel(), _st()), "add")
    return y


def unshuffle(x32, y, f=2):
    """FP32 (N, D, H, W) -> BF16 (N, D/f, H/f, W/f, y.shape[-1]) space-to-depth."""
    n, d, h, w = x32.shape
    _check(L().fdk_nn_unshuffle(_p(x32), _p(y), n, d, h, w, f, y.shape[-1], _st()), "unshuffle")
    return y


def head_shuffle(hid, w, b, x32, out, f=2):
    n, d, h, wd = x32.shape
    _check(L().fdk_nn_head_shuffle(_p(hid), _p(w), _p(b), _p(x32), _p(out), n, d, h, wd, f,
                                   hid.shape[-1], _st()), "head_shuffle")
    return out


def head_shuffle_bwd(dout, hid, w, dh, dw, db, f=2):
    n, d, h, wd = dout.shape
    _check(L().fdk_nn_head_shuffle_bwd(_p(dout), _p(hid), _p(w), _p(dh), _p(dw), _p(db), n, d, h,
                                       wd, f, hid.shape[-1], _st()), "head_shuffle_bwd")
    return dh


def to_bf16(src, dst):
    _check(L().fdk_nn_convert(_p(src), _p(dst), src.numel(), 1, _st()), "to_bf16")
    return dst


def to_f32(src, dst):
    _check(L().fdk_nn_convert(_p(src), _p(dst), src.numel(), 0, _st()), "to_f32")
    return dst


def mse(out, target, dout, loss_sum):
    _check(L().fdk_nn_mse(_p(out), _p(target), _p(dout), _p(loss_sum), out.numel(), _st()), "mse")


def box_copy(src, src_org, dst, dst_org, box, scale=1.0):
    sd = (C.c_int * 3)(*src.shape[-3:])
    so = (C.c_int * 3)(*src_org)
    dd = (C.c_int * 3)(*dst.shape[-3:])
    do = (C.c_int * 3)(*dst_org)
    bx = (C.c_int * 3)(*box)
    _check(L().fdk_nn_box_copy(_p(src), sd, so, _p(dst), dd, do, bx, float(scale), _st()),
           "box_copy")


def poisson_noise(proj, i0, mu, seed):
    _check(L().fdk_poisson_noise(_p(proj), proj.numel(), float(i0), float(mu), int(seed) & 0xffffffff,
                                 _st()), "poisson_noise")


# -------------------------------------------------------------------- layers
class Conv3:
    """3x3x3 'same' convolution (+ bias, activation) with Adam state.  relu: the
    activation mode, 0 none, 1 ReLU, 2 leaky ReLU."""

    def __init__(self, cin, cout, relu, rng, dev, live_cin=None):
        """live_cin: input channels that carry data (the rest are zero padding
        and get zero weights)."""
        torch = _torch()
        self.cin, self.cout, self.relu = cin, cout, int(relu)
        live = cin if live_cin is None else live_cin
        std = np.sqrt(2.0 / (27 * live))
        if cin == 1:
            w = np.zeros((cout, 32), np.float32)
            w[:, :27] = rng.standard_normal((cout, 27)) * std
        else:
            w = (rng.standard_normal((27, cout, cin)) * std).astype(np.float32)
            w[:, :, live:] = 0.0
        self.w = torch.as_tensor(w).to(dev)
        self.b = torch.zeros(cout, device=dev)
        self.wbf = torch.empty(self.w.shape, dtype=I16, device=dev)
        self.wflip = torch.empty((27, cin, cout), dtype=I16, device=dev) if cin > 1 else None
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b)
        self.state = [torch.zeros_like(self.w), torch.zeros_like(self.w),
                      torch.zeros_like(self.b), torch.zeros_like(self.b)]
        self.refresh()

    def refresh(self):
        to_bf16(self.w, self.wbf)
        if self.wflip is not None:
            flip_weights(self.wbf, self.wflip, self.cin, self.cout)

    def forward(self, x, shape, out_f32=False):
        torch = _torch()
        y = torch.empty(shape + (self.cout,), dtype=torch.float32 if out_f32 else I16,
                        device=x.device)
        return conv3d(x, self.wbf, self.b, y, shape, self.cin, self.cout, self.relu, out_f32)

    def cat_ok(self, cup):
        """Whether forward_cat serves this layer (the fused decoder-input kernel)."""
        return (self.cout % 32 == 0 and self.cin % 32 == 0 and cup % 32 == 0
                and 0 < cup < self.cin)

    def forward_cat(self, up, cup, x, shape):
        """forward() of [upsample2(up), x] (channels concatenated) without forming it."""
        torch = _torch()
        y = torch.empty(shape + (self.cout,), dtype=I16, device=x.device)
        return conv3d_cat(up, cup, x, self.wbf, self.b, y, shape, self.cin, self.cout, self.relu)

    def backward(self, dy, x, y, shape, need_dx=True, cat=None):
        """dy: gradient at the (post-ReLU) output y; x: the layer input, or with
        cat = (up, cup) the skip half of the input [upsample2(up), x] that
        forward_cat read."""
        torch = _torch()
        dz = relu_bwd(dy, y, torch.empty_like(dy), self.relu) if self.relu else dy
        if cat is None:
            conv3d_wgrad(x, dz, self.gw, shape, self.cin, self.cout)
        else:
            conv3d_wgrad_cat(cat[0], cat[1], x, dz, self.gw, shape, self.cin, self.cout)
        bias_grad(dz, self.gb, int(np.prod(shape)), self.cout)
        if not need_dx:
            return None
        dx = torch.empty(shape + (self.cin,), dtype=I16, device=dy.device)
        return conv3d(dz, self.wflip, None, dx, shape, self.cout, self.cin, relu=False)

    def params(self):
        return [(self.w, self.wbf, self.gw, self.state[0], self.state[1]),
                (self.b, None, self.gb, self.state[2], self.state[3])]


def adam(w, wbf, g, m, v, lr, step, b1=0.9, b2=0.999, eps=1e-8):
    n = w.numel()
    _check(L().fdk_nn_adam(_p(w), _p(wbf), _p(g), _p(m), _p(v), n, lr, b1, b2, eps, step, n, n,
                           _st()), "adam")


class UNet3D:
    """Residual U-Net: out = x + head(dec1(...)).  Input / output (N, D, H, W)
    FP32 tensors; D % 4f, H % 16f and W % 128f == 0 for block edge f = s2d
    (1: no space-to-depth)."""

    def __init__(self, base: int = 32, seed: int = 0, device="cuda", s2d: int = 1,
                 act: int = 1):
        """act: 1 ReLU, 2 leaky ReLU after every convolution."""
        torch = _torch()
        rng = np.random.default_rng(seed)
        c1, c2, c3 = base, 2 * base, 4 * base
        self.c = (c1, c2, c3)
        if s2d not in (1, 2, 4):
            raise ValueError(f"s2d must be 1, 2 or 4 (got {s2d})")
        if act not in (1, 2):
            raise ValueError(f"act must be 1 (ReLU) or 2 (leaky ReLU) (got {act})")
        self.s2d, self.act = int(s2d), int(act)
        if self.s2d > 1:
            assert c1 == 32, "the depth-to-space head takes 32 channels"
            k = self.s2d ** 3
            self.e1a = Conv3(max(32, k), c1, act, rng, device, live_cin=k)
        else:
            self.e1a = Conv3(1, c1, act, rng, device)
        self.e1b = Conv3(c1, c1, act, rng, device)
        self.e2a = Conv3(c1, c2, act, rng, device)
        self.e2b = Conv3(c2, c2, act, rng, device)
        self.ba = Conv3(c2, c3, act, rng, device)
        self.bb = Conv3(c3, c3, act, rng, device)
        self.d2a = Conv3(c3 + c2, c2, act, rng, device)
        self.d2b = Conv3(c2, c2, act, rng, device)
        self.d1a = Conv3(c2 + c1, c1, act, rng, device)
        self.d1b = Conv3(c1, c1, act, rng, device)
        nh = self.s2d ** 3
        hw = (rng.standard_normal((nh, c1)) * 1e-3).astype(np.float32)
        self.hw = torch.as_tensor(hw if self.s2d > 1 else hw[0]).to(device)
        self.hb = torch.zeros(nh, device=device)
        self.ghw = torch.zeros_like(self.hw)
        self.ghb = torch.zeros_like(self.hb)
        self.hstate = [torch.zeros_like(self.hw), torch.zeros_like(self.hw),
                       torch.zeros_like(self.hb), torch.zeros_like(self.hb)]
        self.step_count = 0
        self.cache = None
        # decoder inputs read in place (forward and weight gradients); False
        # forms the concatenations instead (same values)
        self.fuse_decoder = True

    @property
    def convs(self):
        return [self.e1a, self.e1b, self.e2a, self.e2b, self.ba, self.bb, self.d2a, self.d2b,
                self.d1a, self.d1b]

    def forward(self, x32, keep=False, out=None):
        """x32 (N, D, H, W) FP32 -> (N, D, H, W) FP32, into ``out`` when given.
        keep: hold every activation for backward().  Otherwise each is released
        once its last reader is queued (the stream orders the reuse), so the
        peak is about the decoder's concatenation plus two 32-channel tensors."""
        torch = _torch()
        c1, c2, c3 = self.c
        N, D, H, W = x32.shape
        self.check_shape(x32.shape)
        dev = x32.device
        k = {}

        def bf(shape, c):
            return torch.empty(shape + (c,), dtype=I16, device=dev)

        def hold(**t):
            if keep:
                k.update(t)

        f = self.s2d
        if f > 1:
            D, H, W = D // f, H // f, W // f
            xb = unshuffle(x32, bf((N, D, H, W), self.e1a.cin), f)
        else:
            xb = to_bf16(x32, torch.empty(x32.shape, dtype=I16, device=dev))
        s1_, s2_, s3_ = (N, D, H, W), (N, D // 2, H // 2, W // 2), (N, D // 4, H // 4, W // 4)
        e1a = self.e1a.forward(xb, s1_)
        hold(xb=xb)
        del xb
        s1 = self.e1b.forward(e1a, s1_)
        hold(e1a=e1a)
        del e1a
        p1 = avgpool(s1, bf(s2_, c1), s1_, c1)
        e2a = self.e2a.forward(p1, s2_)
        hold(p1=p1)
        del p1
        s2 = self.e2b.forward(e2a, s2_)
        hold(e2a=e2a)
        del e2a
        p2 = avgpool(s2, bf(s3_, c2), s2_, c2)
        ba = self.ba.forward(p2, s3_)
        hold(p2=p2)
        del p2
        bb = self.bb.forward(ba, s3_)
        hold(ba=ba)
        del ba
        # Decoder inputs [upsample2(coarse), skip]: read in place by the fused
        # kernels (backward() then reads them in place too), formed only when
        # those kernels do not serve the layer.
        fuse = self.fuse_decoder and self.d2a.cat_ok(c3) and self.d1a.cat_ok(c2)
        if fuse:
            d2a = self.d2a.forward_cat(bb, c3, s2, s2_)
            hold(bb=bb, s2=s2)
            del bb, s2
        else:
            cat2 = bf(s2_, c3 + c2)
            upsample(bb, cat2, s3_, c3, c3 + c2, 0)
            channel_copy(s2, c2, 0, cat2, c3 + c2, c3, c2, int(np.prod(s2_)))
            hold(bb=bb, s2=s2)
            del bb, s2
            d2a = self.d2a.forward(cat2, s2_)
            hold(cat2=cat2)
            del cat2
        d2 = self.d2b.forward(d2a, s2_)
        hold(d2a=d2a)
        del d2a
        if fuse:
            d1a = self.d1a.forward_cat(d2, c2, s1, s1_)
            hold(d2=d2, s1=s1)
            del d2, s1
        else:
            cat1 = bf(s1_, c2 + c1)
            upsample(d2, cat1, s2_, c2, c2 + c1, 0)
            channel_copy(s1, c1, 0, cat1, c2 + c1, c2, c1, int(np.prod(s1_)))
            hold(d2=d2, s1=s1)
            del d2, s1
            d1a = self.d1a.forward(cat1, s1_)
            hold(cat1=cat1)
            del cat1
        d1 = self.d1b.forward(d1a, s1_)
        hold(d1a=d1a)
        del d1a
        if out is None:
            out = torch.empty(x32.shape, dtype=torch.float32, device=dev)
        if f > 1:
            head_shuffle(d1, self.hw, self.hb, x32, out, f)
        else:
            _check(L().fdk_nn_head(_p(d1), _p(self.hw), _p(self.hb), _p(x32), _p(out),
                                   out.numel(), c1, _st()), "head")
        if keep:
            k.update(d1=d1)
            self.cache = dict(shapes=(s1_, s2_, s3_), **k)
        return out

    def zero_grad(self):
        for cv in self.convs:
            cv.gw.zero_()
            cv.gb.zero_()
        self.ghw.zero_()
        self.ghb.zero_()

    def backward(self, dout):
        torch = _torch()
        c1, c2, c3 = self.c
        k = self.cache
        s1_, s2_, s3_ = k["shapes"]
        dev = dout.device
        n1, n2 = int(np.prod(s1_)), int(np.prod(s2_))
        dd1 = torch.empty(s1_ + (c1,), dtype=I16, device=dev)
        if self.s2d > 1:
            head_shuffle_bwd(dout, k["d1"], self.hw, dd1, self.ghw, self.ghb, self.s2d)
        else:
            _check(L().fdk_nn_head_bwd(_p(dout), _p(k["d1"]), _p(self.hw), _p(dd1), _p(self.ghw),
                                       _p(self.ghb), dout.numel(), c1, _st()), "head_bwd")
        dd1a = self.d1b.backward(dd1, k["d1a"], k["d1"], s1_)
        if "cat1" in k:
            dcat1 = self.d1a.backward(dd1a, k["cat1"], k["d1a"], s1_)
        else:  # the fused forward: the weight gradient reads [upsample2(d2), s1] in place
            dcat1 = self.d1a.backward(dd1a, k["s1"], k["d1a"], s1_, cat=(k["d2"], c2))
        dd2 = upsample(dcat1, torch.empty(s2_ + (c2,), dtype=I16, device=dev), s2_, c2, c2 + c1,
                       0, adjoint=True)
        ds1 = torch.empty(s1_ + (c1,), dtype=I16, device=dev)
        channel_copy(dcat1, c2 + c1, c2, ds1, c1, 0, c1, n1)
        dd2a = self.d2b.backward(dd2, k["d2a"], k["d2"], s2_)
        if "cat2" in k:
            dcat2 = self.d2a.backward(dd2a, k["cat2"], k["d2a"], s2_)
        else:
            dcat2 = self.d2a.backward(dd2a, k["s2"], k["d2a"], s2_, cat=(k["bb"], c3))
        dbb = upsample(dcat2, torch.empty(s3_ + (c3,), dtype=I16, device=dev), s3_, c3, c3 + c2,
                       0, adjoint=True)
        ds2 = torch.empty(s2_ + (c2,), dtype=I16, device=dev)
        channel_copy(dcat2, c3 + c2, c3, ds2, c2, 0, c2, n2)
        dba = self.bb.backward(dbb, k["ba"], k["bb"], s3_)
        dp2 = self.ba.backward(dba, k["p2"], k["ba"], s3_)
        add(ds2, avgpool(dp2, torch.empty_like(ds2), s2_, c2, adjoint=True), ds2)
        de2a = self.e2b.backward(ds2, k["e2a"], k["s2"], s2_)
        dp1 = self.e2a.backward(de2a, k["p1"], k["e2a"], s2_)
        add(ds1, avgpool(dp1, torch.empty_like(ds1), s1_, c1, adjoint=True), ds1)
        de1a = self.e1b.backward(ds1, k["e1a"], k["s1"], s1_)
        self.e1a.backward(de1a, k["xb"], k["e1a"], s1_, need_dx=False)

    def adam_step(self, lr=1e-3):
        self.step_count += 1
        t = self.step_count
        for cv in self.convs:
            for (w, wbf, g, m, v) in cv.params():
                adam(w, wbf, g, m, v, lr, t)
            if cv.wflip is not None:
                flip_weights(cv.wbf, cv.wflip, cv.cin, cv.cout)
        adam(self.hw, None, self.ghw, self.hstate[0], self.hstate[1], lr, t)
        adam(self.hb, None, self.ghb, self.hstate[2], self.hstate[3], lr, t)

    def train_step(self, x32, target, lr=1e-3):
        torch = _torch()
        out = self.forward(x32, keep=True)
        dout = torch.empty_like(out)
        loss = torch.zeros(1, device=out.device)
        mse(out, target, dout, loss)
        self.zero_grad()
        self.backward(dout)
        self.adam_step(lr)
        self.cache = None
        return loss  # sum of squared errors (device)

    def check_shape(self, shape):
        """Raise unless an (N, D, H, W) input fits the network's kernels."""
        _, D, H, W = shape
        need = self.min_shape
        if D % need[0] or H % need[1] or W % need[2]:
            raise ValueError(f"U-Net input (D, H, W) = {(D, H, W)} must be multiples of {need}")

    @property
    def min_shape(self):
        """The smallest (D, H, W) input."""
        f = self.s2d
        return (4 * f, 16 * f, 128 * f)

    # ------------------------------------------------------------ persistence
    def state_dict(self):
        d = {"arch.s2d": np.array(self.s2d, np.int32), "arch.act": np.array(self.act, np.int32)}
        for i, cv in enumerate(self.convs):
            d[f"conv{i}.w"] = cv.w.cpu().numpy()
            d[f"conv{i}.b"] = cv.b.cpu().numpy()
        d["head.w"] = self.hw.cpu().numpy()
        d["head.b"] = self.hb.cpu().numpy()
        return d

    def load_state_dict(self, d):
        torch = _torch()
        s2d = int(d["arch.s2d"]) if "arch.s2d" in d else 1
        act = int(d["arch.act"]) if "arch.act" in d else 1
        if (s2d, act) != (self.s2d, self.act):
            raise ValueError(f"weights are for s2d={s2d}, act={act}; the network has "
                             f"s2d={self.s2d}, act={self.act}")
        for i, cv in enumerate(self.convs):
            cv.w.copy_(torch.as_tensor(d[f"conv{i}.w"]))
            cv.b.copy_(torch.as_tensor(d[f"conv{i}.b"]))
            cv.refresh()
        self.hw.copy_(torch.as_tensor(d["head.w"]))
        self.hb.copy_(torch.as_tensor(d["head.b"]))

    def save(self, path):
        np.savez_compressed(path, **self.state_dict())

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.load_state_dict({k: z[k] for k in z.files})

    @classmethod
    def from_file(cls, path, device="cuda"):
        """A network built for the saved weights' architecture, with them loaded."""
        with np.load(path, allow_pickle=False) as z:
            d = {k: z[k] for k in z.files}
        s2d = int(d["arch.s2d"]) if "arch.s2d" in d else 1
        act = int(d["arch.act"]) if "arch.act" in d else 1
        base = int(d["conv1.w"].shape[1])
        net = cls(base=base, seed=0, device=device, s2d=s2d, act=act)
        net.load_state_dict(d)
        return net


# ------------------------------------------------------ native inference runner
def inspect(path) -> tuple:
    """(base, s2d, act) of a weight file, read by the library (no device)."""
    b, f, a = C.c_int(), C.c_int(), C.c_int()
    _check(L().fdk_unet_inspect(str(path).encode(), C.byref(b), C.byref(f), C.byref(a)),
           "fdk_unet_inspect")
    return b.value, f.value, a.value


class NativeUNet:
    """The library's own U-Net inference (fdk_unet_*): weights loaded from a file
    by the C++ runner, the forward pass sequenced natively from the same kernels
    as UNet3D (bit-identical outputs).  Bound to the current device."""

    def __init__(self, path):
        self.h = C.c_void_p()
        _check(L().fdk_unet_load(str(path).encode(), C.byref(self.h)), "fdk_unet_load")
        b, f, a = C.c_int(), C.c_int(), C.c_int()
        _check(L().fdk_unet_info(self.h, C.byref(b), C.byref(f), C.byref(a)), "fdk_unet_info")
        self.base, self.s2d, self.act = b.value, f.value, a.value

    @staticmethod
    def _volume(t, name, ndim):
        torch = _torch()
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.dim() == ndim):
            raise FdkError(f"NativeUNet: {name} must be a contiguous {ndim}-D float32 CUDA tensor")

    def forward(self, x32, out=None):
        """x32 (N, D, H, W) FP32 device tensor -> x + net(x)."""
        torch = _torch()
        out = torch.empty_like(x32) if out is None else out
        self._volume(x32, "input", 4)
        self._volume(out, "output", 4)
        if out.shape != x32.shape or out.data_ptr() == x32.data_ptr():
            raise FdkError("NativeUNet: output must match the input's shape and not alias it")
        n, d, h, w = x32.shape
        _check(L().fdk_unet_forward(self.h, _p(x32), _p(out), n, d, h, w, _st()), "fdk_unet_forward")
        return out

    def infer(self, vol, out=None, slab: int = 0, overlap: int = 16):
        """(D, H, W) FP32 device volume in z-slabs (slab 0: the runner's default)."""
        torch = _torch()
        out = torch.empty_like(vol) if out is None else out
        self._volume(vol, "volume", 3)
        self._volume(out, "output", 3)
        if out.shape != vol.shape:
            raise FdkError("NativeUNet: output must match the volume's shape")
        d, h, w = vol.shape
        _check(L().fdk_unet_infer(self.h, _p(vol), _p(out), d, h, w, int(slab), int(overlap),
                                  _st()), "fdk_unet_infer")
        return out

    def close(self):
        if self.h:
            L().fdk_unet_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
