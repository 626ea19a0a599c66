"""U-Net kernels on the GPU against the FP64 oracle (oracle/nn_ref.py) and the
whole network's gradients against torch float64 autograd on the CPU, all on
the same BF16-rounded weights and inputs."""
import ctypes as C

import numpy as np
import pytest

import fdk_amd.unet as U
from oracle import nn_ref as R

pytestmark = pytest.mark.gpu


def bf(torch, a):
    return torch.as_tensor(R.to_bf16_bits(a).view(np.int16)).cuda()


def from_bf(t):
    return R.from_bf16_bits(t.cpu().numpy().view(np.uint16))


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_mfma_16x16x32_bf16_layout(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(0)
    A = R.bf16_round(rng.standard_normal((16, 32)))
    B = R.bf16_round(rng.standard_normal((32, 16)))
    a = torch.as_tensor(A.astype(np.float32)).cuda()
    b = torch.as_tensor(B.astype(np.float32)).cuda()
    d = torch.zeros((16, 16), device="cuda")
    U.L().fdk_nn_mfma_probe(U._p(a), U._p(b), U._p(d), U._st())
    torch.cuda.synchronize()
    assert rel(d.cpu().numpy(), A @ B) < 1e-6


CONV_CASES = [(32, 32, 64), (32, 64, 64), (64, 128, 32), (96, 32, 128), (192, 64, 64),
              (128, 64, 32)]


@pytest.mark.parametrize("cin,cout,W", CONV_CASES)
def test_conv3d_forward(torch_cuda, cin, cout, W):
    torch = torch_cuda
    rng = np.random.default_rng(cin + cout)
    shape = (2, 3, 8, W)
    x = R.bf16_round(rng.standard_normal(shape + (cin,)))
    w = R.bf16_round(rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    y = torch.empty(shape + (cout,), dtype=torch.int16, device="cuda")
    U.conv3d(bf(torch, x), bf(torch, w), torch.as_tensor(b).cuda(), y, shape, cin, cout, relu=True)
    ref = np.maximum(R.conv3d(x, w, b), 0)
    got = from_bf(y)
    assert rel(got, ref) < 4e-3  # BF16 output rounding
    # FP32 output with a residual: FP32 accumulation only
    res = rng.standard_normal(shape + (cout,)).astype(np.float32)
    y32 = torch.empty(shape + (cout,), dtype=torch.float32, device="cuda")
    U.conv3d(bf(torch, x), bf(torch, w), None, y32, shape, cin, cout, relu=False, out_f32=True,
             resid=torch.as_tensor(res).cuda())
    assert rel(y32.cpu().numpy(), R.conv3d(x, w) + res) < 1e-5


@pytest.mark.parametrize("cin,shape", [(32, (2, 6, 12, 96)), (96, (2, 6, 12, 96)),
                                       (64, (1, 2, 4, 96)), (32, (2, 8, 12, 96)),
                                       (96, (1, 4, 4, 32)), (64, (1, 12, 8, 64))])
def test_conv3d_forward_staged_tiles(torch_cuda, cin, shape):
    """Even-depth shapes, which take the LDS-halo kernels (two tiles per block;
    cout 128 as two 64-channel passes): interior and border tiles, both batch
    items, one to three 32-channel chunks, and an odd tile count (3 tiles: the
    last block has one).  Depth 6 and 2 run the 2-plane tiles; depth 8, 4 and
    12 the row-reuse tiles (4 planes at cout 32, 2 planes per 64-channel pass;
    cout 96 as a 64-channel pass and a 32-channel pass)."""
    torch = torch_cuda
    rng = np.random.default_rng(cin + shape[0])
    for cout in (32, 64, 96, 128):
        x = R.bf16_round(rng.standard_normal(shape + (cin,)))
        w = R.bf16_round(rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin))
        b = rng.standard_normal(cout).astype(np.float32)
        y = torch.empty(shape + (cout,), dtype=torch.int16, device="cuda")
        U.conv3d(bf(torch, x), bf(torch, w), torch.as_tensor(b).cuda(), y, shape, cin, cout,
                 relu=True)
        z = R.conv3d(x, w, b)
        assert rel(from_bf(y), np.maximum(z, 0)) < 4e-3
        U.conv3d(bf(torch, x), bf(torch, w), torch.as_tensor(b).cuda(), y, shape, cin, cout,
                 relu=2)  # leaky ReLU
        assert rel(from_bf(y), np.where(z > 0, z, 0.01 * z)) < 4e-3
        res = rng.standard_normal(shape + (cout,)).astype(np.float32)
        y32 = torch.empty(shape + (cout,), dtype=torch.float32, device="cuda")
        U.conv3d(bf(torch, x), bf(torch, w), None, y32, shape, cin, cout, relu=False,
                 out_f32=True, resid=torch.as_tensor(res).cuda())
        assert rel(y32.cpu().numpy(), R.conv3d(x, w) + res) < 1e-5


@pytest.mark.parametrize("cup,cskip,cout,shape", [(64, 32, 32, (2, 8, 12, 96)),
                                                  (128, 64, 64, (1, 8, 12, 96)),
                                                  (64, 32, 32, (1, 2, 4, 32)),
                                                  (64, 32, 96, (1, 8, 12, 64)),
                                                  (64, 32, 96, (1, 6, 8, 32))])
def test_conv3d_decoder_input_fused(torch_cuda, cup, cskip, cout, shape):
    """conv3d of [nearest 2x upsampling of a coarse tensor, skip] read from the two
    tensors (fdk_nn_conv3d_fwd_cat) is bit-identical to upsample + channel copy +
    conv3d, on interior and border tiles of both batch items, and matches the
    oracle on the formed input."""
    torch = torch_cuda
    rng = np.random.default_rng(cup + cout + shape[0])
    n, d, h, w = shape
    cin = cup + cskip
    small = (n, d // 2, h // 2, w // 2)
    up = R.bf16_round(rng.standard_normal(small + (cup,)))
    skip = R.bf16_round(rng.standard_normal(shape + (cskip,)))
    wt = R.bf16_round(rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    up_d, skip_d, w_d, b_d = bf(torch, up), bf(torch, skip), bf(torch, wt), torch.as_tensor(b).cuda()
    cat = torch.empty(shape + (cin,), dtype=torch.int16, device="cuda")
    U.upsample(up_d, cat, small, cup, cin, 0)
    U.channel_copy(skip_d, cskip, 0, cat, cin, cup, cskip, n * d * h * w)
    ref = torch.empty(shape + (cout,), dtype=torch.int16, device="cuda")
    U.conv3d(cat, w_d, b_d, ref, shape, cin, cout, relu=True)
    got = torch.empty_like(ref)
    U.conv3d_cat(up_d, cup, skip_d, w_d, b_d, got, shape, cin, cout, relu=True)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    full = np.concatenate([up.repeat(2, 1).repeat(2, 2).repeat(2, 3), skip], axis=-1)
    assert rel(from_bf(got), np.maximum(R.conv3d(full, wt, b), 0)) < 4e-3


def test_unet_inference_fused_decoder_matches_formed(torch_cuda):
    """Inference (decoder inputs read from the coarse and skip tensors) gives the
    same output as the forward pass that forms them for backward()."""
    torch = torch_cuda
    net = U.UNet3D(base=32, seed=8)
    with torch.no_grad():
        net.hw.mul_(300.0)
    x = torch.as_tensor(np.random.default_rng(9).standard_normal((1, 8, 16, 128)).astype(np.float32)).cuda()
    fused = net.forward(x).clone()
    kept = net.forward(x, keep=True).clone()
    net.fuse_decoder = False
    formed = net.forward(x, keep=True)
    torch.cuda.synchronize()
    assert torch.equal(fused, formed)
    assert torch.equal(kept, formed)


@pytest.mark.parametrize("s2d", [1, 4])
def test_unet_gradients_with_fused_decoder_inputs(torch_cuda, s2d):
    """Training with the decoder inputs read in place (forward and weight
    gradients through fdk_nn_conv3d_wgrad_cat) gives the gradients of the
    formed-concatenation path: the same kernels on the same values, up to the
    order of the blocks' FP32 atomics."""
    torch = torch_cuda
    shape = (1, 8, 16, 128) if s2d == 1 else (1, 16, 64, 512)
    x = torch.as_tensor(np.random.default_rng(21).standard_normal(shape).astype(np.float32)).cuda()
    t = torch.as_tensor(np.random.default_rng(22).standard_normal(shape).astype(np.float32)).cuda()
    grads = []
    for fuse in (True, False):
        net = U.UNet3D(base=32, seed=5, s2d=s2d)
        net.fuse_decoder = fuse
        net.zero_grad()
        out = net.forward(x, keep=True)
        dout = torch.empty_like(out)
        loss = torch.zeros(1, device="cuda")
        U.mse(out, t, dout, loss)
        net.backward(dout)
        torch.cuda.synchronize()
        grads.append([cv.gw.cpu().numpy().copy() for cv in net.convs])
    for a, b in zip(*grads):
        assert rel(a, b) < 1e-5


@pytest.mark.parametrize("shape,cout", [((1, 4, 8, 64), 32), ((2, 3, 8, 96), 32),
                                         ((1, 5, 4, 128), 64)])
def test_conv3d_forward_single_input_channel(torch_cuda, shape, cout):
    """The network's first convolution (one input channel, the 27 taps as K):
    border and interior x steps of the row kernel, both batch items, FP32 output
    with a residual and BF16 output with bias and activation."""
    torch = torch_cuda
    rng = np.random.default_rng(5 + cout)
    x = R.bf16_round(rng.standard_normal(shape + (1,)))
    w27 = R.bf16_round(rng.standard_normal((27, cout, 1)) / 5)
    wpad = np.zeros((cout, 32))
    wpad[:, :27] = w27[:, :, 0].T
    res = rng.standard_normal(shape + (cout,)).astype(np.float32)
    y = torch.empty(shape + (cout,), dtype=torch.float32, device="cuda")
    U.conv3d(bf(torch, x[..., 0]), bf(torch, wpad), None, y, shape, 1, cout, relu=False,
             out_f32=True, resid=torch.as_tensor(res).cuda())
    assert rel(y.cpu().numpy(), R.conv3d(x, w27) + res) < 1e-5
    b = rng.standard_normal(cout).astype(np.float32)
    yb = torch.empty(shape + (cout,), dtype=torch.int16, device="cuda")
    U.conv3d(bf(torch, x[..., 0]), bf(torch, wpad), torch.as_tensor(b).cuda(), yb, shape, 1, cout,
             relu=2)
    z = R.conv3d(x, w27, b)
    assert rel(from_bf(yb), np.where(z > 0, z, 0.01 * z)) < 4e-3


@pytest.mark.parametrize("cin,cout,W", [(32, 64, 64), (192, 64, 64), (96, 32, 128), (64, 128, 32)])
def test_conv3d_input_and_weight_gradients(torch_cuda, cin, cout, W):
    torch = torch_cuda
    rng = np.random.default_rng(cin * 7 + cout)
    shape = (2, 3, 8, W)
    x = R.bf16_round(rng.standard_normal(shape + (cin,)))
    dy = R.bf16_round(rng.standard_normal(shape + (cout,)))
    w = R.bf16_round(rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin))
    wf = torch.empty((27, cin, cout), dtype=torch.int16, device="cuda")
    U.flip_weights(bf(torch, w), wf, cin, cout)
    dx = torch.empty(shape + (cin,), dtype=torch.float32, device="cuda")
    U.conv3d(bf(torch, dy), wf, None, dx, shape, cout, cin, relu=False, out_f32=True)
    assert rel(dx.cpu().numpy(), R.conv3d_dx(dy, w)) < 1e-5
    dw = torch.zeros((27, cout, cin), device="cuda")
    U.conv3d_wgrad(bf(torch, x), bf(torch, dy), dw, shape, cin, cout)
    U.conv3d_wgrad(bf(torch, x), bf(torch, dy), dw, shape, cin, cout)  # accumulates
    assert rel(dw.cpu().numpy(), 2 * R.conv3d_dw(x, dy)) < 1e-5


@pytest.mark.parametrize("cin,cout,shape", [(32, 32, (2, 4, 8, 64)), (96, 32, (1, 6, 12, 96)),
                                            (64, 128, (2, 2, 4, 32)), (192, 64, (1, 4, 8, 64)),
                                            (32, 64, (1, 6, 12, 96))])
def test_conv3d_weight_gradient_tiled(torch_cuda, cin, cout, shape):
    """The tiled weight gradient (32-channel chunk pairs, all 27 taps per block,
    operands through the transposing LDS read): interior and border tiles, both
    batch items, several chunks of cin and cout, accumulation into dW."""
    torch = torch_cuda
    rng = np.random.default_rng(cin + 3 * cout + shape[0])
    x = R.bf16_round(rng.standard_normal(shape + (cin,)))
    dy = R.bf16_round(rng.standard_normal(shape + (cout,)))
    dw = torch.zeros((27, cout, cin), device="cuda")
    U.conv3d_wgrad(bf(torch, x), bf(torch, dy), dw, shape, cin, cout)
    U.conv3d_wgrad(bf(torch, x), bf(torch, dy), dw, shape, cin, cout)  # accumulates
    assert rel(dw.cpu().numpy(), 2 * R.conv3d_dw(x, dy)) < 1e-5


@pytest.mark.parametrize("shape", [(1, 4, 8, 64), (2, 6, 12, 128), (1, 3, 8, 32)])
def test_conv3d_weight_gradient_single_input_channel(torch_cuda, shape):
    """The tiled form (W % 64, H % 4, D % 2: border and interior tiles, both
    batch items, several tiles per block) and the per-voxel form (odd depth)."""
    torch = torch_cuda
    rng = np.random.default_rng(9 + shape[1])
    x = R.bf16_round(rng.standard_normal(shape + (1,)))
    dy = R.bf16_round(rng.standard_normal(shape + (32,)))
    dw = torch.zeros((32, 32), device="cuda")
    U.conv3d_wgrad(bf(torch, x[..., 0]), bf(torch, dy), dw, shape, 1, 32)
    got = dw.cpu().numpy()
    assert np.abs(got[:, 27:]).max() == 0.0
    assert rel(got[:, :27], R.conv3d_dw(x, dy)[:, :, 0].T) < 1e-5


@pytest.mark.parametrize("c", [16, 32, 64, 128])
def test_bias_gradient(torch_cuda, c):
    """Per-channel sums of a BF16 gradient (16-byte vector form for 32, 64 and
    128 channels, the scalar form for 16), accumulated into db."""
    torch = torch_cuda
    rng = np.random.default_rng(c)
    nvox = 3 * 4099
    dy = R.bf16_round(rng.standard_normal((nvox, c)))
    db = torch.ones(c, device="cuda")
    U.bias_grad(bf(torch, dy), db, nvox, c)
    assert rel(db.cpu().numpy(), 1.0 + dy.sum(0)) < 1e-5


@pytest.mark.parametrize("n,mode", [(8 * 5003, 1), (8 * 5003, 2), (5003, 1), (5003, 2)])
def test_activation_backward(torch_cuda, n, mode):
    """dz = dy where y > 0, else 0 (ReLU) or 0.01 dy rounded to BF16 (leaky
    ReLU); the 16-byte vector form (n % 8 == 0) and the scalar form."""
    torch = torch_cuda
    rng = np.random.default_rng(n + mode)
    dy = R.bf16_round(rng.standard_normal(n))
    y = R.bf16_round(rng.standard_normal(n))
    y[::7] = 0.0
    dz = torch.empty(n, dtype=torch.int16, device="cuda")
    U.relu_bwd(bf(torch, dy), bf(torch, y), dz, mode)
    leak = R.bf16_round(np.float32(0.01) * dy.astype(np.float32))  # FP32 product, then RNE
    ref = np.where(y > 0, dy, 0.0 if mode == 1 else leak)
    assert np.array_equal(from_bf(dz), ref)


@pytest.mark.parametrize("c", [32, 24])
def test_head_backward(torch_cuda, c):
    """The 1x1 head's adjoint (16-byte vector form at 32 features, the scalar
    form otherwise): feature gradients dout * w (BF16) and the weight and bias
    gradients accumulated in FP32."""
    torch = torch_cuda
    rng = np.random.default_rng(c + 1)
    nvox = 2 * 4097
    h = R.bf16_round(rng.standard_normal((nvox, c)))
    dout = rng.standard_normal(nvox).astype(np.float32)
    w = rng.standard_normal(c).astype(np.float32)
    dh = torch.empty((nvox, c), dtype=torch.int16, device="cuda")
    dw = torch.zeros(c, device="cuda")
    db = torch.zeros(1, device="cuda")
    t = lambda a: torch.as_tensor(a).cuda()
    d_out, w_d, h_d = t(dout), t(w), bf(torch, h)  # alive until the kernel has run
    U.L().fdk_nn_head_bwd(U._p(d_out), U._p(h_d), U._p(w_d), U._p(dh), U._p(dw), U._p(db),
                          nvox, c, U._st())
    torch.cuda.synchronize()
    assert rel(from_bf(dh), dout[:, None] * w[None, :]) < 4e-3
    assert rel(dw.cpu().numpy(), dout @ h) < 1e-5
    assert abs(float(db.cpu()[0]) - float(dout.sum())) < 1e-4 * np.abs(dout).sum()


@pytest.mark.parametrize("c", [32, 12])
def test_pool_and_upsample_adjoints_and_add(torch_cuda, c):
    """The adjoints of 2x average pooling and of 2x nearest upsampling, and the
    BF16 add, in their 16-byte vector forms (c % 8 == 0) and their scalar forms
    (c = 12), against the oracle."""
    torch = torch_cuda
    rng = np.random.default_rng(c + 5)
    full = (2, 4, 6, 10)
    s = R.bf16_round(rng.standard_normal((2, 2, 3, 5, c)))
    pa = torch.empty(full + (c,), dtype=torch.int16, device="cuda")
    U.avgpool(bf(torch, s), pa, full, c, adjoint=True)
    assert np.array_equal(from_bf(pa), R.bf16_round(R.avgpool2_adjoint(s)))
    fine = R.bf16_round(rng.standard_normal(full + (c + 8,)))
    up_adj = torch.empty((2, 2, 3, 5, c), dtype=torch.int16, device="cuda")
    U.upsample(bf(torch, fine), up_adj, (2, 2, 3, 5), c, c + 8, 8, adjoint=True)
    assert rel(from_bf(up_adj), R.upsample2_adjoint(fine[..., 8:])) < 4e-3
    a = R.bf16_round(rng.standard_normal(8 * 1001 if c == 32 else 1001))
    b = R.bf16_round(rng.standard_normal(a.shape))
    y = torch.empty(a.shape, dtype=torch.int16, device="cuda")
    a_d, b_d = bf(torch, a), bf(torch, b)  # alive until the kernel has run
    U.add(a_d, b_d, y)
    torch.cuda.synchronize()
    assert np.array_equal(from_bf(y), R.bf16_round((a.astype(np.float32) + b.astype(np.float32))))


def test_pool_upsample_copy_head_mse_adam(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(11)
    full = (1, 4, 8, 16)
    x = R.bf16_round(rng.standard_normal(full + (32,)))
    p = torch.empty((1, 2, 4, 8, 32), dtype=torch.int16, device="cuda")
    U.avgpool(bf(torch, x), p, full, 32)
    assert rel(from_bf(p), R.avgpool2(x)) < 4e-3
    s = R.bf16_round(rng.standard_normal((1, 2, 4, 8, 32)))
    pa = torch.empty(full + (32,), dtype=torch.int16, device="cuda")
    U.avgpool(bf(torch, s), pa, full, 32, adjoint=True)
    assert rel(from_bf(pa), R.avgpool2_adjoint(s)) < 4e-3
    cat = torch.zeros(full + (48,), dtype=torch.int16, device="cuda")
    U.upsample(bf(torch, s), cat, (1, 2, 4, 8), 32, 48, 16)
    assert rel(from_bf(cat)[..., 16:], R.upsample2(s)) == 0.0
    up_adj = torch.empty((1, 2, 4, 8, 32), dtype=torch.int16, device="cuda")
    U.upsample(cat, up_adj, (1, 2, 4, 8), 32, 48, 16, adjoint=True)
    assert rel(from_bf(up_adj), R.upsample2_adjoint(from_bf(cat)[..., 16:])) < 4e-3
    h = R.bf16_round(rng.standard_normal(full + (32,)))
    hw = rng.standard_normal(32).astype(np.float32)
    hb = np.array([0.25], np.float32)
    xin = rng.standard_normal(full).astype(np.float32)
    out = torch.empty(full, device="cuda")
    # keep every argument tensor alive until the kernel has run
    args = [bf(torch, h), torch.as_tensor(hw).cuda(), torch.as_tensor(hb).cuda(),
            torch.as_tensor(xin).cuda()]
    U.L().fdk_nn_head(*[U._p(a) for a in args], U._p(out), out.numel(), 32, U._st())
    ref_out = xin + h @ hw + 0.25
    assert rel(out.cpu().numpy(), ref_out) < 1e-6
    tgt = rng.standard_normal(full).astype(np.float32)
    dout = torch.empty(full, device="cuda")
    loss = torch.zeros(1, device="cuda")
    U.mse(out, torch.as_tensor(tgt).cuda(), dout, loss)
    torch.cuda.synchronize()
    assert loss.item() == pytest.approx(np.sum((ref_out - tgt) ** 2), rel=1e-4)
    assert rel(dout.cpu().numpy(), 2 * (ref_out - tgt) / ref_out.size) < 1e-5
    # Adam, two steps against a numpy restatement
    w = rng.standard_normal(1000).astype(np.float32)
    g1, g2 = rng.standard_normal((2, 1000)).astype(np.float32)
    tw, tm, tv = (torch.as_tensor(w).cuda(), torch.zeros(1000, device="cuda"),
                  torch.zeros(1000, device="cuda"))
    wbf = torch.empty(1000, dtype=torch.int16, device="cuda")
    mw, vw, ww = np.zeros(1000), np.zeros(1000), w.astype(np.float64)
    for t, g in ((1, g1), (2, g2)):
        U.adam(tw, wbf, torch.as_tensor(g).cuda(), tm, tv, 1e-3, t)
        mw = 0.9 * mw + 0.1 * g
        vw = 0.999 * vw + 0.001 * g.astype(np.float64) ** 2
        ww = ww - 1e-3 * (mw / (1 - 0.9 ** t)) / (np.sqrt(vw / (1 - 0.999 ** t)) + 1e-8)
    torch.cuda.synchronize()
    assert np.abs(tw.cpu().numpy() - ww).max() < 1e-6
    assert np.abs(from_bf(wbf) - R.bf16_round(tw.cpu().numpy())).max() == 0.0


@pytest.mark.parametrize("f,full", [(2, (1, 6, 10, 14)), (2, (1, 128, 128, 136)),
                                    (4, (1, 12, 20, 28)), (4, (1, 256, 128, 160))])
def test_space_to_depth_kernels(torch_cuda, f, full):
    """unshuffle bit-exact; the f^3-output head and its adjoint against the FP64
    oracle (the larger shapes run the adjoint's tile loop more than once)."""
    torch = torch_cuda
    rng = np.random.default_rng(12 + f)
    n, D, H, W = full
    k = f ** 3
    coarse = (n, D // f, H // f, W // f)
    x = rng.standard_normal(full).astype(np.float32)
    xd = torch.as_tensor(x).cuda()
    cy = max(32, k)
    y = torch.empty(coarse + (cy,), dtype=torch.int16, device="cuda")
    U.unshuffle(xd, y, f)
    assert np.array_equal(from_bf(y), R.unshuffle(R.bf16_round(x), f, cy))
    h = R.bf16_round(rng.standard_normal(coarse + (32,)))
    w = rng.standard_normal((k, 32)).astype(np.float32)
    b = rng.standard_normal(k).astype(np.float32)
    hd, wd, bd = bf(torch, h), torch.as_tensor(w).cuda(), torch.as_tensor(b).cuda()
    out = torch.empty(full, device="cuda")
    U.head_shuffle(hd, wd, bd, xd, out, f)
    assert rel(out.cpu().numpy(), R.head_shuffle(h, w.astype(np.float64), b, x, f)) < 1e-6
    dout = rng.standard_normal(full).astype(np.float32)
    dh = torch.empty_like(hd)
    dw = torch.zeros((k, 32), device="cuda")
    db = torch.zeros(k, device="cuda")
    U.head_shuffle_bwd(torch.as_tensor(dout).cuda(), hd, wd, dh, dw, db, f)
    torch.cuda.synchronize()
    rdh, rdw, rdb = R.head_shuffle_adjoint(dout.astype(np.float64), h, w.astype(np.float64), f)
    assert rel(from_bf(dh), rdh) < 4e-3  # BF16 output
    assert rel(dw.cpu().numpy(), rdw) < 1e-5
    assert rel(db.cpu().numpy(), rdb) < 1e-5


def _torch_unet(net, x):
    """The same network in torch float64 on the CPU, weights taken from net."""
    import torch
    Fn = torch.nn.functional

    def W(cv):
        w = cv.w.cpu().numpy().astype(np.float64)
        if cv.cin == 1:
            w = w[:, :27].T[:, :, None]                       # (27, cout, 1)
        w = R.bf16_round(w)
        t = torch.as_tensor(w.transpose(1, 2, 0).reshape(cv.cout, cv.cin, 3, 3, 3))
        return t.requires_grad_(True)

    params = {}

    def conv(cv, h):
        w = W(cv)
        b = torch.as_tensor(cv.b.cpu().numpy().astype(np.float64)).requires_grad_(True)
        params[id(cv)] = (w, b)
        z = Fn.conv3d(h, w, b, padding=1)
        return torch.relu(z) if net.act == 1 else Fn.leaky_relu(z, 0.01)

    f = net.s2d
    if f > 1:
        xt = torch.as_tensor(np.ascontiguousarray(
            R.unshuffle(R.bf16_round(x), f, net.e1a.cin).transpose(0, 4, 1, 2, 3)))
    else:
        xt = torch.as_tensor(R.bf16_round(x)[:, None]).requires_grad_(False)
    e1 = conv(net.e1b, conv(net.e1a, xt))
    e2 = conv(net.e2b, conv(net.e2a, Fn.avg_pool3d(e1, 2)))
    bt = conv(net.bb, conv(net.ba, Fn.avg_pool3d(e2, 2)))
    d2 = conv(net.d2b, conv(net.d2a, torch.cat([Fn.interpolate(bt, scale_factor=2), e2], 1)))
    d1 = conv(net.d1b, conv(net.d1a, torch.cat([Fn.interpolate(d2, scale_factor=2), e1], 1)))
    hw = torch.as_tensor(net.hw.cpu().numpy().astype(np.float64)).requires_grad_(True)
    hb = torch.as_tensor(net.hb.cpu().numpy().astype(np.float64)).requires_grad_(True)
    if f > 1:  # f^3 outputs per coarse voxel, block entry k = (iz f + iy) f + ix
        s = torch.einsum("ncdhw,kc->nkdhw", d1, hw) + hb[None, :, None, None, None]
        n, _, d, h, w = s.shape
        s = s.reshape(n, f, f, f, d, h, w).permute(0, 4, 1, 5, 2, 6, 3)
        out = torch.as_tensor(x.astype(np.float64)) + s.reshape(n, f * d, f * h, f * w)
        return out, params, hw, hb
    out = torch.as_tensor(x[:, None].astype(np.float64)) + \
        torch.einsum("ncdhw,c->ndhw", d1, hw)[:, None] + hb
    return out[:, 0], params, hw, hb


@pytest.mark.parametrize("s2d,act", [(1, 1), (2, 1), (4, 1), (4, 2)])
def test_unet_forward_and_gradients_match_torch(torch_cuda, s2d, act):
    torch = torch_cuda
    net = U.UNet3D(base=32, seed=3, s2d=s2d, act=act)
    with torch.no_grad():
        net.hw.mul_(300.0)  # a head large enough that the net's output is not just x
    rng = np.random.default_rng(4)
    shape = (1, 8, 16, 128) if s2d == 1 else (1,) + net.min_shape
    x = rng.standard_normal(shape).astype(np.float32)
    tgt = rng.standard_normal(shape).astype(np.float32)
    xd = torch.as_tensor(x).cuda()
    out = net.forward(xd, keep=True)
    dout = torch.empty_like(out)
    loss = torch.zeros(1, device="cuda")
    U.mse(out, torch.as_tensor(tgt).cuda(), dout, loss)
    net.zero_grad()
    net.backward(dout)
    torch.cuda.synchronize()

    ref_out, params, hw, hb = _torch_unet(net, x)
    ref_loss = ((ref_out - torch.as_tensor(tgt.astype(np.float64))) ** 2).mean()
    ref_loss.backward()
    got = out.cpu().numpy() - x
    want = ref_out.detach().numpy() - x
    assert rel(got, want) < 3e-2  # BF16 activations between layers
    cos = []
    for cv in net.convs:
        w, b = params[id(cv)]
        gw_ref = w.grad.numpy().reshape(cv.cout, cv.cin, 27).transpose(2, 0, 1)
        gw = cv.gw.cpu().numpy()
        if cv.cin == 1:
            gw = gw[:, :27].T[:, :, None]
        c = float(np.sum(gw * gw_ref) / (np.linalg.norm(gw) * np.linalg.norm(gw_ref)))
        cos.append(c)
        assert c > 0.99, (cv.cin, cv.cout, c)
        assert rel(gw, gw_ref) < 0.1
        assert rel(cv.gb.cpu().numpy(), b.grad.numpy()) < 0.1
    assert rel(net.ghw.cpu().numpy(), hw.grad.numpy()) < 3e-2
    print("weight-gradient cosine similarity per conv:", [round(c, 5) for c in cos])


@pytest.mark.parametrize("s2d,act", [(1, 1), (2, 1), (4, 1), (4, 2)])
def test_unet_training_reduces_loss(torch_cuda, s2d, act):
    torch = torch_cuda
    net = U.UNet3D(base=32, seed=5, s2d=s2d, act=act)
    rng = np.random.default_rng(6)
    # a smooth target and a noisy, blurred input: the net must learn a filter
    d, h, w = (16, 16, 128) if s2d == 1 else (max(16, net.min_shape[0]),) + net.min_shape[1:]
    z, y, x = np.meshgrid(np.linspace(-1, 1, d), np.linspace(-1, 1, h), np.linspace(-1, 1, w),
                          indexing="ij")
    clean = (np.sin(3 * x) * np.cos(2 * y) + z)[None].astype(np.float32)
    clean = np.concatenate([clean, -clean])
    if s2d < 4:
        noisy = clean + 0.3 * rng.standard_normal(clean.shape).astype(np.float32)
    else:
        # 32 features per coarse voxel cannot undo 64 independent per-voxel errors:
        # a smooth corruption instead
        noisy = clean + (0.6 * np.cos(4 * x) * np.sin(3 * z) + 0.2)[None].astype(np.float32)
    xd, td = torch.as_tensor(noisy).cuda(), torch.as_tensor(clean).cuda()
    # the coarse-grid nets start further from a smoother; single steps are noisy
    # (atomic weight-gradient sums), so the last ten are averaged
    steps = {1: 40, 2: 240, 4: 120}[s2d]
    losses = [net.train_step(xd, td, lr=1e-3).item() for _ in range(steps)]
    tail = float(np.mean(losses[-10:]))
    print("training loss:", [round(l, 1) for l in losses[::steps // 5]], "-> last ten", round(tail, 1))
    assert tail < 0.5 * losses[0]


@pytest.mark.parametrize("fname", ["unet_sparse1024_s2d4.npz", "unet_sparse1024_s2d2.npz",
                                   "unet_sparse1024.npz"])
def test_native_runner_matches_python_sequencing(torch_cuda, fname):
    """fdk_unet_load / fdk_unet_forward (weights read and the pass sequenced in
    C++) give the same output, bit for bit, as UNet3D.from_file on the same file,
    and so does whole-volume slab inference against learned.infer_slabs."""
    import os
    import fdk_amd.learned as Lr
    torch = torch_cuda
    path = os.path.join(os.path.dirname(U.__file__), "weights", fname)
    net = U.UNet3D.from_file(path)
    nat = U.NativeUNet(path)
    assert (nat.base, nat.s2d, nat.act) == (32, net.s2d, net.act)
    f = net.s2d
    shape = (2, 4 * f, 16 * f, 128 * f)
    x = torch.as_tensor(np.random.default_rng(11).standard_normal(shape).astype(np.float32)).cuda()
    want = net.forward(x)
    got = nat.forward(x)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    # slabs: 8 f deep with 4 f of overlap over a 28 f-deep volume (6 slabs, the
    # last flush with the end), against the Python slab loop
    vol = torch.as_tensor(np.random.default_rng(12).standard_normal((28 * f, 16 * f, 128 * f))
                          .astype(np.float32)).cuda()
    want = Lr.infer_slabs(net, vol, slab=8 * f, overlap=4 * f)
    got = nat.infer(vol, slab=8 * f, overlap=4 * f)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    whole = nat.infer(vol, slab=28 * f)
    torch.cuda.synchronize()
    assert torch.equal(whole, net.forward(vol.unsqueeze(0))[0])
    nat.close()
