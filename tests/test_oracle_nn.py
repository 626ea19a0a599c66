"""Pins the FP64 U-Net operator oracle (oracle/nn_ref.py) on the CPU against
torch.nn.functional (library) and finite differences."""
import numpy as np
import pytest

from oracle import nn_ref as R

torch = pytest.importorskip("torch")
F = torch.nn.functional


def _torch_w(w):
    # [27][cout][cin] -> [cout][cin][3][3][3]
    return torch.as_tensor(w.transpose(1, 2, 0).reshape(w.shape[1], w.shape[2], 3, 3, 3))


def _ncdhw(x):
    return torch.as_tensor(np.ascontiguousarray(x.transpose(0, 4, 1, 2, 3)))


def test_conv3d_matches_torch():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6, 7, 3))
    w = rng.standard_normal((27, 4, 3))
    b = rng.standard_normal(4)
    want = F.conv3d(_ncdhw(x), _torch_w(w), torch.as_tensor(b), padding=1).numpy()
    got = R.conv3d(x, w, b).transpose(0, 4, 1, 2, 3)
    assert np.abs(got - want).max() < 1e-12


def test_conv3d_gradients_match_autograd_and_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 4, 5, 6, 3))
    w = rng.standard_normal((27, 2, 3))
    g = rng.standard_normal((1, 4, 5, 6, 2))
    xt = _ncdhw(x).requires_grad_(True)
    wt = _torch_w(w).requires_grad_(True)
    y = F.conv3d(xt, wt, padding=1)
    y.backward(_ncdhw(g))
    assert np.abs(R.conv3d_dx(g, w).transpose(0, 4, 1, 2, 3) - xt.grad.numpy()).max() < 1e-12
    dw = R.conv3d_dw(x, g)
    want = wt.grad.numpy().reshape(2, 3, 27).transpose(2, 0, 1)
    assert np.abs(dw - want).max() < 1e-11
    # finite difference on one weight and one input entry
    eps = 1e-6
    w2 = w.copy()
    w2[13, 1, 2] += eps
    fd = (np.sum(R.conv3d(x, w2) * g) - np.sum(R.conv3d(x, w) * g)) / eps
    assert fd == pytest.approx(dw[13, 1, 2], rel=1e-5)
    x2 = x.copy()
    x2[0, 2, 3, 4, 1] += eps
    fd = (np.sum(R.conv3d(x2, w) * g) - np.sum(R.conv3d(x, w) * g)) / eps
    assert fd == pytest.approx(R.conv3d_dx(g, w)[0, 2, 3, 4, 1], rel=1e-5)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 4, 6, 8, 3))
    small = rng.standard_normal((1, 2, 3, 4, 3))
    want = F.avg_pool3d(_ncdhw(x), 2).numpy()
    assert np.abs(R.avgpool2(x).transpose(0, 4, 1, 2, 3) - want).max() < 1e-12
    # <pool(x), s> == <x, pool^T(s)> and likewise for upsampling
    assert np.sum(R.avgpool2(x) * small) == pytest.approx(np.sum(x * R.avgpool2_adjoint(small)))
    assert np.sum(R.upsample2(small) * x) == pytest.approx(np.sum(small * R.upsample2_adjoint(x)))
    up = F.interpolate(_ncdhw(small), scale_factor=2, mode="nearest").numpy()
    assert np.abs(R.upsample2(small).transpose(0, 4, 1, 2, 3) - up).max() < 1e-12


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.01171875, 3.14159, -2.5e-3], np.float32)
    r = R.bf16_round(vals)
    # spacing 2^-7 at 1: 1 + 2^-8 is a tie (to even: 1), 1 + 3 2^-9 rounds up,
    # 1 + 3 2^-8 is a tie between odd 1 + 2^-7 and even 1 + 2^-6
    assert list(r[:4]) == [1.0, 1.0, 1.0078125, 1.015625]
    assert np.allclose(R.from_bf16_bits(R.to_bf16_bits(vals)), r)
    assert np.abs(r - vals).max() <= 2 ** -8 * np.abs(vals).max()


@pytest.mark.parametrize("f", [2, 4])
def test_space_to_depth_index_map_and_round_trip(f):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 2 * f, 3 * f, 4 * f))
    k = f ** 3
    y = R.unshuffle(x, f, k + 8)
    assert y.shape == (2, 2, 3, 4, k + 8) and not y[..., k:].any()
    for n, z, yy, xx, iz, iy, ix in np.ndindex(2, 2, 3, 4, f, f, f):
        assert y[n, z, yy, xx, (iz * f + iy) * f + ix] == x[n, f * z + iz, f * yy + iy, f * xx + ix]
    assert np.array_equal(R.shuffle(y, f), x)
    # a head whose weights pick channel j + 3 of h puts it on block entry j
    h = rng.standard_normal((1, 2, 3, 4, 80))
    w = np.zeros((k, 80))
    w[np.arange(k), np.arange(k) + 3] = 1.0
    out = R.head_shuffle(h, w, np.zeros(k), np.zeros((1, 2 * f, 3 * f, 4 * f)), f)
    assert np.array_equal(out, R.shuffle(h[..., 3:3 + k], f))


@pytest.mark.parametrize("f", [2, 4])
def test_space_to_depth_head_adjoint_matches_autograd(f):
    rng = np.random.default_rng(6)
    k = f ** 3
    h = rng.standard_normal((2, 2, 3, 4, 32))
    w = rng.standard_normal((k, 32))
    b = rng.standard_normal(k)
    x = rng.standard_normal((2, 2 * f, 3 * f, 4 * f))
    dout = rng.standard_normal(x.shape)
    th, tw, tb = (torch.as_tensor(a).requires_grad_(True) for a in (h, w, b))
    # depth-to-space through torch's own 2-D pixel_shuffle, one fine plane iz at a time
    s = (th @ tw.T + tb).permute(0, 1, 4, 2, 3)                 # n d k h w, k = iz f^2 + (iy f + ix)
    s = s.reshape(2 * 2 * f, f * f, 3, 4)                       # (n d iz) (iy ix) h w
    planes = F.pixel_shuffle(s, f)                              # (n d iz) 1 (f h) (f w)
    out = torch.as_tensor(x) + planes.reshape(2, 2 * f, 3 * f, 4 * f)
    assert np.abs(out.detach().numpy() - R.head_shuffle(h, w, b, x, f)).max() < 1e-12
    (out * torch.as_tensor(dout)).sum().backward()
    dh, dw, db = R.head_shuffle_adjoint(dout, h, w, f)
    assert np.abs(dh - th.grad.numpy()).max() < 1e-12
    assert np.abs(dw - tw.grad.numpy()).max() < 1e-10
    assert np.abs(db - tb.grad.numpy()).max() < 1e-12
