"""TEST INFRASTRUCTURE ONLY -- FP64 oracle of the U-Net operators.

Only tests/ may import this module, to check the library's BF16 MFMA kernels.
Channels-last tensors [n][z][y][x][c]; 3x3x3 'same' convolution with zero
padding and weights [27][cout][cin] (tap = 9 (dz+1) + 3 (dy+1) + (dx+1)).
The reference has no learned component; these operators are pinned by
tests/test_oracle_nn.py against torch.nn.functional (a CPU library) and by
finite differences.
"""
import numpy as np

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    b = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def to_bf16_bits(a):
    b = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _shift(x, dz, dy, dx):
    """out[z, y, x] = x[z+dz, y+dy, x+dx] with zeros outside (axes 1..3)."""
    out = np.zeros_like(x)
    n, D, H, W = x.shape[:4]
    zs, ys, xs = slice(max(0, -dz), D - max(0, dz)), slice(max(0, -dy), H - max(0, dy)), \
        slice(max(0, -dx), W - max(0, dx))
    zt, yt, xt = slice(max(0, dz), D - max(0, -dz)), slice(max(0, dy), H - max(0, -dy)), \
        slice(max(0, dx), W - max(0, -dx))
    out[:, zs, ys, xs] = x[:, zt, yt, xt]
    return out


def conv3d(x, w, b=None):
    """x (n,D,H,W,cin), w (27,cout,cin) -> (n,D,H,W,cout)."""
    y = np.zeros(x.shape[:4] + (w.shape[1],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += _shift(x, dz, dy, dx) @ w[t].T
    if b is not None:
        y += b
    return y


def conv3d_dx(dy, w):
    dx = np.zeros(dy.shape[:4] + (w.shape[2],))
    for t, (dz, dy_, dx_) in enumerate(TAPS):
        dx += _shift(dy, -dz, -dy_, -dx_) @ w[t]
    return dx


def conv3d_dw(x, dy):
    n, D, H, W, cin = x.shape
    cout = dy.shape[-1]
    dw = np.zeros((27, cout, cin))
    dyf = dy.reshape(-1, cout)
    for t, (dz, dy_, dx_) in enumerate(TAPS):
        dw[t] = dyf.T @ _shift(x, dz, dy_, dx_).reshape(-1, cin)
    return dw


def avgpool2(x):
    n, D, H, W, c = x.shape
    return x.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, c).mean(axis=(2, 4, 6))


def avgpool2_adjoint(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def upsample2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def upsample2_adjoint(dy):
    n, D, H, W, c = dy.shape
    return dy.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, c).sum(axis=(2, 4, 6))


def unshuffle(x, f=2, cy=None):
    """Space-to-depth, block edge f: FP32 [n][f d][f h][f w] -> [n][d][h][w][cy];
    channel (iz f + iy) f + ix holds voxel (f z + iz, f y + iy, f x + ix),
    channels f^3.. zero (cy defaults to f^3)."""
    n, D, H, W = x.shape
    k = f ** 3
    cy = k if cy is None else cy
    b = x.reshape(n, D // f, f, H // f, f, W // f, f).transpose(0, 1, 3, 5, 2, 4, 6)
    out = np.zeros((n, D // f, H // f, W // f, cy))
    out[..., :k] = b.reshape(n, D // f, H // f, W // f, k)
    return out


def shuffle(y, f=2):
    """Depth-to-space, the inverse of unshuffle on channels 0..f^3-1."""
    n, d, h, w = y.shape[:4]
    b = y[..., :f ** 3].reshape(n, d, h, w, f, f, f).transpose(0, 1, 4, 2, 5, 3, 6)
    return b.reshape(n, f * d, f * h, f * w)


def head_shuffle(h, w, b, x, f=2):
    """out = x + shuffle(h w^T + b): h [n][d][h][w][c], w [f^3][c], b [f^3]."""
    return x + shuffle(h @ w.T + b, f)


def head_shuffle_adjoint(dout, h, w, f=2):
    """(dh, dw, db) of head_shuffle for the output gradient dout."""
    k = f ** 3
    g = unshuffle(dout, f)
    c = h.shape[-1]
    return g @ w, g.reshape(-1, k).T @ h.reshape(-1, c), g.reshape(-1, k).sum(0)
