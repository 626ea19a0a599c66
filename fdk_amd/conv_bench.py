"""Time one 3x3x3 BF16 convolution layer of the U-Net on the GPU.

    python -m fdk_amd.conv_bench --cin 32 --cout 32 --shape 1 128 1024 1024 [--wgrad] [--cup 64]

Random channels-last BF16 input and weights of the given shape, FP32 bias, ReLU;
the layer runs `--warmup` times untimed, then `--reps` times between two
events on torch's current stream.  Prints one line per timed form with ms and
TFLOP/s (2 * 27 * cin * cout FLOP per output voxel):

* forward (fdk_nn_conv3d_fwd, or fdk_nn_conv3d_fwd_cat with --cup: the
  decoder input [2x upsampled coarse tensor of cup channels, skip tensor of
  cin - cup channels], read in place);
* with --wgrad, the weight gradient of the same layer (fdk_nn_conv3d_wgrad, or
  fdk_nn_conv3d_wgrad_cat with --cup: the decoder input read in place).

The layer kernels are the library's; nothing here falls back to torch.
"""
from __future__ import annotations

import argparse

import numpy as np

from . import unet as U


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--cin", type=int, default=32)
    ap.add_argument("--cout", type=int, default=32)
    ap.add_argument("--shape", type=int, nargs=4, default=[1, 128, 1024, 1024],
                    metavar=("N", "D", "H", "W"))
    ap.add_argument("--cup", type=int, default=0,
                    help="decoder form: channels of the coarse (2x upsampled) input")
    ap.add_argument("--wgrad", action="store_true", help="also time the weight gradient")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    args = ap.parse_args(argv)
    torch = U._torch()
    shape = tuple(args.shape)
    n, d, h, w = shape
    cin, cout, cup = args.cin, args.cout, args.cup
    g = torch.Generator(device="cuda")
    g.manual_seed(0)

    def bf16(*s):
        return (torch.randn(s, device="cuda", generator=g) * 0.5).bfloat16().view(U.I16)

    wt = (torch.randn((27, cout, cin), device="cuda", generator=g) / np.sqrt(27 * cin)).bfloat16().view(U.I16)
    bias = torch.zeros(cout, device="cuda")
    y = torch.empty(shape + (cout,), dtype=U.I16, device="cuda")
    if cup:
        up = bf16(n, d // 2, h // 2, w // 2, cup)
        skip = bf16(*shape, cin - cup)
        fwd = lambda: U.conv3d_cat(up, cup, skip, wt, bias, y, shape, cin, cout, relu=True)
    else:
        x = bf16(*shape, cin)
        fwd = lambda: U.conv3d(x, wt, bias, y, shape, cin, cout, relu=True)
    flop = 2.0 * 27 * cin * cout * n * d * h * w

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.reps

    form = f"forward{' (decoder input, cup %d)' % cup if cup else ''}"
    ms = timed(fwd)
    print(f"conv3d {cin}->{cout} {shape} {form}: {ms:.3f} ms, {flop / ms / 1e9:.1f} TFLOP/s", flush=True)
    if args.wgrad:
        dy = bf16(*shape, cout)
        dw = torch.zeros((27, cout, cin), device="cuda")
        if cup:
            ms = timed(lambda: U.conv3d_wgrad_cat(up, cup, skip, dy, dw, shape, cin, cout))
        else:
            ms = timed(lambda: U.conv3d_wgrad(x, dy, dw, shape, cin, cout))
        form = f"weight gradient{' (decoder input, cup %d)' % cup if cup else ''}"
        print(f"conv3d {cin}->{cout} {shape} {form}: {ms:.3f} ms, "
              f"{flop / ms / 1e9:.1f} TFLOP/s", flush=True)


if __name__ == "__main__":
    main()
