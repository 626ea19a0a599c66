"""Time U-Net training steps on the GPU (forward, data and weight gradients, Adam).

    python -m fdk_amd.train_bench [--s2d 1 2 4] [--steps 5] [--warmup 2]

For each network (space-to-depth edge 1 = full resolution, 2, 4; base 32
channels, random weights) one batch of random FP32 input and target at the
network's training shape (batch 4: 128^3 at full resolution, 256^3 at edge 2,
256 x 256 x 512 at edge 4) is stepped `--warmup` times untimed, then `--steps`
times between two events on torch's current stream; prints ms per step.
"""
from __future__ import annotations

import argparse

from . import unet as U

SHAPES = {1: (4, 128, 128, 128), 2: (4, 256, 256, 256), 4: (4, 256, 256, 512)}


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--s2d", type=int, nargs="+", default=[1, 2, 4], choices=[1, 2, 4])
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    args = ap.parse_args(argv)
    torch = U._torch()
    for f in args.s2d:
        shape = SHAPES[f]
        net = U.UNet3D(base=32, seed=1, s2d=f, act=2 if f == 4 else 1)
        g = torch.Generator(device="cuda")
        g.manual_seed(f)
        x = torch.randn(shape, device="cuda", generator=g)
        t = torch.randn(shape, device="cuda", generator=g)
        for _ in range(args.warmup):
            net.train_step(x, t, lr=1e-5)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            net.train_step(x, t, lr=1e-5)
        e1.record()
        torch.cuda.synchronize()
        print(f"s2d {f} {shape}: {e0.elapsed_time(e1) / args.steps:.2f} ms per training step", flush=True)
        del net, x, t
        torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
