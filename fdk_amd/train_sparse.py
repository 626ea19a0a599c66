"""Train the sparse-view U-Net on the 1024^3 / 1024-view scan and evaluate it.

    python -m fdk_amd.train_sparse --phantoms 200 --steps 600 --out fdk_amd/weights/unet_sparse1024.npz
    python -m fdk_amd.train_sparse --s2d 2 --phantoms 480 --max-seconds 900 \
        --out fdk_amd/weights/unet_sparse1024_s2d2.npz

Pairs: FDK from every 4th of 1024 noisy views (I0 = 1e5) and FDK from all 1024,
over random boxes of seeded ellipsoid phantoms: 256^3 boxes cut into 128^3
patches, or whole 256^3 (``--s2d 2``) or 256 x 256 x 512 (``--s2d 4``) boxes for
the space-to-depth networks.  Adam (lr 1e-3), batch 4.  The held-out comparison (``--eval-seeds``) runs on the whole
1024^3 volume.  Everything runs on the GPU in the library's kernels.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--views", type=int, default=1024)
    ap.add_argument("--phantoms", type=int, default=200)
    ap.add_argument("--region", type=int, default=256)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--max-seconds", type=float, default=None, help="training time budget")
    ap.add_argument("--init", default=None, help="start from saved weights")
    ap.add_argument("--s2d", type=int, default=1, choices=(1, 2, 4),
                    help="space-to-depth block edge (the network runs on a grid that much coarser)")
    ap.add_argument("--act", type=int, default=1, choices=(1, 2), help="1 ReLU, 2 leaky ReLU")
    ap.add_argument("--warmup", type=float, default=0.0,
                    help="fraction of the run over which the learning rate ramps up linearly")
    ap.add_argument("--patch", default=None,
                    help="training patch, an edge or D,H,W (default 128; 256 for --s2d 2; "
                         "256,256,512 for --s2d 4)")
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[100000])
    ap.add_argument("--report", default=None, help="write the evaluation as JSON")
    a = ap.parse_args(argv)

    import numpy as np
    import torch
    from . import learned as Lr
    from .unet import UNet3D

    def log(m):
        print(f"[train_sparse] {m}", flush=True)

    p = Lr.geometry(a.n, a.views)
    sc = Lr.Scanner(p)
    if a.patch:
        patch = tuple(int(v) for v in a.patch.split(","))
        patch = patch * 3 if len(patch) == 1 else patch
    else:
        patch = {1: (Lr.PATCH,) * 3, 2: (256,) * 3, 4: (256, 256, 512)}[a.s2d]
    if a.init:
        net = UNet3D.from_file(a.init)
        if (net.s2d, net.act) != (a.s2d, a.act):
            raise SystemExit(f"{a.init} holds s2d={net.s2d}, act={net.act} weights")
    else:
        net = UNet3D(base=32, seed=0, s2d=a.s2d, act=a.act)
    report = {"geometry": {"n": a.n, "views": a.views, "voxel_mm": p.dx, "pixel_mm": p.du}}
    if a.steps > 0:
        t0 = time.perf_counter()
        region = patch if a.s2d > 1 else (a.region,) * 3
        xs, ys = Lr.make_pairs(sc, list(range(a.phantoms)), region=region,
                               rng=np.random.default_rng(1), patch=patch)
        log(f"{xs.shape[0]} patch pairs from {a.phantoms} phantoms in {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        losses = Lr.train(net, xs, ys, a.steps, a.batch, a.lr, seed=2, log=log,
                          max_seconds=a.max_seconds, warmup=a.warmup)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(f"{len(losses)} steps in {dt:.1f} s ({dt / len(losses) * 1e3:.1f} ms/step)")
        report["training"] = {"pairs": int(xs.shape[0]), "patch": list(patch), "s2d": a.s2d, "act": a.act, "warmup": a.warmup,
                              "steps": len(losses), "batch": a.batch,
                              "lr": a.lr, "seconds": dt, "mse_first": float(np.mean(losses[:10])),
                              "mse_last": float(np.mean(losses[-10:]))}
        del xs, ys
        if a.out:
            os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
            net.save(a.out)
            log(f"saved {a.out}")
    evals = []
    for s in a.eval_seeds:
        evals.append(Lr.compare(net, sc, s, log=log))
    report["evaluation"] = evals
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(report, fh, indent=1)
    print(json.dumps(report), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
