"""Learned sparse-view correction on top of the FDK path (no reference
counterpart).  The reference's sparse-view mechanism is loading every step-th
view of a scan (cuFDK/cuFDK.cpp:62-94, common/type.h:4847-4851); here FDK from
every 4th view is followed by a 3-D residual U-Net trained to map it to the
full-view FDK.  Plain full-view FDK stays both the fallback and the yardstick: every
comparison is against it.

Pipeline pieces:
* ``Scanner``         seeded ellipsoid phantoms, exact projections on the GPU,
                      photon noise (I0 = 1e5), GPU FDK from all or every k-th view
                      into any sub-box.
* ``make_pairs``      training pairs (FDK from 1/4 of the views, FDK from all)
                      cut into patches (128^3; larger for the space-to-depth
                      networks, whose coarsest level needs 32 voxels in x).
* ``train``           Adam on batches of patches, all in the library's kernels.
* ``infer_tiled``     cubic tiles with a 16-voxel overlap over a whole volume;
                      each tile contributes its centre, borders at the edges.
* ``infer_slabs``     whole x-y slices in overlapping z-slabs, or the whole
                      volume in one pass.
"""
from __future__ import annotations

import time

import numpy as np

from . import (MAX_PROJ, PAD_MEAN4, BackProjectionFDK, Parameter, filter_views,
               phantom_project, phantom_rasterize, random_ellipsoids)
from .unet import NativeUNet, UNet3D, box_copy, mse, poisson_noise

I0 = 1.0e5        # photons per detector pixel without attenuation
MU = 0.02         # attenuation per mm for unit phantom density (about water)
PATCH = 128
OVERLAP = 16
SLAB = 528  # z-slab depth of the full-resolution network's inference


def geometry(n: int = 1024, views: int = 1024) -> Parameter:
    """The 1024^3 / 1024^2-detector scan (0.4 mm pixels at n = 1024), scaled with n."""
    du = 0.4 * 1024 / n
    vox = du * 1000.0 / 1500.0
    return Parameter(nx=n, ny=n, nz=n, dx=vox, dy=vox, dz=vox, nu=n, nv=n, du=du, dv=du,
                     sod=1000.0, sid=1500.0, nproj=views, direction=-1, scan_angle_deg=360.0,
                     volume_position_angle_rad=0.0, filter_type=5)


class Scanner:
    """GPU scan simulation and FDK for one geometry."""

    def __init__(self, p: Parameter):
        import torch
        self.p = p
        self.torch = torch
        self._tables = {}

    def tables(self, views: int):
        if views not in self._tables:
            q = self.p.replace(nproj=views)
            t = self.torch
            self._tables[views] = (t.as_tensor(q.cos_table()).cuda(),
                                   t.as_tensor(q.filter().astype(np.float32)).cuda(),
                                   q.proj_matrices().astype(np.float32))
        return self._tables[views]

    def phantom(self, seed: int, n_ellipsoids: int = 20):
        p = self.p
        fov = p.nx * p.dx
        return random_ellipsoids(n_ellipsoids, seed=seed, radius_mm=0.45 * fov,
                                 half_height_mm=0.45 * p.nz * p.dz,
                                 axes_mm=(5.0 * fov / 273.0, 100.0 * fov / 273.0))

    def project(self, ell, noise_seed: int | None = None, out=None):
        """All views as line integrals (views, nv, nu); with photon noise when
        noise_seed is given."""
        p = self.p
        t = self.torch
        raw = out if out is not None else t.empty((p.nproj, p.nv, p.nu), device="cuda")
        for b in range(0, p.nproj, 128):
            phantom_project(ell, p, b, min(128, p.nproj - b), raw[b:b + 128])
        if noise_seed is not None:
            poisson_noise(raw, I0, MU, noise_seed)
        return raw

    def fdk(self, raw, step: int = 1, box=None, store=None, out=None):
        """FDK from every step-th view of raw into box (x, y, z, w, h, d), the
        whole volume by default (as a mirror slab pair).  Returns a device
        tensor [d][h][w]."""
        p = self.p
        t = self.torch
        views = p.nproj // step
        cos, H, mats = self.tables(views)
        if store is None or store.shape[0] < views:
            store = t.empty((views, p.nu, p.nv), device="cuda")
        filter_views(raw, store, cos, H, PAD_MEAN4, True, nviews=views, src_view_step=step)
        if box is None:
            half = p.nz // 2
            top, bot = (0, 0, 0, p.nx, p.ny, half), (0, 0, half, p.nx, p.ny, half)
        else:
            top, bot = tuple(box), (0, 0, 0, 0, 0, 0)
        bp = BackProjectionFDK(p.nv, p.nu, MAX_PROJ, p.nx, p.ny, p.nz, top, bot)
        bp.set_stream(t.cuda.current_stream().cuda_stream)
        for b in range(0, views, MAX_PROJ):
            cnt = min(MAX_PROJ, views - b)
            bp.bp_device(store, p.nv, p.nu, mats[b:b + cnt], cnt, offset_views=b)
        d = top[5] + bot[5]
        vol = out if out is not None else t.empty((d, top[4], top[3]), device="cuda")
        bp.get_volume_device(vol)
        bp.close()
        return vol


def _dhw(v):
    return tuple(int(a) for a in v) if isinstance(v, (tuple, list)) else (int(v),) * 3


def make_pairs(scanner: Scanner, seeds, region=256, sparse_step: int = 4, rng=None,
               patch=PATCH):
    """For each seed: a noisy scan, FDK from every sparse_step-th view and from all
    views inside a random box of `region` voxels (an edge, or (D, H, W)), cut into
    patches of `patch` voxels (likewise).  Returns (inputs, targets), each a
    (P, pd, ph, pw) FP32 device tensor."""
    t = scanner.torch
    p = scanner.p
    rng = np.random.default_rng(0) if rng is None else rng
    rd, rh, rw = _dhw(region)
    pd, ph, pw = _dhw(patch)
    per = (rd // pd) * (rh // ph) * (rw // pw)
    xs = t.empty((len(seeds) * per, pd, ph, pw), device="cuda")
    ys = t.empty_like(xs)
    raw = t.empty((p.nproj, p.nv, p.nu), device="cuda")
    store = t.empty((p.nproj, p.nu, p.nv), device="cuda")
    q = 0
    for seed in seeds:
        ell = scanner.phantom(seed)
        scanner.project(ell, noise_seed=seed + 7919, out=raw)
        org = [int(rng.integers(p.nx // 8, max(p.nx // 8 + 1, p.nx - p.nx // 8 - r)))
               for r in (rw, rh, rd)]
        box = (org[0], org[1], org[2], rw, rh, rd)
        sparse = scanner.fdk(raw, sparse_step, box, store)
        full = scanner.fdk(raw, 1, box, store)
        for a in range(0, rd, pd):
            for b in range(0, rh, ph):
                for c in range(0, rw, pw):
                    box_copy(sparse, (a, b, c), xs[q], (0, 0, 0), (pd, ph, pw))
                    box_copy(full, (a, b, c), ys[q], (0, 0, 0), (pd, ph, pw))
                    q += 1
    t.cuda.synchronize()
    return xs, ys


def lr_at(progress: float, lr: float, lr_final: float = 0.05, warmup: float = 0.0) -> float:
    """Learning rate at `progress` in [0, 1] of a run: a cosine from lr down to
    lr * lr_final, scaled by a linear ramp from lr_final to 1 over the first
    `warmup` of the run."""
    import math
    p = min(max(progress, 0.0), 1.0)
    r = lr * (lr_final + (1.0 - lr_final) * 0.5 * (1.0 + math.cos(math.pi * p)))
    if p < warmup:
        r *= lr_final + (1.0 - lr_final) * p / warmup
    return r


def train(net: UNet3D, xs, ys, steps: int, batch: int = 4, lr: float = 1e-3, seed: int = 0,
          log=None, max_seconds: float | None = None, lr_final: float = 0.05,
          warmup: float = 0.0):
    """Adam on random batches of patch pairs (stops early after max_seconds),
    the learning rate following a cosine from lr to lr * lr_final over the run
    (its progress is the larger of the step and time fractions), after a linear
    ramp from lr * lr_final over the first `warmup` of the run.  Returns the
    per-step mean losses."""
    import torch
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.perf_counter()
    progress_t = 0.0
    for s in range(steps):
        if max_seconds is not None and s % 10 == 0 and s > 0:
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            progress_t = el / max_seconds
            if el > max_seconds:
                if log:
                    log(f"time budget reached after {s} steps")
                break
        progress = min(max(s / max(steps - 1, 1), progress_t), 1.0)
        lr_s = lr_at(progress, lr, lr_final, warmup)
        idx = torch.as_tensor(rng.choice(xs.shape[0], size=batch, replace=False)).cuda()
        x = xs.index_select(0, idx)
        y = ys.index_select(0, idx)
        loss = net.train_step(x, y, lr_s)
        losses.append(loss)
        if log and (s % 25 == 0 or s == steps - 1):
            log(f"step {s}: mse {loss.item() / x.numel():.4e} ({time.perf_counter() - t0:.1f} s)")
    return [l.item() / (batch * xs[0].numel()) for l in losses]


def tiles(n: int, patch: int = PATCH, overlap: int = OVERLAP):
    """Tile origins along one axis and the [lo, hi) each tile writes back:
    neighbouring tiles split their overlap at its midpoint, so the written
    ranges partition [0, n)."""
    stride = patch - overlap
    starts = list(range(0, max(n - patch, 0) + 1, stride))
    if starts[-1] + patch < n:
        starts.append(n - patch)
    cuts = [0] + [(starts[k] + starts[k - 1] + patch) // 2 for k in range(1, len(starts))] + [n]
    return [(s, cuts[k], cuts[k + 1]) for k, s in enumerate(starts)]


def tile_size(net: UNet3D) -> int:
    """The smallest cubic tile the network takes: 128 x its space-to-depth edge."""
    return PATCH * net.s2d


def infer_tiled(net: UNet3D, vol, out=None, batch: int = 4):
    """Apply the network to a (D, H, W) FP32 device volume in overlapping cubic
    tiles of tile_size(net); D, H, W >= that size."""
    import torch
    D, H, W = vol.shape
    T = tile_size(net)
    out = torch.empty_like(vol) if out is None else out
    tz, ty, tx = tiles(D, T), tiles(H, T), tiles(W, T)
    jobs = [(a, b, c) for a in tz for b in ty for c in tx]
    buf = torch.empty((min(batch, len(jobs)), T, T, T), device=vol.device)
    for i in range(0, len(jobs), batch):
        part = jobs[i:i + batch]
        for j, ((sz, _, _), (sy, _, _), (sx, _, _)) in enumerate(part):
            box_copy(vol, (sz, sy, sx), buf[j], (0, 0, 0), (T, T, T))
        res = net.forward(buf[:len(part)])
        for j, ((sz, z0, z1), (sy, y0, y1), (sx, x0, x1)) in enumerate(part):
            box_copy(res[j], (z0 - sz, y0 - sy, x0 - sx), out, (z0, y0, x0),
                     (z1 - z0, y1 - y0, x1 - x0))
    return out


def infer_slabs(net: UNet3D, vol, out=None, slab: int | None = None, overlap: int = OVERLAP):
    """Apply the network to a (D, H, W) FP32 device volume in z-slabs of whole
    x-y slices (slab deep, neighbouring slabs overlapping by ``overlap`` and
    split at the midpoint), batch 1.  slab=None: SLAB (528) for the
    full-resolution network -- two slabs at 1024^3, 1.03x redundant work
    (1.125x with 128-deep slabs, 1.43x with 128^3 tiles), its activations
    peaking near 144 bytes per slab voxel (80 GB at 1024^2 x 528) -- and the
    whole volume in one pass (no redundant work) for the space-to-depth
    networks.  A volume no deeper than the slab runs in one pass; H, W as
    UNet3D requires.  A NativeUNet runs the same schedule in the library
    (fdk_unet_infer)."""
    import torch
    if isinstance(net, NativeUNet):  # the library's C++ runner, same schedule
        return net.infer(vol, out, 0 if slab is None else slab, overlap)
    D = vol.shape[0]
    if slab is None:
        slab = D if net.s2d > 1 else SLAB
    out = torch.empty_like(vol) if out is None else out
    if slab >= D:
        net.forward(vol.unsqueeze(0), out=out.unsqueeze(0))
        return out
    for sz, z0, z1 in tiles(D, slab, overlap):
        res = net.forward(vol[sz:sz + slab].unsqueeze(0))
        out[z0:z1].copy_(res[0, z0 - sz:z1 - sz])
        del res
    return out


def rmse(a, b) -> float:
    import torch
    return float(torch.sqrt(torch.mean((a - b).double() ** 2)).item())


def ssim_slices(a, b) -> float:
    """Mean SSIM over the three central orthogonal slices (8x8 uniform windows,
    data range of b), computed on the host."""
    from scipy.ndimage import uniform_filter
    A, B = a.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)
    rng_ = float(B.max() - B.min())
    c1, c2 = (0.01 * rng_) ** 2, (0.03 * rng_) ** 2
    vals = []
    for sl in (np.s_[A.shape[0] // 2], np.s_[:, A.shape[1] // 2], np.s_[:, :, A.shape[2] // 2]):
        x, y = A[sl], B[sl]
        mx, my = uniform_filter(x, 8), uniform_filter(y, 8)
        vx = uniform_filter(x * x, 8) - mx * mx
        vy = uniform_filter(y * y, 8) - my * my
        cxy = uniform_filter(x * y, 8) - mx * my
        s = ((2 * mx * my + c1) * (2 * cxy + c2)) / ((mx ** 2 + my ** 2 + c1) * (vx + vy + c2))
        vals.append(float(s.mean()))
    return float(np.mean(vals))


def compare_many(nets: dict, scanner: Scanner, seed: int, sparse_step: int = 4, log=None):
    """Whole-volume comparison on one held-out phantom: plain FDK from all views
    vs FDK from every sparse_step-th view, with and without each network of
    `nets` (name -> UNet3D or NativeUNet).  Times are GPU wall times with the noisy projections
    already in HBM; each network is warmed once on the same volume first."""
    import torch
    p = scanner.p
    ell = scanner.phantom(seed)
    raw = scanner.project(ell, noise_seed=seed + 7919)
    store = torch.empty((p.nproj, p.nu, p.nv), device="cuda")
    scanner.tables(p.nproj)
    scanner.tables(p.nproj // sparse_step)
    scanner.fdk(raw, 1, (0, 0, 0, 64, 64, 64), store)  # warm the FDK kernels on a small box
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    full = scanner.fdk(raw, 1, None, store)
    torch.cuda.synchronize()
    t_full = time.perf_counter() - t0

    t0 = time.perf_counter()
    sparse = scanner.fdk(raw, sparse_step, None, store)
    torch.cuda.synchronize()
    t_sparse = time.perf_counter() - t0
    del raw, store

    phantom = torch.empty_like(full)
    phantom_rasterize(ell, p, phantom)
    torch.cuda.synchronize()
    res = {
        "volume": [p.nz, p.ny, p.nx], "views_full": p.nproj, "views_sparse": p.nproj // sparse_step,
        "t_full_fdk_s": t_full, "t_sparse_fdk_s": t_sparse,
        "rmse_vs_full_fdk": {"sparse_fdk": rmse(sparse, full)},
        "rmse_vs_phantom": {"full_fdk": rmse(full, phantom), "sparse_fdk": rmse(sparse, phantom)},
        "ssim_vs_full_fdk": {"sparse_fdk": ssim_slices(sparse, full)},
        "networks": {},
    }
    learned = torch.empty_like(sparse)
    for name, net in nets.items():
        infer_slabs(net, sparse, learned)  # warm: code objects and the allocator's blocks
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infer_slabs(net, sparse, learned)
        torch.cuda.synchronize()
        t_net = time.perf_counter() - t0
        res["networks"][name] = {
            "s2d": net.s2d, "t_network_s": t_net, "t_learned_s": t_sparse + t_net,
            "speedup_vs_full_fdk": t_full / (t_sparse + t_net),
            "rmse_vs_full_fdk": rmse(learned, full), "rmse_vs_phantom": rmse(learned, phantom),
            "ssim_vs_full_fdk": ssim_slices(learned, full),
        }
        torch.cuda.empty_cache()
    if log:
        log(str(res))
    return res


def compare(net: UNet3D, scanner: Scanner, seed: int, sparse_step: int = 4, log=None):
    """compare_many for one network, flattened: t_network_s, t_learned_s,
    speedup_vs_full_fdk and the 'sparse_fdk+net' entries of the RMSE / SSIM
    dictionaries."""
    r = compare_many({"net": net}, scanner, seed, sparse_step)
    n = r.pop("networks")["net"]
    r["t_network_s"], r["t_learned_s"] = n["t_network_s"], n["t_learned_s"]
    r["speedup_vs_full_fdk"] = n["speedup_vs_full_fdk"]
    r["rmse_vs_full_fdk"]["sparse_fdk+net"] = n["rmse_vs_full_fdk"]
    r["rmse_vs_phantom"]["sparse_fdk+net"] = n["rmse_vs_phantom"]
    r["ssim_vs_full_fdk"]["sparse_fdk+net"] = n["ssim_vs_full_fdk"]
    if log:
        log(str(r))
    return r
