"""Iterative reconstruction on the library's projector pair: the ray-marching
forward projector (fdk_fp_project) and the voxel-driven backprojector
(fdk_bp_*), the two loops of the reference's 3dct experiments.

* ``ratio_reconstruct`` -- the multiplicative update of
  Iterative3dReconstructionRatio (3dct/3dct.cpp:1029-1205):
  x <- x * B(p / A x) / B(1), with p shifted to be non-negative, divisions by
  zero giving zero (type.h:3908-3920, 2760-2772), x0 = 1, and the projections
  restricted to a detector ROI 2 pixels in from every edge.
* ``sart_reconstruct`` -- SART_3dReconstruction (3dct.cpp:1207-1292): views in
  a shuffled order, one view s at a time, lambda = 1/3, x0 = 0; the reference's
  step x <- x - lambda / |volume diagonal in mm| * B_s(A_s x - p_s) is
  available, the default normalizes the step by ray lengths and B_s 1.

A is the forward projector in the reference CPU convention (zero outside the
volume, t in [tnear, tfar]); B is the unfiltered backprojection with the 1/w^2
weight.  The ROI is applied as a pixel mask on the projections (the
reference's CPU backprojector instead rejects samples whose truncated
coordinates leave the ROI, which differs only for samples that straddle the
ROI edge).  All arrays are torch tensors on the GPU; the loops run there.
"""
from __future__ import annotations

import numpy as np

from . import FP_ZERO, BackProjectionFDK, Parameter, forward_project


class Projector:
    """Forward projection A and backprojection B of one geometry.

    ``forward(vol, views, out)``: vol (nz, ny, nx) -> out (len(views), nv, nu);
    ``back(proj, views, out)``: proj (len(views), nv, nu) -> out (nz, ny, nx),
    overwritten with the sum over the views."""

    def __init__(self, para: Parameter, step_mm: float = 0.0, fp_mode: int = FP_ZERO):
        import torch
        self.p = para
        self.step_mm = step_mm
        self.fp_mode = fp_mode
        self.mats = para.proj_matrices().astype(np.float32)
        self.bp = BackProjectionFDK(para.nv, para.nu, 32, para.nx, para.ny, para.nz)
        self.bp.set_stream(torch.cuda.current_stream().cuda_stream)

    def forward(self, vol, views, out):
        views = list(views)
        runs = _runs(views)
        k = 0
        for first, count in runs:
            forward_project(self.p, vol, first, count, out[k:k + count], self.step_mm, self.fp_mode)
            k += count
        return out

    def back(self, proj, views, out):
        views = np.asarray(list(views), np.int64)
        self.bp.zero()
        for b in range(0, len(views), 32):
            n = min(32, len(views) - b)
            self.bp.bp_device(proj, self.p.nu, self.p.nv, self.mats[views[b:b + n]], n, transpose=0,
                              offset_views=b)
        self.bp.get_volume_device(out)
        return out

    def close(self):
        self.bp.close()


def _runs(views):
    """Consecutive runs (first, count) of a view list."""
    out = []
    for v in views:
        if out and out[-1][0] + out[-1][1] == v:
            out[-1][1] += 1
        else:
            out.append([v, 1])
    return [tuple(r) for r in out]


def roi_mask(para: Parameter, border: int = 2, device="cuda"):
    import torch
    m = torch.zeros((para.nv, para.nu), dtype=torch.float32, device=device)
    m[border:para.nv - border, border:para.nu - border] = 1.0
    return m


def _safe_div(a, b):
    import torch
    return torch.where(b != 0, a / torch.where(b != 0, b, torch.ones_like(b)), torch.zeros_like(a))


def ratio_reconstruct(para: Parameter, proj, iters: int, roi_border: int = 2, vol=None,
                      callback=None, projector: Projector | None = None):
    """proj: device (nproj, nv, nu) FP32 measured views.  Returns the volume
    (nz, ny, nx) after ``iters`` multiplicative updates; ``callback(it, vol,
    fp)`` sees each iteration's volume and forward projection."""
    import torch
    A = projector or Projector(para)
    views = range(para.nproj)
    p = proj - proj.min() if float(proj.min()) < 0 else proj  # ShiftToNoneNegative
    mask = roi_mask(para, roi_border, proj.device)
    norm = torch.empty((para.nz, para.ny, para.nx), dtype=torch.float32, device=proj.device)
    A.back(mask.expand(para.nproj, -1, -1).contiguous(), views, norm)
    x = torch.ones_like(norm) if vol is None else vol.clone()
    fp = torch.empty_like(p)
    upd = torch.empty_like(norm)
    for it in range(iters):
        A.forward(x, views, fp)
        ratio = _safe_div(p, fp) * mask
        A.back(ratio, views, upd)
        x *= _safe_div(upd, norm)
        if callback:
            callback(it, x, fp)
    if projector is None:
        A.close()
    return x


def sart_reconstruct(para: Parameter, proj, iters: int, lam: float = 1.0 / 3.0, seed: int = 0,
                     vol=None, callback=None, projector: Projector | None = None,
                     normalized: bool = True):
    """proj: device (nproj, nv, nu).  Returns the volume after ``iters`` sweeps
    over all views in a seeded shuffled order, one view per update.

    ``normalized`` (default) divides the residual by the ray lengths A_s 1 and
    the update by B_s 1, the textbook SART step; ``normalized=False`` keeps the
    reference's step lam / |volume diagonal| on the raw backprojection, which
    with the backprojector's 1/w^2 weight (w = distance along the central ray in
    mm) moves the volume by about 1e-9 of the residual per view at
    SOD = 1000 mm."""
    import torch
    A = projector or Projector(para)
    p = proj - proj.min() if float(proj.min()) < 0 else proj
    diag = float(np.linalg.norm([para.dx * para.nx, para.dy * para.ny, para.dz * para.nz]))
    order = np.random.default_rng(seed).permutation(para.nproj)
    shape = (para.nz, para.ny, para.nx)
    x = torch.zeros(shape, dtype=torch.float32, device=proj.device) if vol is None else vol.clone()
    fp = torch.empty((1, para.nv, para.nu), dtype=torch.float32, device=proj.device)
    upd = torch.empty_like(x)
    if normalized:
        ones_v = torch.ones(shape, dtype=torch.float32, device=proj.device)
        ray = torch.empty_like(p)
        A.forward(ones_v, range(para.nproj), ray)  # A_s 1: path length through the volume
        ones_p = torch.ones((1, para.nv, para.nu), dtype=torch.float32, device=proj.device)
        colsum = torch.empty_like(x)
        del ones_v
    for it in range(iters):
        for s in order:
            s = int(s)
            A.forward(x, [s], fp)
            torch.sub(p[s:s + 1], fp, out=fp)
            if normalized:
                fp.copy_(_safe_div(fp, ray[s:s + 1]))
                A.back(fp, [s], upd)
                A.back(ones_p, [s], colsum)
                x += lam * _safe_div(upd, colsum)
            else:
                fp *= lam / diag
                A.back(fp, [s], upd)
                x += upd
        if callback:
            callback(it, x)
    if projector is None:
        A.close()
    return x
