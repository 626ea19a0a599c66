"""Reconstruction quality figures as the reference reports them.

* ``check_image_quality`` -- CheckImageQuality (common/type.h:4991-5030): mean
  absolute error per voxel, MSE, PSNR against a peak of 2.0
  (20 log10 2 - 10 log10 MSE), QI = (2 - error per voxel) / 2, and the
  reference's pass thresholds (PSNR >= 26 dB, error per voxel <= 0.03), which
  it only prints; here they are returned.
* ``ncc`` -- NCC (type.h:4960-4987): normalised cross-correlation of two
  arrays about their means.

Accumulation is FP64 (the reference accumulates the errors in double and NCC in
float).
"""
from __future__ import annotations

import math

import numpy as np


def check_image_quality(ref, recon, snr_threshold: float = 26.0,
                        error_per_pixel_threshold: float = 0.03) -> dict:
    r = np.asarray(ref, np.float64).ravel()
    x = np.asarray(recon, np.float64).ravel()
    if r.shape != x.shape or r.size == 0:
        raise ValueError("volumes must have the same, non-zero size")
    d = x - r
    err = float(np.abs(d).sum() / r.size)
    mse = float((d * d).sum() / r.size)
    psnr = math.inf if mse == 0 else 20.0 * math.log10(2.0) - 10.0 * math.log10(mse)
    qi = (2.0 - err) / 2.0
    return {"error_per_pixel": err, "mse": mse, "psnr_db": psnr, "qi": qi,
            "passed": err <= error_per_pixel_threshold and psnr >= snr_threshold}


def ncc(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    if a.shape != b.shape or a.size == 0:
        raise ValueError("arrays must have the same, non-zero size")
    s0, s1 = a - a.mean(), b - b.mean()
    return float((s0 * s1).sum() / (math.sqrt((s0 * s0).sum()) * math.sqrt((s1 * s1).sum())))
