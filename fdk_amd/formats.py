"""On-disk formats of the reference's FDK path (host side, numpy).

* projection folder: one raw file per view, ``<folder>img%04d.raw``, nu x nv
  samples with u fastest (common/type.h:1796-1828 LoadProjectionImageFolder,
  3845-3858 SaveToFolder); ``step`` loads views 0, step, 2*step, ... numbered
  from ``start_index`` (type.h:3831-3844, the reference's sparse-view knob);
* volume file: three little-endian int32 (nx, ny, nz) then FP32 samples
  [z][y][x] (type.h:3256-3310 VolumeT::Save / Load);
* per-slice output: one raw nx*ny FP32 file per z (cuFDK/mpiFDK.cpp:548-575),
  named ``%06d.raw`` by global slice index.
"""
from __future__ import annotations

import os

import numpy as np


def frame_path(folder: str, index: int) -> str:
    return os.path.join(folder, f"img{index:04d}.raw")


def load_projection_folder(folder: str, nu: int, nv: int, count: int, step: int = 1,
                           start_index: int = 0, src_dtype=np.float32, out=None) -> np.ndarray:
    """Views [count][nv][nu] FP32; slot k*step holds file k*step + start_index
    and the other slots stay zero.  Raises FileNotFoundError / ValueError on a
    missing file or a file whose size is not nu*nv samples (checked for every
    file before any is read, as the reference does)."""
    if nu <= 0 or nv <= 0 or count <= 0 or step <= 0:
        raise ValueError("nu, nv, count and step must be positive")
    dt = np.dtype(src_dtype)
    paths = [(i, frame_path(folder, i + start_index)) for i in range(0, count, step)]
    for _, p in paths:
        size = os.path.getsize(p)
        if size != nu * nv * dt.itemsize:
            raise ValueError(f"{p}: {size} bytes, expected {nu * nv * dt.itemsize}")
    if out is None:
        out = np.zeros((count, nv, nu), np.float32)
    elif out.shape != (count, nv, nu):
        raise ValueError("out has the wrong shape")
    for i, p in paths:
        out[i] = np.fromfile(p, dtype=dt, count=nu * nv).reshape(nv, nu)
    return out


def save_projection_folder(folder: str, views: np.ndarray, step: int = 1) -> None:
    """Write views [count][nv][nu] as img%04d.raw (FP32), every step-th."""
    os.makedirs(folder, exist_ok=True)
    v = np.asarray(views, dtype=np.float32)
    if v.ndim != 3:
        raise ValueError("views must be [count][nv][nu]")
    for i in range(0, v.shape[0], step):
        np.ascontiguousarray(v[i]).tofile(frame_path(folder, i))


def save_volume(path: str, vol: np.ndarray) -> None:
    """vol [nz][ny][nx] -> header (nx, ny, nz) int32 + FP32 samples."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    if v.ndim != 3:
        raise ValueError("volume must be [nz][ny][nx]")
    nz, ny, nx = v.shape
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "wb") as fh:
        fh.write(np.array([nx, ny, nz], "<i4").tobytes())
        fh.write(v.astype("<f4", copy=False).tobytes())


def load_volume(path: str) -> np.ndarray:
    with open(path, "rb") as fh:
        hdr = np.frombuffer(fh.read(12), "<i4")
        if hdr.size != 3 or np.any(hdr <= 0):
            raise ValueError(f"{path}: bad volume header")
        nx, ny, nz = (int(h) for h in hdr)
        data = np.frombuffer(fh.read(), "<f4")
    if data.size != nx * ny * nz:
        raise ValueError(f"{path}: {data.size} samples, header says {nx * ny * nz}")
    return data.reshape(nz, ny, nx).copy()


def save_slices(folder: str, vol: np.ndarray, z0: int = 0) -> None:
    """One raw FP32 nx*ny file per slice of vol [nz][ny][nx], %06d.raw from z0."""
    os.makedirs(folder, exist_ok=True)
    v = np.asarray(vol, dtype=np.float32)
    for k in range(v.shape[0]):
        np.ascontiguousarray(v[k]).tofile(os.path.join(folder, f"{z0 + k:06d}.raw"))
