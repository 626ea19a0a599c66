"""TEST INFRASTRUCTURE ONLY -- the oracle for the FDK hot path.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker: nothing here is part of
the product, and the product (``fdk_amd``) never imports it.

``liboracle.so`` is the plain-C FP64 restatement of the reference algorithm
(``fdk_oracle.c``, each function citing the reference file:line it follows).
``_ref/libfdkref.so``, when it was built in the container holding the
reference checkout, is the reference's own CPU code compiled in place
(``ref_harness.cpp``); it pins the oracle and serves as the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, asdict

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PAD_ZERO, PAD_MEAN3, PAD_MEAN4 = 0, 3, 4
BP_GPU_STRICT, BP_CPU_TRUNC = 0, 1
FP_GPU_CLAMP, FP_CPU_ZERO = 0, 1
FILTER_TYPES = ["none", "ram-lak", "shepp-logan", "cosine", "hamming", "hann"]


class RefGeom(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("nu", C.c_int), ("nv", C.c_int),
        ("du", C.c_double), ("dv", C.c_double),
        ("offset_nu", C.c_double), ("offset_nv", C.c_double),
        ("sod", C.c_double), ("sid", C.c_double),
        ("direction", C.c_int), ("scan_angle_deg", C.c_double), ("nproj", C.c_int),
        ("volume_position_angle_rad", C.c_double), ("step", C.c_int), ("filter_type", C.c_int),
    ]


class OrcEllipsoid(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
                ("ax", C.c_double), ("ay", C.c_double), ("az", C.c_double),
                ("rot", C.c_double * 9), ("density", C.c_double)]


@dataclass
class Geom:
    """Scan geometry in the reference's Parameter terms (sod: source-object,
    sid: source-detector, both mm)."""
    nx: int = 16
    ny: int = 16
    nz: int = 16
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0
    nu: int = 24
    nv: int = 20
    du: float = 1.6
    dv: float = 1.6
    offset_nu: float = 0.0
    offset_nv: float = 0.0
    sod: float = 1000.0
    sid: float = 1500.0
    direction: int = -1
    scan_angle_deg: float = 360.0
    nproj: int = 8
    volume_position_angle_rad: float = 0.0
    step: int = 1
    filter_type: int = 5

    def c(self) -> RefGeom:
        return RefGeom(**asdict(self))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _load(path):
    lib = C.CDLL(path)
    return lib


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _lib = _load(path)
        _lib.orc_angular_ramp_factor.restype = C.c_double
        _lib.orc_view_angle.restype = C.c_double
        _lib.orc_view_angle.argtypes = [C.c_void_p, C.c_int]
        _lib.orc_proj_matrix.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
        _lib.orc_filter_rows.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_long, C.c_void_p,
                                         C.c_int, C.c_int]
        _lib.orc_bp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_void_p]
        _lib.orc_bp_voxels.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                       C.c_void_p, C.c_long, C.c_void_p]
        _lib.orc_fp.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_void_p]
        _lib.orc_cos_weight.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libfdkref.so"))


def ref():
    """The reference's own CPU code compiled in place (oracle/_ref)."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libfdkref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libfdkref.so is not built")
        _ref = _load(path)
        _ref.ref_filter_design.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]
        _ref.ref_fp_3dct.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        _ref.ref_bp_3dct.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        _ref.ref_bp_3dct_views.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        _ref.ref_filter_rows_3dct.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                              C.c_int]
        _ref.ref_fft.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    return _ref


# ---------------------------------------------------------------- oracle API
def nextpow2(n: int) -> int:
    return lib().orc_nextpow2(int(n))


def filter_design(ftype: int, length: int, d: int = 1) -> np.ndarray:
    n = nextpow2(length)
    out = np.zeros(n, np.float64)
    r = lib().orc_filter_design(int(ftype), int(length), int(d), _p(out), n)
    assert r == n
    return out


def angular_ramp_factor(g: Geom) -> float:
    gc = g.c()
    return lib().orc_angular_ramp_factor(C.byref(gc))


def cos_table(g: Geom) -> np.ndarray:
    out = np.zeros((g.nv, g.nu), np.float32)
    gc = g.c()
    lib().orc_cos_table(C.byref(gc), _p(out))
    return out


def view_angle(g: Geom, s: int) -> float:
    gc = g.c()
    return lib().orc_view_angle(C.byref(gc), int(s))


def proj_matrix(g: Geom, angle: float) -> np.ndarray:
    out = np.zeros(12, np.float64)
    gc = g.c()
    lib().orc_proj_matrix(C.byref(gc), float(angle), _p(out))
    return out.reshape(3, 4)


def proj_matrices(g: Geom) -> np.ndarray:
    out = np.zeros((g.nproj, 12), np.float64)
    gc = g.c()
    lib().orc_proj_matrices(C.byref(gc), _p(out))
    return out


def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    re = np.ascontiguousarray(np.real(x), np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), np.float64).copy()
    m = int(round(np.log2(len(re))))
    lib().orc_fft(-1 if inverse else 1, m, _p(re), _p(im))
    return re + 1j * im


def cos_weight(src: np.ndarray, table: np.ndarray) -> np.ndarray:
    src = np.ascontiguousarray(src, np.float32)
    table = np.ascontiguousarray(table, np.float32)
    n = table.size
    nviews = src.size // n
    dst = np.empty_like(src)
    lib().orc_cos_weight(_p(src), _p(table), _p(dst), n, nviews)
    return dst


def filter_rows(src: np.ndarray, H: np.ndarray, pad_mode: int = PAD_MEAN4) -> np.ndarray:
    """src: (..., nu) float32 rows.  Returns filtered rows, same shape."""
    src = np.ascontiguousarray(src, np.float32)
    H = np.ascontiguousarray(H, np.float64)
    nu = src.shape[-1]
    dst = np.empty_like(src)
    lib().orc_filter_rows(_p(src), _p(dst), nu, src.size // nu, _p(H), len(H), int(pad_mode))
    return dst


def bp(g: Geom, proj: np.ndarray, mats: np.ndarray, mode: int = BP_GPU_STRICT,
       z0: int = 0, nzs: int | None = None) -> np.ndarray:
    """proj: (nviews, nv, nu) float32, untransposed.  Returns (nzs, ny, nx) float64."""
    proj = np.ascontiguousarray(proj, np.float32)
    mats = np.ascontiguousarray(mats, np.float64).reshape(-1, 12)
    nzs = g.nz - z0 if nzs is None else nzs
    vol = np.zeros((nzs, g.ny, g.nx), np.float64)
    gc = g.c()
    lib().orc_bp(C.byref(gc), _p(proj), _p(mats), mats.shape[0], int(mode), int(z0), int(nzs),
                 _p(vol))
    return vol


def bp_voxels(g: Geom, proj: np.ndarray, mats: np.ndarray, ijk: np.ndarray,
              mode: int = BP_GPU_STRICT) -> np.ndarray:
    proj = np.ascontiguousarray(proj, np.float32)
    mats = np.ascontiguousarray(mats, np.float64).reshape(-1, 12)
    ijk = np.ascontiguousarray(ijk, np.int32)
    out = np.zeros(ijk.shape[0], np.float64)
    gc = g.c()
    lib().orc_bp_voxels(C.byref(gc), _p(proj), _p(mats), mats.shape[0], int(mode), _p(ijk),
                        ijk.shape[0], _p(out))
    return out


def fp(g: Geom, vol: np.ndarray, view: int, step_mm: float | None = None,
       mode: int = FP_GPU_CLAMP) -> np.ndarray:
    vol = np.ascontiguousarray(vol, np.float32)
    step = min(g.dx, g.dy, g.dz) if step_mm is None else step_mm
    out = np.zeros((g.nv, g.nu), np.float32)
    gc = g.c()
    lib().orc_fp(C.byref(gc), _p(vol), int(view), float(step), int(mode), _p(out))
    return out


def _ellipsoids(ell) -> C.Array:
    arr = (OrcEllipsoid * len(ell))()
    for q, e in enumerate(ell):
        arr[q].cx, arr[q].cy, arr[q].cz = e["center"]
        arr[q].ax, arr[q].ay, arr[q].az = e["axes"]
        for t, val in enumerate(np.asarray(e["rot"], np.float64).reshape(9)):
            arr[q].rot[t] = val
        arr[q].density = e["density"]
    return arr


def phantom_rasterize(ell, g: Geom) -> np.ndarray:
    arr = _ellipsoids(ell)
    vol = np.zeros((g.nz, g.ny, g.nx), np.float32)
    gc = g.c()
    lib().orc_phantom_rasterize(arr, len(ell), C.byref(gc), _p(vol))
    return vol


def phantom_project(ell, g: Geom, view: int) -> np.ndarray:
    arr = _ellipsoids(ell)
    out = np.zeros((g.nv, g.nu), np.float32)
    gc = g.c()
    lib().orc_phantom_project(arr, len(ell), C.byref(gc), int(view), _p(out))
    return out


# ---------------------------------------------------- compiled reference API
def ref_tables(g: Geom):
    gc = g.c()
    n = nextpow2(g.nu)
    cos = np.zeros((g.nv, g.nu), np.float32)
    filt = np.zeros(n, np.float64)
    mats = np.zeros((g.nproj, 12), np.float32)
    af = C.c_double()
    r = ref().ref_tables(C.byref(gc), _p(cos), _p(filt), n, _p(mats), C.byref(af))
    assert r == n, r
    return {"cos": cos, "filter": filt, "mats": mats, "af": af.value}


def ref_filter_design(ftype: int, length: int, d: int = 1) -> np.ndarray:
    n = nextpow2(length)
    out = np.zeros(n, np.float64)
    r = ref().ref_filter_design(int(ftype), int(length), int(d), _p(out), n)
    assert r == n
    return out


def ref_filter_rows_3dct(src: np.ndarray, H: np.ndarray) -> np.ndarray:
    src = np.ascontiguousarray(src, np.float32)
    H = np.ascontiguousarray(H, np.float64)
    w = src.shape[-1]
    h = src.size // w
    dst = np.empty_like(src)
    ref().ref_filter_rows_3dct(_p(src), _p(dst), w, h, _p(H), len(H))
    return dst


def ref_fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    re = np.ascontiguousarray(np.real(x), np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), np.float64).copy()
    m = int(round(np.log2(len(re))))
    ref().ref_fft(-1 if inverse else 1, m, _p(re), _p(im))
    return re + 1j * im


def ref_bp_3dct(g: Geom, frames: np.ndarray, prefiltered: bool) -> np.ndarray:
    frames = np.ascontiguousarray(frames, np.float32)
    vol = np.zeros((g.nz, g.ny, g.nx), np.float32)
    gc = g.c()
    ref().ref_bp_3dct(C.byref(gc), _p(frames), 1 if prefiltered else 0, _p(vol))
    return vol


def ref_bp_3dct_views(g: Geom, frames: np.ndarray, nviews: int) -> np.ndarray:
    """Reference 3dct BackProjection of prefiltered views [0, nviews) of g."""
    frames = np.ascontiguousarray(frames, np.float32)
    vol = np.zeros((g.nz, g.ny, g.nx), np.float32)
    gc = g.c()
    ref().ref_bp_3dct_views(C.byref(gc), _p(frames), 1, int(nviews), _p(vol))
    return vol


def ref_fp_3dct(g: Geom, vol: np.ndarray, view: int) -> np.ndarray:
    vol = np.ascontiguousarray(vol, np.float32)
    out = np.zeros((g.nv, g.nu), np.float32)
    gc = g.c()
    ref().ref_fp_3dct(C.byref(gc), _p(vol), int(view), _p(out))
    return out


# ------------------------------------------------------------------ photon noise
# No reference counterpart (the reference has no noise model; BASELINE's learned
# configuration asks for Poisson noise at I0 = 1e5).  A numpy restatement of the
# library's sampler: the same counter-based hash uniforms, Knuth's multiplication
# method below 10 expected counts, Hoermann's PTRS transformed rejection (1993)
# from 10 up.  Pinned by the Poisson law itself (tests compare its counts with
# scipy.stats.poisson: mean, variance, chi-square of the histogram).

def _hash32(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint32)
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def _noise_unif(seed: int, idx: np.ndarray, k) -> np.ndarray:
    i64 = idx.astype(np.uint64)
    hi = ((i64 * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(32)).astype(np.uint32)
    lo = (i64 & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    kk = np.asarray(k, dtype=np.uint64) * np.uint64(0x632BE5AB)
    # hash32(hash32(seed ^ hi) ^ hash32(lo + 0x632BE5AB k)), all in 32-bit arithmetic
    a = _hash32(np.uint32(seed) ^ hi)
    c = _hash32((lo.astype(np.uint64) + kk).astype(np.uint32))
    h = _hash32(a ^ c)
    return (h.astype(np.float64) + 0.5) * (1.0 / 4294967296.0)


def poisson_counts(lam: np.ndarray, seed: int, idx: np.ndarray | None = None) -> np.ndarray:
    """Poisson variates of mean lam (float64), element i drawn from the hash
    stream (seed, idx[i]); idx defaults to 0..n-1."""
    from scipy.special import gammaln
    lam = np.asarray(lam, np.float64).ravel()
    idx = np.arange(lam.size, dtype=np.uint64) if idx is None else np.asarray(idx, np.uint64)
    out = np.zeros(lam.size)
    small = np.nonzero(lam < 10.0)[0]
    if small.size:
        L = np.exp(-lam[small])
        prod = _noise_unif(seed, idx[small], 0)
        c = np.zeros(small.size)
        for t in range(1, 1001):
            go = prod > L
            if not go.any():
                break
            c[go] += 1
            prod[go] *= _noise_unif(seed, idx[small][go], t)
        out[small] = c
    big = np.nonzero(lam >= 10.0)[0]
    if big.size:
        lb = lam[big]
        slam, loglam = np.sqrt(lb), np.log(lb)
        b = 0.931 + 2.53 * slam
        a = -0.059 + 0.02483 * b
        lia = np.log(1.1239 + 1.1328 / (b - 3.4))
        vr = 0.9277 - 3.6224 / (b - 2.0)
        k = np.floor(lb)
        done = np.zeros(big.size, bool)
        for t in range(128):
            todo = np.nonzero(~done)[0]
            if todo.size == 0:
                break
            U = _noise_unif(seed, idx[big][todo], 2 * t + 1) - 0.5
            V = _noise_unif(seed, idx[big][todo], 2 * t + 2)
            us = 0.5 - np.abs(U)
            kt = np.floor((2.0 * a[todo] / us + b[todo]) * U + lb[todo] + 0.43)
            k[todo] = kt
            acc = (us >= 0.07) & (V <= vr[todo])
            rest = ~acc & ~((kt < 0) | ((us < 0.013) & (V > us)))
            kk = np.where(rest, kt, 0.0)
            with np.errstate(divide="ignore"):
                acc2 = rest & (np.log(V) + lia[todo] - np.log(a[todo] / (us * us) + b[todo])
                               <= -lb[todo] + kk * loglam[todo] - gammaln(kk + 1.0))
            done[todo[acc | acc2]] = True
        out[big] = np.maximum(k, 0.0)
    return out


def poisson_noise(p: np.ndarray, i0: float, mu: float, seed: int) -> np.ndarray:
    """Line integrals p (float32) -> noisy line integrals, as the library's
    photon-noise step: counts ~ Poisson(i0 exp(-mu p)), -ln(max(counts, 1)/i0)/mu."""
    p = np.asarray(p, np.float32)
    lam = i0 * np.exp(-mu * p.astype(np.float64).ravel())
    k = poisson_counts(lam, seed)
    return (-np.log(np.maximum(k, 1.0) / i0) / mu).astype(np.float32).reshape(p.shape)
