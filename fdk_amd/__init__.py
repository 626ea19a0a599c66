"""fdk_amd -- MI355X-native FDK cone-beam reconstruction.

Python view of the C ABI in ``include/fdk_amd.h`` (``libfdk_amd.so``, built by
``fdk_amd/build.py``).  Names follow the reference's operator interface:

* ``Parameter``            common/type.h:4627-4895 (geometry and tables)
* ``BackProjectionFDK``    cudaBP/backprojection.cuh:246-271
* ``BackProjectionRTK``    cudaBP/backprojection.cuh:180-198
* ``CorrectImage_Real32``, ``FilterRow_Real32``  FilterlibCufft/cudaFftFilter.cuh:7-9
* ``FilterProjImage``      FilterlibCufft/cudaFftFilter.cu:394
* ``FilterEngine``         Filterlib/FilterLib.h:148-203
* ``forward_project``      genProj/cudaForwardProjection.cu:338-487

Every compute call goes to the native library; there is no CPU fallback, and
loading fails loudly when the library is missing.  Device buffers are passed as
torch tensors (PyTorch provides HBM allocations and streams only).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, asdict, replace

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfdk_amd.so")

FDK_F32, FDK_F16, FDK_BF16 = 0, 1, 2
PAD_ZERO, PAD_MEAN3, PAD_MEAN4 = 0, 3, 4
FP_CLAMP, FP_ZERO = 0, 1
MAX_PROJ = 32
FDK_ERR_INVALID, FDK_ERR_OOM, FDK_ERR_HIP, FDK_ERR_UNSUPPORTED = -1, -2, -3, -4
FILTER_TYPES = ["none", "ram-lak", "shepp-logan", "cosine", "hamming", "hann"]


class FdkError(RuntimeError):
    pass


class _Geometry(C.Structure):
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


class _Ellipsoid(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
                ("ax", C.c_double), ("ay", C.c_double), ("az", C.c_double),
                ("rot", C.c_double * 9), ("density", C.c_double)]


_lib = None

_SIGS = {
    "fdk_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "fdk_version": (C.c_int, []),
    "fdk_device_count": (C.c_int, []),
    "fdk_geom_angular_ramp_factor": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fdk_geom_cos_table": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fdk_geom_filter_length": (C.c_int, [C.c_void_p]),
    "fdk_geom_filter": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "fdk_geom_view_angle": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fdk_geom_proj_matrices": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "fdk_bp_create": (C.c_int, [C.c_int] * 6 + [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fdk_bp_destroy": (C.c_int, [C.c_void_p]),
    "fdk_bp_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fdk_bp_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                               C.c_int]),
    "fdk_bp_batch_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                      C.c_int64, C.c_void_p, C.c_void_p, C.c_int]),
    "fdk_bp_batch_device_minmax": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                             C.c_int64, C.c_void_p, C.c_void_p, C.c_int]),
    "fdk_bp_sync": (C.c_int, [C.c_void_p]),
    "fdk_bp_count": (C.c_int, [C.c_void_p]),
    "fdk_bp_kernel_ms": (C.c_float, [C.c_void_p]),
    "fdk_bp_zero": (C.c_int, [C.c_void_p]),
    "fdk_bp_get_volume": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "fdk_bp_get_volume_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "fdk_bp_volume_ptr": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fdk_rtk_create": (C.c_int, [C.c_int] * 6 + [C.c_void_p]),
    "fdk_rtk_destroy": (C.c_int, [C.c_void_p]),
    "fdk_rtk_bp": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                             C.c_int]),
    "fdk_rtk_get_volume": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "fdk_rtk_kernel_ms": (C.c_float, [C.c_void_p]),
    "fdk_reconstruct_conebeam": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "fdk_correct_image": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int] * 5
                          + [C.c_void_p]),
    "fdk_filter_rows": (C.c_int, [C.c_void_p, C.c_void_p] + [C.c_int] * 5
                        + [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "fdk_filter_views": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64] + [C.c_int] * 4
                         + [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                            C.c_void_p]),
    "fdk_filter_views_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
                                 + [C.c_int] * 4 + [C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                                    C.c_int, C.c_void_p, C.c_void_p]),
    "fdk_filter_views_async": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
                               + [C.c_int] * 4 + [C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                                  C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fdk_filter_proj_image": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "fdk_dev_malloc": (C.c_int, [C.c_void_p, C.c_size_t]),
    "fdk_dev_free": (C.c_int, [C.c_void_p]),
    "fdk_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "fdk_memcpy_2d": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                                C.c_size_t]),
    "fdk_dev_memset": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    "fdk_filter_frame_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "fdk_fp_project_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double,
                                      C.c_int, C.c_void_p]),
    "fdk_fp_project": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_void_p, C.c_void_p]),
    "fdk_phantom_rasterize": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "fdk_phantom_project": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                      C.c_void_p, C.c_void_p]),
    "fdk_pipe_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                  C.c_void_p]),
    "fdk_pipe_push_raw": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "fdk_pipe_push_filtered": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                         C.c_void_p]),
    "fdk_pipe_finish": (C.c_int, [C.c_void_p]),
    "fdk_pipe_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fdk_pipe_bp": (C.c_void_p, [C.c_void_p]),
    "fdk_pipe_destroy": (C.c_int, [C.c_void_p]),
}


def lib():
    """The native library; raises FdkError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FdkError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                           "(python -m fdk_amd.build)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().fdk_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise FdkError(f"{what}: {last_error()} (status {rc})")
    return rc


def _np_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dev_ptr(t):
    return C.c_void_p(t.data_ptr())


_F32_NAMES = ("torch.float32",)
_U16_NAMES = ("torch.int16", "torch.uint16", "torch.float16", "torch.bfloat16")


def _dev_buffer(t, what: str, nbytes: int, dtype: int | None = FDK_F32) -> int:
    """Checks that ``t`` is a contiguous CUDA tensor holding at least nbytes
    (and, when dtype is given, of an element type matching that stored dtype:
    float32 for FDK_F32, a 16-bit type for FDK_F16 / FDK_BF16); returns its
    device address.  A launch over a too-small tensor would read or write
    outside its allocation, so every device wrapper checks before launching."""
    if not hasattr(t, "data_ptr") or not hasattr(t, "is_cuda"):
        raise FdkError(f"{what}: expected a torch tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise FdkError(f"{what}: tensor is not on the GPU")
    if not t.is_contiguous():
        raise FdkError(f"{what}: tensor is not contiguous")
    if dtype is not None:
        names = _F32_NAMES if dtype == FDK_F32 else _U16_NAMES
        if str(t.dtype) not in names:
            raise FdkError(f"{what}: element type {t.dtype} does not match the stored dtype "
                           f"({'float32' if dtype == FDK_F32 else '16-bit'})")
    have = t.numel() * t.element_size()
    if have < nbytes:
        raise FdkError(f"{what}: tensor holds {have} bytes, the call needs {nbytes}")
    return t.data_ptr()


def _host_array(a, what: str, nbytes: int, elem: int) -> C.c_void_p:
    """Host buffer (numpy array or CPU tensor), contiguous, of element size
    ``elem`` and at least nbytes; returns its address."""
    if hasattr(a, "data_ptr"):
        if a.is_cuda:
            raise FdkError(f"{what}: expected host memory, got a GPU tensor")
        if not a.is_contiguous():
            raise FdkError(f"{what}: host array is not contiguous")
        have, esz, ptr = a.numel() * a.element_size(), a.element_size(), a.data_ptr()
    elif isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise FdkError(f"{what}: host array is not contiguous")
        have, esz, ptr = a.nbytes, a.itemsize, a.ctypes.data
    else:
        raise FdkError(f"{what}: expected a numpy array or a CPU tensor")
    if esz != elem:
        raise FdkError(f"{what}: element size {esz} bytes, expected {elem}")
    if have < nbytes:
        raise FdkError(f"{what}: host array holds {have} bytes, the call needs {nbytes}")
    return C.c_void_p(ptr)


def _mats(mats, n: int, what: str) -> np.ndarray:
    m = np.ascontiguousarray(mats, np.float32).reshape(-1)
    if m.size < 12 * n:
        raise FdkError(f"{what}: {m.size // 12} matrices for {n} views")
    return m


def _scale_offset(so, n: int, what: str):
    if so is None:
        return None
    a = np.ascontiguousarray(so, np.float32).reshape(-1)
    if a.size < 2 * n:
        raise FdkError(f"{what}: {a.size // 2} scale/offset pairs for {n} views")
    return a


def _stream(stream=None):
    if stream is not None:
        return C.c_void_p(stream)
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


# ------------------------------------------------------------------ geometry
@dataclass
class Parameter:
    """Scan geometry (common/type.h:4627-4895).  Reference naming: ``sod`` is the
    source-to-object distance (reference SOD), ``sid`` the source-to-detector
    distance (reference SID), both in mm."""
    nx: int = 256
    ny: int = 256
    nz: int = 256
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0
    nu: int = 256
    nv: int = 256
    du: float = 1.6
    dv: float = 1.6
    offset_nu: float = 0.0
    offset_nv: float = 0.0
    sod: float = 1000.0
    sid: float = 1500.0
    direction: int = -1
    scan_angle_deg: float = 360.0
    nproj: int = 360
    volume_position_angle_rad: float = 0.0
    step: int = 1
    filter_type: int = 5

    def c(self) -> _Geometry:
        return _Geometry(**asdict(self))

    def replace(self, **kw) -> "Parameter":
        return replace(self, **kw)

    @property
    def str_filter(self) -> str:
        return FILTER_TYPES[self.filter_type]

    def angular_ramp_factor(self) -> float:
        out = C.c_double()
        g = self.c()
        _check(lib().fdk_geom_angular_ramp_factor(C.byref(g), C.byref(out)), "AngularAndrampFactor")
        return out.value

    def cos_table(self) -> np.ndarray:
        out = np.empty((self.nv, self.nu), np.float32)
        g = self.c()
        _check(lib().fdk_geom_cos_table(C.byref(g), _np_ptr(out)), "CalculateCosTable")
        return out

    def filter_length(self) -> int:
        g = self.c()
        return _check(lib().fdk_geom_filter_length(C.byref(g)), "filter length")

    def filter(self) -> np.ndarray:
        n = self.filter_length()
        out = np.empty(n, np.float64)
        g = self.c()
        _check(lib().fdk_geom_filter(C.byref(g), _np_ptr(out), n), "filter")
        return out

    def view_angle(self, view: int) -> float:
        out = C.c_double()
        g = self.c()
        _check(lib().fdk_geom_view_angle(C.byref(g), int(view), C.byref(out)), "view angle")
        return out.value

    def proj_matrices(self, first: int = 0, count: int | None = None) -> np.ndarray:
        """(count, 12) FP64 projection matrices for views first..first+count."""
        count = self.nproj - first if count is None else count
        out = np.empty((count, 12), np.float64)
        g = self.c()
        _check(lib().fdk_geom_proj_matrices(C.byref(g), int(first), int(count), _np_ptr(out)),
               "CalculateProjMat")
        return out


# ------------------------------------------------------------ backprojection
class BackProjectionFDK:
    """Drop-in for BackProjectionFDK (cudaBP/backprojectionFDK.cu:161-302).

    ``sub_top`` / ``sub_bottom`` are (x, y, z, width, height, depth) boxes; a
    bottom box of depth 0 means no mirror slab."""

    def __init__(self, projWidth, projHeight, projCount, nx, ny, nz, sub_top=None,
                 sub_bottom=None):
        sub_top = (0, 0, 0, nx, ny, nz) if sub_top is None else sub_top
        sub_bottom = (0, 0, 0, 0, 0, 0) if sub_bottom is None else sub_bottom
        self.sub_top = tuple(int(v) for v in sub_top)
        self.sub_bottom = tuple(int(v) for v in sub_bottom)
        t = (C.c_int64 * 6)(*self.sub_top)
        b = (C.c_int64 * 6)(*self.sub_bottom)
        h = C.c_void_p()
        _check(lib().fdk_bp_create(int(projWidth), int(projHeight), int(projCount), int(nx),
                                   int(ny), int(nz), t, b, C.byref(h)), "BackProjectionFDK")
        self._h = h
        self.nx, self.ny, self.nz = nx, ny, nz

    @classmethod
    def _borrowed(cls, handle, nx, ny, nz, sub_top, sub_bottom, owner):
        """A view of a handle that another object owns (the streaming pipeline)."""
        self = cls.__new__(cls)
        self._h = C.c_void_p(handle)
        self._owner = owner
        self.sub_top, self.sub_bottom = tuple(sub_top), tuple(sub_bottom)
        self.nx, self.ny, self.nz = nx, ny, nz
        return self

    def close(self):
        if getattr(self, "_owner", None) is not None:
            self._h = None
            return
        if getattr(self, "_h", None):
            lib().fdk_bp_destroy(self._h)
            self._h = None

    __del__ = close

    @property
    def depth(self) -> int:
        return self.sub_top[5] + self.sub_bottom[5]

    def set_stream(self, stream):
        _check(lib().fdk_bp_set_stream(self._h, C.c_void_p(stream)), "set_stream")

    def BP(self, transpose: int, proj: np.ndarray, width: int, height: int, mats: np.ndarray,
           nProj: int) -> bool:
        proj = np.ascontiguousarray(proj, np.float32)
        mats = _mats(mats, nProj, "BackProjectionFDK::BP")
        if proj.size < width * height * nProj:
            raise FdkError(f"BackProjectionFDK::BP: {proj.size} samples for {nProj} views of "
                           f"{width} x {height}")
        _check(lib().fdk_bp_batch(self._h, int(transpose), _np_ptr(proj), int(width), int(height),
                                  _np_ptr(mats), int(nProj)), "BackProjectionFDK::BP")
        return True

    def bp_device(self, proj, width: int, height: int, mats: np.ndarray, nProj: int,
                  transpose: int = 1, dtype: int = FDK_F32, view_stride: int | None = None,
                  scale_offset: np.ndarray | None = None, offset_views: int = 0):
        """Backproject nProj device-resident views (a torch tensor) asynchronously."""
        if dtype not in (FDK_F32, FDK_F16, FDK_BF16):
            raise FdkError(f"bp_device: unknown dtype {dtype}")
        mats = _mats(mats, nProj, "bp_device")
        vs = width * height if view_stride is None else int(view_stride)
        if vs < width * height or offset_views < 0:
            raise FdkError("bp_device: view stride below a view, or negative view offset")
        esize = 4 if dtype == FDK_F32 else 2
        need = ((offset_views + nProj - 1) * vs + width * height) * esize if nProj > 0 else 0
        ptr = C.c_void_p(_dev_buffer(proj, "bp_device proj", need, dtype) +
                         offset_views * vs * esize)
        so = _scale_offset(scale_offset, nProj, "bp_device")
        if dtype != FDK_F32 and so is None and nProj > 0:
            raise FdkError("bp_device: 16-bit views need their scale_offset")
        _check(lib().fdk_bp_batch_device(self._h, int(transpose), ptr, int(dtype), int(width),
                                         int(height), vs, _np_ptr(mats),
                                         _np_ptr(so) if so is not None else None, int(nProj)),
               "bp_device")

    def sync(self):
        _check(lib().fdk_bp_sync(self._h), "sync")

    def zero(self):
        _check(lib().fdk_bp_zero(self._h), "zero")

    @property
    def nBpCount(self) -> int:
        return lib().fdk_bp_count(self._h)

    @property
    def fKernelTime(self) -> float:
        return lib().fdk_bp_kernel_ms(self._h)

    def GetVolumeData(self) -> np.ndarray:
        """[z over top then bottom][y][x] float32."""
        out = np.empty((self.depth, self.sub_top[4], self.sub_top[3]), np.float32)
        _check(lib().fdk_bp_get_volume(self._h, _np_ptr(out), out.size), "GetVolumeData")
        return out

    def volume_ptr(self):
        """(device pointer, (nxs, nys, nzz)) of the working [x][y][zz] volume."""
        p = C.c_void_p()
        dims = (C.c_int64 * 3)()
        _check(lib().fdk_bp_volume_ptr(self._h, C.byref(p), dims), "volume_ptr")
        return p.value, tuple(dims)

    def volume_tensor(self):
        """The working volume as a torch view (no copy), shape (nxs, nys, nzz)."""
        import torch
        ptr, dims = self.volume_ptr()
        n = dims[0] * dims[1] * dims[2]
        t = _wrap_device(ptr, n)
        return t.view(*dims)

    def get_volume_device(self, out):
        _dev_buffer(out, "get_volume_device", 0)
        _check(lib().fdk_bp_get_volume_device(self._h, _dev_ptr(out), out.numel()),
               "get_volume_device")
        return out


class Pipeline:
    """Streaming reconstruction: host views in batches of up to 32 are copied
    on one stream while the previous batch is filtered and backprojected on
    another (fdk_pipe_*; the single-node form of cuFDK/mpiFDK.cpp:8-317).

    ``push_raw`` takes raw views [count][nv][nu] FP32 (cos weighting and the
    ramp filter run on the device); ``push_filtered`` takes filtered,
    transposed views [count][nu][nv] in the store dtype (uint16/int16 bit
    patterns for FP16/BF16, with per-view scale/offset).  Host arrays should be
    page-locked (torch ``pin_memory=True``) for the copy to overlap."""

    def __init__(self, para: Parameter, sub_top=None, sub_bottom=None, store_dtype: int = FDK_F32,
                 pad_mode: int = PAD_MEAN4, slots: int = 2):
        sub_top = (0, 0, 0, para.nx, para.ny, para.nz) if sub_top is None else sub_top
        sub_bottom = (0, 0, 0, 0, 0, 0) if sub_bottom is None else sub_bottom
        self._g = para.c()
        t = (C.c_int64 * 6)(*[int(v) for v in sub_top])
        b = (C.c_int64 * 6)(*[int(v) for v in sub_bottom])
        h = C.c_void_p()
        _check(lib().fdk_pipe_create(C.byref(self._g), t, b, int(store_dtype), int(pad_mode),
                                     int(slots), C.byref(h)), "Pipeline")
        self._h = h
        self.store_dtype = store_dtype
        self._view = para.nu * para.nv
        self.bp = BackProjectionFDK._borrowed(lib().fdk_pipe_bp(h), para.nx, para.ny, para.nz,
                                              sub_top, sub_bottom, self)

    def push_raw(self, views, count: int, first_view: int, view_step: int = 1):
        ptr = _host_array(views, "push_raw", int(count) * self._view * 4, 4)
        _check(lib().fdk_pipe_push_raw(self._h, ptr, int(count), int(first_view), int(view_step)),
               "push_raw")

    def push_filtered(self, views, count: int, first_view: int, view_step: int = 1,
                      scale_offset=None):
        esize = 4 if self.store_dtype == FDK_F32 else 2
        ptr = _host_array(views, "push_filtered", int(count) * self._view * esize, esize)
        so = _scale_offset(scale_offset, int(count), "push_filtered")
        _check(lib().fdk_pipe_push_filtered(self._h, ptr, int(count), int(first_view),
                                            int(view_step), _np_ptr(so) if so is not None else None),
               "push_filtered")

    def finish(self):
        _check(lib().fdk_pipe_finish(self._h), "finish")

    def stats(self) -> dict:
        out = (C.c_double * 4)()
        _check(lib().fdk_pipe_stats(self._h, out), "stats")
        return {"views": int(out[0]), "h2d_bytes": out[1], "filter_ms": out[2], "bp_ms": out[3]}

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "bp", None) is not None:
                self.bp._h = None
            lib().fdk_pipe_destroy(self._h)
            self._h = None

    __del__ = close


class BackProjectionRTK:
    """Drop-in for BackProjectionRTK (rtkCudaFDKBackProjectionImageFilter.cu:340-400)."""

    def __init__(self, projWidth, projHeight, projCount, nx, ny, nz):
        h = C.c_void_p()
        _check(lib().fdk_rtk_create(int(projWidth), int(projHeight), int(projCount), int(nx),
                                    int(ny), int(nz), C.byref(h)), "BackProjectionRTK")
        self._h = h
        self.nx, self.ny, self.nz = nx, ny, nz

    def close(self):
        if getattr(self, "_h", None):
            lib().fdk_rtk_destroy(self._h)
            self._h = None

    __del__ = close

    def BP(self, transpose, proj, width, height, mats, nProj) -> bool:
        proj = np.ascontiguousarray(proj, np.float32)
        mats = np.ascontiguousarray(mats, np.float32)
        _check(lib().fdk_rtk_bp(self._h, int(transpose), _np_ptr(proj), int(width), int(height),
                                _np_ptr(mats), int(nProj)), "BackProjectionRTK::BP")
        return True

    def GetVolumeData(self) -> np.ndarray:
        out = np.empty((self.nz, self.ny, self.nx), np.float32)
        _check(lib().fdk_rtk_get_volume(self._h, _np_ptr(out), out.size), "GetVolumeData")
        return out

    @property
    def fKernelTime(self) -> float:
        return lib().fdk_rtk_kernel_ms(self._h)


def _wrap_device(ptr: int, n: int):
    """A float32 torch tensor over n floats of existing device memory."""
    import torch

    class _Arr:
        __cuda_array_interface__ = {"shape": (n,), "typestr": "<f4", "data": (ptr, False),
                                    "version": 3, "strides": None}
    return torch.as_tensor(_Arr(), device="cuda")


def reconstruct_conebeam(proj_size, vol_size, mats, vol_in, vol_out, proj, stream=None):
    """CUDA_reconstruct_conebeam: vol_out = vol_in + BP(proj); volumes are device
    tensors, proj a device tensor or a host numpy array."""
    ps = (C.c_int * 3)(*proj_size)
    vs = (C.c_int * 3)(*vol_size)
    mats = _mats(mats, int(proj_size[2]), "reconstruct_conebeam")
    nvox = int(vol_size[0]) * int(vol_size[1]) * int(vol_size[2]) * 4
    nproj = int(proj_size[0]) * int(proj_size[1]) * int(proj_size[2]) * 4
    _dev_buffer(vol_in, "reconstruct_conebeam vol_in", nvox)
    _dev_buffer(vol_out, "reconstruct_conebeam vol_out", nvox)
    on_dev = not isinstance(proj, np.ndarray)
    if not on_dev:
        proj = np.ascontiguousarray(proj, np.float32)
        if proj.nbytes < nproj:
            raise FdkError("reconstruct_conebeam: host projections smaller than proj_size")
    else:
        _dev_buffer(proj, "reconstruct_conebeam proj", nproj)
    pp = _dev_ptr(proj) if on_dev else _np_ptr(proj)
    _check(lib().fdk_reconstruct_conebeam(ps, vs, _np_ptr(mats), _dev_ptr(vol_in),
                                          _dev_ptr(vol_out), pp, 1 if on_dev else 0,
                                          _stream(stream)), "reconstruct_conebeam")


# ------------------------------------------------------------------ filtering
def CorrectImage_Real32(src, table, dst, stream=None):
    """dst = src * table over a (depth, height, width) device tensor."""
    depth, height, width = src.shape
    _dev_buffer(src, "CorrectImage src", depth * height * width * 4)
    _dev_buffer(dst, "CorrectImage dst", depth * height * width * 4)
    _dev_buffer(table, "CorrectImage table", height * table.shape[-1] * 4)
    _check(lib().fdk_correct_image(_dev_ptr(src), _dev_ptr(table), _dev_ptr(dst), width, height,
                                   depth, width, table.shape[-1], _stream(stream)), "CorrectImage")


def FilterRow_Real32(src, dst, filt, pad_mode: int = PAD_MEAN3, stream=None):
    """Row filter of a (depth, height, width) device tensor with a length-N
    device filter (the cuFFT path pads with the 3-sample mean)."""
    depth, height, width = src.shape
    _dev_buffer(src, "FilterRow src", depth * height * width * 4)
    _dev_buffer(dst, "FilterRow dst", depth * height * width * 4)
    _dev_buffer(filt, "FilterRow filter", 0)
    _check(lib().fdk_filter_rows(_dev_ptr(src), _dev_ptr(dst), width, height, depth, width, width,
                                 _dev_ptr(filt), int(filt.numel()), int(pad_mode),
                                 _stream(stream)), "FilterRow")


def filter_views(src, dst, cos_table, filt, pad_mode: int = PAD_MEAN4, transpose: bool = True,
                 dtype: int = FDK_F32, stream=None, nviews: int | None = None,
                 src_view_offset: int = 0, dst_view_offset: int = 0, src_view_step: int = 1,
                 no_alloc: bool = False):
    """Fused cos weight + row filter.  src (views, nv, nu) float32 device tensor;
    dst (views, nu, nv) when transposed else (views, nv, nu), of the stored dtype.
    ``src_view_step`` reads every k-th source view (a sparse-view subset) without a
    copy.  ``no_alloc`` runs fdk_filter_views_noalloc (no device allocation; at
    most 1024 views and N <= 16384 per call).  Returns per-view (scale, offset)
    for FP16/BF16, else None."""
    nv, nu = src.shape[-2], src.shape[-1]
    if nviews is None:
        nviews = (src.shape[0] - src_view_offset + src_view_step - 1) // src_view_step
    n = int(nviews)
    if dtype not in (FDK_F32, FDK_F16, FDK_BF16):
        raise FdkError(f"filter_views: unknown dtype {dtype}")
    if n < 0 or src_view_offset < 0 or dst_view_offset < 0 or src_view_step < 1:
        raise FdkError("filter_views: negative view count or offset, or a step below 1")
    view = nu * nv
    esize = 4 if dtype == FDK_F32 else 2
    src_need = ((src_view_offset + (n - 1) * src_view_step) * view + view) * 4 if n else 0
    dst_need = (dst_view_offset + n) * view * esize if n else 0
    sp = C.c_void_p(_dev_buffer(src, "filter_views src", src_need) + src_view_offset * view * 4)
    dp = C.c_void_p(_dev_buffer(dst, "filter_views dst", dst_need, dtype) +
                    dst_view_offset * view * esize)
    if cos_table is not None:
        _dev_buffer(cos_table, "filter_views cos table", view * 4)
    _dev_buffer(filt, "filter_views filter", 0)
    so = np.zeros((max(n, 1), 2), np.float32) if dtype != FDK_F32 else None
    fn = lib().fdk_filter_views_noalloc if no_alloc else lib().fdk_filter_views
    _check(fn(sp, view * int(src_view_step), dp, view, int(dtype), nu, nv, int(n),
                                  _dev_ptr(cos_table) if cos_table is not None else None,
                                  _dev_ptr(filt), int(filt.numel()), int(pad_mode),
                                  1 if transpose else 0,
                                  _np_ptr(so) if so is not None else None, _stream(stream)),
           "filter_views")
    return so[:n] if so is not None else None


def FilterProjImage(para: Parameter, proj: np.ndarray, DEPTH: int = 32,
                    pad_mode: int = PAD_MEAN3) -> float:
    """Host views (nviews, nv, nu) filtered in place (cudaFftFilter.cu:394-452);
    returns the summed milliseconds of the batches' upload, filter and
    download, as the reference does."""
    if not isinstance(proj, np.ndarray) or proj.dtype != np.float32 or not proj.flags.c_contiguous:
        raise FdkError("FilterProjImage: proj must be a C-contiguous float32 numpy array")
    if proj.ndim != 3 or proj.shape[1:] != (para.nv, para.nu):
        raise FdkError(f"FilterProjImage: proj shape {proj.shape} is not (views, {para.nv}, "
                       f"{para.nu})")
    cos = para.cos_table()
    filt = para.filter().astype(np.float32)
    ms = C.c_float(0.0)
    _check(lib().fdk_filter_proj_image(_np_ptr(proj), para.nu, para.nv, proj.shape[0],
                                       _np_ptr(cos), _np_ptr(filt), len(filt), int(pad_mode),
                                       int(DEPTH), C.byref(ms)), "FilterProjImage")
    return float(ms.value)


class FilterEngine:
    """Filterlib FilterEngine (Filterlib/FilterLib.h:148-203) on the GPU: cos
    weighting (src is overwritten with the weighted image, as the reference
    does), mean-4 padded row filter, copy or transpose into dst."""

    def __init__(self, cos_image_table: np.ndarray, fft_kernel_real_f32: np.ndarray):
        import torch
        self.cos = torch.as_tensor(np.ascontiguousarray(cos_image_table, np.float32)).cuda()
        self.filt = torch.as_tensor(np.ascontiguousarray(fft_kernel_real_f32, np.float32)).cuda()
        self.nv, self.nu = cos_image_table.shape

    def Filtering(self, src: np.ndarray, dst: np.ndarray, transpose: bool = False):
        import torch
        assert src.shape == (self.nv, self.nu)
        s = torch.as_tensor(np.ascontiguousarray(src, np.float32)).cuda().reshape(1, self.nv, self.nu)
        d = torch.empty((1, self.nu, self.nv) if transpose else (1, self.nv, self.nu),
                        dtype=torch.float32, device="cuda")
        filter_views(s, d, self.cos, self.filt, PAD_MEAN4, transpose)
        CorrectImage_Real32(s, self.cos, s)
        torch.cuda.synchronize()
        src[...] = s.reshape(self.nv, self.nu).cpu().numpy()
        dst[...] = d.reshape(dst.shape).cpu().numpy()
        return self


# --------------------------------------------------------- forward projection
def forward_project(para: Parameter, vol, first: int, count: int, out, step_mm: float = 0.0,
                    mode: int = FP_CLAMP, stream=None):
    """Ray-march views first..first+count of ``para`` through device volume
    ``vol`` (nz, ny, nx) into device ``out`` (count, nv, nu)."""
    g = para.c()
    _dev_buffer(vol, "forward_project vol", para.nx * para.ny * para.nz * 4)
    _dev_buffer(out, "forward_project out", max(int(count), 0) * para.nu * para.nv * 4)
    _check(lib().fdk_fp_project(C.byref(g), _dev_ptr(vol), int(first), int(count), float(step_mm),
                                int(mode), _dev_ptr(out), _stream(stream)), "forward_project")


# ------------------------------------------------------------ random phantoms
def _ell_array(ell):
    arr = (_Ellipsoid * max(len(ell), 1))()
    for q, e in enumerate(ell):
        arr[q].cx, arr[q].cy, arr[q].cz = e["center"]
        arr[q].ax, arr[q].ay, arr[q].az = e["axes"]
        for t, v in enumerate(np.asarray(e["rot"], np.float64).reshape(9)):
            arr[q].rot[t] = v
        arr[q].density = e["density"]
    return arr


def phantom_rasterize(ell, para: Parameter, vol, stream=None):
    g = para.c()
    _dev_buffer(vol, "phantom_rasterize vol", para.nx * para.ny * para.nz * 4)
    _check(lib().fdk_phantom_rasterize(_ell_array(ell), len(ell), C.byref(g), _dev_ptr(vol),
                                       _stream(stream)), "phantom_rasterize")


def phantom_project(ell, para: Parameter, first: int, count: int, out, stream=None):
    g = para.c()
    _dev_buffer(out, "phantom_project out", max(int(count), 0) * para.nu * para.nv * 4)
    _check(lib().fdk_phantom_project(_ell_array(ell), len(ell), C.byref(g), int(first), int(count),
                                     _dev_ptr(out), _stream(stream)), "phantom_project")


from .phantom import random_ellipsoids  # noqa: E402
