"""Argument checks of the Python device wrappers, on the host: every wrapper
refuses a buffer that is too small, of the wrong element type, not contiguous
or not on the GPU before it launches anything (a launch over a too-small
buffer would read or write outside it).  No GPU is touched: the device
tensors are stand-ins that report their size and type."""
import numpy as np
import pytest

import fdk_amd as F


class FakeDev:
    """Looks like a contiguous CUDA tensor to the checks; never dereferenced."""

    def __init__(self, numel, dtype="torch.float32", esize=4, cuda=True, contiguous=True,
                 shape=None):
        self._n, self.dtype, self._e = numel, dtype, esize
        self.is_cuda, self._c = cuda, contiguous
        self.shape = shape if shape is not None else (numel,)

    def numel(self):
        return self._n

    def element_size(self):
        return self._e

    def is_contiguous(self):
        return self._c

    def data_ptr(self):
        return 0x10000


def bp_stub():
    bp = F.BackProjectionFDK.__new__(F.BackProjectionFDK)
    bp._h = None
    return bp


W = H = 64
VIEW = W * H
MATS = np.zeros((32, 12), np.float32)


def test_bp_device_rejects_small_or_mistyped_views():
    bp = bp_stub()
    with pytest.raises(F.FdkError, match="bytes"):
        bp.bp_device(FakeDev(VIEW * 31), W, H, MATS, 32)          # one view short
    with pytest.raises(F.FdkError, match="bytes"):
        bp.bp_device(FakeDev(VIEW * 32), W, H, MATS, 32, offset_views=1)
    with pytest.raises(F.FdkError, match="element type"):
        bp.bp_device(FakeDev(VIEW * 32, "torch.int16", 2), W, H, MATS, 32)
    with pytest.raises(F.FdkError, match="not on the GPU"):
        bp.bp_device(FakeDev(VIEW * 32, cuda=False), W, H, MATS, 32)
    with pytest.raises(F.FdkError, match="not contiguous"):
        bp.bp_device(FakeDev(VIEW * 32, contiguous=False), W, H, MATS, 32)
    with pytest.raises(F.FdkError, match="matrices"):
        bp.bp_device(FakeDev(VIEW * 32), W, H, MATS[:31], 32)
    with pytest.raises(F.FdkError, match="scale_offset"):
        bp.bp_device(FakeDev(VIEW * 32, "torch.int16", 2), W, H, MATS, 32, dtype=F.FDK_F16)
    with pytest.raises(F.FdkError, match="scale/offset"):
        bp.bp_device(FakeDev(VIEW * 32, "torch.int16", 2), W, H, MATS, 32, dtype=F.FDK_F16,
                     scale_offset=np.ones((31, 2), np.float32))
    with pytest.raises(F.FdkError, match="torch tensor"):
        bp.bp_device(np.zeros(VIEW * 32, np.float32), W, H, MATS, 32)


def test_filter_views_rejects_small_destination():
    src = FakeDev(VIEW * 32, shape=(32, H, W))
    cos = FakeDev(VIEW)
    filt = FakeDev(128)
    with pytest.raises(F.FdkError, match="dst"):
        F.filter_views(src, FakeDev(VIEW * 31), cos, filt, nviews=32)
    with pytest.raises(F.FdkError, match="src"):
        F.filter_views(FakeDev(VIEW * 16, shape=(16, H, W)), FakeDev(VIEW * 32), cos, filt,
                       nviews=32)
    with pytest.raises(F.FdkError, match="src"):
        F.filter_views(src, FakeDev(VIEW * 32), cos, filt, nviews=16, src_view_step=3)
    with pytest.raises(F.FdkError, match="element type"):
        F.filter_views(src, FakeDev(VIEW * 32), cos, filt, nviews=32, dtype=F.FDK_F16)
    with pytest.raises(F.FdkError, match="cos table"):
        F.filter_views(src, FakeDev(VIEW * 32), FakeDev(VIEW - 1), filt, nviews=32)


def test_pipeline_push_checks_host_views():
    pipe = F.Pipeline.__new__(F.Pipeline)
    pipe._h, pipe._view, pipe.store_dtype = None, VIEW, F.FDK_F16
    with pytest.raises(F.FdkError, match="host array holds"):
        pipe.push_raw(np.zeros((31, H, W), np.float32), 32, 0)
    with pytest.raises(F.FdkError, match="element size"):
        pipe.push_raw(np.zeros((32, H, W), np.float64), 32, 0)
    with pytest.raises(F.FdkError, match="element size"):
        pipe.push_filtered(np.zeros((32, H, W), np.float32), 32, 0,
                           scale_offset=np.ones((32, 2), np.float32))
    with pytest.raises(F.FdkError, match="scale/offset"):
        pipe.push_filtered(np.zeros((32, H, W), np.int16), 32, 0,
                           scale_offset=np.ones((8, 2), np.float32))
    with pytest.raises(F.FdkError, match="GPU tensor"):
        pipe.push_raw(FakeDev(VIEW * 32), 32, 0)


def test_projector_and_phantom_outputs_checked():
    p = F.Parameter(nx=16, ny=16, nz=16, nu=24, nv=20, nproj=8)
    with pytest.raises(F.FdkError, match="out"):
        F.forward_project(p, FakeDev(16 ** 3), 0, 8, FakeDev(24 * 20 * 7))
    with pytest.raises(F.FdkError, match="vol"):
        F.forward_project(p, FakeDev(16 ** 3 - 1), 0, 8, FakeDev(24 * 20 * 8))
    with pytest.raises(F.FdkError, match="out"):
        F.phantom_project([], p, 0, 8, FakeDev(24 * 20 * 7))
    with pytest.raises(F.FdkError, match="vol"):
        F.phantom_rasterize([], p, FakeDev(16 ** 3 - 1))


def test_filter_proj_image_checks_host_stack():
    p = F.Parameter(nx=16, ny=16, nz=16, nu=24, nv=20, nproj=8)
    with pytest.raises(F.FdkError, match="shape"):
        F.FilterProjImage(p, np.zeros((8, 24, 20), np.float32))
    with pytest.raises(F.FdkError, match="float32"):
        F.FilterProjImage(p, np.zeros((8, 20, 24), np.float64))
