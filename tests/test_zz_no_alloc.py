"""fdk_filter_views_noalloc (no counterpart in the reference): the filter entry
point that allocates no device memory gives exactly the results of
fdk_filter_views, through the C ABI, on two calls in a row (the second smaller:
fewer views, a shorter row and FFT), for FP32 (both layouts) and FP16 / BF16
storage; sizes beyond its limits are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import fdk_amd as F

pytestmark = pytest.mark.gpu


def case(torch, nviews, nu, nv, seed):
    p = F.Parameter(nx=nu, ny=nu, nz=nv, dx=1.0, dy=1.0, dz=1.0, nu=nu, nv=nv, du=1.5, dv=1.5,
                    nproj=nviews, volume_position_angle_rad=0.0, filter_type=5)
    g = torch.Generator().manual_seed(seed)
    src = torch.rand((nviews, nv, nu), generator=g).cuda()
    cos = torch.as_tensor(p.cos_table()).cuda()
    H = torch.as_tensor(p.filter().astype(np.float32)).cuda()
    return src, cos, H


def run_both(torch, src, cos, H, dtype, transpose):
    nviews, nv, nu = src.shape
    shape = (nviews, nu, nv) if transpose else (nviews, nv, nu)
    tdt = torch.float32 if dtype == F.FDK_F32 else torch.int16
    ref = torch.full(shape, 7, dtype=tdt, device="cuda")
    got = torch.full(shape, 9, dtype=tdt, device="cuda")
    so_ref = F.filter_views(src, ref, cos, H, F.PAD_MEAN4, transpose, dtype=dtype)
    so_got = F.filter_views(src, got, cos, H, F.PAD_MEAN4, transpose, dtype=dtype, no_alloc=True)
    torch.cuda.synchronize()
    return ref, got, so_ref, so_got


@pytest.mark.parametrize("dtype,transpose", [(F.FDK_F32, True), (F.FDK_F32, False),
                                             (F.FDK_F16, True), (F.FDK_BF16, True)])
def test_noalloc_filter_equals_allocating_filter(torch_cuda, dtype, transpose):
    torch = torch_cuda
    # first call: 40 views of 300 x 64 (N = 512); second, smaller: 7 of 120 x 48 (N = 256)
    for nviews, nu, nv, seed in ((40, 300, 64, 1), (7, 120, 48, 2)):
        print(f"noalloc dtype={dtype} transpose={transpose} {nviews} views {nu}x{nv}", flush=True)
        src, cos, H = case(torch, nviews, nu, nv, seed)
        ref, got, so_ref, so_got = run_both(torch, src, cos, H, dtype, transpose)
        assert torch.equal(ref, got)
        if dtype == F.FDK_F32:
            assert so_ref is None and so_got is None
        else:
            assert np.array_equal(so_ref, so_got)


def test_noalloc_filter_through_the_c_abi_at_the_view_limit(torch_cuda):
    """The C entry point itself, at 1024 views (its per-call limit), then 1025
    (rejected) and an FFT length over 16384 (rejected)."""
    torch = torch_cuda
    nviews, nu, nv = 1024, 64, 16
    src, cos, H = case(torch, nviews, nu, nv, 3)
    out = torch.empty((nviews, nu, nv), dtype=torch.int16, device="cuda")
    ref = torch.empty_like(out)
    so = np.zeros(2 * nviews, np.float32)
    L = F.lib()
    rc = L.fdk_filter_views_noalloc(src.data_ptr(), nu * nv, out.data_ptr(), nu * nv, F.FDK_BF16,
                                    nu, nv, nviews, cos.data_ptr(), H.data_ptr(), H.numel(),
                                    F.PAD_MEAN4, 1, so.ctypes.data_as(C.c_void_p), None)
    assert rc == 0, F.last_error()
    so_ref = F.filter_views(src, ref, cos, H, F.PAD_MEAN4, True, dtype=F.FDK_BF16)
    torch.cuda.synchronize()
    assert torch.equal(out, ref) and np.array_equal(so.reshape(-1, 2), so_ref)
    rc = L.fdk_filter_views_noalloc(src.data_ptr(), nu * nv, out.data_ptr(), nu * nv, F.FDK_BF16,
                                    nu, nv, nviews + 1, cos.data_ptr(), H.data_ptr(), H.numel(),
                                    F.PAD_MEAN4, 1, so.ctypes.data_as(C.c_void_p), None)
    assert rc == F.FDK_ERR_INVALID and "1024" in F.last_error()
    rc = L.fdk_filter_views_noalloc(src.data_ptr(), nu * nv, out.data_ptr(), nu * nv, F.FDK_F32,
                                    nu, nv, 1, cos.data_ptr(), H.data_ptr(), 32768, F.PAD_MEAN4,
                                    1, None, None)
    assert rc == F.FDK_ERR_INVALID and "16384" in F.last_error()
