"""Edge cases through the C ABI: empty batches, a single view, ragged batches,
odd volume and detector sizes (which take the direct-gather fallbacks), the
largest filter length, and zero-count calls that must be no-ops."""
import numpy as np
import pytest

import fdk_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def pair(**kw):
    p = F.Parameter(**kw)
    g = O.Geom(**{k: getattr(p, k) for k in O.Geom.__dataclass_fields__})
    return p, g


def test_empty_batch_is_a_no_op(torch_cuda):
    p, _ = pair(nx=16, ny=16, nz=16, nu=24, nv=20, nproj=8)
    bp = F.BackProjectionFDK(p.nv, p.nu, 32, 16, 16, 16)
    empty = np.zeros((1, p.nu, p.nv), np.float32)
    bp.BP(1, empty, p.nv, p.nu, np.zeros((1, 12), np.float32), 0)
    assert bp.nBpCount == 0
    assert not np.any(bp.GetVolumeData())
    bp.close()


@pytest.mark.parametrize("nproj", [1, 33])
@pytest.mark.parametrize("shape", [dict(nx=37, ny=29, nz=23, nu=63, nv=45),
                                   dict(nx=40, ny=36, nz=20, nu=64, nv=48)])
def test_single_and_ragged_batches_odd_sizes(torch_cuda, nproj, shape):
    p, g = pair(dx=2.5, dy=2.5, dz=2.5, du=1.9, dv=1.9, nproj=nproj,
                volume_position_angle_rad=0.4, **shape)
    rng = np.random.default_rng(nproj + shape["nx"])
    proj = rng.standard_normal((nproj, p.nv, p.nu)).astype(np.float32)
    mats = p.proj_matrices().astype(np.float32)
    bp = F.BackProjectionFDK(p.nv, p.nu, 32, p.nx, p.ny, p.nz)
    for b in range(0, nproj, 32):
        n = min(32, nproj - b)
        bp.BP(1, np.ascontiguousarray(proj[b:b + n].transpose(0, 2, 1)), p.nv, p.nu,
              mats[b:b + n], n)
    assert bp.nBpCount == nproj
    vol = bp.GetVolumeData()
    bp.close()
    ref = O.bp(g, proj, mats.astype(np.float64), O.BP_GPU_STRICT)
    assert rel(vol, ref) < 2e-5


def test_zero_count_calls_are_no_ops(torch_cuda):
    torch = torch_cuda
    p, _ = pair(nx=16, ny=16, nz=16, nu=32, nv=16, nproj=4)
    src = torch.randn((1, p.nv, p.nu), device="cuda")
    dst = torch.full((1, p.nu, p.nv), 7.0, device="cuda")
    cos = torch.as_tensor(p.cos_table()).cuda()
    H = torch.as_tensor(p.filter().astype(np.float32)).cuda()
    F.filter_views(src, dst, cos, H, F.PAD_MEAN4, True, nviews=0)
    vol = torch.ones((16, 16, 16), device="cuda")
    out = torch.full((1, p.nv, p.nu), 5.0, device="cuda")
    F.forward_project(p, vol, 0, 0, out)
    torch.cuda.synchronize()
    assert torch.all(dst == 7.0) and torch.all(out == 5.0)


def test_largest_filter_length_on_a_4096_detector(torch_cuda):
    torch = torch_cuda
    p, _ = pair(nu=4096, nv=6, du=0.1, dv=0.1, nproj=2)
    assert p.filter_length() == 8192
    rng = np.random.default_rng(3)
    src = (rng.random((2, p.nv, p.nu)) * 10).astype(np.float32)
    cos = p.cos_table()
    H = p.filter().astype(np.float32)
    dst = torch.empty((2, p.nu, p.nv), device="cuda")
    F.filter_views(torch.as_tensor(src).cuda(), dst, torch.as_tensor(cos).cuda(),
                   torch.as_tensor(H).cuda(), F.PAD_MEAN4, True)
    out = dst.cpu().numpy().transpose(0, 2, 1)
    ref = O.filter_rows(O.cos_weight(src, cos), H.astype(np.float64), O.PAD_MEAN4)
    assert rel(out, ref) < 2e-6
