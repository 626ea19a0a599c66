"""GPU parity: every kernel through the C ABI against the oracle on the same
seeded inputs.  Tolerances are written per test: the backprojector and the
forward projector compute in FP32 against an FP64 oracle; the filter runs an
FP32 FFT against the oracle's FP64 one."""
import numpy as np
import pytest

import fdk_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def pair(**kw):
    p = F.Parameter(**kw)
    g = O.Geom(**{k: getattr(p, k) for k in O.Geom.__dataclass_fields__})
    return p, g


def rel_rmse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# FP32 voxel->detector coordinates carry ~1e-7 relative rounding (a few 1e-6
# pixels here); on white-noise views that moves each bilinear sample by a few
# 1e-6 of its value, hence 2e-5 relative RMSE for random inputs (smooth
# projections land far lower; the reconstruction target is 1e-4).
SMALL = dict(nx=48, ny=40, nz=36, dx=2.0, dy=2.0, dz=2.0, nu=64, nv=56, du=1.6, dv=1.6,
             nproj=40, volume_position_angle_rad=-0.3)


def _proj(p, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((p.nproj, p.nv, p.nu)).astype(np.float32)


def _run_fdk(p, proj, mats32, top, bottom, transpose=1):
    bp = F.BackProjectionFDK(p.nv if transpose else p.nu, p.nu if transpose else p.nv, 32,
                             p.nx, p.ny, p.nz, top, bottom)
    for b in range(0, p.nproj, 32):
        n = min(32, p.nproj - b)
        if transpose:
            views = np.ascontiguousarray(proj[b:b + n].transpose(0, 2, 1))
            bp.BP(1, views, p.nv, p.nu, mats32[b:b + n], n)
        else:
            bp.BP(0, proj[b:b + n], p.nu, p.nv, mats32[b:b + n], n)
    assert bp.nBpCount == p.nproj
    vol = bp.GetVolumeData()
    bp.close()
    return vol


@pytest.mark.parametrize("transpose", [1, 0])
def test_bp_fdk_mirror_full_volume(torch_cuda, transpose):
    p, g = pair(**SMALL)
    proj = _proj(p)
    mats = p.proj_matrices().astype(np.float32)
    top = (0, 0, 0, p.nx, p.ny, p.nz // 2)
    bot = (0, 0, p.nz // 2, p.nx, p.ny, p.nz // 2)
    vol = _run_fdk(p, proj, mats, top, bot, transpose)
    ref = O.bp(g, proj, mats.astype(np.float64), O.BP_GPU_STRICT)
    assert vol.shape == ref.shape
    assert rel_rmse(vol, ref) < 2e-5
    assert float(np.abs(vol - ref).max()) < 1e-4 * float(np.abs(ref).max())


def test_bp_fdk_sub_boxes(torch_cuda):
    p, g = pair(**SMALL)
    proj = _proj(p, 1)
    mats = p.proj_matrices().astype(np.float32)
    m64 = mats.astype(np.float64)
    # plain slab (no mirror), offset in x and y
    top = (4, 6, 5, 30, 20, 17)
    vol = _run_fdk(p, proj, mats, top, (0, 0, 0, 0, 0, 0))
    ref = O.bp(g, proj, m64, O.BP_GPU_STRICT, z0=5, nzs=17)[:, 6:26, 4:34]
    assert rel_rmse(vol, ref) < 2e-5
    # mirror pair of thin slabs: z in [3, 10) and [nz-10, nz-3)
    top = (0, 0, 3, p.nx, p.ny, 7)
    bot = (0, 0, p.nz - 10, p.nx, p.ny, 7)
    vol = _run_fdk(p, proj, mats, top, bot)
    ref = np.concatenate([O.bp(g, proj, m64, O.BP_GPU_STRICT, z0=3, nzs=7),
                          O.bp(g, proj, m64, O.BP_GPU_STRICT, z0=p.nz - 10, nzs=7)])
    assert rel_rmse(vol, ref) < 2e-5


def test_bp_general_matrices(torch_cuda):
    """Matrices with a z term in rows 0 and 2 (a tilted detector) take the
    per-voxel kernel."""
    p, g = pair(**SMALL)
    proj = _proj(p, 2)
    mats = p.proj_matrices()
    mats[:, 2] += 0.3 * mats[:, 0]
    mats[:, 10] += 0.001
    mats = mats.astype(np.float32)
    vol = _run_fdk(p, proj, mats, (0, 0, 0, p.nx, p.ny, p.nz), (0, 0, 0, 0, 0, 0))
    ref = O.bp(g, proj, mats.astype(np.float64), O.BP_GPU_STRICT)
    assert rel_rmse(vol, ref) < 2e-5


def _quantize(proj_t, dtype):
    lo = proj_t.reshape(proj_t.shape[0], -1).min(1)
    hi = proj_t.reshape(proj_t.shape[0], -1).max(1)
    q = (proj_t - lo[:, None, None]) / (hi - lo)[:, None, None]
    if dtype == F.FDK_F16:
        stored = q.astype(np.float16)
        deq = stored.astype(np.float64)
        bits = stored.view(np.uint16)
    else:
        b = q.astype(np.float32).view(np.uint32).astype(np.uint64)
        r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)
        bits = r
        deq = (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    deq = deq * (hi - lo)[:, None, None] + lo[:, None, None]
    so = np.stack([hi - lo, lo], 1).astype(np.float32)
    return bits, deq, so


@pytest.mark.parametrize("dtype", [F.FDK_F32, F.FDK_F16, F.FDK_BF16])
def test_bp_device_dtypes(torch_cuda, dtype):
    torch = torch_cuda
    p, g = pair(**SMALL)
    proj = _proj(p, 3)
    mats = p.proj_matrices().astype(np.float32)
    proj_t = np.ascontiguousarray(proj.transpose(0, 2, 1))  # (views, nu, nv)
    if dtype == F.FDK_F32:
        dev = torch.as_tensor(proj_t).cuda()
        deq_t, so = proj_t.astype(np.float64), None
    else:
        bits, deq_t, so = _quantize(proj_t, dtype)
        dev = torch.as_tensor(bits.view(np.int16)).cuda()
    bp = F.BackProjectionFDK(p.nv, p.nu, 32, p.nx, p.ny, p.nz, (0, 0, 0, p.nx, p.ny, p.nz // 2),
                             (0, 0, p.nz // 2, p.nx, p.ny, p.nz // 2))
    bp.set_stream(torch.cuda.current_stream().cuda_stream)
    for b in range(0, p.nproj, 32):
        n = min(32, p.nproj - b)
        bp.bp_device(dev, p.nv, p.nu, mats[b:b + n], n, dtype=dtype, offset_views=b,
                     scale_offset=None if so is None else so[b:b + n])
    bp.sync()
    vol = bp.GetVolumeData()
    deq = np.ascontiguousarray(deq_t.transpose(0, 2, 1)).astype(np.float32)
    ref = O.bp(g, deq, mats.astype(np.float64), O.BP_GPU_STRICT)
    assert rel_rmse(vol, ref) < 2e-5


def test_rtk_matches_oracle(torch_cuda):
    p, g = pair(**SMALL)
    proj = _proj(p, 4)
    mats = p.proj_matrices().astype(np.float32)
    rtk = F.BackProjectionRTK(p.nu, p.nv, 32, p.nx, p.ny, p.nz)
    for b in range(0, p.nproj, 32):
        n = min(32, p.nproj - b)
        rtk.BP(0, proj[b:b + n], p.nu, p.nv, mats[b:b + n], n)
    vol = rtk.GetVolumeData()
    ref = O.bp(g, proj, mats.astype(np.float64), O.BP_GPU_STRICT)
    assert rel_rmse(vol, ref) < 2e-5


def test_reconstruct_conebeam_device(torch_cuda):
    torch = torch_cuda
    p, g = pair(**SMALL)
    proj = _proj(p, 5)[:32]
    mats = p.proj_matrices()[:32].astype(np.float32)
    vin = torch.full((p.nz, p.ny, p.nx), 0.25, device="cuda")
    vout = torch.zeros_like(vin)
    F.reconstruct_conebeam((p.nu, p.nv, 32), (p.nx, p.ny, p.nz), mats, vin, vout,
                           torch.as_tensor(proj).cuda())
    torch.cuda.synchronize()
    ref = O.bp(g, proj, mats.astype(np.float64), O.BP_GPU_STRICT) + 0.25
    assert rel_rmse(vout.cpu().numpy(), ref) < 2e-5


FILTER_CASES = [(256, 64, O.PAD_MEAN4), (100, 37, O.PAD_MEAN3), (2048, 8, O.PAD_MEAN4),
                (4000, 4, O.PAD_ZERO), (1024, 6, O.PAD_MEAN3), (24, 20, O.PAD_MEAN4),
                (600, 32, O.PAD_MEAN4), (2048, 48, O.PAD_MEAN3)]


@pytest.mark.parametrize("nu,nv,pad", FILTER_CASES)
@pytest.mark.parametrize("transpose", [0, 1])
def test_filter_views_matches_oracle(torch_cuda, nu, nv, pad, transpose):
    torch = torch_cuda
    p, g = pair(nu=nu, nv=nv, du=0.4, dv=0.4, nproj=3)
    rng = np.random.default_rng(nu + nv)
    src = rng.random((3, nv, nu)).astype(np.float32) * 50
    cos = p.cos_table()
    H = p.filter().astype(np.float32)
    d_src = torch.as_tensor(src).cuda()
    d_dst = torch.empty((3, nu, nv) if transpose else (3, nv, nu), device="cuda")
    F.filter_views(d_src, d_dst, torch.as_tensor(cos).cuda(), torch.as_tensor(H).cuda(), pad,
                   bool(transpose))
    torch.cuda.synchronize()
    out = d_dst.cpu().numpy()
    if transpose:
        out = out.transpose(0, 2, 1)
    ref = O.filter_rows(O.cos_weight(src, cos), H.astype(np.float64), pad)
    assert rel_rmse(out, ref) < 2e-6
    assert np.abs(out - ref).max() < 2e-5 * np.abs(ref).max()


def test_filter_row_and_correct(torch_cuda):
    torch = torch_cuda
    p, g = pair(nu=300, nv=30, du=0.5, dv=0.5, filter_type=2)
    rng = np.random.default_rng(7)
    src = rng.standard_normal((4, 30, 300)).astype(np.float32)
    cos = p.cos_table()
    d = torch.as_tensor(src).cuda()
    F.CorrectImage_Real32(d, torch.as_tensor(cos).cuda(), d)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), src * cos[None])  # one float multiply: bit-exact
    H = p.filter().astype(np.float32)
    out = torch.empty_like(d)
    F.FilterRow_Real32(d, out, torch.as_tensor(H).cuda())
    torch.cuda.synchronize()
    ref = O.filter_rows(src * cos[None], H.astype(np.float64), O.PAD_MEAN3)
    assert rel_rmse(out.cpu().numpy(), ref) < 2e-6


def test_filter_proj_image_host_in_place(torch_cuda):
    p, g = pair(nu=128, nv=50, du=0.8, dv=0.8, nproj=40)
    rng = np.random.default_rng(8)
    proj = rng.random((40, 50, 128)).astype(np.float32)
    ref = O.filter_rows(O.cos_weight(proj, p.cos_table()), p.filter().astype(np.float32).astype(np.float64),
                        O.PAD_MEAN3)
    ms = F.FilterProjImage(p, proj, DEPTH=16)
    assert rel_rmse(proj, ref) < 2e-6
    assert 0.0 < ms < 60e3, ms  # summed batch time in milliseconds, as the reference returns


def test_filter_engine_transposed(torch_cuda):
    p, g = pair(nu=96, nv=40, du=0.8, dv=0.8)
    rng = np.random.default_rng(9)
    src = rng.random((40, 96)).astype(np.float32)
    cos = p.cos_table()
    H = p.filter().astype(np.float32)
    eng = F.FilterEngine(cos, H)
    s = src.copy()
    dst = np.empty((96, 40), np.float32)
    eng.Filtering(s, dst, transpose=True)
    assert np.array_equal(s, src * cos)  # src overwritten by the weighted image (FilterLib.h:181)
    ref = O.filter_rows(src * cos, H.astype(np.float64), O.PAD_MEAN4)
    assert rel_rmse(dst.T, ref) < 2e-6


@pytest.mark.parametrize("dtype", [F.FDK_F16, F.FDK_BF16])
def test_filter_views_low_precision_store(torch_cuda, dtype):
    torch = torch_cuda
    p, g = pair(nu=512, nv=64, du=0.4, dv=0.4, nproj=5)
    rng = np.random.default_rng(10)
    src = (rng.random((5, 64, 512)) * 100).astype(np.float32)
    d_src = torch.as_tensor(src).cuda()
    cos = torch.as_tensor(p.cos_table()).cuda()
    H = torch.as_tensor(p.filter().astype(np.float32)).cuda()
    ref = torch.empty((5, 512, 64), device="cuda")
    F.filter_views(d_src, ref, cos, H, F.PAD_MEAN4, True)
    q = torch.empty((5, 512, 64), dtype=torch.int16, device="cuda")
    so = F.filter_views(d_src, q, cos, H, F.PAD_MEAN4, True, dtype=dtype)
    torch.cuda.synchronize()
    r = ref.cpu().numpy()
    bits = q.cpu().numpy().view(np.uint16)
    if dtype == F.FDK_F16:
        qv = bits.view(np.float16).astype(np.float64)
        tol = 2.0 ** -11
    else:
        qv = (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        tol = 2.0 ** -8
    deq = qv * so[:, 0, None, None] + so[:, 1, None, None]
    lo, hi = r.reshape(5, -1).min(1), r.reshape(5, -1).max(1)
    assert np.allclose(so[:, 1], lo) and np.allclose(so[:, 0], hi - lo, rtol=1e-6)
    assert qv.min() >= 0.0 and qv.max() <= 1.0
    err = np.abs(deq - r).reshape(5, -1).max(1) / (hi - lo)
    assert (err <= tol * 1.01).all()


@pytest.mark.parametrize("mode", [F.FP_CLAMP, F.FP_ZERO])
@pytest.mark.parametrize("step", [0.0, 0.5])
def test_forward_projection_matches_oracle(torch_cuda, mode, step):
    torch = torch_cuda
    p, g = pair(nx=40, ny=36, nz=32, nu=72, nv=64, du=1.6, dv=1.6, nproj=12,
                volume_position_angle_rad=0.2)
    rng = np.random.default_rng(11)
    vol = rng.random((p.nz, p.ny, p.nx)).astype(np.float32)
    out = torch.empty((4, p.nv, p.nu), device="cuda")
    F.forward_project(p, torch.as_tensor(vol).cuda(), 3, 4, out, step, mode)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = np.stack([O.fp(g, vol, 3 + s, step if step > 0 else None, mode) for s in range(4)])
    # FP32 march vs FP64 march: a ray may take one sample more or less at its exit
    close = np.abs(got - ref) <= 1e-4 * np.abs(ref).max()
    assert close.mean() > 0.995
    assert rel_rmse(got, ref) < 2e-3


def test_phantom_kernels_match_oracle(torch_cuda):
    torch = torch_cuda
    p, g = pair(nx=64, ny=60, nz=56, nu=80, nv=72, du=1.6, dv=1.6, nproj=8)
    ell = F.random_ellipsoids(6, seed=3, radius_mm=28, half_height_mm=26, axes_mm=(4, 20))
    vol = torch.empty((p.nz, p.ny, p.nx), device="cuda")
    F.phantom_rasterize(ell, p, vol)
    proj = torch.empty((3, p.nv, p.nu), device="cuda")
    F.phantom_project(ell, p, 2, 3, proj)
    torch.cuda.synchronize()
    ref_vol = O.phantom_rasterize(ell, g)
    assert np.mean(vol.cpu().numpy() != ref_vol) < 1e-4
    ref_proj = np.stack([O.phantom_project(ell, g, 2 + s) for s in range(3)])
    assert rel_rmse(proj.cpu().numpy(), ref_proj) < 1e-6
