"""The wave-strip backprojector (bp_wave_kernel) on the shapes that exercise its
paths, against the oracle's strict-bilinear backprojection through the C ABI:

* a tall volume whose z-runs leave the detector (runs partly inside: the
  checked tap loop; runs wholly outside: skipped views);
* a deep slab pair (512-voxel runs, ZPT 8) and shallow ones (ZPT 4, two strip
  quads per lane, runs padded past the slab; one with an odd column count);
* a detector with a v offset, whose matrices are not mirror-symmetric: the
  block-strip kernels run instead, with the same results;
* 16-bit views dequantized in the strip combine, against the FP32 path on the
  same values.

Projections are exact line integrals of seeded ellipsoids: their edges are
steep, and FP32 voxel-to-detector positions carry ~1e-5 px of rounding, so the
bound is 2e-5 relative RMSE for these phantoms (a sharper one reached 1e-4)."""
import numpy as np
import pytest

import fdk_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def rel_rmse(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(p, proj, mats32, top, bottom, dtype=F.FDK_F32):
    """Views (nproj, nv, nu) -> transposed device batches of 32."""
    import torch
    views = torch.as_tensor(np.ascontiguousarray(proj.transpose(0, 2, 1))).cuda()
    so = None
    if dtype != F.FDK_F32:
        lo = views.reshape(p.nproj, -1).min(1).values
        hi = views.reshape(p.nproj, -1).max(1).values
        q = (views - lo[:, None, None]) / (hi - lo)[:, None, None]
        store = (q.half() if dtype == F.FDK_F16 else q.bfloat16()).view(torch.int16).contiguous()
        so = torch.stack([hi - lo, lo], 1).cpu().numpy().astype(np.float32)
        deq = (store.view(torch.float16 if dtype == F.FDK_F16 else torch.bfloat16).float()
               * (hi - lo)[:, None, None] + lo[:, None, None])
        proj = deq.transpose(1, 2).cpu().numpy()  # what the kernel samples
    else:
        store = views
    bp = F.BackProjectionFDK(p.nv, p.nu, 32, p.nx, p.ny, p.nz, top, bottom)
    bp.set_stream(torch.cuda.current_stream().cuda_stream)
    for b in range(0, p.nproj, 32):
        n = min(32, p.nproj - b)
        bp.bp_device(store, p.nv, p.nu, mats32[b:b + n], n, dtype=dtype, offset_views=b,
                     scale_offset=None if so is None else so[b:b + n])
    bp.sync()
    vol = bp.GetVolumeData()
    bp.close()
    return vol, proj


def oracle(p, proj, mats32, top, bottom):
    g = O.Geom(**{k: getattr(p, k) for k in O.Geom.__dataclass_fields__})
    m = mats32.astype(np.float64)
    parts = [O.bp(g, proj, m, O.BP_GPU_STRICT, z0=top[2], nzs=top[5])]
    if bottom[5]:
        parts.append(O.bp(g, proj, m, O.BP_GPU_STRICT, z0=bottom[2], nzs=bottom[5]))
    return np.concatenate(parts)


def smooth_views(p, seed):
    """Exact projections of a seeded random-ellipsoid phantom."""
    import torch
    ell = F.random_ellipsoids(8, seed=seed, radius_mm=0.4 * p.nx * p.dx,
                              half_height_mm=0.45 * p.nz * p.dz,
                              axes_mm=(4 * p.dx, 0.3 * p.nx * p.dx))
    raw = torch.empty((p.nproj, p.nv, p.nu), device="cuda")
    F.phantom_project(ell, p, 0, p.nproj, raw)
    return raw.cpu().numpy()


CASES = {
    # 64-wide detector rows (v) for a 200-deep volume: most runs cross the edge
    "tall": (dict(nx=40, ny=36, nz=200, dx=2.0, dy=2.0, dz=2.0, nu=72, nv=64, du=1.8, dv=1.8,
                  nproj=40, volume_position_angle_rad=0.2), "mirror"),
    # 100-deep slabs (ZPT 4) with an odd column count along y
    "odd": (dict(nx=40, ny=37, nz=200, dx=2.0, dy=2.0, dz=2.0, nu=72, nv=64, du=1.8, dv=1.8,
                 nproj=33, volume_position_angle_rad=0.3), "mirror"),
    # 1024 deep: 512-voxel runs (ZPT 8)
    "deep": (dict(nx=32, ny=32, nz=1024, dx=0.5, dy=0.5, dz=0.5, nu=48, nv=800, du=0.75, dv=0.75,
                  nproj=36, volume_position_angle_rad=0.0), "mirror"),
    # a thin mirror pair z in [40, 72) and [nz-72, nz-40): ZPT 4 runs padded past it
    "thin": (dict(nx=48, ny=40, nz=160, dx=1.0, dy=1.0, dz=1.0, nu=64, nv=240, du=1.5, dv=1.5,
                  nproj=33, volume_position_angle_rad=-0.4), "thin"),
    # detector v offset: not mirror-symmetric, block-strip kernels
    "offset": (dict(nx=48, ny=40, nz=36, dx=2.0, dy=2.0, dz=2.0, nu=64, nv=56, du=1.6, dv=1.6,
                    offset_nv=2.5, nproj=40, volume_position_angle_rad=0.1), "mirror"),
}


def boxes(p, kind):
    if kind == "thin":
        return (0, 0, 40, p.nx, p.ny, 32), (0, 0, p.nz - 72, p.nx, p.ny, 32)
    return (0, 0, 0, p.nx, p.ny, p.nz // 2), (0, 0, p.nz // 2, p.nx, p.ny, p.nz // 2)


@pytest.mark.parametrize("name", list(CASES))
def test_wave_paths_match_oracle(torch_cuda, name):
    kw, kind = CASES[name]
    p = F.Parameter(**kw, filter_type=5)
    print(f"wave kernel case {name}: {p.nx}x{p.ny}x{p.nz}, {p.nu}x{p.nv}, {p.nproj} views",
          flush=True)
    proj = smooth_views(p, 7)
    mats = p.proj_matrices().astype(np.float32)
    top, bot = boxes(p, kind)
    vol, _ = run(p, proj, mats, top, bot)
    ref = oracle(p, proj, mats, top, bot)
    err = rel_rmse(vol, ref)
    print(f"  rel RMSE {err:.2e}", flush=True)
    assert err < 2e-5
    assert float(np.abs(vol - ref).max()) < 2e-4 * float(np.abs(ref).max())


@pytest.mark.parametrize("dtype", [F.FDK_F16, F.FDK_BF16])
def test_wave_16bit_views_match_fp32_path(torch_cuda, dtype):
    """16-bit views against the FP32 path on the values they dequantize to
    (the strip combine folds scale and offset in), and against the oracle."""
    kw, kind = CASES["tall"]
    p = F.Parameter(**kw, filter_type=5)
    proj = smooth_views(p, 7)
    mats = p.proj_matrices().astype(np.float32)
    top, bot = boxes(p, kind)
    vol, deq = run(p, proj, mats, top, bot, dtype)
    vol32, _ = run(p, np.ascontiguousarray(deq), mats, top, bot)
    same = rel_rmse(vol, vol32)
    err = rel_rmse(vol, oracle(p, deq, mats, top, bot))
    print(f"  dtype {dtype}: vs the FP32 path on the stored values {same:.2e}, vs the oracle "
          f"{err:.2e}", flush=True)
    assert same < 2e-6
    assert err < 2e-5
