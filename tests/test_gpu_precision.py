"""Projection storage precision at the 2048^3 geometry (2048^2 views of
0.2 mm, 0.1333 mm voxels): FP32, FP16 and BF16 stored views (16-bit ones
scaled per view to [0, 1]), FP32 interpolation and accumulation, compared with
an FP64 reconstruction (oracle FP64 filter + FP64 backprojection with FP64
matrices) of the central 256^3 sub-volume from every 8th of the 4096 views."""
import numpy as np
import pytest

import fdk_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def test_storage_precision_against_fp64(torch_cuda):
    torch = torch_cuda
    n_full, step, sub = 2048, 8, 256
    p = F.Parameter(nx=sub, ny=sub, nz=sub, dx=0.2 / 1.5, dy=0.2 / 1.5, dz=0.2 / 1.5,
                    nu=n_full, nv=n_full, du=0.2, dv=0.2, sod=1000.0, sid=1500.0,
                    nproj=4096 // step, volume_position_angle_rad=0.0, filter_type=5)
    V = p.nproj
    # same phantom as the 2048^3 benchmark; the central 256^3 is inside it
    ell = F.random_ellipsoids(40, seed=2, radius_mm=120.0, half_height_mm=120.0, axes_mm=(5, 100))
    raw = torch.empty((V, p.nv, p.nu), device="cuda")
    for b in range(0, V, 32):
        F.phantom_project(ell, p, b, 32, raw[b:b + 32])
    cos = torch.as_tensor(p.cos_table()).cuda()
    H32 = p.filter().astype(np.float32)
    H = torch.as_tensor(H32).cuda()
    mats = p.proj_matrices().astype(np.float32)
    vols = {}
    for name, dt in (("f32", F.FDK_F32), ("f16", F.FDK_F16), ("bf16", F.FDK_BF16)):
        store = torch.empty((V, p.nu, p.nv), device="cuda",
                            dtype=torch.float32 if dt == F.FDK_F32 else torch.int16)
        so = F.filter_views(raw, store, cos, H, F.PAD_MEAN4, True, dtype=dt)
        bp = F.BackProjectionFDK(p.nv, p.nu, 32, sub, sub, sub, (0, 0, 0, sub, sub, sub // 2),
                                 (0, 0, sub // 2, sub, sub, sub // 2))
        bp.set_stream(torch.cuda.current_stream().cuda_stream)
        for b in range(0, V, 32):
            bp.bp_device(store, p.nv, p.nu, mats[b:b + 32], 32, dtype=dt, offset_views=b,
                         scale_offset=None if so is None else so[b:b + 32])
        bp.sync()
        vols[name] = bp.GetVolumeData().astype(np.float64)
        bp.close()
        del store
    proj = raw.cpu().numpy()
    del raw
    g = O.Geom(**{k: getattr(p, k) for k in O.Geom.__dataclass_fields__})
    # FP64 reference: FP64 filter (FP64 H), FP64 matrices, FP64 accumulation
    f64 = O.filter_rows(O.cos_weight(proj, p.cos_table()), p.filter(), O.PAD_MEAN4)
    ref = O.bp(g, f64, p.proj_matrices(), O.BP_GPU_STRICT)
    rows = []
    for name, v in vols.items():
        rel = float(np.linalg.norm(v - ref) / np.linalg.norm(ref))
        mx = float(np.abs(v - ref).max())
        rows.append((name, rel, mx))
        print(f"precision {name:5s}: relative RMSE {rel:.3e}, max abs error {mx:.3e} "
              f"(volume max {np.abs(ref).max():.3f})")
    err = {r[0]: r[1] for r in rows}
    assert err["f32"] < 1e-5
    assert err["f16"] < 2e-4
    assert err["bf16"] < 2e-3
