"""Streaming pipeline (fdk_pipe_*): host batches copied on one stream while the
previous batch is filtered and backprojected on another.

* raw views pushed in 32-view batches with a ragged last batch, and a sparse
  scan (every 3rd view, the reference's ``step``), against the oracle FDK of
  the same views, to the 1e-4 relative-RMSE bar;
* FP16 filtered views pushed from page-locked memory give exactly the volume
  of the same views backprojected from device memory;
* argument errors are rejected before any device work.
"""
import numpy as np
import pytest

import fdk_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def rel_rmse(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def geom(n=64, views=72):
    return F.Parameter(nx=n, ny=n, nz=n, dx=256.0 / n, dy=256.0 / n, dz=256.0 / n, nu=n, nv=n,
                       du=384.0 / n, dv=384.0 / n, nproj=views, volume_position_angle_rad=0.0,
                       filter_type=5)


def oracle_fdk(p, proj, views):
    g = O.Geom(**{k: getattr(p, k) for k in O.Geom.__dataclass_fields__})
    H = p.filter().astype(np.float32).astype(np.float64)
    f = O.filter_rows(O.cos_weight(proj, p.cos_table()), H, O.PAD_MEAN4)
    mats = p.proj_matrices().astype(np.float32).astype(np.float64)[views]
    return O.bp(g, f, mats, O.BP_GPU_STRICT)


@pytest.mark.parametrize("step", [1, 3])
def test_raw_pipeline_matches_oracle(torch_cuda, step):
    torch = torch_cuda
    p = geom()
    n = p.nx
    print(f"pipeline raw step={step} {n}^3 {p.nproj} views", flush=True)
    ell = F.random_ellipsoids(10, seed=4, radius_mm=110, half_height_mm=110, axes_mm=(20, 100))
    raw = torch.empty((p.nproj, p.nv, p.nu), device="cuda")
    F.phantom_project(ell, p, 0, p.nproj, raw)
    host = raw.cpu().pin_memory()
    views = list(range(0, p.nproj, step))
    half = n // 2
    pipe = F.Pipeline(p, (0, 0, 0, n, n, half), (0, 0, half, n, n, half))
    for b in range(0, len(views), 32):
        cnt = min(32, len(views) - b)
        batch = host[views[b]:views[b] + cnt * step:step].contiguous() if step > 1 else \
            host[views[b]:views[b] + cnt]
        pipe.push_raw(batch, cnt, views[b], step)
    pipe.finish()
    vol = pipe.bp.GetVolumeData()
    st = pipe.stats()
    pipe.close()
    assert st["views"] == len(views) and st["bp_ms"] > 0 and st["filter_ms"] > 0
    assert st["h2d_bytes"] == len(views) * p.nu * p.nv * 4
    ref = oracle_fdk(p, raw.cpu().numpy()[views], views)
    err = rel_rmse(vol, ref)
    print(f"  rel RMSE vs oracle {err:.2e}", flush=True)
    assert err < 1e-4


def test_filtered_fp16_pipeline_equals_device_path(torch_cuda):
    torch = torch_cuda
    p = geom(128, 64)
    n = p.nx
    print(f"pipeline fp16 {n}^3 {p.nproj} views", flush=True)
    ell = F.random_ellipsoids(12, seed=5, radius_mm=110, half_height_mm=110, axes_mm=(20, 100))
    raw = torch.empty((p.nproj, p.nv, p.nu), device="cuda")
    F.phantom_project(ell, p, 0, p.nproj, raw)
    cos = torch.as_tensor(p.cos_table()).cuda()
    H = torch.as_tensor(p.filter().astype(np.float32)).cuda()
    store = torch.empty((p.nproj, p.nu, p.nv), dtype=torch.int16, device="cuda")
    so = F.filter_views(raw, store, cos, H, F.PAD_MEAN4, True, dtype=F.FDK_F16)
    mats = p.proj_matrices().astype(np.float32)
    half = n // 2
    boxes = ((0, 0, 0, n, n, half), (0, 0, half, n, n, half))
    bp = F.BackProjectionFDK(p.nv, p.nu, 32, n, n, n, *boxes)
    bp.set_stream(torch.cuda.current_stream().cuda_stream)
    for b in range(0, p.nproj, 32):
        bp.bp_device(store, p.nv, p.nu, mats[b:b + 32], 32, dtype=F.FDK_F16, offset_views=b,
                     scale_offset=so[b:b + 32])
    bp.sync()
    ref = bp.GetVolumeData()
    bp.close()
    host = store.cpu().pin_memory()
    pipe = F.Pipeline(p, *boxes, store_dtype=F.FDK_F16, slots=3)
    for b in range(0, p.nproj, 32):
        pipe.push_filtered(host[b:b + 32], 32, b, 1, so[b:b + 32])
    pipe.finish()
    vol = pipe.bp.GetVolumeData()
    pipe.close()
    assert np.array_equal(vol, ref)


def test_pipeline_rejects_bad_pushes(torch_cuda):
    p = geom(32, 40)
    pipe = F.Pipeline(p, store_dtype=F.FDK_BF16)
    x = np.zeros((40, p.nv, p.nu), np.float32)
    with pytest.raises(F.FdkError, match="outside 1..32"):
        pipe.push_raw(x, 33, 0)
    with pytest.raises(F.FdkError, match="outside"):
        pipe.push_raw(x, 8, 36)
    with pytest.raises(F.FdkError, match="scale_offset"):
        pipe.push_filtered(np.zeros((4, p.nu, p.nv), np.int16), 4, 0)
    with pytest.raises(F.FdkError, match="slots"):
        F.Pipeline(p, slots=7)
    pipe.push_raw(x[:8], 8, 32)  # a good push still works after rejected ones
    pipe.finish()
    assert pipe.stats()["views"] == 8
    pipe.close()
