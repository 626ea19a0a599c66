"""Iterative reconstruction (ratio / MLEM-like and SART loops of
3dct/3dct.cpp:1029-1292) on the GPU projector pair.

* the projector pair against the oracle's forward projector (zero outside,
  t <= tfar) and backprojector (1/w^2, strict bounds);
* two ratio iterations against the same loop run on the oracle's operators;
* SART and the ratio update reduce the projection residual from sweep to sweep.
"""
import numpy as np
import pytest

import fdk_amd as F
import oracle as O
from fdk_amd import iterative as it

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def geom(n=32, views=40):
    return F.Parameter(nx=n, ny=n, nz=n, dx=128.0 / n, dy=128.0 / n, dz=128.0 / n, nu=48, nv=48,
                       du=4.0 * 32 / n, dv=4.0 * 32 / n, nproj=views, volume_position_angle_rad=0.0,
                       filter_type=0)


def phantom(torch, p, seed=7):
    ell = F.random_ellipsoids(6, seed=seed, radius_mm=56, half_height_mm=56, axes_mm=(10, 40),
                              density=(0.2, 1.0))
    vol = torch.empty((p.nz, p.ny, p.nx), device="cuda")
    F.phantom_rasterize(ell, p, vol)
    proj = torch.empty((p.nproj, p.nv, p.nu), device="cuda")
    F.phantom_project(ell, p, 0, p.nproj, proj)
    return vol, proj


def test_projector_pair_matches_oracle(torch_cuda):
    torch = torch_cuda
    p = geom()
    g = O.Geom(**{k: getattr(p, k) for k in O.Geom.__dataclass_fields__})
    vol, proj = phantom(torch, p)
    A = it.Projector(p)
    views = [0, 1, 2, 17, 39]
    fp = torch.empty((len(views), p.nv, p.nu), device="cuda")
    A.forward(vol, views, fp)
    v = vol.cpu().numpy()
    ref = np.stack([O.fp(g, v, s, None, O.FP_CPU_ZERO) for s in views])
    assert rel(fp.cpu().numpy(), ref) < 1e-5
    back = torch.empty_like(vol)
    A.back(proj[views].contiguous(), views, back)
    mats = p.proj_matrices().astype(np.float32).astype(np.float64)[views]
    refb = O.bp(g, proj[views].cpu().numpy(), mats, O.BP_GPU_STRICT)
    assert rel(back.cpu().numpy(), refb) < 2e-5
    A.close()


def test_ratio_iterations_match_oracle_loop(torch_cuda):
    torch = torch_cuda
    p = geom()
    g = O.Geom(**{k: getattr(p, k) for k in O.Geom.__dataclass_fields__})
    _, proj = phantom(torch, p)
    x = it.ratio_reconstruct(p, proj, 2).cpu().numpy()

    P = proj.cpu().numpy().astype(np.float64)
    P = P - min(P.min(), 0.0)
    mask = np.zeros((p.nv, p.nu))
    mask[2:-2, 2:-2] = 1.0
    mats = p.proj_matrices().astype(np.float32).astype(np.float64)

    def B(q):
        return O.bp(g, q.astype(np.float32), mats, O.BP_GPU_STRICT).astype(np.float64)

    def div(a, b):
        return np.where(b != 0, a / np.where(b != 0, b, 1.0), 0.0)

    norm = B(np.broadcast_to(mask, P.shape))
    xr = np.ones((p.nz, p.ny, p.nx))
    for _ in range(2):
        fp = np.stack([O.fp(g, xr.astype(np.float32), s, None, O.FP_CPU_ZERO)
                       for s in range(p.nproj)]).astype(np.float64)
        xr = xr * div(B(div(P, fp) * mask), norm)
    assert rel(x, xr) < 1e-4


@pytest.mark.parametrize("algo", ["ratio", "sart"])
def test_residual_decreases(torch_cuda, algo):
    torch = torch_cuda
    p = geom(32, 48)
    vol, proj = phantom(torch, p, seed=9)
    A = it.Projector(p)
    fp = torch.empty_like(proj)
    res = []

    def residual(i, x, *_):
        A.forward(x, range(p.nproj), fp)
        res.append(float(torch.linalg.norm(fp - proj) / torch.linalg.norm(proj)))

    if algo == "ratio":
        it.ratio_reconstruct(p, proj, 6, callback=residual, projector=A)
    else:
        it.sart_reconstruct(p, proj, 3, callback=residual, projector=A)
    A.close()
    print(algo, ["%.4f" % r for r in res], flush=True)
    assert all(b < a for a, b in zip(res, res[1:])), res
    assert res[-1] < 0.5 * res[0] or res[-1] < 0.05
