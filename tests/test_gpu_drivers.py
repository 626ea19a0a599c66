"""The C++ drop-in layer (include/fdk_amd.hpp) driven the way the reference's
own drivers drive it, checked against the oracle on the projections the driver
wrote."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fdk_amd", "bin", "fdk_drivers")


def geom(n, views):
    du = 409.6 / n
    dx = du * 1000.0 / 1500.0
    return O.Geom(nx=n, ny=n, nz=n, dx=dx, dy=dx, dz=dx, nu=n, nv=n, du=du, dv=du, nproj=views,
                  volume_position_angle_rad=0.0, filter_type=5)


def run(mode, n, views, prefix):
    assert os.path.exists(DRIVER), "C++ driver not built"
    r = subprocess.run([DRIVER, mode, str(n), str(views), prefix], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr


def oracle_fdk(g, proj):
    H = O.filter_design(5, g.nu).astype(np.float32).astype(np.float64)
    f = O.filter_rows(O.cos_weight(proj, O.cos_table(g)), H, O.PAD_MEAN4)
    m = O.proj_matrices(g).astype(np.float32).astype(np.float64)
    return O.bp(g, f, m, O.BP_GPU_STRICT)


@pytest.mark.parametrize("mode", ["fdk", "rtk"])
def test_cpp_fdk_driver_matches_oracle(tmp_path, mode):
    n, views = 64, 72
    prefix = str(tmp_path / mode)
    run(mode, n, views, prefix)
    proj = np.fromfile(prefix + ".proj.raw", np.float32).reshape(views, n, n)
    vol = np.fromfile(prefix + ".vol.raw", np.float32).reshape(n, n, n)
    ref = oracle_fdk(geom(n, views), proj)
    err = float(np.linalg.norm(vol - ref) / np.linalg.norm(ref))
    assert err < 1e-5, err


def test_cpp_forward_projection_driver_matches_oracle(tmp_path):
    n, views = 48, 6
    prefix = str(tmp_path / "fp")
    run("fp", n, views, prefix)
    vol = np.fromfile(prefix + ".vol.raw", np.float32).reshape(n, n, n)
    fp = np.fromfile(prefix + ".fp.raw", np.float32).reshape(views, n, n)
    g = geom(n, views)
    # the reference signature marches at min(dx, dy, dz) (genProj's projector step)
    step = min(g.dx, g.dy, g.dz)
    ref = np.stack([O.fp(g, vol, s, step, O.FP_GPU_CLAMP) for s in range(views)])
    close = np.abs(fp - ref) <= 1e-4 * np.abs(ref).max()
    assert close.mean() > 0.995


def test_cpp_filter_proj_image_driver_matches_oracle(tmp_path):
    """FilterlibCufft's test_filter call forms: LoadFolder, FilterProjImage
    (which returns the summed milliseconds), SaveToFolder."""
    n, views = 64, 40
    prefix = str(tmp_path / "filter")
    run("filter", n, views, prefix)
    g = geom(n, views)
    raw = np.stack([np.fromfile(f"{prefix}_raw/img{i:04d}.raw", np.float32).reshape(n, n)
                    for i in range(views)])
    flt = np.stack([np.fromfile(f"{prefix}_flt/img{i:04d}.raw", np.float32).reshape(n, n)
                    for i in range(views)])
    H = O.filter_design(5, g.nu).astype(np.float32).astype(np.float64)
    ref = O.filter_rows(O.cos_weight(raw, O.cos_table(g)), H, O.PAD_MEAN3)
    err = float(np.linalg.norm(flt - ref) / np.linalg.norm(ref))
    assert err < 2e-6, err
