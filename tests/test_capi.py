"""The C-ABI library without a GPU: it loads, exports every entry point that
include/fdk_amd.h declares, computes the host geometry tables like the oracle,
and rejects invalid arguments with a status and a message before touching a
device."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import fdk_amd as F
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fdk_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(F.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every binding the Python layer declares is part of the header
    assert set(F.exported_symbols()) <= set(names)
    import fdk_amd.unet as U
    U.L()  # binds the fdk_nn_* signatures without touching a device


def test_device_count_without_gpu_is_not_an_error():
    assert F.lib().fdk_device_count() >= 0
    assert F.lib().fdk_version() >= 100


GEOMS = [dict(nx=16, ny=16, nz=16, nu=24, nv=20, nproj=8, volume_position_angle_rad=-0.3),
         dict(nx=64, ny=48, nz=40, nu=100, nv=90, du=0.7, dv=0.9, offset_nu=1.5, offset_nv=-2.0,
              volume_position_angle_rad=-0.85289, nproj=33, scan_angle_deg=200.0, direction=1,
              filter_type=2),
         dict(nx=256, ny=256, nz=256, nu=256, nv=256, nproj=360)]


@pytest.mark.parametrize("kw", GEOMS)
def test_geometry_tables_match_oracle(kw):
    p = F.Parameter(**kw)
    g = O.Geom(**{k: getattr(p, k) for k in O.Geom.__dataclass_fields__})
    assert p.angular_ramp_factor() == O.angular_ramp_factor(g)
    assert np.array_equal(p.cos_table(), O.cos_table(g))
    assert np.array_equal(p.filter(), O.filter_design(p.filter_type, p.nu))
    m, mo = p.proj_matrices(), O.proj_matrices(g)
    assert np.abs(m - mo).max() <= 1e-12 * np.abs(mo).max()
    assert p.proj_matrices(5, 3).shape == (3, 12)
    assert np.array_equal(p.proj_matrices(5, 3), m[5:8])
    assert p.view_angle(7) == O.view_angle(g, 7)


def test_geometry_matches_reference_fixture():
    z = np.load(os.path.join(ROOT, "tests", "golden", "tables_256.npz"), allow_pickle=False)
    import json
    p = F.Parameter(**json.loads(str(z["geom"])))
    assert np.array_equal(p.cos_table(), z["cos"])
    assert np.array_equal(p.filter(), z["filter"])
    assert np.abs(p.proj_matrices() - z["mats"]).max() <= 2e-7 * np.abs(z["mats"]).max()


def _status(rc):
    return rc, F.last_error()


def test_invalid_geometry_is_rejected():
    g = F.Parameter(nu=0).c()
    out = np.zeros(4)
    rc, msg = _status(F.lib().fdk_geom_filter(C.byref(g), out.ctypes.data_as(C.c_void_p), 4))
    assert rc == -1 and "positive" in msg
    g = F.Parameter(nu=256, filter_type=9).c()
    rc, msg = _status(F.lib().fdk_geom_filter(C.byref(g), out.ctypes.data_as(C.c_void_p), 512))
    assert rc == -1 and "filter_type" in msg
    g = F.Parameter(nu=256).c()
    rc, msg = _status(F.lib().fdk_geom_filter(C.byref(g), out.ctypes.data_as(C.c_void_p), 4))
    assert rc == -1 and "512" in msg


def test_bp_create_rejects_bad_boxes_before_allocating():
    h = C.c_void_p()
    L = F.lib()
    top = (C.c_int64 * 6)(0, 0, 0, 64, 64, 32)
    bad_bottom = (C.c_int64 * 6)(0, 0, 30, 64, 64, 32)   # does not mirror the top slab
    rc, msg = _status(L.fdk_bp_create(64, 64, 32, 64, 64, 64, top, bad_bottom, C.byref(h)))
    assert rc == -1 and "mirror" in msg and not h.value
    too_big = (C.c_int64 * 6)(0, 0, 0, 65, 64, 32)
    rc, msg = _status(L.fdk_bp_create(64, 64, 32, 64, 64, 64, too_big, None, C.byref(h)))
    assert rc == -1 and "exceeds" in msg
    rc, msg = _status(L.fdk_bp_create(1, 64, 32, 64, 64, 64, top, None, C.byref(h)))
    assert rc == -1
    with pytest.raises(F.FdkError, match="mirror"):
        F.BackProjectionFDK(64, 64, 32, 64, 64, 64, (0, 0, 0, 64, 64, 32), (0, 0, 30, 64, 64, 32))


def test_filter_rejects_bad_lengths_and_pads():
    L = F.lib()
    fake = C.c_void_p(16)  # never dereferenced: validation fails first
    rc, msg = _status(L.fdk_filter_rows(fake, fake, 256, 4, 1, 256, 256, fake, 500, 4, None))
    assert rc == -1 and "power of two" in msg
    rc, msg = _status(L.fdk_filter_rows(fake, fake, 256, 4, 1, 256, 256, fake, 512, 5, None))
    assert rc == -1 and "pad_mode" in msg
    rc, msg = _status(L.fdk_filter_rows(fake, fake, 600, 4, 1, 600, 600, fake, 512, 4, None))
    assert rc == -1 and "row length" in msg
    rc, msg = _status(L.fdk_filter_views(fake, 100, fake, 100, 7, 10, 10, 1, None, fake, 16, 4, 1,
                                         None, None))
    assert rc == -1 and "dtype" in msg


def test_forward_projector_rejects_bad_geometry():
    g = F.Parameter(sid=500.0, sod=1000.0).c()
    fake = C.c_void_p(16)
    rc, msg = _status(F.lib().fdk_fp_project(C.byref(g), fake, 0, 1, 0.5, 0, fake, None))
    assert rc == -1 and "sod" in msg


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_random_ellipsoids_are_seeded_and_inside():
    a = F.random_ellipsoids(10, seed=1, radius_mm=120, half_height_mm=120)
    b = F.random_ellipsoids(10, seed=1, radius_mm=120, half_height_mm=120)
    assert [e["center"] for e in a] == [e["center"] for e in b]
    for e in a:
        r = max(e["axes"])
        c = np.array(e["center"])
        assert np.hypot(c[0], c[1]) + r <= 120 + 1e-9
        assert abs(c[2]) + r <= 120 + 1e-9
        R = np.asarray(e["rot"])
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert -0.5 <= e["density"] <= 1.0


def test_pipeline_rejects_bad_arguments_without_a_device():
    L = F.lib()
    g = F.Parameter(nx=16, ny=16, nz=16, nu=24, nv=20, nproj=8).c()
    h = C.c_void_p()
    top = (C.c_int64 * 6)(0, 0, 0, 16, 16, 16)
    rc, msg = _status(L.fdk_pipe_create(C.byref(g), top, None, 7, F.PAD_MEAN4, 2, C.byref(h)))
    assert rc == -1 and "store dtype" in msg and not h.value
    rc, msg = _status(L.fdk_pipe_create(C.byref(g), top, None, F.FDK_F32, F.PAD_MEAN4, 9, C.byref(h)))
    assert rc == -1 and "slots" in msg and not h.value
    rc, msg = _status(L.fdk_pipe_push_raw(None, None, 4, 0, 1))
    assert rc == -1 and "null" in msg
    rc, msg = _status(L.fdk_pipe_finish(None))
    assert rc == -1
