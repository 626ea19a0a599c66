"""On-disk formats (projection folder with step, volume file, per-slice files):
byte layout as the reference writes it (common/type.h:3256-3310,
3845-3858; cuFDK/mpiFDK.cpp:548-575), Python round trips, and the C++ drop-in
header reading and writing the same files."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from fdk_amd import formats as fm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_volume_file_layout_and_round_trip(tmp_path):
    vol = np.arange(4 * 3 * 2, dtype=np.float32).reshape(4, 3, 2) * 0.25  # [z][y][x]
    p = str(tmp_path / "v.raw")
    fm.save_volume(p, vol)
    raw = open(p, "rb").read()
    assert len(raw) == 12 + vol.size * 4
    assert np.array_equal(np.frombuffer(raw[:12], "<i4"), [2, 3, 4])  # nx, ny, nz
    assert np.array_equal(np.frombuffer(raw[12:], "<f4"), vol.ravel())  # x fastest, then y, z
    assert np.array_equal(fm.load_volume(p), vol)
    open(p, "ab").write(b"\0\0\0\0")
    with pytest.raises(ValueError):
        fm.load_volume(p)


def test_projection_folder_step_and_start(tmp_path):
    rng = np.random.default_rng(0)
    views = rng.standard_normal((10, 5, 6)).astype(np.float32)  # [count][nv][nu]
    d = str(tmp_path / "p")
    fm.save_projection_folder(d, views)
    assert sorted(os.listdir(d))[:2] == ["img0000.raw", "img0001.raw"]
    assert os.path.getsize(os.path.join(d, "img0003.raw")) == 5 * 6 * 4
    assert np.array_equal(np.fromfile(os.path.join(d, "img0003.raw"), np.float32),
                          views[3].ravel())  # u fastest
    got = fm.load_projection_folder(d, 6, 5, 7, step=3)
    for k in range(7):
        assert np.array_equal(got[k], views[k] if k % 3 == 0 else 0 * views[k])
    got = fm.load_projection_folder(d, 6, 5, 4, step=2, start_index=1)
    assert np.array_equal(got[0], views[1]) and np.array_equal(got[2], views[3])
    with pytest.raises(ValueError):
        fm.load_projection_folder(d, 6, 4, 4)
    with pytest.raises(FileNotFoundError):
        fm.load_projection_folder(d, 6, 5, 12)


def test_slices(tmp_path):
    vol = np.arange(24, dtype=np.float32).reshape(4, 3, 2)
    d = str(tmp_path / "s")
    fm.save_slices(d, vol, z0=100)
    assert sorted(os.listdir(d)) == ["000100.raw", "000101.raw", "000102.raw", "000103.raw"]
    assert np.array_equal(np.fromfile(os.path.join(d, "000102.raw"), np.float32), vol[2].ravel())


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_header_reads_and_writes_the_same_files(tmp_path):
    import fdk_amd as F
    exe = str(tmp_path / "formats_check")
    libdir = os.path.dirname(F.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "formats_check.cpp"), "-o", exe, "-L", libdir,
                    "-lfdk_amd", f"-Wl,-rpath,{libdir}"], check=True)
    views = (np.arange(7 * 30, dtype=np.float32) * 0.5).reshape(7, 5, 6)
    fm.save_projection_folder(str(tmp_path / "proj"), views)
    vol = np.linspace(-1, 1, 2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    fm.save_volume(str(tmp_path / "vol.raw"), vol)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    lines = dict((ln.split()[0], ln.split()[1:]) for ln in r.stdout.splitlines() if ln.strip())
    loaded = sum(float(views[k].sum()) for k in (0, 3, 6))
    assert float(lines["proj_sum"][0]) == pytest.approx(loaded, rel=1e-9)
    assert float(lines["proj_sum"][2]) == views[3, 0, 0]
    assert float(lines["proj_sum"][4]) == 0.0  # frame 1 is not loaded with step 3
    assert lines["vol"][:3] == ["4", "3", "2"]
    assert float(lines["vol"][4]) == pytest.approx(float(vol.sum()), abs=1e-5)
    w = fm.load_volume(str(tmp_path / "vol_cpp.raw"))
    assert w.shape == (4, 2, 3) and np.array_equal(w.ravel(), 0.5 * np.arange(24, dtype=np.float32))
    s = np.fromfile(str(tmp_path / "slices_cpp" / "000011.raw"), np.float32)
    assert np.array_equal(s, w[1].ravel())
    back = fm.load_projection_folder(str(tmp_path / "proj_cpp"), 6, 5, 7, step=3)
    assert np.array_equal(back[3], views[3]) and np.array_equal(back[6], views[6])
    # the C++ CheckImageQuality / NCC agree with the Python restatement
    from fdk_amd.quality import check_image_quality, ncc
    wv = 0.5 * np.arange(24, dtype=np.float32)
    q = check_image_quality(wv, wv + np.float32(0.05))
    got = [float(t) for t in lines["quality"][:4]]
    assert got == pytest.approx([q["error_per_pixel"], q["mse"], q["psnr_db"], q["qi"]], rel=1e-6)
    assert int(lines["quality"][4]) == int(q["passed"])
    assert float(lines["ncc"][0]) == pytest.approx(ncc(wv, wv + np.float32(0.05)), abs=1e-9)
