"""The native U-Net runner's weight-file reader (fdk_unet_inspect), on the CPU:
the architecture of each committed weight file as the library reads it from the
compressed .npz, against numpy's reading of the same file, and the errors for
files that are not weight sets."""
import os

import numpy as np
import pytest

import fdk_amd as F
import fdk_amd.unet as U

WEIGHTS = os.path.join(os.path.dirname(F.__file__), "weights")


@pytest.mark.parametrize("fname", ["unet_sparse1024_s2d4.npz", "unet_sparse1024_s2d2.npz",
                                   "unet_sparse1024.npz"])
def test_inspect_committed_weights(fname):
    path = os.path.join(WEIGHTS, fname)
    with np.load(path, allow_pickle=False) as z:
        want = (int(z["conv1.w"].shape[1]),
                int(z["arch.s2d"]) if "arch.s2d" in z.files else 1,
                int(z["arch.act"]) if "arch.act" in z.files else 1)
    assert U.inspect(path) == want


def test_inspect_stored_and_compressed_archives(tmp_path):
    rng = np.random.default_rng(0)
    d = {"conv1.w": rng.standard_normal((27, 16, 16)).astype(np.float32),
         "arch.s2d": np.array(1, np.int32), "arch.act": np.array(2, np.int32)}
    np.savez(tmp_path / "stored.npz", **d)
    np.savez_compressed(tmp_path / "deflated.npz", **d)
    assert U.inspect(tmp_path / "stored.npz") == (16, 1, 2)
    assert U.inspect(tmp_path / "deflated.npz") == (16, 1, 2)


def test_inspect_rejects_other_files(tmp_path):
    (tmp_path / "text.npz").write_text("not a zip archive")
    with pytest.raises(F.FdkError):
        U.inspect(tmp_path / "text.npz")
    with pytest.raises(F.FdkError):
        U.inspect(tmp_path / "missing.npz")
    np.savez(tmp_path / "other.npz", a=np.zeros(3, np.float32))
    with pytest.raises(F.FdkError, match="conv1.w"):
        U.inspect(tmp_path / "other.npz")
    np.savez(tmp_path / "bad_s2d.npz", **{"conv1.w": np.zeros((27, 32, 32), np.float32),
                                         "arch.s2d": np.array(3, np.int32)})
    with pytest.raises(F.FdkError, match="s2d"):
        U.inspect(tmp_path / "bad_s2d.npz")
