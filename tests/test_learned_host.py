"""Host logic of the learned path (no GPU): tiling of a volume into
overlapping 128^3 tiles and the scan geometry."""
import pytest

import fdk_amd.learned as Lr


@pytest.mark.parametrize("n", [128, 200, 256, 512, 1024])
def test_tiles_partition_the_axis(n):
    t = Lr.tiles(n)
    assert t[0][1] == 0 and t[-1][2] == n
    for (s, lo, hi), (s2, lo2, hi2) in zip(t, t[1:]):
        assert hi == lo2                      # written ranges partition [0, n)
        assert s2 - s <= Lr.PATCH - Lr.OVERLAP
    for s, lo, hi in t:
        assert s <= lo < hi <= s + Lr.PATCH and 0 <= s and s + Lr.PATCH <= n
    if n == 1024:
        assert len(t) == 9                    # 9^3 = 729 tiles for the 1024^3 volume


def test_sparse_view_geometry():
    p = Lr.geometry(1024, 1024)
    assert (p.nx, p.nu, p.nproj) == (1024, 1024, 1024)
    assert p.du == pytest.approx(0.4) and p.dx == pytest.approx(0.4 / 1.5)
    q = Lr.geometry(256, 64)
    assert q.nu * q.du == pytest.approx(409.6)


def test_weight_files_declare_their_architecture():
    """Each committed weight file loads without pickle and its arrays fit the
    architecture it declares (space-to-depth edge, activation)."""
    import os

    import numpy as np
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fdk_amd",
                        "weights")
    seen = 0
    for fname in sorted(os.listdir(root)):
        if not fname.endswith(".npz"):
            continue
        with np.load(os.path.join(root, fname), allow_pickle=False) as z:
            f = int(z["arch.s2d"]) if "arch.s2d" in z.files else 1
            act = int(z["arch.act"]) if "arch.act" in z.files else 1
            assert f in (1, 2, 4) and act in (1, 2)
            k = f ** 3
            cin0 = 1 if f == 1 else max(32, k)
            assert z["conv0.w"].shape == ((32, 32) if f == 1 else (27, 32, cin0))
            assert z["head.w"].shape == ((32,) if f == 1 else (k, 32))
            assert z["head.b"].shape == (k,)
            for i in range(10):
                assert np.all(np.isfinite(z[f"conv{i}.w"]))
        seen += 1
    assert seen >= 2


def test_learning_rate_schedule():
    lr = 1e-3
    assert Lr.lr_at(0.0, lr) == pytest.approx(lr)
    assert Lr.lr_at(1.0, lr) == pytest.approx(0.05 * lr)
    assert Lr.lr_at(0.5, lr) == pytest.approx(lr * (0.05 + 0.95 * 0.5))
    # warm-up: starts at lr_final x the cosine value, ramps linearly, then joins the cosine
    assert Lr.lr_at(0.0, lr, warmup=0.1) == pytest.approx(0.05 * lr)
    assert Lr.lr_at(0.05, lr, warmup=0.1) == pytest.approx(Lr.lr_at(0.05, lr) * (0.05 + 0.95 * 0.5))
    assert Lr.lr_at(0.1, lr, warmup=0.1) == pytest.approx(Lr.lr_at(0.1, lr))
    vals = [Lr.lr_at(p / 100, lr, warmup=0.1) for p in range(101)]
    assert all(b >= a for a, b in zip(vals[:10], vals[1:11]))      # rising through the warm-up
    assert all(b <= a for a, b in zip(vals[10:-1], vals[11:]))     # then falling


def test_patch_shapes():
    assert Lr._dhw(128) == (128, 128, 128)
    assert Lr._dhw((256, 256, 512)) == (256, 256, 512)
    assert Lr._dhw([64, 32, 16]) == (64, 32, 16)
