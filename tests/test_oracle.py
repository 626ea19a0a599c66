"""Pins the oracle (oracle/fdk_oracle.c) without a GPU: against the golden
fixtures written from the reference's own CPU code (tests/golden), against
closed forms, and against invariants chosen so that a plausible mistake (a
dropped factor, a transposed operand, a wrong sign or padding) fails."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLD, name), allow_pickle=False)
    geom = json.loads(str(z["geom"])) if "geom" in z.files else None
    return z, geom


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


# ------------------------------------------------------------ geometry tables
@pytest.mark.parametrize("name", ["tables_tiny.npz", "tables_256.npz", "tables_odd.npz"])
def test_tables_match_reference(name):
    z, geom = load(name)
    g = O.Geom(**geom)
    assert O.angular_ramp_factor(g) == float(z["af"])
    assert np.array_equal(O.cos_table(g), z["cos"])          # float-exact restatement
    assert np.array_equal(O.filter_design(g.filter_type, g.nu), z["filter"])
    m = O.proj_matrices(g)
    # the reference multiplies six float matrices; ours is FP64
    assert np.abs(m - z["mats"]).max() <= 2e-7 * np.abs(m).max()


def test_filter_designs_match_reference():
    z, _ = load("filter_designs.npz")
    for key in z.files:
        t, n = key[1:].split("_n")
        assert np.array_equal(O.filter_design(int(t), int(n)), z[key]), key


def test_nextpow2_is_strictly_greater():
    assert [O.nextpow2(n) for n in (1, 2, 3, 255, 256, 257, 2048, 4096)] == \
        [2, 4, 4, 256, 512, 512, 4096, 8192]


def test_filter_closed_forms():
    N = 512
    ramlak = O.filter_design(1, 256)
    k = np.arange(N // 2 + 1)
    assert np.allclose(ramlak[: N // 2 + 1], 2 * k / N, rtol=0, atol=1e-15)
    hann = O.filter_design(5, 256)
    assert np.allclose(hann[: N // 2 + 1], 2 * k / N * 0.5 * (1 + np.cos(2 * np.pi * k / N)))
    assert hann[N // 2] == pytest.approx(0.0, abs=1e-15)
    for t in range(6):
        h = O.filter_design(t, 256)
        assert np.array_equal(h[1:], h[1:][::-1])   # even: H[N-k] = H[k]
        assert h[0] == 0.0                           # ramp kills DC


def test_angular_ramp_factor_closed_form():
    g = O.Geom(nu=256, du=0.37, sod=812.0, sid=1234.0, nproj=999)
    af = math.pi * g.sid / (2 * g.du * g.nproj * g.sod)
    assert O.angular_ramp_factor(g) == pytest.approx(af, rel=1e-14)


def test_cos_table_closed_form():
    g = O.Geom(nu=33, nv=21, du=0.5, dv=0.7, offset_nu=0.0, sod=900.0, sid=1300.0, nproj=100)
    c = O.cos_table(g).astype(np.float64)
    u = (np.arange(33) - 16) * 0.5
    v = (10 - np.arange(21)) * 0.7
    w = g.sid / np.sqrt(u[None, :] ** 2 + v[:, None] ** 2 + g.sid ** 2) * g.sod ** 2 * \
        O.angular_ramp_factor(g)
    assert np.allclose(c, w, rtol=2e-7, atol=0)
    assert c[10, 16] == pytest.approx(g.sod ** 2 * O.angular_ramp_factor(g), rel=1e-7)


@pytest.mark.parametrize("vpa,direction", [(0.0, -1), (-0.8528925, -1), (0.4, 1)])
def test_projection_matrix_closed_forms(vpa, direction):
    g = O.Geom(nx=31, ny=27, nz=23, dx=0.9, dy=1.1, dz=1.3, nu=41, nv=37, du=1.7, dv=1.5,
               nproj=17, volume_position_angle_rad=vpa, direction=direction)
    cx, cy, cz = 15.0, 13.0, 11.0
    for s in range(g.nproj):
        m = O.proj_matrices(g)[s].reshape(3, 4)
        p = m @ np.array([cx, cy, cz, 1.0])
        # the isocentre lands on the detector centre at depth SOD
        assert p[2] == pytest.approx(g.sod)
        assert p[0] / p[2] == pytest.approx((g.nu - 1) / 2)
        assert p[1] / p[2] == pytest.approx((g.nv - 1) / 2)
        # rows 0 and 2 do not depend on z; v is mirror symmetric about cz
        assert m[0, 2] == 0.0 and m[2, 2] == 0.0
        for (i, j, k) in [(3, 5, 2), (20, 9, 7), (30, 26, 0)]:
            a = m @ np.array([i, j, k, 1.0])
            b = m @ np.array([i, j, g.nz - 1 - k, 1.0])
            assert a[1] / a[2] + b[1] / b[2] == pytest.approx(g.nv - 1, abs=1e-9)
        # a voxel offset along x by a voxels at theta: projected offset is SID/SOD-scaled
        th = O.view_angle(g, s)
        x_mm, y_mm = 4 * g.dx, 0.0
        xr, yr = x_mm * math.cos(th) - y_mm * math.sin(th), x_mm * math.sin(th) + y_mm * math.cos(th)
        q = m @ np.array([cx + 4, cy, cz, 1.0])
        assert q[2] == pytest.approx(g.sod + yr)
        assert q[0] / q[2] - (g.nu - 1) / 2 == pytest.approx(g.sid * xr / (g.du * (g.sod + yr)))


# ------------------------------------------------------------------------ FFT
def test_fft_matches_reference_and_numpy():
    z, _ = load("fft.npz")
    for n in (64, 512):
        x = z[f"x{n}"]
        assert np.abs(O.fft(x) - z[f"fwd{n}"]).max() < 1e-11
        assert np.abs(O.fft(x, True) - z[f"inv{n}"]).max() < 1e-13
        assert np.abs(O.fft(x) - np.fft.fft(x)).max() < 1e-11
        assert np.abs(O.fft(x, True) - np.fft.ifft(x)).max() < 1e-13
        assert np.sum(np.abs(O.fft(x)) ** 2) == pytest.approx(n * np.sum(np.abs(x) ** 2))


# ----------------------------------------------------------------- filtering
def test_filter_rows_mean3_matches_reference():
    z, _ = load("filter_rows_mean3.npz")
    assert np.abs(O.filter_rows(z["img"], z["H24"], O.PAD_MEAN3) - z["out24"]).max() <= \
        1e-6 * np.abs(z["out24"]).max()
    assert np.abs(O.filter_rows(z["rows"], z["H256"], O.PAD_MEAN3) - z["out256"]).max() <= \
        1e-6 * np.abs(z["out256"]).max()


def test_filter_rows_delta_response_is_idft_of_H():
    nu, N = 200, 256
    H = O.filter_design(5, nu)
    h = np.real(np.fft.ifft(H))
    off = (N - nu) // 2
    for p in (0, 57, 199):
        row = np.zeros((1, nu), np.float32)
        row[0, p] = 1.0
        out = O.filter_rows(row, H, O.PAD_ZERO)[0]
        expect = np.array([h[(off + i - (off + p)) % N] for i in range(nu)])
        assert np.abs(out - expect).max() < 1e-7


def test_filter_rows_pad_modes():
    nu = 100
    H = O.filter_design(1, nu)
    const = np.full((2, nu), 3.5, np.float32)
    # mean padding of a constant row is the constant: the ramp (H[0] = 0) returns 0
    assert np.abs(O.filter_rows(const, H, O.PAD_MEAN4)).max() < 1e-6
    assert np.abs(O.filter_rows(const, H, O.PAD_MEAN3)).max() < 1e-6
    assert np.abs(O.filter_rows(const, H, O.PAD_ZERO)).max() > 0.1
    # mean4 and mean3 differ exactly through the pad values
    rng = np.random.default_rng(1)
    r = rng.random((1, nu)).astype(np.float32)
    a = O.filter_rows(r, H, O.PAD_MEAN4)
    b = O.filter_rows(r, H, O.PAD_MEAN3)
    assert np.abs(a - b).max() > 1e-6


def test_filter_rows_linearity():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((4, 300)).astype(np.float32)
    b = rng.standard_normal((4, 300)).astype(np.float32)
    H = O.filter_design(5, 300)
    lhs = O.filter_rows((2 * a + b).astype(np.float32), H, O.PAD_MEAN4)
    rhs = 2 * O.filter_rows(a, H, O.PAD_MEAN4) + O.filter_rows(b, H, O.PAD_MEAN4)
    assert np.abs(lhs - rhs).max() < 1e-5


# ------------------------------------------------------------ backprojection
@pytest.mark.parametrize("name", ["bp_3dct_tiny.npz", "bp_3dct_odd.npz"])
def test_bp_cpu_trunc_matches_reference(name):
    z, geom = load(name)
    g = O.Geom(**geom)
    tables = {"bp_3dct_tiny.npz": "tables_tiny.npz", "bp_3dct_odd.npz": "tables_odd.npz"}[name]
    m = load(tables)[0]["mats"]  # the reference's float matrices for this geometry
    vol = O.bp(g, z["proj"], m.astype(np.float64), O.BP_CPU_TRUNC)
    # the reference accumulates in float with float matrices
    assert rel(vol, z["vol_prefiltered"]) < 2e-5


def test_full_3dct_fdk_matches_reference():
    """The 3dct CPU FDK: cos table times AngularAndrampFactor (applied a second
    time, 3dct.cpp:39-43), 3-sample mean padding in FP64, truncating bilinear."""
    z, geom = load("bp_3dct_tiny.npz")
    g = O.Geom(**geom)
    af = O.angular_ramp_factor(g)
    w = (O.cos_weight(z["raw"], O.cos_table(g)).astype(np.float64) * af).astype(np.float32)
    f = O.filter_rows(w, O.filter_design(g.filter_type, g.nu), O.PAD_MEAN3)
    m = O.proj_matrices(g).astype(np.float32).astype(np.float64)
    vol = O.bp(g, f, m, O.BP_CPU_TRUNC)
    assert rel(vol, z["vol_fdk"]) < 2e-5


def _bp_bruteforce(g, proj, mats, ijk):
    """Independent pure-Python restatement of the strict GPU sample."""
    out = []
    for (i, j, k) in ijk:
        acc = 0.0
        for s in range(mats.shape[0]):
            m = mats[s].reshape(3, 4)
            uw, vw, w = m @ np.array([i, j, k, 1.0])
            u, v = uw / w, vw / w
            if 0 <= u < g.nu - 1 and 0 <= v < g.nv - 1:
                iu, iv = math.floor(u), math.floor(v)
                fu, fv = u - iu, v - iv
                P = proj[s].astype(np.float64)
                val = ((1 - fv) * ((1 - fu) * P[iv, iu] + fu * P[iv, iu + 1])
                       + fv * ((1 - fu) * P[iv + 1, iu] + fu * P[iv + 1, iu + 1]))
                acc += val / (w * w)
        out.append(acc)
    return np.array(out)


def test_bp_gpu_strict_bruteforce():
    g = O.Geom(nx=12, ny=10, nz=9, dx=3.0, dy=3.0, dz=3.0, nu=17, nv=13, du=2.0, dv=2.5,
               nproj=5, volume_position_angle_rad=0.3)
    rng = np.random.default_rng(3)
    proj = rng.standard_normal((5, 13, 17)).astype(np.float32)
    mats = O.proj_matrices(g)
    ijk = [(0, 0, 0), (11, 9, 8), (5, 4, 4), (2, 7, 1), (9, 1, 6), (6, 5, 8)]
    want = _bp_bruteforce(g, proj, mats, ijk)
    got = O.bp_voxels(g, proj, mats, np.array(ijk), O.BP_GPU_STRICT)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15)
    vol = O.bp(g, proj, mats, O.BP_GPU_STRICT)
    assert np.allclose([vol[k, j, i] for (i, j, k) in ijk], want, rtol=1e-12, atol=1e-15)


def test_bp_constant_projection_closed_form():
    g = O.Geom(nx=10, ny=10, nz=10, dx=2.0, dy=2.0, dz=2.0, nu=32, nv=32, du=1.6, dv=1.6, nproj=7)
    proj = np.full((7, 32, 32), 2.5, np.float32)
    mats = O.proj_matrices(g)
    vol = O.bp(g, proj, mats, O.BP_GPU_STRICT)
    for (i, j, k) in [(0, 0, 0), (4, 5, 6), (9, 9, 9)]:
        want = sum(2.5 / (mats[s].reshape(3, 4) @ np.array([i, j, k, 1.0]))[2] ** 2
                   for s in range(7))
        assert vol[k, j, i] == pytest.approx(want, rel=1e-12)


def test_bp_mirror_symmetry():
    g = O.Geom(nx=9, ny=8, nz=10, dx=2.0, dy=2.0, dz=2.0, nu=20, nv=18, du=1.6, dv=1.6, nproj=6)
    rng = np.random.default_rng(4)
    proj = rng.standard_normal((6, 18, 20)).astype(np.float32)
    flipped = np.ascontiguousarray(proj[:, ::-1, :])   # v -> nv-1-v
    mats = O.proj_matrices(g)
    a = O.bp(g, proj, mats, O.BP_GPU_STRICT)
    b = O.bp(g, flipped, mats, O.BP_GPU_STRICT)
    assert np.allclose(a, b[::-1], rtol=1e-9, atol=1e-12)


def test_bp_linearity_and_batch_split():
    g = O.Geom(nx=8, ny=8, nz=8, dx=2.5, dy=2.5, dz=2.5, nu=20, nv=20, nproj=10)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((10, 20, 20)).astype(np.float32)
    b = rng.standard_normal((10, 20, 20)).astype(np.float32)
    mats = O.proj_matrices(g)
    lhs = O.bp(g, (a + 3 * b).astype(np.float32), mats)
    rhs = O.bp(g, a, mats) + 3 * O.bp(g, b, mats)
    assert np.allclose(lhs, rhs, rtol=1e-5, atol=1e-9)
    split = O.bp(g, a[:4], mats[:4]) + O.bp(g, a[4:], mats[4:])
    assert np.allclose(split, O.bp(g, a, mats), rtol=1e-12, atol=1e-15)


# -------------------------------------------------------- forward projection
def test_fp_cpu_zero_matches_reference():
    z, geom = load("fp_3dct_tiny.npz")
    g = O.Geom(**geom)
    for s in range(3):
        got = O.fp(g, z["vol"], s, mode=O.FP_CPU_ZERO)
        # float vs double marching: a ray can gain or lose its last sample
        close = np.abs(got - z["views"][s]) <= 1e-4 * np.abs(z["views"][s]).max()
        assert close.mean() > 0.99
        assert rel(got, z["views"][s]) < 1e-3


def test_fp_sphere_against_analytic_chords():
    g = O.Geom(nx=64, ny=64, nz=64, dx=1.0, dy=1.0, dz=1.0, nu=48, nv=48, du=2.0, dv=2.0, nproj=4)
    sphere = [{"center": (0.0, 0.0, 0.0), "axes": (20.0, 20.0, 20.0), "rot": np.eye(3),
               "density": 1.0}]
    vol = O.phantom_rasterize(sphere, g)
    exact = O.phantom_project(sphere, g, 1)
    for mode in (O.FP_GPU_CLAMP, O.FP_CPU_ZERO):
        fp = O.fp(g, vol, 1, 0.5, mode)
        inside = exact > 5.0
        core = exact > 20.0
        # voxelised sphere + trilinear sampling: away from grazing chords the
        # ray-march stays within ~1.5 voxels of the chord, and the integral over
        # the shadow within 2 %
        assert np.abs(fp - exact)[core].max() < 1.6
        assert abs(fp[inside].sum() / exact[inside].sum() - 1) < 0.02


def test_fp_constant_volume_counts_samples():
    """Inside the volume a constant volume samples to exactly its value, so each
    ray sums step * (number of samples), which brackets the chord length."""
    g = O.Geom(nx=24, ny=20, nz=16, dx=1.0, dy=1.0, dz=1.0, nu=30, nv=24, du=1.6, dv=1.6, nproj=3)
    vol = np.ones((16, 20, 24), np.float32)
    for mode in (O.FP_GPU_CLAMP, O.FP_CPU_ZERO):
        fp = O.fp(g, vol, 0, 0.5, mode)
        hit = fp > 0
        assert hit.any()
        # values are multiples of the step (within float rounding on the zero mode edge)
        frac = fp[hit] / 0.5 - np.round(fp[hit] / 0.5)
        if mode == O.FP_GPU_CLAMP:
            assert np.abs(frac).max() < 1e-4
        # view 0 looks along +y: the rays next to the centre cross the 19 mm
        # (ny - 1 voxel pitches) box, so they sum 38 or 39 samples of 0.5 mm
        assert (fp[11:13, 14:16] >= 19.0).all() and (fp[11:13, 14:16] <= 19.5 + 1e-4).all()
        assert fp.max() <= 19.5 + 1e-4


# ------------------------------------------------------------------- phantom
def test_phantom_volume_and_central_chord():
    g = O.Geom(nx=80, ny=80, nz=80, dx=1.0, dy=1.0, dz=1.0, nu=64, nv=64, du=2.0, dv=2.0, nproj=4)
    rot = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    e = [{"center": (3.0, -2.0, 1.0), "axes": (25.0, 15.0, 10.0), "rot": rot, "density": 0.7}]
    vol = O.phantom_rasterize(e, g)
    assert vol.sum() / 0.7 == pytest.approx(4 / 3 * math.pi * 25 * 15 * 10, rel=0.02)
    s = [{"center": (0.0, 0.0, 0.0), "axes": (12.0, 12.0, 12.0), "rot": np.eye(3), "density": 2.0}]
    g2 = O.Geom(nx=8, ny=8, nz=8, nu=65, nv=65, du=1.0, dv=1.0, nproj=4)
    proj = O.phantom_project(s, g2, 0)
    assert proj[32, 32] == pytest.approx(2.0 * 24.0, rel=1e-12)   # through the centre


@pytest.mark.parametrize("pad_mode", [O.PAD_MEAN4, O.PAD_MEAN3])
@pytest.mark.parametrize("nu", [6, 37, 64, 200, 256])
def test_mean_pads_on_random_rows_match_an_explicit_dft(pad_mode, nu):
    """Non-constant rows: the oracle's padded row filter equals a numpy DFT of
    the explicitly padded row.  Filterlib (FilterLib.h:37-46) fills the left
    (N - nu) / 2 samples with (x0 + x1 + x2 + x3) / 4 and the right ones with
    (x[-1] + x[-2] + x[-3] + x[-4]) / 4, summed in float; the cuFFT path
    (cudaFftFilter.cu:22-30) uses the 3-sample means.  Odd nu: the one extra
    right sample takes the right mean (a deliberate deviation from the
    reference, whose stale sample is undefined)."""
    rng = np.random.default_rng(nu + 10 * pad_mode)
    N = O.nextpow2(nu)
    H = O.filter_design(5, nu)
    rows = (rng.standard_normal((3, 4, nu)) + 2.0).astype(np.float32)
    out = O.filter_rows(rows, H, pad_mode)
    off = (N - nu) // 2
    k = 4 if pad_mode == O.PAD_MEAN4 else 3
    flat = rows.reshape(-1, nu)
    ref = np.empty_like(flat, dtype=np.float64)
    for i, x in enumerate(flat):
        left = np.float32(0.0)
        right = np.float32(0.0)
        for j in range(k):  # summed in float, in the reference's order
            left = np.float32(left + x[j])
            right = np.float32(right + x[nu - 1 - j])
        left, right = np.float32(left / np.float32(k)), np.float32(right / np.float32(k))
        padded = np.empty(N)
        padded[:off] = left
        padded[off:off + nu] = x
        padded[off + nu:] = right
        ref[i] = np.fft.ifft(np.fft.fft(padded) * H).real[off:off + nu]
    err = np.abs(out.reshape(-1, nu) - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err
    # the pad values matter: the same rows with the other mode differ
    other = O.filter_rows(rows, H, O.PAD_MEAN3 if pad_mode == O.PAD_MEAN4 else O.PAD_MEAN4)
    assert np.abs(other.reshape(-1, nu) - ref).max() / np.abs(ref).max() > 1e-4
