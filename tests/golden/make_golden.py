"""Writes the golden fixtures of tests/golden/ from the reference's own CPU code.

Runs only in the container that holds the reference checkout: it needs
oracle/_ref/libfdkref.so, which oracle/build_ref.sh compiles from the reference
sources in place (common/type.h, 3dct/3dct.cpp).  Every fixture is data: seeded
inputs and the reference's outputs for them.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

TINY = dict(nx=16, ny=16, nz=16, nu=24, nv=20, du=1.6, dv=1.6, nproj=8,
            volume_position_angle_rad=-0.3, filter_type=5)
GEOM_256 = dict(nx=256, ny=256, nz=256, dx=1.0, dy=1.0, dz=1.0, nu=256, nv=256, du=1.6, dv=1.6,
          nproj=360, volume_position_angle_rad=0.0, filter_type=5)
ODD = dict(nx=20, ny=18, nz=14, dx=1.3, dy=1.1, dz=0.9, nu=30, nv=26, du=1.2, dv=1.4,
           offset_nu=1.5, offset_nv=-2.0, nproj=12, scan_angle_deg=200.0, direction=1,
           volume_position_angle_rad=-0.8528925, filter_type=2)


def save(name, geom=None, **arrays):
    if geom is not None:
        arrays["geom"] = np.array(json.dumps(geom))
    np.savez_compressed(os.path.join(HERE, name), **arrays)
    print("wrote", name, {k: v.shape for k, v in arrays.items()})


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libfdkref.so is missing (run oracle/build_ref.sh)")
    rng = np.random.default_rng(2024)
    for name, geom in (("tables_tiny.npz", TINY), ("tables_256.npz", GEOM_256), ("tables_odd.npz", ODD)):
        g = O.Geom(**geom)
        r = O.ref_tables(g)
        save(name, geom, cos=r["cos"], filter=r["filter"], mats=r["mats"], af=np.array(r["af"]))
    designs = {}
    for t in range(6):
        for L in (24, 256, 300, 2048):
            designs[f"t{t}_n{L}"] = O.ref_filter_design(t, L)
    save("filter_designs.npz", **designs)
    x64 = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    x512 = rng.standard_normal(512) + 1j * rng.standard_normal(512)
    save("fft.npz", x64=x64, fwd64=O.ref_fft(x64), inv64=O.ref_fft(x64, True),
         x512=x512, fwd512=O.ref_fft(x512), inv512=O.ref_fft(x512, True))
    img = rng.standard_normal((20, 24)).astype(np.float32)
    H24 = O.ref_filter_design(5, 24)
    rows = (rng.random((8, 256)) * 10).astype(np.float32)
    H256 = O.ref_filter_design(2, 256)
    save("filter_rows_mean3.npz", img=img, H24=H24, out24=O.ref_filter_rows_3dct(img, H24),
         rows=rows, H256=H256, out256=O.ref_filter_rows_3dct(rows, H256))
    g = O.Geom(**TINY)
    proj = rng.standard_normal((g.nproj, g.nv, g.nu)).astype(np.float32)
    raw = (rng.random((g.nproj, g.nv, g.nu)) * 20).astype(np.float32)
    save("bp_3dct_tiny.npz", TINY, proj=proj, vol_prefiltered=O.ref_bp_3dct(g, proj, True),
         raw=raw, vol_fdk=O.ref_bp_3dct(g, raw, False))
    go = O.Geom(**ODD)
    proj_o = rng.standard_normal((go.nproj, go.nv, go.nu)).astype(np.float32)
    save("bp_3dct_odd.npz", ODD, proj=proj_o, vol_prefiltered=O.ref_bp_3dct(go, proj_o, True))
    vol = rng.random((g.nz, g.ny, g.nx)).astype(np.float32)
    save("fp_3dct_tiny.npz", TINY, vol=vol,
         views=np.stack([O.ref_fp_3dct(g, vol, s) for s in range(3)]))


if __name__ == "__main__":
    main()
