"""fdk_amd.counter_summary on synthetic rocprofv3 counter CSVs: per-dispatch sums
over the named kernel only (template instantiations included, other kernels
excluded), averages over dispatches, and the derived per-update and busy
figures the bench's roofline object reads."""
import csv
import json

import pytest

from fdk_amd import counter_summary as CS

FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_pass(d, rows):
    d.mkdir()
    with open(d / "pmc_counter_collection.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(FIELDS, r)))


def test_summary_figures(tmp_path, capsys):
    k = "bp_wave_kernel<0, 8>(BpArgs, BpMats, BpScale, int)"
    other = "filter_rows_kernel(FilterArgs)"
    # pass 1: two dispatches, each counter split over two rows (summed per dispatch)
    p1 = []
    for di, scale in ((1, 1.0), (2, 3.0)):
        for c, v in (("SQ_INSTS_VALU", 640.0), ("SQ_INSTS_LDS", 64.0), ("SQ_INSTS_SALU", 32.0),
                     ("SQ_ACTIVE_INST_VALU", 512.0), ("GRBM_GUI_ACTIVE", 8.0 * 1024)):
            p1.append((di, k, c, v * scale / 2))
            p1.append((di, k, c, v * scale / 2))
        p1.append((di + 10, other, "SQ_INSTS_VALU", 1e9))  # another kernel: ignored
    write_pass(tmp_path / "p1", p1)
    write_pass(tmp_path / "p2", [(1, k, "SQ_ACTIVE_INST_LDS", 256.0), (1, k, "FETCH_SIZE", 4.0),
                                 (1, k, "WRITE_SIZE", 2.0), (1, k, "TCC_HIT_sum", 3.0),
                                 (1, k, "TCC_MISS_sum", 1.0), (1, k, "TCP_TCC_READ_REQ_sum", 5.0)])
    CS.main([str(tmp_path), "bp_wave_kernel", "w", "128", "p1", "p2"])
    r = json.loads(capsys.readouterr().out)
    avg = r["counters_avg_per_dispatch"]
    assert avg["SQ_INSTS_VALU"] == pytest.approx(640.0 * 2)  # mean of 640 and 1920
    assert r["dispatches"]["SQ_INSTS_VALU"] == 2
    assert r["valu_insts_per_update"] == pytest.approx(1280.0 / 128)
    assert r["valu_insts_per_wave_update"] == pytest.approx(1280.0 / 2)
    assert r["lds_insts_per_update"] == pytest.approx(128.0 / 128)
    cycles = 2 * 8.0 * 1024 / 8  # GRBM_GUI_ACTIVE averaged, per XCD
    assert r["gpu_cycles"] == pytest.approx(cycles)
    assert r["valu_busy"] == pytest.approx(1024.0 * 4 / 1024 / cycles)
    assert r["lds_busy"] == pytest.approx(256.0 * 4 / 1024 / cycles)
    assert r["hbm_fetch_bytes_per_launch"] == 4.0 * 1024 * 2.0
    assert r["hbm_write_bytes_per_launch"] == 2.0 * 1024
    assert r["hbm_bytes_per_launch"] == 4.0 * 1024 * 2.0 + 2.0 * 1024
    assert r["l2_bytes_per_launch"] == 5.0 * 64
    assert r["tcc_hit_rate"] == pytest.approx(0.75)
