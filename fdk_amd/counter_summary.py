"""Summarise rocprofv3 counter passes of one kernel as the JSON the bench reads.

    python -m fdk_amd.counter_summary <out_dir> <kernel> <workload> <updates_per_launch> <pass_dir>...

Each ``<out_dir>/<pass_dir>`` holds one ``rocprofv3 --pmc ... -o pmc -f csv`` run
(``pmc_counter_collection.csv``).  Counter values are summed per dispatch over
the rows of the named kernel and averaged over its dispatches.  HBM traffic
scales FETCH_SIZE by 2.0 and WRITE_SIZE by 1.0 (KiB per dispatch; the factors a
plain 1 GiB device copy reads on this part, ``bench.py --calibrate-copy``).
Busy fractions divide by GRBM_GUI_ACTIVE / 8 (cycles per XCD).  The JSON goes
to stdout; ``bench.py`` reads ``profiles/<kernel>_<dtype>_<n>cube_counters.json``
for its roofline object.
"""
from __future__ import annotations

import collections
import csv
import json
import os
import sys

FETCH_FACTOR, WRITE_FACTOR = 2.0, 1.0


def summarise(out_dir: str, kernel: str, workload: str, updates: float, passes: list[str]) -> dict:
    vals = collections.defaultdict(list)
    for d in passes:
        with open(os.path.join(out_dir, d, "pmc_counter_collection.csv")) as fh:
            rows = list(csv.DictReader(fh))
        per = collections.defaultdict(float)
        for r in rows:
            name = r["Kernel_Name"]
            if not name.startswith(kernel):  # kernel or kernel<template args>(...)
                continue
            per[(r["Dispatch_Id"], r["Counter_Name"])] += float(r["Counter_Value"])
        for (_, c), v in per.items():
            vals[c].append(v)
    avg = {c: sum(v) / len(v) for c, v in vals.items()}
    fetch = avg["FETCH_SIZE"] * 1024 * FETCH_FACTOR
    write = avg["WRITE_SIZE"] * 1024 * WRITE_FACTOR
    cycles = avg["GRBM_GUI_ACTIVE"] / 8
    return {
        "kernel": kernel, "workload": workload, "voxel_updates_per_launch": updates,
        "counters_avg_per_dispatch": {c: avg[c] for c in sorted(avg)},
        "dispatches": {c: len(v) for c, v in vals.items()},
        "valu_insts_per_update": avg["SQ_INSTS_VALU"] / updates,
        "valu_insts_per_wave_update": avg["SQ_INSTS_VALU"] / (updates / 64),
        "lds_insts_per_update": avg["SQ_INSTS_LDS"] / updates,
        "salu_insts_per_update": avg["SQ_INSTS_SALU"] / updates,
        "valu_busy": avg["SQ_ACTIVE_INST_VALU"] * 4 / 1024 / cycles,
        "lds_busy": avg["SQ_ACTIVE_INST_LDS"] * 4 / 1024 / cycles,
        "gpu_cycles": cycles,
        "hbm_fetch_bytes_per_launch": fetch, "hbm_write_bytes_per_launch": write,
        "hbm_bytes_per_launch": fetch + write,
        "l2_bytes_per_launch": avg["TCP_TCC_READ_REQ_sum"] * 64,
        "tcc_hit_rate": avg["TCC_HIT_sum"] / (avg["TCC_HIT_sum"] + avg["TCC_MISS_sum"]),
        "fetch_factor": FETCH_FACTOR, "write_factor": WRITE_FACTOR,
    }


def main(argv=None) -> None:
    a = sys.argv[1:] if argv is None else argv
    if len(a) < 5:
        raise SystemExit(__doc__)
    print(json.dumps(summarise(a[0], a[1], a[2], float(a[3]), a[4:]), indent=1))


if __name__ == "__main__":
    main()
