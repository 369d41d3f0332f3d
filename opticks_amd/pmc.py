"""
Summarize rocprofv3 counter-collection runs of the propagation kernel into the
per-launch figures bench.py reports beside its roofline.

Three runs of `bench.py --pmc-pass <workload> --copy-calibration-mib 256`, one
counter group each (rocprofv3 collects one group per run):

  rocprofv3 --pmc FETCH_SIZE -d <dir>/fetch -o fetch -f csv -- python3 -u bench.py --pmc-pass ...
  rocprofv3 --pmc WRITE_SIZE -d <dir>/write -o write -f csv -- python3 -u bench.py --pmc-pass ...
  rocprofv3 --pmc SQ_INSTS_VALU SQ_THREAD_CYCLES_VALU SQ_ACTIVE_INST_VALU SQ_WAVES SQ_BUSY_CYCLES \\
      SQ_WAVE_CYCLES GRBM_GUI_ACTIVE -d <dir>/valu -o valu -f csv -- python3 -u bench.py --pmc-pass ...

  python -m opticks_amd.pmc <dir> <workload> <out.json>

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what the same counter
reads for the device copy of known size in the same process (the rocclr copy
kernel, 256 MiB read and 256 MiB written), so `hbm_bytes_per_launch` is in bytes
of HBM traffic. Lane utilisation = SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU)
as the counters define them (a copy kernel reads 64 active lanes). The last
dispatch of the propagation kernel is used (the first one warms caches).
"""
from __future__ import annotations

import csv
import glob
import json
import os
import sys
from collections import defaultdict


def read_dispatches(run_dir):
    """{dispatch id: {"kernel", "counters": {name: value}, resources}} of every
    *counter_collection.csv under run_dir"""
    out = {}
    files = sorted(glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True))
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {run_dir}")
    for path in files:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                key = (path, int(row["Dispatch_Id"]))
                d = out.setdefault(key, {"kernel": row["Kernel_Name"], "counters": defaultdict(float),
                                         "vgpr": int(row.get("VGPR_Count", 0) or 0),
                                         "scratch": int(row.get("Scratch_Size", 0) or 0),
                                         "lds": int(row.get("LDS_Block_Size", 0) or 0),
                                         "workgroup": int(row.get("Workgroup_Size", 0) or 0)})
                d["counters"][row["Counter_Name"]] += float(row["Counter_Value"])
    return [out[k] for k in sorted(out)]


def last_of(dispatches, prefix):
    sel = [d for d in dispatches if d["kernel"].replace("void ", "").startswith(prefix)]
    if not sel:
        names = sorted({d["kernel"] for d in dispatches})
        raise KeyError(f"no dispatch of {prefix}; kernels seen: {names[:20]}")
    return sel[-1]


def summarize(base, workload, n_photon, copy_mib=256, kernel="k_sim_wave"):
    fetch = read_dispatches(os.path.join(base, "fetch"))
    write = read_dispatches(os.path.join(base, "write"))
    valu = read_dispatches(os.path.join(base, "valu"))
    kf, kw, kv = last_of(fetch, kernel), last_of(write, kernel), last_of(valu, kernel)
    cf, cw = last_of(fetch, "__amd_rocclr_copyBuffer"), last_of(write, "__amd_rocclr_copyBuffer")
    moved_kib = copy_mib * 1024.0
    ff = moved_kib / cf["counters"]["FETCH_SIZE"]
    wf = moved_kib / cw["counters"]["WRITE_SIZE"]
    fb = kf["counters"]["FETCH_SIZE"] * ff * 1024.0
    wb = kw["counters"]["WRITE_SIZE"] * wf * 1024.0
    c = kv["counters"]
    util = c["SQ_THREAD_CYCLES_VALU"] / (64.0 * c["SQ_ACTIVE_INST_VALU"]) if c.get("SQ_ACTIVE_INST_VALU") else None
    return {
        "workload": workload, "n_photon": int(n_photon), "kernel": kernel, "instantiation": kv["kernel"],
        "fetch_kib_raw": kf["counters"]["FETCH_SIZE"], "write_kib_raw": kw["counters"]["WRITE_SIZE"],
        "copy_mib": copy_mib, "copy_fetch_kib": cf["counters"]["FETCH_SIZE"],
        "copy_write_kib": cw["counters"]["WRITE_SIZE"], "fetch_factor": ff, "write_factor": wf,
        "hbm_fetch_bytes": fb, "hbm_write_bytes": wb, "hbm_bytes_per_launch": fb + wb,
        "valu_insts_per_launch": c.get("SQ_INSTS_VALU"), "valu_active_inst": c.get("SQ_ACTIVE_INST_VALU"),
        "valu_thread_cycles": c.get("SQ_THREAD_CYCLES_VALU"), "valu_lane_utilisation": util,
        "sq_waves": c.get("SQ_WAVES"), "sq_wave_cycles": c.get("SQ_WAVE_CYCLES"),
        "sq_busy_cycles": c.get("SQ_BUSY_CYCLES"), "grbm_gui_active": c.get("GRBM_GUI_ACTIVE"),
        "block": kv["workgroup"], "vgpr_granule": kv["vgpr"], "scratch_bytes_per_lane": kv["scratch"],
        "lds_bytes": kv["lds"],
    }


def workload_photons(workload):
    """photons of one launch of the --pmc-pass workload"""
    from . import configs
    if workload == "csg10M":
        return configs.config1().num_photon
    return (configs.config2() if workload == "batched64" else configs.config4()).num_photon


if __name__ == "__main__":
    base, workload, out = sys.argv[1], sys.argv[2], sys.argv[3]
    r = summarize(base, workload, workload_photons(workload))
    r["command"] = (f"rocprofv3 --pmc <group> -d {base}/<group> -o <group> -f csv -- python3 -u bench.py "
                    f"--pmc-pass {workload} --copy-calibration-mib 256; python -m opticks_amd.pmc {base} "
                    f"{workload} {out}")
    with open(out, "w") as f:
        json.dump(r, f, indent=1)
    print(json.dumps({k: r[k] for k in ("instantiation", "hbm_bytes_per_launch", "valu_insts_per_launch",
                                        "valu_lane_utilisation")}))
