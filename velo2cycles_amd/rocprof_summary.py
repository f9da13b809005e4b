"""Short text and JSON summaries of rocprofv3 result databases (rocpd SQLite).

    python -m velo2cycles_amd.rocprof_summary trace DB
        per-kernel calls, average and total duration (the --kernel-trace --stats run)
    python -m velo2cycles_amd.rocprof_summary pmc DB [DB ...]
        per-kernel counter values averaged over dispatches (--pmc runs)
    python -m velo2cycles_amd.rocprof_summary traffic --fetch DB --write DB --copy-bytes N
                                                       --workload NAME [--out JSON]
        HBM bytes per launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
        scaled by the factor that makes the plain copy kernel k_copy_calib (N bytes
        read and N bytes written) read exactly N bytes
    python -m velo2cycles_amd.rocprof_summary requests --reads DB --write DB --copy-bytes N
                                                        --workload NAME [--out JSON]
        HBM bytes per launch from the L2-to-memory request counts by size:
        reads 128 TCC_EA0_RDREQ_128B + 32 TCC_EA0_RDREQ_32B + 64 x the rest of
        TCC_EA0_RDREQ, writes WRITE_SIZE; the plain copy kernel's figures give
        the check factors (1.0 when every request is counted at its size).
        (FETCH_SIZE counts a request as 128 B only through TCC_BUBBLE, which on
        gfx950 misses the 128-byte reads: a streaming copy reads as half.)
    python -m velo2cycles_amd.rocprof_summary valu DB --workload NAME [--out JSON]
        VALU and f64-transcendental wavefront instructions per launch
        (SQ_INSTS_VALU, SQ_INSTS_VALU_TRANS_F64)

Kernel names are shortened to the library's own names (v2c::k_...).
"""
from __future__ import annotations

import argparse
import collections
import json
import re
import sqlite3

# library kernel -> the name bench.py reports it under
# (the fused update's two kernels, one dispatch each per step, are one bench
# entry: their per-dispatch figures add)
BENCH_NAME = {"k_m2n_props": "m2n_props", "k_m2n_cells": "m2n_cells", "k_n2m_scatter": "n2m_cells",
              "k_n2m_update<true>": "n2m_update", "k_n2m_update_fast": "n2m_update",
              "k_n2m_update_list": "n2m_update", "k_m2n_nodes": "m2n_nodes", "k_n2m_nodes": "n2m_nodes"}


def short(name: str) -> str:
    n = re.sub(r"\(.*", "", name).replace("void ", "")
    if "rocprim" in n:
        m = re.search(r"(radix_sort_onesweep_\w+|onesweep_histogram\w*|\w+_kernel)", n)
        return "rocprim::" + (m.group(1) if m else "kernel")
    return n.replace("v2c::", "")[:48]


def trace(path: str) -> str:
    db = sqlite3.connect(path)
    rows = db.execute("select name, count(*), sum(duration) from kernels group by name").fetchall()
    agg = collections.OrderedDict()
    total = sum(r[2] for r in rows)
    for name, calls, ns in sorted(rows, key=lambda r: -r[2]):
        a = agg.setdefault(short(name), [0, 0.0])
        a[0] += calls
        a[1] += ns
    lines = [f"{'kernel':40s} {'calls':>6s} {'avg_us':>10s} {'total_us':>11s} {'pct':>6s}"]
    for k, (c, t) in agg.items():
        lines.append(f"{k:40s} {c:6d} {t / c / 1e3:10.1f} {t / 1e3:11.1f} {100.0 * t / total:6.2f}")
    return "\n".join(lines)


def pmc_table(path: str) -> dict:
    """{kernel: {counter: mean value per dispatch}}"""
    db = sqlite3.connect(path)
    rows = db.execute("select kernel_name, counter_name, dispatch_id, value from counters_collection").fetchall()
    tot = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    for k, c, d, v in rows:
        tot[short(k)][c] += v
        disp[short(k)].add(d)
    return {k: {c: v / len(disp[k]) for c, v in cs.items()} for k, cs in tot.items()}


def pmc(paths) -> str:
    lines = []
    for p in paths:
        t = pmc_table(p)
        for k in sorted(t, key=lambda k: -sum(t[k].values())):
            lines.append(f"{k:40s} " + " ".join(f"{c}={v:.6g}" for c, v in sorted(t[k].items())))
    return "\n".join(lines)


def traffic(fetch_db: str, write_db: str, copy_bytes: int, workload: str) -> dict:
    f, w = pmc_table(fetch_db), pmc_table(write_db)
    cf = copy_bytes / (f["k_copy_calib"]["FETCH_SIZE"] * 1024.0)
    cw = copy_bytes / (w["k_copy_calib"]["WRITE_SIZE"] * 1024.0)
    out = {"_calibration": {"copy_bytes": copy_bytes, "fetch_factor": cf, "write_factor": cw}}
    for k, name in BENCH_NAME.items():
        if k in f and k in w:
            rd = f[k]["FETCH_SIZE"] * 1024.0 * cf
            wr = w[k]["WRITE_SIZE"] * 1024.0 * cw
            e = out.setdefault(name, {"hbm_read_bytes": 0.0, "hbm_write_bytes": 0.0, "hbm_bytes_per_launch": 0.0,
                                      "kernels": []})
            e["hbm_read_bytes"] += rd
            e["hbm_write_bytes"] += wr
            e["hbm_bytes_per_launch"] += rd + wr
            e["kernels"].append(k)
    return {workload: out}


def _req_bytes(c: dict) -> float:
    r, r32, r128 = c["TCC_EA0_RDREQ_sum"], c["TCC_EA0_RDREQ_32B_sum"], c["TCC_EA0_RDREQ_128B_sum"]
    return 128.0 * r128 + 32.0 * r32 + 64.0 * (r - r128 - r32)


def requests(reads_db: str, write_db: str, copy_bytes: int, workload: str) -> dict:
    f, w = pmc_table(reads_db), pmc_table(write_db)
    cf = copy_bytes / _req_bytes(f["k_copy_calib"])
    cw = copy_bytes / (w["k_copy_calib"]["WRITE_SIZE"] * 1024.0)
    cb = f["k_copy_calib"]
    fetch_size_formula = (cb["TCC_BUBBLE_sum"] * 128.0 + (cb["TCC_EA0_RDREQ_sum"] - cb["TCC_BUBBLE_sum"]
                          - cb["TCC_EA0_RDREQ_32B_sum"]) * 64.0 + cb["TCC_EA0_RDREQ_32B_sum"] * 32.0)
    out = {"_calibration": {"copy_bytes": copy_bytes, "read_factor": cf, "write_factor": cw,
                            "copy_read_by_fetch_size_formula": fetch_size_formula,
                            "method": "reads from TCC_EA0_RDREQ by request size, writes WRITE_SIZE"}}
    per_kernel = {}
    for k in f:
        if k in w and "TCC_EA0_RDREQ_sum" in f[k]:
            per_kernel[k] = {"read_bytes": _req_bytes(f[k]) * cf,
                             "write_bytes": w[k]["WRITE_SIZE"] * 1024.0 * cw}
    for k, name in BENCH_NAME.items():
        if k in per_kernel:
            rd, wr = per_kernel[k]["read_bytes"], per_kernel[k]["write_bytes"]
            e = out.setdefault(name, {"hbm_read_bytes": 0.0, "hbm_write_bytes": 0.0, "hbm_bytes_per_launch": 0.0,
                                      "kernels": []})
            e["hbm_read_bytes"] += rd
            e["hbm_write_bytes"] += wr
            e["hbm_bytes_per_launch"] += rd + wr
            e["kernels"].append(k)
    out["_all_kernels"] = per_kernel
    return {workload: out}


def valu(db: str, workload: str) -> dict:
    t = pmc_table(db)
    out = {}
    for k, name in BENCH_NAME.items():
        if k in t and "SQ_INSTS_VALU" in t[k]:
            e = out.setdefault(name, {"valu_per_launch": 0.0, "trans_f64_per_launch": 0.0, "kernels": []})
            e["valu_per_launch"] += t[k]["SQ_INSTS_VALU"]
            e["trans_f64_per_launch"] += t[k].get("SQ_INSTS_VALU_TRANS_F64", 0.0)
            e["kernels"].append(k)
    return {workload: out}


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("trace")
    a.add_argument("db")
    a = sub.add_parser("pmc")
    a.add_argument("db", nargs="+")
    a = sub.add_parser("traffic")
    a.add_argument("--fetch", required=True)
    a.add_argument("--write", required=True)
    a.add_argument("--copy-bytes", type=int, default=1 << 30)
    a.add_argument("--workload", required=True)
    a.add_argument("--out")
    a = sub.add_parser("requests")
    a.add_argument("--reads", required=True)
    a.add_argument("--write", required=True)
    a.add_argument("--copy-bytes", type=int, default=1 << 30)
    a.add_argument("--workload", required=True)
    a.add_argument("--out")
    a = sub.add_parser("valu")
    a.add_argument("db")
    a.add_argument("--workload", required=True)
    a.add_argument("--out")
    args = ap.parse_args()
    if args.cmd == "trace":
        print(trace(args.db))
    elif args.cmd == "pmc":
        print(pmc(args.db))
    else:
        if args.cmd == "traffic":
            res = traffic(args.fetch, args.write, args.copy_bytes, args.workload)
        elif args.cmd == "requests":
            res = requests(args.reads, args.write, args.copy_bytes, args.workload)
        else:
            res = valu(args.db, args.workload)
        text = json.dumps(res, indent=1)
        if args.out:
            open(args.out, "w").write(text + "\n")
        print(text)


if __name__ == "__main__":
    main()
