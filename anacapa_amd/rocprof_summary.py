"""Summarise rocprofv3 CSV output of a bench run into a short text table.

Inputs: the ``*_kernel_stats.csv`` of a ``--kernel-trace --stats`` run and, optionally, the
``*_counter_collection.csv`` files of one ``--pmc FETCH_SIZE`` and one ``--pmc WRITE_SIZE`` run
of a bench invocation that started with ``--copy-probe-mb M`` (one elementwise kernel that reads
M MiB and writes M MiB).  The probe calibrates both counters: factor = known KiB / counted KiB
for that kernel, applied to every other kernel's per-dispatch figure.

    python -m anacapa_amd.rocprof_summary --stats S.csv [--fetch F.csv --write W.csv --probe-mb M]
        [--family k_extend,k_shadow,k_mis] [--json out.json]
"""
from __future__ import annotations

import argparse
import collections
import csv
import json

PROBE = "vectorized_elementwise_kernel"
FILL = "FillFunctor"


def short(name: str) -> str:
    base = name.split("(")[0]
    return base.replace("void ", "").strip()


def kernel_stats(path):
    rows = list(csv.DictReader(open(path)))
    return [(short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
             float(r["Percentage"])) for r in rows]


def counters(path, counter):
    """Per-kernel list of per-dispatch counter values (KiB)."""
    out = collections.defaultdict(list)
    probe = []
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        v = float(r["Counter_Value"])
        if PROBE in r["Kernel_Name"]:
            if FILL not in r["Kernel_Name"]:
                probe.append(v)
            continue
        out[short(r["Kernel_Name"])].append(v)
    return out, probe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--probe-mb", type=float, default=0.0)
    ap.add_argument("--family", default="k_extend,k_shadow,k_mis")
    ap.add_argument("--sq", help="counter CSV of a run with SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU "
                                 "SQ_WAVE_CYCLES SQ_INSTS_VALU SQ_WAVES GRBM_GUI_ACTIVE")
    ap.add_argument("--cus", type=int, default=256)
    ap.add_argument("--xcds", type=int, default=8,
                    help="GRBM_GUI_ACTIVE is reported summed over the XCDs; the derived metrics take its max")
    ap.add_argument("--json")
    a = ap.parse_args(argv)
    fam = a.family.split(",")
    in_fam = lambda n: any(n.split("::")[-1].startswith(f) for f in fam)  # noqa: E731

    st = kernel_stats(a.stats)
    print("| kernel | calls | total ms | avg us | % |")
    print("|---|---:|---:|---:|---:|")
    for n, c, tot, avg, pct in st:
        print(f"| {n} | {c} | {tot / 1e6:.2f} | {avg / 1e3:.1f} | {pct:.2f} |")
    fcalls = sum(c for n, c, *_ in st if in_fam(n))
    ftot = sum(t for n, _, t, *_ in st if in_fam(n))
    res = {"family": fam, "family_dispatches": fcalls,
           "family_avg_dispatch_us": round(ftot / max(1, fcalls) / 1e3, 2),
           "kernels": {n: {"calls": c, "total_ms": round(tot / 1e6, 3), "avg_us": round(avg / 1e3, 2)}
                       for n, c, tot, avg, _ in st}}
    print(f"\ntraversal family ({', '.join(fam)}): {fcalls} dispatches, mean {res['family_avg_dispatch_us']} us")

    if a.fetch and a.write and a.probe_mb:
        known = a.probe_mb * 1024.0
        fam_kib = {}
        for tag, path, cname in (("fetch", a.fetch, "FETCH_SIZE"), ("write", a.write, "WRITE_SIZE")):
            per, probe = counters(path, cname)
            if not probe:
                raise SystemExit(f"no probe kernel in {path}")
            factor = known / probe[-1]
            res[f"{tag}_probe_kib"] = probe[-1]
            res[f"{tag}_factor"] = round(factor, 4)
            print(f"\n{cname}: probe kernel counted {probe[-1]:.0f} KiB for {known:.0f} KiB -> factor {factor:.4f}")
            print(f"| kernel | dispatches | raw KiB/dispatch | calibrated MB/dispatch |")
            print("|---|---:|---:|---:|")
            fv = []
            for n, vals in sorted(per.items()):
                m = sum(vals) / len(vals)
                print(f"| {n} | {len(vals)} | {m:.1f} | {m * factor * 1024 / 1e6:.3f} |")
                if in_fam(n):
                    fv += vals
            fam_kib[tag] = sum(fv) / max(1, len(fv)) * factor
        traffic = (fam_kib["fetch"] + fam_kib["write"]) * 1024.0
        res["family_traffic_bytes_per_dispatch"] = round(traffic)
        print(f"\ntraversal family HBM traffic per dispatch (calibrated FETCH + WRITE): {traffic / 1e6:.2f} MB")
    if a.sq:
        # rocprofv3's derived-metric definitions for this agent (rocprofv3 -L), with
        # GRBM_GUI_ACTIVE = max over XCDs (the CSV holds the sum; all XCDs run the whole kernel):
        #   VALUBusy = 100 * SQ_ACTIVE_INST_VALU / CU_NUM / GRBM_GUI_ACTIVE
        #   VALUUtilization = 100 * SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU * 64)
        #   OccupancyPercent = 400 * SQ_WAVE_CYCLES / GRBM_GUI_ACTIVE / CU_NUM / 32
        per = collections.defaultdict(lambda: collections.defaultdict(list))
        for r in csv.DictReader(open(a.sq)):
            per[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        print("\n| kernel | dispatches | VALU busy % | VALU utilization % | occupancy % (waves/SIMD of 8) | VALU insts/wave |")
        print("|---|---:|---:|---:|---:|---:|")
        res["sq"] = {}
        for n, c in sorted(per.items()):
            if "GRBM_GUI_ACTIVE" not in c or "SQ_ACTIVE_INST_VALU" not in c:
                continue
            g = sum(c["GRBM_GUI_ACTIVE"]) / a.xcds
            act = sum(c["SQ_ACTIVE_INST_VALU"])
            busy = 100.0 * act / a.cus / max(g, 1.0)
            util = 100.0 * sum(c.get("SQ_THREAD_CYCLES_VALU", [0.0])) / max(act * 64.0, 1.0)
            occ = 400.0 * sum(c.get("SQ_WAVE_CYCLES", [0.0])) / max(g, 1.0) / a.cus / 32.0
            ipw = sum(c.get("SQ_INSTS_VALU", [0.0])) / max(sum(c.get("SQ_WAVES", [1.0])), 1.0)
            res["sq"][n] = {"valu_busy": round(busy, 1), "valu_util": round(util, 1), "occupancy": round(occ, 1)}
            print(f"| {n} | {len(c['GRBM_GUI_ACTIVE'])} | {busy:.1f} | {util:.1f} | {occ:.1f} ({occ * 0.08:.1f}) | {ipw:.0f} |")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
