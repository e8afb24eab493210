"""Summaries of rocprofv3 result databases (SQLite, `-o run` -> run_results.db).

    python -m tval3_amd.profiling kernels DIR [--top N]
        per-kernel calls, total ms, mean us and share of a --kernel-trace run

    python -m tval3_amd.profiling traffic --fetch DIR --write DIR
            [--copy-mib 512] [--kernels k1,k2] [--json FILE]
        HBM bytes per dispatch from one FETCH_SIZE and one WRITE_SIZE --pmc
        run.  The two counters are in KiB per dispatch; each is scaled by what
        it reads for the runs' device-to-device copies of --copy-mib MiB
        (bench.py --copy-mb), whose true read and write volumes are known:
        factor = copy bytes / (counter KiB * 1024).

Kernel names are grouped by their base name (template arguments and
parameter lists dropped), so all instances of a template share one line.
"""
import argparse
import glob
import json
import os
import re
import sqlite3
import sys
from collections import defaultdict


def _db(path):
    if os.path.isdir(path):
        hits = sorted(glob.glob(os.path.join(path, "**", "*_results.db"), recursive=True))
        if not hits:
            raise SystemExit(f"no *_results.db under {path}")
        path = hits[0]
    return sqlite3.connect(path)


def base_name(name):
    """'void k_descend<64, float, false, true>(KA, ...)' -> 'k_descend'."""
    n = name.replace("(anonymous namespace)::", "").split("(")[0]
    n = re.sub(r"<.*>", "", n)
    n = n.replace("void ", "").strip()
    return n.split("::")[-1]


def full_name(name):
    """Kernel name without its parameter list (template arguments kept)."""
    return name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()


def kernel_stats(path, top=30):
    c = _db(path)
    rows = c.execute("select name, duration from kernels").fetchall()
    acc = defaultdict(lambda: [0, 0.0])
    for name, dur in rows:
        a = acc[full_name(name)]
        a[0] += 1
        a[1] += dur * 1e-6  # ns -> ms
    tot = sum(v[1] for v in acc.values())
    out = sorted(acc.items(), key=lambda kv: -kv[1][1])[:top]
    return out, tot


def counter_means(path, counter):
    c = _db(path)
    rows = c.execute("select kernel_name, value from counters_collection where counter_name = ?",
                     (counter,)).fetchall()
    acc = defaultdict(list)
    for name, v in rows:
        acc[name].append(v)
    return acc


def traffic(fetch_dir, write_dir, copy_mib=512, kernels=None):
    fr = counter_means(fetch_dir, "FETCH_SIZE")
    wr = counter_means(write_dir, "WRITE_SIZE")
    copy_kib = copy_mib * 1024.0
    # copies of exactly copy_mib MiB (a run may hold other copies of about
    # that size, e.g. the last chunk of a tensor): FETCH_SIZE counts half
    # their bytes and WRITE_SIZE all of them, within 0.2 %
    cf = [v for n, vs in fr.items() if "copyBuffer" in n for v in vs
          if abs(v - 0.5 * copy_kib) <= 0.002 * 0.5 * copy_kib]
    cw = [v for n, vs in wr.items() if "copyBuffer" in n for v in vs
          if abs(v - copy_kib) <= 0.002 * copy_kib]
    if not cf or not cw:
        raise SystemExit("no calibration copies of the given size in the runs")
    f_fac = copy_kib / (sum(cf) / len(cf))
    w_fac = copy_kib / (sum(cw) / len(cw))
    per = defaultdict(lambda: {"fetch": [], "write": []})
    for n, vs in fr.items():
        per[base_name(n)]["fetch"] += vs
    for n, vs in wr.items():
        per[base_name(n)]["write"] += vs
    res = {}
    for k, d in per.items():
        if kernels and k not in kernels:
            continue
        if not d["fetch"] or not d["write"]:
            continue
        fk = sum(d["fetch"]) / len(d["fetch"])
        wk = sum(d["write"]) / len(d["write"])
        res[k] = {"dispatches": len(d["fetch"]), "fetch_kib": fk, "write_kib": wk,
                  "hbm_bytes_per_launch": (fk * f_fac + wk * w_fac) * 1024.0}
    return res, f_fac, w_fac, cf, cw


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    k.add_argument("--top", type=int, default=30)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-mib", type=int, default=512)
    t.add_argument("--kernels", default=None)
    t.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        out, tot = kernel_stats(a.dir, a.top)
        print(f"{'kernel':64s} {'calls':>7s} {'total ms':>11s} {'avg us':>10s} {'%':>6s}")
        for name, (n, ms) in out:
            print(f"{name[:64]:64s} {n:7d} {ms:11.2f} {1e3 * ms / n:10.2f} {100 * ms / tot:6.2f}")
        print(f"{'all kernels':64s} {'':7s} {tot:11.2f}")
        return
    ks = set(a.kernels.split(",")) if a.kernels else None
    res, ff, wf, cf, cw = traffic(a.fetch, a.write, a.copy_mib, ks)
    def show(vs):
        if len(vs) <= 6:
            return ", ".join(f"{v:.1f}" for v in vs)
        return f"{len(vs)} copies, mean {sum(vs) / len(vs):.1f} (min {min(vs):.1f}, max {max(vs):.1f})"
    print(f"calibration: {a.copy_mib} MiB device copies; FETCH_SIZE {show(cf)} KiB -> factor {ff:.4f}; "
          f"WRITE_SIZE {show(cw)} KiB -> factor {wf:.4f}")
    print(f"{'kernel':40s} {'n':>6s} {'FETCH KiB':>14s} {'WRITE KiB':>14s} {'HBM bytes (cal.)':>17s}")
    for name, d in sorted(res.items(), key=lambda kv: -kv[1]["hbm_bytes_per_launch"]):
        print(f"{name[:40]:40s} {d['dispatches']:6d} {d['fetch_kib']:14.1f} {d['write_kib']:14.1f} "
              f"{d['hbm_bytes_per_launch']:17.4g}")
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"calibration": {"copy_mib": a.copy_mib, "fetch_factor": ff, "write_factor": wf},
                       "kernels": res}, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
