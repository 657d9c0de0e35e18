"""Summaries of rocprofv3 output for the H*v measurements.

    python dmft-ed_amd/dmed/rocprof.py stats   --dir DIR --out FILE.md --cmd "..."
    python dmft-ed_amd/dmed/rocprof.py traffic --fetch DIR --write DIR --dim D
                                              --out FILE.json --md FILE.md --cmd "..."
    python dmft-ed_amd/dmed/rocprof.py l2      --dir DIR --out FILE.md --cmd "..."

`stats` condenses a `--kernel-trace --stats` run (kernel_stats.csv) into a
markdown table.  `traffic` turns two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
runs into HBM bytes per launch of the H*v kernels.  `l2` turns a
`--pmc TCC_HIT_sum TCC_MISS_sum` run into L2 requests and hit rates per
launch of the H*v kernels and of the vector copy.  The traffic counters are scaled
by what they read for the FP64 vector copy kernel (k_scale_copy<double,
double>, which reads and writes exactly 8*dim bytes), so the factors between
counter units and bytes come from the same run.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys

COPY = "k_scale_copy<double, double>"
HXV = {"fp64": "k_hxv_col<double, 0", "fp32": "k_hxv_col<float, 0"}


def _find(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not hits:
        raise SystemExit(f"no *{suffix} under {d}")
    return hits[0]


def _short(name, n=70):
    name = re.sub(r"\s+", " ", name)
    return name if len(name) <= n else name[:n - 3] + "..."


def counters(d, counter):
    """{kernel name: [values per dispatch]} of one counter."""
    out = {}
    with open(_find(d, "counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return out


def _pick(table, key):
    names = [k for k in table if key in k]
    if not names:
        raise SystemExit(f"kernel {key!r} not in the counter table")
    vals = [v for k in names for v in table[k]]
    return names[0], sum(vals) / len(vals), len(vals)


def traffic(a):
    fetch = counters(a.fetch, "FETCH_SIZE")
    write = counters(a.write, "WRITE_SIZE")
    nbytes = 8 * a.dim
    _, fc, _ = _pick(fetch, COPY)
    _, wc, _ = _pick(write, COPY)
    ff, wf = nbytes / fc, nbytes / wc          # bytes per counter unit
    res = {"dim": a.dim, "fetch_bytes_per_unit": ff, "write_bytes_per_unit": wf,
           "copy_fetch_counter": fc, "copy_write_counter": wc}
    rows = []
    for prec, key in HXV.items():
        try:
            name, fv, nf = _pick(fetch, key)
            _, wv, nw = _pick(write, key)
        except SystemExit:
            continue
        rb, wb = fv * ff, wv * wf
        esz = 8 if prec == "fp64" else 4
        res[f"{prec}_bytes_per_launch"] = rb + wb
        res[f"{prec}_read_bytes"] = rb
        res[f"{prec}_write_bytes"] = wb
        res[f"{prec}_algorithmic_bytes"] = 2 * esz * a.dim
        rows.append((prec, _short(name), nf, fv, wv, rb, wb, 2 * esz * a.dim))
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
        f.write("\n")
    lines = [
        "# HBM traffic of the H*v kernel, Ns=16 half-filled sector (dim %d)" % a.dim,
        "",
        "Commands (two runs, one counter each):",
        "",
    ] + ["    " + c for c in a.cmd] + [
        "",
        "Calibration: the FP64 vector copy `%s` moves 8*dim = %.4g bytes each way;" % (COPY, nbytes),
        "it reads FETCH_SIZE = %.6g and WRITE_SIZE = %.6g per dispatch, so one counter" % (fc, wc),
        "unit is %.1f bytes (FETCH_SIZE) and %.1f bytes (WRITE_SIZE)." % (ff, wf),
        "",
        "| precision | kernel | dispatches | FETCH_SIZE | WRITE_SIZE | read bytes | written bytes | algorithmic bytes | traffic / algorithmic |",
        "|---|---|---|---|---|---|---|---|---|",
    ]
    for prec, name, n, fv, wv, rb, wb, alg in rows:
        lines.append(f"| {prec} | `{name}` | {n} | {fv:.6g} | {wv:.6g} | {rb:.4g} | {wb:.4g} | "
                     f"{alg:.4g} | {(rb + wb) / alg:.2f} |")
    with open(a.md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(json.dumps(res))


def l2(a):
    hit = counters(a.dir, "TCC_HIT_sum")
    miss = counters(a.dir, "TCC_MISS_sum")
    keys = [("copy", COPY)] + [(p, k) for p, k in HXV.items()] + [
        ("fp64 Lanczos", "k_hxv_col<double, 1"), ("fp32 Lanczos", "k_hxv_col<float, 1")]
    lines = ["# L2 hits and misses of the H*v kernel, Ns=16 half-filled sector", "",
             "Command:", ""] + ["    " + c for c in a.cmd] + [
        "", "TCC_HIT_sum and TCC_MISS_sum count L2 requests of all 8 XCDs per dispatch "
        "(reads and writes).", "",
        "| kernel | form | dispatches | hits | misses | hit rate |", "|---|---|---|---|---|---|"]
    for form, key in keys:
        try:
            name, h, n = _pick(hit, key)
            _, m, _ = _pick(miss, key)
        except SystemExit:
            continue
        lines.append(f"| `{_short(name)}` | {form} | {n} | {h:.4g} | {m:.4g} | {h / (h + m):.3f} |")
    with open(a.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def stats(a):
    path = _find(a.dir, "kernel_stats.csv")
    with open(path) as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = ["# Kernel trace: " + a.title, "", "Command:", ""] + ["    " + c for c in a.cmd] + [
        "", "| kernel | calls | total ms | average ms | min ms | max ms | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:a.top]:
        lines.append("| `%s` | %s | %.3f | %.4f | %.4f | %.4f | %.1f |" % (
            _short(r["Name"], 90), r["Calls"], float(r["TotalDurationNs"]) / 1e6,
            float(r["AverageNs"]) / 1e6, float(r["MinNs"]) / 1e6, float(r["MaxNs"]) / 1e6,
            float(r["Percentage"])))
    with open(a.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd_name", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--dim", type=int, required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--md", required=True)
    t.add_argument("--cmd", action="append", default=[])
    s = sub.add_parser("stats")
    s.add_argument("--dir", required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--title", default="")
    s.add_argument("--top", type=int, default=12)
    s.add_argument("--cmd", action="append", default=[])
    c = sub.add_parser("l2")
    c.add_argument("--dir", required=True)
    c.add_argument("--out", required=True)
    c.add_argument("--cmd", action="append", default=[])
    a = ap.parse_args(argv)
    {"traffic": traffic, "stats": stats, "l2": l2}[a.cmd_name](a)


if __name__ == "__main__":
    sys.exit(main())
