"""Summaries of rocprofv3 output for this library's kernels.

  python -m lagrbf_amd.rocprof_summary stats DIR
      per-kernel table from the *_kernel_stats.csv that `--kernel-trace --stats`
      writes (calls, average / total duration, share).
  python -m lagrbf_amd.rocprof_summary pmc DIR [--kernel SUBSTR]
      mean per-dispatch value of every counter of a `--pmc` run, per kernel.
  python -m lagrbf_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR --copy-mib M [--kernel SUBSTR] [--json OUT --key K]
      HBM bytes per dispatch (median over a kernel's dispatches) from two
      `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs (counter values in KiB per dispatch).  The runs must include a plain copy of
      M MiB (bench.py --copy-calibration M launches torch.neg over float64, which
      reads M MiB and writes M MiB); each counter is scaled by known bytes /
      counter bytes of that copy.
"""
import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict


def _find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        raise SystemExit(f"no {pattern} under {d}")
    return hits


def short_name(name):
    n = name.strip()
    if n.startswith("void "):
        n = n[5:]
    p = n.find("(")
    return n[:p] if p > 0 else n


def stats(d, top=25, out=sys.stdout):
    rows = []
    for f in _find(d, "*kernel_stats.csv"):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    print(f"{'kernel':<58} {'calls':>7} {'avg us':>11} {'total ms':>10} {'%':>6}", file=out)
    for r in rows[:top]:
        print(f"{short_name(r['Name'])[:58]:<58} {int(r['Calls']):>7} {float(r['AverageNs']) / 1e3:>11.2f} "
              f"{float(r['TotalDurationNs']) / 1e6:>10.2f} {100 * float(r['TotalDurationNs']) / tot:>6.2f}", file=out)
    print(f"{'all kernels':<58} {sum(int(r['Calls']) for r in rows):>7} {'':>11} {tot / 1e6:>10.2f}", file=out)


def counters(d, counter):
    """kernel short name -> list of per-dispatch counter values (KiB)."""
    per = defaultdict(lambda: defaultdict(float))
    for f in _find(d, "*counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                per[(short_name(r["Kernel_Name"]), r["Dispatch_Id"])]["v"] += float(r["Counter_Value"])
    out = defaultdict(list)
    for (k, _), v in per.items():
        out[k].append(v["v"])
    return out


def pmc(d, kernel_sub="", out=sys.stdout, last=0):
    """Mean per-dispatch value of every collected counter, per kernel (over the
    last `last` dispatches of each kernel when last > 0: the steady state after
    the autotuner's trial launches)."""
    per = defaultdict(lambda: defaultdict(float))
    names = set()
    for f in _find(d, "*counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                k = short_name(r["Kernel_Name"])
                if kernel_sub and kernel_sub not in k:
                    continue
                per[(k, r["Dispatch_Id"])][r["Counter_Name"]] += float(r["Counter_Value"])
                names.add(r["Counter_Name"])
    agg = defaultdict(lambda: defaultdict(list))
    for (k, _), cv in sorted(per.items(), key=lambda kv: (kv[0][0], int(kv[0][1]))):
        for n, v in cv.items():
            agg[k][n].append(v)
    if last > 0:
        for k in agg:
            for n in agg[k]:
                agg[k][n] = agg[k][n][-last:]
    names = sorted(names)
    for k in sorted(agg):
        print(k, file=out)
        for n in names:
            v = agg[k].get(n)
            if v:
                print(f"    {n:<28} {sum(v) / len(v):>18.1f}   ({len(v)} dispatches)", file=out)


def traffic(fdir, wdir, copy_mib, kernel_sub):
    fetch, write = counters(fdir, "FETCH_SIZE"), counters(wdir, "WRITE_SIZE")
    copy_bytes = copy_mib * (1 << 20)
    ck = [k for k in fetch if "neg" in k.lower()]
    if not ck:
        raise SystemExit("copy calibration kernel (torch.neg) not found in the FETCH_SIZE run")
    cf = max(fetch[ck[0]]) * 1024.0
    cw = max(write[ck[0]]) * 1024.0
    sf, sw = copy_bytes / cf, copy_bytes / cw
    res = {"copy_kernel": ck[0], "copy_bytes": copy_bytes, "fetch_scale": sf, "write_scale": sw, "kernels": {}}
    for k in sorted(fetch):
        if kernel_sub and kernel_sub not in k:
            continue
        fv, wv = fetch[k], write.get(k, [0.0])
        # median over dispatches: steady state (a kernel's first dispatch may see a cold
        # or unsorted input)
        res["kernels"][k] = {"dispatches": len(fv), "fetch_bytes": sf * 1024.0 * median(fv),
                             "write_bytes": sw * 1024.0 * median(wv)}
    return res


def median(v):
    v = sorted(v)
    n = len(v)
    return v[n // 2] if n % 2 else 0.5 * (v[n // 2 - 1] + v[n // 2])


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=25)
    c = sub.add_parser("pmc")
    c.add_argument("dir")
    c.add_argument("--kernel", default="")
    c.add_argument("--last", type=int, default=0, help="average the last N dispatches of each kernel only")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--kernel", default="")
    t.add_argument("--json", default="")
    t.add_argument("--key", default="")
    t.add_argument("--shape", type=int, default=-1, help="launch shape the counter runs used (stored per shape)")
    a = ap.parse_args()
    if a.cmd == "stats":
        stats(a.dir, a.top)
        return
    if a.cmd == "pmc":
        pmc(a.dir, a.kernel, last=a.last)
        return
    res = traffic(a.fetch_dir, a.write_dir, a.copy_mib, a.kernel)
    print(f"copy kernel {res['copy_kernel']}: scale fetch {res['fetch_scale']:.4f}, write {res['write_scale']:.4f}")
    for k, v in res["kernels"].items():
        print(f"{k:<58} {v['dispatches']:>6} fetch {v['fetch_bytes'] / 1e6:>12.3f} MB  write {v['write_bytes'] / 1e6:>12.3f} MB")
    if a.json:
        data = {}
        if os.path.exists(a.json):
            with open(a.json) as fh:
                data = json.load(fh)
        # the evaluation kernel of the timed steps: the most dispatched one (the
        # autotuner's one-off trial launches of the other shapes are left out)
        ev = [(k, v) for k, v in res["kernels"].items() if "k_eval_" in k]
        if ev:
            name, big = max(ev, key=lambda kv: kv[1]["dispatches"])
            entry = {"kernel": name, "fetch_bytes": big["fetch_bytes"], "write_bytes": big["write_bytes"],
                     "fetch_scale": res["fetch_scale"], "write_scale": res["write_scale"],
                     "copy_bytes": res["copy_bytes"]}
            data.setdefault(a.key, {}).setdefault("shapes", {})[str(a.shape)] = entry
            with open(a.json, "w") as fh:
                json.dump(data, fh, indent=1)


if __name__ == "__main__":
    main()
