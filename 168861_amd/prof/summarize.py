"""Reduces the rocprofv3 output of profile_sweep.sh to a small JSON.

Timed region: bench.py times the last `timed_steps` steps of each sweep form in
launches of `steps_per_launch` steps; an implicit launch is one k_batch_tree
dispatch per size class of slots, running concurrently (a "group"). So:
  * k_batch_tree<double|float, *>: the last timed_steps / steps_per_launch groups
    of that precision (a group: dispatches overlapping in time); their span (first
    start to last end) and the summed counters of their dispatches;
  * k_explicit_batch: the last timed_steps dispatches (one per step).
Counters: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), and the same counters on
the copy kernels of prof/copy_calib, which move a known number of bytes: the
factor (known bytes / counted bytes) of the copy with the same access width
(8-byte lanes) scales the kernels' counts to HBM bytes.
"""
import argparse
import csv
import glob
import json
import os
import re


def rows(pattern):
    out = []
    for f in sorted(glob.glob(pattern, recursive=True)):
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    return out


def kname(r):
    return r.get("Kernel_Name") or r.get("Kernel-Name") or r.get("Name")


FAMILIES = {"k_batch_tree<double,*>": lambda n: "k_batch_tree<double" in n,
            "k_batch_tree<float,*>": lambda n: "k_batch_tree<float" in n,
            "k_explicit_batch": lambda n: core(n) == "k_explicit_batch"}


def core(name):
    """Bare function name: no namespace, template arguments or parameters."""
    n = re.sub(r"<.*>", "", re.sub(r"\(.*", "", name or ""))
    return n.strip().split(" ")[-1].split("::")[-1]


def counters(raw, tag, counter):
    """Per dispatch, in dispatch order: (kernel name, value in bytes)."""
    out = []
    for r in rows(os.path.join(raw, tag, "**", "*counter_collection.csv")):
        if r["Counter_Name"] != counter:
            continue
        out.append((int(r.get("Dispatch_Id", 0)), kname(r), float(r["Counter_Value"]) * 1024.0))
    out.sort()
    return [(n, v) for _, n, v in out]


def groups(dispatches):
    """Dispatches (start, end) sorted by start, merged while they overlap."""
    g = []
    for s, e in sorted(dispatches):
        if g and s <= g[-1][1]:
            g[-1] = (g[-1][0], max(g[-1][1], e), g[-1][2] + 1)
        else:
            g.append((s, e, 1))
    return g


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("raw")
    ap.add_argument("out")
    ap.add_argument("--timed-steps", type=int, default=40)
    ap.add_argument("--steps-per-launch", type=int, default=10)
    a = ap.parse_args()
    nlaunch = a.timed_steps // a.steps_per_launch
    res = {"timed_steps": a.timed_steps, "steps_per_launch": a.steps_per_launch, "stats": {}}
    for r in rows(os.path.join(a.raw, "trace", "**", "*kernel_stats.csv")):
        if any(f(r["Name"]) for f in FAMILIES.values()) or "copy_f64" in r["Name"]:
            res["stats"][r["Name"]] = {k: r[k] for k in ("Calls", "TotalDurationNs",
                                                         "AverageNs", "MinNs", "MaxNs")}
    trace = rows(os.path.join(a.raw, "trace", "**", "*kernel_trace.csv"))
    # copy calibration: known bytes / counted bytes, per access width and counter
    cal = {}
    for tag, counter in (("copy_fetch", "FETCH_SIZE"), ("copy_write", "WRITE_SIZE")):
        per = {}
        for n, v in counters(a.raw, tag, counter):
            per.setdefault(core(n), []).append(v)
        for n, vs in per.items():
            known = 8.0 * (1 << 27)
            cal.setdefault(n, {})[counter] = {"counted_bytes_mean": sum(vs) / len(vs),
                                              "known_bytes": known,
                                              "factor": known / (sum(vs) / len(vs))}
    res["copy_calibration"] = cal
    fx = {c: cal.get("copy_f64", {}).get(c, {}).get("factor") for c in ("FETCH_SIZE", "WRITE_SIZE")}
    fetch = counters(a.raw, "fetch", "FETCH_SIZE")
    write = counters(a.raw, "write", "WRITE_SIZE")
    res["kernels"] = {}
    for label, match in FAMILIES.items():
        mine = [r for r in trace if match(kname(r))]
        if not mine:
            continue
        d = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in mine]
        explicit = label == "k_explicit_batch"
        if explicit:
            last = sorted(d)[-a.timed_steps:]
            spans = [e - s for s, e in last]
            per_group, steps_per = 1, 1
        else:
            g = groups(d)[-nlaunch:]
            spans = [e - s for s, e, _ in g]
            per_group = max(n for _, _, n in g)
            steps_per = a.steps_per_launch
        k = {"timed_launches": len(spans), "dispatches_per_launch": per_group,
             "steps_per_launch": steps_per, "mean_launch_ns": sum(spans) / len(spans)}
        ntake = len(spans) * per_group
        for counter, vals in (("FETCH_SIZE", fetch), ("WRITE_SIZE", write)):
            mv = [v for n, v in vals if match(n)][-ntake:]
            if mv:
                k[counter + "_bytes_per_launch"] = sum(mv) / len(spans)
        if fx["FETCH_SIZE"] and fx["WRITE_SIZE"] and "FETCH_SIZE_bytes_per_launch" in k and \
                "WRITE_SIZE_bytes_per_launch" in k:
            hbm = (k["FETCH_SIZE_bytes_per_launch"] * fx["FETCH_SIZE"] +
                   k["WRITE_SIZE_bytes_per_launch"] * fx["WRITE_SIZE"])
            k["hbm_bytes_per_launch"] = hbm
            k["hbm_bytes_per_step"] = hbm / steps_per
        res["kernels"][label] = k
    json.dump(res, open(a.out, "w"), indent=1)
    print(json.dumps(res["kernels"], indent=1))


if __name__ == "__main__":
    main()
