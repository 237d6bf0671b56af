"""Reduces the rocprofv3 output of profile_sweep.sh to small JSON files.

  summarize.py reduce RAW OUT [--timed-steps N] [--steps-per-launch S] [--classes C]
      whatever RAW holds (kernel trace, FETCH_SIZE, WRITE_SIZE, copy calibration);
  summarize.py merge OUT IN...
      merges reduced files (each run on a box of its own) and scales the counters.

Timed region: bench.py times the last N steps of each sweep form in launches of S
steps; an implicit launch is one k_batch_tree dispatch per size class of slots (C
of them in the sweep), running concurrently (a "group"). So:
  * k_batch_tree<double|float, *>: the last N / S groups of that value type (a
    group: dispatches overlapping in time), their span (first start to last end),
    and the counters of their N / S * C dispatches (the last ones of that type);
  * k_explicit_batch: the last N dispatches (one per step).
Wave states: SQ_WAVE_CYCLES and the cycles waves spent waiting (SQ_WAIT_ANY: on a
memory, LDS or barrier dependency; SQ_WAIT_INST_ANY: for an instruction to issue) or
issuing (SQ_ACTIVE_INST_*), summed over the same dispatches.
Counters: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), the HBM-side request
counters TCC_EA0_RDREQ_DRAM_32B_sum and TCC_EA0_WRREQ_WRITE_DRAM_32B_sum (32-byte
units), and the same counters on the copy kernels of copy_calib, which move a known
number of bytes: the factor (known bytes / counted bytes) of the copy with the
kernels' access width (8-byte lanes, copy_f64) scales the kernels' counts to HBM
bytes.
"""
import csv
import glob
import json
import os
import re
import sys

FAMILIES = {"k_batch_tree<double,*>": lambda n: "k_batch_tree<double" in n,
            "k_batch_tree<float,*>": lambda n: "k_batch_tree<float" in n,
            "k_explicit_batch": lambda n: core(n) == "k_explicit_batch"}


def core(name):
    """Bare function name: no namespace, template arguments or parameters."""
    n = re.sub(r"<.*>", "", re.sub(r"\(.*", "", name or ""))
    return n.strip().split(" ")[-1].split("::")[-1]


def rows(pattern):
    out = []
    for f in sorted(glob.glob(pattern, recursive=True)):
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    return out


def kname(r):
    return r.get("Kernel_Name") or r.get("Kernel-Name") or r.get("Name")


# counter: (run tag, bytes per unit): FETCH_SIZE / WRITE_SIZE count KiB, the DRAM
# request counters 32-byte units
UNITS = {"FETCH_SIZE": ("fetch", 1024), "WRITE_SIZE": ("write", 1024),
         "TCC_EA0_RDREQ_DRAM_32B_sum": ("dram", 32),
         "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum": ("dram", 32)}
# wave-state counters (counts, summed over the timed dispatches)
SQ = ["SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY",
      "SQ_ACTIVE_INST_ANY", "SQ_ACTIVE_INST_VALU", "SQ_INSTS_VALU"]


_loaded = {}


def _run(raw, tag):
    """Counter name -> [(dispatch id, kernel name, summed value)], one pass per run."""
    key = (raw, tag)
    if key not in _loaded:
        acc = {}
        for f in sorted(glob.glob(os.path.join(raw, tag, "**", "*counter_collection.csv"),
                                  recursive=True)):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    k = (r["Counter_Name"], int(r.get("Dispatch_Id", 0)))
                    if k in acc:
                        acc[k][1] += float(r["Counter_Value"])
                    else:
                        acc[k] = [kname(r), float(r["Counter_Value"])]
        per = {}
        for (c, d), (n, v) in acc.items():
            per.setdefault(c, []).append((d, n, v))
        for v in per.values():
            v.sort()
        _loaded[key] = per
    return _loaded[key]


def counters(raw, tag, counter):
    """Per dispatch, in dispatch order: (kernel name, value in bytes, or a count)."""
    unit = UNITS[counter][1] if counter in UNITS else 1
    return [(n, v * unit) for _, n, v in _run(raw, tag).get(counter, [])]


def groups(dispatches):
    """Dispatches (start, end), merged while they overlap."""
    g = []
    for s, e in sorted(dispatches):
        if g and s <= g[-1][1]:
            g[-1] = (g[-1][0], max(g[-1][1], e), g[-1][2] + 1)
        else:
            g.append((s, e, 1))
    return g


def reduce(raw, out, timed_steps=40, per_launch=10, classes=3):
    nlaunch = timed_steps // per_launch
    res = {"timed_steps": timed_steps, "steps_per_launch": per_launch, "kernels": {}}
    stats = {}
    for r in rows(os.path.join(raw, "trace", "**", "*kernel_stats.csv")):
        if any(f(r["Name"]) for f in FAMILIES.values()):
            stats[r["Name"]] = {k: r[k] for k in ("Calls", "TotalDurationNs", "AverageNs",
                                                  "MinNs", "MaxNs")}
    if stats:
        res["stats_all_dispatches"] = stats
    trace = rows(os.path.join(raw, "trace", "**", "*kernel_trace.csv"))
    for label, match in FAMILIES.items():
        k = {}
        explicit = label == "k_explicit_batch"
        mine = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in trace
                if match(kname(r))]
        if mine:
            if explicit:
                spans = [e - s for s, e in sorted(mine)[-timed_steps:]]
            else:
                g = groups(mine)[-nlaunch:]
                spans = [e - s for s, e, _ in g]
                k["dispatches_per_launch_seen"] = [n for _, _, n in g]
            k["timed_launches"] = len(spans)
            k["steps_per_launch"] = 1 if explicit else per_launch
            k["mean_launch_ns"] = sum(spans) / len(spans)
        take = timed_steps if explicit else nlaunch * classes
        for counter, (tag, _) in UNITS.items():
            mv = [v for n, v in counters(raw, tag, counter) if match(n)][-take:]
            if mv:  # per launch: one dispatch (explicit) or a group of `classes`
                k[counter + "_bytes_per_launch"] = sum(mv) / (len(mv) if explicit else
                                                              len(mv) / classes)
                k["steps_per_launch"] = 1 if explicit else per_launch
        sq = {}
        for counter in SQ:
            mv = [v for n, v in counters(raw, "sq", counter) if match(n)][-take:]
            if mv:
                sq[counter] = sum(mv)
        if sq.get("SQ_WAVE_CYCLES"):
            wc = sq["SQ_WAVE_CYCLES"]
            sq["share_of_wave_cycles"] = {c: sq[c] / wc for c in (
                "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_ANY", "SQ_ACTIVE_INST_VALU")
                if c in sq}
            k["wave_states"] = sq
        if k:
            res["kernels"][label] = k
    cal = {}
    for counter, (tag, _) in UNITS.items():
        tag = "copy_" + tag
        per = {}
        for n, v in counters(raw, tag, counter):
            per.setdefault(core(n), []).append(v)
        for n, vs in per.items():
            known = 8.0 * (1 << 27)
            mean = sum(vs) / len(vs)
            cal.setdefault(n, {})[counter] = {"counted_bytes_mean": mean, "known_bytes": known,
                                              "factor": known / mean}
    if cal:
        res["copy_calibration"] = cal
    json.dump(res, open(out, "w"), indent=1)
    print(json.dumps(res, indent=1))


def merge(out, ins):
    res = {}

    def deep(a, b):
        for k, v in b.items():
            if isinstance(v, dict) and isinstance(a.get(k), dict):
                deep(a[k], v)
            else:
                a[k] = v
    for f in ins:
        deep(res, json.load(open(f)))
    cal = res.get("copy_calibration", {}).get("copy_f64", {})
    ff = cal.get("FETCH_SIZE", {}).get("factor")
    fw = cal.get("WRITE_SIZE", {}).get("factor")
    res["scale"] = {"FETCH_SIZE": ff, "WRITE_SIZE": fw, "from": "copy_f64 (8-byte lanes)"}
    dr = cal.get("TCC_EA0_RDREQ_DRAM_32B_sum", {}).get("factor")
    dw = cal.get("TCC_EA0_WRREQ_WRITE_DRAM_32B_sum", {}).get("factor")
    res["scale"].update({"DRAM_READ": dr, "DRAM_WRITE": dw})
    for k in res.get("kernels", {}).values():
        if dr and dw and "TCC_EA0_RDREQ_DRAM_32B_sum_bytes_per_launch" in k:
            k["dram_bytes_per_step"] = (
                k["TCC_EA0_RDREQ_DRAM_32B_sum_bytes_per_launch"] * dr +
                k["TCC_EA0_WRREQ_WRITE_DRAM_32B_sum_bytes_per_launch"] * dw) / k["steps_per_launch"]
        if ff and fw and "FETCH_SIZE_bytes_per_launch" in k and "WRITE_SIZE_bytes_per_launch" in k:
            k["hbm_read_bytes_per_launch"] = k["FETCH_SIZE_bytes_per_launch"] * ff
            k["hbm_write_bytes_per_launch"] = k["WRITE_SIZE_bytes_per_launch"] * fw
            k["hbm_bytes_per_launch"] = k["hbm_read_bytes_per_launch"] + k["hbm_write_bytes_per_launch"]
            k["hbm_bytes_per_step"] = k["hbm_bytes_per_launch"] / k["steps_per_launch"]
    json.dump(res, open(out, "w"), indent=1)
    print(json.dumps(res.get("kernels", {}), indent=1))


def main():
    if len(sys.argv) >= 4 and sys.argv[1] == "reduce":
        kw = {}
        args = sys.argv[4:]
        for i in range(0, len(args) - 1, 2):
            key = {"--timed-steps": "timed_steps", "--steps-per-launch": "per_launch",
                   "--classes": "classes"}[args[i]]
            kw[key] = int(args[i + 1])
        reduce(sys.argv[2], sys.argv[3], **kw)
    elif len(sys.argv) >= 4 and sys.argv[1] == "merge":
        merge(sys.argv[2], sys.argv[3:])
    else:
        sys.exit(__doc__)


if __name__ == "__main__":
    main()
