"""The profile summarizer (168861_amd/prof/summarize.py) on synthetic rocprofv3 CSV
output: timed-region selection (the last groups of overlapping dispatches), counter
units, the copy calibration and the merge of runs from separate boxes (CPU only)."""
import csv
import importlib.util
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SUMMARIZE = os.path.join(os.path.dirname(HERE), "168861_amd", "prof", "summarize.py")


def _load():
    spec = importlib.util.spec_from_file_location("summarize", SUMMARIZE)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _write(path, rows, fields):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        w.writerows(rows)


TREE = ["void nrt::k_batch_tree<double, 256, 0>(nrt::BatchSlot<double> const*)",
        "void nrt::k_batch_tree<double, 1024, 0>(nrt::BatchSlot<double> const*)"]
EXPL = "nrt::k_explicit_batch(nrt::ExSlot const*, HIP_vector_type<int, 2u> const*, long long const*, int, int)"


def test_reduce_and_merge(tmp_path):
    S = _load()
    assert S.core(TREE[0]) == "k_batch_tree" and S.core(EXPL) == "k_explicit_batch"
    assert S.FAMILIES["k_explicit_batch"](EXPL)
    assert not S.FAMILIES["k_explicit_batch"]("nrt::k_explicit_batch_trace(x)")
    raw_a, raw_b = str(tmp_path / "a"), str(tmp_path / "b")
    # trace: 6 launch groups of 2 overlapping tree dispatches (10 steps each), then 12
    # explicit dispatches; the last 4 groups and the last 10 explicit ones are timed
    trace, t, disp = [], 0, 0
    for g in range(6):
        for k, name in enumerate(TREE):
            disp += 1
            trace.append({"Dispatch_Id": disp, "Kernel_Name": name, "Start_Timestamp": t + k,
                          "End_Timestamp": t + 1000 + 100 * g + k})
        t += 10000
    for e in range(12):
        disp += 1
        trace.append({"Dispatch_Id": disp, "Kernel_Name": EXPL, "Start_Timestamp": t,
                      "End_Timestamp": t + 50 + e})
        t += 100
    _write(os.path.join(raw_a, "trace", "x_kernel_trace.csv"), trace,
           ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
    counters = []
    for r in trace:  # FETCH_SIZE in KiB, reported in two instance rows per dispatch
        for half in range(2):
            counters.append({"Dispatch_Id": r["Dispatch_Id"], "Kernel_Name": r["Kernel_Name"],
                             "Counter_Name": "FETCH_SIZE", "Counter_Value": 512})
    _write(os.path.join(raw_a, "fetch", "x_counter_collection.csv"), counters,
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
    # the copy calibration: 1 GiB read, FETCH_SIZE counts half of it
    _write(os.path.join(raw_a, "copy_fetch", "c_counter_collection.csv"),
           [{"Dispatch_Id": 1, "Kernel_Name": "copy_f64(double const*, double*, long long)",
             "Counter_Name": "FETCH_SIZE", "Counter_Value": 524288}],
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
    # another box: WRITE_SIZE and its calibration
    wr = [{"Dispatch_Id": r["Dispatch_Id"], "Kernel_Name": r["Kernel_Name"],
           "Counter_Name": "WRITE_SIZE", "Counter_Value": 256} for r in trace]
    _write(os.path.join(raw_b, "write", "x_counter_collection.csv"), wr,
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
    _write(os.path.join(raw_b, "copy_write", "c_counter_collection.csv"),
           [{"Dispatch_Id": 1, "Kernel_Name": "copy_f64(double const*, double*, long long)",
             "Counter_Name": "WRITE_SIZE", "Counter_Value": 1048576}],
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
    a, b, m = str(tmp_path / "a.json"), str(tmp_path / "b.json"), str(tmp_path / "m.json")
    S.reduce(raw_a, a, timed_steps=40, per_launch=10, classes=2)
    S.reduce(raw_b, b, timed_steps=40, per_launch=10, classes=2)
    S.merge(m, [a, b])
    res = json.load(open(m))
    tree = res["kernels"]["k_batch_tree<double,*>"]
    assert tree["timed_launches"] == 4 and tree["dispatches_per_launch_seen"] == [2, 2, 2, 2]
    assert tree["mean_launch_ns"] == sum(1000 + 100 * g + 1 for g in range(2, 6)) / 4
    # per launch: 2 dispatches x 1 MiB counted; x2 calibration for reads, x1 for writes
    assert tree["FETCH_SIZE_bytes_per_launch"] == 2 * 1024 * 1024
    assert abs(tree["hbm_read_bytes_per_launch"] - 4 * 1024 * 1024) < 1
    assert abs(tree["hbm_bytes_per_step"] - (4 * 1024 * 1024 + 2 * 256 * 1024) / 10) < 1
    ex = res["kernels"]["k_explicit_batch"]
    assert ex["timed_launches"] == 12 and ex["steps_per_launch"] == 1
    assert ex["mean_launch_ns"] == sum(50 + e for e in range(12)) / 12
    assert abs(ex["hbm_bytes_per_step"] - (2 * 1024 * 1024 + 256 * 1024)) < 1
    assert res["scale"]["FETCH_SIZE"] == 2.0 and res["scale"]["WRITE_SIZE"] == 1.0
